"""CPU emulation of the device Jacobi eigensolver's work split
(jacobi_packed in jax-hfbfft_amd/csrc/hfb_linalg.hip), without a GPU.

Each phase of a round is emulated with the kernel's own index formulas for
every work item at once: the round-robin pairs, the packed lower-triangle
offsets ploff/prow, the block decode (ka, kb) of the in-place update and the
(pair, row) decode of the eigenvector update.  The emulation asserts that
every access stays inside the arrays the kernel sizes for the block
(n (n + 1) / 2 packed entries, n x n eigenvector scratch, NMAXJ / 2 pair
slots) and that within one phase no two work items write the same entry or
read an entry another item writes (the in-place update's correctness
condition).  The rotations then run in FP64 and the eigenvalues and
eigenvectors are compared with numpy.linalg.eigh (lower triangle), for the
block sizes of every kernel instance (LDS, 64-, 128- and 256-state)."""
import math

import numpy as np
import pytest

NMAXJ = 256


def ploff(i):
    return (i * (i + 1)) >> 1


def prow(e):
    e = np.asarray(e)
    i = ((np.sqrt(8.0 * e + 1.0) - 1.0) * 0.5).astype(np.int64)
    i = np.where(ploff(i + 1) <= e, i + 1, i)
    i = np.where(ploff(i) > e, i - 1, i)
    return i


def pairs_of_round(n, t):
    m = n + (n & 1)
    m1 = m - 1
    npair = m // 2
    k = np.arange(npair)
    xa = k - 1 + t
    xa = np.where(xa >= m1, xa - m1, xa)
    xb = m1 - k - 1 + t
    xb = np.where(xb >= m1, xb - m1, xb)
    a0 = np.where(k == 0, 0, xa + 1)
    b0 = xb + 1
    p = np.minimum(a0, b0)
    q = np.maximum(a0, b0)
    q = np.where(q < n, q, -1)
    return p, q


def packed_index(i, j):
    """(storage index, conjugate?) of entry (i, j)."""
    i, j = np.asarray(i), np.asarray(j)
    lo = i >= j
    return np.where(lo, ploff(i) + j, ploff(j) + i), ~lo


def rsq(x):
    return 1.0 / math.sqrt(x)


def jacobi_emulated(M, sort=True, max_sweeps=40):
    n = M.shape[0]
    tri = ploff(n)
    # load_lower_packed
    e = np.arange(tri)
    i = prow(e)
    j = e - ploff(i)
    assert (i >= 0).all() and (i < n).all() and (j >= 0).all() and (j <= i).all()
    A = M[i, j].copy()
    A[i == j] = A[i == j].real
    Vt = np.eye(n, dtype=complex)
    m = n + (n & 1)
    npair = m // 2
    assert npair <= NMAXJ // 2
    nblk = ploff(npair)
    sweeps = 0
    for sweep in range(max_sweeps):
        dia = np.sum(np.abs(A[i == j]) ** 2)
        off = 2.0 * np.sum(np.abs(A[i != j]) ** 2)
        if off <= 1e-26 * dia or off == 0.0:  # the kernel's JACOBI_TOL2
            break
        sweeps += 1
        for t in range(m - 1):
            p, q = pairs_of_round(n, t)
            assert (p >= 0).all() and (p < n).all() and (q < n).all()
            used = np.concatenate([p, q[q >= 0]])
            assert len(np.unique(used)) == len(used)  # disjoint pairs
            c = np.ones(npair)
            s = np.zeros(npair)
            ea = np.ones(npair, dtype=complex)
            for k in range(npair):
                if q[k] < 0:
                    continue
                app, aqq = A[ploff(p[k]) + p[k]].real, A[ploff(q[k]) + q[k]].real
                apq = np.conj(A[ploff(q[k]) + p[k]])
                r2 = abs(apq) ** 2
                den = abs(app) + abs(aqq)
                if r2 > 0.0 and r2 > 1e-36 * den * den:
                    ri = rsq(r2)
                    ea[k] = apq * ri
                    th = 0.5 * (aqq - app) * ri
                    h = 1.0 + th * th
                    rt = h * rsq(h)
                    tt = (1.0 if th >= 0.0 else -1.0) / (abs(th) + rt)
                    c[k] = rsq(1.0 + tt * tt)
                    s[k] = tt * c[k]
            # block update: all items of the phase at once
            it = np.arange(nblk)
            ka = prow(it)
            kb = it - ploff(ka)
            assert (ka < npair).all() and (kb >= 0).all() and (kb <= ka).all()
            pa, qa, pb, qb = p[ka], q[ka], p[kb], q[kb]
            has = [np.ones(nblk, bool), qb >= 0, qa >= 0, (qa >= 0) & (qb >= 0)]
            rows = [pa, pa, qa, qa]
            cols = [pb, qb, pb, qb]
            X = []
            idx = []
            for h, r_, c_ in zip(has, rows, cols):
                k_, cj = packed_index(np.where(h, r_, 0), np.where(h, c_, 0))
                assert ((k_[h] >= 0) & (k_[h] < tri)).all()
                v = np.where(cj, np.conj(A[k_]), A[k_])
                X.append(np.where(h, v, 0.0))
                idx.append((k_, cj, h))
            if sweep == 0 and t < 4:
                # the in-place condition: the entries written by one block are
                # read and written by no other block of the round
                owner = {}
                for (k_, cj, h) in idx:
                    for bb in np.nonzero(h)[0]:
                        prev = owner.setdefault(int(k_[bb]), int(bb))
                        assert prev == bb, "two blocks touch one entry"
            eca, esa = ea[ka] * c[ka], ea[ka] * s[ka]
            ecb, esb = ea[kb] * c[kb], ea[kb] * s[kb]
            ca, sa, cb, sb = c[ka], s[ka], c[kb], s[kb]
            y00 = X[0] * ecb - X[1] * sb
            y01 = X[0] * esb + X[1] * cb
            y10 = X[2] * ecb - X[3] * sb
            y11 = X[2] * esb + X[3] * cb
            n00 = np.conj(eca) * y00 - y10 * sa
            n01 = np.conj(eca) * y01 - y11 * sa
            n10 = np.conj(esa) * y00 + y10 * ca
            n11 = np.conj(esa) * y01 + y11 * ca
            diag = ka == kb
            n10 = np.where(diag & (qa >= 0) & (sa != 0.0), 0.0, n10)
            n00 = np.where(diag, n00.real, n00)
            n11 = np.where(diag, n11.real, n11)
            newA = A.copy()
            # off-diagonal blocks: pst of all four; diagonal blocks: (p,p), (q,p), (q,q)
            for (k_, cj, h), v, use in zip(idx, [n00, n01, n10, n11], [h_ for h_ in has]):
                sel = use & ~diag
                newA[k_[sel]] = np.where(cj[sel], np.conj(v[sel]), v[sel])
            dsel = diag
            newA[ploff(pa[dsel]) + pa[dsel]] = n00[dsel]
            dq = dsel & (qa >= 0)
            newA[ploff(qa[dq]) + pa[dq]] = n10[dq]
            newA[ploff(qa[dq]) + qa[dq]] = n11[dq]
            A = newA
            # eigenvector update, items (pair k, row r)
            itv = np.arange(n * npair)
            kk = itv // n
            rr = itv - kk * n
            assert (kk < npair).all() and (rr < n).all()
            pk, qk = p[kk], q[kk]
            act = (qk >= 0) & (s[kk] != 0.0)
            assert (pk[act] * n + rr[act] < n * n).all() and (qk[act] * n + rr[act] < n * n).all()
            xp = Vt[pk[act], rr[act]]
            xq = Vt[qk[act], rr[act]]
            e_ = ea[kk[act]]
            Vt[pk[act], rr[act]] = xp * e_ * c[kk[act]] - xq * s[kk[act]]
            Vt[qk[act], rr[act]] = xp * e_ * s[kk[act]] + xq * c[kk[act]]
    ev = np.array([A[ploff(x) + x].real for x in range(n)])
    perm = np.arange(n)
    if sort:
        perm = np.empty(n, dtype=int)
        for x in range(n):
            r = sum(1 for y in range(n) if ev[y] < ev[x] or (ev[y] == ev[x] and y < x))
            perm[r] = x
    V = Vt[perm, :].T
    return ev[perm], V, sweeps


@pytest.mark.parametrize("n", [1, 2, 3, 8, 16, 20, 33, 36, 40, 64, 65, 82, 126, 128, 167])
def test_jacobi_work_split_and_result(n):
    rng = np.random.default_rng(n)
    H = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = H + H.conj().T
    # a degenerate cluster, as in spherical shells
    if n >= 8:
        H[:4, :4] = np.diag([3.0, 3.0, 3.0, 3.0])
        H[:4, 4:] *= 1e-3
        H[4:, :4] *= 1e-3
    lo = np.tril(H) + np.triu(rng.standard_normal((n, n)), 1) * 7.0  # upper triangle must not be read
    ev, V, sweeps = jacobi_emulated(lo, sort=True)
    w = np.linalg.eigvalsh(H, UPLO="L")
    assert np.allclose(ev, w, atol=1e-10 * max(1.0, np.max(np.abs(w))))
    assert np.allclose(V.conj().T @ V, np.eye(n), atol=1e-12)
    assert np.allclose(H @ V, V * ev[None, :], atol=1e-9 * max(1.0, np.max(np.abs(w))))
    assert sweeps < 40


# ---------------------------------------------------------------------------
# jacobi_wave: the rounds in the first jw_nrw(n) wavefronts (one 2x2 block of
# the lower block triangle per lane, the matrix double-buffered between
# rounds), the eigenvector rotations in the other wavefronts one sweep
# behind, from a double-buffered rotation log.

NDIAG_LDS = 36
LDS_BYTES = 160 * 1024
DIAG_STATIC_LDS = 2848  # diag_kernel's static LDS for blocks in LDS (compiler resource report)


def rr_pair(k, t, m):
    m1 = m - 1
    xa = k - 1 + t
    xa = np.where(xa >= m1, xa - m1, xa)
    xb = m1 - k - 1 + t
    xb = np.where(xb >= m1, xb - m1, xb)
    a0 = np.where(k == 0, 0, xa + 1)
    b0 = xb + 1
    return np.minimum(a0, b0), np.maximum(a0, b0)


def instance(n):
    """(NP, NT) of the diag_kernel instance a block of n states runs in."""
    if n <= 16:
        return 16, 256
    if n <= 32:
        return 32, 512
    return 64, 512


def jw_m(n):
    return n + (n & 1)


def jw_log_bytes(n):
    m = jw_m(n)
    return 2 * (m - 1) * (m // 2) * 24


def jw_extra_bytes(n):
    m = jw_m(n)
    return jw_log_bytes(n) + (m * m * 16 if n & 1 else 0)


def jw_nrw(n):
    npair = jw_m(n) // 2
    return (npair * (npair + 1) // 2 + 63) // 64


def jw_rot(app, aqq, apq):
    r2 = abs(apq) ** 2
    den = abs(app) + abs(aqq)
    if r2 > 0.0 and r2 > 1e-36 * den * den:
        d = 0.5 * (aqq - app)
        rho2 = d * d + r2
        rho = rho2 * rsq(rho2)
        w = rho + abs(d)
        f = rsq(2.0 * rho * w)
        g = f if d >= 0.0 else -f
        return w * f, apq * g
    return 1.0, 0.0j


def jacobi_wave_emulated(M, max_sweeps=40):
    n = M.shape[0]
    NP, NT = instance(n)
    if n == 1:
        return np.array([M[0, 0].real]), np.ones((1, 1), complex), 0
    m = n + (n & 1)
    npair, nrnd = m // 2, m - 1
    nblk = npair * (npair + 1) // 2
    na = jw_nrw(n)
    assert na <= 3 and na < NT // 64  # jw_sweep instances 1..3, at least one eigenvector wavefront
    lsz = nrnd * npair
    # LDS: five n x n work matrices (B0..B4), the log after B4, for odd n the
    # second matrix buffer after the log (for even n it is B0); the first
    # buffer spans B2..B3
    assert m * m <= 2 * n * n
    if n % 2 == 0:
        assert m * m <= n * n
    assert 16 * 5 * n * n + jw_extra_bytes(n) + DIAG_STATIC_LDS <= LDS_BYTES
    A = [np.zeros((m, m), dtype=complex), np.full((m, m), np.nan, dtype=complex)]
    il, jl = np.tril_indices(n)
    A[0][il, jl] = M[il, jl]
    A[0][jl, il] = np.conj(M[il, jl])
    A[0][np.arange(n), np.arange(n)] = np.real(np.diag(M))
    V = np.eye(n, dtype=complex)
    lc = np.full(2 * lsz, np.nan)
    ls = np.full(2 * lsz, np.nan, dtype=complex)
    # (wavefront, lane) -> block, the kernel's decode
    blocks = []
    for wave in range(na):
        for lane in range(64):
            blk = lane + 64 * wave
            a = int((math.sqrt(8.0 * blk + 1.0) - 1.0) * 0.5)
            while a * (a + 1) // 2 > blk:
                a -= 1
            while (a + 1) * (a + 2) // 2 <= blk:
                a += 1
            if blk < nblk:
                blocks.append(((wave, lane), a, blk - a * (a + 1) // 2))
    assert sorted((a, b) for _, a, b in blocks) == [(a, b) for a in range(npair) for b in range(a + 1)]
    # eigenvector items: row r in wavefront na + r % (NT/64 - na)
    nvw = NT // 64 - na
    vitems = {}
    for w in range(nvw):
        nrow = (n - w + nvw - 1) // nvw if w < n else 0
        items = []
        for it in range(nrow * npair):
            lr, k = divmod(it, npair)
            items.append((it % 64, w + nvw * lr, k))
        vitems[w] = items
    assert sorted((r, k) for w in vitems for _, r, k in vitems[w]) == [(r, k) for r in range(n) for k in range(npair)]

    def norms(Acur):  # as the kernel sums them: off-diagonal entries apart from the diagonal
        od = Acur - np.diag(np.diag(Acur))
        return np.sum(np.abs(od) ** 2), np.sum(np.abs(np.diag(Acur)) ** 2)

    def apply_v(buf, rounds=None):
        # one rounds wavefront: a whole sweep from the sweep's log slots
        # (buf * lsz + t * npair + k); several: one round R from the two-slot
        # log ((R & 1) * npair + k)
        for t in (range(nrnd) if rounds is None else [rounds % nrnd]):
            for w in range(nvw):
                touched = set()
                for _, r, k in vitems[w]:
                    p, q = (int(x) for x in rr_pair(k, t, m))
                    if q < n:
                        idx = buf * lsz + t * npair + k if rounds is None else (rounds & 1) * npair + k
                        assert 0 <= idx < 2 * lsz and not np.isnan(lc[idx])
                        assert 0 <= r * n + p < n * n and 0 <= r * n + q < n * n
                        assert (r, p) not in touched and (r, q) not in touched
                        touched |= {(r, p), (r, q)}
                        c, sg = lc[idx], ls[idx]
                        vp, vq = V[r, p], V[r, q]
                        V[r, p] = vp * c - vq * np.conj(sg)
                        V[r, q] = vp * sg + vq * c

    off, dia = norms(A[0])
    sweep = 0
    R = 0
    while True:
        buf = sweep & 1
        stop = sweep >= max_sweeps or off <= 1e-26 * dia or off == 0.0
        if not stop:
            if na == 1:
                lc[buf * lsz:(buf + 1) * lsz] = np.nan
            for t in range(nrnd):
                cur, nxt = A[R & 1], A[(R + 1) & 1]
                nxt[:] = np.nan
                writes = {}
                for (wl, a, b) in blocks:
                    pa, qa = (int(x) for x in rr_pair(a, t, m))
                    pb, qb = (int(x) for x in rr_pair(b, t, m))
                    assert 0 <= pa < qa < m and 0 <= pb < qb < m
                    X = cur[np.ix_([pa, qa], [pb, qb])]
                    ca, sa = jw_rot(cur[pa, pa].real, cur[qa, qa].real, cur[pa, qa])
                    cb, sb = jw_rot(cur[pb, pb].real, cur[qb, qb].real, cur[pb, qb])
                    assert not np.isnan(X).any()
                    Y = np.empty((2, 2), complex)
                    Y[:, 0] = X[:, 0] * cb - X[:, 1] * np.conj(sb)
                    Y[:, 1] = X[:, 0] * sb + X[:, 1] * cb
                    N = np.empty((2, 2), complex)
                    N[0] = Y[0] * ca - sa * Y[1]
                    N[1] = np.conj(sa) * Y[0] + Y[1] * ca
                    if a == b:
                        if sa != 0.0:
                            N[0, 1] = N[1, 0] = 0.0
                            N[0, 0] = N[0, 0].real
                            N[1, 1] = N[1, 1].real
                        idx = buf * lsz + t * npair + a if na == 1 else (R & 1) * npair + a
                        assert 0 <= idx < 2 * lsz
                        lc[idx], ls[idx] = ca, sa
                    ent = [((pa, pb), N[0, 0]), ((pa, qb), N[0, 1]), ((qa, pb), N[1, 0]), ((qa, qb), N[1, 1])]
                    if a != b:
                        ent += [((y, x), np.conj(v)) for (x, y), v in ent]
                    for (x, y), v in ent:
                        assert 0 <= x < m and 0 <= y < m
                        assert writes.setdefault((x, y), wl) == wl, "two lanes write one entry"
                        nxt[x, y] = v
                assert len(writes) == m * m  # a round rewrites the whole matrix once
                if na > 1:  # the eigenvector wavefronts, one round behind (barrier per round)
                    if R > 0:
                        apply_v(None, R - 1)
                R += 1
            off, dia = norms(A[R & 1])
        if na == 1 and sweep > 0:
            apply_v(buf ^ 1)
        if na > 1 and stop and R > 0:
            apply_v(None, R - 1)
        if stop:
            break
        sweep += 1
    Af = A[R & 1]
    assert R == sweep * nrnd
    if n & 1:  # the extra index stays a zero row and column
        assert np.all(Af[n, :] == 0) and np.all(Af[:, n] == 0)
    ev = np.real(np.diag(Af))[:n]
    perm = np.empty(n, dtype=int)
    for x in range(n):
        perm[sum(1 for y in range(n) if ev[y] < ev[x] or (ev[y] == ev[x] and y < x))] = x
    return ev[perm], V[:, perm], sweep


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 8, 12, 16, 17, 20, 21, 24, 25, 31, 32, 33, 35, 36])
def test_wave_jacobi_work_split_and_result(n):
    rng = np.random.default_rng(100 + n)
    H = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = H + H.conj().T
    if n >= 8:  # a degenerate cluster, as in spherical shells
        H[:4, :4] = np.diag([3.0, 3.0, 3.0, 3.0])
        H[:4, 4:] *= 1e-3
        H[4:, :4] *= 1e-3
    lo = np.tril(H) + np.triu(rng.standard_normal((n, n)), 1) * 7.0  # upper triangle must not be read
    ev, V, sweeps = jacobi_wave_emulated(lo)
    w = np.linalg.eigvalsh(H, UPLO="L")
    sc = max(1.0, np.max(np.abs(w)))
    assert np.allclose(ev, w, atol=1e-12 * sc)
    assert np.allclose(V.conj().T @ V, np.eye(n), atol=1e-13)
    assert np.allclose(H @ V, V * ev[None, :], atol=1e-11 * sc)
    assert (n == 1 or 0 < sweeps) and sweeps < 40


def test_wave_jacobi_nearly_diagonal_sweeps():
    """Near convergence of the SCF iteration lambda_lin is nearly diagonal in
    the previous eigenbasis, with a degenerate shell: the wavefront rounds
    (rotation of jw_rot) need no more sweeps than the packed-triangle rounds
    with the same pair ordering and stopping rule."""
    n = 20
    rng = np.random.default_rng(7)
    d = np.sort(rng.uniform(-50.0, 0.0, n))
    d[4:10] = d[4]  # a six-fold shell
    E = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    for eps in (1e-5, 1e-7):
        H = np.diag(d) + eps * (E + E.conj().T)
        ev, V, sweeps = jacobi_wave_emulated(np.tril(H))
        _, _, sweeps_packed = jacobi_emulated(np.tril(H))
        assert sweeps <= sweeps_packed
        assert np.allclose(H @ V, V * ev[None, :], atol=1e-11 * 50)
