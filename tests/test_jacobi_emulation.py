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
