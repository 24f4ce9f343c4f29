"""Typed torch-tensor wrappers over the C ABI (include/hfb_hip.h).

Each function validates shapes/dtypes, passes raw device pointers and the
current HIP stream, and raises on a nonzero status.  Outputs are allocated
by the caller or here with torch (the library itself never allocates).
"""
from __future__ import annotations

import ctypes

import torch

from ctypes import c_void_p

from . import _lib
from ._lib import check, lib, need, ptr, stream_handle

# Optional timing of library calls with HIP events on the launching stream:
# set TIMERS to a dict name -> list and every timed call appends an
# (start, end) event pair.  While an iteration is being captured into HIP
# graphs, SEGMENTER splits the capture around timed calls so that replays can
# be bracketed by events outside the graphs.
TIMERS = None
SEGMENTER = None


class _timed:
    def __init__(self, name):
        self.name = name
        self.ev = None

    def __enter__(self):
        if self.name is None:
            return self
        if SEGMENTER is not None:
            SEGMENTER.cut(self.name)
        elif TIMERS is not None:
            self.ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            self.ev[0].record()
        return self

    def __exit__(self, *exc):
        if self.name is None:
            return False
        if SEGMENTER is not None:
            SEGMENTER.cut(None)
        elif self.ev is not None:
            self.ev[1].record()
            TIMERS.setdefault(self.name, []).append(self.ev)
        return False


def cderv(ctx, axis: int, kind: int, x: torch.Tensor, bmass: torch.Tensor | None = None):
    """One-axis derivative of fields shaped (..., nx, ny, nz), complex.

    kind 0 -> d1; kind 1 -> (d1, d2); kind 2 -> (d1, d(B d1)) (levels.py:131-395)."""
    prec = _lib.prec_of(x.dtype)
    need(x, name="psin")
    shp = tuple(x.shape)
    if shp[-3:] != (ctx.nx, ctx.ny, ctx.nz):
        raise ValueError(f"trailing shape {shp[-3:]} does not match the grid")
    nf = x.numel() // ctx.nvox
    o1 = torch.empty_like(x)
    o2 = torch.empty_like(x) if kind >= 1 else None
    if kind == 2:
        need(bmass, _lib.real_of(prec), (ctx.nx, ctx.ny, ctx.nz), "pos_func")
    check(lib().hfb_cderv(ctx.handle, prec, axis, kind, ptr(x), ptr(bmass), ptr(o1), ptr(o2), nf,
                          stream_handle()), "cderv")
    return o1 if kind == 0 else (o1, o2)


def hpsi(ctx, psi, npsi0, upot, bmass, aq, spot, wlspot, v_pair, w, wuv,
         want_pout=True, want_del=True, want_partial=True, out=None, timer=None, reduce_sums=True, weights_job=None):
    """Batched hpsi01 (meanfield.py:40-159) on psi (S, 2, nx, ny, nz).

    Returns (pout, pout_mf, pout_del, sums) where sums (S, 2) double holds
    (Re<psi|h psi>, Re<psi|pout>) without the wxyz factor, or None.  With
    reduce_sums=False, sums is an HSum of the unreduced tile partials, which
    precond_step and grstep_finish add up themselves.  ``weights_job``
    (from weights_job()) runs grstep_weights in the same launch; it must
    write the w, wuv given here."""
    prec = _lib.prec_of(psi.dtype)
    rt = _lib.real_of(prec)
    S = psi.shape[0]
    need(psi, _lib.cplx_of(prec), (S, 2, ctx.nx, ctx.ny, ctx.nz), "pinn")
    s4 = (2, ctx.nx, ctx.ny, ctx.nz)
    s5 = (2, 3, ctx.nx, ctx.ny, ctx.nz)
    for t, s, nm in ((upot, s4, "upot"), (bmass, s4, "bmass"), (v_pair, s4, "v_pair"),
                     (aq, s5, "aq"), (spot, s5, "spot"), (wlspot, s5, "wlspot")):
        need(t, rt, s, nm)
    need(w, torch.float64, (S,), "weight")
    need(wuv, torch.float64, (S,), "weightuv")
    if out is None:
        out = {}
    pmf = out.get("pout_mf")
    if pmf is None:
        pmf = torch.empty_like(psi)
    pout = out.get("pout")
    if pout is None and want_pout:
        pout = torch.empty_like(psi)
    pdel = out.get("pout_del")
    if pdel is None and want_del:
        pdel = torch.empty_like(psi)
    ntile = lib().hfb_hpsi_ntiles(ctx.handle, prec)
    if ntile < 0:
        check(ntile, "hpsi_ntiles")
    part = ctx.work(("hpsi_part", prec, S), (S, ntile, 2), torch.float64) if want_partial else None
    nt = ctypes.c_int()
    with _timed(timer):
        check(lib().hfb_hpsi_ex(ctx.handle, prec, ptr(psi), S, int(npsi0), ptr(upot), ptr(bmass), ptr(aq), ptr(spot),
                                ptr(wlspot), ptr(v_pair), ptr(w), ptr(wuv), ptr(pout), ptr(pmf), ptr(pdel),
                                ptr(part), ctypes.byref(nt),
                                ctypes.byref(weights_job) if weights_job is not None else None, stream_handle()),
              "hpsi")
    sums = None
    if want_partial:
        sums = HSum(part, nt.value)
        if reduce_sums:
            sums = reduce_hsum(ctx, sums)
    return pout, pmf, pdel, sums


class HSum:
    """Per-state (Re<psi|h psi>, Re<psi|pout>) as tile partials (S, ntile, 2)
    not yet added up."""

    def __init__(self, part, ntile):
        self.part, self.ntile = part, int(ntile)


def reduce_hsum(ctx, hs):
    """(S, 2) totals of an HSum (a tensor is returned unchanged)."""
    if not isinstance(hs, HSum):
        return hs
    S = hs.part.shape[0]
    sums = ctx.work(("hpsi_sums", S), (S, 2), torch.float64)
    check(lib().hfb_reduce_partials(ptr(hs.part), S, hs.ntile, 2, 1.0, ptr(sums), stream_handle()),
          "reduce_partials")
    return sums


def _hsum_args(hs):
    return (hs.part, hs.ntile) if isinstance(hs, HSum) else (hs, 0)


def poisson(ctx, rho: torch.Tensor, e2: float, wcoul: torch.Tensor | None = None):
    """Coulomb potential of rho[1] on the doubled grid (coulomb.py:54-75)."""
    need(rho, torch.float64, (2, ctx.nx, ctx.ny, ctx.nz), "rho")
    if wcoul is None:
        wcoul = torch.empty((ctx.nx, ctx.ny, ctx.nz), dtype=torch.float64, device=rho.device)
    work = ctx.work("poisson", tuple(ctx.n2), torch.complex128)
    check(lib().hfb_poisson(ctx.handle, ptr(rho), ptr(wcoul), float(e2), ptr(work), stream_handle()), "poisson")
    return wcoul


def poisson_noalloc(ctx, rho: torch.Tensor, e2: float, wcoul: torch.Tensor):
    """poisson with the library's static device scratch (no allocation in the
    library; wcoul is the caller's)."""
    need(rho, torch.float64, (2, ctx.nx, ctx.ny, ctx.nz), "rho")
    need(wcoul, torch.float64, (ctx.nx, ctx.ny, ctx.nz), "wcoul")
    check(lib().hfb_poisson_noalloc(ctx.handle, ptr(rho), ptr(wcoul), float(e2), stream_handle()), "poisson_noalloc")
    return wcoul


def selftest_mfma_f32(a: torch.Tensor, b: torch.Tensor):
    """Raw (64 lanes, 4 registers) result of one 16x16x4 FP32 matrix-core product."""
    need(a, torch.float32, (16, 4), "a")
    need(b, torch.float32, (4, 16), "b")
    d = torch.empty((64, 4), dtype=torch.float32, device=a.device)
    check(lib().hfb_selftest_mfma_f32(ptr(a), ptr(b), ptr(d), stream_handle()), "selftest_mfma_f32")
    return d


def selftest_mfma_f64(a: torch.Tensor, b: torch.Tensor):
    need(a, torch.float64, (16, 4), "a")
    need(b, torch.float64, (4, 16), "b")
    d = torch.empty(512, dtype=torch.float64, device=a.device)
    check(lib().hfb_selftest_mfma_f64(ptr(a), ptr(b), ptr(d), stream_handle()), "selftest_mfma")
    return d[:256].reshape(16, 16), d[256:].reshape(64, 4)


# ---------------------------------------------------------------------------
# Gradient step (static.py:93-259)
# ---------------------------------------------------------------------------
def weights_job(psi_states, npsi0, tbcs, wocc, wstates, wguv, pairwg, v_pair, nvox, out, vpair_max=None,
                tail=None):
    """The arguments of grstep_weights as a record for hpsi(..., weights_job=).
    ``tail``: a tail_job() the same workgroup runs first (vpair_max may then
    be the E[VMAX] slots it writes)."""
    j = _lib.WeightsJob(int(psi_states), int(npsi0), int(bool(tbcs)), ptr(wocc), ptr(wstates), ptr(wguv),
                        ptr(pairwg), ptr(v_pair), int(nvox), ptr(vpair_max), ptr(out["w"]), ptr(out["wuv"]),
                        ptr(out["pw"]), ptr(out["pshift"]))
    if tail is not None:
        j.tail = ctypes.pointer(tail)
        j._keep = tail
    return j


def tail_job(ctx, fc, dens, levels, E):
    """energy_tail's arguments as a record (for weights_job(tail=) and
    run_tail); keeps the force record alive."""
    nb = lib().hfb_skyrme_nblocks(ctx.handle)
    sky = ctx.work("sky_part", (3 * nb,), torch.float64)
    en = ctx.work("en_part", (nb * 20,), torch.float64)
    t = _lib.TailJob(ctypes.addressof(fc), ptr(sky), ptr(en), ptr(dens.current), int(levels.nstmax),
                     ptr(levels.wocc), ptr(levels.wstates), ptr(levels.sp_kinetic), ptr(levels.sp_energy),
                     ptr(levels.sp_orbital), ptr(levels.sp_spin), ptr(E))
    t._keep = fc
    return t


def run_tail(ctx, t):
    """energy_tail from a tail_job() record."""
    check(lib().hfb_energy_tail(ctx.handle, ctypes.c_void_p(t.force), ctypes.c_void_p(t.sky_part),
                                ctypes.c_void_p(t.en_part), ctypes.c_void_p(t.current), int(t.nstates),
                                ctypes.c_void_p(t.wocc), ctypes.c_void_p(t.wstates), ctypes.c_void_p(t.sp_kinetic),
                                ctypes.c_void_p(t.sp_energy), ctypes.c_void_p(t.sp_orbital),
                                ctypes.c_void_p(t.sp_spin), ctypes.c_void_p(t.E), stream_handle()), "energy_tail")


def grstep_weights(psi_states, npsi0, tbcs, wocc, wstates, wguv, pairwg, v_pair, nvox, out, vpair_max=None):
    """Per-state hpsi weights and preconditioner parameters into
    out['w','wuv','pw','pshift']; vpair_max (2,) device maxima of v_pair
    per isospin when known (else v_pair is scanned)."""
    S = int(psi_states)
    check(lib().hfb_grstep_weights(S, int(npsi0), int(bool(tbcs)), ptr(wocc), ptr(wstates), ptr(wguv), ptr(pairwg),
                                   ptr(v_pair), int(nvox), ptr(vpair_max), ptr(out["w"]), ptr(out["wuv"]),
                                   ptr(out["pw"]), ptr(out["pshift"]), stream_handle()), "grstep_weights")


def precond_step(ctx, psi, ps1, lagrange, hsum, wxyz, use_lagrange, pw, pshift, e0dmp, h2ma, x0dmp,
                 fluct=None):
    """psi <- psi - x0dmp * laplace(residual) in place (static.py:93-146).

    hsum: (S, 2) totals or an HSum.  If ``fluct`` is given it receives the
    fluctuation sums: a (S,2) float64 tensor gets (sum|r|^2, sum|lag -
    h_exp psi|^2)*wxyz; a dict gets {"part": tile partials, "ntile": n}
    for grstep_finish to add up."""
    prec = _lib.prec_of(psi.dtype)
    S = psi.shape[0]
    need(psi, _lib.cplx_of(prec), (S, 2, ctx.nx, ctx.ny, ctx.nz), "psi")
    need(ps1, psi.dtype, psi.shape, "ps1")
    need(lagrange, psi.dtype, psi.shape, "lagrange")
    work = ctx.work(("step", prec, S), tuple(psi.shape), psi.dtype)
    fpart = None
    nt = lib().hfb_step_ntiles(ctx.handle, prec)
    if fluct is not None:
        fpart = ctx.work(("step_part", prec, S), (2 * S, nt, 2), torch.float64)
    hp, hnt = _hsum_args(hsum)
    check(lib().hfb_precond_step(ctx.handle, prec, ptr(psi), ptr(ps1), ptr(lagrange), S, ptr(hp), hnt, float(wxyz),
                                 int(bool(use_lagrange)), ptr(pw), ptr(pshift), float(e0dmp), float(h2ma),
                                 float(x0dmp), ptr(fpart), ptr(work), stream_handle()), "precond_step")
    if isinstance(fluct, dict):
        fluct["part"], fluct["ntile"] = fpart, 2 * nt
    elif fluct is not None:
        check(lib().hfb_reduce_partials(ptr(fpart), S, 2 * nt, 2, float(wxyz), ptr(fluct), stream_handle()),
              "reduce_partials")
    return psi


def finish_job(hsum, wxyz, sp_energy, wocc, wstates, fluct, sp_efluct1, sp_efluct2, acc):
    """The arguments of grstep_finish as a record (gram2(..., finish_job=));
    the referenced tensors are kept alive on the record."""
    S = sp_energy.shape[0]
    hp, hnt = _hsum_args(hsum)
    if isinstance(fluct, dict):
        fp, fnt = fluct["part"], fluct["ntile"]
    else:
        fp, fnt = fluct, 0
    job = _lib.FinishJob(int(S), ptr(hp), int(hnt), float(wxyz), ptr(sp_energy), ptr(wocc), ptr(wstates), ptr(fp),
                         int(fnt), ptr(sp_efluct1), ptr(sp_efluct2), ptr(acc))
    job._keep = (hp, fp)
    return job


def grstep_finish(hsum, wxyz, sp_energy, wocc, wstates, fluct, sp_efluct1, sp_efluct2, acc):
    """hsum: (S, 2) totals or an HSum; fluct: None, (S, 2) totals or the dict
    filled by precond_step."""
    j = finish_job(hsum, wxyz, sp_energy, wocc, wstates, fluct, sp_efluct1, sp_efluct2, acc)
    check(lib().hfb_grstep_finish(j.nstates, j.hsum, j.hsum_ntile, j.wxyz, j.sp_energy, j.wocc, j.wstates, j.fsum,
                                  j.fsum_ntile, j.sp_efluct1, j.sp_efluct2, j.acc, stream_handle()), "grstep_finish")


# ---------------------------------------------------------------------------
# diagstep (static.py:261-577)
# ---------------------------------------------------------------------------
def gram(ctx, A, Bs, scale, out):
    """out[m] = scale * A^H Bs[m] for a state block A (nst, 2, nx, ny, nz)."""
    prec = _lib.prec_of(A.dtype)
    nst = A.shape[0]
    K = A[0].numel()
    nm = len(Bs)
    for b in Bs:
        need(b, A.dtype, A.shape, "B")
    need(out, torch.complex128, (nm, nst, nst), "gram_out")
    nwork = lib().hfb_gram_workspace(nst, K)
    if ctx is not None:
        part = ctx.work(("gram_part", nst), (nwork,), torch.float64)
    else:
        part = torch.empty(nwork, dtype=torch.float64, device=A.device)
    barr = (c_void_p * nm)(*[b.data_ptr() for b in Bs])
    check(lib().hfb_gram(prec, ptr(A), barr, nm, nst, K, float(scale), ptr(part), ptr(out), stream_handle()), "gram")
    return out


def gram2(ctx, blocks, scale, outs, finish_job=None):
    """Both isospin blocks in one launch: blocks[iq] = (A, [B_m]) and
    outs[iq] (nm, nst, nst) complex128 receives scale * A^H B_m.
    ``finish_job`` (from finish_job()) runs grstep_finish in the same launch."""
    A0 = blocks[0][0]
    prec = _lib.prec_of(A0.dtype)
    K = A0[0].numel()
    nm = len(blocks[0][1])
    As = (c_void_p * 2)()
    Bs = (c_void_p * 8)()
    parts = (c_void_p * 2)()
    op = (c_void_p * 2)()
    ns = (ctypes.c_int * 2)()
    for iq, (A, Bl) in enumerate(blocks):
        nst = A.shape[0]
        ns[iq] = nst
        if nst == 0:
            continue
        if len(Bl) != nm:
            raise ValueError("both isospin blocks need the same number of right-hand blocks")
        for b in Bl:
            need(b, A.dtype, A.shape, "B")
        need(outs[iq], torch.complex128, (nm, nst, nst), "gram_out")
        As[iq] = A.data_ptr()
        for m, b in enumerate(Bl):
            Bs[iq * 4 + m] = b.data_ptr()
        part = ctx.work(("gram_part", iq, nst), (lib().hfb_gram_workspace(nst, K),), torch.float64)
        parts[iq] = part.data_ptr()
        op[iq] = outs[iq].data_ptr()
    check(lib().hfb_gram2_ex(prec, As, Bs, nm, ns, K, float(scale), parts, op,
                             ctypes.byref(finish_job) if finish_job is not None else None, stream_handle()), "gram2")
    return outs


def _diag_work(nst, device):
    """Work matrices of the subspace kernel for one block above 36 states."""
    return torch.empty(5 * nst * nst, dtype=torch.complex128, device=device) if nst > 36 else None


def loewdin(S, U=None, sp_norm=None):
    """U = S^{-1/2} and sp_norm = Re diag S of one block's overlap S
    (loewdin_orthonormalize, static.py:463-482)."""
    nst = S.shape[0]
    need(S, torch.complex128, (nst, nst), "rhomatr")
    dev = S.device
    if U is None:
        U = torch.empty((nst, nst), dtype=torch.complex128, device=dev)
    if sp_norm is None:
        sp_norm = torch.empty(nst, dtype=torch.float64, device=dev)
    check(lib().hfb_loewdin(nst, ptr(S), ptr(U), ptr(sp_norm), ptr(_diag_work(nst, dev)), stream_handle()),
          "loewdin")
    return U, sp_norm


def hfb_matrices(ctx, iq, start, nst, U, Hold, Gold, ipair, tbcs, levels, static, E, lam):
    """construct_hfb_matrices (static.py:484-577) with the caller's unitary;
    lam receives lambda+ (lagrange = psi_new lam)."""
    nmax = static.hmatrix.shape[-1]
    for t, nm in ((U, "unitary"), (Hold, "h_matrix_old"), (lam, "lambda")):
        need(t, torch.complex128, (nst, nst), nm)
    if ipair:
        need(Gold, torch.complex128, (nst, nst), "gap_matrix_old")
    Ep = E.data_ptr()
    sz = 8
    check(lib().hfb_hfb_matrices(ctx.handle, int(iq), int(start), int(nst), int(nmax), ptr(U), ptr(Hold),
                                 ptr(Gold) if ipair else None, int(bool(ipair)), int(bool(tbcs)), ptr(levels.wocc),
                                 ptr(levels.wstates), ptr(levels.wguv), ptr(levels.pairwg), ptr(levels.sp_energy),
                                 ptr(levels.deltaf), ptr(static.hmatrix), ptr(static.gapmatrix), ptr(static.symcond),
                                 ptr(static.lambda_save), c_void_p(Ep + sz * EN["EFL1Q"]),
                                 c_void_p(Ep + sz * EN["EFL2Q"]), c_void_p(Ep + sz * EN["EFL1"]),
                                 c_void_p(Ep + sz * EN["EFL2"]), ptr(lam), ptr(_diag_work(nst, U.device)),
                                 stream_handle()), "hfb_matrices")
    return lam


def natural(phase, iq, start, nst, iternat, ipair, delstepbas, e0bas, npart, pair_cut, state_cut, softcut, levels,
            static, pairs, work, cm_out=None, m2_out=None):
    """One phase of hfb_natural's configuration-space sub-iterations."""
    nmax = static.hmatrix.shape[-1]
    check(lib().hfb_natural(int(phase), int(iq), int(start), int(nst), int(nmax), int(iternat), int(bool(ipair)),
                            float(delstepbas), float(e0bas), float(npart), float(pair_cut), float(state_cut),
                            float(softcut), ptr(static.hmatrix), ptr(static.gapmatrix), ptr(static.symcond),
                            ptr(static.lambda_save), ptr(levels.sp_energy), ptr(levels.deltaf), ptr(levels.wocc),
                            ptr(levels.wguv), ptr(levels.pairwg), ptr(levels.wstates), ptr(pairs.eferm),
                            ptr(pairs.epair), ptr(pairs.avdelt), ptr(pairs.avdeltv2), ptr(pairs.avg), ptr(work),
                            ptr(cm_out), ptr(m2_out), stream_handle()), "natural")


def lastdiag(iq, start, nst, eferm, static, wocc):
    """Quasi-particle energies and lower-component norms of one isospin."""
    nmax = static.hmatrix.shape[-1]
    dev = wocc.device
    work = torch.empty(nst * (nst + 1) // 2 + 2 * nst * nst, dtype=torch.complex128, device=dev)
    qe = torch.empty(nst, dtype=torch.float64, device=dev)
    qn = torch.empty(nst, dtype=torch.float64, device=dev)
    check(lib().hfb_lastdiag(int(iq), int(start), int(nst), int(nmax), float(eferm), ptr(static.hmatrix),
                             ptr(static.gapmatrix), ptr(wocc), ptr(qe), ptr(qn), ptr(work), stream_handle()),
          "lastdiag")
    return qe, qn


def rotate(A, U1, U2, out1, out2):
    """out1 = A U1, out2 = A U2 over states (psi_2d @ unitary, static.py:316, 557)."""
    prec = _lib.prec_of(A.dtype)
    nst = A.shape[0]
    K = A[0].numel()
    check(lib().hfb_rotate(prec, ptr(A), ptr(U1), ptr(U2), ptr(out1), ptr(out2), nst, K, stream_handle()), "rotate")


def rotate2(blocks, K):
    """rotate for both isospin blocks in one launch; blocks[iq] = (A, U1, U2,
    out1, out2) or None for an empty block.  K elements per state."""
    arrs = [(c_void_p * 2)() for _ in range(5)]
    ns = (ctypes.c_int * 2)()
    prec = None
    for iq, blk in enumerate(blocks):
        if blk is None:
            continue
        A = blk[0]
        p = _lib.prec_of(A.dtype)
        if prec is not None and p != prec:
            raise ValueError("both isospin blocks need the same precision")
        prec = p
        if A[0].numel() != K:
            raise ValueError("block length differs from K")
        ns[iq] = A.shape[0]
        for a, t in zip(arrs, blk):
            a[iq] = t.data_ptr() if t is not None else None
    if prec is None:
        return
    check(lib().hfb_rotate2(prec, *arrs, ns, K, stream_handle()), "rotate2")


def diag_small(ctx, nst0, nst1, nmax, mats, tdiag, ipair, tbcs, construct, levels, static, E, uo, work=None,
               jacobi_tol=0.0):
    """Loewdin, optional diagonalisation and HFB matrices for both isospins;
    ``jacobi_tol``: relative off-diagonal norm at which the lambda_lin
    eigensolve stops (0: the FP64 default 1e-13)."""
    m = (c_void_p * 8)(*[(t.data_ptr() if t is not None else None) for t in mats])
    u = (c_void_p * 4)(*[t.data_ptr() for t in uo])
    Ep = E.data_ptr()
    sz = 8
    check(lib().hfb_diag_small(ctx.handle, int(nst0), int(nst1), int(nmax), m, int(bool(tdiag)), int(bool(ipair)),
                               int(bool(tbcs)), int(bool(construct)), ptr(levels.wocc), ptr(levels.wstates),
                               ptr(levels.wguv), ptr(levels.pairwg), ptr(levels.sp_norm), ptr(levels.sp_energy),
                               ptr(levels.deltaf), ptr(static.hmatrix), ptr(static.gapmatrix), ptr(static.symcond),
                               ptr(static.lambda_save), c_void_p(Ep + sz * EN["EFL1Q"]),
                               c_void_p(Ep + sz * EN["EFL2Q"]), c_void_p(Ep + sz * EN["EFL1"]),
                               c_void_p(Ep + sz * EN["EFL2"]), u, ptr(work), c_void_p(Ep + sz * EN["JACOBI"]),
                               float(jacobi_tol), stream_handle()), "diag_small")


# ---------------------------------------------------------------------------
# Densities and single-particle properties
# ---------------------------------------------------------------------------
def density(ctx, psi, nst0, nst1, wocc, wstates, wguv, pairwg, cmtot, want_dens, want_sp):
    """Returns (dout, spart) scratch tensors filled by the three axis passes."""
    prec = _lib.prec_of(psi.dtype)
    nd = ctypes.c_longlong()
    nsp = ctypes.c_longlong()
    check(lib().hfb_density_sizes(ctx.handle, prec, ctypes.byref(nd), ctypes.byref(nsp)), "density_sizes")
    S = nst0 + nst1
    dout = ctx.work("dens_out", (nd.value,), torch.float64) if want_dens else None
    spart = ctx.work(("sp_part", S), (S * nsp.value,), torch.float64) if want_sp else None
    check(lib().hfb_density(ctx.handle, prec, ptr(psi), int(nst0), int(nst1), ptr(wocc), ptr(wstates), ptr(wguv),
                            ptr(pairwg), ptr(cmtot), int(bool(want_dens)), int(bool(want_sp)), ptr(dout), ptr(spart),
                            stream_handle()), "density")
    return dout, spart


def density_finish(ctx, dout, dens, mix, addnew=0.2, psi_dtype=torch.complex128):
    """dout as filled by density() for states of psi_dtype (its layout
    depends on the precision)."""
    prec = _lib.prec_of(psi_dtype)
    check(lib().hfb_density_finish(ctx.handle, prec, ptr(dout), ptr(dens.rho), ptr(dens.chi), ptr(dens.tau),
                                   ptr(dens.current), ptr(dens.sdens), ptr(dens.sodens), int(bool(mix)),
                                   float(addnew), stream_handle()), "density_finish")


def density_sp_finish(ctx, dout, dens, mix, addnew, spart, nst0, nst1, h2m, levels):
    """density_finish and sp_finish in one launch (the precision of levels.psi)."""
    prec = _lib.prec_of(levels.psi.dtype)
    check(lib().hfb_density_sp_finish(ctx.handle, prec, ptr(dout), ptr(dens.rho), ptr(dens.chi), ptr(dens.tau),
                                      ptr(dens.current), ptr(dens.sdens), ptr(dens.sodens), int(bool(mix)),
                                      float(addnew), ptr(spart), int(nst0), int(nst1), float(h2m[0]), float(h2m[1]),
                                      ptr(levels.sp_kinetic), ptr(levels.sp_orbital), ptr(levels.sp_spin),
                                      ptr(levels.sp_parity), stream_handle()), "density_sp_finish")


def sp_finish(ctx, spart, nst0, nst1, h2m, levels):
    check(lib().hfb_sp_finish(ctx.handle, ptr(spart), int(nst0), int(nst1), float(h2m[0]), float(h2m[1]),
                              ptr(levels.sp_kinetic), ptr(levels.sp_orbital), ptr(levels.sp_spin),
                              ptr(levels.sp_parity), stream_handle()), "sp_finish")


# ---------------------------------------------------------------------------
# Potentials and energies
# ---------------------------------------------------------------------------
EN = {"EHFT": 0, "EHF0": 1, "EHF1": 2, "EHF2": 3, "EHF3": 4, "EHFLS": 5, "EHFLSODD": 6, "EHFC": 7, "ECORC": 8,
      "EHFINT": 9, "E3CORR": 10, "EZPE": 11, "CRHO0": 12, "CRHO1": 13, "CDRHO0": 14, "CDRHO1": 15, "CTAU0": 16,
      "CTAU1": 17, "CDJ0": 18, "CDJ1": 19, "CJ0": 20, "CJ1": 21, "EHF": 22, "TKE": 23, "ORB": 24, "SPIN": 27,
      "JTOT": 30, "EFL1": 33, "EFL2": 34, "EFL1Q": 35, "EFL2Q": 37, "ECORRP": 39, "SUMFLU": 40, "DELESUM": 41,
      "EPAIR": 42, "JACOBI": 44, "NSITER": 46, "VMAX": 48}
EN_LEN = 50


def _deriv_work(ctx):
    n = lib().hfb_deriv_workspace(ctx.handle)
    if n < 0:
        check(int(n), "deriv_workspace")
    return ctx.work(("deriv_work",), (n,), torch.float64)


def density_derivs(ctx, dens, der):
    check(lib().hfb_density_derivs(ctx.handle, ptr(dens.rho), ptr(dens.sdens), ptr(dens.current), ptr(dens.sodens),
                                   ptr(der["divj"]), ptr(der["lap"]), ptr(der["grad"]), ptr(der["curls"]),
                                   ptr(der["curlj"]), ptr(_deriv_work(ctx)), stream_handle()), "density_derivs")


def density_derivs_coulomb(ctx, dens, der, e2, wcoul):
    """density_derivs and poisson(rho, e2, wcoul) in one call (fewer launches)."""
    need(dens.rho, torch.float64, (2, ctx.nx, ctx.ny, ctx.nz), "rho")
    need(wcoul, torch.float64, (ctx.nx, ctx.ny, ctx.nz), "wcoul")
    pwork = ctx.work("poisson", tuple(ctx.n2), torch.complex128)
    check(lib().hfb_density_derivs_coulomb(ctx.handle, ptr(dens.rho), ptr(dens.sdens), ptr(dens.current),
                                           ptr(dens.sodens), ptr(der["divj"]), ptr(der["lap"]), ptr(der["grad"]),
                                           ptr(der["curls"]), ptr(der["curlj"]), ptr(_deriv_work(ctx)), float(e2),
                                           ptr(wcoul), ptr(pwork), stream_handle()), "density_derivs_coulomb")


def meanfield_derivs(ctx, mf):
    check(lib().hfb_meanfield_derivs(ctx.handle, ptr(mf.aq), ptr(mf.bmass), ptr(mf.divaq_buf), ptr(mf.dbmass_buf),
                                     ptr(_deriv_work(ctx)), stream_handle()), "meanfield_derivs")


def skyrme(ctx, fc, dens, der, wcoul, mf, E, f32=None, finish=True):
    """finish=False leaves ecorrp and the v_pair maxima as partials (the
    "sky_part" workspace) for energy_tail."""
    nb = lib().hfb_skyrme_nblocks(ctx.handle)
    epart = ctx.work("sky_part", (3 * nb,), torch.float64)
    g = f32 or {}
    check(lib().hfb_skyrme(ctx.handle, ctypes.byref(fc), ptr(dens.rho), ptr(dens.tau), ptr(dens.chi),
                           ptr(dens.current), ptr(dens.sdens), ptr(der["divj"]), ptr(der["lap"]), ptr(der["grad"]),
                           ptr(der["curls"]), ptr(der["curlj"]), ptr(wcoul), ptr(mf.upot), ptr(mf.bmass),
                           ptr(mf.wlspot), ptr(mf.aq), ptr(mf.spot), ptr(mf.v_pair), ptr(g.get("upot")),
                           ptr(g.get("bmass")), ptr(g.get("wlspot")), ptr(g.get("aq")), ptr(g.get("spot")),
                           ptr(g.get("v_pair")), ptr(epart), ptr(E), int(bool(finish)), stream_handle()), "skyrme")


def integ_energy(ctx, fc, dens, der, wcoul, E, finish=True):
    """finish=False leaves the block partials (the "en_part" workspace) for
    energy_tail."""
    nb = lib().hfb_skyrme_nblocks(ctx.handle)
    epart = ctx.work("en_part", (nb * 20,), torch.float64)
    check(lib().hfb_integ_energy(ctx.handle, ctypes.byref(fc), ptr(dens.rho), ptr(dens.tau), ptr(dens.current),
                                 ptr(dens.sdens), ptr(der["divj"]), ptr(der["lap"]), ptr(der["curlj"]), ptr(wcoul),
                                 ptr(epart), ptr(E), int(bool(finish)),
                                 stream_handle()), "integ_energy")


def skyrme_energy(ctx, fc, dens, der, wcoul, mf, f32=None):
    """skyrme(..., finish=False) and integ_energy(..., finish=False) in one
    launch; energy_tail finishes both."""
    nb = lib().hfb_skyrme_nblocks(ctx.handle)
    sky = ctx.work("sky_part", (3 * nb,), torch.float64)
    en = ctx.work("en_part", (nb * 20,), torch.float64)
    g = f32 or {}
    check(lib().hfb_skyrme_energy(ctx.handle, ctypes.byref(fc), ptr(dens.rho), ptr(dens.tau), ptr(dens.chi),
                                  ptr(dens.current), ptr(dens.sdens), ptr(der["divj"]), ptr(der["lap"]),
                                  ptr(der["grad"]), ptr(der["curls"]), ptr(der["curlj"]), ptr(wcoul), ptr(mf.upot),
                                  ptr(mf.bmass), ptr(mf.wlspot), ptr(mf.aq), ptr(mf.spot), ptr(mf.v_pair),
                                  ptr(g.get("upot")), ptr(g.get("bmass")), ptr(g.get("wlspot")), ptr(g.get("aq")),
                                  ptr(g.get("spot")), ptr(g.get("v_pair")), ptr(sky), ptr(en), stream_handle()),
          "skyrme_energy")


def energy_tail(ctx, fc, dens, levels, E):
    """ecorrp, the integrals and sum_energy from the partials that
    skyrme(..., finish=False) and integ_energy(..., finish=False) left."""
    run_tail(ctx, tail_job(ctx, fc, dens, levels, E))


def sum_energy(levels, E):
    check(lib().hfb_sum_energy(int(levels.nstmax), ptr(levels.wocc), ptr(levels.wstates), ptr(levels.sp_kinetic),
                               ptr(levels.sp_energy), ptr(levels.sp_orbital), ptr(levels.sp_spin), ptr(E),
                               stream_handle()), "sum_energy")


# ---------------------------------------------------------------------------
# Energy-functional backward (no reference counterpart)
# ---------------------------------------------------------------------------
def energy_bwd(ctx, forces_c, tx, gout, dens, der, wcoul, want_fields=True, want_couplings=True):
    """Gradients of ehfint for a batch of coupling sets sharing the densities.

    forces_c: list of HfbForce (host), tx: (B, 10) float64 device raw
    couplings, gout: (B,) float64 device upstream gradients.  Returns
    (fields dict or None, d_coupling (B, 10) or None); the density gradients
    are summed over the batch with gout."""
    B = len(forces_c)
    arr = (type(forces_c[0]) * B)(*forces_c)
    need(tx, torch.float64, (B, 10), "couplings")
    need(gout, torch.float64, (B,), "grad_output")
    dev = dens.rho.device
    s4 = tuple(dens.rho.shape)
    s5 = tuple(dens.sodens.shape)
    fields = None
    if want_fields:
        fields = {k: torch.empty(s, dtype=torch.float64, device=dev)
                  for k, s in (("rho", s4), ("tau", s4), ("sodens", s5), ("current", s5), ("sdens", s5))}
    dcpl = torch.empty((B, 10), dtype=torch.float64, device=dev) if want_couplings else None
    part = ctx.work("adj_part", (lib().hfb_energy_bwd_nblocks(ctx.handle) * 72,), torch.float64)
    g = fields or {}
    check(lib().hfb_energy_bwd(ctx.handle, ctypes.cast(arr, c_void_p), ptr(tx), ptr(gout), B, ptr(dens.rho),
                               ptr(dens.tau), ptr(dens.sodens), ptr(dens.current), ptr(dens.sdens),
                               ptr(der["divj"]), ptr(der["lap"]), ptr(der["grad"]), ptr(der["curls"]),
                               ptr(der["curlj"]), ptr(wcoul), ptr(g.get("rho")), ptr(g.get("tau")),
                               ptr(g.get("sodens")), ptr(g.get("current")), ptr(g.get("sdens")), ptr(dcpl),
                               ptr(part), stream_handle()), "energy_bwd")
    return fields, dcpl
