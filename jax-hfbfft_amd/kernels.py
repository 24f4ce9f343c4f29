"""Typed torch-tensor wrappers over the C ABI (include/hfb_hip.h).

Each function validates shapes/dtypes, passes raw device pointers and the
current HIP stream, and raises on a nonzero status.  Outputs are allocated
by the caller or here with torch (the library itself never allocates).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import check, lib, need, ptr, stream_handle


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
         want_pout=True, want_del=True, want_partial=True, out=None):
    """Batched hpsi01 (meanfield.py:40-159) on psi (S, 2, nx, ny, nz).

    Returns (pout, pout_mf, pout_del, sums) where sums (S, 2) double holds
    (Re<psi|h psi>, Re<psi|pout>) without the wxyz factor, or None."""
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
    pout = out.get("pout", torch.empty_like(psi) if want_pout else None)
    pdel = out.get("pout_del", torch.empty_like(psi) if want_del else None)
    ntile = lib().hfb_hpsi_ntiles(ctx.handle, prec)
    if ntile < 0:
        check(ntile, "hpsi_ntiles")
    part = ctx.work(("hpsi_part", prec, S), (S, ntile, 2), torch.float64) if want_partial else None
    nt = ctypes.c_int()
    check(lib().hfb_hpsi(ctx.handle, prec, ptr(psi), S, int(npsi0), ptr(upot), ptr(bmass), ptr(aq), ptr(spot),
                         ptr(wlspot), ptr(v_pair), ptr(w), ptr(wuv), ptr(pout), ptr(pmf), ptr(pdel), ptr(part),
                         ctypes.byref(nt), stream_handle()), "hpsi")
    sums = None
    if want_partial:
        sums = torch.empty((S, 2), dtype=torch.float64, device=psi.device)
        check(lib().hfb_reduce_partials(ptr(part), S, nt.value, 2, 1.0, ptr(sums), stream_handle()),
              "reduce_partials")
    return pout, pmf, pdel, sums


def poisson(ctx, rho: torch.Tensor, e2: float, wcoul: torch.Tensor | None = None):
    """Coulomb potential of rho[1] on the doubled grid (coulomb.py:54-75)."""
    need(rho, torch.float64, (2, ctx.nx, ctx.ny, ctx.nz), "rho")
    if wcoul is None:
        wcoul = torch.empty((ctx.nx, ctx.ny, ctx.nz), dtype=torch.float64, device=rho.device)
    work = ctx.work("poisson", tuple(ctx.n2), torch.complex128)
    check(lib().hfb_poisson(ctx.handle, ptr(rho), ptr(wcoul), float(e2), ptr(work), stream_handle()), "poisson")
    return wcoul


def selftest_mfma_f64(a: torch.Tensor, b: torch.Tensor):
    need(a, torch.float64, (16, 4), "a")
    need(b, torch.float64, (4, 16), "b")
    d = torch.empty(512, dtype=torch.float64, device=a.device)
    check(lib().hfb_selftest_mfma_f64(ptr(a), ptr(b), ptr(d), stream_handle()), "selftest_mfma")
    return d[:256].reshape(16, 16), d[256:].reshape(64, 4)
