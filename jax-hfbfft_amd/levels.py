"""Single-particle states and the spectral derivative entry points
(levels.py:6-437).

``Levels`` holds the spinor block ``psi`` (S, 2, nx, ny, nz) and its
companions on the HIP device; their dtype (complex128 or complex64) selects
the FP64 or FP32 kernels.  Per-state vectors are float64 device tensors."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, kernels


@dataclass
class Levels:
    nstmax: int
    nneut: int
    nprot: int
    npmin: np.ndarray
    npsi: np.ndarray
    charge_number: float
    mass_number: float
    psi: torch.Tensor
    hampsi: torch.Tensor
    lagrange: torch.Tensor
    hmfpsi: torch.Tensor
    delpsi: torch.Tensor
    sp_orbital: torch.Tensor
    sp_spin: torch.Tensor
    isospin: torch.Tensor
    sp_energy: torch.Tensor
    sp_efluct1: torch.Tensor
    sp_kinetic: torch.Tensor
    sp_norm: torch.Tensor
    sp_efluct2: torch.Tensor
    sp_parity: torch.Tensor
    wocc: torch.Tensor
    wguv: torch.Tensor
    pairwg: torch.Tensor
    wstates: torch.Tensor
    deltaf: torch.Tensor

    def block(self, iq):
        n0 = int(self.npsi[0])
        return slice(0, n0) if iq == 0 else slice(n0, n0 + int(self.npsi[1]))


def init_levels(grids, device=None, dtype=torch.complex128, **kwargs) -> Levels:
    """levels.py:37-130 (basis sizes, isospin map, occupations)."""
    nneut = kwargs.get("nneut", 82)
    nprot = kwargs.get("nprot", 50)
    if kwargs.get("npsi") is not None:
        npsi = np.array(kwargs["npsi"], dtype=np.int64)
    else:
        npsi = np.array([max(126, nneut + int(1.65 * nneut ** 0.666667)),
                         max(82, nprot + int(1.65 * nprot ** 0.666667))], dtype=np.int64)
    S = int(npsi[0] + npsi[1])
    dev = torch.device("cuda") if device is None else torch.device(device)
    s5 = (S, 2, grids.nx, grids.ny, grids.nz)
    z5 = lambda: torch.zeros(s5, dtype=dtype, device=dev)
    zf = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
    isospin = torch.zeros(S, dtype=torch.int32, device=dev)
    isospin[int(npsi[0]):] = 1
    wocc = zf(S)
    wocc[:nneut] = 1.0
    wocc[int(npsi[0]):int(npsi[0]) + nprot] = 1.0
    return Levels(S, nneut, nprot, np.array([0, int(npsi[0])]), npsi, nprot, nneut + nprot,
                  z5(), z5(), z5(), z5(), z5(), zf(S, 3), zf(S, 3), isospin, zf(S), zf(S), zf(S), zf(S), zf(S),
                  zf(S), wocc, zf(S), torch.ones(S, dtype=torch.float64, device=dev),
                  torch.ones(S, dtype=torch.float64, device=dev), zf(S))


# ---------------------------------------------------------------------------
# cderv{x,y,z}{00,01,02}: same names and argument order as levels.py:131-395.
# psin: (..., nx, ny, nz) complex on the device.
# ---------------------------------------------------------------------------
def _ctx_for(psin, d):
    nx, ny, nz = psin.shape[-3:]
    key = (int(nx), int(ny), int(nz), float(d))
    c = _CTX_D.get(key)
    if c is None:
        c = _lib.Context(nx, ny, nz, d, d, d)
        _CTX_D[key] = c
    return c


_CTX_D = {}


def _d00(axis):
    def f(d, psin):
        return kernels.cderv(_ctx_for(psin, d), axis, 0, psin.contiguous())
    return f


def _d01(axis):
    def f(d, psin):
        return kernels.cderv(_ctx_for(psin, d), axis, 1, psin.contiguous())
    return f


def _d02(axis):
    def f(d, psin, pos_func):
        rt = torch.float64 if psin.dtype == torch.complex128 else torch.float32
        return kernels.cderv(_ctx_for(psin, d), axis, 2, psin.contiguous(), pos_func.to(rt).contiguous())
    return f


cdervx00, cdervy00, cdervz00 = _d00(0), _d00(1), _d00(2)
cdervx01, cdervy01, cdervz01 = _d01(0), _d01(1), _d01(2)
cdervx02, cdervy02, cdervz02 = _d02(0), _d02(1), _d02(2)


def laplace(nst, iq, grids, forces, levels, wg, wguv, psin, v_pairmax, e0inv):
    """FFT-diagonal preconditioner applied to psin (levels.py:397-437).

    Runs the step kernel with psi = 0, residual = psin and x0dmp = -1, which
    returns exactly IFFT[FFT(psin) / ((max(wg,.1)(E0 + h2ma k^2) + .5 max(wguv,.1) vmax) N)]."""
    ctx = grids.ctx
    x = psin.reshape(-1, 2, grids.nx, grids.ny, grids.nz).contiguous()
    ns = x.shape[0]
    out = torch.zeros_like(x)
    zero = torch.zeros_like(x)
    dev = x.device
    pw = torch.full((ns,), max(float(wg), 0.1), dtype=torch.float64, device=dev)
    ps = torch.full((ns,), 0.5 * max(float(wguv), 0.1) * float(v_pairmax), dtype=torch.float64, device=dev)
    hsum = torch.zeros((ns, 2), dtype=torch.float64, device=dev)
    kernels.precond_step(ctx, out, x, zero, hsum, grids.wxyz, True, pw, ps, float(e0inv), forces.h2ma, -1.0)
    return out.reshape(psin.shape)
