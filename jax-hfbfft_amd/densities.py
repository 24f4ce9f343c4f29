"""Densities (densities.py:6-337): rho, chi, tau, current, spin and
spin-orbit densities per isospin, accumulated on the device without
materialising per-state fields."""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import kernels


@dataclass
class Densities:
    rho: torch.Tensor
    chi: torch.Tensor
    tau: torch.Tensor
    current: torch.Tensor
    sdens: torch.Tensor
    sodens: torch.Tensor


def init_densities(grids, device=None) -> Densities:
    dev = torch.device("cuda") if device is None else torch.device(device)
    s4 = (2, grids.nx, grids.ny, grids.nz)
    s5 = (2, 3, grids.nx, grids.ny, grids.nz)
    z = lambda s: torch.zeros(s, dtype=torch.float64, device=dev)
    return Densities(z(s4), z(s4), z(s4), z(s5), z(s5), z(s5))


def add_density(densities, grids, levels):
    """Sum of per-state contributions per isospin (densities.py:177-234);
    returns the updated densities like the reference."""
    return add_density_sweep(densities, grids, levels)[0]


def add_density_sweep(densities, grids, levels, mix=False, addnew=0.2, moment=None, want_sp=False, sp_h2m=None):
    """add_density in one sweep over the states, optionally fused with the
    loop's linear mixing of rho, tau, chi (``mix``, static.py:1817-1823) and
    with the per-state kinetic energy, <L>, <S> and parity of sp_properties
    (``want_sp``, inout.py:64-149) from the same transforms.  With
    ``sp_h2m`` (forces.h2m) the single-particle properties are also
    finished into levels (what sp_properties(..., spart) does).  Returns
    (densities, per-state partial sums or None)."""
    ctx = grids.ctx
    n0, n1 = int(levels.npsi[0]), int(levels.npsi[1])
    cm = moment.cmtot.to(device=levels.psi.device, dtype=torch.float64) if moment is not None else \
        torch.zeros(3, dtype=torch.float64, device=levels.psi.device)
    dout, spart = kernels.density(ctx, levels.psi, n0, n1, levels.wocc, levels.wstates, levels.wguv, levels.pairwg,
                                  cm, True, want_sp)
    if want_sp and sp_h2m is not None:
        kernels.density_sp_finish(ctx, dout, densities, mix, addnew, spart, n0, n1, sp_h2m, levels)
    else:
        kernels.density_finish(ctx, dout, densities, mix, addnew, psi_dtype=levels.psi.dtype)
    return densities, spart
