"""Single-particle properties (inout.py:64-149): kinetic energy, <L>, <S>,
parity of every state, from one transform per axis."""
from __future__ import annotations

import torch

from . import kernels


def sp_properties(forces, grids, levels, moment, spart=None):
    """Fills levels.sp_kinetic/orbital/spin/parity; ``spart`` may hold the
    per-state sums already produced by add_density(..., want_sp=True)."""
    ctx = grids.ctx
    n0, n1 = int(levels.npsi[0]), int(levels.npsi[1])
    if spart is None:
        cm = moment.cmtot.to(device=levels.psi.device, dtype=torch.float64)
        _, spart = kernels.density(ctx, levels.psi, n0, n1, levels.wocc, levels.wstates, levels.wguv, levels.pairwg,
                                   cm, False, True)
    kernels.sp_finish(ctx, spart, n0, n1, forces.h2m, levels)
    return levels
