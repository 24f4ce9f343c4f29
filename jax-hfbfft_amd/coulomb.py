"""Coulomb field by FFT Green's function on the doubled grid (coulomb.py:5-75).

The transformed kernel q-hat lives in the library context (computed on the
device at context creation); ``Coulomb`` holds the grid sizes and wcoul."""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import kernels


@dataclass
class Coulomb:
    nx2: int
    ny2: int
    nz2: int
    wcoul: torch.Tensor


def init_coulomb(grids, device=None) -> Coulomb:
    """coulomb.py:23-52; the kernel transform is done once per grid context."""
    dev = torch.device("cuda") if device is None else torch.device(device)
    grids.ctx  # builds the context (and the transformed kernel) now
    n2 = (grids.nx, grids.ny, grids.nz) if grids.periodic else (2 * grids.nx, 2 * grids.ny, 2 * grids.nz)
    return Coulomb(*n2, torch.zeros((grids.nx, grids.ny, grids.nz), dtype=torch.float64, device=dev))


def poisson(grids, params, coulomb, rho, out=None):
    """Coulomb potential of the proton density rho[1] (coulomb.py:54-75)."""
    return kernels.poisson(grids.ctx, rho.contiguous(), params.e2, out)
