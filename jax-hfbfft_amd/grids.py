"""Cartesian grid (grids.py:5-154).

Coordinates and the dense derivative matrices live on the host for
inspection; the device copies of the derivative matrices and the FFT tables
are owned by the library context (``grids.ctx``)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib


def sder(nmax: int, d: float) -> np.ndarray:
    """Dense Fourier first-derivative matrix (grids.py:36-46)."""
    icn = (nmax + 1) // 2
    afac = math.pi / icn
    i = np.arange(1, nmax + 1)
    j = np.arange(1, icn)[:, None, None]
    g = i[None, :] - i[:, None]
    return -afac * (np.sum(-j * np.sin(j * afac * g), axis=0) - 0.5 * icn * np.sin(icn * afac * g)) / (icn * d)


def sder2(nmax: int, d: float) -> np.ndarray:
    """Dense Fourier second-derivative matrix (grids.py:51-61)."""
    icn = (nmax + 1) // 2
    afac = math.pi / icn
    i = np.arange(1, nmax + 1)
    j = np.arange(1, icn)[:, None, None]
    g = i[None, :] - i[:, None]
    return -(afac * afac) * (np.sum((j ** 2) * np.cos(j * afac * g), axis=0)
                             + 0.5 * (icn ** 2) * np.cos(icn * afac * g)) / (icn * d * d)


@dataclass
class Grids:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    periodic: bool
    wxyz: float
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    der1x: np.ndarray
    der1y: np.ndarray
    der1z: np.ndarray
    der2x: np.ndarray
    der2y: np.ndarray
    der2z: np.ndarray
    bangx: float = 0.0
    bangy: float = 0.0
    bangz: float = 0.0
    tabc_x: int = 0
    tabc_y: int = 0
    tabc_z: int = 0

    @property
    def ctx(self):
        return _lib.context_for(self)


def init_grids(params, **kwargs) -> Grids:
    """grids.py:66-138 with the same defaults, checks and error messages."""
    kw = dict(nx=48, ny=48, nz=48, dx=0.8, dy=0.8, dz=0.8, bangx=0.0, bangy=0.0, bangz=0.0,
              tabc_x=0, tabc_y=0, tabc_z=0, periodic=False)
    kw.update(kwargs)
    if not params.tfft and (abs(kw["bangx"]) > 1e-5 or abs(kw["bangy"]) > 1e-5 or abs(kw["bangz"]) > 1e-5):
        raise ValueError("bloch boundaries cannot be used without tfft.")
    if params.tabc_nprocs == 1 and (kw["tabc_x"] != 0 or kw["tabc_y"] != 0 or kw["tabc_z"] != 0):
        raise ValueError("no tabc possible with tabc_nprocs=1.")
    if kw["nx"] % 2 or kw["ny"] % 2 or kw["nz"] % 2:
        raise ValueError("grids: nx, ny, and nz must be even.")
    if kw["dx"] * kw["dy"] * kw["dz"] <= 0.0:
        if kw["dx"] <= 0.0:
            raise ValueError("grid spacing given as 0.")
        kw["dy"] = kw["dz"] = kw["dx"]

    def coord(n, d):
        return (np.arange(n, dtype=np.float64) - (n - 1) / 2) * d

    nx, ny, nz, dx, dy, dz = kw["nx"], kw["ny"], kw["nz"], kw["dx"], kw["dy"], kw["dz"]
    return Grids(nx, ny, nz, dx, dy, dz, bool(kw["periodic"]), dx * dy * dz,
                 coord(nx, dx), coord(ny, dy), coord(nz, dz),
                 sder(nx, dx), sder(ny, dy), sder(nz, dz), sder2(nx, dx), sder2(ny, dy), sder2(nz, dz),
                 kw["bangx"] * math.pi, kw["bangy"] * math.pi, kw["bangz"] * math.pi,
                 kw["tabc_x"], kw["tabc_y"], kw["tabc_z"])
