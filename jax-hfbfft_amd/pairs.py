"""BCS pairing (pairs.py:13-220), solved on the device.

The Fermi level is found by a bracketed parallel bisection on the GPU
(the reference runs scipy brentq on the host with a device sync per
evaluation); the result is within 1e-15 of the root, inside brentq's
xtol of 1e-14."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, kernels
from ._lib import check, lib, ptr, stream_handle


def _z2():
    return torch.zeros(2, dtype=torch.float64, device="cuda")


@dataclass
class Pairs:
    eferm: torch.Tensor = field(default_factory=_z2)
    epair: torch.Tensor = field(default_factory=_z2)
    avdelt: torch.Tensor = field(default_factory=_z2)
    avdeltv2: torch.Tensor = field(default_factory=_z2)
    avg: torch.Tensor = field(default_factory=_z2)
    firstcall: bool = True
    status: torch.Tensor = field(default_factory=lambda: torch.zeros(1, dtype=torch.int32, device="cuda"))


def pairgap_mean_field(levels, meanfield, grids):
    """deltaf = <v_pair> * pairwg (pairs.py:148-157): the branch taken only
    when the first pair() call comes after iteration 10."""
    dens = (levels.psi.conj() * levels.psi).real.sum(dim=1).to(torch.float64)
    vp = meanfield.v_pair[levels.isospin.long()]
    return (vp * dens).sum(dim=(1, 2, 3)) * grids.wxyz * levels.pairwg


def pair(levels, meanfield, forces, params, grids, pairs_data, E=None, check_status=True):
    """pairs.py:161-220 (the summary print is left to the caller)."""
    const_gap = -1.0
    if params.iteration <= 10:
        const_gap = 11.2 / np.sqrt(levels.mass_number)
    elif pairs_data.firstcall:
        levels.deltaf.copy_(pairgap_mean_field(levels, meanfield, grids))
    pairs_data.firstcall = False
    pc = (ctypes.c_double * 2)(*[float(v) for v in forces.pair_cutoff])
    sc = (ctypes.c_double * 2)(*[float(v) for v in forces.state_cutoff])
    check(lib().hfb_bcs(int(levels.npsi[0]), int(levels.npsi[1]), float(levels.nneut), float(levels.nprot), pc, sc,
                        float(forces.softcut_range), float(const_gap), ptr(levels.sp_energy), ptr(levels.deltaf),
                        ptr(levels.wocc), ptr(levels.wguv), ptr(levels.pairwg), ptr(levels.wstates),
                        ptr(pairs_data.eferm), ptr(pairs_data.epair), ptr(pairs_data.avdelt), ptr(pairs_data.avdeltv2),
                        ptr(pairs_data.avg), ptr(E), ptr(pairs_data.status), stream_handle()), "bcs")
    if check_status:
        raise_on_status(pairs_data)
    return levels, pairs_data


def raise_on_status(pairs_data):
    if int(pairs_data.status.item()) != 0:
        pairs_data.status.zero_()
        raise ValueError("f(a) and f(b) must have different signs")
