"""Run driver: builds a static HF(+BCS) calculation from a configuration and
runs the reference's loop on the device.

Precision modes
  "f64":   complex128 states, every kernel in FP64.
  "mixed": complex64 states; FFT derivatives, h psi, the step and the
           density transforms in FP32 (FP64 partial sums); Gram matrices,
           rotations, Jacobi eigensolves, potentials, energies and BCS in
           FP64.  ``promote()`` switches a run to complex128 for the final
           FP64 refinement iterations.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import static as st_mod
from .coulomb import init_coulomb
from .densities import init_densities
from .energies import init_energies
from .forces import init_forces
from .grids import init_grids
from .levels import init_levels
from .meanfield import init_meanfield
from .moment import init_moment
from .output import observables
from .pairs import Pairs
from .params import init_params


def add_noise(psi, amplitude=1e-3, seed=0):
    """psi_i += a * max|psi_i| * (xi + i eta)/sqrt2 with xi then eta drawn as
    whole arrays from numpy PCG64(seed) (build-defined; the reference's
    harmosc has no noise)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xi = rng.standard_normal(psi.shape)
    eta = rng.standard_normal(psi.shape)
    amax = np.max(np.abs(psi).reshape(psi.shape[0], -1), axis=1)[:, None, None, None, None]
    return psi + amplitude * amax * (xi + 1j * eta) / np.sqrt(2.0)


def deformed_radii(b, beta2):
    """Gaussian widths radin = sqrt2 b, stretched along z for an axial
    quadrupole bias beta2 to first order (build-defined)."""
    r = math.sqrt(2.0) * b
    c = math.sqrt(5.0 / (4.0 * math.pi)) * beta2
    return r * (1.0 - 0.5 * c), r * (1.0 - 0.5 * c), r * (1.0 + c)


# Workloads named by the benchmark configurations (nucleus, grid, states).
CONFIGS = {
    "O16": dict(name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, b=1.8),
    "Ca40": dict(name="SV-bas", n=32, d=1.0, nneut=20, nprot=20, b=2.0),
    "Cr48": dict(name="SV-bas", n=32, d=1.0, nneut=24, nprot=24, npsi=[36, 36], b=2.0, beta2=0.2, ipair=5,
                 tbcs=True),
    "Pb208": dict(name="SV-bas", n=48, d=0.8, nneut=126, nprot=82, b=2.5),
}


class HFRun:
    def __init__(self, params, forces, grids, levels, meanfield, densities, coulomb, static, energies, pairs, moment):
        self.params, self.forces, self.grids, self.levels = params, forces, grids, levels
        self.meanfield, self.densities, self.coulomb, self.static = meanfield, densities, coulomb, static
        self.energies, self.pairs, self.moment = energies, pairs, moment
        self.derivs = None
        self.tbcssav = forces.tbcs
        self.iteration_no = 0

    @classmethod
    def from_config(cls, precision="f64", device="cuda:0", name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, npsi=None,
                    b=1.8, beta2=0.0, ipair=0, tbcs=False, v0=None, noise=1e-3, seed=0, mprint=10, psi0=None,
                    **static_kw):
        if precision not in ("f64", "mixed"):
            raise ValueError(f"precision must be 'f64' or 'mixed', not {precision!r}")
        dev = torch.device(device)
        torch.cuda.set_device(dev)
        params = init_params(imode=1, mprint=mprint)
        forces = init_forces(params, name=name, ipair=ipair, tbcs=tbcs, v0prot=v0, v0neut=v0)
        grids = init_grids(params, nx=n, ny=n, nz=n, dx=d, dy=d, dz=d)
        dtype = torch.complex128 if precision == "f64" else torch.complex64
        levels = init_levels(grids, device=dev, dtype=dtype, nneut=nneut, nprot=nprot,
                             npsi=npsi if npsi is not None else [nneut, nprot])
        rx, ry, rz = deformed_radii(b, beta2)
        forces, static = st_mod.init_static(forces, levels, radinx=rx, radiny=ry, radinz=rz, **static_kw)
        coulomb = init_coulomb(grids, dev)
        if psi0 is None:
            psi0 = st_mod.harmosc_host(grids, levels.npsi, nneut, nprot, (rx, ry, rz))
            if noise:
                psi0 = add_noise(psi0, noise, seed)
        levels.psi.copy_(torch.from_numpy(np.ascontiguousarray(psi0)).to(dtype))
        meanfield = init_meanfield(grids, dev, f32=(precision == "mixed"))
        return cls(params, forces, grids, levels, meanfield, init_densities(grids, dev), coulomb, static,
                   init_energies(dev), Pairs(), init_moment([0.0, 0.0, 0.0]))

    @property
    def state(self):
        return (self.coulomb, self.densities, self.energies, self.forces, self.grids, self.levels, self.meanfield,
                self.moment, self.params, self.static, self.pairs)

    def prelude(self):
        self.derivs = st_mod.prelude(*self.state)
        self.tbcssav = self.forces.tbcs

    def iteration(self, i, check=False):
        self.iteration_no = i
        conv, self.derivs = st_mod.iterate(i, *self.state, self.tbcssav, self.derivs, check=check)
        return conv

    def run(self, maxiter, start=1, check_every=10, callback=None):
        for i in range(start, start + maxiter):
            chk = check_every > 0 and (i % check_every == 0)
            conv = self.iteration(i, check=chk)
            if callback is not None:
                callback(i, self)
            if conv:
                return i
        return start + maxiter - 1

    def promote(self):
        """Switch a mixed run to complex128 states and FP64 kernels."""
        lv = self.levels
        if lv.psi.dtype == torch.complex128:
            return
        for name in ("psi", "hampsi", "lagrange", "hmfpsi", "delpsi"):
            setattr(lv, name, getattr(lv, name).to(torch.complex128))

    def energies_host(self):
        torch.cuda.synchronize()
        return self.energies.host()

    def observables(self):
        return observables(self.densities, self.grids, self.levels.nneut, self.levels.nprot)
