"""Energy functional integrals and single-particle sums (energies.py:6-371).

``Energies`` keeps every quantity in one float64 device vector (slot
layout of include/hfb_hip.h, HFB_EN_*); the reference's attribute names
read it (one device-to-host copy per read)."""
from __future__ import annotations

import torch

from . import kernels

EN = kernels.EN
_SCALARS = {"ehft": "EHFT", "ehf0": "EHF0", "ehf1": "EHF1", "ehf2": "EHF2", "ehf3": "EHF3", "ehfls": "EHFLS",
            "ehflsodd": "EHFLSODD", "ehfc": "EHFC", "ecorc": "ECORC", "ehfint": "EHFINT", "e3corr": "E3CORR",
            "e_zpe": "EZPE", "ehfCrho0": "CRHO0", "ehfCrho1": "CRHO1", "ehfCdrho0": "CDRHO0",
            "ehfCdrho1": "CDRHO1", "ehfCtau0": "CTAU0", "ehfCtau1": "CTAU1", "ehfCdJ0": "CDJ0", "ehfCdJ1": "CDJ1",
            "ehfCj0": "CJ0", "ehfCj1": "CJ1", "ehf": "EHF", "tke": "TKE"}


class Energies:
    def __init__(self, device=None):
        dev = torch.device("cuda") if device is None else torch.device(device)
        self.vec = torch.zeros(kernels.EN_LEN, dtype=torch.float64, device=dev)
        self.efluct1prev = 0.0
        self.efluct2prev = 0.0
        self.ehfprev = 0.0
        self.e_extern = 0.0
        # energy tail left by the last SCF iteration (kernels.tail_job); the
        # next iteration's first h psi launch runs it, reads run it here
        self.tail = None
        self.stale = False
        self._tail_ctx = None

    def defer_tail(self, ctx, tail):
        self._tail_ctx, self.tail, self.stale = ctx, tail, True

    def flush(self):
        """Run the deferred energy tail, if any (rerunning it is harmless: it
        rewrites the same slots from the same partials)."""
        if self.stale:
            kernels.run_tail(self._tail_ctx, self.tail)
            self.stale = False

    def __getattr__(self, name):
        if name in _SCALARS:
            self.flush()
            return float(self.vec[EN[_SCALARS[name]]].item())
        raise AttributeError(name)

    @property
    def efluct1(self):
        return self.vec[EN["EFL1"]:EN["EFL1"] + 1]

    @property
    def efluct2(self):
        return self.vec[EN["EFL2"]:EN["EFL2"] + 1]

    @property
    def efluct1q(self):
        return self.vec[EN["EFL1Q"]:EN["EFL1Q"] + 2]

    @property
    def efluct2q(self):
        return self.vec[EN["EFL2Q"]:EN["EFL2Q"] + 2]

    @property
    def orbital(self):
        self.flush()
        return self.vec[EN["ORB"]:EN["ORB"] + 3]

    @property
    def spin(self):
        self.flush()
        return self.vec[EN["SPIN"]:EN["SPIN"] + 3]

    @property
    def total_angmom(self):
        self.flush()
        return self.vec[EN["JTOT"]:EN["JTOT"] + 3]

    def host(self):
        """All slots as a host dict (one synchronising copy)."""
        self.flush()
        v = self.vec.cpu().numpy()
        d = {k: float(v[EN[s]]) for k, s in _SCALARS.items()}
        d.update(efluct1=float(v[EN["EFL1"]]), efluct2=float(v[EN["EFL2"]]), ecorrp=float(v[EN["ECORRP"]]),
                 sumflu=float(v[EN["SUMFLU"]]), delesum=float(v[EN["DELESUM"]]),
                 epair=v[EN["EPAIR"]:EN["EPAIR"] + 2].copy(),
                 jacobi_sweeps=v[EN["JACOBI"]:EN["JACOBI"] + 2].copy(),
                 ns_iterations=v[EN["NSITER"]:EN["NSITER"] + 2].copy())
        return d


def init_energies(device=None):
    return Energies(device)


def _set_slots(E, lo, src):
    """E[lo:lo+len(src)] = src, skipped when src already is that slice (the
    driver keeps pairs.epair and meanfield.ecorrp_dev as views of E)."""
    dst = E[lo:lo + src.numel()]
    if src.data_ptr() != dst.data_ptr():
        dst.copy_(src.reshape(dst.shape))


def integ_energy(coulomb, densities, energies, forces, grids, levels, params, pairs, derivs=None, finish=True):
    """energies.py:90-306.  ``derivs``: DensityDerivs of these densities, if
    already computed for skyrme.  Pairing energies come from pairs.epair.
    ``finish=False`` leaves the integrals to a later kernels.energy_tail."""
    from .meanfield import density_derivs
    if derivs is None:
        derivs = density_derivs(grids, densities)
    E = energies.vec
    _set_slots(E, EN["EPAIR"], pairs.epair)
    fc = forces.to_c(params, levels.mass_number)
    kernels.integ_energy(grids.ctx, fc, densities, derivs, coulomb.wcoul, E, finish=finish)
    return energies


def sum_energy(energies, levels, meanfield, pairs):
    """energies.py:309-371."""
    E = energies.vec
    _set_slots(E, EN["EPAIR"], pairs.epair)
    _set_slots(E, EN["ECORRP"], meanfield.ecorrp_dev)
    kernels.sum_energy(levels, E)
    return energies
