"""Nuclear observables and the Sky3D-style result files (output.py of the
reference).

observables() / calculate_complete_observables() reduce the densities on the
device in FP64 (particle numbers, centres of mass, rms radii, second moments,
quadrupole tensor, isovector dipole, momenta from the current) and finish on
the host (3x3 eigenvalues, deformation, angular momenta from the
single-particle properties), output.py:177-293.  FortranOutputWriter writes
conver.res, energies.res, dipoles.res, spin.res, monopoles.res,
quadrupoles.res and momenta.res with the reference's headers and column
formats (output.py:49-174)."""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np
import torch


def _moments(densities, grids):
    """Device reductions: pnr (2), cm (3,2), x2m (3,2), qmat (3,3,2) about
    each isospin's c.m., rms (2), isovector dipole (3), pcm (3,2)."""
    rho = densities.rho
    dev = rho.device
    w = grids.wxyz
    X = torch.as_tensor(grids.x, device=dev)[:, None, None]
    Y = torch.as_tensor(grids.y, device=dev)[None, :, None]
    Z = torch.as_tensor(grids.z, device=dev)[None, None, :]
    xyz = (X, Y, Z)
    pnr = torch.stack([rho[iq].sum() * w for iq in range(2)])
    cm = torch.zeros((3, 2), dtype=torch.float64, device=dev)
    for iq in range(2):
        for i, c in enumerate(xyz):
            cm[i, iq] = (c * rho[iq]).sum() * w / pnr[iq]
    rms = torch.zeros(2, dtype=torch.float64, device=dev)
    x2m = torch.zeros((3, 2), dtype=torch.float64, device=dev)
    qmat = torch.zeros((3, 3, 2), dtype=torch.float64, device=dev)
    for iq in range(2):
        co = (X - cm[0, iq], Y - cm[1, iq], Z - cm[2, iq])
        r2 = co[0] ** 2 + co[1] ** 2 + co[2] ** 2
        rms[iq] = torch.sqrt((r2 * rho[iq]).sum() * w / pnr[iq])
        for i in range(3):
            x2m[i, iq] = (co[i] ** 2 * rho[iq]).sum() * w / pnr[iq]
            for j in range(3):
                q = (3 * co[i] ** 2 - r2) if i == j else 3 * co[i] * co[j]
                qmat[i, j, iq] = (q * rho[iq]).sum() * w
    dip = torch.zeros(3, dtype=torch.float64, device=dev)
    dd = rho[0] / pnr[0] - rho[1] / pnr[1]
    for i, c in enumerate(xyz):
        dip[i] = (c * dd).sum() * w
    pcm = torch.zeros((3, 2), dtype=torch.float64, device=dev)
    cur = getattr(densities, "current", None)
    if cur is not None:
        pcm = cur.sum(dim=(2, 3, 4)).transpose(0, 1) * w
    return [t.cpu().numpy() for t in (pnr, cm, rms, x2m, qmat, dip, pcm)]


def _quadrupole(qmat, q):
    ev = np.linalg.eigvalsh(q)
    return math.sqrt(5.0 / (16.0 * math.pi)) * ev[2], math.sqrt(5.0 / (96.0 * math.pi)) * (ev[1] - ev[0])


def observables(densities, grids, nneut, nprot):
    """N, c.m., rms radii, Q20/Q22 and deformation (output.py:177-293)."""
    pnr, cm, rms, x2m, qmat, _, _ = _moments(densities, grids)
    pnrtot = pnr[0] + pnr[1]
    rmstot = np.sqrt((pnr[0] * rms[0] ** 2 + pnr[1] * rms[1] ** 2) / pnrtot)
    q20 = np.zeros(2)
    q22 = np.zeros(2)
    for iq in range(2):
        q20[iq], q22[iq] = _quadrupole(qmat, qmat[:, :, iq])
    q20tot, q22tot = _quadrupole(qmat, qmat[:, :, 0] + qmat[:, :, 1])
    beta20 = q20tot * (4.0 * np.pi / (5.0 * rmstot ** 2 * pnrtot))
    beta22 = q22tot * (4.0 * np.pi / (5.0 * rmstot ** 2 * pnrtot))
    return dict(pnr=pnr, cm=cm, rms=rms, rmstot=rmstot, q20=q20, q22=q22, q20tot=q20tot, q22tot=q22tot,
                beta20tot=beta20, beta22tot=beta22, beta=np.sqrt(beta20 ** 2 + 2.0 * beta22 ** 2))


@dataclass
class CompleteNuclearObservables:
    """output.py:8-46."""
    pnr: np.ndarray = field(default_factory=lambda: np.zeros(2))
    pnrtot: float = 0.0
    cm: np.ndarray = field(default_factory=lambda: np.zeros((3, 2)))
    cmtot: np.ndarray = field(default_factory=lambda: np.zeros(3))
    dipole_isovector: np.ndarray = field(default_factory=lambda: np.zeros(3))
    rms: np.ndarray = field(default_factory=lambda: np.zeros(2))
    rmstot: float = 0.0
    x2m: np.ndarray = field(default_factory=lambda: np.zeros((3, 2)))
    x2mtot: np.ndarray = field(default_factory=lambda: np.zeros(3))
    q20: np.ndarray = field(default_factory=lambda: np.zeros(2))
    q22: np.ndarray = field(default_factory=lambda: np.zeros(2))
    q20tot: float = 0.0
    q22tot: float = 0.0
    q20T1: float = 0.0
    beta20tot: float = 0.0
    beta22tot: float = 0.0
    beta: float = 0.0
    gamma: float = 0.0
    orbital_momentum: np.ndarray = field(default_factory=lambda: np.zeros(3))
    spin_momentum: np.ndarray = field(default_factory=lambda: np.zeros(3))
    total_momentum: np.ndarray = field(default_factory=lambda: np.zeros(3))
    pcm: np.ndarray = field(default_factory=lambda: np.zeros((3, 2)))


def calculate_angular_momentum(levels, grids=None):
    """Sum over states with wocc > 1e-10 of wocc wstates <L>, <S>
    (output.py:300-327)."""
    wocc = levels.wocc.cpu().numpy() if torch.is_tensor(levels.wocc) else np.asarray(levels.wocc)
    ws = levels.wstates.cpu().numpy() if torch.is_tensor(levels.wstates) else np.asarray(levels.wstates)
    orb = levels.sp_orbital.cpu().numpy() if torch.is_tensor(levels.sp_orbital) else np.asarray(levels.sp_orbital)
    spn = levels.sp_spin.cpu().numpy() if torch.is_tensor(levels.sp_spin) else np.asarray(levels.sp_spin)
    sel = wocc > 1e-10
    wt = (wocc * ws)[sel]
    return (wt[:, None] * orb[sel]).sum(axis=0), (wt[:, None] * spn[sel]).sum(axis=0)


def calculate_complete_observables(densities, levels, grids, params, current_densities=None):
    """Every observable of output.py:177-293 (current_densities: give
    densities.current for the momenta, as the reference's dynamic mode does)."""
    pnr, cm, rms, x2m, qmat, dip, pcm = _moments(densities, grids)
    if current_densities is None:
        pcm = np.zeros((3, 2))
    pnrtot = pnr[0] + pnr[1]
    cmtot = (pnr[0] * cm[:, 0] + pnr[1] * cm[:, 1]) / pnrtot
    rmstot = math.sqrt((pnr[0] * rms[0] ** 2 + pnr[1] * rms[1] ** 2) / pnrtot)
    x2mtot = (pnr[0] * x2m[:, 0] + pnr[1] * x2m[:, 1]) / pnrtot
    q20 = np.zeros(2)
    q22 = np.zeros(2)
    for iq in range(2):
        q20[iq], q22[iq] = _quadrupole(qmat, qmat[:, :, iq])
    q20tot, q22tot = _quadrupole(qmat, qmat[:, :, 0] + qmat[:, :, 1])
    q20T1 = -q20[0] / levels.nneut + q20[1] / levels.nprot
    beta20 = q20tot * (4.0 * math.pi / (5.0 * rmstot ** 2 * pnrtot))
    beta22 = q22tot * (4.0 * math.pi / (5.0 * rmstot ** 2 * pnrtot))
    beta = math.sqrt(beta20 ** 2 + 2.0 * beta22 ** 2)
    gamma = abs(math.atan2(math.sqrt(2.0) * beta22, beta20)) * 180.0 / math.pi
    if gamma > 120.0:
        gamma -= 120.0
    if 60.0 < gamma <= 120.0:
        gamma = 120.0 - gamma
    orb, spn = calculate_angular_momentum(levels, grids)
    return CompleteNuclearObservables(pnr=pnr, pnrtot=pnrtot, cm=cm, cmtot=cmtot, dipole_isovector=dip, rms=rms,
                                      rmstot=rmstot, x2m=x2m, x2mtot=x2mtot, q20=q20, q22=q22, q20tot=q20tot,
                                      q22tot=q22tot, q20T1=q20T1, beta20tot=beta20, beta22tot=beta22, beta=beta,
                                      gamma=gamma, orbital_momentum=orb, spin_momentum=spn,
                                      total_momentum=orb + spn, pcm=pcm)


class FortranOutputWriter:
    """The seven Sky3D result files of output.py:49-174, same names, headers
    and column formats.  Energies may be the package's Energies (device
    vector) or any object with the named attributes."""

    NAMES = ("conver", "energies", "dipoles", "spin", "monopoles", "quadrupoles", "momenta")
    HEADERS = {
        "conver": "# Iter   Energy  d_Energy    sp_fluct    max(lam-)     rms(lam-)      rms    "
                  "beta2  gamma      x_0     e_pair(1)    e_pair(2)\n",
        "energies": "# Iter    N(n)    N(p)       E(sum)         E(integ)       Ekin         "
                    "E_Coul         ehfCrho0       ehfCrho1       ehfCdrho0      ehfCdrho1     "
                    "ehfCtau0       ehfCtau1       ehfCdJ0        ehfCdJ1       e_pair(1)    "
                    "e_pair(2)      e_zpe\n",
        "dipoles": "# Iter    c.m. x-y-z                                  Isovector dipoles x-y-z\n",
        "spin": "# Iter      Lx        Ly        Lz        Sx        Sy        Sz        Jx        Jy        Jz\n",
        "monopoles": "# Iter    RMS_n     RMS_p     RMS_tot   RMS_diff   N_dens    Z_dens    A_dens\n",
        "quadrupoles": "# Iter    Q20_n     Q20_p     Q20_tot   Q22_tot   <x²>_tot  <y²>_tot  "
                       "<z²>_tot  Beta      Gamma\n",
        "momenta": "# Iter    Px_n      Py_n      Pz_n      Px_p      Py_p      Pz_p      Px_tot    Py_tot    "
                   "Pz_tot\n",
    }

    def __init__(self, output_dir: str):
        self.output_dir = output_dir
        os.makedirs(output_dir, exist_ok=True)
        self.files = {k: os.path.join(output_dir, f"{k}.res") for k in self.NAMES}
        for k in self.NAMES:
            with open(self.files[k], "w", encoding="utf-8") as f:
                f.write(self.HEADERS[k])

    def write_iteration_data(self, iteration, energies, observables, static, pairs):
        e = energies.host() if hasattr(energies, "host") else energies
        get = (lambda k: e[k]) if isinstance(e, dict) else (lambda k: getattr(e, k))
        efl1 = float(np.ravel(get("efluct1"))[0])
        efl2 = float(np.ravel(get("efluct2"))[0])
        ehfprev = getattr(energies, "ehfprev", 0.0)
        dE = get("ehf") - ehfprev if (ehfprev != 0.0 and iteration > 0) else 0.0
        ep = pairs.epair.cpu().numpy() if torch.is_tensor(pairs.epair) else np.asarray(pairs.epair)
        o = observables
        with open(self.files["conver"], "a", encoding="utf-8") as f:
            f.write(f"{iteration:4d} {get('ehf'):12.6f} {dE:12.6f} {efl1:12.6e} {efl1:12.6e} {efl2:12.6e} "
                    f"{o.rmstot:8.4f} {o.beta:8.4f} {o.gamma:8.2f} {static.x0dmp:8.4f} {ep[0]:12.6f} "
                    f"{ep[1]:12.6f}\n")
        with open(self.files["energies"], "a", encoding="utf-8") as f:
            f.write(f"{iteration:4d} {o.pnr[0]:8.2f} {o.pnr[1]:8.2f} {get('ehf'):14.6f} {get('ehfint'):14.6f} "
                    f"{get('tke'):12.6f} {get('ehfc'):12.6f} {get('ehfCrho0'):12.6f} {get('ehfCrho1'):12.6f} "
                    f"{get('ehfCdrho0'):12.6f} {get('ehfCdrho1'):12.6f} {get('ehfCtau0'):12.6f} "
                    f"{get('ehfCtau1'):12.6f} {get('ehfCdJ0'):12.6f} {get('ehfCdJ1'):12.6f} {ep[0]:12.6f} "
                    f"{ep[1]:12.6f} {get('e_zpe'):12.6f}\n")
        with open(self.files["dipoles"], "a", encoding="utf-8") as f:
            f.write(f"{iteration:4d} {o.cmtot[0]:12.6f} {o.cmtot[1]:12.6f} {o.cmtot[2]:12.6f} "
                    f"{o.dipole_isovector[0]:12.6f} {o.dipole_isovector[1]:12.6f} {o.dipole_isovector[2]:12.6f}\n")
        with open(self.files["spin"], "a", encoding="utf-8") as f:
            L, S, J = o.orbital_momentum, o.spin_momentum, o.total_momentum
            f.write(f"{iteration:4d} {L[0]:10.6f} {L[1]:10.6f} {L[2]:10.6f} {S[0]:10.6f} {S[1]:10.6f} "
                    f"{S[2]:10.6f} {J[0]:10.6f} {J[1]:10.6f} {J[2]:10.6f}\n")
        with open(self.files["monopoles"], "a", encoding="utf-8") as f:
            f.write(f"{iteration:4d} {o.rms[0]:9.6f} {o.rms[1]:9.6f} {o.rmstot:9.6f} {o.rms[0] - o.rms[1]:9.6f} "
                    f"{o.pnr[0]:9.3f} {o.pnr[1]:9.3f} {o.pnrtot:9.3f}\n")
        with open(self.files["quadrupoles"], "a", encoding="utf-8") as f:
            f.write(f"{iteration:4d} {o.q20[0]:9.6f} {o.q20[1]:9.6f} {o.q20tot:9.6f} {o.q22tot:9.6f} "
                    f"{o.x2mtot[0]:9.6f} {o.x2mtot[1]:9.6f} {o.x2mtot[2]:9.6f} {o.beta:9.6f} {o.gamma:9.2f}\n")
        pt = o.pcm[:, 0] + o.pcm[:, 1]
        with open(self.files["momenta"], "a", encoding="utf-8") as f:
            f.write(f"{iteration:4d} {o.pcm[0, 0]:10.6f} {o.pcm[1, 0]:10.6f} {o.pcm[2, 0]:10.6f} "
                    f"{o.pcm[0, 1]:10.6f} {o.pcm[1, 1]:10.6f} {o.pcm[2, 1]:10.6f} {pt[0]:10.6f} {pt[1]:10.6f} "
                    f"{pt[2]:10.6f}\n")


def complete_sinfo_with_fortran_output(coulomb, densities, energies, forces, grids, levels, meanfield, moment,
                                       params, static, pairs, output_writer):
    """Observables of the current iteration into the result files
    (output.py:331-355); returns (energies, observables)."""
    ob = calculate_complete_observables(densities, levels, grids, params)
    if moment is not None:
        # in place: launches recorded earlier keep reading the same buffer
        moment.cmtot.copy_(torch.as_tensor(ob.cmtot, dtype=torch.float64))
    output_writer.write_iteration_data(params.iteration, energies, ob, static, pairs)
    return energies, ob
