"""Nuclear observables from the densities (output.py:177-293): particle
numbers, centre of mass, rms radii, quadrupole moments and deformation.
Reductions run on the device in FP64; the 3x3 eigenvalues on the host."""
from __future__ import annotations

import numpy as np
import torch


def observables(densities, grids, nneut, nprot):
    rho = densities.rho
    dev = rho.device
    w = grids.wxyz
    X = torch.as_tensor(grids.x, device=dev)[:, None, None]
    Y = torch.as_tensor(grids.y, device=dev)[None, :, None]
    Z = torch.as_tensor(grids.z, device=dev)[None, None, :]
    pnr = torch.stack([rho[iq].sum() * w for iq in range(2)])
    cm = torch.zeros((3, 2), dtype=torch.float64, device=dev)
    for iq in range(2):
        for i, c in enumerate((X, Y, Z)):
            cm[i, iq] = (c * rho[iq]).sum() * w / pnr[iq]
    rms = torch.zeros(2, dtype=torch.float64, device=dev)
    qmat = torch.zeros((3, 3, 2), dtype=torch.float64, device=dev)
    for iq in range(2):
        co = (X - cm[0, iq], Y - cm[1, iq], Z - cm[2, iq])
        r2 = co[0] ** 2 + co[1] ** 2 + co[2] ** 2
        rms[iq] = torch.sqrt((r2 * rho[iq]).sum() * w / pnr[iq])
        for i in range(3):
            for j in range(3):
                q = (3 * co[i] ** 2 - r2) if i == j else 3 * co[i] * co[j]
                qmat[i, j, iq] = (q * rho[iq]).sum() * w
    pnr, cm, rms, qmat = (t.cpu().numpy() for t in (pnr, cm, rms, qmat))
    pnrtot = pnr[0] + pnr[1]
    rmstot = np.sqrt((pnr[0] * rms[0] ** 2 + pnr[1] * rms[1] ** 2) / pnrtot)
    q20 = np.zeros(2)
    q22 = np.zeros(2)
    for iq in range(2):
        ev = np.linalg.eigvalsh(qmat[:, :, iq])
        q20[iq] = np.sqrt(5.0 / (16.0 * np.pi)) * ev[2]
        q22[iq] = np.sqrt(5.0 / (96.0 * np.pi)) * (ev[1] - ev[0])
    evt = np.linalg.eigvalsh(qmat[:, :, 0] + qmat[:, :, 1])
    q20tot = np.sqrt(5.0 / (16.0 * np.pi)) * evt[2]
    q22tot = np.sqrt(5.0 / (96.0 * np.pi)) * (evt[1] - evt[0])
    beta20 = q20tot * (4.0 * np.pi / (5.0 * rmstot ** 2 * pnrtot))
    beta22 = q22tot * (4.0 * np.pi / (5.0 * rmstot ** 2 * pnrtot))
    return dict(pnr=pnr, cm=cm, rms=rms, rmstot=rmstot, q20=q20, q22=q22, q20tot=q20tot, q22tot=q22tot,
                beta20tot=beta20, beta22tot=beta22, beta=np.sqrt(beta20 ** 2 + 2.0 * beta22 ** 2))
