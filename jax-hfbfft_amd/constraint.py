"""Constraints on the quadrupole moments and the principal axes
(constraint.py of the reference).

The reference's module is never called (meanfield.py:246-248 imports a
module named ``constraints``; its potential sum and its isospin index do not
match the arrays).  Here the constraint fields are isoscalar (one field for
both isospins); skyrme subtracts sum_i lambda_i Q_i from U_q; the SCF loop
runs tune_constraints after the subspace step and the pairing (the
multipliers and the exponential correction of the states), so that the
following orthonormalisation and densities see the corrected states.  The
loop uses the fused device kernels hfb_constraint_tune / _correct /
_potential; the separate helpers below keep the reference's interface."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import check, lib, ptr, stream_handle, prec_of

NC_MAX = 8


@dataclass
class Constraints:
    tconstraint: bool = False
    tq_prin_axes: bool = False
    alpha20_wanted: float = -1e99
    alpha22_wanted: float = -1e99
    c0constr: float = 0.8
    d0constr: float = 1e-4
    qepsconstr: float = 0.3
    dampgamma: float = 1.0
    damprad: float = 6.0
    r0rms: float = 0.93
    numconstraint: int = 0
    constr_field: torch.Tensor = None  # (numconstraint, nx, ny, nz) float64, isoscalar
    lambda_crank: torch.Tensor = None
    dlambda_crank: torch.Tensor = None
    goal_crank: torch.Tensor = None
    actual_crank: torch.Tensor = None
    old_crank: torch.Tensor = None
    actual_crank2: torch.Tensor = None
    qcorr: torch.Tensor = None


def init_constraints(params, mass_number, grids, device=None) -> Constraints:
    """Fields and multipliers (constraint.py:49-196): Q20 ~ 2z^2-x^2-y^2
    (goal alpha20_wanted), Q22 ~ x^2-y^2 (alpha22_wanted), with tq_prin_axes
    xy, xz, yz, x, y, z (goal 0); normalised by r0rms^2 A^(5/3), masked by
    1/(1+exp((r-damprad)/dampgamma)); multipliers start at -0.2."""
    c = Constraints()
    for k in ("alpha20_wanted", "alpha22_wanted", "tq_prin_axes", "c0constr", "d0constr", "qepsconstr", "dampgamma",
              "damprad", "r0rms"):
        if k in params and (not k.startswith("alpha") or params[k] > -1e90):
            setattr(c, k, params[k])
    X = np.asarray(grids.x)[:, None, None]
    Y = np.asarray(grids.y)[None, :, None]
    Z = np.asarray(grids.z)[None, None, :]
    mask = 1.0 / (1.0 + np.exp((np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - c.damprad) / c.dampgamma))
    norm = c.r0rms ** 2 * mass_number ** (5.0 / 3.0)
    fields, goals = [], []
    if c.alpha20_wanted > -1e90:
        fields.append(math.sqrt(math.pi / 5.0) / norm * (2.0 * Z ** 2 - X ** 2 - Y ** 2) * mask)
        goals.append(c.alpha20_wanted)
    if c.alpha22_wanted > -1e90:
        fields.append(math.sqrt(1.2 * math.pi) / norm * (X ** 2 - Y ** 2) * mask)
        goals.append(c.alpha22_wanted)
    if c.tq_prin_axes:
        fxy = 0.5 * math.sqrt(math.pi / 5.0) / norm
        for f in (X * Y, X * Z, Y * Z, X, Y, Z):
            fields.append(fxy * f * mask)
            goals.append(0.0)
    c.numconstraint = len(fields)
    if not fields:
        return c
    dev = torch.device("cuda") if device is None else torch.device(device)
    shape = (grids.nx, grids.ny, grids.nz)
    c.tconstraint = True
    c.constr_field = torch.tensor(np.stack([np.broadcast_to(f, shape) for f in fields]), dtype=torch.float64,
                                  device=dev).contiguous()
    n = c.numconstraint
    z = lambda: torch.zeros(n, dtype=torch.float64, device=dev)
    c.lambda_crank = torch.full((n,), -0.2, dtype=torch.float64, device=dev)
    c.dlambda_crank = z()
    c.goal_crank = torch.tensor(goals, dtype=torch.float64, device=dev)
    c.actual_crank, c.old_crank, c.actual_crank2, c.qcorr = z(), z(), z(), z()
    return c


def add_constraint_potential(potential, constraints, potential32=None):
    """U_q -= sum_i lambda_i Q_i for both isospins, in place on the device;
    potential32: the float32 copy of U the mixed-precision h psi reads,
    refreshed from the result."""
    c = constraints
    if not c.tconstraint:
        return potential
    nvox = c.constr_field[0].numel()
    check(lib().hfb_constraint_potential(c.numconstraint, nvox, ptr(c.constr_field), ptr(c.lambda_crank),
                                         ptr(potential), ptr(potential32), stream_handle()), "constraint_potential")
    return potential


def calculate_constraint_expectation_values(densities, constraints, grids):
    """(<Q_i>, |<Q_i^2> - <Q_i>^2/N|) over rho_n + rho_p (constraint.py:216-245)."""
    c = constraints
    rho = densities.rho[0] + densities.rho[1]
    numb = grids.wxyz * rho.sum()
    ex = grids.wxyz * (c.constr_field * rho).sum(dim=(1, 2, 3))
    m2 = grids.wxyz * (c.constr_field ** 2 * rho).sum(dim=(1, 2, 3))
    return ex, (m2 - ex ** 2 / numb).abs()


def update_lagrange_multipliers(constraints, expectations, variances, e0act, x0act):
    """constraint.py:248-276."""
    c = constraints
    den = 2.0 * variances + c.d0constr
    c.lambda_crank += -c.qepsconstr * (max(e0act, 1.0) / x0act) * (expectations - c.old_crank) / den
    c.qcorr.copy_(c.c0constr * (expectations - c.goal_crank) / den)
    c.actual_crank.copy_(expectations)
    c.actual_crank2.copy_(variances)
    return c


def apply_wavefunction_corrections(psi, constraints, levels):
    """psi_s *= exp(-wocc_s wstates_s sum_i qcorr_i Q_i) in place
    (constraint.py:279-318)."""
    c = constraints
    if not c.tconstraint:
        return psi
    nvox = c.constr_field[0].numel()
    check(lib().hfb_constraint_correct(prec_of(psi.dtype), ptr(psi), psi.shape[0], nvox, c.numconstraint,
                                       ptr(c.constr_field), ptr(c.qcorr), ptr(levels.wocc), ptr(levels.wstates),
                                       stream_handle()), "constraint_correct")
    return psi


def before_constraint_step(constraints, densities=None):
    """old <- actual (constraint.py:321-325).  tune_constraints does this
    itself as the first step of its update."""
    if constraints.tconstraint:
        constraints.old_crank.copy_(constraints.actual_crank)
    return constraints


def tune_constraints(constraints, densities, levels, grids, e0act, x0act, apply_corrections=True):
    """Expectation values, multipliers and wave-function corrections
    (constraint.py:328-344), one launch for the update (the previous
    expectation values become old_crank first) and one for the states."""
    c = constraints
    if not c.tconstraint:
        return c, levels.psi
    nvox = c.constr_field[0].numel()
    check(lib().hfb_constraint_tune(c.numconstraint, nvox, float(grids.wxyz), ptr(c.constr_field), ptr(densities.rho),
                                    float(c.qepsconstr), float(c.d0constr), float(c.c0constr), float(e0act),
                                    float(x0act), ptr(c.lambda_crank), ptr(c.goal_crank), ptr(c.actual_crank),
                                    ptr(c.old_crank), ptr(c.actual_crank2), ptr(c.qcorr), stream_handle()),
          "constraint_tune")
    if apply_corrections:
        apply_wavefunction_corrections(levels.psi, c, levels)
    return c, levels.psi


def check_constraint_convergence(constraints, tolerance=0.01):
    """constraint.py:379-390."""
    c = constraints
    if not c.tconstraint:
        return True
    act = c.actual_crank.cpu().numpy()
    goal = c.goal_crank.cpu().numpy()
    return bool(np.all(np.abs(act - goal) <= np.maximum(tolerance, np.abs(goal) * 0.01)))


def get_constraint_summary(constraints):
    """constraint.py:393-430."""
    c = constraints
    if not c.tconstraint:
        return {"enabled": False}
    out = {"enabled": True, "numconstraint": c.numconstraint, "converged": check_constraint_convergence(c)}
    g, a, l = (t.cpu().numpy() for t in (c.goal_crank, c.actual_crank, c.lambda_crank))
    i = 0
    for name, want in (("Q20", c.alpha20_wanted), ("Q22", c.alpha22_wanted)):
        if want > -1e90:
            out[name] = {"target": float(g[i]), "actual": float(a[i]), "lambda": float(l[i])}
            i += 1
    if c.tq_prin_axes:
        out["principal_axes"] = True
    return out
