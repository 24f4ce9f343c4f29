"""Backward of the Skyrme energy functional on the GPU.

``EnergyFunctional`` is a ``torch.autograd.Function`` whose forward is the
functional integral ``ehfint`` of energies.py:90-306 (summed at
energies.py:267) for a batch of coupling sets sharing one set of densities,
and whose backward is the hand-written kernel ``hfb_energy_bwd``: gradients
with respect to the densities rho, tau, J (sodens), j (current), s (sdens)
and to the ten couplings t0..t3, x0..x3, alpha (= power) and W0 (= t4).

The reference has no derivative of the functional; what it has are the mean
fields of meanfield.py:163-391, which are those derivatives (times wxyz)
for the time-even part: dE/drho_q = wxyz U_q, dE/dtau_q = wxyz B_q,
dE/dJ_q = wxyz W_q, dE/ds_q = wxyz S_q, and dE/dj_q = wxyz times the spin
part of A_q (ehfint has no j^2 term).  The derived fields (Laplacian,
gradient, divergence, curls) and the Coulomb potential are recomputed in the
backward pass from the saved densities rather than stored.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass
from types import SimpleNamespace

import torch

from . import kernels
from .meanfield import density_derivs

COUPLINGS = ("t0", "t1", "t2", "t3", "x0", "x1", "x2", "x3", "power", "t4")
MAX_BATCH = 8


@dataclass
class FunctionalEnv:
    """What the functional needs besides densities and couplings: the grid
    (and its library context), the force whose other parameters (b4p, h2m,
    pairing, ex, zpe) stay fixed, the run parameters (e2, tcoul) and the mass
    number (zero-point term)."""
    grids: object
    forces: object
    params: object
    mass_number: float


def coupling_vector(forces) -> torch.Tensor:
    """(t0, t1, t2, t3, x0, x1, x2, x3, power, t4) of a force, float64 on the host."""
    return torch.tensor([float(getattr(forces, k)) for k in COUPLINGS], dtype=torch.float64)


def forces_at(template, values, params):
    """A copy of ``template`` with the ten couplings replaced and the b- and
    C-couplings derived again (forces.py:95-122)."""
    f = copy.copy(template)
    for k, v in zip(COUPLINGS, values.tolist()):
        setattr(f, k, float(v))
    f.derive(params.pi, params.e2)
    return f


def _densities(rho, tau, sodens, current, sdens):
    return SimpleNamespace(rho=rho.contiguous(), tau=tau.contiguous(), sodens=sodens.contiguous(),
                           current=current.contiguous(), sdens=sdens.contiguous())


def _derived(env, dens):
    der = density_derivs(env.grids, dens)
    wcoul = kernels.poisson(env.grids.ctx, dens.rho, env.params.e2) if env.params.tcoul else None
    return der, wcoul


def _force_structs(env, couplings):
    return [forces_at(env.forces, couplings[b], env.params).to_c(env.params, env.mass_number)
            for b in range(couplings.shape[0])]


class EnergyFunctional(torch.autograd.Function):
    """E_b = ehfint(couplings[b]; rho, tau, J, j, s) for b < B <= 8.

    couplings: (B, 10) float64 on the host (the order of COUPLINGS);
    densities: float64 HIP tensors (2, n, n, n) / (2, 3, n, n, n).
    Returns E (B,) on the device."""

    @staticmethod
    def forward(ctx, couplings, rho, tau, sodens, current, sdens, env):
        if couplings.dim() != 2 or couplings.shape[1] != len(COUPLINGS):
            raise ValueError(f"couplings must be (B, {len(COUPLINGS)})")
        B = couplings.shape[0]
        if not 1 <= B <= MAX_BATCH:
            raise ValueError(f"1 <= batch <= {MAX_BATCH}")
        dens = _densities(rho, tau, sodens, current, sdens)
        der, wcoul = _derived(env, dens)
        E = torch.zeros((B, kernels.EN_LEN), dtype=torch.float64, device=rho.device)
        for b, fc in enumerate(_force_structs(env, couplings.detach().cpu())):
            kernels.integ_energy(env.grids.ctx, fc, dens, der, wcoul, E[b])
        ctx.save_for_backward(couplings, rho, tau, sodens, current, sdens)
        ctx.env = env
        return E[:, kernels.EN["EHFINT"]].clone()

    @staticmethod
    def backward(ctx, grad_e):
        couplings, rho, tau, sodens, current, sdens = ctx.saved_tensors
        env = ctx.env
        dens = _densities(rho, tau, sodens, current, sdens)
        der, wcoul = _derived(env, dens)  # recomputed, not stored
        host = couplings.detach().cpu()
        tx = host.to(device=rho.device, dtype=torch.float64).contiguous()
        gout = grad_e.detach().to(device=rho.device, dtype=torch.float64).contiguous()
        want_fields = any(ctx.needs_input_grad[1:6])
        want_c = ctx.needs_input_grad[0]
        fields, dc = kernels.energy_bwd(env.grids.ctx, _force_structs(env, host), tx, gout, dens, der, wcoul,
                                        want_fields=want_fields, want_couplings=want_c)
        g = fields or {}
        d_c = dc.to(couplings.device) if dc is not None else None
        return d_c, g.get("rho"), g.get("tau"), g.get("sodens"), g.get("current"), g.get("sdens"), None


def energy(couplings, densities, env):
    """ehfint for each row of ``couplings`` (B, 10) on ``densities`` (a
    Densities object), differentiable in both."""
    d = densities
    return EnergyFunctional.apply(couplings, d.rho, d.tau, d.sodens, d.current, d.sdens, env)


def functional_gradients(densities, env, couplings=None):
    """dE/d(density fields) and dE/d(couplings) at one coupling set (the
    force's own when ``couplings`` is None), as a dict of tensors."""
    c = coupling_vector(env.forces) if couplings is None else couplings.reshape(-1).detach().cpu()
    c = c.reshape(1, -1).clone().requires_grad_(True)
    leaves = [t.detach().clone().requires_grad_(True) for t in (densities.rho, densities.tau, densities.sodens,
                                                              densities.current, densities.sdens)]
    E = EnergyFunctional.apply(c, *leaves, env)
    grads = torch.autograd.grad(E.sum(), [c] + leaves)
    out = dict(zip(("couplings", "rho", "tau", "sodens", "current", "sdens"), grads))
    out["couplings"] = out["couplings"].reshape(-1)
    out["energy"] = E.detach()
    return out
