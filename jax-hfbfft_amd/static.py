"""Static Hartree-Fock(+BCS) solver (static.py of the reference).

Same entry points and argument order as the reference: init_static,
harmosc, grstep, diagstep, statichf, statichf_with_benchmark.  The state
objects hold device tensors; every function launches the library's kernels
on the current HIP stream and returns the (mutated) objects like the
reference does."""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import kernels
from .densities import add_density, add_density_sweep
from .energies import _set_slots, integ_energy, sum_energy
from .inout import sp_properties
from .meanfield import density_derivs, skyrme
from .pairs import pair


@dataclass
class Static:
    tdiag: bool = False
    tlarge: bool = False
    tvaryx_0: bool = False
    ttime: bool = False
    tsort: bool = False
    maxiter: int = 200
    iternat: int = 100
    iternat_start: int = 40
    iteranneal: int = 0
    pairenhance: float = 0.0
    inibcs: int = 30
    inidiag: int = 30
    delstepbas: float = 2.0
    e0bas: float = 10.0
    outerpot: int = 0
    radinx: float = 3.1
    radiny: float = 3.1
    radinz: float = 3.1
    serr: float = 1.0e-20
    x0dmp: float = 0.45
    e0dmp: float = 100.0
    x0dmpmin: float = 0.45
    outertype: str = "N"
    hmatrix: torch.Tensor = None
    gapmatrix: torch.Tensor = None
    symcond: torch.Tensor = None
    lambda_save: torch.Tensor = None
    acc: torch.Tensor = None  # device (sumflu, delesum)
    qp: list = None  # [(quasi-particle energies, lower norms)] per isospin after the loop (lastdiag)
    constraints: object = None  # constraint.Constraints of a constrained run

    @property
    def sumflu(self):
        return float(self.acc[0].item())

    @property
    def delesum(self):
        return float(self.acc[1].item())



# Relative off-diagonal norm at which the lambda_lin eigensolve of diagstep
# stops: below the accuracy of its entries, 1e-13 for FP64 states and 1e-8
# for FP32 states, whose Gram entries carry about 1e-7 (measured on the
# MI355X: Ca40, Cr48 and O16 at iteration 25 move by less than 0.01 keV
# against 1e-10, with 2.0 instead of 2.2 sweeps per call at Ca40 and 2.3
# instead of 3.0 at Cr48).
JACOBI_TOL_F64 = 1e-13
JACOBI_TOL_F32 = 1e-8

def init_static(forces, levels, **kwargs):
    """Damping defaults and c.m. scaling of h2m (static.py:53-90)."""
    nst = max(int(levels.npsi[0]), int(levels.npsi[1]))
    known = {k: v for k, v in kwargs.items() if k in Static.__dataclass_fields__}
    st = Static(**known)
    st.x0dmpmin = st.x0dmp
    dev = levels.psi.device
    z = lambda: torch.zeros((2, nst, nst), dtype=torch.complex128, device=dev)
    st.hmatrix, st.gapmatrix, st.symcond, st.lambda_save = z(), z(), z(), z()
    st.acc = torch.zeros(2, dtype=torch.float64, device=dev)
    if forces.zpe == 0:
        forces.h2m = forces.h2m * ((levels.mass_number - 1.0) / levels.mass_number)
    return forces, st


# ---------------------------------------------------------------------------
# Initial states (static.py:580-738)
# ---------------------------------------------------------------------------
def shell_quantum_numbers(nps):
    """(i, j, k) of the states of one isospin: shells ka = i+j+k in the
    k-outer / j / i-inner order, each entry twice (static.py:613-659)."""
    out = []
    ka = 0
    while len(out) < nps:
        for k in range(ka + 1):
            for j in range(ka + 1):
                for i in range(ka + 1):
                    if ka == i + j + k:
                        out.extend([(i, j, k)] * 2)
        ka += 1
    return out[:nps]


def harmosc_host(grids, npsi, nneut, nprot, radin):
    """Gaussian times monomial initial states, normalised, as a host array
    (static.py:580-738)."""
    X = grids.x[:, None, None]
    Y = grids.y[None, :, None]
    Z = grids.z[None, None, :]
    w = grids.wxyz
    g = np.exp(-((X / radin[0]) ** 2 + (Y / radin[1]) ** 2 + (Z / radin[2]) ** 2))
    g = g / np.sqrt(w * np.sum(g * g))
    S = int(npsi[0] + npsi[1])
    psi = np.zeros((S, 2, grids.nx, grids.ny, grids.nz), dtype=np.complex128)
    start = 0
    for iq in range(2):
        nps = int(npsi[iq])
        qn = shell_quantum_numbers(nps)
        for m in range(nps):
            s = start + m
            if m == 0:
                psi[s, 0] = g
            else:
                i, j, k = qn[m]
                poly = (grids.x ** i)[:, None, None] * (grids.y ** j)[None, :, None] * (grids.z ** k)[None, None, :]
                psi[s, m % 2] = g * poly
            an = w * np.sum(np.abs(psi[s]) ** 2)
            if m == 0 or an > 1e-12:
                psi[s] /= np.sqrt(an)
        start += nps
    return psi


def harmosc(grids, levels, params, static):
    """static.py:580-738: initial states and occupations on the device."""
    psi = harmosc_host(grids, levels.npsi, levels.nneut, levels.nprot, (static.radinx, static.radiny, static.radinz))
    levels.psi.copy_(torch.from_numpy(psi).to(levels.psi.dtype))
    levels.wocc.zero_()
    levels.wocc[:levels.nneut] = 1.0
    n0 = int(levels.npsi[0])
    levels.wocc[n0:n0 + levels.nprot] = 1.0
    levels.wguv.zero_()
    return levels


# ---------------------------------------------------------------------------
# Gradient step (static.py:93-259)
# ---------------------------------------------------------------------------
def _ws(ctx, levels):
    S = levels.nstmax
    dt = levels.psi.dtype
    shp = tuple(levels.psi.shape)
    f64 = torch.float64
    return dict(w=ctx.work(("w", S), (S,), f64), wuv=ctx.work(("wuv", S), (S,), f64),
                pw=ctx.work(("pw", S), (S,), f64), pshift=ctx.work(("pshift", S), (S,), f64),
                pout=ctx.work(("pout", dt, S), shp, dt))


def grstep(forces, grids, levels, meanfield, params, static, defer_finish=False, tail=None):
    """Damped gradient step on all states (static.py:215-259).  With
    ``defer_finish`` the scalar tail (sp_energy, fluctuations, sumflu,
    delesum) is returned as a job for diagstep(..., finish_job=) to run in
    its first launch: (levels, static, job).  ``tail``: the previous
    iteration's deferred energy tail (kernels.tail_job), run by the first h psi
    launch before the step weights that read its v_pair maxima."""
    ctx = grids.ctx
    S = levels.nstmax
    n0 = int(levels.npsi[0])
    ws = _ws(ctx, levels)
    # the step's weights are computed by an extra workgroup of the first h psi
    # launch (its x pass does not read them)
    wjob = kernels.weights_job(S, n0, forces.tbcs, levels.wocc, levels.wstates, levels.wguv, levels.pairwg,
                               meanfield.v_pair, grids.nx * grids.ny * grids.nz, ws, vpair_max=meanfield.vpair_max,
                               tail=tail)
    fields = meanfield.fields(levels.psi.dtype)
    # h psi on the old states: ps1 (weighted) and psi_mf -> hampsi (static.py:152-173)
    # the per-state sums stay as tile partials, added up by the step and the
    # finishing kernel themselves
    _, _, _, hsum = kernels.hpsi(ctx, levels.psi, n0, *fields, ws["w"], ws["wuv"], want_del=False,
                                 out={"pout": ws["pout"], "pout_mf": levels.hampsi}, reduce_sums=False,
                                 weights_job=wjob)
    want_fl = params.mprint > 0 and params.iteration % params.mprint == 0
    fluct = {} if want_fl else None
    if static.e0dmp > 0.0:
        use_lag = (params.iteration > 1) and (not forces.tbcs)
        kernels.precond_step(ctx, levels.psi, ws["pout"], levels.lagrange, hsum, grids.wxyz, use_lag, ws["pw"],
                             ws["pshift"], static.e0dmp, forces.h2ma, static.x0dmp, fluct=fluct)
    else:  # static.py:181-185
        hsum = kernels.reduce_hsum(ctx, hsum)
        spe_new = grids.wxyz * hsum[:, 0]
        fac = (1.0 + static.x0dmp * (spe_new - levels.sp_energy)).to(levels.psi.real.dtype)
        levels.psi.mul_(fac.view(-1, 1, 1, 1, 1)).sub_(static.x0dmp * ws["pout"])
        want_fl = False
    # h psi on the new states -> hmfpsi, delpsi (static.py:195-208); with no
    # pairing v_pair == 0 and delpsi stays identically zero
    out = {"pout_mf": levels.hmfpsi}
    if forces.ipair != 0:
        out["pout_del"] = levels.delpsi
    kernels.hpsi(ctx, levels.psi, n0, *fields, ws["w"], ws["wuv"], want_pout=False, want_del=forces.ipair != 0,
                 want_partial=False, out=out)
    fin = (hsum, grids.wxyz, levels.sp_energy, levels.wocc, levels.wstates, fluct if want_fl else None,
           levels.sp_efluct1, levels.sp_efluct2, static.acc)
    if defer_finish:
        return levels, static, kernels.finish_job(*fin)
    kernels.grstep_finish(*fin)
    return levels, static


def _one_state_weights(forces, grids, levels, meanfield, s, iq):
    """hpsi and preconditioner weights of state s (isospin iq) as 1-element
    device vectors: the HFB weights, or (1, 0) in the BCS branch
    (static.py:152-170, 128-137)."""
    f64 = torch.float64
    dev = levels.psi.device
    out = {k: torch.empty(1, dtype=f64, device=dev) for k in ("w", "wuv", "pw", "pshift")}
    kernels.grstep_weights(1, 1 if iq == 0 else 0, forces.tbcs, levels.wocc[s:s + 1], levels.wstates[s:s + 1],
                           levels.wguv[s:s + 1], levels.pairwg[s:s + 1], meanfield.v_pair,
                           grids.nx * grids.ny * grids.nz, out)
    return out


def _one_state_step(forces, grids, levels, params, static, s, psi1, ps1, lag1, hsum, wts):
    """psi1 <- psi1 - x0dmp * laplace(residual) in place for one state, and
    its fluctuation norms (static.py:93-146)."""
    want_fl = params.mprint > 0 and params.iteration % params.mprint == 0
    fl = torch.empty((1, 2), dtype=torch.float64, device=psi1.device) if want_fl else None
    use_lag = (params.iteration > 1) and (not forces.tbcs)
    kernels.precond_step(grids.ctx, psi1, ps1, lag1, hsum, grids.wxyz, use_lag, wts["pw"], wts["pshift"],
                         static.e0dmp, forces.h2ma, static.x0dmp, fluct=fl)
    if want_fl:
        return fl[0, 0].sqrt(), fl[0, 1].sqrt()
    return levels.sp_efluct1[s].clone(), levels.sp_efluct2[s].clone()


def _one(t, shape):
    return t.reshape((1,) + tuple(shape[-4:])).contiguous()


def e0dmp_gt_zero(args):
    """Damped, preconditioned step of one state (static.py:93-146):
    args = (nst, iq, psin, ps1, lagrange, forces, grids, levels, meanfield,
    params, static) with nst the state index, psin/ps1/lagrange that state's
    (2, nx, ny, nz) spinors; returns (psin_new, sp_efluct1, sp_efluct2)."""
    nst, iq, psin, ps1, lagrange, forces, grids, levels, meanfield, params, static = args
    s, iq = int(nst), int(iq)
    psi1 = psin.reshape((1,) + tuple(psin.shape[-4:])).clone()
    p1 = _one(ps1, psin.shape)
    l1 = _one(lagrange, psin.shape)
    # h_exp = Re<psi|ps1>: the unscaled overlap enters the step kernel's sums
    g = torch.empty((1, 1, 1), dtype=torch.complex128, device=psi1.device)
    kernels.gram(grids.ctx, psi1, [p1], 1.0, g)
    hsum = torch.zeros((1, 2), dtype=torch.float64, device=psi1.device)
    hsum[0, 1] = g.real.reshape(())
    wts = _one_state_weights(forces, grids, levels, meanfield, s, iq)
    efl1, efl2 = _one_state_step(forces, grids, levels, params, static, s, psi1, p1, l1, hsum, wts)
    return psi1[0], efl1, efl2


def grstep_helper(nst, iq, spe_mf, psin, lagrange, forces, grids, levels, meanfield, params, static):
    """Gradient step of one state (static.py:148-213): h psi with the weights
    of the BCS or HFB branch, the damped step (e0dmp_gt_zero, or the simple
    step when e0dmp <= 0), h psi of the new state.  Returns (psin, psi_mf,
    hmfpsi, delpsi, denerg, spe_mf_new, sp_efluct1, sp_efluct2)."""
    ctx = grids.ctx
    s, iq = int(nst), int(iq)
    psi1 = psin.reshape((1,) + tuple(psin.shape[-4:])).clone()
    l1 = _one(lagrange, psin.shape)
    n0 = 1 if iq == 0 else 0
    wts = _one_state_weights(forces, grids, levels, meanfield, s, iq)
    fields = meanfield.fields(psi1.dtype)
    ps1, psi_mf, _, hsum = kernels.hpsi(ctx, psi1, n0, *fields, wts["w"], wts["wuv"], want_del=False)
    hsum = hsum.clone()
    spe_new = grids.wxyz * hsum[0, 0]
    spe_old = torch.as_tensor(spe_mf, dtype=torch.float64, device=psi1.device)
    if static.e0dmp > 0.0:
        efl1, efl2 = _one_state_step(forces, grids, levels, params, static, s, psi1, ps1, l1, hsum, wts)
    else:  # static.py:181-185
        fac = (1.0 + static.x0dmp * (spe_new - spe_old)).to(psi1.real.dtype)
        psi1.mul_(fac).sub_(static.x0dmp * ps1)
        efl1, efl2 = levels.sp_efluct1[s].clone(), levels.sp_efluct2[s].clone()
    _, hmf, dl, _ = kernels.hpsi(ctx, psi1, n0, *fields, wts["w"], wts["wuv"], want_pout=False, want_del=True,
                                 want_partial=False)
    denerg = (spe_old - spe_new) / spe_new.abs()
    return psi1[0], psi_mf[0], hmf[0], dl[0], denerg, spe_new, efl1, efl2


# ---------------------------------------------------------------------------
# Orthonormalisation / subspace (static.py:261-577)
# ---------------------------------------------------------------------------
def _state_major(psi_2d, nst):
    """The (nst, K) state-major block behind a reference-style psi_2d (K, nst)
    (static.py:277-284), or a state block (nst, 2, nx, ny, nz) as it is.
    A transposed view of a state block is used without a copy."""
    if psi_2d.dim() == 2 and psi_2d.shape[1] == nst:
        a = psi_2d.transpose(0, 1)
        return a if a.is_contiguous() else a.contiguous()
    if psi_2d.shape[0] != nst:
        raise ValueError(f"psi_2d of shape {tuple(psi_2d.shape)} does not hold {nst} states")
    a = psi_2d.reshape(nst, -1)
    return a if a.is_contiguous() else a.contiguous()


def loewdin_orthonormalize(psi_2d, wxyz, nst):
    """S = wxyz psi_2d^H psi_2d, U = S^{-1/2} with the eigenvalue clamp, and
    sp_norm = Re diag S (static.py:463-482).  psi_2d: (2N, nst) as in the
    reference (e.g. block.reshape(nst, -1).T) or a state block; U is
    (nst, nst) complex128 with psi_new = psi_2d U."""
    nst = int(nst)
    A = _state_major(psi_2d, nst)
    S = torch.empty((1, nst, nst), dtype=torch.complex128, device=A.device)
    kernels.gram(None, A, [A], float(wxyz), S)
    return kernels.loewdin(S[0])


def construct_hfb_matrices(unitary, psi_2d, energies, forces, grids, levels, static, iq, start, end, nst):
    """H = U^H (wxyz psi_2d^H hmfpsi) U, the gap matrix likewise, sp_energy,
    deltaf, lambda and its Hermitian/anti-Hermitian parts, the fluctuation
    measures and lagrange = psi_new lambda+ for isospin iq, states
    start:end (static.py:484-577).  psi_2d holds the states before the
    rotation; levels.psi[start:end] the rotated ones."""
    nst = int(nst)
    start, end = int(start), int(end)
    if end - start != nst:
        raise ValueError("end - start must equal nst")
    ctx = grids.ctx
    A = _state_major(psi_2d, nst)
    ipair = forces.ipair != 0
    Bs = [levels.hmfpsi[start:end].reshape(nst, -1)]
    if ipair:
        Bs.append(levels.delpsi[start:end].reshape(nst, -1))
    for b in Bs:
        if b.dtype != A.dtype:
            raise ValueError("psi_2d and the level arrays need the same dtype")
    M = torch.empty((len(Bs), nst, nst), dtype=torch.complex128, device=A.device)
    kernels.gram(ctx, A, Bs, grids.wxyz, M)
    U = unitary.to(torch.complex128).contiguous()
    lam = torch.empty((nst, nst), dtype=torch.complex128, device=A.device)
    kernels.hfb_matrices(ctx, int(iq), start, nst, U, M[0], M[1] if ipair else None, ipair, forces.tbcs, levels,
                         static, energies.vec, lam)
    kernels.rotate(levels.psi[start:end], lam, None, levels.lagrange[start:end], None)
    return energies, levels, static



def diagstep(energies, forces, grids, levels, static, diagonalize=False, construct=True, npsi_neutron=None,
             npsi_proton=None, finish_job=None):
    """Loewdin orthonormalisation, optional subspace diagonalisation and the
    HFB matrices / Lagrange field (static.py:261-461).  ``finish_job``: a
    gradient step's deferred tail, run by the Gram launch."""
    ctx = grids.ctx
    n0 = int(levels.npsi[0]) if npsi_neutron is None else int(npsi_neutron)
    n1 = int(levels.npsi[1]) if npsi_proton is None else int(npsi_proton)
    nmax = static.hmatrix.shape[-1]
    use_gap = forces.ipair != 0 and construct
    mats, uo, blocks, outs = [], [], [], []
    for iq, (a, b) in enumerate(((0, n0), (n0, n0 + n1))):
        nst = b - a
        A = levels.psi[a:b]
        Bs = [A, levels.hmfpsi[a:b]]
        if diagonalize:
            Bs.append(levels.hampsi[a:b])
        if use_gap:
            Bs.append(levels.delpsi[a:b])
        M = ctx.work(("gram_out", iq, len(Bs), nst), (len(Bs), nst, nst), torch.complex128)
        blocks.append((A, Bs))
        outs.append(M)
        mats += [M[0], M[1], M[2] if diagonalize else None, M[len(Bs) - 1] if use_gap else None]
        uo += [ctx.work(("U", iq, nst), (nst, nst), torch.complex128), ctx.work(("M2", iq, nst), (nst, nst),
                                                                              torch.complex128)]
    # overlap, H_old, lambda_lin and G_old of both isospins in one pass over the states
    kernels.gram2(ctx, blocks, grids.wxyz, outs, finish_job=finish_job)
    # blocks above 36 states keep the subspace work matrices in global memory
    work = ctx.work(("diag_work", nmax), (10 * nmax * nmax,), torch.complex128) if max(n0, n1) > 36 else None
    jtol = JACOBI_TOL_F32 if levels.psi.dtype == torch.complex64 else JACOBI_TOL_F64
    kernels.diag_small(ctx, n0, n1, nmax, mats, diagonalize, forces.ipair != 0, forces.tbcs, construct, levels, static,
                       energies.vec, uo, work, jacobi_tol=jtol)
    # psi <- psi U and lagrange <- psi M2 in place (every workgroup reads its
    # columns of psi before it writes them)
    rot = []
    for iq, (a, b) in enumerate(((0, n0), (n0, n0 + n1))):
        rot.append((levels.psi[a:b], uo[2 * iq], uo[2 * iq + 1], levels.psi[a:b], levels.lagrange[a:b])
                   if b > a else None)
    kernels.rotate2(rot, levels.psi[0].numel())
    return energies, levels, static


# ---------------------------------------------------------------------------
# Drivers (static.py:1394-1927)
# ---------------------------------------------------------------------------
class LoopState:
    """Per-run device scratch shared by the loop's calls."""

    def __init__(self, grids, densities):
        self.derivs = density_derivs(grids, densities)


def prelude(coulomb, densities, energies, forces, grids, levels, meanfield, moment, params, static, pairs):
    """Steps before the loop (static.py:1700-1744).  With params.trestart
    (a run continued from load_checkpoint) the iteration number is kept and
    the first orthonormalisation is skipped (static.py:1700-1706)."""
    if not params.trestart:
        params.iteration = 0
        diagstep(energies, forces, grids, levels, static, False, True)
    add_density(densities, grids, levels)
    derivs = density_derivs(grids, densities)
    skyrme(coulomb, densities, forces, grids, meanfield, params, static, constraints=static.constraints,
           derivs=derivs, E=energies.vec, mass_number=levels.mass_number)
    grstep(forces, grids, levels, meanfield, params, static)
    if forces.ipair != 0:
        pair(levels, meanfield, forces, params, grids, pairs, E=energies.vec, check_status=False)
    diagstep(energies, forces, grids, levels, static, False, True)
    sp_properties(forces, grids, levels, moment)
    integ_energy(coulomb, densities, energies, forces, grids, levels, params, pairs, derivs=derivs)
    sum_energy(energies, levels, meanfield, pairs)
    if static.tvaryx_0:
        static.x0dmp *= 3.0
    return derivs


def iterate(i, coulomb, densities, energies, forces, grids, levels, meanfield, moment, params, static, pairs,
            tbcssav, derivs=None, check=True, defer_tail=True):
    """One SCF iteration: the loop body of static.py:1755-1882 without output.
    With ``defer_tail`` the energy tail (ecorrp, the integrals, sum_energy) is
    left to the next iteration's first launch, or to the first read of the
    energies (Energies.flush)."""
    params.iteration = i
    forces.tbcs = True if i <= static.inibcs else tbcssav
    static.tdiag = not (i > static.inidiag)
    _, _, fin = grstep(forces, grids, levels, meanfield, params, static, defer_finish=True, tail=energies.tail)
    if forces.tbcs:
        static.tdiag = True
    diagstep(energies, forces, grids, levels, static, static.tdiag, True, finish_job=fin)
    if forces.ipair != 0:
        pair(levels, meanfield, forces, params, grids, pairs, E=energies.vec, check_status=False)
    if forces.ipair != 0 and static.iternat > 0 and i > static.iternat_start and not forces.tbcs:
        # configuration-space sub-iterations (static.py:1795-1797)
        for iq in range(2):
            hfb_natural(forces, grids, levels, meanfield, params, static, iq, static.iternat, pairs=pairs)
    c = static.constraints
    if c is not None and c.tconstraint:
        # multipliers and the correction of the states (constraint.py:328-344)
        from .constraint import tune_constraints
        tune_constraints(c, densities, levels, grids, static.e0dmp, static.x0dmp)
    # density with the 0.2/0.8 mixing of rho, tau, chi and the per-state
    # properties from the same transforms (static.py:1800-1833)
    # (sp_properties, static.py:1831, is finished in the same launch)
    _, spart = add_density_sweep(densities, grids, levels, mix=True, addnew=0.2, moment=moment, want_sp=True,
                                 sp_h2m=forces.h2m)
    # the derived fields and the Coulomb potential share their launches
    derivs = density_derivs(grids, densities, out=derivs, coulomb=coulomb if params.tcoul else None, e2=params.e2)
    # skyrme, sp_properties, integ_energy and sum_energy (static.py:1810-1833)
    # with the scalar finishing steps of the first, third and fourth in one
    # launch at the end
    skyrme(coulomb, densities, forces, grids, meanfield, params, static, constraints=c, derivs=derivs,
           E=energies.vec, mass_number=levels.mass_number, finish=False, with_energy=True,
           wcoul_ready=bool(params.tcoul))
    if static.tsort:  # static.py:1829-1831
        for iq in range(2):
            sort_states(forces, grids, levels, iq)
    # integ_energy's integrals come from the same launch as the potentials;
    # its pairing energies go to the energy vector as integ_energy does
    _set_slots(energies.vec, kernels.EN["EPAIR"], pairs.epair)
    tail = kernels.tail_job(grids.ctx, forces.to_c(params, levels.mass_number), densities, levels, energies.vec)
    if defer_tail:
        energies.defer_tail(grids.ctx, tail)
    else:
        kernels.run_tail(grids.ctx, tail)
        energies.tail, energies.stale = None, False
    converged = False
    if check or static.tvaryx_0:
        e = energies.host()
        converged = e["efluct1"] < static.serr and i > 1
        if static.tvaryx_0:
            if ((e["ehf"] < energies.ehfprev and e["efluct1"] < energies.efluct1prev * (1.0 - 1.0e-5))
                    or e["efluct2"] < energies.efluct2prev * (1.0 - 1.0e-5)):
                static.x0dmp *= 1.005
            else:
                static.x0dmp *= 0.8
            static.x0dmp = float(np.clip(static.x0dmp, static.x0dmpmin, static.x0dmpmin * 5.0))
            energies.efluct1prev, energies.efluct2prev, energies.ehfprev = e["efluct1"], e["efluct2"], e["ehf"]
    return converged, derivs


def statichf(coulomb, densities, energies, forces, grids, levels, meanfield, moment, params, static, pairs,
             output_writer=None, output_interval=5, callback=None):
    """Prelude + SCF loop (static.py:1394-1601).  ``output_writer`` is not
    required (the reference needs one, Q1); ``callback(i, state)`` runs after
    every iteration."""
    firstiter = params.iteration + 1 if params.trestart else 1
    derivs = prelude(coulomb, densities, energies, forces, grids, levels, meanfield, moment, params, static, pairs)
    tbcssav = forces.tbcs
    if output_writer is not None:
        from .output import complete_sinfo_with_fortran_output
        complete_sinfo_with_fortran_output(coulomb, densities, energies, forces, grids, levels, meanfield, moment,
                                           params, static, pairs, output_writer)
    for i in range(firstiter, static.maxiter + 1):
        conv, derivs = iterate(i, coulomb, densities, energies, forces, grids, levels, meanfield, moment, params,
                               static, pairs, tbcssav, derivs)
        if output_writer is not None and i % output_interval == 0:
            complete_sinfo_with_fortran_output(coulomb, densities, energies, forces, grids, levels, meanfield,
                                               moment, params, static, pairs, output_writer)
        if params.mrest > 0 and params.wffile not in ("", "none") and i % params.mrest == 0:
            save_checkpoint(params.wffile, levels, params, static, pairs)
        if callback is not None:
            callback(i, levels, energies)
        if conv:
            break
    # final quasi-particle diagonalisation (static.py:1912-1925)
    static.qp = [lastdiag(forces, grids, levels, params, static, pairs, iq) for iq in range(2)]
    return coulomb, densities, energies, forces, grids, levels, meanfield, moment, params, static


def statichf_with_benchmark(coulomb, densities, energies, forces, grids, levels, meanfield, moment, params, static,
                            pairs, output_writer=None, output_interval=10):
    """statichf with per-iteration wall times (static.py:1666-1927); returns
    the reference's tuple plus the list of iteration durations."""
    firstiter = params.iteration + 1 if params.trestart else 1
    derivs = prelude(coulomb, densities, energies, forces, grids, levels, meanfield, moment, params, static, pairs)
    tbcssav = forces.tbcs
    times = []
    for i in range(firstiter, static.maxiter + 1):
        t0 = time.perf_counter()
        conv, derivs = iterate(i, coulomb, densities, energies, forces, grids, levels, meanfield, moment, params,
                               static, pairs, tbcssav, derivs)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        if conv:
            break
    static.qp = [lastdiag(forces, grids, levels, params, static, pairs, iq) for iq in range(2)]
    return (coulomb, densities, energies, forces, grids, levels, meanfield, moment, params, static), times


# ---------------------------------------------------------------------------
# State ordering, plane-wave start, checkpoints
# (static.py:807-918, 1338-1392, 1700-1701)
# ---------------------------------------------------------------------------
SORTED_FIELDS = ("psi", "lagrange", "sp_energy", "deltaf", "sp_norm", "sp_efluct1", "sp_efluct2", "wocc", "wguv",
                 "wstates", "pairwg")
# per-state results of sp_properties that this build computes before the
# sort (in the density sweep); moving them along is the same as computing
# them after it, as the reference does
SP_FIELDS = ("sp_kinetic", "sp_orbital", "sp_spin", "sp_parity")


def sort_states(forces, grids, levels, isospin):
    """Order the states of one isospin block by sp_energy, stable, moving
    psi, lagrange and the per-state arrays along (static.py:1338-1392).  The
    block is the isospin's npsi states (the reference slices by nneut/nprot,
    the same block only when npsi = (N, Z)).  Arrays are permuted in place on
    the device."""
    n0 = int(levels.npsi[0])
    a, b = (0, n0) if isospin == 0 else (n0, n0 + int(levels.npsi[1]))
    if b - a <= 1:
        return levels
    perm = torch.argsort(levels.sp_energy[a:b], stable=True) + a
    for name in SORTED_FIELDS + SP_FIELDS:
        t = getattr(levels, name)
        t[a:b].copy_(t.index_select(0, perm))
    return levels


def planewaves_host(grids, npsi, nneut, nprot, nucleon_mass, hbc=197.328910, nmax=7):
    """Plane-wave states (static.py:807-918) as a host array: per isospin
    block the wave numbers k_a in -nmax//2 .. nmax//2 ordered (stable) by
    hbc^2/2m sum_a ((2 pi k_a + bang_a)/(n_a d_a))^2, each first with spin up,
    then spin down, normalised."""
    r = range(-nmax // 2, nmax // 2 + 1)
    ks = [(kx, ky, kz) for kx in r for ky in r for kz in r]
    bang = (grids.bangx, grids.bangy, grids.bangz)
    ns = (grids.nx, grids.ny, grids.nz)
    ds = (grids.dx, grids.dy, grids.dz)
    en = np.array([hbc ** 2 / (2.0 * nucleon_mass) * sum(((2.0 * math.pi * k[c] + bang[c]) / (ns[c] * ds[c])) ** 2
                                                         for c in range(3)) for k in ks])
    order = [ks[i] for i in np.argsort(en, kind="stable")]
    I = np.arange(grids.nx)[:, None, None]
    J = np.arange(grids.ny)[None, :, None]
    K = np.arange(grids.nz)[None, None, :]
    S = int(npsi[0] + npsi[1])
    psi = np.zeros((S, 2, grids.nx, grids.ny, grids.nz), dtype=np.complex128)
    start = 0
    for iq in range(2):
        for m in range(int(npsi[iq])):
            k = order[m // 2]
            ph = (I * ((2.0 * math.pi * k[0] + bang[0]) / grids.nx) + J * ((2.0 * math.pi * k[1] + bang[1]) / grids.ny)
                  + K * ((2.0 * math.pi * k[2] + bang[2]) / grids.nz))
            psi[start + m, m % 2] = np.cos(ph) + 1j * np.sin(ph)
            psi[start + m] /= math.sqrt(np.sum(np.abs(psi[start + m]) ** 2) * grids.wxyz)
        start += int(npsi[iq])
    return psi


def planewaves(grids, levels, params, static, forces):
    """Plane-wave initial states and occupations on the device
    (static.py:807-918; the reference reads an undefined global for the
    nucleon mass, here it is forces.nucleon_mass)."""
    psi = planewaves_host(grids, levels.npsi, levels.nneut, levels.nprot, forces.nucleon_mass, params.hbc)
    levels.psi.copy_(torch.from_numpy(psi).to(levels.psi.dtype))
    levels.wocc.zero_()
    levels.wocc[:levels.nneut] = 1.0
    n0 = int(levels.npsi[0])
    levels.wocc[n0:n0 + levels.nprot] = 1.0
    levels.wguv.zero_()
    return levels


CHECKPOINT_ARRAYS = ("psi", "lagrange", "hampsi", "hmfpsi", "delpsi", "sp_energy", "sp_norm", "sp_efluct1",
                     "sp_efluct2", "sp_kinetic", "sp_orbital", "sp_spin", "sp_parity", "wocc", "wguv", "pairwg",
                     "wstates", "deltaf")
CHECKPOINT_STATIC = ("hmatrix", "gapmatrix", "symcond", "lambda_save", "acc")
CHECKPOINT_PAIRS = ("eferm", "epair", "avdelt", "avdeltv2", "avg")


def save_checkpoint(path, levels, params, static, pairs):
    """Write the wave functions and the solver state a restart needs (the
    reference has the restart flag params.trestart and the file name
    params.wffile, static.py:1700-1701, but no writer): one .npz with the
    level arrays, the subspace matrices, the pairing results, the iteration
    number and x0dmp."""
    out = {f"levels.{k}": getattr(levels, k).detach().cpu().numpy() for k in CHECKPOINT_ARRAYS}
    out.update({f"static.{k}": getattr(static, k).detach().cpu().numpy() for k in CHECKPOINT_STATIC})
    out.update({f"pairs.{k}": getattr(pairs, k).detach().cpu().numpy() for k in CHECKPOINT_PAIRS})
    out["npsi"] = np.asarray(levels.npsi, dtype=np.int64)
    out["iteration"] = np.int64(params.iteration)
    out["x0dmp"] = np.float64(static.x0dmp)
    out["pairs_firstcall"] = np.int64(bool(pairs.firstcall))
    tmp = f"{path}.tmp.npz"
    np.savez(tmp, **out)
    import os
    os.replace(tmp, path)


def load_checkpoint(path, levels, params, static, pairs):
    """Read a save_checkpoint file into the existing device arrays (shapes and
    precision of this run; complex128 data is narrowed for a mixed run) and
    set params.trestart and params.iteration for the restart."""
    with np.load(path, allow_pickle=False) as d:
        if tuple(d["npsi"]) != tuple(int(v) for v in levels.npsi):
            raise ValueError(f"checkpoint has npsi {tuple(d['npsi'])}, this run {tuple(levels.npsi)}")
        for pre, obj, names in (("levels", levels, CHECKPOINT_ARRAYS), ("static", static, CHECKPOINT_STATIC),
                                ("pairs", pairs, CHECKPOINT_PAIRS)):
            for k in names:
                t = getattr(obj, k)
                a = d[f"{pre}.{k}"]
                if tuple(a.shape) != tuple(t.shape):
                    raise ValueError(f"checkpoint array {pre}.{k} has shape {a.shape}, expected {tuple(t.shape)}")
                t.copy_(torch.from_numpy(a).to(t.dtype))
        params.iteration = int(d["iteration"])
        static.x0dmp = float(d["x0dmp"])
        pairs.firstcall = bool(int(d["pairs_firstcall"]))
    params.trestart = True
    return levels, params, static, pairs


# ---------------------------------------------------------------------------
# HFB in the canonical basis (static.py:920-1168)
# ---------------------------------------------------------------------------
def _block(levels, isospin):
    n0 = int(levels.npsi[0])
    return (0, n0) if isospin == 0 else (n0, n0 + int(levels.npsi[1]))


def hfb_natural(forces, grids, levels, meanfield, params, static, isospin, iternat, pairs=None):
    """Sub-iterations of the canonical basis of one isospin in configuration
    space (static.py:920-1070) over its npsi states: damped steps along
    symcond with QR re-orthonormalisation, transformed h and gap matrices,
    s.p. energies and gaps, the one-isospin pairing update between
    sub-iterations and the full pair at the last one, then psi <- psi cm and
    lagrange <- psi_new htran (as the reference does), hmatrix, gapmatrix,
    lambda_save.  The reference's version slices by nneut/nprot and calls the
    pairing with no pairing data; here ``pairs`` is that data (required with
    pairing)."""
    a, b = _block(levels, isospin)
    nst = b - a
    if nst <= 0:
        return levels, static
    ipair = forces.ipair != 0
    if pairs is None:
        if ipair:
            raise ValueError("hfb_natural with pairing needs the pairing data (pairs=)")
        from .pairs import Pairs
        pairs = Pairs()
    ctx = grids.ctx
    c128 = torch.complex128
    work = ctx.work(("natural", nst), (10 * nst * nst + nst,), c128)
    npart = levels.nneut if isospin == 0 else levels.nprot
    args = (isospin, a, nst, int(iternat), ipair, static.delstepbas, static.e0bas, npart,
            forces.pair_cutoff[isospin], forces.state_cutoff[isospin], forces.softcut_range, levels, static, pairs,
            work)
    kernels.natural(0, *args)
    if ipair:
        pair(levels, meanfield, forces, params, grids, pairs, check_status=False)
    cm = ctx.work(("natural_cm", nst), (nst, nst), c128)
    m2 = ctx.work(("natural_m2", nst), (nst, nst), c128)
    kernels.natural(1, *args, cm_out=cm, m2_out=m2)
    kernels.rotate(levels.psi[a:b], cm, m2, levels.psi[a:b], levels.lagrange[a:b])
    return levels, static


def lastdiag(forces, grids, levels, params, static, pairs, isospin):
    """Quasi-particle energies (ascending) and the norms of their lower
    components for one isospin (static.py:1072-1168) over its npsi states;
    device tensors (nst,), (nst,)."""
    a, b = _block(levels, isospin)
    return kernels.lastdiag(isospin, a, b - a, float(pairs.eferm[isospin]), static, levels.wocc)
