"""FP64 CPU restatement of jax-hfbfft's static Skyrme-HF(+BCS) iteration.

TEST INFRASTRUCTURE ONLY.  This module is the oracle the HIP path is checked
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline -- never as a product code path.

Every function restates the reference algorithm (ascsn/jax-hfbfft) from its
source text; the ``file:line`` in each docstring names what it follows.
Arithmetic is NumPy/SciPy FP64 throughout.  FFTs use ``scipy.fft`` with the
same normalisation arguments as the reference; Hermitian eigenproblems use
``numpy.linalg.eigh`` with ``UPLO='L'`` (the reference's XLA ``eigh`` reads
the lower triangle); the BCS root uses ``scipy.optimize.brentq`` with the
reference's bracket and tolerance.

Arrays follow the reference's layout: a state block is ``(S, 2, nx, ny, nz)``
complex128 (spin second, z fastest); isospin fields are ``(2, nx, ny, nz)``
and vector fields ``(2, 3, nx, ny, nz)`` float64.

Parity pins (what checks each function independently of this text) are
listed in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as sfft
from scipy.optimize import brentq

WORKERS = max(1, len(os.sched_getaffinity(0))) if hasattr(os, "sched_getaffinity") else 1

# ---------------------------------------------------------------------------
# Constants and force table                               params.py:53-54
# ---------------------------------------------------------------------------
HBC = 197.328910
E2 = 1.4399784085965135

# SLy4 is the reference's only entry (_forces.yml:1-25).  SV-bas is not in the
# reference; its couplings are the published ones (P. Kluepfel et al., PRC 79,
# 034310 (2009)), in the t/x parametrisation Sky3D tabulates.  Sky3D gives
# SV-bas zpe=2, which the reference's c.m. handling ignores (no h2m scaling,
# no e_zpe): static.py:87-88, energies.py:259-263.
FORCE_TABLE = {
    "SLy4": dict(ex=1, zpe=0, h2m=[20.735530, 20.735530],
                 t0=-2488.913, t1=486.818, t2=-546.395, t3=13777.0, t4=123.0,
                 x0=0.834, x1=-0.344, x2=-1.0, x3=1.354, b4p=61.5, power=0.1666667,
                 vdi=dict(v0prot=295.369, v0neut=286.669, rho0pr=0.0),
                 dddi=dict(v0prot=-1.0, v0neut=-1.0, rho0pr=0.16)),
    "SV-bas": dict(ex=1, zpe=2, h2m=[20.7212601, 20.7498207],
                   t0=-1879.64001, t1=313.749063, t2=112.676224, t3=12527.3890,
                   t4=124.633260, x0=0.258546602, x1=-0.381689970, x2=-2.82364431,
                   x3=0.123068743, b4p=34.1161, power=0.3,
                   vdi=dict(v0prot=300.0, v0neut=300.0, rho0pr=0.0),
                   dddi=dict(v0prot=674.5227, v0neut=606.6355, rho0pr=0.2011)),
}


def rfft_n(x, axis):
    return sfft.fft(x, axis=axis, workers=WORKERS)


def rifft_fwdnorm(x, axis):
    """Unnormalised inverse DFT (numpy/jax ``norm="forward"`` on ifft)."""
    return sfft.ifft(x, axis=axis, norm="forward", workers=WORKERS)


# ---------------------------------------------------------------------------
# Forces                                                   forces.py:60-143
# ---------------------------------------------------------------------------
@dataclass
class Forces:
    name: str
    ex: int
    zpe: int
    h2m: np.ndarray
    t0: float
    t1: float
    t2: float
    t3: float
    t4: float
    x0: float
    x1: float
    x2: float
    x3: float
    b4p: float
    power: float
    v0prot: float
    v0neut: float
    rho0pr: float
    ipair: int
    tbcs: bool
    pair_cutoff: np.ndarray
    state_cutoff: np.ndarray
    softcut_range: float
    h2ma: float = 0.0
    nucleon_mass: float = 0.0
    b0: float = 0.0
    b0p: float = 0.0
    b1: float = 0.0
    b1p: float = 0.0
    b2: float = 0.0
    b2p: float = 0.0
    b3: float = 0.0
    b3p: float = 0.0
    b4: float = 0.0
    slate: float = 0.0
    Crho0: float = 0.0
    Crho1: float = 0.0
    Crho0D: float = 0.0
    Crho1D: float = 0.0
    Cdrho0: float = 0.0
    Cdrho1: float = 0.0
    Ctau0: float = 0.0
    Ctau1: float = 0.0
    CdJ0: float = 0.0
    CdJ1: float = 0.0


def derive_couplings(f: Forces) -> Forces:
    """b- and C-couplings from t, x, power (forces.py:95-122, 141-142)."""
    t0, t1, t2, t3, t4 = f.t0, f.t1, f.t2, f.t3, f.t4
    x0, x1, x2, x3 = f.x0, f.x1, f.x2, f.x3
    f.b0 = t0 * (1 + 0.5 * x0)
    f.b0p = t0 * (0.5 + x0)
    f.b1 = (t1 + 0.5 * x1 * t1 + t2 + 0.5 * x2 * t2) / 4
    f.b1p = (t1 * (0.5 + x1) - t2 * (0.5 + x2)) / 4
    f.b2 = (3 * t1 * (1 + 0.5 * x1) - t2 * (1 + 0.5 * x2)) / 8
    f.b2p = (3 * t1 * (0.5 + x1) + t2 * (0.5 + x2)) / 8
    f.b3 = t3 * (1 + 0.5 * x3) / 4
    f.b3p = t3 * (0.5 + x3) / 4
    f.b4 = t4 / 2
    f.slate = (3 / math.pi) ** (1 / 3) * E2
    f.Crho0 = 0.5 * f.b0 - 0.25 * f.b0p
    f.Crho1 = -0.25 * f.b0p
    f.Crho0D = (1 / 3) * f.b3 - (1 / 6) * f.b3p
    f.Crho1D = -(1 / 6) * f.b3p
    f.Cdrho0 = -0.5 * f.b2 + 0.25 * f.b2p
    f.Cdrho1 = 0.25 * f.b2p
    f.Ctau0 = f.b1 - 0.5 * f.b1p
    f.Ctau1 = -0.5 * f.b1p
    f.CdJ0 = -f.b4 - 0.5 * f.b4p
    f.CdJ1 = -0.5 * f.b4p
    return f


def init_forces(name="SLy4", ipair=0, tbcs=False, v0prot=None, v0neut=None, rho0pr=None,
                pair_cutoff=(0.0, 0.0), state_cutoff=(0.0, 0.0), softcut_range=0.1,
                overrides=None) -> Forces:
    """forces.py:60-143.  ``overrides`` replaces t/x/power/b4p entries (adjoint tests)."""
    if name not in FORCE_TABLE:
        raise KeyError(f"Force '{name}' not found in the force table.")
    tab = dict(FORCE_TABLE[name])
    if overrides:
        tab.update(overrides)
    f = Forces(name=name, ex=tab["ex"], zpe=tab["zpe"], h2m=np.array(tab["h2m"], dtype=np.float64),
               t0=tab["t0"], t1=tab["t1"], t2=tab["t2"], t3=tab["t3"], t4=tab["t4"],
               x0=tab["x0"], x1=tab["x1"], x2=tab["x2"], x3=tab["x3"], b4p=tab["b4p"],
               power=tab["power"], v0prot=0.0, v0neut=0.0, rho0pr=0.16, ipair=ipair, tbcs=tbcs,
               pair_cutoff=np.array(pair_cutoff, dtype=np.float64),
               state_cutoff=np.array(state_cutoff, dtype=np.float64),
               softcut_range=softcut_range)
    derive_couplings(f)
    if ipair == 5:
        f.v0prot, f.v0neut, f.rho0pr = tab["vdi"]["v0prot"], tab["vdi"]["v0neut"], tab["vdi"]["rho0pr"]
    if ipair == 6:
        f.v0prot, f.v0neut, f.rho0pr = tab["dddi"]["v0prot"], tab["dddi"]["v0neut"], tab["dddi"]["rho0pr"]
    if v0prot is not None:
        f.v0prot = v0prot
    if v0neut is not None:
        f.v0neut = v0neut
    if rho0pr is not None:
        f.rho0pr = rho0pr
    f.h2ma = float(0.5 * np.sum(f.h2m))
    f.nucleon_mass = HBC ** 2 / (2.0 * f.h2ma)
    return f


# ---------------------------------------------------------------------------
# Grid                                                      grids.py:36-137
# ---------------------------------------------------------------------------
def sder(nmax: int, d: float) -> np.ndarray:
    """Dense Fourier first-derivative matrix (grids.py:36-46)."""
    icn = (nmax + 1) // 2
    afac = math.pi / icn
    i = np.arange(1, nmax + 1)
    j = np.arange(1, icn)[:, None, None]
    grid = i[None, :] - i[:, None]
    return -afac * (np.sum(-j * np.sin(j * afac * grid), axis=0)
                    - 0.5 * icn * np.sin(icn * afac * grid)) / (icn * d)


def sder2(nmax: int, d: float) -> np.ndarray:
    """Dense Fourier second-derivative matrix (grids.py:51-61)."""
    icn = (nmax + 1) // 2
    afac = math.pi / icn
    i = np.arange(1, nmax + 1)
    j = np.arange(1, icn)[:, None, None]
    grid = i[None, :] - i[:, None]
    return -(afac * afac) * (np.sum((j ** 2) * np.cos(j * afac * grid), axis=0)
                             + 0.5 * (icn ** 2) * np.cos(icn * afac * grid)) / (icn * d * d)


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


def init_grids(nx=48, ny=48, nz=48, dx=0.8, dy=0.8, dz=0.8, periodic=False) -> Grids:
    """grids.py:66-138 (Bloch/TABC options are not on this path)."""
    if nx % 2 or ny % 2 or nz % 2:
        raise ValueError("grids: nx, ny, and nz must be even.")
    if dx * dy * dz <= 0.0:
        if dx <= 0.0:
            raise ValueError("grid spacing given as 0.")
        dy = dz = dx

    def coord(n, d):
        return (np.arange(n, dtype=np.float64) - (n - 1) / 2) * d

    return Grids(nx, ny, nz, dx, dy, dz, periodic, dx * dy * dz,
                 coord(nx, dx), coord(ny, dy), coord(nz, dz),
                 sder(nx, dx), sder(ny, dy), sder(nz, dz),
                 sder2(nx, dx), sder2(ny, dy), sder2(nz, dz))


# Axis-wise dense-matrix application, the reference's einsum forms
# ('ij,jkl->ikl', 'jl,ilk->ijk', 'kl,ijl->ijk').
def apply_x(m, f):
    return np.einsum("ij,jkl->ikl", m, f, optimize=True)


def apply_y(m, f):
    return np.einsum("jl,ilk->ijk", m, f, optimize=True)


def apply_z(m, f):
    return np.einsum("kl,ijl->ijk", m, f, optimize=True)


# ---------------------------------------------------------------------------
# Spectral derivatives                                    levels.py:131-437
# ---------------------------------------------------------------------------
def _bcast(vec, axis, ndim):
    shape = [1] * ndim
    shape[axis] = vec.shape[0]
    return vec.reshape(shape)


def d1_factors(n, d):
    """Multipliers of the first derivative: i m kfac / n, Nyquist and m=0 zero.

    levels.py:134-146.  Index n-m gets -(i m kfac)/n for m=1..n/2-1; the
    scatter to index n (m=0) is dropped by JAX (out of bounds)."""
    kfac = 2.0 * math.pi / (d * n)
    fac = np.zeros(n, dtype=np.complex128)
    m = np.arange(1, n // 2)
    fac[m] = 1j * m * kfac / n
    fac[n - m] = -(1j * m * kfac) / n
    return fac


def d2_factors(n, d):
    """Second-derivative multipliers of cderv*01, Nyquist kept (levels.py:201-203)."""
    kfac = 2.0 * math.pi / (d * n)
    fac = np.zeros(n, dtype=np.float64)
    idx = np.arange(0, n // 2)
    fac[idx] = -(idx * kfac) ** 2 / n
    fac[n - 1 - idx] = -((idx + 1) * kfac) ** 2 / n
    return fac


def cderv00(d, psin, axis):
    """First derivative along ``axis`` (levels.py:131-189, cderv{x,y,z}00)."""
    n = psin.shape[axis]
    f = rfft_n(psin, axis)
    f = f * _bcast(d1_factors(n, d), axis, psin.ndim)
    return rifft_fwdnorm(f, axis)


def cderv01(d, psin, axis):
    """(d1, d2) along ``axis`` (levels.py:190-281, cderv{x,y,z}01)."""
    n = psin.shape[axis]
    f = rfft_n(psin, axis)
    d2 = rifft_fwdnorm(f * _bcast(d2_factors(n, d), axis, psin.ndim), axis)
    d1 = rifft_fwdnorm(f * _bcast(d1_factors(n, d), axis, psin.ndim), axis)
    return d1, d2


def cderv02(d, psin, pos_func, axis):
    """(d1, d(B d1)) along ``axis`` with the Nyquist rule (levels.py:283-395).

    ``psin`` has shape (..., nx, ny, nz) with the spatial axes last;
    ``pos_func`` (the effective mass B) has shape (nx, ny, nz).  The Nyquist
    entry of the second transform is -(pi/d)^2 * mean_axis(B) * psi_hat_N / n.
    """
    n = psin.shape[axis]
    nd = psin.ndim
    f = rfft_n(psin, axis)
    nyq = np.take(f, [n // 2], axis=axis)            # FFT(psi) at Nyquist, kept as a slab
    d1fac = _bcast(d1_factors(n, d), axis, nd)
    d1 = rifft_fwdnorm(f * d1fac, axis)
    g = rfft_n(d1 * pos_func, axis)
    g = g * d1fac
    sp_axis = (axis % nd) - nd + 3                   # axis of pos_func (3-D)
    bmean = np.mean(pos_func, axis=sp_axis, keepdims=True)
    sl = [slice(None)] * nd
    sl[axis] = slice(n // 2, n // 2 + 1)
    g[tuple(sl)] = -(math.pi / d) ** 2 * bmean * nyq / n
    d2 = rifft_fwdnorm(g, axis)
    return d1, d2


def k2_table(n, d):
    """-(m kfac)^2 for m = 0..n/2-1 then n/2..1 (levels.py:407-420)."""
    kfac = 2.0 * math.pi / (d * n)
    return np.concatenate((-(np.arange(0, n // 2) * kfac) ** 2, -(np.arange(n // 2, 0, -1) * kfac) ** 2))


def laplace(grids, h2ma, wg, wguv, psin, v_pairmax, e0inv):
    """FFT-diagonal preconditioner (levels.py:397-437).

    ``psin`` has shape (..., nx, ny, nz); ``wg``, ``wguv``, ``v_pairmax`` are
    scalars or arrays broadcasting over the leading axes."""
    nx, ny, nz = grids.nx, grids.ny, grids.nz
    weight = np.maximum(wg, 0.1)
    weightuv = np.maximum(wguv, 0.1)
    k2 = (k2_table(nx, grids.dx)[:, None, None] + k2_table(ny, grids.dy)[None, :, None]
          + k2_table(nz, grids.dz)[None, None, :])
    f = sfft.fftn(psin, axes=(-3, -2, -1), workers=WORKERS)
    den = (weight * (e0inv - h2ma * k2) + 0.5 * weightuv * v_pairmax) * (nx * ny * nz)
    return sfft.ifftn(f / den, axes=(-3, -2, -1), norm="forward", workers=WORKERS)


# ---------------------------------------------------------------------------
# Norms                                                        trivial.py:5-12
# ---------------------------------------------------------------------------
def rpsnorm(ps, wxyz):
    return wxyz * np.sum(np.real(np.conj(ps) * ps))


def overlap(pl, pr, wxyz):
    return wxyz * np.sum(np.conj(pl) * pr)


# ---------------------------------------------------------------------------
# Mean field                                              meanfield.py:7-391
# ---------------------------------------------------------------------------
@dataclass
class Meanfield:
    upot: np.ndarray
    bmass: np.ndarray
    divaq: np.ndarray
    v_pair: np.ndarray
    aq: np.ndarray
    spot: np.ndarray
    wlspot: np.ndarray
    dbmass: np.ndarray
    ecorrp: float = 0.0


def init_meanfield(grids) -> Meanfield:
    s4 = (2, grids.nx, grids.ny, grids.nz)
    s5 = (2, 3, grids.nx, grids.ny, grids.nz)
    z = np.zeros
    return Meanfield(z(s4), z(s4), z(s4), z(s4), z(s5), z(s5), z(s5), z(s5), 0.0)


def hpsi01(grids, meanfield, iq, weight, weightuv, pinn):
    """Single-particle Hamiltonian on two-spinors (meanfield.py:40-159).

    ``pinn`` is (..., 2, nx, ny, nz); the spin axis is -4.  ``weight`` and
    ``weightuv`` broadcast against ``pinn`` (scalars or (S,1,1,1,1)).
    Returns (w*h psi - wuv*v_pair psi, h psi, v_pair psi)."""
    U = meanfield.upot[iq]
    B = meanfield.bmass[iq]
    A = meanfield.aq[iq]
    S = meanfield.spot[iq]
    W = meanfield.wlspot[iq]
    p0 = pinn[..., 0, :, :, :]
    p1 = pinn[..., 1, :, :, :]
    s0, s1 = 0.5, -0.5
    # steps 1-2: U and spin-current coupling
    o0 = U * p0 + (S[0] - 1j * S[1]) * p1 + S[2] * p0
    o1 = U * p1 + (S[0] + 1j * S[1]) * p0 - S[2] * p1
    # step 3: x
    d1, d2 = cderv02(grids.dx, pinn, B, -3)
    o0 = o0 - (1j * (0.5 * A[0] - s0 * W[1])) * d1[..., 0, :, :, :] - s0 * W[2] * d1[..., 1, :, :, :] - d2[..., 0, :, :, :]
    o1 = o1 - (1j * (0.5 * A[0] - s1 * W[1])) * d1[..., 1, :, :, :] - s1 * W[2] * d1[..., 0, :, :, :] - d2[..., 1, :, :, :]
    q0 = (-0.5j) * (A[0] - W[1]) * p0 - 0.5 * W[2] * p1
    q1 = (-0.5j) * (A[0] + W[1]) * p1 + 0.5 * W[2] * p0
    d = cderv00(grids.dx, np.stack([q0, q1], axis=-4), -3)
    o0 = o0 + d[..., 0, :, :, :]
    o1 = o1 + d[..., 1, :, :, :]
    # step 4: y
    d1, d2 = cderv02(grids.dy, pinn, B, -2)
    o0 = o0 - (1j * (0.5 * A[1] + s0 * W[0])) * d1[..., 0, :, :, :] + (1j * 0.5 * W[2]) * d1[..., 1, :, :, :] - d2[..., 0, :, :, :]
    o1 = o1 - (1j * (0.5 * A[1] + s1 * W[0])) * d1[..., 1, :, :, :] + (1j * 0.5 * W[2]) * d1[..., 0, :, :, :] - d2[..., 1, :, :, :]
    q0 = (-0.5j) * (A[1] + W[0]) * p0 + (0.5j) * W[2] * p1
    q1 = (-0.5j) * (A[1] - W[0]) * p1 + (0.5j) * W[2] * p0
    d = cderv00(grids.dy, np.stack([q0, q1], axis=-4), -2)
    o0 = o0 + d[..., 0, :, :, :]
    o1 = o1 + d[..., 1, :, :, :]
    # step 5: z
    d1, d2 = cderv02(grids.dz, pinn, B, -1)
    o0 = o0 - (1j * 0.5 * A[2]) * d1[..., 0, :, :, :] + (s0 * W[0] - 1j * 0.5 * W[1]) * d1[..., 1, :, :, :] - d2[..., 0, :, :, :]
    o1 = o1 - (1j * 0.5 * A[2]) * d1[..., 1, :, :, :] + (s1 * W[0] - 1j * 0.5 * W[1]) * d1[..., 0, :, :, :] - d2[..., 1, :, :, :]
    q0 = (-0.5j) * A[2] * p0 + (0.5 * W[0] - 0.5j * W[1]) * p1
    q1 = (-0.5j) * A[2] * p1 + (-0.5 * W[0] - 0.5j * W[1]) * p0
    d = cderv00(grids.dz, np.stack([q0, q1], axis=-4), -1)
    o0 = o0 + d[..., 0, :, :, :]
    o1 = o1 + d[..., 1, :, :, :]
    pout_mf = np.stack([o0, o1], axis=-4)
    pout_del = pinn * meanfield.v_pair[iq]
    pout = weight * pout_mf - weightuv * pout_del
    return pout, pout_mf, pout_del


# ---------------------------------------------------------------------------
# Coulomb                                                    coulomb.py:5-75
# ---------------------------------------------------------------------------
@dataclass
class Coulomb:
    nx2: int
    ny2: int
    nz2: int
    wcoul: np.ndarray
    q: np.ndarray


def init_iq(n, d, periodic):
    """coulomb.py:14-21."""
    idx = np.arange(n)
    idx = np.where(idx <= n // 2, idx, idx - n)
    if periodic:
        return (2.0 * math.pi * idx / (n * d)) ** 2
    return (d * idx) ** 2


def init_coulomb(grids) -> Coulomb:
    """Green's function on the (2n)^3 doubled grid, FFT'd once (coulomb.py:23-52)."""
    nx2 = grids.nx if grids.periodic else 2 * grids.nx
    ny2 = grids.ny if grids.periodic else 2 * grids.ny
    nz2 = grids.nz if grids.periodic else 2 * grids.nz
    q = (init_iq(nx2, grids.dx, grids.periodic)[:, None, None] + init_iq(ny2, grids.dy, grids.periodic)[None, :, None]
         + init_iq(nz2, grids.dz, grids.periodic)[None, None, :]).astype(np.complex128)
    q[0, 0, 0] = 1.0
    if grids.periodic:
        q = 1.0 / np.real(q) + 0j
        q[0, 0, 0] = 0.0
    else:
        q = 1.0 / np.sqrt(np.real(q)) + 0j
        q[0, 0, 0] = 2.84 / (grids.dx * grids.dy * grids.dz) ** (1.0 / 3.0)
        q = sfft.fftn(q, workers=WORKERS)
    return Coulomb(nx2, ny2, nz2, np.zeros((grids.nx, grids.ny, grids.nz)), q)


def poisson(grids, coulomb, rho):
    """Coulomb potential of the proton density (coulomb.py:54-75)."""
    rho2 = np.zeros((coulomb.nx2, coulomb.ny2, coulomb.nz2), dtype=np.complex128)
    rho2[:grids.nx, :grids.ny, :grids.nz] = rho[1]
    rho2 = sfft.fftn(rho2, workers=WORKERS)
    if grids.periodic:
        rho2 = rho2 * (4.0 * math.pi * E2 * np.real(coulomb.q))
    else:
        rho2 = rho2 * (E2 * grids.wxyz * coulomb.q)
    w = sfft.ifftn(rho2, workers=WORKERS)
    return np.real(w[:grids.nx, :grids.ny, :grids.nz])


# ---------------------------------------------------------------------------
# Densities                                               densities.py:6-337
# ---------------------------------------------------------------------------
@dataclass
class Densities:
    rho: np.ndarray
    chi: np.ndarray
    tau: np.ndarray
    current: np.ndarray
    sdens: np.ndarray
    sodens: np.ndarray


def init_densities(grids) -> Densities:
    s4 = (2, grids.nx, grids.ny, grids.nz)
    s5 = (2, 3, grids.nx, grids.ny, grids.nz)
    return Densities(np.zeros(s4), np.zeros(s4), np.zeros(s4), np.zeros(s5), np.zeros(s5), np.zeros(s5))


def density_contributions(grids, weight, weightuv, psin):
    """Per-state contributions (densities.py:237-330), vectorised over states.

    ``psin`` (S, 2, nx, ny, nz); ``weight``, ``weightuv`` (S,).
    Returns summed (rho, chi, tau, current(3), sdens(3), sodens(3))."""
    w = weight[:, None, None, None]
    wuv = weightuv[:, None, None, None]
    p0 = psin[:, 0]
    p1 = psin[:, 1]
    c0, c1 = np.conj(p0), np.conj(p1)
    dens = np.real(p0 * c0 + p1 * c1)
    rho = np.sum(w * dens, axis=0)
    chi = np.sum(0.5 * wuv * dens, axis=0)
    sd = np.stack([np.sum(2.0 * w * np.real(c0 * p1), axis=0),
                   np.sum(2.0 * w * np.imag(c0 * p1), axis=0),
                   np.sum(w * np.real(c0 * p0 - c1 * p1), axis=0)])
    dx = cderv00(grids.dx, psin, -3)
    dy = cderv00(grids.dy, psin, -2)
    dz = cderv00(grids.dz, psin, -1)
    x0, x1 = dx[:, 0], dx[:, 1]
    y0, y1 = dy[:, 0], dy[:, 1]
    z0, z1 = dz[:, 0], dz[:, 1]
    tau = np.sum(w * (np.real(x0 * np.conj(x0) + x1 * np.conj(x1))
                      + np.real(y0 * np.conj(y0) + y1 * np.conj(y1))
                      + np.real(z0 * np.conj(z0) + z1 * np.conj(z1))), axis=0)
    cur = np.stack([np.sum(w * np.imag(x0 * c0 + x1 * c1), axis=0),
                    np.sum(w * np.imag(y0 * c0 + y1 * c1), axis=0),
                    np.sum(w * np.imag(z0 * c0 + z1 * c1), axis=0)])
    sodens_y_x = -w * np.imag(x0 * c0 - x1 * c1)
    sodens_z_x = -w * np.real(p0 * np.conj(x1) - p1 * np.conj(x0))
    sodens_x_y = w * np.imag(y0 * c0 - y1 * c1)
    sodens_z_y = -w * np.imag(y1 * c0 + y0 * c1)
    sodens_x_z = w * np.real(z1 * c0 - z0 * c1)
    sodens_y_z = w * np.imag(z1 * c0 + z0 * c1)
    so = np.stack([np.sum(sodens_x_y + sodens_x_z, axis=0),
                   np.sum(sodens_y_x + sodens_y_z, axis=0),
                   np.sum(sodens_z_x + sodens_z_y, axis=0)])
    return rho, chi, tau, cur, sd, so


def add_density(densities, grids, levels):
    """Sum of per-state densities per isospin (densities.py:177-234)."""
    weights = levels.wocc * levels.wstates
    weightuvs = levels.wguv * levels.pairwg * levels.wstates
    out = init_densities(grids)
    for iq in range(2):
        sl = levels.block(iq)
        rho, chi, tau, cur, sd, so = density_contributions(grids, weights[sl], weightuvs[sl], levels.psi[sl])
        out.rho[iq], out.chi[iq], out.tau[iq] = rho, chi, tau
        out.current[iq], out.sdens[iq], out.sodens[iq] = cur, sd, so
    densities.rho, densities.chi, densities.tau = out.rho, out.chi, out.tau
    densities.current, densities.sdens, densities.sodens = out.current, out.sdens, out.sodens
    return densities


# ---------------------------------------------------------------------------
# Skyrme potentials                                      meanfield.py:163-391
# ---------------------------------------------------------------------------
def div3(grids, v):
    return apply_x(grids.der1x, v[0]) + apply_y(grids.der1y, v[1]) + apply_z(grids.der1z, v[2])


def grad3(grids, f):
    return np.stack([apply_x(grids.der1x, f), apply_y(grids.der1y, f), apply_z(grids.der1z, f)])


def curl3(grids, v):
    return np.stack([apply_y(grids.der1y, v[2]) - apply_z(grids.der1z, v[1]),
                     apply_z(grids.der1z, v[0]) - apply_x(grids.der1x, v[2]),
                     apply_x(grids.der1x, v[1]) - apply_y(grids.der1y, v[0])])


def lap3(grids, f):
    return apply_x(grids.der2x, f) + apply_y(grids.der2y, f) + apply_z(grids.der2z, f)


def skyrme(coulomb, densities, forces, grids, meanfield, tcoul=True, outertype="N", constraints=None):
    """Mean-field potentials from the densities (meanfield.py:163-391); with
    active constraints U_q -= sum_i lambda_i Q_i (meanfield.py:246-248,
    constraint.py:199-213, the same field for both isospins)."""
    f = forces
    rho = densities.rho
    eps = 1.0e-25
    upot = np.zeros_like(rho)
    rho_tot = rho[0] + rho[1]
    rho_tot_pow = rho_tot ** f.power
    for iq in range(2):
        ic = 1 - iq
        upot[iq] = rho_tot_pow * ((f.b3 * (f.power + 2.0) / 3.0 - 2.0 * f.b3p / 3.0) * rho[iq]
                                  + f.b3 * (f.power + 2.0) / 3.0 * rho[ic]
                                  - (f.b3p * f.power / 3.0) * (rho[0] ** 2 + rho[1] ** 2) / (rho_tot + eps))
    divj = np.stack([div3(grids, densities.sodens[iq]) for iq in range(2)])
    for iq in range(2):
        ic = 1 - iq
        upot[iq] += -(f.b4 + f.b4p) * divj[iq] - f.b4 * divj[ic]
    if tcoul:
        coulomb.wcoul = poisson(grids, coulomb, rho)
        upot[1] += coulomb.wcoul
        if f.ex != 0:
            upot[1] += -f.slate * rho[1] ** (1.0 / 3.0)
    lap = np.stack([lap3(grids, rho[iq]) for iq in range(2)])
    for iq in range(2):
        ic = 1 - iq
        upot[iq] += ((f.b0 - f.b0p) * rho[iq] + f.b0 * rho[ic]
                     + (f.b1 - f.b1p) * densities.tau[iq] + f.b1 * densities.tau[ic]
                     - (f.b2 - f.b2p) * lap[iq] - f.b2 * lap[ic])
    bmass = np.stack([f.h2m[iq] + (f.b1 - f.b1p) * rho[iq] + f.b1 * rho[1 - iq] for iq in range(2)])
    grad = np.stack([grad3(grids, rho[iq]) for iq in range(2)])
    wl = np.stack([(f.b4 + f.b4p) * grad[iq] + f.b4 * grad[1 - iq] for iq in range(2)])
    curls = np.stack([curl3(grids, densities.sdens[iq]) for iq in range(2)])
    aq = np.stack([-2.0 * (f.b1 - f.b1p) * densities.current[iq] - 2.0 * f.b1 * densities.current[1 - iq]
                   - (f.b4 + f.b4p) * curls[iq] - f.b4 * curls[1 - iq] for iq in range(2)])
    curlj = np.stack([curl3(grids, densities.current[iq]) for iq in range(2)])
    spot = np.stack([-(f.b4 + f.b4p) * curlj[iq] - f.b4 * curlj[1 - iq] for iq in range(2)])
    divaq = np.stack([div3(grids, aq[iq]) for iq in range(2)])
    dbmass = np.stack([grad3(grids, bmass[iq]) for iq in range(2)])
    if f.ipair == 6:
        rearr = (f.v0neut / f.rho0pr) * densities.chi[0] ** 2 + (f.v0prot / f.rho0pr) * densities.chi[1] ** 2
        upot[0] += rearr
        upot[1] += rearr
        meanfield.ecorrp = -np.sum((rho[0] + rho[1]) * rearr) * grids.wxyz / 2.0
    else:
        meanfield.ecorrp = 0.0
    v_pair = np.zeros_like(rho)
    for iq in range(2):
        v0act = f.v0prot if iq == 1 else f.v0neut
        if f.ipair == 6:
            v_pair[iq] = v0act * densities.chi[iq] * (1.0 - (rho[0] + rho[1]) / f.rho0pr)
        else:
            v_pair[iq] = v0act * densities.chi[iq]
    if outertype != "N":
        raise NotImplementedError(f"outertype '{outertype}' not implemented.")
    if constraints is not None and constraints.tconstraint:
        cpot = np.tensordot(constraints.lambda_crank, constraints.constr_field, axes=(0, 0))
        upot = upot - cpot[None]
    meanfield.upot, meanfield.bmass, meanfield.wlspot = upot, bmass, wl
    meanfield.aq, meanfield.spot, meanfield.divaq, meanfield.dbmass = aq, spot, divaq, dbmass
    meanfield.v_pair = v_pair
    return meanfield, coulomb


# ---------------------------------------------------------------------------
# Levels                                                      levels.py:6-130
# ---------------------------------------------------------------------------
@dataclass
class Levels:
    nneut: int
    nprot: int
    npsi: np.ndarray
    psi: np.ndarray
    hampsi: np.ndarray
    lagrange: np.ndarray
    hmfpsi: np.ndarray
    delpsi: np.ndarray
    isospin: np.ndarray
    sp_energy: np.ndarray
    sp_efluct1: np.ndarray
    sp_efluct2: np.ndarray
    sp_kinetic: np.ndarray
    sp_norm: np.ndarray
    sp_parity: np.ndarray
    sp_orbital: np.ndarray
    sp_spin: np.ndarray
    wocc: np.ndarray
    wguv: np.ndarray
    pairwg: np.ndarray
    wstates: np.ndarray
    deltaf: np.ndarray

    @property
    def nstmax(self):
        return int(self.npsi[0] + self.npsi[1])

    @property
    def mass_number(self):
        return self.nneut + self.nprot

    @property
    def charge_number(self):
        return self.nprot

    def block(self, iq):
        n0 = int(self.npsi[0])
        return slice(0, n0) if iq == 0 else slice(n0, n0 + int(self.npsi[1]))


def init_levels(grids, nneut=82, nprot=50, npsi=None) -> Levels:
    """levels.py:37-130."""
    if npsi is None:
        npsi = [max(126, nneut + int(1.65 * nneut ** 0.666667)), max(82, nprot + int(1.65 * nprot ** 0.666667))]
    npsi = np.array(npsi, dtype=np.int64)
    S = int(npsi[0] + npsi[1])
    s5 = (S, 2, grids.nx, grids.ny, grids.nz)
    isospin = np.zeros(S, dtype=np.int32)
    isospin[int(npsi[0]):] = 1
    wocc = np.zeros(S)
    wocc[:nneut] = 1.0
    wocc[int(npsi[0]):int(npsi[0]) + nprot] = 1.0
    z = lambda: np.zeros(S)
    return Levels(nneut, nprot, npsi, np.zeros(s5, np.complex128), np.zeros(s5, np.complex128),
                  np.zeros(s5, np.complex128), np.zeros(s5, np.complex128), np.zeros(s5, np.complex128),
                  isospin, z(), z(), z(), z(), z(), z(), np.zeros((S, 3)), np.zeros((S, 3)),
                  wocc, z(), np.ones(S), np.ones(S), z())


# ---------------------------------------------------------------------------
# Static solver state                                         static.py:19-90
# ---------------------------------------------------------------------------
@dataclass
class Static:
    tdiag: bool = False
    tvaryx_0: bool = False
    tsort: bool = False
    maxiter: int = 200
    iternat: int = 100
    iternat_start: int = 40
    delstepbas: float = 2.0
    e0bas: float = 10.0
    iteranneal: int = 0
    pairenhance: float = 0.0
    inibcs: int = 30
    inidiag: int = 30
    radinx: float = 3.1
    radiny: float = 3.1
    radinz: float = 3.1
    serr: float = 1.0e-20
    delesum: float = 0.0
    sumflu: float = 0.0
    x0dmp: float = 0.45
    e0dmp: float = 100.0
    x0dmpmin: float = 0.45
    outertype: str = "N"
    hmatrix: np.ndarray = None
    gapmatrix: np.ndarray = None
    symcond: np.ndarray = None
    lambda_save: np.ndarray = None


@dataclass
class Energies:
    ehft: float = 0.0
    ehf0: float = 0.0
    ehf1: float = 0.0
    ehf2: float = 0.0
    ehf3: float = 0.0
    ehfls: float = 0.0
    ehflsodd: float = 0.0
    ehfc: float = 0.0
    ecorc: float = 0.0
    ehfint: float = 0.0
    efluct1: np.ndarray = field(default_factory=lambda: np.zeros(1))
    efluct1q: np.ndarray = field(default_factory=lambda: np.zeros(2))
    efluct1prev: float = 0.0
    efluct2: np.ndarray = field(default_factory=lambda: np.zeros(1))
    efluct2q: np.ndarray = field(default_factory=lambda: np.zeros(2))
    efluct2prev: float = 0.0
    tke: float = 0.0
    ehf: float = 0.0
    ehfprev: float = 0.0
    e3corr: float = 0.0
    e_zpe: float = 0.0
    orbital: np.ndarray = field(default_factory=lambda: np.zeros(3))
    spin: np.ndarray = field(default_factory=lambda: np.zeros(3))
    total_angmom: np.ndarray = field(default_factory=lambda: np.zeros(3))
    ehfCrho0: float = 0.0
    ehfCrho1: float = 0.0
    ehfCdrho0: float = 0.0
    ehfCdrho1: float = 0.0
    ehfCtau0: float = 0.0
    ehfCtau1: float = 0.0
    ehfCdJ0: float = 0.0
    ehfCdJ1: float = 0.0
    ehfCj0: float = 0.0
    ehfCj1: float = 0.0


@dataclass
class Pairs:
    eferm: np.ndarray = field(default_factory=lambda: np.zeros(2))
    epair: np.ndarray = field(default_factory=lambda: np.zeros(2))
    avdelt: np.ndarray = field(default_factory=lambda: np.zeros(2))
    avdeltv2: np.ndarray = field(default_factory=lambda: np.zeros(2))
    avg: np.ndarray = field(default_factory=lambda: np.zeros(2))
    firstcall: bool = True


@dataclass
class Params:
    iteration: int = 0
    mprint: int = 10
    tcoul: bool = True


def init_static(forces, levels, **kwargs):
    """Damping defaults and c.m. scaling of h2m (static.py:53-90)."""
    nst = int(max(levels.npsi[0], levels.npsi[1]))
    st = Static(**kwargs)
    st.x0dmpmin = st.x0dmp
    st.hmatrix = np.zeros((2, nst, nst), np.complex128)
    st.gapmatrix = np.zeros((2, nst, nst), np.complex128)
    st.symcond = np.zeros((2, nst, nst), np.complex128)
    st.lambda_save = np.zeros((2, nst, nst), np.complex128)
    if forces.zpe == 0:
        forces.h2m = forces.h2m * ((levels.mass_number - 1.0) / levels.mass_number)
    return forces, st


# ---------------------------------------------------------------------------
# Gradient step                                              static.py:93-259
# ---------------------------------------------------------------------------
def grstep(forces, grids, levels, meanfield, params, static):
    """Damped gradient step on every state (static.py:93-259), per isospin block."""
    psi_new = np.empty_like(levels.psi)
    hampsi = np.empty_like(levels.psi)
    hmfpsi = np.empty_like(levels.psi)
    delpsi = np.empty_like(levels.psi)
    spe_new_all = np.empty(levels.nstmax)
    efl1_all = levels.sp_efluct1.copy()
    efl2_all = levels.sp_efluct2.copy()
    denerg_all = np.empty(levels.nstmax)
    b5 = lambda v: v[:, None, None, None, None]
    for iq in range(2):
        sl = levels.block(iq)
        psin = levels.psi[sl]
        lag = levels.lagrange[sl]
        if forces.tbcs:
            w1, wuv1 = 1.0, 0.0
        else:
            w1 = b5(levels.wstates[sl] * levels.wocc[sl])
            wuv1 = b5(levels.wstates[sl] * levels.wguv[sl] * levels.pairwg[sl])
        ps1, psi_mf, _ = hpsi01(grids, meanfield, iq, w1, wuv1, psin)
        spe_new = grids.wxyz * np.real(np.sum(np.conj(psin) * psi_mf, axis=(1, 2, 3, 4)))
        spe_old = levels.sp_energy[sl]
        if static.e0dmp > 0.0:
            h_exp = grids.wxyz * np.real(np.sum(np.conj(psin) * ps1, axis=(1, 2, 3, 4)))
            if (params.iteration > 1) and (not forces.tbcs):
                ps1o = ps1 - lag
            else:
                ps1o = ps1 - b5(h_exp) * psin
            if params.mprint > 0 and params.iteration % params.mprint == 0:
                efl1_all[sl] = np.sqrt(grids.wxyz * np.sum(np.abs(ps1o) ** 2, axis=(1, 2, 3, 4)))
                efl2_all[sl] = np.sqrt(grids.wxyz * np.sum(np.abs(lag - b5(h_exp) * psin) ** 2, axis=(1, 2, 3, 4)))
            if forces.tbcs:
                ps2 = laplace(grids, forces.h2ma, 1.0, 0.0, ps1o, 0.0, static.e0dmp)
            else:
                ps2 = laplace(grids, forces.h2ma, b5(levels.wocc[sl]), b5(levels.wguv[sl] * levels.pairwg[sl]),
                              ps1o, np.max(meanfield.v_pair[iq]), static.e0dmp)
            pnew = psin - static.x0dmp * ps2
        else:
            pnew = b5(1.0 + static.x0dmp * (spe_new - spe_old)) * psin - static.x0dmp * ps1
        _, hmf, dl = hpsi01(grids, meanfield, iq, w1, wuv1, pnew)
        psi_new[sl], hampsi[sl], hmfpsi[sl], delpsi[sl] = pnew, psi_mf, hmf, dl
        spe_new_all[sl] = spe_new
        denerg_all[sl] = (spe_old - spe_new) / np.abs(spe_new)
    levels.psi, levels.hampsi, levels.hmfpsi, levels.delpsi = psi_new, hampsi, hmfpsi, delpsi
    levels.sp_energy = spe_new_all
    levels.sp_efluct1, levels.sp_efluct2 = efl1_all, efl2_all
    static.sumflu = static.sumflu + np.sum(levels.sp_efluct1)
    static.delesum = static.delesum + np.sum(levels.wocc * levels.wstates * denerg_all)
    return levels, static


def e0dmp_gt_zero(args):
    """Preconditioned damped step of one state (static.py:93-146)."""
    nst, iq, psin, ps1, lagrange, forces, grids, levels, meanfield, params, static = args
    h_exp = np.real(overlap(psin, ps1, grids.wxyz))
    if (params.iteration > 1) and (not forces.tbcs):
        ps1o = ps1 - lagrange
    else:
        ps1o = ps1 - h_exp * psin
    if params.mprint > 0 and params.iteration % params.mprint == 0:
        efl1 = np.sqrt(rpsnorm(ps1o, grids.wxyz))
        efl2 = np.sqrt(rpsnorm(lagrange - h_exp * psin, grids.wxyz))
    else:
        efl1, efl2 = levels.sp_efluct1[nst], levels.sp_efluct2[nst]
    if forces.tbcs:
        ps2 = laplace(grids, forces.h2ma, 1.0, 0.0, ps1o, 0.0, static.e0dmp)
    else:
        ps2 = laplace(grids, forces.h2ma, levels.wocc[nst], levels.wguv[nst] * levels.pairwg[nst], ps1o,
                      np.max(meanfield.v_pair[iq]), static.e0dmp)
    return psin - static.x0dmp * ps2, efl1, efl2


def grstep_helper(nst, iq, spe_mf, psin, lagrange, forces, grids, levels, meanfield, params, static):
    """Gradient step of one state (static.py:148-213); returns (psin, psi_mf,
    hmfpsi, delpsi, denerg, spe_mf_new, sp_efluct1, sp_efluct2)."""
    if forces.tbcs:
        w, wuv = 1.0, 0.0
    else:
        w = levels.wstates[nst] * levels.wocc[nst]
        wuv = levels.wstates[nst] * levels.wguv[nst] * levels.pairwg[nst]
    ps1, psi_mf, _ = hpsi01(grids, meanfield, iq, w, wuv, psin)
    spe_new = np.real(overlap(psin, psi_mf, grids.wxyz))
    if static.e0dmp > 0.0:
        pnew, efl1, efl2 = e0dmp_gt_zero((nst, iq, psin, ps1, lagrange, forces, grids, levels, meanfield, params,
                                          static))
    else:
        pnew = (1.0 + static.x0dmp * (spe_new - spe_mf)) * psin - static.x0dmp * ps1
        efl1, efl2 = levels.sp_efluct1[nst], levels.sp_efluct2[nst]
    _, hmf, dl = hpsi01(grids, meanfield, iq, w, wuv, pnew)
    denerg = (spe_mf - spe_new) / np.abs(spe_new)
    return pnew, psi_mf, hmf, dl, denerg, spe_new, efl1, efl2


# ---------------------------------------------------------------------------
# Orthonormalisation / subspace                             static.py:261-577
# ---------------------------------------------------------------------------
def as2d(block):
    """(nst, 2, nx, ny, nz) -> (2N, nst).  Row order differs from the
    reference's Fortran-order reshape; every use is an inner product over rows."""
    return block.reshape(block.shape[0], -1).T


def from2d(mat, shape):
    return mat.T.reshape(shape)


def loewdin_orthonormalize(psi_2d, wxyz, nst):
    """S^{-1/2} with eigenvalue clamp (static.py:463-482)."""
    rhomatr = (np.conj(psi_2d.T) @ psi_2d) * wxyz
    sp_norm = np.real(np.diagonal(rhomatr)).copy()
    w, v = np.linalg.eigh(rhomatr, UPLO="L")
    thr = max(1e-12, np.finfo(np.float64).eps * np.max(w) * nst)
    w_safe = np.maximum(w, thr)
    v_scaled = v * (1.0 / np.sqrt(w_safe))[None, :]
    return v @ np.conj(v_scaled.T), sp_norm


def diagstep(energies, forces, grids, levels, static, diagonalize=False, construct=True):
    """Loewdin + optional subspace diagonalisation + HFB matrices (static.py:261-461)."""
    for iq in range(2):
        sl = levels.block(iq)
        nst = sl.stop - sl.start
        shape = levels.psi[sl].shape
        psi_2d = as2d(levels.psi[sl]).copy()      # the pre-rotation basis, kept for H_old
        unitary, sp_norm = loewdin_orthonormalize(psi_2d, grids.wxyz, nst)
        levels.sp_norm[sl] = sp_norm
        if diagonalize:
            lam_lin = (np.conj(psi_2d.T) @ as2d(levels.hampsi[sl])) * grids.wxyz
            _, u_lam = np.linalg.eigh(lam_lin, UPLO="L")
            unitary = unitary @ u_lam
        levels.psi[sl] = from2d(psi_2d @ unitary, shape)
        if construct:
            construct_hfb_matrices(unitary, psi_2d, energies, forces, grids, levels, static, iq, sl, nst)
        else:
            lag = as2d(levels.psi[sl]) @ static.lambda_save[iq, :nst, :nst]
            levels.lagrange[sl] = from2d(lag, shape)
    return energies, levels, static


def construct_hfb_matrices(unitary, psi_2d, energies, forces, grids, levels, static, iq, sl, nst):
    """static.py:484-577."""
    h_old = (np.conj(psi_2d.T) @ as2d(levels.hmfpsi[sl])) * grids.wxyz
    h = np.conj(unitary.T) @ (h_old @ unitary)
    g = np.zeros_like(h)
    if forces.ipair != 0:
        g_old = (np.conj(psi_2d.T) @ as2d(levels.delpsi[sl])) * grids.wxyz
        g = np.conj(unitary.T) @ (g_old @ unitary)
    static.hmatrix[iq, :nst, :nst] = h
    static.gapmatrix[iq, :nst, :nst] = g
    levels.sp_energy[sl] = np.real(np.diagonal(h))
    if forces.ipair != 0:
        levels.deltaf[sl] = np.real(np.diagonal(g)) * levels.pairwg[sl]
    if forces.tbcs:
        static.gapmatrix[iq, :nst, :nst] = g * np.eye(nst)
    weight = levels.wocc[sl] * levels.wstates[sl]
    weightuv = levels.wguv[sl] * levels.pairwg[sl] * levels.wstates[sl]
    lam = weight[None, :] * static.hmatrix[iq, :nst, :nst] - weightuv[None, :] * static.gapmatrix[iq, :nst, :nst]
    lam_asym = 0.5 * (lam - np.conj(lam.T))
    lam_sym = 0.5 * (lam + np.conj(lam.T))
    energies.efluct1q[iq] = np.max(np.abs(lam_asym))
    energies.efluct2q[iq] = np.sqrt(np.sum(np.abs(lam_asym) ** 2) / nst ** 2)
    static.symcond[iq, :nst, :nst] = lam_asym
    static.lambda_save[iq, :nst, :nst] = lam_sym
    shape = levels.psi[sl].shape
    levels.lagrange[sl] = from2d(as2d(levels.psi[sl]) @ lam_sym, shape)
    energies.efluct1[0] = np.max(energies.efluct1q)
    energies.efluct2[0] = np.mean(energies.efluct2q)


# ---------------------------------------------------------------------------
# BCS pairing                                                  pairs.py:25-220
# ---------------------------------------------------------------------------
def bcs_particle_number(eferm, sp_energy, deltaf, wstates):
    """pairs.py:25-45."""
    edif = sp_energy - eferm
    equasi = np.maximum(np.sqrt(edif ** 2 + deltaf ** 2), 1.0e-20)
    v2 = np.clip(0.5 * (1.0 - edif / equasi), 1.0e-10, 1.0 - 1.0e-10)
    return np.sum(v2 * wstates)


def find_eferm(target, sp_energy, deltaf, wstates):
    """brentq on [-100, 100], xtol=1e-14 (pairs.py:47-61)."""
    return brentq(lambda e: float(bcs_particle_number(e, sp_energy, deltaf, wstates)) - target,
                  a=-100.0, b=100.0, xtol=1e-14)


def soft_cutoff(sp_energy, ecut, cutwid):
    return 1.0 / (1.0 + np.exp((sp_energy - ecut) / cutwid))


def pairdn(iq, particle_number, levels, forces, pairs):
    """pairs.py:66-131."""
    mask = levels.isospin == iq
    e = levels.sp_energy[mask]
    dl = levels.deltaf[mask]
    ws = levels.wstates[mask]
    ef = find_eferm(particle_number, e, dl, ws)
    pairwg_new = levels.pairwg.copy()
    wstates_new = levels.wstates.copy()
    ws_calc = levels.wstates[mask]
    if forces.pair_cutoff[iq] > 0.0:
        pairwg_new[mask] = soft_cutoff(e, ef + forces.pair_cutoff[iq], forces.softcut_range * forces.pair_cutoff[iq])
    if forces.state_cutoff[iq] > 0.0:
        sc = soft_cutoff(e, ef + forces.state_cutoff[iq], forces.softcut_range * forces.state_cutoff[iq])
        wstates_new[mask] = sc
        ws_calc = sc
    edif = e - ef
    equasi = np.maximum(np.sqrt(edif ** 2 + dl ** 2), 1.0e-20)
    v2 = 0.5 * (1.0 - edif / equasi)
    uv = np.sqrt(np.maximum(v2 - v2 ** 2, 1.0e-6))
    wocc = np.clip(v2, 1e-10, 1.0 - 1e-10)
    vol = 0.5 * uv * ws_calc
    sumuv = np.sum(vol)
    sumduv = np.sum(vol * dl)
    sumv2 = np.sum(wocc * ws_calc)
    sumdv2 = np.sum(dl * wocc * ws_calc)
    sumuv_s = max(sumuv, 1.0e-20)
    sumv2_s = max(sumv2, 1.0e-20)
    pairs.eferm[iq] = ef
    pairs.epair[iq] = sumduv
    pairs.avdelt[iq] = sumduv / sumuv_s
    pairs.avdeltv2[iq] = sumdv2 / sumv2_s
    pairs.avg[iq] = sumduv / sumuv_s ** 2
    return wocc, uv, pairwg_new, wstates_new


def pairgap(levels, meanfield, params, grids, firstcall):
    """pairs.py:133-157."""
    if params.iteration <= 10:
        return 11.2 / np.sqrt(levels.mass_number) * np.ones_like(levels.deltaf)
    if not firstcall:
        return levels.deltaf
    dens = np.sum(np.real(levels.psi * np.conj(levels.psi)), axis=1)
    vp = np.where(levels.isospin[:, None, None, None] == 0, meanfield.v_pair[0][None], meanfield.v_pair[1][None])
    return np.sum(vp * dens, axis=(1, 2, 3)) * grids.wxyz * levels.pairwg


def pair(levels, meanfield, forces, params, grids, pairs):
    """pairs.py:161-220 (the summary print is omitted)."""
    levels.deltaf = pairgap(levels, meanfield, params, grids, pairs.firstcall)
    pairs.firstcall = False
    for iq, pn in ((0, float(levels.nneut)), (1, float(levels.nprot))):
        wocc, wguv, pairwg_full, wstates_full = pairdn(iq, pn, levels, forces, pairs)
        mask = levels.isospin == iq
        levels.wocc = levels.wocc.copy()
        levels.wguv = levels.wguv.copy()
        levels.wocc[mask] = wocc
        levels.wguv[mask] = wguv
        if forces.pair_cutoff[iq] > 0.0:
            levels.pairwg = pairwg_full
        if forces.state_cutoff[iq] > 0.0:
            levels.wstates = wstates_full
    return levels, pairs


# ---------------------------------------------------------------------------
# HFB sub-iterations in configuration space                 static.py:920-1336
# ---------------------------------------------------------------------------
def lambda_matrix(hmat, gapmat, levels, sl):
    """Hermitian part of w_j H_ij - wuv_j G_ij (static.py:1036-1053)."""
    w = levels.wocc[sl] * levels.wstates[sl]
    wuv = levels.wguv[sl] * levels.pairwg[sl] * levels.wstates[sl]
    lt = w[None, :] * hmat - wuv[None, :] * gapmat
    return 0.5 * (lt + np.conj(lt.T))


def pair_subiter(levels, forces, pairs, iq):
    """BCS of one isospin inside the sub-iterations (static.py:1170-1336).
    The reference calls brentq with rtol=1e-16, below SciPy's minimum of
    4 eps, so brentq raises and the reference always takes its fallback:
    the midpoint of levels N-1 and N of the block (in block order)."""
    sl = levels.block(iq)
    pn = float(levels.nneut if iq == 0 else levels.nprot)
    e = levels.sp_energy[sl].copy()
    dl = levels.deltaf[sl].copy()
    ws = levels.wstates[sl].copy()
    pw = levels.pairwg[sl].copy()
    nf = int(round(pn)) - 1
    ef0 = 0.5 * (e[nf] + e[nf + 1]) if 0 <= nf < len(e) - 1 else float(np.mean(e))

    def f(x):
        edif = e - x
        eq = np.maximum(np.sqrt(edif ** 2 + dl ** 2), 1e-20)
        return float(np.sum(np.clip(0.5 * (1.0 - edif / eq), 1e-10, 1.0 - 1e-10) * ws)) - pn
    try:
        ef = brentq(f, a=-100.0, b=100.0, xtol=1e-14, rtol=1e-16)
    except ValueError:
        ef = float(ef0)
    edif = e - ef
    dsq = dl * pw
    eq = np.maximum(np.sqrt(edif ** 2 + dsq ** 2), 1e-20)
    v2 = np.clip(0.5 * (1.0 - edif / eq), 1e-6, 1.0 - 1e-6)
    wguv = np.sqrt(np.maximum(v2 * (1.0 - v2), 1e-6))
    wsn, pwn = ws, pw
    if forces.pair_cutoff[iq] > 0.0:
        pwn = pw * soft_cutoff(e, ef + forces.pair_cutoff[iq], forces.softcut_range * forces.pair_cutoff[iq])
    if forces.state_cutoff[iq] > 0.0:
        wsn = ws * soft_cutoff(e, ef + forces.state_cutoff[iq], forces.softcut_range * forces.state_cutoff[iq])
    vol = 0.5 * wguv * wsn
    sumuv = max(float(np.sum(vol)), 1e-20)
    sumduv = float(np.sum(vol * dl))
    sumv2 = float(np.sum(v2 * wsn))
    sumdv2 = float(np.sum(dl * v2 * wsn))
    pairs.eferm[iq] = ef
    pairs.avdelt[iq] = sumduv / sumuv
    pairs.avdeltv2[iq] = sumdv2 / sumv2
    pairs.epair[iq] = sumduv
    pairs.avg[iq] = sumduv / (sumuv * sumuv)
    levels.wocc[sl] = v2
    levels.wguv[sl] = wguv
    if forces.pair_cutoff[iq] > 0.0:
        levels.pairwg[sl] = pwn
    if forces.state_cutoff[iq] > 0.0:
        levels.wstates[sl] = wsn
    return levels


def hfb_natural(forces, grids, levels, meanfield, params, static, pairs, isospin, iternat):
    """Gradient sub-iterations of the canonical basis in configuration space
    (static.py:920-1070) for the npsi states of one isospin (the reference
    slices by nneut/nprot and passes no pairing data, Q3): damped step along
    symcond, QR re-orthonormalisation (numpy.linalg.qr, LAPACK Householder
    as jnp.linalg.qr), transformed matrices, s.p. energies and gaps, pairing
    (a one-isospin sub-solve, the full pair at the last sub-iteration), new
    symcond; then psi <- psi cm, lagrange <- psi_new htran (the reference
    passes htran where its helper expects lambda), hmatrix, gapmatrix,
    lambda_save."""
    sl = levels.block(isospin)
    nst = sl.stop - sl.start
    cm = np.eye(nst, dtype=np.complex128)
    e = levels.sp_energy[sl]
    if static.e0bas == 0.0:
        dampstep = np.full(nst, static.delstepbas / (np.max(e) - np.min(e)))
    else:
        dampstep = static.delstepbas / (e - levels.sp_energy[sl.start] + static.e0bas)
    hmat = static.hmatrix[isospin, :nst, :nst].copy()
    gapmat = static.gapmatrix[isospin, :nst, :nst].copy()
    htran = gaptran = None
    for it in range(iternat):
        cm = cm - dampstep[:, None] * static.symcond[isospin, :nst, :nst]
        cm, _ = np.linalg.qr(cm)
        htran = np.conj(cm.T) @ hmat @ cm
        gaptran = np.conj(cm.T) @ gapmat @ cm
        hcm = hmat @ cm
        gapcm = gapmat @ cm
        last = it == iternat - 1
        if forces.ipair != 0 or last:
            levels.sp_energy[sl] = np.real(np.diag(htran))
            if forces.ipair != 0:
                levels.deltaf[sl] = np.real(np.diag(gaptran)) * levels.pairwg[sl]
                if not last:
                    pair_subiter(levels, forces, pairs, isospin)
                else:
                    pair(levels, meanfield, forces, params, grids, pairs)
        if not last:
            lam = lambda_matrix(htran, gaptran, levels, sl)
            w = levels.wocc[sl] * levels.wstates[sl]
            wuv = levels.wguv[sl] * levels.pairwg[sl] * levels.wstates[sl]
            static.symcond[isospin, :nst, :nst] = w[None, :] * hcm - wuv[None, :] * gapcm - cm @ lam
    shape = levels.psi[sl].shape
    psi_new = as2d(levels.psi[sl]) @ cm
    levels.psi[sl] = from2d(psi_new, shape)
    levels.lagrange[sl] = from2d(psi_new @ htran, shape)
    static.hmatrix[isospin, :nst, :nst] = htran
    static.gapmatrix[isospin, :nst, :nst] = gaptran
    static.lambda_save[isospin, :nst, :nst] = lambda_matrix(htran, gaptran, levels, sl)
    return levels, static


def lastdiag(forces, grids, levels, params, static, pairs, isospin):
    """Quasi-particle energies and lower-component norms of one isospin
    (static.py:1072-1168) over its npsi states (Q3): the matrix
    (h - eferm) xi + Delta eta in the canonical basis, xi_ij =
    -sqrt(v_i v_j) + sqrt((1-v_i)(1-v_j)), eta_ij = sqrt(v_i (1-v_j)) +
    sqrt(v_j (1-v_i)), v = wocc; eigh; the lower components from the
    coefficient matrix of time-reversed pairs (global index parity)."""
    sl = levels.block(isospin)
    nst = sl.stop - sl.start
    v = levels.wocc[sl]
    xi = -np.sqrt(np.outer(v, v)) + np.sqrt(np.outer(1.0 - v, 1.0 - v))
    eta = np.sqrt(np.outer(v, 1.0 - v)) + np.sqrt(np.outer(1.0 - v, v))
    h = static.hmatrix[isospin, :nst, :nst] - pairs.eferm[isospin] * np.eye(nst)
    H = h * xi + static.gapmatrix[isospin, :nst, :nst] * eta
    ev, U = np.linalg.eigh(0.5 * (H + np.conj(H.T)))
    gi = sl.start + np.arange(nst)
    coeff = np.zeros((nst, nst), dtype=np.complex128)
    for i in range(nst):
        for j in range(nst):
            if gi[i] - gi[j] == 1 and gi[i] % 2 == 0:
                coeff[i, j] = -np.sqrt(levels.wocc[gi[i]])
            if gi[j] - gi[i] == 1 and gi[i] % 2 == 1:
                coeff[i, j] = np.sqrt(levels.wocc[gi[i]])
    low = coeff @ U
    norm = np.sqrt(np.sum(np.abs(low) ** 2, axis=0))
    order = np.argsort(ev, kind="stable")
    return ev[order], norm[order]


# ---------------------------------------------------------------------------
# Constraints on the quadrupole moments and the principal axes
#                                                         constraint.py:1-430
# ---------------------------------------------------------------------------
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
    constr_field: np.ndarray = None   # (numconstraint, nx, ny, nz), the same for both isospins
    lambda_crank: np.ndarray = None
    goal_crank: np.ndarray = None
    actual_crank: np.ndarray = None
    old_crank: np.ndarray = None
    actual_crank2: np.ndarray = None
    qcorr: np.ndarray = None


def init_constraints(params, mass_number, grids):
    """Constraint fields with the reference's normalisations and masking
    (constraint.py:49-196): Q20 ~ 2z^2-x^2-y^2, Q22 ~ x^2-y^2, and with
    tq_prin_axes xy, xz, yz, x, y, z (goal 0), each times
    1/(1+exp((r-damprad)/dampgamma)); multipliers start at -0.2."""
    c = Constraints()
    for k in ("alpha20_wanted", "alpha22_wanted", "tq_prin_axes", "c0constr", "d0constr", "qepsconstr", "dampgamma",
              "damprad", "r0rms"):
        if k in params and (not k.startswith("alpha") or params[k] > -1e90):
            setattr(c, k, params[k])
    X = grids.x[:, None, None]
    Y = grids.y[None, :, None]
    Z = grids.z[None, None, :]
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
            fields.append(fxy * f * mask * np.ones((grids.nx, grids.ny, grids.nz)))
            goals.append(0.0)
    c.numconstraint = len(fields)
    if not fields:
        return c
    c.tconstraint = True
    c.constr_field = np.stack([np.broadcast_to(f, (grids.nx, grids.ny, grids.nz)) for f in fields])
    n = c.numconstraint
    c.lambda_crank = np.full(n, -0.2)
    c.goal_crank = np.array(goals, dtype=float)
    c.actual_crank = np.zeros(n)
    c.old_crank = np.zeros(n)
    c.actual_crank2 = np.zeros(n)
    c.qcorr = np.zeros(n)
    return c


def constraint_expectation_values(densities, constraints, grids):
    """<Q_i> = wxyz sum Q_i (rho_n + rho_p) and the variances
    |<Q_i^2> - <Q_i>^2 / N| (constraint.py:216-245)."""
    rho = densities.rho[0] + densities.rho[1]
    numb = grids.wxyz * np.sum(rho)
    f = constraints.constr_field
    ex = grids.wxyz * np.sum(f * rho[None], axis=(1, 2, 3))
    m2 = grids.wxyz * np.sum(f ** 2 * rho[None], axis=(1, 2, 3))
    return ex, np.abs(m2 - ex ** 2 / numb)


def tune_constraints(constraints, densities, levels, grids, e0act, x0act, apply_corrections=True):
    """Multiplier update and wave-function correction (constraint.py:248-344):
    lambda_i += -qeps max(e0,1)/x0 (<Q_i> - old_i)/(2 var_i + d0), qcorr_i =
    c0 (<Q_i> - goal_i)/(2 var_i + d0), psi_s *= exp(-wocc_s wstates_s sum_i
    qcorr_i Q_i) (the reference indexes the isospin as isospin - 1, which
    for 0/1 labels picks the wrong field; the field is isoscalar, so the
    isospin does not matter here)."""
    c = constraints
    ex, var = constraint_expectation_values(densities, c, grids)
    c.lambda_crank = c.lambda_crank - c.qepsconstr * (max(e0act, 1.0) / x0act) * (ex - c.old_crank) / (
        2.0 * var + c.d0constr)
    c.qcorr = c.c0constr * (ex - c.goal_crank) / (2.0 * var + c.d0constr)
    c.actual_crank, c.actual_crank2 = ex, var
    if apply_corrections:
        corr = -np.tensordot(c.qcorr, c.constr_field, axes=(0, 0))
        w = levels.wstates * levels.wocc
        levels.psi = levels.psi * np.exp(w[:, None, None, None, None] * corr[None, None])
    return c


# ---------------------------------------------------------------------------
# Single-particle properties                                  inout.py:64-149
# ---------------------------------------------------------------------------
def sp_properties(forces, grids, levels, cmtot=(0.0, 0.0, 0.0)):
    """Kinetic energy, <L>, <S>, parity per state (inout.py:64-149)."""
    psi = levels.psi
    xx = (grids.x - cmtot[0])[None, None, :, None, None]
    yy = (grids.y - cmtot[1])[None, None, None, :, None]
    zz = (grids.z - cmtot[2])[None, None, None, None, :]
    psx, _ = cderv01(grids.dx, psi, -3)
    psy, _ = cderv01(grids.dy, psi, -2)
    psz, _ = cderv01(grids.dz, psi, -1)
    k2 = (k2_table(grids.nx, grids.dx)[:, None, None] + k2_table(grids.ny, grids.dy)[None, :, None]
          + k2_table(grids.nz, grids.dz)[None, None, :])
    psw = sfft.ifftn(sfft.fftn(psi, axes=(-3, -2, -1), workers=WORKERS) * k2, axes=(-3, -2, -1), workers=WORKERS)
    ax = (1, 2, 3, 4)
    re, im = np.real(psi), np.imag(psi)
    ox = np.sum(re * (yy * np.imag(psz) - zz * np.imag(psy)) + im * (zz * np.real(psy) - yy * np.real(psz)), axis=ax)
    oy = np.sum(re * (zz * np.imag(psx) - xx * np.imag(psz)) + im * (xx * np.real(psz) - zz * np.real(psx)), axis=ax)
    oz = np.sum(re * (xx * np.imag(psy) - yy * np.imag(psx)) + im * (yy * np.real(psx) - xx * np.real(psy)), axis=ax)
    p0, p1 = psi[:, 0], psi[:, 1]
    a3 = (1, 2, 3)
    sx = np.sum(np.real(np.conj(p0) * p1) + np.real(np.conj(p1) * p0), axis=a3)
    sy = np.sum(np.real(np.conj(p0) * p1 * (-1j)) + np.real(np.conj(p1) * p0 * 1j), axis=a3)
    sz = np.sum(np.real(np.conj(p0) * p0) - np.real(np.conj(p1) * p1), axis=a3)
    kin_raw = -np.sum(np.real(np.conj(psi) * psw), axis=ax)
    h2m_f = np.where(levels.isospin == 0, forces.h2m[0], forces.h2m[1])
    flipped = psi[:, :, ::-1, ::-1, ::-1]
    par = np.sum(re * np.real(flipped) + im * np.imag(flipped), axis=ax)
    levels.sp_orbital = grids.wxyz * np.stack([ox, oy, oz], axis=1)
    levels.sp_spin = grids.wxyz * 0.5 * np.stack([sx, sy, sz], axis=1)
    levels.sp_kinetic = grids.wxyz * h2m_f * kin_raw
    levels.sp_parity = grids.wxyz * par
    return levels


# ---------------------------------------------------------------------------
# Energies                                                   energies.py:90-371
# ---------------------------------------------------------------------------
def integ_energy(coulomb, densities, energies, forces, grids, levels, pairs, tcoul=True):
    """Energy-functional integrals (energies.py:90-306)."""
    f, w = forces, grids.wxyz
    rho, tau = densities.rho, densities.tau
    d2 = np.stack([lap3(grids, rho[iq]) for iq in range(2)])
    rhot = rho[0] + rho[1]
    rho0, rho1 = rhot, -rho[0] + rho[1]
    rhon, rhop = rho[0], rho[1]
    d2rhon, d2rhop = d2[0], d2[1]
    d2rho = d2rhon + d2rhop
    d2rho0, d2rho1 = d2rho, -d2rhon + d2rhop
    tau0, tau1 = tau[0] + tau[1], -tau[0] + tau[1]
    e = energies
    e.ehf0 = w * np.sum((f.b0 * rhot ** 2 - f.b0p * (rhop ** 2 + rhon ** 2)) / 2.0)
    e.ehfCrho0 = w * np.sum((f.Crho0 + f.Crho0D * rho0 ** f.power) * rho0 ** 2)
    e.ehfCrho1 = w * np.sum((f.Crho1 + f.Crho1D * rho0 ** f.power) * rho1 ** 2)
    e.ehf3 = w * np.sum(rhot ** f.power * (f.b3 * rhot ** 2 - f.b3p * (rhop ** 2 + rhon ** 2)) / 3.0)
    e.ehf2 = w * np.sum((-f.b2 * rhot * d2rho + f.b2p * (rhop * d2rhop + rhon * d2rhon)) / 2.0)
    e.ehfCdrho0 = w * np.sum(f.Cdrho0 * rho0 * d2rho0)
    e.ehfCdrho1 = w * np.sum(f.Cdrho1 * rho1 * d2rho1)
    e.ehfCtau0 = w * np.sum(f.Ctau0 * tau0 * rho0)
    e.ehfCtau1 = w * np.sum(f.Ctau1 * tau1 * rho1)
    e.e3corr = -f.power * e.ehf3 / 2.0
    e.ehf1 = w * np.sum(f.b1 * (rho[0] + rho[1]) * (tau[0] + tau[1])
                        - f.b1p * (rho[0] * tau[0] + rho[1] * tau[1]))
    c = densities.current
    # energies.py:157-167 index current[:,:,:,0,0] etc. (Q6): reproduced as written.
    e.ehfCj0 = -w * np.sum(f.Ctau0 * ((c[:, :, :, 0, 0] + c[:, :, :, 0, 1]) ** 2
                                      + (c[:, :, :, 1, 0] + c[:, :, :, 1, 1]) ** 2
                                      + (c[:, :, :, 2, 0] + c[:, :, :, 2, 1]) ** 2))
    e.ehfCj1 = -w * np.sum(f.Ctau1 * ((-c[:, :, :, 0, 0] + c[:, :, :, 0, 1]) ** 2
                                      + (-c[:, :, :, 1, 0] + c[:, :, :, 1, 1]) ** 2
                                      + (-c[:, :, :, 2, 0] + c[:, :, :, 2, 1]) ** 2))
    divj = np.stack([div3(grids, densities.sodens[iq]) for iq in range(2)])
    ehfls = w * np.sum(-f.b4 * (rho[0] + rho[1]) * (divj[0] + divj[1])
                       - f.b4p * (rho[1] * divj[1] + rho[0] * divj[0]))
    e.ehfCdJ0 = w * np.sum(f.CdJ0 * (rho[0] + rho[1]) * (divj[0] + divj[1]))
    e.ehfCdJ1 = w * np.sum(f.CdJ1 * (-rho[0] + rho[1]) * (-divj[0] + divj[1]))
    cc = np.stack([curl3(grids, c[iq]) for iq in range(2)])
    sd = densities.sdens
    ctot = cc[0] + cc[1]
    e.ehflsodd = w * np.sum(-f.b4 * (np.sum(sd[0] * ctot, axis=0) + np.sum(sd[1] * ctot, axis=0))
                            - f.b4p * (np.sum(sd[0] * cc[0], axis=0) + np.sum(sd[1] * cc[1], axis=0)))
    e.ehfls = ehfls + e.ehflsodd
    e.ehfc = 0.0
    e.ecorc = 0.0
    if tcoul:
        sc = -3.0 / 4.0 * f.slate if f.ex != 0 else 0.0
        e.ehfc = w * np.sum(0.5 * rho[1] * coulomb.wcoul + sc * rho[1] ** (4.0 / 3.0))
        e.ecorc = w * np.sum(sc / 3.0 * rho[1] ** (4.0 / 3.0))
    e.ehft = w * np.sum(f.h2m[0] * tau[0] + f.h2m[1] * tau[1])
    e.e_zpe = 17.3 / levels.mass_number ** 0.2 if f.zpe == 1 else 0.0
    e.ehfint = (e.ehft + e.ehf0 + e.ehf1 + e.ehf2 + e.ehf3 + e.ehfls + e.ehfc
                - np.sum(pairs.epair) - e.e_zpe)
    return e


def sum_energy(energies, levels, meanfield, pairs):
    """Single-particle-sum energy ehf (energies.py:309-371)."""
    e = energies
    wt = levels.wocc * levels.wstates
    e.ehf = (np.sum(wt * (levels.sp_kinetic + levels.sp_energy)) / 2.0 + e.e3corr + meanfield.ecorrp
             + e.ecorc - np.sum(pairs.epair) - e.e_zpe)
    e.tke = np.sum(wt * levels.sp_kinetic)
    e.orbital = np.sum(wt[:, None] * levels.sp_orbital, axis=0)
    e.spin = np.sum(wt[:, None] * levels.sp_spin, axis=0)
    e.total_angmom = e.orbital + e.spin
    return e


# ---------------------------------------------------------------------------
# Observables                                                output.py:177-293
# ---------------------------------------------------------------------------
def observables(densities, grids, nneut, nprot):
    """Particle numbers, c.m., rms radii, quadrupole moments (output.py:177-293)."""
    rho = densities.rho
    w = grids.wxyz
    X = grids.x[:, None, None]
    Y = grids.y[None, :, None]
    Z = grids.z[None, None, :]
    pnr = np.array([np.sum(rho[iq]) * w for iq in range(2)])
    cm = np.zeros((3, 2))
    for iq in range(2):
        for i, c in enumerate((X, Y, Z)):
            cm[i, iq] = np.sum(c * rho[iq]) * w / pnr[iq]
    rms = np.zeros(2)
    qmat = np.zeros((3, 3, 2))
    for iq in range(2):
        xx, yy, zz = X - cm[0, iq], Y - cm[1, iq], Z - cm[2, iq]
        r2 = xx ** 2 + yy ** 2 + zz ** 2
        rms[iq] = np.sqrt(np.sum(r2 * rho[iq]) * w / pnr[iq])
        co = (xx, yy, zz)
        for i in range(3):
            for j in range(3):
                if i == j:
                    qmat[i, j, iq] = np.sum((3 * co[i] ** 2 - r2) * rho[iq]) * w
                else:
                    qmat[i, j, iq] = np.sum(3 * co[i] * co[j] * rho[iq]) * w
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
    return dict(pnr=pnr, cm=cm, rms=rms, rmstot=rmstot, q20=q20, q22=q22, q20tot=q20tot,
                q22tot=q22tot, beta20tot=beta20, beta22tot=beta22,
                beta=np.sqrt(beta20 ** 2 + 2.0 * beta22 ** 2))


def complete_observables(densities, levels, grids, current=None):
    """Every observable of output.py:177-354 as a dict: N, c.m. (per isospin
    and total), isovector dipole, rms radii, second moments, quadrupole
    moments (per isospin, total, isovector q20T1), beta20/22, beta, gamma,
    <L>, <S>, <J> summed over states with wocc > 1e-10, and the momenta from
    the current when it is given."""
    ob = observables(densities, grids, levels.nneut, levels.nprot)
    rho = densities.rho
    w = grids.wxyz
    X = grids.x[:, None, None]
    Y = grids.y[None, :, None]
    Z = grids.z[None, None, :]
    pnr, cm = ob["pnr"], ob["cm"]
    pnrtot = pnr[0] + pnr[1]
    cmtot = (pnr[0] * cm[:, 0] + pnr[1] * cm[:, 1]) / pnrtot
    x2m = np.zeros((3, 2))
    for iq in range(2):
        for i, c in enumerate((X, Y, Z)):
            x2m[i, iq] = np.sum((c - cm[i, iq]) ** 2 * rho[iq]) * w / pnr[iq]
    x2mtot = (pnr[0] * x2m[:, 0] + pnr[1] * x2m[:, 1]) / pnrtot
    q20T1 = -ob["q20"][0] / levels.nneut + ob["q20"][1] / levels.nprot
    b20, b22 = ob["beta20tot"], ob["beta22tot"]
    gamma = abs(math.atan2(math.sqrt(2.0) * b22, b20)) * 180.0 / math.pi
    if gamma > 120.0:
        gamma -= 120.0
    if 60.0 < gamma <= 120.0:
        gamma = 120.0 - gamma
    dip = np.array([np.sum(c * (rho[0] / pnr[0] - rho[1] / pnr[1])) * w for c in (X, Y, Z)])
    sel = levels.wocc > 1e-10
    wt = (levels.wocc * levels.wstates)[sel]
    orb = (wt[:, None] * levels.sp_orbital[sel]).sum(axis=0)
    spn = (wt[:, None] * levels.sp_spin[sel]).sum(axis=0)
    pcm = np.zeros((3, 2))
    if current is not None:
        for iq in range(2):
            for i in range(3):
                pcm[i, iq] = np.sum(current[iq, i]) * w
    out = dict(ob)
    out.update(pnrtot=pnrtot, cmtot=cmtot, x2m=x2m, x2mtot=x2mtot, q20T1=q20T1, gamma=gamma,
               dipole_isovector=dip, orbital_momentum=orb, spin_momentum=spn, total_momentum=orb + spn, pcm=pcm)
    return out


# ---------------------------------------------------------------------------
# State ordering and plane-wave start              static.py:807-918, 1338-1392
# ---------------------------------------------------------------------------
SORTED_FIELDS = ("psi", "lagrange", "sp_energy", "deltaf", "sp_norm", "sp_efluct1", "sp_efluct2", "wocc", "wguv",
                 "wstates", "pairwg")


def sort_states(forces, grids, levels, isospin):
    """Order the states of one isospin block by sp_energy (stable), with the
    eleven per-state arrays the reference moves along (static.py:1338-1392).
    The block is the isospin's npsi states (the reference slices by
    nneut/nprot, which is the same block only when npsi = (N, Z))."""
    sl = levels.block(isospin)
    perm = sl.start + np.argsort(levels.sp_energy[sl], kind="stable")
    for name in SORTED_FIELDS:
        a = getattr(levels, name)
        a[sl] = a[perm]
    return levels


def planewave_order(grids, nucleon_mass, nmax=7):
    """Wave-number triples (kx, ky, kz), each in -nmax//2 .. nmax//2 with kx
    outermost, ordered by the kinetic energy hbc^2/2m sum_a ((2 pi k_a +
    bang_a) / (n_a d_a))^2, stable (static.py:807-845)."""
    r = range(-nmax // 2, nmax // 2 + 1)
    ks = [(kx, ky, kz) for kx in r for ky in r for kz in r]
    bang = (getattr(grids, "bangx", 0.0), getattr(grids, "bangy", 0.0), getattr(grids, "bangz", 0.0))
    ns = (grids.nx, grids.ny, grids.nz)
    ds = (grids.dx, grids.dy, grids.dz)
    en = np.array([HBC ** 2 / (2.0 * nucleon_mass) * sum(((2.0 * math.pi * k[a] + bang[a]) / (ns[a] * ds[a])) ** 2
                                                         for a in range(3)) for k in ks])
    return [ks[i] for i in np.argsort(en, kind="stable")]


def planewaves(grids, levels, forces):
    """Plane-wave initial states (static.py:807-918): per isospin block, the
    wave numbers in the order of planewave_order, each first with spin up
    then spin down, normalised; wocc = 1 for the first N (Z) states of each
    block, wguv = 0."""
    order = planewave_order(grids, forces.nucleon_mass)
    I = np.arange(grids.nx)[:, None, None]
    J = np.arange(grids.ny)[None, :, None]
    K = np.arange(grids.nz)[None, None, :]
    bang = (getattr(grids, "bangx", 0.0), getattr(grids, "bangy", 0.0), getattr(grids, "bangz", 0.0))
    psi = np.zeros_like(levels.psi)
    for iq in range(2):
        sl = levels.block(iq)
        for m in range(sl.stop - sl.start):
            k = order[m // 2]
            ph = (I * ((2.0 * math.pi * k[0] + bang[0]) / grids.nx) + J * ((2.0 * math.pi * k[1] + bang[1]) / grids.ny)
                  + K * ((2.0 * math.pi * k[2] + bang[2]) / grids.nz))
            s = sl.start + m
            psi[s, m % 2] = np.cos(ph) + 1j * np.sin(ph)
            psi[s] /= math.sqrt(np.sum(np.abs(psi[s]) ** 2) * grids.wxyz)
    levels.psi = psi
    levels.wocc = np.zeros(levels.nstmax)
    levels.wocc[:levels.nneut] = 1.0
    n0 = int(levels.npsi[0])
    levels.wocc[n0:n0 + levels.nprot] = 1.0
    levels.wguv = np.zeros(levels.nstmax)
    return levels


# ---------------------------------------------------------------------------
# Initial states                                             static.py:580-738
# ---------------------------------------------------------------------------
def shell_quantum_numbers(nps):
    """(i, j, k) per state of one isospin block, two spins per shell entry,
    k-outer / j / i-inner order with ka = i+j+k (static.py:613-659)."""
    out = []
    ka = 0
    while len(out) < nps:
        for k in range(ka + 1):
            for j in range(ka + 1):
                for i in range(ka + 1):
                    if ka == i + j + k:
                        for _ in range(2):
                            if len(out) < nps:
                                out.append((i, j, k))
        ka += 1
    return out


def harmosc(grids, levels, static):
    """Gaussian x monomial initial states, normalised (static.py:580-738)."""
    X = grids.x[:, None, None]
    Y = grids.y[None, :, None]
    Z = grids.z[None, None, :]
    gauss = np.exp(-((X / static.radinx) ** 2 + (Y / static.radiny) ** 2 + (Z / static.radinz) ** 2))
    gauss = gauss / np.sqrt(rpsnorm(np.stack([gauss, np.zeros_like(gauss)]), grids.wxyz))
    psi = np.zeros_like(levels.psi)
    for iq in range(2):
        sl = levels.block(iq)
        nps = sl.stop - sl.start
        qn = shell_quantum_numbers(nps)
        for m in range(nps):
            s = sl.start + m
            if m == 0:
                psi[s, 0] = gauss
            else:
                i, j, k = qn[m]
                poly = (grids.x ** i)[:, None, None] * (grids.y ** j)[None, :, None] * (grids.z ** k)[None, None, :]
                psi[s, m % 2] = gauss * poly
            an = rpsnorm(psi[s], grids.wxyz)
            if m == 0 or an > 1e-12:
                psi[s] = psi[s] / np.sqrt(an)
    wocc = np.zeros(levels.nstmax)
    wocc[:levels.nneut] = 1.0
    wocc[int(levels.npsi[0]):int(levels.npsi[0]) + levels.nprot] = 1.0
    levels.psi = psi
    levels.wocc = wocc
    levels.wguv = np.zeros_like(wocc)
    return levels


def add_noise(psi, amplitude=1e-3, seed=0):
    """Seeded noise psi_i += a * max|psi_i| * (xi + i eta)/sqrt2 with xi, eta
    ~ N(0,1) from PCG64(seed), xi drawn for the whole array before eta.
    Build-defined: the reference's harmosc has no noise."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xi = rng.standard_normal(psi.shape)
    eta = rng.standard_normal(psi.shape)
    amax = np.max(np.abs(psi).reshape(psi.shape[0], -1), axis=1)[:, None, None, None, None]
    return psi + amplitude * amax * (xi + 1j * eta) / np.sqrt(2.0)


def deformed_radii(b, beta2):
    """Gaussian widths radin = sqrt2*b, stretched along z by an axial
    quadrupole bias beta2 to first order (build-defined)."""
    r = math.sqrt(2.0) * b
    c = math.sqrt(5.0 / (4.0 * math.pi)) * beta2
    return r * (1.0 - 0.5 * c), r * (1.0 - 0.5 * c), r * (1.0 + c)


# ---------------------------------------------------------------------------
# Driver                                                   static.py:1666-1882
# ---------------------------------------------------------------------------
@dataclass
class HFState:
    forces: Forces
    grids: Grids
    levels: Levels
    meanfield: Meanfield
    densities: Densities
    coulomb: Coulomb
    static: Static
    params: Params
    energies: Energies
    pairs: Pairs
    constraints: Constraints = None


def setup(name="SLy4", n=24, d=1.0, nneut=8, nprot=8, npsi=None, b=1.8, beta2=0.0,
          ipair=0, tbcs=False, v0=None, noise=1e-3, seed=0, mprint=10, constraints=None, **static_kw) -> HFState:
    """Build every state container and the initial wave functions
    (init sequence of test.py:145-163)."""
    forces = init_forces(name, ipair=ipair, tbcs=tbcs,
                         v0prot=v0, v0neut=v0)
    grids = init_grids(n, n, n, d, d, d)
    levels = init_levels(grids, nneut, nprot, npsi if npsi is not None else [nneut, nprot])
    rx, ry, rz = deformed_radii(b, beta2)
    forces, static = init_static(forces, levels, radinx=rx, radiny=ry, radinz=rz, **static_kw)
    coulomb = init_coulomb(grids)
    levels = harmosc(grids, levels, static)
    if noise:
        levels.psi = add_noise(levels.psi, noise, seed)
    return HFState(forces, grids, levels, init_meanfield(grids), init_densities(grids), coulomb, static,
                   Params(mprint=mprint), Energies(), Pairs(),
                   init_constraints(constraints, nneut + nprot, grids) if constraints else None)


def prelude(st: HFState, restart=False):
    """Steps before the loop (static.py:1700-1740).  restart: the reference's
    params.trestart branch (static.py:1700-1706), which keeps the iteration
    number and skips the first orthonormalisation."""
    if not restart:
        st.params.iteration = 0
        diagstep(st.energies, st.forces, st.grids, st.levels, st.static, False, True)
    add_density(st.densities, st.grids, st.levels)
    skyrme(st.coulomb, st.densities, st.forces, st.grids, st.meanfield, constraints=getattr(st, "constraints", None))
    grstep(st.forces, st.grids, st.levels, st.meanfield, st.params, st.static)
    if st.forces.ipair != 0:
        pair(st.levels, st.meanfield, st.forces, st.params, st.grids, st.pairs)
    diagstep(st.energies, st.forces, st.grids, st.levels, st.static, False, True)
    sp_properties(st.forces, st.grids, st.levels)
    integ_energy(st.coulomb, st.densities, st.energies, st.forces, st.grids, st.levels, st.pairs)
    sum_energy(st.energies, st.levels, st.meanfield, st.pairs)
    if st.static.tvaryx_0:
        st.static.x0dmp *= 3.0
    st._tbcssav = st.forces.tbcs


def iteration(st: HFState, i: int):
    """One SCF iteration, the loop body of static.py:1755-1882 without output."""
    f, s = st.forces, st.static
    c = getattr(st, "constraints", None)
    if c is not None and c.tconstraint:  # constraint.py:321-325
        c.old_crank = c.actual_crank.copy()
    st.params.iteration = i
    f.tbcs = True if i <= s.inibcs else st._tbcssav
    s.tdiag = not (i > s.inidiag)
    grstep(f, st.grids, st.levels, st.meanfield, st.params, s)
    if f.tbcs:
        s.tdiag = True
    diagstep(st.energies, f, st.grids, st.levels, s, s.tdiag, True)
    if f.ipair != 0:
        pair(st.levels, st.meanfield, f, st.params, st.grids, st.pairs)
    if f.ipair != 0 and s.iternat > 0 and i > s.iternat_start and not f.tbcs:  # static.py:1795-1797
        for iq in range(2):
            hfb_natural(f, st.grids, st.levels, st.meanfield, st.params, s, st.pairs, iq, s.iternat)
    if c is not None and c.tconstraint:
        tune_constraints(c, st.densities, st.levels, st.grids, s.e0dmp, s.x0dmp)
    d = st.densities
    old_rho, old_tau, old_chi = d.rho.copy(), d.tau.copy(), d.chi.copy()
    add_density(d, st.grids, st.levels)
    d.rho = 0.2 * d.rho + 0.8 * old_rho
    d.tau = 0.2 * d.tau + 0.8 * old_tau
    d.chi = 0.2 * d.chi + 0.8 * old_chi
    skyrme(st.coulomb, d, f, st.grids, st.meanfield, constraints=c)
    if s.tsort:  # static.py:1829-1831
        for iq in range(2):
            sort_states(f, st.grids, st.levels, iq)
    sp_properties(f, st.grids, st.levels)
    integ_energy(st.coulomb, d, st.energies, f, st.grids, st.levels, st.pairs)
    sum_energy(st.energies, st.levels, st.meanfield, st.pairs)
    e = st.energies
    if s.tvaryx_0:
        if ((e.ehf < e.ehfprev and e.efluct1[0] < e.efluct1prev * (1.0 - 1.0e-5))
                or e.efluct2[0] < e.efluct2prev * (1.0 - 1.0e-5)):
            s.x0dmp *= 1.005
        else:
            s.x0dmp *= 0.8
        s.x0dmp = float(np.clip(s.x0dmp, s.x0dmpmin, s.x0dmpmin * 5.0))
        e.efluct1prev, e.efluct2prev, e.ehfprev = e.efluct1[0], e.efluct2[0], e.ehf
    return e.efluct1[0] < s.serr and i > 1


def statichf(st: HFState, maxiter=None, callback=None):
    """Prelude + loop (static.py:1666-1882); returns the per-iteration ehf."""
    prelude(st)
    hist = []
    for i in range(1, (maxiter or st.static.maxiter) + 1):
        conv = iteration(st, i)
        hist.append(st.energies.ehf)
        if callback is not None:
            callback(i, st)
        if conv:
            break
    return hist
