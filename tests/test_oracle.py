"""Pins of the FP64 oracle (oracle/hfb_oracle.py), CPU only.

Each test checks an oracle function against something its own text does not
restate: a closed form, an invariant, a brute-force sum, a library routine,
or another oracle function restated from a different reference file (for
example the dense sder matrices of grids.py against the FFT derivatives of
levels.py).  The mapping function -> pins is:

  sder, sder2, cderv00/01 ...... test_dense_matrices_equal_fft_derivatives,
                                 test_cderv_plane_waves
  cderv02 (Nyquist mean rule) .. test_cderv02_constant_mass, test_cderv02_variable_mass,
                                 test_cderv02_nyquist_uses_axis_mean
  laplace, k2_table ............ test_laplace_plane_wave
  hpsi01 ....................... test_hpsi_hermitian, test_hpsi_plane_wave_eigenvalue,
                                 test_energy_functional_derivative_is_hpsi
  skyrme, density_contributions,
  integ_energy (ehfint) ........ test_energy_functional_derivative_is_hpsi,
                                 test_sum_energy_equals_ehfint_for_consistent_potentials
  integ_energy C-split ......... test_c_coupling_split_reproduces_b_form
  density_contributions ........ test_density_plane_wave
  poisson, init_coulomb ........ test_poisson_direct_sum
  loewdin_orthonormalize ....... test_loewdin
  diagstep, construct_hfb ...... test_diagstep_diagonalises_subspace
  find_eferm, pairdn, pair ..... test_bcs_degenerate_closed_form, test_bcs_matches_root
  sp_properties ................ test_sp_properties_closed_forms
  sum_energy ................... test_sum_energy_equals_ehfint_for_consistent_potentials
  observables .................. test_observables_gaussian
  shell_quantum_numbers, harmosc test_harmosc_shells
  init_forces / derive ......... test_c_coupling_split_reproduces_b_form, test_force_constants
  iteration / statichf ......... test_oracle_golden_o16 (regression against the committed
                                 fixture), test_energy_decreases_o16
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def rng(seed=1):
    return np.random.Generator(np.random.PCG64(seed))


def cplx(r, shape):
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def kvec(n, d):
    """Physical wave numbers of the DFT bins (Nyquist as +pi/d)."""
    m = np.arange(n)
    m = np.where(m < n // 2, m, m - n)
    k = 2 * math.pi * m / (n * d)
    k[n // 2] = math.pi / d
    return k


def grid_coords(G):
    return np.meshgrid(G.x, G.y, G.z, indexing="ij")


# ---------------------------------------------------------------------------
# Derivatives
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,d", [(6, 0.9), (8, 1.0), (16, 0.7)])
def test_dense_matrices_equal_fft_derivatives(O, n, d):
    """grids.py:36-63 dense matrices vs levels.py:131-281 FFT derivatives."""
    G = O.init_grids(n, n, n, d, d, d)
    f = cplx(rng(n), (n, n, n))
    for axis, app, m1, m2 in ((-3, O.apply_x, G.der1x, G.der2x), (-2, O.apply_y, G.der1y, G.der2y),
                              (-1, O.apply_z, G.der1z, G.der2z)):
        d1 = O.cderv00(d, f, axis)
        d1b, d2 = O.cderv01(d, f, axis)
        scale = np.max(np.abs(f)) / d ** 2
        assert np.max(np.abs(app(m1, f) - d1)) < 1e-12 * scale
        assert np.max(np.abs(d1b - d1)) < 1e-12 * scale
        assert np.max(np.abs(app(m2, f) - d2)) < 1e-11 * scale


@pytest.mark.parametrize("n,d", [(8, 1.0), (10, 0.8)])
def test_cderv_plane_waves(O, n, d):
    """exp(i k x): d1 = i k (0 at m=0 and Nyquist); d2 = -k^2 (Nyquist kept)."""
    j = np.arange(n)
    k = kvec(n, d)
    for m in range(n):
        f = np.exp(2j * math.pi * m * j / n)
        d1, d2 = O.cderv01(d, f, -1)
        e1 = 0.0 if m in (0, n // 2) else 1j * k[m]
        assert np.allclose(d1, e1 * f, atol=1e-12)
        assert np.allclose(d2, -(k[m] ** 2) * f, atol=1e-11)
        assert np.allclose(O.cderv00(d, f, -1), e1 * f, atol=1e-12)


def test_cderv02_constant_mass(O):
    """Constant B: d(B d psi) = B * d2 psi exactly, Nyquist included."""
    n, d = 8, 0.9
    psi = cplx(rng(2), (2, n, n, n))
    B = np.full((n, n, n), 1.7)
    for axis in (-3, -2, -1):
        d1, d2 = O.cderv02(d, psi, B, axis)
        e1, e2 = O.cderv01(d, psi, axis)
        assert np.allclose(d1, e1, atol=1e-12)
        assert np.allclose(d2, 1.7 * e2, atol=1e-10)


def test_cderv02_variable_mass(O):
    """Band-limited psi and B: d/dx(B dpsi/dx) analytic."""
    n, d = 16, 1.0
    x = np.arange(n) * d
    k1, k2 = 2 * math.pi * 2 / (n * d), 2 * math.pi * 3 / (n * d)
    psi1 = np.exp(1j * k1 * x)
    B1 = 2.0 + 0.3 * np.cos(k2 * x)
    dB1 = -0.3 * k2 * np.sin(k2 * x)
    expect1 = 1j * k1 * dB1 * psi1 - k1 ** 2 * B1 * psi1
    psi = np.broadcast_to(psi1[:, None, None], (n, n, n)).copy()
    B = np.broadcast_to(B1[:, None, None], (n, n, n)).copy()
    _, d2 = O.cderv02(d, psi, B, -3)
    assert np.max(np.abs(d2 - expect1[:, None, None])) < 1e-11


def test_cderv02_nyquist_uses_axis_mean(O):
    """Nyquist mode along x: d2 = -(pi/d)^2 * mean_x(B)(y, z) * psi."""
    n, d = 8, 1.1
    r = rng(3)
    B = 1.0 + r.random((n, n, n))
    psi = np.broadcast_to(np.cos(math.pi * np.arange(n))[:, None, None], (n, n, n)) * (1 + 0j)
    d1, d2 = O.cderv02(d, psi, B, -3)
    assert np.max(np.abs(d1)) < 1e-12
    expect = -(math.pi / d) ** 2 * B.mean(axis=0)[None] * psi
    assert np.max(np.abs(d2 - expect)) < 1e-11


def test_laplace_plane_wave(O):
    """The preconditioner divides each plane wave by w(E0 + h2ma k^2) + wuv vmax/2."""
    n, d = 8, 1.0
    G = O.init_grids(n, n, n, d, d, d)
    k = kvec(n, d)
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for (mx, my, mz), w, wuv, vmax in (((1, 2, 3), 1.0, 0.0, 0.0), ((4, 0, 7), 0.05, 0.3, 2.0),
                                       ((4, 4, 4), 0.7, 0.02, 5.0)):
        psi = np.exp(2j * math.pi * (mx * I + my * J + mz * K) / n)
        k2 = k[mx] ** 2 + k[my] ** 2 + k[mz] ** 2
        out = O.laplace(G, 20.7, w, wuv, psi, vmax, 100.0)
        den = max(w, 0.1) * (100.0 + 20.7 * k2) + 0.5 * max(wuv, 0.1) * vmax
        assert np.allclose(out, psi / den, atol=1e-14)


# ---------------------------------------------------------------------------
# h psi
# ---------------------------------------------------------------------------
def random_meanfield(O, G, r, scale=1.0):
    n = G.nx
    mf = O.init_meanfield(G)
    mf.upot = -50 + 10 * r.standard_normal((2, n, n, n)) * scale
    mf.bmass = 20 + 3 * r.random((2, n, n, n))
    mf.aq = r.standard_normal((2, 3, n, n, n)) * scale
    mf.spot = r.standard_normal((2, 3, n, n, n)) * scale
    mf.wlspot = r.standard_normal((2, 3, n, n, n)) * scale
    mf.v_pair = r.standard_normal((2, n, n, n))
    return mf


def test_hpsi_hermitian(O):
    """<phi|h psi> = conj <psi|h phi> for arbitrary real potential fields."""
    n, d = 8, 1.0
    G = O.init_grids(n, n, n, d, d, d)
    r = rng(4)
    mf = random_meanfield(O, G, r)
    psi = cplx(r, (2, n, n, n))
    phi = cplx(r, (2, n, n, n))
    for iq in (0, 1):
        _, hpsi, _ = O.hpsi01(G, mf, iq, 1.0, 0.0, psi)
        _, hphi, _ = O.hpsi01(G, mf, iq, 1.0, 0.0, phi)
        a = np.vdot(phi, hpsi)
        b = np.conj(np.vdot(psi, hphi))
        assert abs(a - b) < 1e-12 * abs(a)


def test_hpsi_plane_wave_eigenvalue(O):
    """Constant fields: h e^{ikr} chi = (U + B k^2 + A.k + S.sigma + W.(k x sigma)) e^{ikr} chi."""
    n, d = 8, 1.0
    G = O.init_grids(n, n, n, d, d, d)
    k = kvec(n, d)
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.array([[1, 0], [0, -1]], complex)
    sig = (sx, sy, sz)
    r = rng(5)
    U, B = -40.0, 21.0
    A = r.standard_normal(3)
    S = r.standard_normal(3)
    W = r.standard_normal(3)
    mf = O.init_meanfield(G)
    mf.upot[:] = U
    mf.bmass[:] = B
    for c in range(3):
        mf.aq[:, c] = A[c]
        mf.spot[:, c] = S[c]
        mf.wlspot[:, c] = W[c]
    chi = cplx(r, 2)
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1
    for m in ((1, 2, 3), (0, 5, 7), (2, 0, 0)):
        kk = np.array([k[m[0]], k[m[1]], k[m[2]]])
        phase = np.exp(2j * math.pi * (m[0] * I + m[1] * J + m[2] * K) / n)
        psi = chi[:, None, None, None] * phase
        # W.(k x sigma) = sum_cab eps_cab W_c k_a sigma_b
        so = sum(eps[c, a, b] * W[c] * kk[a] * sig[b] for c in range(3) for a in range(3) for b in range(3))
        hmat = (U + B * kk @ kk + A @ kk) * np.eye(2) + sum(S[c] * sig[c] for c in range(3)) + so
        _, h, _ = O.hpsi01(G, mf, 0, 1.0, 0.0, psi)
        expect = (hmat @ chi)[:, None, None, None] * phase
        assert np.max(np.abs(h - expect)) < 1e-10 * np.max(np.abs(expect))


def test_hpsi_nyquist_kinetic(O):
    """Constant B: the Nyquist plane wave along x has kinetic eigenvalue B (pi/d)^2."""
    n, d = 8, 1.2
    G = O.init_grids(n, n, n, d, d, d)
    mf = O.init_meanfield(G)
    mf.bmass[:] = 20.0
    I = np.arange(n)[:, None, None] * np.ones((n, n, n))
    psi = np.stack([np.cos(math.pi * I), np.zeros((n, n, n))]) + 0j
    _, h, _ = O.hpsi01(G, mf, 1, 1.0, 0.0, psi)
    assert np.allclose(h, 20.0 * (math.pi / d) ** 2 * psi, atol=1e-9)


# ---------------------------------------------------------------------------
# Energy functional <-> mean field <-> h psi
# ---------------------------------------------------------------------------
def smooth_real_states(O, G, nst, seed):
    """Real two-spinors (so the current j vanishes): Gaussian x low-order
    polynomials with fixed real spin mixtures."""
    X, Y, Z = grid_coords(G)
    r = rng(seed)
    g = np.exp(-(X ** 2 + 1.2 * Y ** 2 + 0.9 * Z ** 2) / 2.0 ** 2)
    polys = [np.ones_like(X), X, Y, Z, X * Y, Z * Z - 1.0]
    out = np.zeros((nst, 2, G.nx, G.ny, G.nz), complex)
    for s in range(nst):
        th = r.uniform(0, math.pi)
        f = g * (polys[s % len(polys)] + 0.2 * r.standard_normal() * X)
        out[s, 0] = math.cos(th) * f
        out[s, 1] = math.sin(th) * f
        out[s] /= math.sqrt(O.rpsnorm(out[s], G.wxyz))
    return out


def band_limited(phi):
    """Zero every Nyquist plane of a (..., n, n, n) field."""
    f = np.fft.fftn(phi, axes=(-3, -2, -1))
    n = phi.shape[-1]
    f[..., n // 2, :, :] = 0
    f[..., :, n // 2, :] = 0
    f[..., :, :, n // 2] = 0
    return np.fft.ifftn(f, axes=(-3, -2, -1))


def _functional_setup(O, n=16, d=1.0, nst=3):
    G = O.init_grids(n, n, n, d, d, d)
    F = O.init_forces("SV-bas")
    C = O.init_coulomb(G)
    psi = {0: smooth_real_states(O, G, nst, 11), 1: smooth_real_states(O, G, nst, 12)}
    return G, F, C, psi


def _densities(O, G, psi):
    D = O.init_densities(G)
    for iq in (0, 1):
        nst = psi[iq].shape[0]
        rho, chi, tau, cur, sd, so = O.density_contributions(G, np.ones(nst), np.zeros(nst), psi[iq])
        D.rho[iq], D.chi[iq], D.tau[iq] = rho, chi, tau
        D.current[iq], D.sdens[iq], D.sodens[iq] = cur, sd, so
    return D


def _ehfint(O, G, F, C, psi):
    D = _densities(O, G, psi)
    C.wcoul = O.poisson(G, C, D.rho)
    lv = O.init_levels(O.init_grids(2, 2, 2, 1.0, 1.0, 1.0), 8, 8, [8, 8])
    return O.integ_energy(C, D, O.Energies(), F, G, lv, O.Pairs()).ehfint


def test_energy_functional_derivative_is_hpsi(O):
    """d ehfint[psi + eps phi]/d eps = 2 Re sum_i wxyz <phi_i| h_q psi_i>.

    Pins hpsi01 (every term), skyrme (U, B, W, A, S from meanfield.py) and
    integ_energy (energies.py) against each other: they are restated from
    different reference files, and a dropped term, a wrong sign or a wrong
    coupling in any of them breaks the identity.  The base states are real
    so that the current j vanishes (integ_energy's ehfint has no j^2 term);
    the directions phi are complex and band-limited (the Nyquist rule of
    cderv02 is not the derivative of a functional)."""
    G, F, C, psi = _functional_setup(O)
    D = _densities(O, G, psi)
    mf = O.init_meanfield(G)
    O.skyrme(C, D, F, G, mf)
    r = rng(13)
    for trial in range(3):
        phi = {iq: band_limited(cplx(r, psi[iq].shape)) * 0.05 for iq in (0, 1)}
        expect = 0.0
        for iq in (0, 1):
            _, h, _ = O.hpsi01(G, mf, iq, 1.0, 0.0, psi[iq])
            expect += 2.0 * G.wxyz * np.real(np.vdot(phi[iq], h))
        eps = 1e-4
        ep = _ehfint(O, G, F, C, {iq: psi[iq] + eps * phi[iq] for iq in (0, 1)})
        em = _ehfint(O, G, F, C, {iq: psi[iq] - eps * phi[iq] for iq in (0, 1)})
        fd = (ep - em) / (2 * eps)
        assert abs(fd - expect) < 1e-6 * abs(expect), (trial, fd, expect)


def test_sum_energy_equals_ehfint_for_consistent_potentials(O):
    """With potentials computed from the same densities, the single-particle
    sum of sum_energy (energies.py:309-371) equals the functional integral
    ehfint (energies.py:267): the e3corr and ecorc corrections of the
    reference make the rearrangement terms agree.  The states are
    band-limited: the Nyquist rows of cderv02 and of the FFT Laplacian in
    sp_properties are the only parts without a functional counterpart."""
    G, F, C, psi = _functional_setup(O)
    psi = {iq: np.real(band_limited(psi[iq])) + 0j for iq in (0, 1)}
    D = _densities(O, G, psi)
    mf = O.init_meanfield(G)
    O.skyrme(C, D, F, G, mf)
    nst = psi[0].shape[0]
    lv = O.init_levels(G, nst, nst, [nst, nst])
    lv.psi = np.concatenate([psi[0], psi[1]])
    for iq in (0, 1):
        _, h, _ = O.hpsi01(G, mf, iq, 1.0, 0.0, psi[iq])
        lv.sp_energy[lv.block(iq)] = G.wxyz * np.real(np.sum(np.conj(psi[iq]) * h, axis=(1, 2, 3, 4)))
    O.sp_properties(F, G, lv)
    E = O.integ_energy(C, D, O.Energies(), F, G, lv, O.Pairs())
    O.sum_energy(E, lv, mf, O.Pairs())
    assert abs(E.ehf - E.ehfint) < 1e-11 * abs(E.ehfint), (E.ehf, E.ehfint)
    # the kinetic energy of sp_properties equals the functional's ehft
    assert abs(E.tke - E.ehft) < 1e-11 * abs(E.ehft)


def test_c_coupling_split_reproduces_b_form(O):
    """energies.py's C-coupling split equals the b-coupling integrals
    (isoscalar/isovector algebra of forces.py:95-122)."""
    n, d = 12, 1.0
    G = O.init_grids(n, n, n, d, d, d)
    r = rng(6)
    X, Y, Z = grid_coords(G)
    base = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 9.0)
    D = O.init_densities(G)
    for iq in (0, 1):
        D.rho[iq] = 0.08 * base * (1 + 0.1 * iq + 0.05 * r.random((n, n, n)))
        D.tau[iq] = 0.1 * base * (1 + 0.05 * r.random((n, n, n)))
        D.sodens[iq] = 0.01 * r.standard_normal((3, n, n, n)) * base
        D.current[iq] = 0.01 * r.standard_normal((3, n, n, n)) * base
        D.sdens[iq] = 0.01 * r.standard_normal((3, n, n, n)) * base
    for name in ("SV-bas", "SLy4"):
        F = O.init_forces(name)
        C = O.init_coulomb(G)
        C.wcoul = O.poisson(G, C, D.rho)
        lv = O.init_levels(O.init_grids(2, 2, 2, 1.0, 1.0, 1.0), 8, 8, [8, 8])
        E = O.integ_energy(C, D, O.Energies(), F, G, lv, O.Pairs())
        tol = 1e-12 * (abs(E.ehf0) + abs(E.ehf3))
        assert abs((E.ehf0 + E.ehf3) - (E.ehfCrho0 + E.ehfCrho1)) < tol
        assert abs(E.ehf1 - (E.ehfCtau0 + E.ehfCtau1)) < 1e-12 * abs(E.ehf1)
        assert abs(E.ehf2 - (E.ehfCdrho0 + E.ehfCdrho1)) < 1e-12 * abs(E.ehf2)
        assert abs((E.ehfls - E.ehflsodd) - (E.ehfCdJ0 + E.ehfCdJ1)) < 1e-12 * abs(E.ehfls - E.ehflsodd) + 1e-15


def test_force_constants(O):
    """Nucleon mass from h2m and hbar c; Slater constant (3/pi)^(1/3) e2."""
    F = O.init_forces("SLy4")
    assert abs(F.nucleon_mass - 938.9) < 0.1
    assert abs(F.slate - (3 / math.pi) ** (1 / 3) * 1.4399784085965135) < 1e-15
    assert F.h2ma == pytest.approx(20.735530)
    with pytest.raises(KeyError):
        O.init_forces("nope")


# ---------------------------------------------------------------------------
# Densities
# ---------------------------------------------------------------------------
def test_density_plane_wave(O):
    """a e^{ikr} chi with weight w: rho = w|a|^2, tau = k^2 rho, j = k rho,
    s = w|a|^2 chi^+ sigma chi, spin-orbit density J = k x s."""
    n, d = 8, 1.0
    G = O.init_grids(n, n, n, d, d, d)
    k = kvec(n, d)
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = rng(7)
    for m in ((1, 2, 3), (3, 0, 6), (0, 7, 0)):
        chi = cplx(r, 2)
        chi /= np.linalg.norm(chi)
        a, w = 0.37, 1.3
        kk = np.array([k[m[0]], k[m[1]], k[m[2]]])
        psi = (a * chi[:, None, None, None] * np.exp(2j * math.pi * (m[0] * I + m[1] * J + m[2] * K) / n))[None]
        rho, chi_d, tau, cur, sd, so = O.density_contributions(G, np.array([w]), np.array([0.4]), psi)
        rho0 = w * a * a
        s = rho0 * np.array([2 * np.real(np.conj(chi[0]) * chi[1]), 2 * np.imag(np.conj(chi[0]) * chi[1]),
                             abs(chi[0]) ** 2 - abs(chi[1]) ** 2])
        assert np.allclose(rho, rho0, atol=1e-14)
        assert np.allclose(chi_d, 0.5 * 0.4 * a * a, atol=1e-14)
        assert np.allclose(tau, kk @ kk * rho0, atol=1e-12)
        Jv = np.cross(kk, s)
        for c in range(3):
            assert np.allclose(cur[c], kk[c] * rho0, atol=1e-12)
            assert np.allclose(sd[c], s[c], atol=1e-14)
            assert np.allclose(so[c], Jv[c], atol=1e-12)


def test_density_norm_sum(O):
    """wxyz * sum rho = sum_i w_i ||psi_i||^2 for any states."""
    n, d = 8, 1.0
    G = O.init_grids(n, n, n, d, d, d)
    r = rng(8)
    psi = cplx(r, (4, 2, n, n, n))
    w = r.random(4)
    rho = O.density_contributions(G, w, np.zeros(4), psi)[0]
    assert abs(np.sum(rho) - np.sum(w * np.sum(np.abs(psi) ** 2, axis=(1, 2, 3, 4)))) < 1e-10 * np.sum(rho)


# ---------------------------------------------------------------------------
# Coulomb
# ---------------------------------------------------------------------------
def test_poisson_direct_sum(O):
    """FFT on the doubled grid = direct sum e2 wxyz sum rho(r') G(r - r'),
    G(0) = 2.84 / (dx dy dz)^(1/3), G(r) = 1/|r| (coulomb.py:14-75)."""
    n, d = 6, 0.9
    G = O.init_grids(n, n, n, d, d, d)
    C = O.init_coulomb(G)
    rho = np.zeros((2, n, n, n))
    rho[1] = rng(9).random((n, n, n))
    w = O.poisson(G, C, rho)
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    diff = (idx[:, None, :] - idx[None, :, :]) * d
    dist = np.sqrt(np.sum(diff ** 2, axis=-1))
    Gm = np.where(dist > 0, 1.0 / np.where(dist > 0, dist, 1.0), 2.84 / d)
    direct = (1.4399784085965135 * d ** 3 * Gm @ rho[1].reshape(-1)).reshape(n, n, n)
    assert np.max(np.abs(w - direct)) < 1e-12 * np.max(np.abs(direct))


# ---------------------------------------------------------------------------
# Subspace algebra
# ---------------------------------------------------------------------------
def test_loewdin(O):
    """U = S^{-1/2}: Hermitian, (psi U)^H (psi U) wxyz = I, U^2 S = I."""
    r = rng(10)
    K, nst, wxyz = 60, 6, 0.7
    psi2 = cplx(r, (K, nst))
    U, norm = O.loewdin_orthonormalize(psi2, wxyz, nst)
    S = psi2.conj().T @ psi2 * wxyz
    assert np.allclose(norm, np.real(np.diag(S)))
    assert np.allclose(U, U.conj().T, atol=1e-12)
    q = psi2 @ U
    assert np.allclose(q.conj().T @ q * wxyz, np.eye(nst), atol=1e-12)
    assert np.allclose(U @ U @ S, np.eye(nst), atol=1e-10)


def test_loewdin_clamps_null_direction(O):
    """A linearly dependent state is clamped, not amplified to inf/nan."""
    r = rng(11)
    psi2 = cplx(r, (40, 4))
    psi2[:, 3] = psi2[:, 1]
    U, _ = O.loewdin_orthonormalize(psi2, 1.0, 4)
    assert np.all(np.isfinite(U))


def test_diagstep_diagonalises_subspace(O):
    """Orthonormal states with h psi = psi M (M Hermitian): diagstep with
    diagonalize=True returns sp_energy = eigvalsh(M) ascending, a diagonal
    H, Hermitian lambda+ and lagrange = psi_new lambda+ (static.py:261-577)."""
    n, d, nst = 4, 1.0, 5
    G = O.init_grids(n, n, n, d, d, d)
    F = O.init_forces("SV-bas")
    lv = O.init_levels(G, 3, 2, [nst, nst])
    F, st = O.init_static(F, lv)
    r = rng(12)
    Ms = {}
    for iq in (0, 1):
        sl = lv.block(iq)
        Q, _ = np.linalg.qr(cplx(r, (2 * n ** 3, nst)))
        q = Q / math.sqrt(G.wxyz)
        M = cplx(r, (nst, nst))
        M = M + M.conj().T
        Ms[iq] = M
        shape = lv.psi[sl].shape
        lv.psi[sl] = O.from2d(q, shape)
        lv.hampsi[sl] = O.from2d(q @ M, shape)
        lv.hmfpsi[sl] = O.from2d(q @ M, shape)
    E = O.Energies()
    O.diagstep(E, F, G, lv, st, diagonalize=True, construct=True)
    for iq in (0, 1):
        sl = lv.block(iq)
        ev = np.linalg.eigvalsh(Ms[iq])
        assert np.allclose(lv.sp_energy[sl], ev, atol=1e-10)
        H = st.hmatrix[iq, :nst, :nst]
        assert np.allclose(H, np.diag(ev), atol=1e-10)
        p2 = O.as2d(lv.psi[sl])
        assert np.allclose(p2.conj().T @ p2 * G.wxyz, np.eye(nst), atol=1e-10)
        lam = st.lambda_save[iq, :nst, :nst]
        assert np.allclose(lam, lam.conj().T)
        assert np.allclose(O.as2d(lv.lagrange[sl]), p2 @ lam, atol=1e-10)
        # lambda = H diag(w): with occupations 1 and 0 the asymmetric part is
        # half the off-diagonal coupling between occupied and empty states (0 here)
        assert E.efluct1q[iq] < 1e-9


# ---------------------------------------------------------------------------
# BCS
# ---------------------------------------------------------------------------
def _levels_for_bcs(O, e, dl, nneut, nprot):
    S = len(e)
    lv = O.init_levels(O.init_grids(2, 2, 2, 1.0, 1.0, 1.0), nneut, nprot, [S // 2, S - S // 2])
    lv.sp_energy = np.array(e, float)
    lv.deltaf = np.array(dl, float)
    return lv


def test_bcs_degenerate_closed_form(O):
    """nst degenerate levels at e with gap D and N particles: v^2 = N/nst, so
    lambda = e + D (2v^2-1)/sqrt(1-(2v^2-1)^2) and epair = nst D sqrt(v2-v4)/2."""
    nst, N, e0, D = 6, 4.0, -10.0, 1.5
    v2 = N / nst
    c = 2 * v2 - 1
    lam = e0 + D * c / math.sqrt(1 - c * c)
    ef = O.find_eferm(N, np.full(nst, e0), np.full(nst, D), np.ones(nst))
    assert abs(ef - lam) < 1e-12
    lv = _levels_for_bcs(O, [e0] * nst + [e0 - 3] * nst, [D] * (2 * nst), 4, 4)
    F = O.init_forces("SV-bas", ipair=5)
    P = O.Pairs()
    wocc, uv, _, _ = O.pairdn(0, N, lv, F, P)
    assert np.allclose(wocc, v2, atol=1e-12)
    assert np.allclose(uv, math.sqrt(v2 - v2 * v2), atol=1e-12)
    assert abs(P.epair[0] - nst * D * math.sqrt(v2 - v2 * v2) / 2) < 1e-11
    assert abs(P.avdelt[0] - D) < 1e-12


def test_bcs_matches_root(O):
    """Non-degenerate spectrum: the particle number at eferm is N."""
    r = rng(14)
    e = np.sort(r.uniform(-40, 5, 20))
    dl = r.uniform(0.5, 2.0, 20)
    ef = O.find_eferm(8.0, e, dl, np.ones(20))
    assert abs(O.bcs_particle_number(ef, e, dl, np.ones(20)) - 8.0) < 1e-10


# ---------------------------------------------------------------------------
# Single-particle properties, observables, initial states
# ---------------------------------------------------------------------------
def test_sp_properties_closed_forms(O):
    n, d, a = 24, 0.6, 1.5
    G = O.init_grids(n, n, n, d, d, d)
    F = O.init_forces("SV-bas")
    lv = O.init_levels(G, 3, 3, [3, 3])
    X, Y, Z = grid_coords(G)
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    k = kvec(n, d)
    g = np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / a ** 2)
    lv.psi[0, 0] = np.exp(2j * math.pi * (2 * I + 3 * J + 8 * K) / n)
    # neutron 1: Gaussian along +x spin, x*y*z odd parity
    lv.psi[1, 0] = g * X * Y * Z
    lv.psi[1, 1] = g * X * Y * Z
    # neutron 2: (x + i y) g, L_z = 1, spin along +y
    lv.psi[2, 0] = (X + 1j * Y) * g
    lv.psi[2, 1] = 1j * (X + 1j * Y) * g
    # protons: even-parity Gaussian spin down
    lv.psi[3, 1] = g
    lv.psi[4, 1] = g * X * X
    lv.psi[5, 0] = g * Z
    for s in range(6):
        lv.psi[s] /= math.sqrt(O.rpsnorm(lv.psi[s], G.wxyz))
    O.sp_properties(F, G, lv)
    kk2 = k[2] ** 2 + k[3] ** 2 + k[8] ** 2
    assert abs(lv.sp_kinetic[0] - F.h2m[0] * kk2) < 1e-10
    assert np.allclose(lv.sp_spin[0], [0, 0, 0.5], atol=1e-12)
    assert np.allclose(lv.sp_spin[1], [0.5, 0, 0], atol=1e-12)
    assert np.allclose(lv.sp_spin[2], [0, 0.5, 0], atol=1e-12)
    assert np.allclose(lv.sp_spin[3], [0, 0, -0.5], atol=1e-12)
    assert abs(lv.sp_parity[1] + 1) < 1e-12 and abs(lv.sp_parity[3] - 1) < 1e-12
    assert abs(lv.sp_parity[4] - 1) < 1e-12 and abs(lv.sp_parity[5] + 1) < 1e-12
    assert np.allclose(lv.sp_orbital[2], [0, 0, 1], atol=1e-8)
    assert np.allclose(lv.sp_orbital[3], [0, 0, 0], atol=1e-10)
    # Gaussian exp(-r^2/a^2): <p^2> = 3/a^2, with the proton h2m
    assert lv.sp_kinetic[3] / F.h2m[1] == pytest.approx(3.0 / a ** 2, rel=1e-8)


def test_observables_gaussian(O):
    """Deformed, shifted Gaussian densities: particle number, c.m., rms and
    Q20 = sqrt(5/16pi) N (az^2 - a^2), Q22 = 0."""
    n, d = 24, 0.8
    G = O.init_grids(n, n, n, d, d, d)
    X, Y, Z = grid_coords(G)
    D = O.init_densities(G)
    pars = [(0.1, 1.5, 1.9, 0.3), (0.08, 1.6, 1.6, -0.2)]
    for iq, (A, a, az, c) in enumerate(pars):
        D.rho[iq] = A * np.exp(-((X - c) ** 2 + Y ** 2) / a ** 2 - Z ** 2 / az ** 2)
    ob = O.observables(D, G, 8, 8)
    for iq, (A, a, az, c) in enumerate(pars):
        N = A * math.pi ** 1.5 * a * a * az
        assert ob["pnr"][iq] == pytest.approx(N, rel=1e-12)
        assert ob["cm"][0, iq] == pytest.approx(c, abs=1e-12)
        # box truncation of the tails is ~1e-11 relative at this box size
        assert ob["rms"][iq] == pytest.approx(math.sqrt(a * a + az * az / 2), rel=1e-9)
        assert ob["q20"][iq] == pytest.approx(math.sqrt(5 / (16 * math.pi)) * N * abs(az * az - a * a), abs=1e-8)
        assert abs(ob["q22"][iq]) < 1e-10


def test_harmosc_shells(O):
    """Shell order by brute-force enumeration: ka = i+j+k, then k, j, i
    ascending, each entry twice (two spins); harmosc states normalised with
    alternating spin, the first 8 of each block orthonormal."""
    trip = sorted(((i, j, k) for i in range(6) for j in range(6) for k in range(6) if i + j + k <= 4),
                  key=lambda t: (sum(t), t[2], t[1], t[0]))
    expect = [t for t in trip for _ in range(2)]
    assert O.shell_quantum_numbers(60) == expect[:60]
    G = O.init_grids(16, 16, 16, 1.0, 1.0, 1.0)
    lv = O.init_levels(G, 8, 8, [10, 10])
    F, st = O.init_static(O.init_forces("SV-bas"), lv, radinx=2.5, radiny=2.5, radinz=2.5)
    lv = O.harmosc(G, lv, st)
    p = lv.psi.reshape(20, -1)
    gram = p.conj() @ p.T * G.wxyz
    assert np.allclose(np.diag(gram), 1.0)
    assert np.allclose(gram[:8, :8], np.eye(8), atol=1e-12)
    assert np.all(np.abs(lv.psi[1, 0]) == 0) and np.all(np.abs(lv.psi[2, 1]) == 0)
    assert lv.wocc.sum() == 16


def test_add_noise_seeded(O):
    psi = np.ones((3, 2, 4, 4, 4), complex)
    a = O.add_noise(psi, 1e-3, 0)
    b = O.add_noise(psi, 1e-3, 0)
    c = O.add_noise(psi, 1e-3, 1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    dev = (a - psi) / 1e-3
    assert 0.5 < np.std(dev.real) * math.sqrt(2) < 1.5


# ---------------------------------------------------------------------------
# Whole iterations
# ---------------------------------------------------------------------------
def test_oracle_golden_o16(O):
    """Regression against the fixture written by tests/golden/make_golden.py."""
    ref = json.loads((GOLDEN / "o16_oracle.json").read_text())
    st = O.setup(**ref["config"])
    O.prelude(st)
    got = [dict(ehf=st.energies.ehf, ehfint=st.energies.ehfint)]
    for i in range(1, len(ref["iterations"])):
        O.iteration(st, i)
        got.append(dict(ehf=st.energies.ehf, ehfint=st.energies.ehfint, efluct1=float(st.energies.efluct1[0])))
    for a, b in zip(got, ref["iterations"]):
        for k in a:
            assert abs(a[k] - b[k]) < 1e-9 * max(1.0, abs(b[k])), (k, a[k], b[k])


def test_energy_decreases_o16(O):
    """The damped gradient iteration lowers ehf monotonically once it is past
    the first steps, and ehf approaches ehfint."""
    st = O.setup(name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, b=1.8)
    hist = O.statichf(st, maxiter=25)
    assert all(b <= a + 1e-6 for a, b in zip(hist[5:], hist[6:]))
    assert abs(st.energies.ehf - st.energies.ehfint) < 0.5
