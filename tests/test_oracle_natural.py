"""CPU pins of the canonical-basis HFB functions of the oracle
(static.py:920-1336: hfb_natural, pair_subiter, lastdiag) and of the
Householder convention the device kernel follows.

  QR convention (device) ... test_householder_matches_lapack_qr
  lastdiag ................. test_lastdiag_no_pairing, test_lastdiag_bcs_closed_form
  pair_subiter ............. test_pair_subiter_takes_the_brentq_fallback
  hfb_natural .............. test_natural_identity_without_gradient, test_natural_is_unitary_and_descends
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def householder_q(a):
    """The device kernel's algorithm (householder_q in csrc/hfb_linalg.hip):
    zlarfg reflectors, H_k^H from the left, Q = H_0 ... H_{n-1} backwards."""
    A = a.astype(np.complex128).copy()
    n = A.shape[0]
    tau = np.zeros(n, complex)
    for k in range(n):
        xn2 = np.sum(np.abs(A[k + 1:, k]) ** 2)
        alpha = A[k, k]
        if not (xn2 == 0.0 and alpha.imag == 0.0):
            nrm = math.sqrt(alpha.real ** 2 + alpha.imag ** 2 + xn2)
            beta = -nrm if alpha.real >= 0.0 else nrm
            tau[k] = complex((beta - alpha.real) / beta, -alpha.imag / beta)
            A[k + 1:, k] *= 1.0 / (alpha - beta)
            A[k, k] = beta
        v = np.concatenate([[1.0], A[k + 1:, k]])
        w = v.conj() @ A[k:, k + 1:]
        A[k:, k + 1:] -= np.conj(tau[k]) * np.outer(v, w)
    Q = np.eye(n, dtype=complex)
    for k in range(n - 1, -1, -1):
        v = np.concatenate([[1.0], A[k + 1:, k]])
        w = v.conj() @ Q[k:, k:]
        Q[k:, k:] -= tau[k] * np.outer(v, w)
    return Q


@pytest.mark.parametrize("n", [1, 2, 5, 12, 33])
def test_householder_matches_lapack_qr(n):
    r = _rng(n)
    a = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    a[0, 0] = -abs(a[0, 0])  # both signs of Re alpha
    q_ref, _ = np.linalg.qr(a)
    assert np.allclose(householder_q(a), q_ref, atol=1e-12)
    # already orthonormal input: the identity is kept exactly
    assert np.allclose(householder_q(np.eye(n)), np.eye(n), atol=0)


def _levels(O, npsi=(8, 6), nneut=4, nprot=3):
    G = O.init_grids(4, 4, 4, 1.0, 1.0, 1.0)
    F = O.init_forces("SV-bas", ipair=5)
    lv = O.init_levels(G, nneut, nprot, list(npsi))
    F, st = O.init_static(F, lv)
    return G, F, lv, st


def test_lastdiag_no_pairing(O):
    """Occupations 1/0, no gap: quasi-particle energies |e_i - eferm| of a
    diagonal h (hole states e < eferm, particle states above)."""
    G, F, lv, st = _levels(O)
    e = np.array([-30., -25., -20., -18., -5., -3., 2., 4.])
    st.hmatrix[0, :8, :8] = np.diag(e)
    lv.wocc[:8] = [1, 1, 1, 1, 0, 0, 0, 0]
    P = O.Pairs()
    P.eferm[0] = -10.0
    ev, nrm = O.lastdiag(F, G, lv, None, st, P, 0)
    assert np.allclose(ev, np.sort(np.abs(e + 10.0)), atol=1e-12)


def test_lastdiag_bcs_closed_form(O):
    """Diagonal h and gap with BCS occupations v^2 = (1 - (e-l)/E)/2: the
    canonical-basis matrix is diagonal with E_i = sqrt((e_i-l)^2 + D_i^2),
    and the lower norm of eigenvector j is |c_i| for the row i that maps
    onto component j (time-reversed partner)."""
    G, F, lv, st = _levels(O)
    r = _rng(4)
    e = np.sort(r.uniform(-30, 5, 8))
    D = r.uniform(0.5, 2.0, 8)
    lam = -12.0
    E = np.sqrt((e - lam) ** 2 + D ** 2)
    v2 = 0.5 * (1.0 - (e - lam) / E)
    st.hmatrix[0, :8, :8] = np.diag(e)
    st.gapmatrix[0, :8, :8] = np.diag(D)
    lv.wocc[:8] = v2
    P = O.Pairs()
    P.eferm[0] = lam
    ev, nrm = O.lastdiag(F, G, lv, None, st, P, 0)
    order = np.argsort(E, kind="stable")
    assert np.allclose(ev, E[order], atol=1e-11)
    # component j gets the coefficient of row i = partner(j): even rows take
    # -sqrt(v2) from j = i-1, odd rows +sqrt(v2) from j = i+1
    expect = np.zeros(8)
    for i in range(8):
        j = i - 1 if i % 2 == 0 else i + 1
        if 0 <= j < 8:
            expect[j] = math.sqrt(v2[i])
    assert np.allclose(nrm, expect[order], atol=1e-12)


def test_pair_subiter_takes_the_brentq_fallback(O):
    """The reference's brentq call (rtol=1e-16) is rejected by SciPy, so the
    sub-iteration pairing uses the midpoint of block levels N-1 and N."""
    with pytest.raises(ValueError):
        brentq(lambda x: x, -1.0, 1.0, xtol=1e-14, rtol=1e-16)
    G, F, lv, st = _levels(O)
    r = _rng(6)
    lv.sp_energy[:] = np.concatenate([np.sort(r.uniform(-30, 0, 8)), np.sort(r.uniform(-25, 5, 6))])
    lv.deltaf[:] = r.uniform(0.5, 1.5, 14)
    P = O.Pairs()
    O.pair_subiter(lv, F, P, 0)
    ef = 0.5 * (lv.sp_energy[3] + lv.sp_energy[4])
    assert P.eferm[0] == ef
    edif = lv.sp_energy[:8] - ef
    v2 = np.clip(0.5 * (1 - edif / np.sqrt(edif ** 2 + lv.deltaf[:8] ** 2)), 1e-6, 1 - 1e-6)
    assert np.allclose(lv.wocc[:8], v2, atol=1e-14)


def test_natural_identity_without_gradient(O):
    """One sub-iteration from symcond = 0 without pairing: QR of the identity is
    the identity, so the
    states are unchanged, htran = h and lagrange = psi h (static.py:1055-1068)."""
    G, F, lv, st = _levels(O)
    F.ipair = 0
    r = _rng(8)
    n = 8
    h = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    h = h + h.conj().T
    st.hmatrix[0, :n, :n] = h
    lv.psi[:n] = r.standard_normal(lv.psi[:n].shape) + 0j
    psi0 = lv.psi.copy()
    O.hfb_natural(F, G, lv, None, None, st, O.Pairs(), 0, 1)
    assert np.allclose(lv.psi, psi0, atol=1e-14)
    assert np.allclose(O.as2d(lv.lagrange[:n]), O.as2d(psi0[:n]) @ h, atol=1e-12)
    assert np.allclose(st.hmatrix[0, :n, :n], h, atol=1e-13)


def test_natural_is_unitary_and_descends(O):
    """With pairing: the configuration transform is unitary (orthonormal
    states stay orthonormal) and the Lagrange asymmetry symcond shrinks over
    the sub-iterations (a descent in the canonical basis)."""
    G, F, lv, st = _levels(O)
    r = _rng(9)
    n = 8
    e = np.sort(r.uniform(-30, 5, n))
    h = np.diag(e) + 0.3 * (lambda m: m + m.conj().T)(r.standard_normal((n, n)) + 1j * r.standard_normal((n, n)))
    g = np.diag(r.uniform(0.8, 1.2, n)) + 0.05 * (lambda m: m + m.conj().T)(r.standard_normal((n, n)) + 0j)
    st.hmatrix[0, :n, :n] = h
    st.gapmatrix[0, :n, :n] = g
    lv.sp_energy[:n] = np.real(np.diag(h))
    lv.deltaf[:n] = np.real(np.diag(g))
    P = O.Pairs()
    O.pair_subiter(lv, F, P, 0)
    lam = O.lambda_matrix(h, g, lv, slice(0, n))
    w = lv.wocc[:n] * lv.wstates[:n]
    wuv = lv.wguv[:n] * lv.pairwg[:n] * lv.wstates[:n]
    st.symcond[0, :n, :n] = w[None, :] * h - wuv[None, :] * g - lam
    s0 = np.linalg.norm(st.symcond[0, :n, :n])
    Q, _ = np.linalg.qr(r.standard_normal((2 * 64, n)) + 1j * r.standard_normal((2 * 64, n)))
    lv.psi[:n] = O.from2d(Q / math.sqrt(G.wxyz), lv.psi[:n].shape)
    lv.sp_energy[n:] = np.sort(r.uniform(-25, 5, 6))
    lv.deltaf[n:] = r.uniform(0.5, 1.5, 6)
    P.firstcall = False
    params = O.Params(iteration=50)
    O.hfb_natural(F, G, lv, None, params, st, P, 0, 30)
    p2 = O.as2d(lv.psi[:n])
    assert np.allclose(p2.conj().T @ p2 * G.wxyz, np.eye(n), atol=1e-12)
    assert np.linalg.norm(st.symcond[0, :n, :n]) < 0.5 * s0


# ---------------------------------------------------------------------------
# Constraints (constraint.py:49-344)
# ---------------------------------------------------------------------------
def test_constraint_fields_and_potential(O):
    """Q20 field = sqrt(pi/5)/(r0^2 A^(5/3)) (2z^2-x^2-y^2) mask: zero
    expectation for a spherical density, the principal-axis moments of a
    reflection-symmetric density vanish, and skyrme's U_q drops by exactly
    sum_i lambda_i Q_i for both isospins."""
    n, d = 16, 1.0
    G = O.init_grids(n, n, n, d, d, d)
    c = O.init_constraints(dict(alpha20_wanted=0.3, alpha22_wanted=0.1, tq_prin_axes=True), 16, G)
    assert c.numconstraint == 8 and c.tconstraint
    X, Y, Z = np.meshgrid(G.x, G.y, G.z, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    mask = 1.0 / (1.0 + np.exp(r - 6.0))
    assert np.allclose(c.constr_field[0], math.sqrt(math.pi / 5) / (0.93 ** 2 * 16 ** (5 / 3))
                       * (2 * Z ** 2 - X ** 2 - Y ** 2) * mask, atol=1e-15)
    D = O.init_densities(G)
    D.rho[0] = 0.08 * np.exp(-r ** 2 / 4.0)
    D.rho[1] = 0.07 * np.exp(-r ** 2 / 4.0)
    ex, var = O.constraint_expectation_values(D, c, G)
    assert np.all(np.abs(ex) < 1e-14) and np.all(var > 0)
    F = O.init_forces("SV-bas")
    mf0 = O.init_meanfield(G)
    mf1 = O.init_meanfield(G)
    C = O.init_coulomb(G)
    O.skyrme(C, D, F, G, mf0)
    c.lambda_crank = np.linspace(-1.0, 1.0, 8)
    O.skyrme(C, D, F, G, mf1, constraints=c)
    diff = mf1.upot - mf0.upot
    expect = -np.tensordot(c.lambda_crank, c.constr_field, axes=(0, 0))
    assert np.allclose(diff[0], expect, atol=1e-12) and np.allclose(diff[1], expect, atol=1e-12)


def test_constrained_iteration_approaches_the_goal(O):
    """A spherical O16 start driven towards Q20 = 0.3: the constrained moment
    rises towards the goal over the first iterations (constraint.py:248-318)."""
    st = O.setup(name="SV-bas", n=16, d=1.25, nneut=8, nprot=8, constraints=dict(alpha20_wanted=0.3))
    O.prelude(st)
    hist = []
    for i in range(1, 6):
        O.iteration(st, i)
        hist.append(float(st.constraints.actual_crank[0]))
    assert all(b > a for a, b in zip(hist, hist[1:])), hist
    assert 0.15 < hist[-1] < 0.35, hist
