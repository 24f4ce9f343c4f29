"""CPU pins of the oracle's plane-wave start, state ordering and complete
observables (static.py:807-918, 1338-1392; output.py:177-354).

  planewave_order, planewaves .... test_planewave_order, test_planewaves_are_kinetic_eigenstates
  sort_states .................... test_sort_states_moves_every_field
  complete_observables ........... test_complete_observables_gaussians
"""
import math

import numpy as np
import pytest


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def test_planewave_order(O):
    """Energies non-decreasing; (0,0,0) first, then the six |k| = 1 vectors in
    generation order (kx outermost), then the twelve |k|^2 = 2 ones."""
    G = O.init_grids(8, 8, 8, 1.0, 1.0, 1.0)
    F = O.init_forces("SV-bas")
    order = O.planewave_order(G, F.nucleon_mass)
    assert len(order) == 8 ** 3  # range(-7 // 2, 7 // 2 + 1): -4 .. 3 per axis (static.py:818-821)
    k2 = [sum(c * c for c in k) for k in order]
    assert all(a <= b for a, b in zip(k2, k2[1:]))
    assert order[0] == (0, 0, 0)
    assert order[1:7] == [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 1), (0, 1, 0), (1, 0, 0)]
    assert all(k2[i] == 2 for i in range(7, 19))


def test_planewaves_are_kinetic_eigenstates(O):
    """Orthonormal states, spin up then down per wave number, each an
    eigenstate of the kinetic operator: sp_kinetic = h2m_q |k|^2 with
    k_a = 2 pi m_a / (n d) (inout.py:64-149)."""
    n, d = 8, 1.1
    G = O.init_grids(n, n, n, d, d, d)
    F = O.init_forces("SV-bas")
    lv = O.init_levels(G, 6, 5, [8, 7])
    lv = O.planewaves(G, lv, F)
    p = lv.psi.reshape(lv.nstmax, -1)
    for iq in range(2):
        sl = lv.block(iq)
        g = p[sl].conj() @ p[sl].T * G.wxyz
        assert np.allclose(g, np.eye(sl.stop - sl.start), atol=1e-12)
    assert np.all(lv.psi[0, 1] == 0) and np.all(lv.psi[1, 0] == 0) and np.all(lv.psi[8, 1] == 0)
    O.sp_properties(F, G, lv)
    order = O.planewave_order(G, F.nucleon_mass)
    for iq in range(2):
        sl = lv.block(iq)
        for m in range(sl.stop - sl.start):
            k = order[m // 2]
            k2 = sum((2 * math.pi * c / (n * d)) ** 2 for c in k)
            assert lv.sp_kinetic[sl.start + m] == pytest.approx(F.h2m[iq] * k2, abs=1e-10)
    assert lv.wocc.sum() == 11 and lv.wocc[6] == 0 and lv.wocc[8 + 4] == 1 and lv.wocc[8 + 5] == 0


def test_sort_states_moves_every_field(O):
    """Per block, sp_energy ascending (stable for ties) and every moved field
    follows the same permutation; the other block and hampsi stay as they were."""
    G = O.init_grids(4, 4, 4, 1.0, 1.0, 1.0)
    F = O.init_forces("SV-bas")
    lv = O.init_levels(G, 4, 3, [6, 5])
    r = _rng(3)
    S = lv.nstmax
    for name in O.SORTED_FIELDS:
        a = getattr(lv, name)
        setattr(lv, name, r.standard_normal(a.shape) + (1j * r.standard_normal(a.shape) if a.dtype == complex else 0))
    lv.sp_energy[2] = lv.sp_energy[4]  # a tie
    lv.hampsi = r.standard_normal(lv.psi.shape) + 0j
    before = {name: getattr(lv, name).copy() for name in O.SORTED_FIELDS + ("hampsi",)}
    O.sort_states(F, G, lv, 0)
    e = lv.sp_energy[:6]
    assert np.all(np.diff(e) >= 0)
    # the tie keeps its original order: state 2's psi comes before state 4's
    j2 = [j for j in range(6) if np.array_equal(lv.psi[j], before["psi"][2])][0]
    j4 = [j for j in range(6) if np.array_equal(lv.psi[j], before["psi"][4])][0]
    assert j4 == j2 + 1
    perm = np.argsort(before["sp_energy"][:6], kind="stable")
    for name in O.SORTED_FIELDS:
        assert np.array_equal(getattr(lv, name)[:6], before[name][:6][perm]), name
        assert np.array_equal(getattr(lv, name)[6:], before[name][6:]), name
    assert np.array_equal(lv.hampsi, before["hampsi"])


def test_complete_observables_gaussians(O):
    """Shifted axial Gaussians: second moments a^2/2, isovector dipole
    cm_n - cm_p along x, gamma = 0 for a prolate shape, q20T1 from the
    per-isospin q20, <L>, <S> as occupation-weighted sums."""
    n, d = 24, 0.8
    G = O.init_grids(n, n, n, d, d, d)
    X, Y, Z = np.meshgrid(G.x, G.y, G.z, indexing="ij")
    D = O.init_densities(G)
    pars = [(0.1, 1.5, 1.9, 0.3), (0.08, 1.6, 1.9, -0.2)]
    for iq, (A, a, az, c) in enumerate(pars):
        D.rho[iq] = A * np.exp(-((X - c) ** 2 + Y ** 2) / a ** 2 - Z ** 2 / az ** 2)
    lv = O.init_levels(G, 8, 8, [10, 10])
    r = _rng(5)
    lv.sp_orbital = r.standard_normal((20, 3))
    lv.sp_spin = r.standard_normal((20, 3))
    lv.wocc = np.where(np.arange(20) % 3 == 0, 0.0, r.uniform(0.1, 1.0, 20))
    lv.wstates = r.uniform(0.5, 1.0, 20)
    ob = O.complete_observables(D, lv, G)
    for iq, (A, a, az, c) in enumerate(pars):
        assert ob["x2m"][0, iq] == pytest.approx(a * a / 2, rel=1e-9)
        assert ob["x2m"][2, iq] == pytest.approx(az * az / 2, rel=1e-9)
    assert ob["dipole_isovector"][0] == pytest.approx(0.3 - (-0.2), rel=1e-9)
    assert abs(ob["dipole_isovector"][1]) < 1e-12 and abs(ob["dipole_isovector"][2]) < 1e-12
    assert ob["gamma"] == pytest.approx(0.0, abs=1e-6)
    assert ob["q20T1"] == pytest.approx(-ob["q20"][0] / 8 + ob["q20"][1] / 8)
    w = lv.wocc * lv.wstates
    assert np.allclose(ob["orbital_momentum"], (w[:, None] * lv.sp_orbital).sum(axis=0))
    assert np.allclose(ob["total_momentum"], ob["orbital_momentum"] + ob["spin_momentum"])
    cmtot = (ob["pnr"][0] * 0.3 + ob["pnr"][1] * -0.2) / ob["pnrtot"]
    assert ob["cmtot"][0] == pytest.approx(cmtot, abs=1e-12)
