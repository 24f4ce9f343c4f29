"""The reference's per-state and per-block entry points of the step and the
subspace algebra (static.py:93-213 grstep_helper / e0dmp_gt_zero,
static.py:463-482 loewdin_orthonormalize, static.py:484-577
construct_hfb_matrices) on the HIP path against the oracle functions of the
same names, FP64, on identical inputs.  Tolerances: 1e-10 relative for
states and matrices, 1e-10 absolute (MeV) for energies and fluctuations."""
import copy

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def T(x, dtype=None):
    return torch.tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


def _case(O, pkg, npsi=None, ipair=0, tbcs=False, nneut=8, nprot=8):
    cfg = dict(name="SV-bas", n=24, d=1.0, nneut=nneut, nprot=nprot, npsi=npsi, b=1.8, ipair=ipair, tbcs=tbcs)
    ref = O.setup(**cfg)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, **cfg)
    return ref, run


def _sync_state(ref, run):
    """Copy the device state (states, Lagrange field, potentials, levels
    vectors) into the oracle so both sides start from identical inputs."""
    lv, mf = run.levels, run.meanfield
    for nm in ("psi", "lagrange", "hampsi", "hmfpsi", "delpsi"):
        setattr(ref.levels, nm, getattr(lv, nm).cpu().numpy().copy())
    for nm in ("sp_energy", "sp_efluct1", "sp_efluct2", "wocc", "wguv", "pairwg", "wstates"):
        setattr(ref.levels, nm, getattr(lv, nm).cpu().numpy().copy())
    for nm in ("upot", "bmass", "aq", "spot", "wlspot", "v_pair"):
        setattr(ref.meanfield, nm, getattr(mf, nm).cpu().numpy().copy())


@pytest.mark.parametrize("phase", ["late", "bcs", "hfb-weights", "simple"])
def test_grstep_helper_matches_oracle(pkg, O, dev, phase):
    ref, run = _case(O, pkg, npsi=[10, 9])
    run.prelude()
    lv = run.levels
    rng = np.random.default_rng(3)
    if phase == "hfb-weights":
        S = lv.nstmax
        for nm, lo, hi in (("wocc", 0.1, 0.9), ("wguv", 0.05, 0.5), ("pairwg", 0.3, 1.0), ("wstates", 0.5, 1.0)):
            getattr(lv, nm).copy_(T(rng.uniform(lo, hi, S)))
        run.meanfield.v_pair.copy_(T(rng.uniform(0.0, 3.0, tuple(run.meanfield.v_pair.shape))))
    _sync_state(ref, run)
    it = 1 if phase == "bcs" else 35
    for P in (ref.params, run.params):
        P.iteration, P.mprint = it, 5
    tbcs = phase == "bcs"
    ref.forces.tbcs = run.forces.tbcs = tbcs
    if phase == "simple":
        ref.static.e0dmp = run.static.e0dmp = 0.0
        ref.static.x0dmp = run.static.x0dmp = 0.01
    for s in (0, 7, 12, 18):
        iq = 0 if s < 10 else 1
        got = pkg.static.grstep_helper(s, iq, lv.sp_energy[s], lv.psi[s], lv.lagrange[s], run.forces, run.grids, lv,
                                       run.meanfield, run.params, run.static)
        exp = O.grstep_helper(s, iq, ref.levels.sp_energy[s], ref.levels.psi[s], ref.levels.lagrange[s], ref.forces,
                              ref.grids, ref.levels, ref.meanfield, ref.params, ref.static)
        for k, (g, e) in enumerate(zip(got, exp)):
            g = g.cpu().numpy() if torch.is_tensor(g) else g
            if np.ndim(e) > 0:
                assert rel(g, e) < 1e-10, (phase, s, k, rel(g, e))
            else:
                assert abs(float(g) - float(e)) < 1e-10 * max(1.0, abs(float(e))), (phase, s, k, g, e)
        # e0dmp_gt_zero on the same state, given the weighted h psi
        if phase != "simple":
            w = 1.0 if tbcs else ref.levels.wstates[s] * ref.levels.wocc[s]
            wuv = 0.0 if tbcs else ref.levels.wstates[s] * ref.levels.wguv[s] * ref.levels.pairwg[s]
            ps1, _, _ = O.hpsi01(ref.grids, ref.meanfield, iq, w, wuv, ref.levels.psi[s])
            e_new, e1, e2 = O.e0dmp_gt_zero((s, iq, ref.levels.psi[s], ps1, ref.levels.lagrange[s], ref.forces,
                                              ref.grids, ref.levels, ref.meanfield, ref.params, ref.static))
            g_new, g1, g2 = pkg.static.e0dmp_gt_zero((s, iq, lv.psi[s], T(ps1), lv.lagrange[s], run.forces, run.grids,
                                                      lv, run.meanfield, run.params, run.static))
            assert rel(g_new.cpu().numpy(), e_new) < 1e-10
            assert abs(float(g1) - e1) < 1e-10 * max(1.0, e1) and abs(float(g2) - e2) < 1e-10 * max(1.0, e2)


@pytest.mark.parametrize("nst", [8, 30, 40, 130])
def test_loewdin_orthonormalize_matches_oracle(pkg, O, dev, nst):
    """S^{-1/2} of a non-orthonormal block (oscillator start plus noise) in the
    LDS (<= 36) and global-memory (<= 128, <= 256) instances of the solver;
    psi_2d passed as the reference's (2N, nst) view."""
    rng = np.random.default_rng(nst)
    n = 8
    K = 2 * n ** 3
    psi = rng.standard_normal((nst, K)) + 1j * rng.standard_normal((nst, K))
    psi += 0.5 * np.roll(psi, 1, axis=0)  # correlated columns: S far from I
    wxyz = 0.9
    U_ref, sp_ref = O.loewdin_orthonormalize(psi.T.copy(), wxyz, nst)
    P = T(psi)
    U, sp = pkg.static.loewdin_orthonormalize(P.T, wxyz, nst)
    assert rel(sp.cpu().numpy(), sp_ref) < 1e-12
    assert rel(U.cpu().numpy(), U_ref) < 1e-10
    # the rotated block is orthonormal
    q = (P.T @ U).cpu().numpy()
    assert np.max(np.abs(q.conj().T @ q * wxyz - np.eye(nst))) < 1e-10


@pytest.mark.parametrize("ipair,tbcs", [(0, False), (5, False), (5, True)])
def test_construct_hfb_matrices_matches_oracle(pkg, O, dev, ipair, tbcs):
    """construct_hfb_matrices with a given unitary (Loewdin times a random
    unitary) against the oracle's: matrices, sp_energy, deltaf, the
    fluctuation measures and lagrange = psi_new lambda+."""
    ref, run = _case(O, pkg, npsi=[12, 10], ipair=ipair, tbcs=tbcs)
    rng = np.random.default_rng(17)
    lv = run.levels
    S = lv.nstmax
    shp = ref.levels.psi.shape
    for nm in ("hmfpsi", "delpsi"):
        a = 0.1 * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
        setattr(ref.levels, nm, a)
        getattr(lv, nm).copy_(T(a))
    for nm, lo, hi in (("wocc", 0.1, 1.0), ("wguv", 0.0, 0.5), ("pairwg", 0.5, 1.0), ("wstates", 0.7, 1.0)):
        v = rng.uniform(lo, hi, S)
        setattr(ref.levels, nm, v)
        getattr(lv, nm).copy_(T(v))
    for iq in range(2):
        sl = ref.levels.block(iq)
        nst = sl.stop - sl.start
        psi_2d = O.as2d(ref.levels.psi[sl]).copy()
        U0, _ = O.loewdin_orthonormalize(psi_2d, ref.grids.wxyz, nst)
        q, r = np.linalg.qr(rng.standard_normal((nst, nst)) + 1j * rng.standard_normal((nst, nst)))
        Ug = U0 @ q
        old = lv.psi[sl].clone()
        ref.levels.psi[sl] = O.from2d(psi_2d @ Ug, ref.levels.psi[sl].shape)
        lv.psi[sl].copy_(T(ref.levels.psi[sl]))
        O.construct_hfb_matrices(Ug, psi_2d, ref.energies, ref.forces, ref.grids, ref.levels, ref.static, iq, sl, nst)
        pkg.static.construct_hfb_matrices(T(Ug), old.reshape(nst, -1).T, run.energies, run.forces, run.grids, lv,
                                          run.static, iq, sl.start, sl.stop, nst)
        torch.cuda.synchronize()
        for nm in ("hmatrix", "gapmatrix", "symcond", "lambda_save"):
            g = getattr(run.static, nm)[iq, :nst, :nst].cpu().numpy()
            e = getattr(ref.static, nm)[iq, :nst, :nst]
            assert np.max(np.abs(g - e)) < 1e-10 * max(1.0, np.max(np.abs(e))), (nm, iq)
        assert rel(lv.sp_energy[sl].cpu().numpy(), ref.levels.sp_energy[sl]) < 1e-10
        if ipair:
            assert rel(lv.deltaf[sl].cpu().numpy(), ref.levels.deltaf[sl]) < 1e-10
        assert rel(lv.lagrange[sl].cpu().numpy(), ref.levels.lagrange[sl]) < 1e-10
        e = run.energies.host()
        assert abs(e["efluct1"] - ref.energies.efluct1[0]) < 1e-10 * max(1.0, ref.energies.efluct1[0])
        assert abs(e["efluct2"] - ref.energies.efluct2[0]) < 1e-10 * max(1.0, ref.energies.efluct2[0])


def test_bcs_default_pairing_basis(pkg, O, dev):
    """pair on blocks of 126 and 82 states (the reference's smallest default
    pairing basis, levels.py:47-49) against the oracle's brentq solve."""
    npsi = [126, 82]
    cfg = dict(name="SV-bas", n=24, d=1.0, nneut=50, nprot=40, npsi=npsi, b=1.8, ipair=5, tbcs=True, noise=0)
    ref = O.setup(**cfg)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, **cfg)
    rng = np.random.default_rng(23)
    S = sum(npsi)
    e = np.concatenate([np.sort(rng.uniform(-45, 20, n)) for n in npsi])
    ref.levels.sp_energy = e.copy()
    run.levels.sp_energy.copy_(T(e))
    for it in (3, 12):
        ref.params.iteration = run.params.iteration = it
        d = rng.uniform(0.5, 2.0, S)
        ref.levels.deltaf = d.copy()
        run.levels.deltaf.copy_(T(d))
        O.pair(ref.levels, ref.meanfield, ref.forces, ref.params, ref.grids, ref.pairs)
        pkg.pairs.pair(run.levels, run.meanfield, run.forces, run.params, run.grids, run.pairs)
        assert rel(run.levels.wocc.cpu().numpy(), ref.levels.wocc) < 1e-12
        assert rel(run.levels.wguv.cpu().numpy(), ref.levels.wguv) < 1e-12
        for nm in ("eferm", "epair", "avdelt", "avdeltv2", "avg"):
            assert rel(getattr(run.pairs, nm).cpu().numpy(), getattr(ref.pairs, nm)) < 1e-12, nm
