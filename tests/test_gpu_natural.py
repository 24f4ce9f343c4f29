"""hfb_natural (static.py:920-1070) and lastdiag (static.py:1072-1168) on the
HIP path against the oracle, FP64: the configuration-space sub-iterations on
identical inputs (states, matrices, pairing data; the QR follows LAPACK's
Householder convention on both sides, so the states compare directly), whole
HFB iterations that run them, and the quasi-particle energies after them.
Tolerances: 1e-10 relative for states and matrices, 1e-7 MeV for energies."""
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


CFG = dict(name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, npsi=[12, 12], b=1.8, ipair=5, tbcs=False)
LEVELS = ("psi", "lagrange", "hampsi", "hmfpsi", "delpsi", "sp_energy", "deltaf", "wocc", "wguv", "pairwg", "wstates",
          "sp_efluct1", "sp_efluct2")


def _copy_to_oracle(ref, run):
    for nm in LEVELS:
        setattr(ref.levels, nm, getattr(run.levels, nm).cpu().numpy().copy())
    for nm in ("hmatrix", "gapmatrix", "symcond", "lambda_save"):
        setattr(ref.static, nm, getattr(run.static, nm).cpu().numpy().copy())
    for nm in ("eferm", "epair", "avdelt", "avdeltv2", "avg"):
        setattr(ref.pairs, nm, getattr(run.pairs, nm).cpu().numpy().copy())
    ref.pairs.firstcall = run.pairs.firstcall
    for nm in ("upot", "bmass", "aq", "spot", "wlspot", "v_pair"):
        setattr(ref.meanfield, nm, getattr(run.meanfield, nm).cpu().numpy().copy())


def test_hfb_natural_matches_oracle(pkg, O, dev):
    ref = O.setup(**CFG)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, **CFG)
    run.prelude()
    for i in range(1, 4):
        run.iteration(i)
    torch.cuda.synchronize()
    _copy_to_oracle(ref, run)
    ref.params.iteration = run.params.iteration = 50
    ref.forces.tbcs = run.forces.tbcs = False
    for iq in range(2):
        O.hfb_natural(ref.forces, ref.grids, ref.levels, ref.meanfield, ref.params, ref.static, ref.pairs, iq, 6)
        pkg.static.hfb_natural(run.forces, run.grids, run.levels, run.meanfield, run.params, run.static, iq, 6,
                               pairs=run.pairs)
    torch.cuda.synchronize()
    for nm in ("psi", "lagrange", "sp_energy", "deltaf", "wocc", "wguv"):
        assert rel(getattr(run.levels, nm).cpu().numpy(), getattr(ref.levels, nm)) < 1e-10, nm
    for nm in ("hmatrix", "gapmatrix", "lambda_save"):
        assert rel(getattr(run.static, nm).cpu().numpy(), getattr(ref.static, nm)) < 1e-10, nm
    # symcond is the small antisymmetric remainder of terms of the size of h:
    # round-off on the scale of h (absolute, MeV)
    scale = np.max(np.abs(ref.static.hmatrix))
    d = np.max(np.abs(run.static.symcond.cpu().numpy() - ref.static.symcond))
    assert d < 1e-13 * scale * 12, (d, scale)
    for nm in ("eferm", "epair", "avdelt", "avdeltv2", "avg"):
        assert rel(getattr(run.pairs, nm).cpu().numpy(), getattr(ref.pairs, nm)) < 1e-10, nm


def test_hfb_iterations_with_natural_orbitals(pkg, O, dev):
    """HFB mode past iternat_start (inibcs = inidiag = iternat_start = 2,
    iternat = 4): every iteration from the third runs the sub-iterations;
    then lastdiag's quasi-particle energies and norms."""
    kw = dict(CFG, inibcs=2, inidiag=2, iternat_start=2, iternat=4)
    ref = O.setup(**kw)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, **kw)
    O.prelude(ref)
    run.prelude()
    for i in range(1, 6):
        O.iteration(ref, i)
        run.iteration(i)
        e = run.energies_host()
        print("natural it", i, e["ehf"], ref.energies.ehf, flush=True)
        assert abs(e["ehf"] - ref.energies.ehf) < 1e-7, (i, e["ehf"], ref.energies.ehf)
    assert rel(run.levels.wocc.cpu().numpy(), ref.levels.wocc) < 1e-8
    for iq in range(2):
        qe, qn = pkg.static.lastdiag(run.forces, run.grids, run.levels, run.params, run.static, run.pairs, iq)
        ee, en = O.lastdiag(ref.forces, ref.grids, ref.levels, ref.params, ref.static, ref.pairs, iq)
        qe, qn = qe.cpu().numpy(), qn.cpu().numpy()
        assert np.allclose(qe, ee, atol=1e-7), (iq, qe - ee)
        # eigenvectors of (near-)degenerate levels (time-reversed pairs) are
        # free within their cluster: compare the summed squared norms
        cuts = np.nonzero(np.diff(ee) > 1e-5)[0] + 1
        for a, b in zip(np.r_[0, cuts], np.r_[cuts, len(ee)]):
            assert abs(np.sum(qn[a:b] ** 2) - np.sum(en[a:b] ** 2)) < 1e-6, (iq, a, b)
