"""Long FP64 trajectories on the HIP path against the oracle's, through the
committed fixtures of tests/golden/make_golden.py (the oracle needs about a
second per O16 iteration on the host, the device about a millisecond).

* O16, 200 iterations with the default switches inibcs = inidiag = 30: the
  first phase (BCS-branch weights, subspace diagonalisation) and, from
  iteration 31, 170 iterations of the Lagrange residual without
  diagonalisation (static.py:99-104, 1763-1771).  ehf and ehfint to 1e-7
  MeV every iteration, the largest sp_efluct1 of every mprint-th iteration
  to 1e-6 relative, observables (output.py:177-293) and the sorted
  single-particle energies at iteration 200.
* O16 with two empty states per isospin and the adaptive step size
  (tvaryx_0, static.py:1743-1744, 1870-1882), 12 iterations: ehf and the
  x0dmp chosen after every iteration; eager and graph-enabled runs (the
  latter runs such iterations eagerly).
* output.observables against oracle.observables on identical densities."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _psi0(O, cfg):
    st = O.setup(**cfg)
    return st.levels.psi


def test_o16_200_iterations_match_oracle(pkg, O, dev):
    ref = json.loads((GOLDEN / "o16_trajectory.json").read_text())
    cfg = ref["config"]
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=_psi0(O, cfg), **cfg)
    run.prelude()
    for i, r in enumerate(ref["iterations"], start=1):
        run.iteration(i)
        e = run.energies_host()
        assert abs(e["ehf"] - r["ehf"]) < 1e-7, (i, e["ehf"], r["ehf"])
        assert abs(e["ehfint"] - r["ehfint"]) < 1e-7, (i, e["ehfint"], r["ehfint"])
        if i % run.params.mprint == 0:
            m = float(run.levels.sp_efluct1.max())
            assert abs(m - r["sp_efluct1_max"]) < 1e-6 * r["sp_efluct1_max"] + 1e-12, (i, m, r["sp_efluct1_max"])
    ob = run.observables()
    ro = ref["observables"]
    assert np.allclose(ob["pnr"], ro["pnr"], rtol=0, atol=1e-9)
    assert np.allclose(ob["rms"], ro["rms"], rtol=1e-9, atol=0)
    assert abs(ob["rmstot"] - ro["rmstot"]) < 1e-9 * ro["rmstot"]
    assert np.allclose(ob["q20"], ro["q20"], rtol=1e-6, atol=1e-7)
    assert np.allclose(ob["cm"], ro["cm"], rtol=0, atol=1e-9)
    spe = np.sort(run.levels.sp_energy.cpu().numpy())
    assert np.allclose(spe, ref["sp_energy"], rtol=0, atol=1e-7)


@pytest.mark.parametrize("graphs", [False, True])
def test_adaptive_step_size_matches_oracle(pkg, O, dev, graphs):
    ref = json.loads((GOLDEN / "o16_tvaryx.json").read_text())
    cfg = dict(ref["config"])
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=_psi0(O, cfg), **cfg)
    run.prelude()
    assert abs(run.static.x0dmp - ref["iterations"][0]["x0dmp"]) < 1e-15
    if graphs:
        run.enable_graphs(True)
        run.precapture(range(1, 13))
    for i, r in enumerate(ref["iterations"][1:], start=1):
        run.iteration(i)
        e = run.energies_host()
        assert abs(e["ehf"] - r["ehf"]) < 1e-7, (i, e["ehf"], r["ehf"])
        assert abs(run.static.x0dmp - r["x0dmp"]) < 1e-12 * r["x0dmp"], (i, run.static.x0dmp, r["x0dmp"])


def test_observables_match_oracle(pkg, O, dev):
    """Particle numbers, c.m., rms radii, Q20/Q22 and deformation of a
    deformed BCS start (Cr48 shape on a 24^3 box) on identical densities."""
    cfg = dict(name="SV-bas", n=24, d=1.0, nneut=24, nprot=24, npsi=[30, 30], b=2.0, beta2=0.3, ipair=5,
               tbcs=True)
    ref = O.setup(**cfg)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, **cfg)
    run.prelude()
    for i in (1, 2):
        run.iteration(i)
    torch.cuda.synchronize()
    ref.densities.rho = run.densities.rho.cpu().numpy().copy()
    exp = O.observables(ref.densities, ref.grids, 24, 24)
    got = run.observables()
    for k, v in exp.items():
        assert np.allclose(got[k], v, rtol=1e-11, atol=1e-11), (k, got[k], v)
    assert got["q20tot"] > 1.0  # the quadrupole bias survives: a deformed density
