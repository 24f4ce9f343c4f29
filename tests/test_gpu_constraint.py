"""Constrained HF runs (constraint.py:49-344; the potential in skyrme,
meanfield.py:246-248) on the HIP path against the oracle: Q20 and Q22 goals
with the principal-axis constraints, FP64 energies to 1e-7 MeV, multipliers
and moments to 1e-9 relative; the mixed-precision run (FP32 copy of U kept
in step) follows the FP64 run."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CON = dict(alpha20_wanted=0.3, alpha22_wanted=0.05, tq_prin_axes=True)
CFG = dict(name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, b=1.8)


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_constrained_iterations_match_oracle(pkg, O, dev):
    ref = O.setup(**CFG, constraints=CON)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, constraints=CON, **CFG)
    assert rel(run.static.constraints.constr_field.cpu().numpy(), ref.constraints.constr_field) < 1e-14
    O.prelude(ref)
    run.prelude()
    for i in range(1, 6):
        O.iteration(ref, i)
        run.iteration(i)
        e = run.energies_host()
        c = run.static.constraints
        print("constrained", i, e["ehf"], ref.energies.ehf, c.actual_crank.cpu().numpy()[:2], flush=True)
        assert abs(e["ehf"] - ref.energies.ehf) < 1e-7, (i, e["ehf"], ref.energies.ehf)
        assert rel(c.actual_crank.cpu().numpy(), ref.constraints.actual_crank) < 1e-9
        assert rel(c.lambda_crank.cpu().numpy(), ref.constraints.lambda_crank) < 1e-9
    s = pkg.constraint.get_constraint_summary(run.static.constraints)
    assert s["enabled"] and abs(s["Q20"]["actual"] - ref.constraints.actual_crank[0]) < 1e-9


def test_constrained_mixed_precision_follows_f64(pkg, O, dev):
    runs = {}
    psi0 = O.setup(**CFG).levels.psi
    for prec in ("f64", "mixed"):
        r = pkg.driver.HFRun.from_config(precision=prec, psi0=psi0, constraints=CON, **CFG)
        r.prelude()
        for i in range(1, 6):
            r.iteration(i)
        runs[prec] = r
    e = {p: r.energies_host()["ehf"] for p, r in runs.items()}
    q = {p: r.static.constraints.actual_crank.cpu().numpy() for p, r in runs.items()}
    print("constrained mixed - f64:", e["mixed"] - e["f64"], q["mixed"] - q["f64"])
    assert abs(e["mixed"] - e["f64"]) < 1e-3
    assert np.max(np.abs(q["mixed"] - q["f64"])) < 1e-5
