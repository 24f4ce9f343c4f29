"""The plane-wave start, state ordering, result files and checkpoint/restart
(static.py:807-918, 1338-1392, 1394-1420, 1700-1701; output.py:49-354) on
the HIP path against the oracle, FP64.  Energies to 1e-7 MeV, observables to
1e-11 relative, permutations exact."""
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


CFG = dict(name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, b=1.8)


def test_planewave_start_matches_oracle(pkg, O, dev):
    """planewaves on both sides, then the prelude and three iterations."""
    ref = O.setup(**CFG, npsi=[10, 10])
    O.planewaves(ref.grids, ref.levels, ref.forces)
    run = pkg.driver.HFRun.from_config(precision="f64", **CFG, npsi=[10, 10])
    pkg.static.planewaves(run.grids, run.levels, run.params, run.static, run.forces)
    assert rel(run.levels.psi.cpu().numpy(), ref.levels.psi) < 1e-14
    assert np.array_equal(run.levels.wocc.cpu().numpy(), ref.levels.wocc)
    O.prelude(ref)
    run.prelude()
    for i in range(1, 4):
        O.iteration(ref, i)
        run.iteration(i)
        e = run.energies_host()
        assert abs(e["ehf"] - ref.energies.ehf) < 1e-7, (i, e["ehf"], ref.energies.ehf)


def test_sort_states_matches_oracle(pkg, O, dev):
    ref = O.setup(**CFG, npsi=[11, 9])
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, **CFG, npsi=[11, 9])
    rng = np.random.default_rng(7)
    lv = run.levels
    for name in O.SORTED_FIELDS:
        t = getattr(lv, name)
        a = rng.standard_normal(tuple(t.shape))
        if t.is_complex():
            a = a + 1j * rng.standard_normal(tuple(t.shape))
        t.copy_(torch.tensor(a, dtype=t.dtype, device="cuda"))
        setattr(ref.levels, name, a)
    lv.sp_energy[3] = lv.sp_energy[7]
    ref.levels.sp_energy[3] = ref.levels.sp_energy[7]
    for iq in range(2):
        O.sort_states(ref.forces, ref.grids, ref.levels, iq)
        pkg.static.sort_states(run.forces, run.grids, lv, iq)
    for name in O.SORTED_FIELDS:
        assert np.array_equal(getattr(lv, name).cpu().numpy(), getattr(ref.levels, name)), name


def test_sorted_iterations_match_oracle(pkg, O, dev):
    """tsort: the states are ordered after every potential update
    (static.py:1829-1831); two empty states per isospin make the order change."""
    ref = O.setup(**CFG, npsi=[10, 10], tsort=True)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, **CFG, npsi=[10, 10], tsort=True)
    O.prelude(ref)
    run.prelude()
    for i in range(1, 5):
        O.iteration(ref, i)
        run.iteration(i)
        e = run.energies_host()
        assert abs(e["ehf"] - ref.energies.ehf) < 1e-7, (i, e["ehf"], ref.energies.ehf)
    assert np.all(np.diff(run.levels.sp_energy[:10].cpu().numpy()) >= 0)
    assert rel(run.levels.sp_energy.cpu().numpy(), ref.levels.sp_energy) < 1e-9


def test_complete_observables_and_result_files(pkg, O, dev, tmp_path):
    """statichf with the result-file writer every second iteration: the
    observables match the oracle's on the same densities and levels, and the
    files hold them with the reference's formats."""
    cfg = dict(name="SV-bas", n=24, d=1.0, nneut=12, nprot=10, npsi=[14, 12], b=1.9, beta2=0.3)
    ref = O.setup(**cfg)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, maxiter=6, **cfg)
    out = pkg.output.FortranOutputWriter(str(tmp_path / "res"))
    pkg.static.statichf(*run.state, output_writer=out, output_interval=2)
    torch.cuda.synchronize()
    lv = run.levels
    for nm in ("wocc", "wstates", "sp_orbital", "sp_spin"):
        setattr(ref.levels, nm, getattr(lv, nm).cpu().numpy())
    ref.densities.rho = run.densities.rho.cpu().numpy()
    exp = O.complete_observables(ref.densities, ref.levels, ref.grids)
    got = pkg.output.calculate_complete_observables(run.densities, lv, run.grids, run.params)
    for k, v in exp.items():
        assert np.allclose(getattr(got, k), v, rtol=1e-11, atol=1e-11), (k, getattr(got, k), v)
    rows = {k: [l.split() for l in open(tmp_path / "res" / f"{k}.res") if not l.startswith("#")]
            for k in out.NAMES}
    assert [int(r[0]) for r in rows["conver"]] == [0, 2, 4, 6]
    e = run.energies_host()
    last = rows["energies"][-1]
    assert abs(float(last[3]) - e["ehf"]) < 1e-6 and abs(float(last[4]) - e["ehfint"]) < 1e-6
    mono = rows["monopoles"][-1]
    assert abs(float(mono[1]) - got.rms[0]) < 1e-6 and abs(float(mono[3]) - got.rmstot) < 1e-6
    quad = rows["quadrupoles"][-1]
    assert abs(float(quad[3]) - got.q20tot) < 1e-6 and abs(float(quad[9]) - got.gamma) < 1e-2
    spin = rows["spin"][-1]
    assert abs(float(spin[9]) - got.total_momentum[2]) < 1e-6


def test_checkpoint_restart_matches_oracle(pkg, O, dev, tmp_path):
    """save_checkpoint after 6 iterations; a new run loads it and restarts
    (params.trestart: prelude without the first orthonormalisation, iterations
    from 7, static.py:1700-1706); the oracle restarts from the same arrays."""
    cfg = dict(CFG, npsi=[10, 10])
    ref = O.setup(**cfg)
    run = pkg.driver.HFRun.from_config(precision="f64", psi0=ref.levels.psi, **cfg)
    run.prelude()
    for i in range(1, 7):
        run.iteration(i)
    f = str(tmp_path / "wf.npz")
    pkg.static.save_checkpoint(f, run.levels, run.params, run.static, run.pairs)
    run2 = pkg.driver.HFRun.from_config(precision="f64", **cfg)
    pkg.static.load_checkpoint(f, run2.levels, run2.params, run2.static, run2.pairs)
    assert run2.params.trestart and run2.params.iteration == 6
    for nm in ("psi", "lagrange", "wocc", "sp_energy"):
        assert torch.equal(getattr(run2.levels, nm), getattr(run.levels, nm)), nm
    d = np.load(f, allow_pickle=False)
    for nm in pkg.static.CHECKPOINT_ARRAYS:
        setattr(ref.levels, nm, d[f"levels.{nm}"].copy())
    for nm in ("hmatrix", "gapmatrix", "symcond", "lambda_save"):
        setattr(ref.static, nm, d[f"static.{nm}"].copy())
    ref.params.iteration = 6
    ref.static.x0dmp = float(d["x0dmp"])
    O.prelude(ref, restart=True)
    run2.prelude()
    e = run2.energies_host()
    assert abs(e["ehf"] - ref.energies.ehf) < 1e-8, (e["ehf"], ref.energies.ehf)
    for i in range(7, 10):
        O.iteration(ref, i)
        run2.iteration(i)
        e = run2.energies_host()
        assert abs(e["ehf"] - ref.energies.ehf) < 1e-7, (i, e["ehf"], ref.energies.ehf)
