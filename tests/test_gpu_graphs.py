"""HIP-graph replay of SCF iterations reproduces eager execution.

Each replay is followed by a host read of the energies (a synchronisation)
and compared with an eager run from the same initial states; the runs cross
the inibcs/inidiag switch and the mprint fluctuation iterations so several
captured signatures are replayed repeatedly."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def make(pkg, precision, ipair=0):
    cfg = dict(name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, b=1.8, mprint=3, inibcs=5, inidiag=5)
    if ipair:
        cfg.update(ipair=5, tbcs=True, npsi=[12, 12])
    r = pkg.driver.HFRun.from_config(precision=precision, **cfg)
    r.prelude()
    return r


@pytest.mark.parametrize("mode", ["second", "first", "pre"])
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_graph_replay_matches_eager(pkg, dev, mode, precision):
    eager = make(pkg, precision)
    graphed = make(pkg, precision)
    graphed.enable_graphs(True, capture_first=(mode == "first"))
    if mode == "pre":
        graphed.precapture(range(1, 15))
    for i in range(1, 15):
        eager.iteration(i)
        graphed.iteration(i)
        a, b = eager.energies_host(), graphed.energies_host()
        for k in ("ehf", "ehfint", "efluct1", "efluct2", "sumflu", "delesum"):
            assert a[k] == b[k], (mode, i, k, a[k], b[k])
    assert torch.equal(eager.levels.psi, graphed.levels.psi)
    assert len(graphed._graphs) >= 3


def test_graph_replay_pairing(pkg, dev):
    eager = make(pkg, "f64", ipair=5)
    graphed = make(pkg, "f64", ipair=5)
    graphed.enable_graphs(True)
    for i in range(1, 14):
        eager.iteration(i)
        graphed.iteration(i)
        a, b = eager.energies_host(), graphed.energies_host()
        assert a["ehf"] == b["ehf"], (i, a["ehf"], b["ehf"])
    assert torch.equal(eager.levels.wocc, graphed.levels.wocc)


@pytest.mark.parametrize("precision,ipair", [("mixed", 0), ("f64", 5)])
def test_deferred_energy_tail(pkg, dev, precision, ipair):
    """The energy tail left to the next iteration's first h psi launch (and
    run again by the read at the end) gives the trajectory of running it at
    the end of every iteration, bit for bit."""
    runs = {}
    for defer in (False, True):
        r = make(pkg, precision, ipair=ipair)
        for i in range(1, 9):
            _, r.derivs = pkg.static.iterate(i, *r.state, r.tbcssav, r.derivs, check=False, defer_tail=defer)
        assert r.energies.stale == defer
        runs[defer] = (r.energies_host(), r.levels.psi.clone(), r.levels.sp_energy.clone())
        assert not r.energies.stale
    a, b = runs[False], runs[True]
    for k, v in a[0].items():
        assert np.array_equal(np.asarray(v), np.asarray(b[0][k])), (k, v, b[0][k])
    assert torch.equal(a[1], b[1])
    assert torch.equal(a[2], b[2])


def test_grouped_graph_replay_matches_eager(pkg, dev):
    """Runs of iterations captured into one graph chain (precapture_groups,
    run(..., group=)) reproduce eager execution bit for bit."""
    eager = make(pkg, "mixed")
    graphed = make(pkg, "mixed")
    graphed.enable_graphs(True)
    graphed.precapture(range(1, 4))
    graphed.precapture_groups(4, 12, 4)
    eager.run(15, start=1, check_every=0)
    graphed.run(3, start=1, check_every=0)
    graphed.run(12, start=4, check_every=0, group=4)
    assert len(graphed._groups_n) >= 2
    a, b = eager.energies_host(), graphed.energies_host()
    for k in ("ehf", "ehfint", "efluct1", "efluct2", "sumflu", "delesum"):
        assert a[k] == b[k], (k, a[k], b[k])
    assert torch.equal(eager.levels.psi, graphed.levels.psi)
