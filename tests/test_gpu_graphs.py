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
