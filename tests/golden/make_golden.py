"""Writes o16_oracle.json: the oracle's ehf/ehfint after the prelude and
after each of the first SCF iterations of O16 (SV-bas, 24^3, dx = 1 fm,
b = 1.8 fm, seeded noise 1e-3).  A regression fixture for the oracle itself
(tests/test_oracle.py::test_oracle_golden_o16) and a CPU-free reference for
the GPU parity tests.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import hfb_oracle as O  # noqa: E402

CONFIG = dict(name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, b=1.8)
NITER = 5


def main():
    st = O.setup(**CONFIG)
    O.prelude(st)
    its = [dict(ehf=st.energies.ehf, ehfint=st.energies.ehfint)]
    for i in range(1, NITER + 1):
        O.iteration(st, i)
        e = st.energies
        its.append(dict(ehf=e.ehf, ehfint=e.ehfint, efluct1=float(e.efluct1[0])))
    ob = O.observables(st.densities, st.grids, 8, 8)
    out = dict(config=CONFIG, iterations=its,
               observables_after=dict(rms=ob["rms"].tolist(), q20=ob["q20"].tolist(), pnr=ob["pnr"].tolist()),
               sp_energy_after=st.levels.sp_energy.tolist())
    (Path(__file__).resolve().parent / "o16_oracle.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
