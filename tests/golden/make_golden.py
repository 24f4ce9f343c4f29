"""Writes o16_oracle.json: the oracle's ehf/ehfint after the prelude and
after each of the first SCF iterations of O16 (SV-bas, 24^3, dx = 1 fm,
b = 1.8 fm, seeded noise 1e-3).  A regression fixture for the oracle itself
(tests/test_oracle.py::test_oracle_golden_o16) and a CPU-free reference for
the GPU parity tests; also o16_trajectory.json and o16_tvaryx.json (see
trajectory()).  Run from the repository root:
    python tests/golden/make_golden.py
"""
import json
import sys

import numpy as np
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




TRAJ_ITER = 200
TVARY_ITER = 12


def trajectory():
    """o16_trajectory.json: iterations 1..200 of O16 with the reference's
    default switches (inibcs = inidiag = 30), i.e. across the change to the
    Lagrange residual without diagonalisation at iteration 31
    (static.py:1763-1771, 99-104): ehf, ehfint, efluct1/2 and the largest
    sp_efluct1 per iteration, the observables (output.py:177-293) and the
    sorted single-particle energies at the end.  Then o16_tvaryx.json: 12
    iterations with the adaptive step size (tvaryx_0, static.py:1743-1744,
    1870-1882), ehf and x0dmp per iteration."""
    st = O.setup(**CONFIG)
    O.prelude(st)
    its = []
    for i in range(1, TRAJ_ITER + 1):
        O.iteration(st, i)
        e = st.energies
        its.append(dict(ehf=e.ehf, ehfint=e.ehfint, efluct1=float(e.efluct1[0]), efluct2=float(e.efluct2[0]),
                        sp_efluct1_max=float(st.levels.sp_efluct1.max())))
    ob = O.observables(st.densities, st.grids, 8, 8)
    out = dict(config=CONFIG, iterations=its,
               observables=dict(rms=ob["rms"].tolist(), rmstot=float(ob["rmstot"]), q20=ob["q20"].tolist(),
                                q20tot=float(ob["q20tot"]), pnr=ob["pnr"].tolist(), cm=ob["cm"].tolist(),
                                beta=float(ob["beta"])),
               sp_energy=np.sort(st.levels.sp_energy).tolist())
    (Path(__file__).resolve().parent / "o16_trajectory.json").write_text(json.dumps(out, indent=1) + "\n")

    # two empty states per isospin: the Lagrange matrix is then not Hermitian
    # and efluct1/2, which steer the step size, are well above round-off
    tcfg = dict(CONFIG, npsi=[10, 10], tvaryx_0=True)
    st = O.setup(**tcfg)
    O.prelude(st)
    its = [dict(x0dmp=st.static.x0dmp)]
    for i in range(1, TVARY_ITER + 1):
        O.iteration(st, i)
        its.append(dict(ehf=st.energies.ehf, x0dmp=st.static.x0dmp, efluct1=float(st.energies.efluct1[0]),
                        efluct2=float(st.energies.efluct2[0])))
    out = dict(config=tcfg, iterations=its)
    (Path(__file__).resolve().parent / "o16_tvaryx.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
    trajectory()
