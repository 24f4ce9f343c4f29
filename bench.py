"""Benchmark: static Skyrme-HF iterations per second on one MI355X.

Workload: Ca40 (20N, 20Z) on a 32^3 grid, dx = 1.0 fm, SV-bas couplings, no
pairing, 20+20 two-spinor states from the oscillator start (b = 2.0 fm) with
seeded noise 1e-3.  One step = one SCF iteration (the loop body of the
reference's statichf: gradient step, orthonormalisation/diagonalisation,
densities with mixing, potentials incl. Coulomb, single-particle properties,
energies).  Two runs from the same initial states: mixed precision (FP32
transforms and h psi, FP64 subspace algebra, potentials and energies) and
all-FP64; `value` is the mixed run's iterations/s over the timed steps.

Accuracy: both runs continue to the iteration where the FP64 run has
converged (ACCURACY_ITERS, or --accuracy-iters), the mixed run in mixed
precision; E_B = ehf, the rms radii, Q20 and (with pairing) the mean gaps
and Fermi energies are compared there ("unrefined"), then the mixed run is
promoted to FP64 and both runs do `--refine` more FP64 iterations before the
reported comparison.

GEMM efficiency: `gemm` times the loop's Gram and rotation calls on the
states of each run (HIP events around a replayed graph) and reports
F_gemm / (t * peak) with F_gemm = 4 * 8 * K * sum_q nst_q^2.

Energy gradient: `coupling_gradient` times the backward of the energy
functional w.r.t. the ten Skyrme couplings (a batch of 8 seeded +-2%
coupling sets) on the refined densities of the mixed run and reports its
largest deviation from central differences.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
METRIC = "HF iter/s at E_B err<1 keV vs FP64, mixed vs FP64 GPU; HBM % of peak"
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--refine", type=int, default=50)
    ap.add_argument("--workload", default="Ca40")
    ap.add_argument("--cpu-iters", type=int, default=10, help="oracle iterations timed for the CPU baseline")
    ap.add_argument("--accuracy-iters", type=int, default=None,
                    help="iterations of both runs before the observables are compared (default per workload: "
                         "where the FP64 run has converged)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--eager", action="store_true", help="launch kernels from Python instead of HIP graphs")
    ap.add_argument("--group", type=int, default=10,
                    help="timed iterations replayed per HIP graph chain (1: one graph per iteration)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy of known size: each launch reads and writes mib MiB."""
    import torch
    n = mib * 2**20 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "launches": 5}))


# iterations after which the FP64 run of each workload has converged on the
# MI355X (|delta ehf| below 1e-8 MeV per iteration, largest single-particle
# residual below 1e-5 MeV); Cr48's BCS iteration drifts slowly instead and is
# compared at the same iteration count
ACCURACY_ITERS = {"O16": 800, "Ca40": 1000, "Cr48": 1500, "Pb208": 1500}


def iter_bytes(S, nvox, c, r, tdiag=False, ipair=False):
    """Algorithmic HBM bytes of one iteration: P spinor-array passes plus
    density, potential and Coulomb traffic (B_iter = P S 2N c + F), with
    P = 11, +1 for the hampsi read of the subspace diagonalisation and +2 for
    the delpsi write and read with pairing."""
    P = 11 + (1 if tdiag else 0) + (2 if ipair else 0)
    F = (2 * 12 + 2 * 16 + 2 * 12) * nvox * r + 2 * 2 * 8 * nvox * 16
    return P * S * 2 * nvox * c + F


def window_bytes(run, args, c):
    """Algorithmic bytes of the timed iterations warmup+1 .. warmup+steps
    (subspace diagonalisation while i <= inidiag or BCS is on)."""
    st, f = run.static, run.forces
    S = run.levels.nstmax
    nvox = run.grids.nx * run.grids.ny * run.grids.nz
    tot = 0
    for i in range(args.warmup + 1, args.warmup + args.steps + 1):
        tbcs = True if i <= st.inibcs else run.tbcssav
        tdiag = tbcs or i <= st.inidiag
        tot += iter_bytes(S, nvox, c, 8, tdiag=tdiag, ipair=f.ipair != 0)
    return tot


def hpsi_bytes(S, nvox, c, r):
    """Compulsory bytes of one h psi over S states: read psi, write h psi and
    the weighted result, read 12 potential fields per isospin once."""
    return S * 2 * nvox * c * 3 + 2 * 12 * nvox * r


def measured_traffic(workload):
    """HBM bytes of one hfb_hpsi call (FP32) measured with rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE and scaled by a plain copy, as recorded in
    profiles/hpsi_<workload>_traffic.md (line `hpsi_traffic_bytes = ...`);
    None when that workload has no such record."""
    import re
    f = ROOT / "profiles" / f"hpsi_{workload.lower()}_traffic.md"
    if not f.exists():
        return None
    m = re.search(r"hpsi_traffic_bytes\s*=\s*([0-9]+)", f.read_text())
    return int(m.group(1)) if m else None


def hpsi_burst(pkg, run, reps=20):
    """Average duration of the gradient step's h psi call (same arguments,
    same states and potentials as in the loop; it rewrites the same two
    buffers with the same values) over `reps` calls captured back to back in
    one HIP graph, timed with HIP events around its replay on the launching
    stream.  This is the kernels' time without the launch of a separate graph
    segment that the in-loop timing includes."""
    import torch
    st, kernels = pkg.static, pkg.kernels
    ctx, lv = run.grids.ctx, run.levels
    ws = st._ws(ctx, lv)
    fields = run.meanfield.fields(lv.psi.dtype)
    n0 = int(lv.npsi[0])

    def calls():
        for _ in range(reps):
            kernels.hpsi(ctx, lv.psi, n0, *fields, ws["w"], ws["wuv"], want_del=False,
                         out={"pout": ws["pout"], "pout_mf": lv.hampsi}, reduce_sums=False)
    calls()  # workspaces exist before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        g.capture_begin()
        calls()
        g.capture_end()
    torch.cuda.current_stream().wait_stream(side)
    g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    g.replay()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def gemm_burst(pkg, run, reps=10):
    """Average time of the loop's GEMM-shaped work, the four kinds of
    F_gemm (static.py:301-303, 466, 496, 562): the Gram matrices
    psi^H psi and psi^H hmfpsi of both isospin blocks (one hfb_gram2 call)
    and the rotations psi U and psi (U lambda+) (one hfb_rotate2 call, into
    scratch so the states stay unchanged), on the run's own states, over
    `reps` repetitions captured in one HIP graph and timed with HIP events
    around its replay.  Returns (ms per repetition, F_gemm)."""
    import torch
    kernels = pkg.kernels
    ctx, lv = run.grids.ctx, run.levels
    n0, n1 = int(lv.npsi[0]), int(lv.npsi[1])
    sl = [slice(0, n0), slice(n0, n0 + n1)]
    blocks = [(lv.psi[s], [lv.psi[s], lv.hmfpsi[s]]) for s in sl]
    outs = [torch.empty((2, n, n), dtype=torch.complex128, device=lv.psi.device) for n in (n0, n1)]
    U = [torch.eye(n, dtype=torch.complex128, device=lv.psi.device) for n in (n0, n1)]
    o1, o2 = torch.empty_like(lv.psi), torch.empty_like(lv.psi)
    rot = [(lv.psi[s], u, u, o1[s], o2[s]) for s, u in zip(sl, U)]
    K = lv.psi[0].numel()
    scale = float(run.grids.wxyz)

    def calls():
        for _ in range(reps):
            kernels.gram2(ctx, blocks, scale, outs)
            kernels.rotate2(rot, K)
    calls()  # workspaces exist before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        g.capture_begin()
        calls()
        g.capture_end()
    torch.cuda.current_stream().wait_stream(side)
    g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    g.replay()
    b.record()
    torch.cuda.synchronize()
    flops = 4 * 8 * K * (n0 * n0 + n1 * n1)
    return a.elapsed_time(b) / reps, flops


def snapshot(run):
    """E_B and the observables of a run (output.py:177-293, pairs.py:125-127)."""
    e = run.energies_host()
    ob = run.observables()
    out = dict(ehf=e["ehf"], ehfint=e["ehfint"], rms=[float(x) for x in ob["rms"]], rmstot=float(ob["rmstot"]),
               q20=[float(x) for x in ob["q20"]], q20tot=float(ob["q20tot"]), beta=float(ob["beta"]),
               sp_efluct1_max=float(run.levels.sp_efluct1.max()))
    if run.forces.ipair != 0:
        out["avdelt"] = [float(x) for x in run.pairs.avdelt.cpu()]
        out["eferm"] = [float(x) for x in run.pairs.eferm.cpu()]
    return out


def compare(unref, ref, acc_it, refine, pairing):
    """Errors of the mixed run against the FP64 run after acc_it iterations
    (mixed precision throughout: `unrefined`) and after `refine` more FP64
    iterations of both (the reported figures)."""
    def errs(m, f):
        d = {"eb_err_kev": abs(m["ehf"] - f["ehf"]) * 1e3,
             "rms_err_fm": max(abs(a - b) for a, b in zip(m["rms"] + [m["rmstot"]], f["rms"] + [f["rmstot"]])),
             "q20_err_fm2": abs(m["q20tot"] - f["q20tot"])}
        if pairing:
            d["gap_err_mev"] = max(abs(a - b) for a, b in zip(m["avdelt"], f["avdelt"]))
            d["eferm_err_mev"] = max(abs(a - b) for a, b in zip(m["eferm"], f["eferm"]))
        return d
    out = errs(ref["mixed"], ref["f64"])
    out["unrefined"] = errs(unref["mixed"], unref["f64"])
    out["iterations"] = acc_it
    out["refine_iters"] = refine
    f = ref["f64"]
    out["f64"] = {k: f[k] for k in f}
    # convergence of the FP64 run: energy change per iteration over the
    # refinement window and the largest single-particle residual norm
    out["f64_dehf_per_iter"] = abs(f["ehf"] - unref["f64"]["ehf"]) / max(refine, 1)
    out["f64_converged"] = bool(out["f64_dehf_per_iter"] < 1e-8 and f["sp_efluct1_max"] < 1e-4)
    return out


def coupling_gradient(pkg, run, nb=8, reps=5):
    """Energy gradient w.r.t. the ten Skyrme couplings on the run's densities
    for nb sets of seeded +-2% perturbations (EnergyFunctional forward and
    backward, hfb_energy_bwd): average time of one batched evaluation with
    HIP events on torch's stream, and the largest deviation of the responses
    c_k dE/dc_k from central differences (h = 1e-5 relative) over all sets,
    relative to each set's largest response."""
    import numpy as np
    import torch
    A = pkg.adjoint
    d = run.densities
    env = A.FunctionalEnv(run.grids, run.forces, run.params, run.levels.mass_number)
    c0 = A.coupling_vector(env.forces)
    u = np.random.Generator(np.random.PCG64(7)).uniform(-1.0, 1.0, size=(nb, c0.numel()))
    C0 = c0[None, :] * (1.0 + 0.02 * torch.from_numpy(u))

    def grad():
        C = C0.clone().requires_grad_(True)
        return torch.autograd.grad(A.energy(C, d, env).sum(), [C])[0]
    g = grad()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        grad()
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    fd = torch.zeros_like(C0)
    for k in range(C0.shape[1]):
        h = 1e-5 * C0[:, k].abs()
        cp, cm = C0.clone(), C0.clone()
        cp[:, k] += h
        cm[:, k] -= h
        fd[:, k] = (A.energy(cp, d, env).detach().cpu() - A.energy(cm, d, env).detach().cpu()) / (2.0 * h)
    rg, rf = g.cpu() * C0, fd * C0
    err = ((rg - rf).abs().max(dim=1).values / rf.abs().max(dim=1).values).max().item()
    return {"sets": nb, "avg_ms": ms, "fd_rel_err": err,
            "densities": "the promoted mixed run after its FP64 refinement iterations",
            "timing": f"HIP events around {reps} batched forward+backward evaluations"}


def cpu_baseline(cfg, iters):
    sys.path.insert(0, str(ROOT))
    from oracle import hfb_oracle as O
    cores = len(os.sched_getaffinity(0))
    O.WORKERS = cores
    try:
        import threadpoolctl
        threadpoolctl.threadpool_limits(cores)
    except Exception:
        pass
    st = O.setup(**cfg)
    O.prelude(st)
    O.iteration(st, 1)  # first iteration excluded like the GPU warmup
    t0 = time.perf_counter()
    for i in range(2, 2 + iters):
        O.iteration(st, i)
    dt = time.perf_counter() - t0
    return {"value": iters / dt, "unit": "iter/s", "cores": cores, "kind": "port",
            "sample": f"{iters} SCF iterations (2..{1 + iters}) of the same workload on the FP64 NumPy/SciPy "
                      f"oracle, scipy.fft workers={cores}"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import __graft_entry__ as ge
    pkg = ge.load_package()
    drv = pkg.driver
    kernels = pkg.kernels
    cfg = dict(drv.CONFIGS[args.workload])
    dev = "cuda:0"

    runs = {}
    for prec in ("mixed", "f64"):
        r = drv.HFRun.from_config(precision=prec, device=dev, **cfg)
        r.prelude()
        if not args.eager:
            # every iteration replays a HIP graph captured ahead of time
            r.enable_graphs(True)
            r.precapture(range(1, args.warmup + args.steps + 1))
            r.precapture_groups(args.warmup + 1, args.steps, args.group)
        runs[prec] = r
    # warmup (untimed) -- iterations 1..W
    for prec, r in runs.items():
        r.run(args.warmup, start=1, check_every=0)
    torch.cuda.synchronize()

    timing = {}
    for prec, r in runs.items():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.run(args.steps, start=args.warmup + 1, check_every=0, group=1 if args.eager else args.group)
        torch.cuda.synchronize()
        timing[prec] = time.perf_counter() - t0

    burst_ms = hpsi_burst(pkg, runs["mixed"])
    gemm = {}
    for prec, peak in (("mixed", 157.3e12), ("f64", 78.6e12)):
        ms, flops = gemm_burst(pkg, runs[prec])
        gemm[prec] = {"avg_ms": ms, "frac": flops / (ms * 1e-3 * peak), "peak_tflops": peak / 1e12}
    gemm["flops"] = flops
    gemm["kinds"] = ("psi^H psi, psi^H hmfpsi (hfb_gram2), psi U, psi U lambda+ (hfb_rotate2); "
                     "F_gemm = 4 * 8 * K * sum_q nst_q^2, frac = F_gemm / (t * peak); mixed: FP32 matrix cores, "
                     "f64: FP64 peak")

    # accuracy: both runs continue (the mixed run in mixed precision) to the
    # iteration where the FP64 run has converged; then the mixed run is
    # promoted to FP64 for `refine` more iterations and the FP64 run continues
    # in step
    done = args.warmup + args.steps
    acc_it = max(done, args.accuracy_iters if args.accuracy_iters is not None
                 else ACCURACY_ITERS.get(args.workload, 1000))
    for prec, r in runs.items():
        if acc_it > done:
            r.run(acc_it - done, start=done + 1, check_every=0)
    unref = {prec: snapshot(r) for prec, r in runs.items()}
    start = acc_it + 1
    runs["mixed"].promote()
    if not args.eager:
        runs["mixed"].precapture(range(start, start + args.refine))
    for prec, r in runs.items():
        r.run(args.refine, start=start, check_every=0)
    ref = {prec: snapshot(r) for prec, r in runs.items()}
    accuracy = compare(unref, ref, acc_it, args.refine, bool(cfg.get("ipair")))
    e_mix, e_64 = ref["mixed"], ref["f64"]
    cgrad = coupling_gradient(pkg, runs["mixed"])
    err_kev = accuracy["eb_err_kev"]

    lv = runs["f64"].levels
    S = lv.nstmax
    nvox = runs["f64"].grids.nx * runs["f64"].grids.ny * runs["f64"].grids.nz
    mixed_ips = args.steps / timing["mixed"]
    f64_ips = args.steps / timing["f64"]
    ach = hpsi_bytes(S, nvox, 8, 4) / (burst_ms * 1e-3) / 1e9
    g = runs["f64"].grids
    xpass = ("hpsi_mfma_x_kernel x pass on the matrix cores" if g.nx == g.ny == g.nz == 32
             else "hpsi_lane_kernel x pass")
    roof = {"bound": "hbm", "kernel": f"hpsi ({xpass} + hpsi_plane_kernel y-z, FP32)", "achieved": ach,
            "peak": 8000.0,
            "unit": "GB/s", "frac": (ach / 8000.0) if ach else None, "traffic": measured_traffic(args.workload),
            "avg_launch_ms": burst_ms, "launches": 20,
            "timing": "HIP events around a graph of 20 back-to-back calls on the loop's states, after the timed region",
            "algorithmic_bytes": hpsi_bytes(S, nvox, 8, 4)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, args.cpu_iters)
    out = {
        "metric": METRIC, "value": mixed_ips, "unit": "iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 / mixed_ips, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{args.workload} {cfg['n']}^3 dx={cfg['d']} {cfg['name']} "
                               f"{cfg['nneut']}N+{cfg['nprot']}Z states, HO start b={cfg['b']} + noise 1e-3",
                   "states": S, "grid": cfg["n"], "refine_iters": args.refine,
                   "launch": "eager" if args.eager else f"hip-graph replay, {args.group} iterations per graph"},
        "fp64_iter_s": f64_ips, "speedup_vs_fp64": mixed_ips / f64_ips,
        "eb_err_kev": err_kev, "eb_err_kev_unrefined": accuracy["unrefined"]["eb_err_kev"],
        "ehf_mixed": e_mix["ehf"], "ehf_f64": e_64["ehf"], "accuracy": accuracy,
        "hbm_frac_iter_mixed": window_bytes(runs["f64"], args, 8) * mixed_ips / args.steps / HBM_PEAK,
        "hbm_frac_iter_f64": window_bytes(runs["f64"], args, 16) * f64_ips / args.steps / HBM_PEAK,
        "roofline": roof, "gemm": gemm, "coupling_gradient": cgrad, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
