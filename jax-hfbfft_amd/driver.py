"""Run driver: builds a static HF(+BCS) calculation from a configuration and
runs the reference's loop on the device.

Precision modes
  "f64":   complex128 states, every kernel in FP64.
  "mixed": complex64 states; FFT derivatives, h psi, the step and the
           density transforms in FP32 (FP64 partial sums); Gram matrices,
           rotations, Jacobi eigensolves, potentials, energies and BCS in
           FP64.  ``promote()`` switches a run to complex128 for the final
           FP64 refinement iterations.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from . import static as st_mod
from .coulomb import init_coulomb
from .densities import init_densities
from .energies import init_energies
from .forces import init_forces
from .grids import init_grids
from .levels import init_levels
from .meanfield import init_meanfield
from .moment import init_moment
from .output import observables
from .pairs import Pairs
from .params import init_params


def add_noise(psi, amplitude=1e-3, seed=0):
    """psi_i += a * max|psi_i| * (xi + i eta)/sqrt2 with xi then eta drawn as
    whole arrays from numpy PCG64(seed) (build-defined; the reference's
    harmosc has no noise)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xi = rng.standard_normal(psi.shape)
    eta = rng.standard_normal(psi.shape)
    amax = np.max(np.abs(psi).reshape(psi.shape[0], -1), axis=1)[:, None, None, None, None]
    return psi + amplitude * amax * (xi + 1j * eta) / np.sqrt(2.0)


def deformed_radii(b, beta2):
    """Gaussian widths radin = sqrt2 b, stretched along z for an axial
    quadrupole bias beta2 to first order (build-defined)."""
    r = math.sqrt(2.0) * b
    c = math.sqrt(5.0 / (4.0 * math.pi)) * beta2
    return r * (1.0 - 0.5 * c), r * (1.0 - 0.5 * c), r * (1.0 + c)


# Workloads named by the benchmark configurations (nucleus, grid, states).
CONFIGS = {
    "O16": dict(name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, b=1.8),
    "Ca40": dict(name="SV-bas", n=32, d=1.0, nneut=20, nprot=20, b=2.0),
    "Cr48": dict(name="SV-bas", n=32, d=1.0, nneut=24, nprot=24, npsi=[36, 36], b=2.0, beta2=0.2, ipair=5,
                 tbcs=True),
    "Pb208": dict(name="SV-bas", n=48, d=0.8, nneut=126, nprot=82, b=2.5),
}


class _Segmenter:
    """Captures one iteration into a chain of HIP graphs, cut around timed
    library calls (kernels._timed) so replays can be bracketed by events."""

    def __init__(self, pool):
        self.pool = pool
        self.graphs = []
        self.labels = []
        self._begin(None)

    def _begin(self, label):
        g = torch.cuda.CUDAGraph()
        g.capture_begin(pool=self.pool, capture_error_mode="relaxed")
        self.graphs.append(g)
        self.labels.append(label)

    def cut(self, label):
        self.graphs[-1].capture_end()
        self._begin(label)

    def close(self):
        self.graphs[-1].capture_end()


class HFRun:
    def __init__(self, params, forces, grids, levels, meanfield, densities, coulomb, static, energies, pairs, moment):
        self.params, self.forces, self.grids, self.levels = params, forces, grids, levels
        self.meanfield, self.densities, self.coulomb, self.static = meanfield, densities, coulomb, static
        self.energies, self.pairs, self.moment = energies, pairs, moment
        self.derivs = None
        self.tbcssav = forces.tbcs
        self.iteration_no = 0
        self.graphs_enabled = False
        self.capture_first = False

    @classmethod
    def from_config(cls, precision="f64", device="cuda:0", name="SV-bas", n=24, d=1.0, nneut=8, nprot=8, npsi=None,
                    b=1.8, beta2=0.0, ipair=0, tbcs=False, v0=None, noise=1e-3, seed=0, mprint=10, psi0=None,
                    constraints=None, **static_kw):
        if precision not in ("f64", "mixed"):
            raise ValueError(f"precision must be 'f64' or 'mixed', not {precision!r}")
        dev = torch.device(device)
        if dev.type != "cuda":
            raise RuntimeError(f"HFRun runs on a HIP device (got {dev}); there is no CPU path")
        _lib.require_gpu()
        torch.cuda.set_device(dev)
        params = init_params(imode=1, mprint=mprint)
        forces = init_forces(params, name=name, ipair=ipair, tbcs=tbcs, v0prot=v0, v0neut=v0)
        grids = init_grids(params, nx=n, ny=n, nz=n, dx=d, dy=d, dz=d)
        dtype = torch.complex128 if precision == "f64" else torch.complex64
        levels = init_levels(grids, device=dev, dtype=dtype, nneut=nneut, nprot=nprot,
                             npsi=npsi if npsi is not None else [nneut, nprot])
        rx, ry, rz = deformed_radii(b, beta2)
        forces, static = st_mod.init_static(forces, levels, radinx=rx, radiny=ry, radinz=rz, **static_kw)
        coulomb = init_coulomb(grids, dev)
        if constraints:
            from .constraint import init_constraints
            static.constraints = init_constraints(constraints, nneut + nprot, grids, dev)
        if psi0 is None:
            psi0 = st_mod.harmosc_host(grids, levels.npsi, nneut, nprot, (rx, ry, rz))
            if noise:
                psi0 = add_noise(psi0, noise, seed)
        levels.psi.copy_(torch.from_numpy(np.ascontiguousarray(psi0)).to(dtype))
        meanfield = init_meanfield(grids, dev, f32=(precision == "mixed"))
        energies = init_energies(dev)
        pairs = Pairs()
        # the pairing energies live in their slots of the energy vector, so the
        # energy kernels read them without a copy
        from .kernels import EN
        pairs.epair = energies.vec[EN["EPAIR"]:EN["EPAIR"] + 2]
        return cls(params, forces, grids, levels, meanfield, init_densities(grids, dev), coulomb, static,
                   energies, pairs, init_moment(torch.zeros(3, dtype=torch.float64, device=dev)))

    @property
    def state(self):
        return (self.coulomb, self.densities, self.energies, self.forces, self.grids, self.levels, self.meanfield,
                self.moment, self.params, self.static, self.pairs)

    def prelude(self):
        self.derivs = st_mod.prelude(*self.state)
        self.tbcssav = self.forces.tbcs

    def signature(self, i):
        """Host-side switches of iteration i; iterations with equal signatures
        launch identical kernel sequences (static.py:1759-1786, 99-122, pairs.py:135)."""
        s, p, f = self.static, self.params, self.forces
        tbcs = True if i <= s.inibcs else self.tbcssav
        tdiag = tbcs or not (i > s.inidiag)
        return (tbcs, tdiag, i > 1 and not tbcs, p.mprint > 0 and i % p.mprint == 0,
                f.ipair != 0 and i <= 10, self.levels.psi.dtype, s.x0dmp, self.energies.tail is not None)

    def iteration(self, i, check=False):
        self.iteration_no = i
        # the adaptive step size (tvaryx_0) reads the energies on the host
        # after every iteration and changes x0dmp: such iterations run eagerly
        if not self.graphs_enabled or self.static.tvaryx_0:
            conv, self.derivs = st_mod.iterate(i, *self.state, self.tbcssav, self.derivs, check=check)
            return conv
        sig = self.signature(i)
        g = self._graphs.get(sig)
        if g is None and (sig in self._seen or self.capture_first):
            g = self._capture(i)
            self._graphs[sig] = g
        if g is None:
            self._seen.add(sig)
            conv, self.derivs = st_mod.iterate(i, *self.state, self.tbcssav, self.derivs, check=check)
            return conv
        # replay: keep the host-side switches in step with the captured iteration
        self.params.iteration = i
        self.forces.tbcs = sig[0]
        self.static.tdiag = sig[1]
        # the captured iteration left its energy tail to the next one
        self.energies.defer_tail(*g.tail)
        from . import kernels
        for graph, label in zip(g.graphs, g.labels):
            if label is not None and kernels.TIMERS is not None:
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record()
                graph.replay()
                ev[1].record()
                kernels.TIMERS.setdefault(label, []).append(ev)
            else:
                graph.replay()
        if check:
            e = self.energies.host()
            return e["efluct1"] < self.static.serr and i > 1
        return False

    def enable_graphs(self, on=True, capture_first=False):
        """Replay each distinct iteration signature from HIP graphs.  By
        default a signature is captured on its second occurrence (the first
        runs eagerly and allocates its workspaces); with capture_first the
        first occurrence is captured directly (new workspaces then come from
        the graph pool).  With the adaptive step size (static.tvaryx_0) every
        iteration runs eagerly: it reads the energies on the host and
        changes x0dmp after each one (static.py:1870-1882)."""
        self.graphs_enabled = on
        self.capture_first = capture_first
        if not hasattr(self, "_graphs"):
            self._graphs, self._seen, self._groups_n = {}, set(), {}
            self._pool = torch.cuda.graph_pool_handle()

    def precapture(self, iterations):
        """Capture the graphs of the given iteration numbers ahead of time
        (capture records kernels without running them, so no state changes)."""
        if self.static.tvaryx_0:
            return
        saved = (self.params.iteration, self.forces.tbcs, self.static.tdiag)
        tail = (self.energies.tail, self.energies.stale, self.energies._tail_ctx)
        for i in iterations:
            sig = self.signature(i)
            if sig not in self._graphs:
                self._graphs[sig] = self._capture(i)
        self.params.iteration, self.forces.tbcs, self.static.tdiag = saved
        self.energies.tail, self.energies.stale, self.energies._tail_ctx = tail

    def _capture(self, i, count=1):
        """Capture iterations i .. i+count-1 into one chain of HIP graphs."""
        from . import kernels
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            seg = _Segmenter(self._pool)
            kernels.SEGMENTER = seg
            try:
                for j in range(i, i + count):
                    _, self.derivs = st_mod.iterate(j, *self.state, self.tbcssav, self.derivs, check=False)
            finally:
                kernels.SEGMENTER = None
                seg.close()
        seg.tail = (self.energies._tail_ctx, self.energies.tail)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        return seg

    def _group_key(self, i, count):
        return tuple(self.signature(j) for j in range(i, i + count))

    def precapture_groups(self, start, count, group):
        """Capture the graphs that run(count, start, check_every=0,
        group=group) replays: runs of `group` consecutive iterations, one
        graph chain each (fewer graph launches; the same kernels)."""
        if group <= 1 or self.static.tvaryx_0:
            return
        saved = (self.params.iteration, self.forces.tbcs, self.static.tdiag)
        tail = (self.energies.tail, self.energies.stale, self.energies._tail_ctx)
        if self.energies.tail is None:  # the state every iteration after the first starts in
            from . import kernels
            ctx = self.grids.ctx
            self.energies.defer_tail(ctx, kernels.tail_job(ctx, self.forces.to_c(self.params, self.levels.mass_number),
                                                           self.densities, self.levels, self.energies.vec))
        for i in range(max(start, 2), start + count - group + 1, group):
            key = self._group_key(i, group)
            if key not in self._groups_n:
                self._groups_n[key] = self._capture(i, group)
        self.params.iteration, self.forces.tbcs, self.static.tdiag = saved
        self.energies.tail, self.energies.stale, self.energies._tail_ctx = tail

    def run(self, maxiter, start=1, check_every=10, callback=None, group=1):
        """Iterations start .. start+maxiter-1.  With graphs enabled,
        check_every == 0 and no callback, runs of `group` iterations whose
        graph chain was captured (precapture_groups) replay as one."""
        i, end = start, start + maxiter
        while i < end:
            if (group > 1 and self.graphs_enabled and check_every == 0 and callback is None and i + group <= end
                    and self.energies.tail is not None):
                g = self._groups_n.get(self._group_key(i, group))
                if g is not None:
                    last = i + group - 1
                    sig = self.signature(last)
                    self.iteration_no = last
                    self.params.iteration = last
                    self.forces.tbcs = sig[0]
                    self.static.tdiag = sig[1]
                    for graph in g.graphs:
                        graph.replay()
                    self.energies.defer_tail(*g.tail)
                    i += group
                    continue
            chk = check_every > 0 and (i % check_every == 0)
            conv = self.iteration(i, check=chk)
            if callback is not None:
                callback(i, self)
            if conv:
                return i
            i += 1
        return end - 1

    def promote(self):
        """Switch a mixed run to complex128 states and FP64 kernels."""
        lv = self.levels
        if lv.psi.dtype == torch.complex128:
            return
        for name in ("psi", "hampsi", "lagrange", "hmfpsi", "delpsi"):
            setattr(lv, name, getattr(lv, name).to(torch.complex128))
        if hasattr(self, "_graphs"):  # captured graphs refer to the old buffers
            # retired graphs stay referenced: dropping every graph of a pool
            # would release the pool that later captures still use
            self._retired = getattr(self, "_retired", []) + list(self._graphs.values()) + list(
                self._groups_n.values())
            self._graphs, self._seen, self._groups_n = {}, set(), {}

    def energies_host(self):
        torch.cuda.synchronize()
        return self.energies.host()

    def observables(self):
        return observables(self.densities, self.grids, self.levels.nneut, self.levels.nprot)
