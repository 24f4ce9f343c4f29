"""ctypes binding of libhfbhip.so (declared in include/hfb_hip.h).

The library is mandatory: every entry raises if it is missing or if no HIP
device is present.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_void_p
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().with_name("libhfbhip.so")

HFB_F64 = 0
HFB_F32 = 1

_lib = None


class TailJob(ctypes.Structure):
    """hfb_tail_job (include/hfb_hip.h)."""
    _fields_ = [("force", c_void_p), ("sky_part", c_void_p), ("en_part", c_void_p), ("current", c_void_p),
                ("nstates", c_int), ("wocc", c_void_p), ("wstates", c_void_p), ("sp_kinetic", c_void_p),
                ("sp_energy", c_void_p), ("sp_orbital", c_void_p), ("sp_spin", c_void_p), ("E", c_void_p)]


class WeightsJob(ctypes.Structure):
    """hfb_weights_job (include/hfb_hip.h)."""
    _fields_ = [("nstates", c_int), ("npsi0", c_int), ("tbcs", c_int), ("wocc", c_void_p), ("wstates", c_void_p),
                ("wguv", c_void_p), ("pairwg", c_void_p), ("v_pair", c_void_p), ("nvox", ctypes.c_longlong),
                ("vpair_max", c_void_p), ("w", c_void_p), ("wuv", c_void_p), ("pw", c_void_p),
                ("pshift", c_void_p), ("tail", ctypes.POINTER(TailJob))]


class FinishJob(ctypes.Structure):
    """hfb_finish_job (include/hfb_hip.h)."""
    _fields_ = [("nstates", c_int), ("hsum", c_void_p), ("hsum_ntile", c_int), ("wxyz", c_double),
                ("sp_energy", c_void_p), ("wocc", c_void_p), ("wstates", c_void_p), ("fsum", c_void_p),
                ("fsum_ntile", c_int), ("sp_efluct1", c_void_p), ("sp_efluct2", c_void_p), ("acc", c_void_p)]

# name -> (restype, argtypes); every entry returns int status
_SIGS = {
    "hfb_strerror": (ctypes.c_char_p, [c_int]),
    "hfb_ctx_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_int, c_int, c_double, c_double, c_double, c_int]),
    "hfb_ctx_destroy": (c_int, [c_void_p]),
    "hfb_cderv": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "hfb_hpsi": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                         POINTER(c_int), c_void_p]),
    "hfb_hpsi_ex": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            POINTER(c_int), c_void_p, c_void_p]),
    "hfb_hpsi_ntiles": (c_int, [c_void_p, c_int]),
    "hfb_reduce_partials": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_void_p, c_void_p]),
    "hfb_poisson": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "hfb_selftest_mfma_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "hfb_selftest_mfma_f32": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "hfb_step_ntiles": (c_int, [c_void_p, c_int]),
    "hfb_grstep_weights": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   ctypes.c_longlong, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hfb_precond_step": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_double,
                                 c_int, c_void_p, c_void_p, c_double, c_double, c_double, c_void_p, c_void_p,
                                 c_void_p]),
    "hfb_grstep_finish": (c_int, [c_int, c_void_p, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "hfb_gram_workspace": (c_int, [c_int, ctypes.c_longlong]),
    "hfb_gram": (c_int, [c_int, c_void_p, POINTER(c_void_p), c_int, c_int, ctypes.c_longlong, c_double, c_void_p,
                         c_void_p, c_void_p]),
    "hfb_gram2": (c_int, [c_int, POINTER(c_void_p), POINTER(c_void_p), c_int, POINTER(c_int), ctypes.c_longlong,
                          c_double, POINTER(c_void_p), POINTER(c_void_p), c_void_p]),
    "hfb_gram2_ex": (c_int, [c_int, POINTER(c_void_p), POINTER(c_void_p), c_int, POINTER(c_int), ctypes.c_longlong,
                             c_double, POINTER(c_void_p), POINTER(c_void_p), c_void_p, c_void_p]),
    "hfb_rotate": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, ctypes.c_longlong,
                           c_void_p]),
    "hfb_rotate2": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_longlong,
                            c_void_p]),
    "hfb_diag_small": (c_int, [c_void_p, c_int, c_int, c_int, POINTER(c_void_p), c_int, c_int, c_int, c_int, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_void_p),
                               c_void_p, c_void_p, c_double, c_void_p]),
    "hfb_poisson_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "hfb_loewdin": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hfb_hfb_matrices": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int]
                         + [c_void_p] * 15 + [c_void_p]),
    "hfb_constraint_potential": (c_int, [c_int, ctypes.c_longlong, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hfb_constraint_tune": (c_int, [c_int, ctypes.c_longlong, c_double, c_void_p, c_void_p] + [c_double] * 5 +
                            [c_void_p] * 6 + [c_void_p]),
    "hfb_constraint_correct": (c_int, [c_int, c_void_p, c_int, ctypes.c_longlong, c_int] + [c_void_p] * 4 +
                               [c_void_p]),
    "hfb_natural": (c_int, [c_int] * 7 + [c_double] * 6 + [c_void_p] * 18 + [c_void_p]),
    "hfb_lastdiag": (c_int, [c_int, c_int, c_int, c_int, c_double] + [c_void_p] * 6 + [c_void_p]),
    "hfb_density_sizes": (c_int, [c_void_p, c_int, POINTER(ctypes.c_longlong), POINTER(ctypes.c_longlong)]),
    "hfb_density": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "hfb_density_finish": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_int, c_double, c_void_p]),
    "hfb_density_sp_finish": (c_int, [c_void_p, c_int] + [c_void_p] * 7 + [c_int, c_double, c_void_p, c_int, c_int,
                                                                           c_double, c_double] + [c_void_p] * 5),
    "hfb_sp_finish": (c_int, [c_void_p, c_void_p, c_int, c_int, c_double, c_double, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p]),
    "hfb_deriv_workspace": (ctypes.c_longlong, [c_void_p]),
    "hfb_density_derivs": (c_int, [c_void_p] + [c_void_p] * 9 + [c_void_p, c_void_p]),
    "hfb_density_derivs_coulomb": (c_int, [c_void_p] + [c_void_p] * 9 + [c_void_p, c_double, c_void_p, c_void_p,
                                                                          c_void_p]),
    "hfb_meanfield_derivs": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hfb_skyrme_nblocks": (c_int, [c_void_p]),
    "hfb_skyrme": (c_int, [c_void_p, c_void_p] + [c_void_p] * 11 + [c_void_p] * 6 + [c_void_p] * 6 +
                   [c_void_p, c_void_p, c_int, c_void_p]),
    "hfb_integ_energy": (c_int, [c_void_p, c_void_p] + [c_void_p] * 8 + [c_void_p, c_void_p, c_int, c_void_p]),
    "hfb_skyrme_energy": (c_int, [c_void_p, c_void_p] + [c_void_p] * 11 + [c_void_p] * 6 + [c_void_p] * 6 +
                          [c_void_p, c_void_p, c_void_p]),
    "hfb_energy_tail": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hfb_sum_energy": (c_int, [c_int] + [c_void_p] * 7 + [c_void_p]),
    "hfb_energy_bwd_nblocks": (c_int, [c_void_p]),
    "hfb_energy_bwd": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int] + [c_void_p] * 11 + [c_void_p] * 6 +
                       [c_void_p, c_void_p]),
    "hfb_bcs": (c_int, [c_int, c_int, c_double, c_double, POINTER(c_double), POINTER(c_double), c_double, c_double] +
                [c_void_p] * 13 + [c_void_p]),
}

EXPORTED = tuple(_SIGS)


def lib():
    """Load libhfbhip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(python jax-hfbfft_amd/_build.py)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().hfb_strerror(rc).decode()
        raise RuntimeError(f"libhfbhip {what}: {msg} ({rc})")


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("libhfbhip needs a HIP device; torch.cuda.is_available() is False")


def ptr(t):
    return None if t is None else c_void_p(t.data_ptr())


def stream_handle():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def prec_of(dtype) -> int:
    if dtype in (torch.complex128, torch.float64):
        return HFB_F64
    if dtype in (torch.complex64, torch.float32):
        return HFB_F32
    raise TypeError(f"unsupported dtype {dtype}")


def real_of(prec: int):
    return torch.float64 if prec == HFB_F64 else torch.float32


def cplx_of(prec: int):
    return torch.complex128 if prec == HFB_F64 else torch.complex64


def need(t, dtype=None, shape=None, name="array"):
    """Validate a device tensor argument before it crosses the C ABI."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be on a HIP device")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} has dtype {t.dtype}, expected {dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t


class Context:
    """Owns one hfb_ctx (twiddles, derivative matrices, Coulomb kernel)."""

    def __init__(self, nx, ny, nz, dx, dy, dz, periodic=False, device=None):
        require_gpu()
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.periodic = bool(periodic)
        h = c_void_p()
        check(lib().hfb_ctx_create(ctypes.byref(h), self.device.index, self.nx, self.ny, self.nz,
                                   self.dx, self.dy, self.dz, int(self.periodic)), "ctx_create")
        self.handle = h
        self.n2 = (self.nx, self.ny, self.nz) if self.periodic else (2 * self.nx, 2 * self.ny, 2 * self.nz)
        self._work = {}

    @property
    def nvox(self):
        return self.nx * self.ny * self.nz

    def work(self, key, shape, dtype):
        """Persistent scratch tensors owned by the Python side (allocated once)."""
        t = self._work.get(key)
        if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
            t = torch.empty(shape, dtype=dtype, device=self.device)
            self._work[key] = t
        return t

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            try:
                _lib.hfb_ctx_destroy(h)
            except Exception:
                pass
            self.handle = None


_CTX = {}


def context_for(grids) -> Context:
    """One Context per (device, grid geometry)."""
    require_gpu()
    key = (torch.cuda.current_device(), grids.nx, grids.ny, grids.nz, float(grids.dx), float(grids.dy),
           float(grids.dz), bool(grids.periodic))
    c = _CTX.get(key)
    if c is None:
        c = Context(grids.nx, grids.ny, grids.nz, grids.dx, grids.dy, grids.dz, grids.periodic)
        _CTX[key] = c
    return c
