"""hfb_poisson_noalloc (the Coulomb solve with the library's static device
scratch instead of a caller workspace) against hfb_poisson on the same
densities, through the C ABI: bit-identical potentials, on calls in a row on
one GPU with decreasing grids (48^3, then 32^3, then 24^3), and on 48^3
again after the smaller calls left their data in the scratch."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _rho(shape, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x = np.linspace(-1, 1, nx)[:, None, None]
    y = np.linspace(-1, 1, ny)[None, :, None]
    z = np.linspace(-1, 1, nz)[None, None, :]
    g = np.exp(-4 * (x * x + y * y + z * z))
    r = np.stack([g, g * (1.0 + 0.1 * rng.random(shape))]) * 0.08
    return torch.tensor(r, dtype=torch.float64, device="cuda")


def _solve_both(pkg, shape, d, seed):
    from jax_hfbfft_amd import _lib, kernels
    ctx = _lib.Context(*shape, d, d, d)
    rho = _rho(shape, seed)
    e2 = 1.4399784085965135
    ref = kernels.poisson(ctx, rho, e2)
    out = torch.full(shape, float("nan"), dtype=torch.float64, device="cuda")
    rc = _lib.lib().hfb_poisson_noalloc(ctx.handle, _lib.ptr(rho), _lib.ptr(out), ctypes.c_double(e2),
                                        _lib.stream_handle())
    assert rc == 0, _lib.lib().hfb_strerror(rc)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def test_noalloc_poisson_matches_workspace_entry(pkg, dev):
    for shape, d, seed in (((48, 48, 48), 0.8, 1), ((32, 32, 32), 1.0, 2), ((24, 24, 24), 0.9, 3)):
        ref, got = _solve_both(pkg, shape, d, seed)
        print("noalloc poisson", shape, float(np.max(np.abs(ref))), flush=True)
        assert np.isfinite(got).all()
        assert np.array_equal(ref, got), (shape, float(np.max(np.abs(ref - got))))
    # again on the largest grid after the smaller calls left their data behind
    ref, got = _solve_both(pkg, (48, 48, 48), 0.8, 4)
    assert np.array_equal(ref, got)


def test_noalloc_poisson_rejects_bad_arguments(pkg, dev):
    from jax_hfbfft_amd import _lib
    ctx = _lib.Context(24, 24, 24, 1.0, 1.0, 1.0)
    rho = _rho((24, 24, 24), 5)
    L = _lib.lib()
    assert L.hfb_poisson_noalloc(None, _lib.ptr(rho), _lib.ptr(rho), ctypes.c_double(1.0), None) != 0
    assert L.hfb_poisson_noalloc(ctx.handle, None, _lib.ptr(rho), ctypes.c_double(1.0), None) != 0
