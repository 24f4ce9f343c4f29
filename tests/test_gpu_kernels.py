"""Kernel-level parity of the HIP path against the FP64 oracle.

Tolerances: FP64 kernels agree with the oracle to 1e-11 relative to the
array's max magnitude; FP32 kernels to 2e-5 (FFT round-off of a 24-48 point
transform in single precision)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = {torch.complex128: 1e-11, torch.complex64: 2e-5}


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def test_mfma_f64_layout(pkg, dev):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((16, 4))
    b = rng.standard_normal((4, 16))
    d, raw = pkg.kernels.selftest_mfma_f64(torch.tensor(a, device=dev), torch.tensor(b, device=dev))
    torch.cuda.synchronize()
    ref = a @ b
    err = relerr(d.cpu().numpy(), ref)
    if err > 1e-12:
        print("raw lane layout:\n", raw.cpu().numpy())
        print("reference:\n", ref)
    assert err < 1e-12


@pytest.mark.parametrize("n", [24, 32, 48])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_cderv_all_kinds(pkg, O, dev, n, dtype):
    from jax_hfbfft_amd import _lib
    d = 0.8 if n == 48 else 1.0
    ctx = _lib.Context(n, n, n, d, d, d)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, 2, n, n, n)) + 1j * rng.standard_normal((2, 2, n, n, n))
    B = 20.0 + rng.standard_normal((n, n, n))
    rt = torch.float64 if dtype == torch.complex128 else torch.float32
    xt = torch.tensor(x, dtype=dtype, device=dev)
    Bt = torch.tensor(B, dtype=rt, device=dev)
    for axis in range(3):
        ax = axis - 3
        d1 = pkg.kernels.cderv(ctx, axis, 0, xt)
        assert relerr(d1.cpu().numpy(), O.cderv00(d, x, ax)) < TOL[dtype], ("00", axis)
        e1, e2 = pkg.kernels.cderv(ctx, axis, 1, xt)
        r1, r2 = O.cderv01(d, x, ax)
        assert relerr(e1.cpu().numpy(), r1) < TOL[dtype], ("01 d1", axis)
        assert relerr(e2.cpu().numpy(), r2) < TOL[dtype], ("01 d2", axis)
        f1, f2 = pkg.kernels.cderv(ctx, axis, 2, xt, Bt)
        r1, r2 = O.cderv02(d, x, B, ax)
        assert relerr(f1.cpu().numpy(), r1) < TOL[dtype], ("02 d1", axis)
        assert relerr(f2.cpu().numpy(), r2) < TOL[dtype], ("02 d2", axis)


def random_meanfield(O, g, rng):
    mf = O.init_meanfield(g)
    s4 = (2, g.nx, g.ny, g.nz)
    s5 = (2, 3, g.nx, g.ny, g.nz)
    mf.upot = -50.0 + 10.0 * rng.standard_normal(s4)
    mf.bmass = 20.0 + rng.standard_normal(s4)
    mf.v_pair = rng.standard_normal(s4)
    mf.aq = rng.standard_normal(s5)
    mf.spot = rng.standard_normal(s5)
    mf.wlspot = rng.standard_normal(s5)
    return mf


@pytest.mark.parametrize("n", [24, 32])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_hpsi_matches_oracle(pkg, O, dev, n, dtype):
    from jax_hfbfft_amd import _lib
    g = O.init_grids(n, n, n, 1.0, 1.0, 1.0)
    ctx = _lib.Context(n, n, n, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(7)
    mf = random_meanfield(O, g, rng)
    S, npsi0 = 3, 2
    psi = rng.standard_normal((S, 2, n, n, n)) + 1j * rng.standard_normal((S, 2, n, n, n))
    w = np.array([1.0, 0.7, 0.3])
    wuv = np.array([0.0, 0.2, 0.5])
    rt = torch.float64 if dtype == torch.complex128 else torch.float32
    T = lambda a, t=rt: torch.tensor(a, dtype=t, device=dev)
    pout, pmf, pdel, sums = pkg.kernels.hpsi(
        ctx, T(psi, dtype), npsi0, T(mf.upot), T(mf.bmass), T(mf.aq), T(mf.spot), T(mf.wlspot), T(mf.v_pair),
        T(w, torch.float64), T(wuv, torch.float64))
    for s in range(S):
        iq = 0 if s < npsi0 else 1
        r_out, r_mf, r_del = O.hpsi01(g, mf, iq, w[s], wuv[s], psi[s])
        assert relerr(pmf[s].cpu().numpy(), r_mf) < TOL[dtype], ("mf", s)
        assert relerr(pout[s].cpu().numpy(), r_out) < TOL[dtype], ("out", s)
        assert relerr(pdel[s].cpu().numpy(), r_del) < TOL[dtype], ("del", s)
        ov = np.real(np.vdot(psi[s], r_mf))
        ovo = np.real(np.vdot(psi[s], r_out))
        tol = 1e-10 if dtype == torch.complex128 else 1e-4
        assert abs(sums[s, 0].item() - ov) <= tol * max(1.0, np.abs(psi[s]).max() * np.abs(r_mf).max() * psi[s].size)
        assert abs(sums[s, 1].item() - ovo) <= tol * max(1.0, np.abs(psi[s]).max() * np.abs(r_out).max() * psi[s].size)


@pytest.mark.parametrize("n", [24, 32, 48])
def test_poisson_matches_oracle(pkg, O, dev, n):
    from jax_hfbfft_amd import _lib
    d = 0.8 if n == 48 else 1.0
    g = O.init_grids(n, n, n, d, d, d)
    coul = O.init_coulomb(g)
    ctx = _lib.Context(n, n, n, d, d, d)
    X, Y, Z = np.meshgrid(g.x, g.y, g.z, indexing="ij")
    rho = np.stack([0.08 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 8.0),
                    0.07 * np.exp(-((X - 0.5) ** 2 + Y ** 2 + (Z + 0.3) ** 2) / 9.0)])
    ref = O.poisson(g, coul, rho)
    out = pkg.kernels.poisson(ctx, torch.tensor(rho, device=dev), O.E2)
    assert relerr(out.cpu().numpy(), ref) < 1e-12
