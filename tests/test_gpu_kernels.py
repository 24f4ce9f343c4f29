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


def test_mfma_f32_layout(pkg, dev):
    """The FP32 h.psi kernels read register r of lane l of a 16x16x4 product
    as row 4*(l//16) + r, column l % 16."""
    rng = np.random.default_rng(1)
    a = rng.standard_normal((16, 4)).astype(np.float32)
    b = rng.standard_normal((4, 16)).astype(np.float32)
    raw = pkg.kernels.selftest_mfma_f32(torch.tensor(a, device=dev), torch.tensor(b, device=dev)).cpu().numpy()
    ref = a.astype(np.float64) @ b.astype(np.float64)
    want = np.array([[ref[4 * (l // 16) + r, l % 16] for r in range(4)] for l in range(64)])
    err = relerr(raw, want)
    if err > 1e-6:
        print("raw lane layout:\n", raw)
        print("reference:\n", ref)
    assert err < 1e-6


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


@pytest.mark.parametrize("n", [24, 32, 48])
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


@pytest.mark.parametrize("n", [24, 32, 48])
def test_derivs_coulomb_matches_separate(pkg, O, dev, n):
    """The combined derivative + Coulomb call (fused launches on 32^3 and
    48^3) gives the separate calls' fields bit for bit, and the Coulomb
    potential of the oracle."""
    from types import SimpleNamespace
    from jax_hfbfft_amd import _lib
    d = 0.8 if n == 48 else 1.0
    g = O.init_grids(n, n, n, d, d, d)
    coul = O.init_coulomb(g)
    ctx = _lib.Context(n, n, n, d, d, d)
    rng = np.random.default_rng(n)
    X, Y, Z = np.meshgrid(g.x, g.y, g.z, indexing="ij")
    rho = np.stack([0.08 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 8.0),
                    0.07 * np.exp(-((X - 0.5) ** 2 + Y ** 2 + (Z + 0.3) ** 2) / 9.0)])
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    dens = SimpleNamespace(rho=T(rho), sdens=T(rng.standard_normal((2, 3, n, n, n))),
                           current=T(rng.standard_normal((2, 3, n, n, n))),
                           sodens=T(rng.standard_normal((2, 3, n, n, n))))
    s4, s5 = (2, n, n, n), (2, 3, n, n, n)
    mk = lambda: {k: torch.empty(s, dtype=torch.float64, device=dev)
                  for k, s in (("divj", s4), ("lap", s4), ("grad", s5), ("curls", s5), ("curlj", s5))}
    a, b = mk(), mk()
    pkg.kernels.density_derivs(ctx, dens, a)
    wa = pkg.kernels.poisson(ctx, dens.rho, O.E2)
    wb = torch.empty((n, n, n), dtype=torch.float64, device=dev)
    pkg.kernels.density_derivs_coulomb(ctx, dens, b, O.E2, wb)
    torch.cuda.synchronize()
    for k in a:
        assert torch.equal(a[k], b[k]), k
    assert torch.equal(wa, wb)
    assert relerr(wb.cpu().numpy(), O.poisson(g, coul, rho)) < 1e-12


@pytest.mark.parametrize("nsts", [(8, 8), (20, 20), (21, 7), (36, 30), (40, 0), (64, 50), (126, 82)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_gram2_matches_matmul(pkg, dev, nsts, dtype):
    """scale * A^H B_m for both isospin blocks, with B_0 = A (Hermitian, half
    computed) and two other blocks, against numpy in FP64 on the same
    inputs.  complex128: 1e-12 of the largest entry.  complex64 (FP32 matrix
    cores, FP32 sums of at most 256 products per chunk and of 16 block sums,
    FP64 over chunks): entry (i, j) within 2e-5 |a_i| |b_j| (the worst-case
    bound of those sums), and the RMS error below 1e-6 of the RMS bound."""
    g = pkg.grids.init_grids(pkg.params.init_params(), nx=24, ny=24, nz=24, dx=1.0, dy=1.0, dz=1.0)
    ctx = g.ctx
    rng = np.random.default_rng(sum(nsts))
    blocks, outs, ref, norms = [], [], [], []
    for iq, nst in enumerate(nsts):
        arrs = [rng.standard_normal((nst, 2, 24, 24, 24)) + 1j * rng.standard_normal((nst, 2, 24, 24, 24))
                for _ in range(3)]
        ts = [torch.tensor(a, dtype=dtype, device=dev) for a in arrs]
        blocks.append((ts[0], [ts[0], ts[1], ts[2]]))
        outs.append(torch.empty((3, nst, nst), dtype=torch.complex128, device=dev))
        A2 = ts[0].cpu().numpy().astype(np.complex128).reshape(nst, 2 * 24 ** 3)
        ref.append([0.7 * A2.conj() @ t.cpu().numpy().astype(np.complex128).reshape(nst, 2 * 24 ** 3).T
                    for t in ts])
        norms.append([np.linalg.norm(t.cpu().numpy().astype(np.complex128).reshape(nst, 2 * 24 ** 3), axis=1) for t in ts])
    pkg.kernels.gram2(ctx, blocks, 0.7, outs)
    torch.cuda.synchronize()
    for iq in range(2):
        if nsts[iq] == 0:
            continue
        for m in range(3):
            r = ref[iq][m]
            got = outs[iq][m].cpu().numpy()
            if dtype == torch.complex128:
                assert np.abs(got - r).max() <= 1e-12 * np.abs(r).max(), (iq, m)
            else:
                bound = 0.7 * np.outer(norms[iq][0], norms[iq][m])
                err = np.abs(got - r)
                assert np.all(err <= 2e-5 * bound), (iq, m, float(np.max(err / bound)))
                assert np.sqrt(np.mean(err ** 2)) <= 1e-6 * np.sqrt(np.mean(bound ** 2)), (iq, m)


@pytest.mark.parametrize("nst", [8, 20, 30, 36, 64, 126])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_rotate_matches_matmul(pkg, dev, nst, dtype):
    """out1 = psi U1 and out2 = psi U2 in place against numpy on the same
    inputs.  complex128: 1e-12 of the largest entry.  complex64 (FP32
    matrix cores, U rounded to FP32, FP32 sums of the 2 nst real products):
    every entry within the recursive-summation bound (2 nst + 2) 2^-24
    sum_i |psi_i| |U_ij|."""
    rng = np.random.default_rng(nst)
    K = 2 * 24 ** 3
    psi = rng.standard_normal((nst, K)) + 1j * rng.standard_normal((nst, K))
    U1 = rng.standard_normal((nst, nst)) + 1j * rng.standard_normal((nst, nst))
    U2 = rng.standard_normal((nst, nst)) + 1j * rng.standard_normal((nst, nst))
    tp = torch.tensor(psi, dtype=dtype, device=dev)
    tu1 = torch.tensor(U1, dtype=torch.complex128, device=dev)
    tu2 = torch.tensor(U2, dtype=torch.complex128, device=dev)
    out2 = torch.empty_like(tp)
    pkg.kernels.rotate(tp, tu1, tu2, tp, out2)
    torch.cuda.synchronize()
    p64 = torch.tensor(psi, dtype=dtype).numpy().astype(np.complex128)
    for got, U in ((tp, U1), (out2, U2)):
        r = (p64.T @ U).T
        err = np.abs(got.cpu().numpy() - r)
        if dtype == torch.complex128:
            assert err.max() <= 1e-12 * np.abs(r).max()
        else:
            bound = (2 * nst + 2) * 2.0 ** -24 * (np.abs(p64).T @ np.abs(U)).T * np.sqrt(2)
            assert np.all(err <= bound), float(np.max(err / bound))


@pytest.mark.parametrize("nsts", [(20, 20), (36, 30), (126, 82), (12, 0)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_rotate2_matches_matmul(pkg, dev, nsts, dtype):
    """Both isospin blocks in one launch, in place, against numpy (tolerances
    of test_rotate_matches_matmul)."""
    rng = np.random.default_rng(sum(nsts))
    K = 2 * 24 ** 3
    blocks, refs = [], []
    for nst in nsts:
        if nst == 0:
            blocks.append(None)
            refs.append(None)
            continue
        psi = rng.standard_normal((nst, K)) + 1j * rng.standard_normal((nst, K))
        U1 = rng.standard_normal((nst, nst)) + 1j * rng.standard_normal((nst, nst))
        U2 = rng.standard_normal((nst, nst)) + 1j * rng.standard_normal((nst, nst))
        tp = torch.tensor(psi, dtype=dtype, device=dev)
        out2 = torch.empty_like(tp)
        blocks.append((tp, torch.tensor(U1, dtype=torch.complex128, device=dev),
                       torch.tensor(U2, dtype=torch.complex128, device=dev), tp, out2))
        p64 = torch.tensor(psi, dtype=dtype).numpy().astype(np.complex128)
        refs.append((p64, U1, U2))
    pkg.kernels.rotate2(blocks, K)
    torch.cuda.synchronize()
    for blk, ref, nst in zip(blocks, refs, nsts):
        if blk is None:
            continue
        p64 = ref[0]
        for got, U in ((blk[3], ref[1]), (blk[4], ref[2])):
            r = (p64.T @ U).T
            err = np.abs(got.cpu().numpy() - r)
            if dtype == torch.complex128:
                assert err.max() <= 1e-12 * np.abs(r).max()
            else:  # the FP32 recursive-summation bound, as in test_rotate_matches_matmul
                bound = (2 * nst + 2) * 2.0 ** -24 * (np.abs(p64).T @ np.abs(U)).T * np.sqrt(2)
                assert np.all(err <= bound), float(np.max(err / bound))
