"""Energy-functional backward (hfb_energy_bwd behind EnergyFunctional).

Checks, on the converged densities of a Ca40 run (32^3, FP64, 1000 SCF
iterations from the oscillator start, the energy then changing by < 1e-8
MeV per iteration) and on the densities of a Pb208 run (48^3, 60 mixed-
precision iterations then 10 FP64 refinement iterations of the promoted
states):
  * coupling gradients for a batch of 8 seeded +-2% perturbations of the ten
    couplings against central finite differences (h = 1e-5 relative), as
    responses c_k dE/dc_k relative to the set's largest response;
  * density gradients against central finite differences along random
    directions;
  * density gradients against the mean fields of skyrme (meanfield.py:
    dE/drho = wxyz U, dE/dtau = wxyz B, dE/dJ = wxyz W, dE/ds = wxyz S,
    dE/dj = wxyz times A_q without its j terms).
Tolerance: 1e-6 relative for the finite differences (the requirement),
1e-11 for the mean-field identities (same arithmetic, different order)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["Ca40", "Pb208"])
def case(pkg, request):
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    cfg = dict(pkg.driver.CONFIGS[request.param])
    if request.param == "Ca40":  # converged FP64 densities
        run = pkg.driver.HFRun.from_config(precision="f64", **cfg)
        run.prelude()
        run.enable_graphs(True)
        run.run(1000, start=1, check_every=0)
        e0 = run.energies_host()["ehf"]
        run.run(1, start=1001, check_every=0)
        assert abs(run.energies_host()["ehf"] - e0) < 1e-8
    else:  # FP32 bulk, then FP64 refinement of the promoted states
        run = pkg.driver.HFRun.from_config(precision="mixed", **cfg)
        run.prelude()
        run.enable_graphs(True)
        run.run(60, start=1, check_every=0)
        run.promote()
        run.run(10, start=61, check_every=0)
    torch.cuda.synchronize()
    print(request.param, "ehf", run.energies_host()["ehf"], flush=True)
    env = pkg.adjoint.FunctionalEnv(run.grids, run.forces, run.params, run.levels.mass_number)
    return run, env


def perturbed_couplings(pkg, env, nb=8, seed=7):
    c0 = pkg.adjoint.coupling_vector(env.forces)
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(-1.0, 1.0, size=(nb, c0.numel()))
    return c0[None, :] * (1.0 + 0.02 * torch.from_numpy(u))


def test_coupling_gradient_matches_finite_differences(pkg, case):
    run, env = case
    A = pkg.adjoint
    d = run.densities
    C = perturbed_couplings(pkg, env).requires_grad_(True)
    E = A.energy(C, d, env)
    (g,) = torch.autograd.grad(E.sum(), [C])
    Cd = C.detach()
    fd = torch.zeros_like(Cd)
    for k in range(Cd.shape[1]):
        h = 1e-5 * Cd[:, k].abs()
        cp, cm = Cd.clone(), Cd.clone()
        cp[:, k] += h
        cm[:, k] -= h
        ep = A.energy(cp, d, env).detach().cpu()
        em = A.energy(cm, d, env).detach().cpu()
        fd[:, k] = (ep - em) / (2.0 * h)
    # compare the responses to relative changes, c_k dE/dc_k (the finite
    # difference is taken in relative steps), against the largest response
    # of the set: the rounding noise of E limits single small components
    resp_g = g * Cd
    resp_fd = fd * Cd
    rel = ((resp_g - resp_fd).abs().max(dim=1).values / resp_fd.abs().max(dim=1).values).max().item()
    assert rel < 1e-6, (rel, g[0].tolist(), fd[0].tolist())


def test_density_gradient_matches_finite_differences(pkg, case):
    run, env = case
    A = pkg.adjoint
    d = run.densities
    c = A.coupling_vector(env.forces).reshape(1, -1)
    out = A.functional_gradients(d, env)
    gen = torch.Generator(device="cuda").manual_seed(11)
    names = ("rho", "tau", "sodens", "current", "sdens")
    base = {k: getattr(d, k).clone() for k in names}
    for name in names:
        x = base[name]
        r = torch.rand(x.shape, generator=gen, device=x.device, dtype=torch.float64) * 2 - 1
        # multiplicative direction for rho keeps it positive (rho^alpha, rho^(1/3));
        # E is linear in j and in s (through s.curl j only), so those take
        # unit directions and a large step to rise above the rounding of E
        if name == "rho":
            direction, eps = x * r, 1e-5
        elif name in ("current", "sdens"):
            direction, eps = r, 1e-2
        else:
            direction, eps = r * x.abs().max(), 1e-5
        vals = []
        for sgn in (1.0, -1.0):
            fields = dict(base)
            fields[name] = (x + sgn * eps * direction).contiguous()
            E = A.EnergyFunctional.apply(c, fields["rho"], fields["tau"], fields["sodens"], fields["current"],
                                         fields["sdens"], env)
            vals.append(E.item())
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = torch.sum(out[name] * direction).item()
        assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-6 * abs(out["energy"].item())), (name, an, fd)


def test_density_gradient_is_the_mean_field(pkg, case):
    run, env = case
    A = pkg.adjoint
    d = run.densities
    out = A.functional_gradients(d, env)
    mf = pkg.meanfield.init_meanfield(run.grids)
    pkg.meanfield.skyrme(run.coulomb, d, run.forces, run.grids, mf, run.params, run.static,
                         mass_number=run.levels.mass_number)
    w = run.grids.wxyz
    f = run.forces

    def close(a, b):
        return (a - b).abs().max().item() <= 1e-11 * max(b.abs().max().item(), 1e-30)

    assert close(out["rho"], w * mf.upot)
    assert close(out["tau"], w * mf.bmass)
    assert close(out["sodens"], w * mf.wlspot)
    assert close(out["sdens"], w * mf.spot)
    j = d.current
    a_spin = torch.stack([mf.aq[q] + 2.0 * (f.b1 - f.b1p) * j[q] + 2.0 * f.b1 * j[1 - q] for q in range(2)])
    assert close(out["current"], w * a_spin)


def test_backward_is_linear_in_grad_output(pkg, case):
    """Two sets with upstream weights (2, -1): density gradients combine
    linearly, coupling gradients scale per set."""
    run, env = case
    A = pkg.adjoint
    d = run.densities
    C = perturbed_couplings(pkg, env, nb=2, seed=3).requires_grad_(True)
    rho = d.rho.detach().clone().requires_grad_(True)
    E = A.EnergyFunctional.apply(C, rho, d.tau, d.sodens, d.current, d.sdens, env)
    gC, grho = torch.autograd.grad((E * torch.tensor([2.0, -1.0], device=E.device, dtype=E.dtype)).sum(), [C, rho])
    singles = []
    for b in range(2):
        cb = C.detach()[b:b + 1].clone().requires_grad_(True)
        rb = d.rho.detach().clone().requires_grad_(True)
        Eb = A.EnergyFunctional.apply(cb, rb, d.tau, d.sodens, d.current, d.sdens, env)
        singles.append(torch.autograd.grad(Eb.sum(), [cb, rb]))
    assert torch.allclose(gC[0], 2.0 * singles[0][0][0], rtol=1e-12, atol=0)
    assert torch.allclose(gC[1], -singles[1][0][0], rtol=1e-12, atol=0)
    ref = 2.0 * singles[0][1] - singles[1][1]
    assert (grho - ref).abs().max().item() <= 1e-12 * ref.abs().max().item()
