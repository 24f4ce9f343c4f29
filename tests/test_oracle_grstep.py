"""Closed-form pins of the oracle's gradient step (static.py:93-259), CPU only.

States are built in a basis of plane waves phi_k, which are exact eigenvectors
of h for constant potentials (U + B k^2, checked by
test_hpsi_plane_wave_eigenvalue) and of the preconditioner (divisor
max(w,.1)(E0 + h2ma k^2) + max(wuv,.1) vmax/2, checked by
test_laplace_plane_wave).  With psi = phi C (C unitary, nst x nst) and a
Lagrange field lagrange = psi M, every quantity grstep produces has a closed
form in the coefficients:

  ps1_j        = sum_k (w_j eps_k - wuv_j vp) C_kj phi_k
  h_exp_j      = sum_k |C_kj|^2 (w_j eps_k - wuv_j vp)
  residual_j   = ps1_j - (C M)_j            iteration > 1 and not tbcs
               = ps1_j - h_exp_j psi_j      otherwise
  psi'_j       = psi_j - x0 sum_k r_kj / den_k(j) phi_k
  efluct1_j    = |r_j|,  efluct2_j = |C M_j - h_exp_j C_j|  (every mprint-th iteration)

with the hpsi weights w_j = wstates_j wocc_j, wuv_j = wstates_j wguv_j pairwg_j
(HFB branch) or (1, 0) (BCS branch), and the preconditioner's own weights
(wocc_j, wguv_j pairwg_j, max v_pair) or (1, 0, 0).  These pin the choice of
residual (Lagrange field vs expectation value), the weights of both branches
and the e0dmp <= 0 step, which the whole-iteration tests cannot tell apart.
"""
import math

import numpy as np
import pytest

N, D = 8, 1.0
U_Q = (-40.0, -35.0)
B_Q = (20.0, 21.5)
VP_Q = (3.0, 2.2)


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _kphys(m):
    m = m if m < N // 2 else m - N
    return 2 * math.pi * m / (N * D)


# distinct non-Nyquist wave vectors (Nyquist index N/2 is avoided)
MODES = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (7, 0, 0), (1, 1, 0), (0, 7, 1), (2, 1, 7), (3, 0, 5)]


def _unitary(r, n):
    a = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    q, rr = np.linalg.qr(a)
    return q * (np.diag(rr) / np.abs(np.diag(rr)))[None, :]


def _setup(O, npsi=(6, 5), seed=0, mixed=True):
    G = O.init_grids(N, N, N, D, D, D)
    F = O.init_forces("SV-bas")
    lv = O.init_levels(G, 4, 3, list(npsi))
    F, st = O.init_static(F, lv)
    mf = O.init_meanfield(G)
    for iq in range(2):
        mf.upot[iq] = U_Q[iq]
        mf.bmass[iq] = B_Q[iq]
        mf.v_pair[iq] = VP_Q[iq]
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    r = _rng(seed)
    basis = {}
    for iq in range(2):
        sl = lv.block(iq)
        nst = sl.stop - sl.start
        phis, eps, k2 = [], [], []
        for s in range(nst):
            m = MODES[s]
            spin = r.standard_normal(2) + 1j * r.standard_normal(2)
            spin /= np.linalg.norm(spin)
            ph = np.exp(2j * math.pi * (m[0] * I + m[1] * J + m[2] * K) / N)
            phi = spin[:, None, None, None] * ph / math.sqrt(G.wxyz * N ** 3)
            kk = sum(_kphys(c) ** 2 for c in m)
            phis.append(phi)
            eps.append(U_Q[iq] + B_Q[iq] * kk)
            k2.append(kk)
        phis = np.array(phis)
        C = _unitary(r, nst) if mixed else np.eye(nst, dtype=complex)
        lv.psi[sl] = np.einsum("kj,k...->j...", C, phis)
        Mm = r.standard_normal((nst, nst)) + 1j * r.standard_normal((nst, nst))
        Mm = 5.0 * (Mm + Mm.conj().T)
        lv.lagrange[sl] = np.einsum("kj,k...->j...", C @ Mm, phis)
        basis[iq] = dict(phi=phis, C=C, M=Mm, eps=np.array(eps), k2=np.array(k2))
    S = lv.nstmax
    lv.sp_energy = r.uniform(-30, 0, S)
    lv.sp_efluct1 = r.uniform(0, 1, S)
    lv.sp_efluct2 = r.uniform(0, 1, S)
    return G, F, lv, st, mf, basis


def _set_weights(O, lv, r):
    S = lv.nstmax
    lv.wocc = r.uniform(0.05, 0.95, S)
    lv.wguv = r.uniform(0.02, 0.5, S)
    lv.pairwg = r.uniform(0.3, 1.0, S)
    lv.wstates = r.uniform(0.5, 1.0, S)


def _expected(F, lv, st, basis, iteration, mprint, tbcs, iq):
    """psi', hampsi, hmfpsi, delpsi coefficients and the scalars of one block."""
    sl = lv.block(iq)
    b = basis[iq]
    C, M, eps, k2 = b["C"], b["M"], b["eps"], b["k2"]
    nst = C.shape[0]
    vp = VP_Q[iq]
    if tbcs:
        w = np.ones(nst)
        wuv = np.zeros(nst)
        pw, puv, vmax = np.ones(nst), np.zeros(nst), 0.0
    else:
        w = lv.wstates[sl] * lv.wocc[sl]
        wuv = lv.wstates[sl] * lv.wguv[sl] * lv.pairwg[sl]
        pw, puv, vmax = lv.wocc[sl], lv.wguv[sl] * lv.pairwg[sl], vp
    lam = w[None, :] * eps[:, None] - wuv[None, :] * vp          # (k, j) eigenvalue of the weighted h
    ps1 = lam * C
    h_exp = np.sum(np.abs(C) ** 2 * lam, axis=0)
    spe_new = np.sum(np.abs(C) ** 2 * eps[:, None], axis=0)
    CM = C @ M
    if iteration > 1 and not tbcs:
        res = ps1 - CM
    else:
        res = ps1 - h_exp[None, :] * C
    out = dict(spe_new=spe_new, hampsi=eps[:, None] * C)
    if st.e0dmp > 0.0:
        den = np.maximum(pw, 0.1)[None, :] * (st.e0dmp + F.h2ma * k2[:, None]) \
            + 0.5 * np.maximum(puv, 0.1)[None, :] * vmax
        new = C - st.x0dmp * res / den
        if mprint > 0 and iteration % mprint == 0:
            out["efl1"] = np.sqrt(np.sum(np.abs(res) ** 2, axis=0))
            out["efl2"] = np.sqrt(np.sum(np.abs(CM - h_exp[None, :] * C) ** 2, axis=0))
        else:
            out["efl1"], out["efl2"] = lv.sp_efluct1[sl].copy(), lv.sp_efluct2[sl].copy()
    else:
        new = (1.0 + st.x0dmp * (spe_new - lv.sp_energy[sl]))[None, :] * C - st.x0dmp * ps1
        out["efl1"], out["efl2"] = lv.sp_efluct1[sl].copy(), lv.sp_efluct2[sl].copy()
    out["psi"] = new
    out["hmfpsi"] = eps[:, None] * new
    out["delpsi"] = vp * new
    out["denerg"] = (lv.sp_energy[sl] - spe_new) / np.abs(spe_new)
    return out


def _coef(block, phis):
    """Coefficients of states (nst, 2, n, n, n) in the orthonormal basis phis."""
    a = phis.reshape(phis.shape[0], -1)
    b = block.reshape(block.shape[0], -1)
    return (np.conj(a) @ b.T) * (D ** 3)


def _check(O, G, F, lv, st, mf, basis, iteration, mprint, tbcs):
    exp = {iq: _expected(F, lv, st, basis, iteration, mprint, tbcs, iq) for iq in range(2)}
    old_w = lv.wocc * lv.wstates
    sumflu0, delesum0 = st.sumflu, st.delesum
    F.tbcs = tbcs
    P = O.Params(iteration=iteration, mprint=mprint)
    lv2, st2 = O.grstep(F, G, lv, mf, P, st)
    efl1 = np.concatenate([exp[0]["efl1"], exp[1]["efl1"]])
    dsum = 0.0
    for iq in range(2):
        sl = lv2.block(iq)
        phis = basis[iq]["phi"]
        e = exp[iq]
        for name in ("psi", "hampsi", "hmfpsi", "delpsi"):
            got = getattr(lv2, name)[sl]
            c = _coef(got, phis)
            assert np.allclose(c, e[name], atol=1e-11), (name, iq, np.max(np.abs(c - e[name])))
            # nothing outside the span of the plane waves
            resid = got - np.einsum("kj,k...->j...", c, phis)
            assert np.max(np.abs(resid)) < 1e-11
        assert np.allclose(lv2.sp_energy[sl], e["spe_new"], atol=1e-11)
        assert np.allclose(lv2.sp_efluct1[sl], e["efl1"], atol=1e-10)
        assert np.allclose(lv2.sp_efluct2[sl], e["efl2"], atol=1e-10)
        dsum += np.sum(old_w[sl] * e["denerg"])
    assert st2.sumflu == pytest.approx(sumflu0 + np.sum(efl1), rel=1e-12, abs=1e-12)
    assert st2.delesum == pytest.approx(delesum0 + dsum, rel=1e-12, abs=1e-12)


def test_grstep_late_phase_residual_uses_lagrange_field(O):
    """iteration > 1, not tbcs, occupations 1/0: the residual is ps1 - lagrange
    (static.py:99-104), not ps1 - h_exp psi."""
    G, F, lv, st, mf, basis = _setup(O, seed=1)
    _check(O, G, F, lv, st, mf, basis, iteration=7, mprint=7, tbcs=False)


def test_grstep_hfb_weights(O):
    """Fractional wocc, wguv, pairwg, wstates in the HFB branch: h psi weighted by
    (wstates wocc, wstates wguv pairwg) (static.py:159-166) and the
    preconditioner by (wocc, wguv pairwg, max v_pair) (static.py:128-137)."""
    G, F, lv, st, mf, basis = _setup(O, seed=2)
    _set_weights(O, lv, _rng(20))
    _check(O, G, F, lv, st, mf, basis, iteration=12, mprint=4, tbcs=False)


def test_grstep_hfb_weights_first_iteration(O):
    """iteration 1 without tbcs: HFB weights, but the residual against h_exp psi."""
    G, F, lv, st, mf, basis = _setup(O, seed=3)
    _set_weights(O, lv, _rng(30))
    _check(O, G, F, lv, st, mf, basis, iteration=1, mprint=1, tbcs=False)


def test_grstep_bcs_branch(O):
    """tbcs: weights (1, 0) in h psi and the preconditioner, residual ps1 - h_exp psi,
    whatever the occupations and the iteration number; fluctuations kept
    from the last mprint-th iteration otherwise."""
    G, F, lv, st, mf, basis = _setup(O, seed=4)
    _set_weights(O, lv, _rng(40))
    _check(O, G, F, lv, st, mf, basis, iteration=9, mprint=10, tbcs=True)


def test_grstep_eigenstates_with_consistent_lagrange_are_fixed(O):
    """Plane-wave eigenstates with lagrange = psi diag(w eps): the step is the
    identity and efluct1 = 0 (fixed point of the late phase)."""
    G, F, lv, st, mf, basis = _setup(O, seed=5, mixed=False)
    for iq in range(2):
        sl = lv.block(iq)
        b = basis[iq]
        w = lv.wstates[sl] * lv.wocc[sl]
        b["M"] = np.diag(w * b["eps"]).astype(complex)
        lv.lagrange[sl] = np.einsum("kj,k...->j...", b["C"] @ b["M"], b["phi"])
    psi0 = lv.psi.copy()
    _check(O, G, F, lv, st, mf, basis, iteration=3, mprint=1, tbcs=False)
    assert np.max(np.abs(lv.psi - psi0)) < 1e-12
    assert np.max(lv.sp_efluct1) < 1e-10


def test_grstep_simple_step_without_preconditioner(O):
    """e0dmp <= 0: psi' = (1 + x0 (spe' - spe)) psi - x0 ps1 (static.py:181-185)."""
    G, F, lv, st, mf, basis = _setup(O, seed=6)
    _set_weights(O, lv, _rng(60))
    st.e0dmp = 0.0
    st.x0dmp = 0.01
    _check(O, G, F, lv, st, mf, basis, iteration=5, mprint=5, tbcs=False)


def test_energy_decreases_late_phase(O):
    """With the switches at iteration 2 (inibcs = inidiag = 2) the late-phase
    step (Lagrange residual, no diagonalisation) lowers ehf and the largest
    single-particle residual norm sp_efluct1 over 24 iterations (small O16
    box; with every state occupied the Lagrange matrix is Hermitian and the
    asymmetry efluct1 stays zero)."""
    st = O.setup(name="SV-bas", n=16, d=1.25, nneut=8, nprot=8, b=1.8, inibcs=2, inidiag=2, mprint=1)
    e1 = []

    def cb(i, s):
        e1.append(float(np.max(s.levels.sp_efluct1)))

    hist = O.statichf(st, maxiter=24, callback=cb)
    assert all(b <= a + 1e-6 for a, b in zip(hist[3:], hist[4:])), hist
    assert e1[-1] < 0.5 * e1[3], e1


@pytest.mark.parametrize("case", ["late", "hfb", "bcs", "simple"])
def test_grstep_helper_equals_batched_grstep(O, case):
    """The per-state restatement (grstep_helper / e0dmp_gt_zero,
    static.py:93-213) gives the batched grstep's state for every state."""
    G, F, lv, st, mf, basis = _setup(O, seed=7)
    if case != "late":
        _set_weights(O, lv, _rng(70))
    if case == "simple":
        st.e0dmp, st.x0dmp = 0.0, 0.01
    F.tbcs = case == "bcs"
    P = O.Params(iteration=6, mprint=3)
    import copy
    lv0 = copy.deepcopy(lv)
    per = [O.grstep_helper(s, int(lv0.isospin[s]), lv0.sp_energy[s], lv0.psi[s], lv0.lagrange[s], F, G, lv0, mf, P,
                           st) for s in range(lv0.nstmax)]
    O.grstep(F, G, lv, mf, P, st)
    for s, (pn, pmf, hmf, dl, den, spe, e1, e2) in enumerate(per):
        assert np.max(np.abs(pn - lv.psi[s])) < 1e-12
        assert np.max(np.abs(pmf - lv.hampsi[s])) < 1e-10
        assert np.max(np.abs(hmf - lv.hmfpsi[s])) < 1e-10
        assert np.max(np.abs(dl - lv.delpsi[s])) < 1e-12
        assert abs(spe - lv.sp_energy[s]) < 1e-11
        assert abs(e1 - lv.sp_efluct1[s]) < 1e-11 and abs(e2 - lv.sp_efluct2[s]) < 1e-11
