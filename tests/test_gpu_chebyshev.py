"""Chebyshev moments on one bond (mpsa_chebyshev_moments: the recurrence of
ChebychevEngine::simulateSSCompress, src/container.hpp:2343-2470, with the
exact H_eff of the plan).

Bars: the moments mu_k and the reconstructed doubled moments (2<t_k|t_k> -
mu_0, 2<t_k|t_{k-1}> - mu_1) against the closed form
sum_n |<n|t_0>|^2 T_k((E_n - shift) / scale) from the dense H_eff's
eigendecomposition (1e-11 of mu_0); the complex plan of the same bond in a
random complex gauge (unitary on the bond left of the centre) gives the real
plan's moments (1e-11 of mu_0); at the full Hubbard D = 1024 size, where the
dense form is out of reach, |mu_k| <= mu_0 and the doubling identities of a
Hermitian H_eff (1e-9 of mu_0).
"""
import numpy as np
import pytest
import torch

import complexcase as cc
import problems
from mpstools_amd import native

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _setup(P, Lb, Rb, psi):
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    dL, dR = torch.from_numpy(Lt.pack(Lb)).cuda(), torch.from_numpy(Rt.pack(Rb)).cuda()
    plan.bind_env(dL, dR)
    return plan, vt, (dL, dR)


@pytest.mark.parametrize("model,L,N,D,seed", [("tv", 12, 6, 12, 71), ("hubbard", 6, 6, 4, 72)])
def test_chebyshev_moments_match_spectral_sum(gpu, oracle, model, L, N, D, seed):
    _say("chebyshev", model, L, N, D)
    P = problems.build(oracle, model, L, N, D, seed)
    c = P["c"]
    plan, vt, keep = _setup(P, P["Ls"][c - 1], P["Rs"][c + 2], P["psi"])
    H = problems.dense_heff(oracle, P, vt)
    E, U = np.linalg.eigh(0.5 * (H + H.T))
    shift, scale = 0.5 * (E[-1] + E[0]), 0.55 * (E[-1] - E[0])
    x0 = vt.pack(P["psi"])
    w = (U.T @ x0) ** 2
    th = np.arccos(np.clip((E - shift) / scale, -1, 1))
    nmom = 60
    want = np.array([np.sum(w * np.cos(k * th)) for k in range(2 * nmom - 1)])
    mu, rc = native.chebyshev_moments(plan, torch.from_numpy(x0).cuda(), nmom, scale, shift)
    assert np.max(np.abs(mu - want[:nmom])) <= 1e-11 * want[0]
    assert np.max(np.abs(rc - want)) <= 1e-11 * want[0]


def test_chebyshev_moments_complex_gauge_invariant(gpu, oracle):
    _say("chebyshev complex gauge")
    P = problems.build(oracle, "tv", 20, 10, 60, 73)
    c, mps, mpo, itk = P["c"], P["mps"], P["mpo"], P["itk"]
    plan, vt, keep = _setup(P, P["Ls"][c - 1], P["Rs"][c + 2], P["psi"])
    x0 = torch.from_numpy(vt.pack(P["psi"])).cuda()
    G = cc.bond_unitaries(mps[c - 1], 73)
    Lg = native.add_contraction_layer(P["Ls"][c - 2], cc.gauge_right(mps[c - 1], G), mpo[c - 1], itk[c - 1], 1)
    pz, vz, keepz = _setup(P, Lg, cc.with_imag(P["Rs"][c + 2], 0, 0.0), cc.gauge_left(P["psi"], G))
    xz = torch.from_numpy(vz.pack(cc.gauge_left(P["psi"], G))).cuda()
    # spectral bounds from a short Lanczos run, widened
    e = native.lanczos(plan, x0, 40, 1e-12, True, want_state=False)["energies"]
    shift, scale = 0.5 * (e[-1] + e[0]), 0.6 * (e[-1] - e[0]) + 1.0
    mu, rc = native.chebyshev_moments(plan, x0, 80, scale, shift)
    muz, rcz = native.chebyshev_moments(pz, xz, 80, scale, shift)
    assert np.max(np.abs(muz - mu)) <= 1e-11 * mu[0]
    assert np.max(np.abs(rcz - rc)) <= 1e-11 * mu[0]
    assert np.max(np.abs(mu)) <= 1.0 + 1e-12  # |T_k| <= 1 on the rescaled spectrum, <t0|t0> = 1


def test_chebyshev_moments_full_size_complex_hubbard(gpu):
    """The Hubbard bond at full size (L = 80, N = 80, D = 1024, environments
    built on the GPU) as a complex128 plan (the real environments as complex
    tables: a Hermitian H_eff) on a start vector with random phases:
    size-independent properties of the recurrence.  On the rescaled spectrum
    |T_k| <= 1, so |mu_k| <= mu_0; and for a Hermitian H_eff the reconstructed
    doubled moments equal the directly computed ones where both exist
    (T_2k = 2 T_k^2 - 1, T_{2k-1} = 2 T_k T_{k-1} - T_1), to 1e-9 of mu_0."""
    _say("chebyshev full-size complex hubbard")
    from mpstools_amd import bench_problem
    prob = bench_problem.u1_centre("hubbard", 80, 80, 1024, 20260005, precisions=("f64",))
    c, mpo, itk = prob["c"], prob["mpo"], prob["itk"]
    vz = prob["vt"].as_complex()
    plan = native.HeffPlan(prob["Lt"].as_complex(), prob["Rt"].as_complex(), vz, vz, (mpo[c], mpo[c + 1]),
                           (itk, itk), native.F64)
    dL = torch.cat([prob["dL"], torch.zeros_like(prob["dL"])])
    dR = torch.cat([prob["dR"], torch.zeros_like(prob["dR"])])
    plan.bind_env(dL, dR)
    g = torch.Generator(device="cuda")
    g.manual_seed(81)
    th = 2 * np.pi * torch.rand(prob["psi"].numel(), dtype=torch.float64, device="cuda", generator=g)
    x0 = torch.cat([prob["psi"] * torch.cos(th), prob["psi"] * torch.sin(th)])
    x0 /= x0.norm()
    e = native.lanczos(plan, x0, 60, 1e-12, True, want_state=False)["energies"]
    shift, half = 0.5 * (e[-1] + e[0]), 0.5 * (e[-1] - e[0])
    nmom = 64
    mu, rc = native.chebyshev_moments(plan, x0, nmom, 2.0 * half, shift)  # Ritz values lie inside: widen 2x
    assert abs(mu[0] - 1.0) <= 1e-12
    assert np.all(np.abs(mu) <= mu[0] * (1 + 1e-9))
    assert np.max(np.abs(rc[:nmom] - mu)) <= 1e-9 * mu[0]
    # the real plan on the real part of the same recurrence: moments of a real start vector
    pr = prob["plan_f64"]
    xr = prob["psi"] / prob["psi"].norm()
    mur, rcr = native.chebyshev_moments(pr, xr, nmom, 2.0 * half, shift)
    assert np.max(np.abs(rcr[:nmom] - mur)) <= 1e-9 * mur[0]
