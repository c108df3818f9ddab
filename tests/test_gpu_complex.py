"""Complex128 plans on the GPU (planar re/im storage, real MPO; the
complex128 instantiation of contractTwosite / contractSingle /
addContractionLayer, src/MPSfunctions.hpp:949-981, 857-885, 442-555).

Bars:
- against the complex oracle (pinned against real contractions in
  test_complex.py): 1e-12 relative, two-site and one-site H_eff, both
  environment updates, plain and mixed;
- gauge covariance, which fixes the conjugation conventions without the
  oracle: a random unitary G per sector on the bond left of the centre
  (A -> A G, psi -> G^H psi) makes the environment built on the GPU
  G^H L G and H_eff psi G^H (H_eff psi) (1e-12), and the complex Lanczos run
  from G^H psi has the real run's ground-state energy (1e-10) and state
  G^H gs (overlap 1 within 1e-6).
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


def _complex_case(oracle, model, L, N, D, seed, dense=False):
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c = P["c"]
    return P, cc.with_imag(P["Ls"][c - 1], seed + 1), cc.with_imag(P["Rs"][c + 2], seed + 2), cc.with_imag(
        P["psi"], seed + 3)


@pytest.mark.parametrize("model,L,N,D,seed,dense", [("tv", 20, 10, 60, 108, False), ("xxz", 16, 0, 128, 109, True),
                                                    ("hubbard", 12, 12, 64, 110, False)])
def test_complex_heff2_matches_oracle(gpu, oracle, model, L, N, D, seed, dense):
    _say("complex heff2", model, L, N, D)
    P, Lz, Rz, xz = _complex_case(oracle, model, L, N, D, seed, dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    got = native.contract_twosite(Lz, Rz, xz, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    want = oracle.contract_twosite(Lz, Rz, xz, mpo[c], mpo[c + 1], itk[c], itk[c + 1], cplx=True)
    assert cc.rel(got, {k: np.asarray(want[k]) for k in want.keys()}) <= 1e-12


@pytest.mark.parametrize("model,L,N,D,seed", [("tv", 20, 10, 60, 114), ("hubbard", 12, 12, 64, 115)])
def test_complex_heff1_and_env_match_oracle(gpu, oracle, model, L, N, D, seed):
    _say("complex heff1/env", model, L, N, D)
    P, Lz, _, _ = _complex_case(oracle, model, L, N, D, seed)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    R1 = cc.with_imag(P["Rs"][c + 1], seed + 4)
    Az = cc.with_imag(P["mps"][c], seed + 5)
    got = native.contract_single(Lz, Az, mpo[c], R1, itk[c])
    want = oracle.contract_single(Lz, Az, mpo[c], R1, itk[c], cplx=True)
    assert cc.rel(got, {k: np.asarray(want[k]) for k in want.keys()}) <= 1e-12
    for direction, E, site in ((1, Lz, c), (0, cc.with_imag(P["Rs"][c + 1], seed + 6), c)):
        A = cc.with_imag(P["mps"][site], seed + 7)
        for up in (None, cc.with_imag(P["mps"][site], seed + 8)):
            got = native.add_contraction_layer(E, A, mpo[site], itk[site], direction, up)
            want = oracle.add_layer(E, A, mpo[site], itk[site], direction, Ap=up, cplx=True)
            assert got.keys() == want.keys()
            assert cc.rel(got, {k: np.asarray(want[k]) for k in want.keys()}) <= 1e-12, (direction, up is None)


def _lanczos(L, R, psi, mpo, itk, c):
    Lt, Rt, xt = native.Table.from_blocks(L), native.Table.from_blocks(R), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    dL, dR = torch.from_numpy(Lt.pack(L)).cuda(), torch.from_numpy(Rt.pack(R)).cuda()
    plan.bind_env(dL, dR)
    g = native.lanczos(plan, torch.from_numpy(vt.pack(psi)).cuda(), 300, 1e-10, True)
    return g, vt.unpack(g["state"].cpu().numpy())


@pytest.mark.parametrize("model,L,N,D,seed", [("tv", 20, 10, 60, 112), ("hubbard", 10, 10, 48, 113)])
def test_complex_gauge_covariance_and_lanczos(gpu, oracle, model, L, N, D, seed):
    _say("complex gauge", model, L, N, D)
    P = problems.build(oracle, model, L, N, D, seed)
    c, mps, mpo, itk = P["c"], P["mps"], P["mpo"], P["itk"]
    assert c >= 2
    G = cc.bond_unitaries(mps[c - 1], seed)
    L0, R0, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    # the environment built on the GPU from the gauged tensor
    Lg = native.add_contraction_layer(P["Ls"][c - 2], cc.gauge_right(mps[c - 1], G), mpo[c - 1], itk[c - 1], 1)
    assert cc.rel(Lg, {k: G[k[0]].conj().T @ np.asarray(L0[k]) @ G[k[1]] for k in L0.keys()}) <= 1e-12
    # H_eff in the gauged basis
    psig = cc.gauge_left(psi, G)
    y = native.contract_twosite(L0, R0, psi, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    yg = native.contract_twosite(Lg, R0, psig, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    want = cc.gauge_left(y, G)
    assert cc.rel(yg, {k: np.asarray(want[k]) for k in want.keys()}) <= 1e-12
    # complex Lanczos: same spectrum, gauged ground state
    gr, sr = _lanczos(L0, R0, psi, mpo, itk, c)
    gz, sz = _lanczos(Lg, cc.with_imag(R0, 0, scale=0.0), psig, mpo, itk, c)
    assert abs(gz["energies"][0] - gr["energies"][0]) <= 1e-10
    sw = cc.gauge_left(sr, G)
    ov = sum(complex(np.vdot(np.asarray(sw[k]), np.asarray(sz[k]))) for k in sw.keys())
    assert abs(abs(ov) - 1.0) <= 1e-6


def test_complex_batched_lanczos_matches_single_runs(gpu, oracle):
    """Several complex start vectors through mpsa_lanczos_batched (one batched
    complex matvec per step, graph replay) against one mpsa_lanczos run per
    vector: the same iteration counts and Ritz values (1e-10)."""
    _say("complex batched lanczos")
    P = problems.build(oracle, "tv", 20, 10, 60, 116)
    c, mps, mpo, itk = P["c"], P["mps"], P["mpo"], P["itk"]
    # a Hermitian complex H_eff: the left environment of the gauged MPS
    G = cc.bond_unitaries(mps[c - 1], 116)
    Lz = oracle.add_layer(cc.with_imag(P["Ls"][c - 2], 0, 0.0), cc.gauge_right(mps[c - 1], G), mpo[c - 1],
                          itk[c - 1], 1, cplx=True)
    Rz = cc.with_imag(P["Rs"][c + 2], 0, 0.0)
    Lt, Rt = native.Table.from_blocks(Lz), native.Table.from_blocks(Rz)
    xs = [cc.with_imag(cc.gauge_left(P["psi"], G), 200 + v) for v in range(4)]
    xt = native.Table.from_blocks(xs[0])
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    dL, dR = torch.from_numpy(Lt.pack(Lz)).cuda(), torch.from_numpy(Rt.pack(Rz)).cuda()
    plan.bind_env(dL, dR)
    n = vt.length
    x0 = torch.cat([torch.from_numpy(vt.pack(x)) for x in xs]).cuda()
    b = native.lanczos_batched(plan, x0, 4, n, 40, 1e-10, want_state=False)
    for v in range(4):
        s = native.lanczos(plan, x0[v * n:(v + 1) * n].clone(), 40, 1e-10, True, want_state=False)
        assert b["it"][v] == s["it"]
        assert np.max(np.abs(b["energies"][v] - s["energies"])) <= 1e-10 * max(1.0, np.max(np.abs(s["energies"])))


@pytest.mark.parametrize("prec,tol", [(native.F32, 1e-5), (native.BF16X3, 1e-4)])
@pytest.mark.parametrize("model,L,N,D,seed,dense", [("tv", 20, 10, 60, 117, False), ("xxz", 16, 0, 256, 118, True),
                                                    ("hubbard", 12, 12, 64, 119, False)])
def test_complex_reduced_precision_within_bound(gpu, oracle, model, L, N, D, seed, dense, prec, tol):
    """Reduced-precision complex plans (L and R converted as [re | im | -im]
    planes, x as [re | im]) against the complex oracle, within the
    precision's bound for real plans, with every admissible tile class."""
    _say("complex reduced precision", model, L, N, D, prec)
    P, Lz, Rz, xz = _complex_case(oracle, model, L, N, D, seed, dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lt, Rt, xt = native.Table.from_blocks(Lz), native.Table.from_blocks(Rz), native.Table.from_blocks(xz)
    yt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]).as_complex()
    plan = native.HeffPlan(Lt, Rt, xt, yt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), prec)
    dL, dR = torch.from_numpy(Lt.pack(Lz)).cuda(), torch.from_numpy(Rt.pack(Rz)).cuda()
    plan.bind_env(dL, dR)
    x = torch.from_numpy(xt.pack(xz)).cuda()
    want = oracle.contract_twosite(Lz, Rz, xz, mpo[c], mpo[c + 1], itk[c], itk[c + 1], cplx=True)
    want = {k: np.asarray(want[k]) for k in want.keys()}
    ncls = native.TILE_CLASSES if prec == native.BF16X3 else 3
    for c1 in range(ncls):
        for c2 in range(ncls):
            try:
                plan.set_tile_class(1, c1)
                plan.set_tile_class(2, c2)
            except native.MpsaError:
                continue
            y = torch.zeros(yt.length, dtype=torch.float64, device="cuda")
            plan.apply(x, y)
            torch.cuda.synchronize()
            assert cc.rel(yt.unpack(y.cpu().numpy()), want) <= tol, (c1, c2)
    if prec == native.BF16X3 and dense:  # unit segment scales: the 256x256 class is admitted
        plan.set_tile_class(1, 3)
        plan.set_tile_class(2, 3)
