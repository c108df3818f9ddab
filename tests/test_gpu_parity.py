"""GPU parity: the HIP grouped-GEMM path (through the C ABI) against the CPU
oracle on identical seeded inputs.

Bar: identical output key sets (block structure) and, in FP64, a norm-wise
relative difference <= 1e-12 (per the north-star parity target).  Reduced
precision paths are held to the bounds stated in each test.
"""
import numpy as np
import pytest

import problems
from mpstools_amd import native
from mpstools_amd.blocks import Blocks

pytestmark = pytest.mark.gpu


def _rel(a: Blocks, b: Blocks):
    assert a.keys() == b.keys()
    num = sum(np.sum(np.abs(a[k] - b[k]) ** 2) for k in a.keys())
    den = sum(np.sum(np.abs(b[k]) ** 2) for k in b.keys())
    return float(np.sqrt(num / max(den, 1e-300)))


def _say(*a):
    print("[case]", *a, flush=True)


CASES = [
    ("xxz", 10, 5, 12, 101, False),
    ("asym", 10, 4, 16, 102, False),
    ("tv", 16, 8, 40, 103, False),
    ("xxz", 12, 0, 48, 104, True),
    ("tv", 24, 12, 150, 105, False),
]
# Fermi-Hubbard (d = 4, U(1) x U(1) packed charges), U = 4, 6 up + 6 down
HUB = [("hubbard", 12, 12, 64, 108, False)]


@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES + HUB)
def test_heff2_f64_matches_oracle(gpu, oracle, model, L, N, D, seed, dense):
    _say("heff2 f64", model, L, N, D, seed)
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    args = (P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    y_ref = oracle.contract_twosite(*args)
    y = native.contract_twosite(*args, precision=native.F64)
    assert y.keys() == y_ref.keys()
    for k in y.keys():
        assert y[k].shape == y_ref[k].shape
    assert _rel(y, y_ref) <= 1e-12


@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES[:4] + HUB)
def test_heff1_f64_matches_oracle(gpu, oracle, model, L, N, D, seed, dense):
    _say("heff1 f64", model, L, N, D, seed)
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk, mps = P["c"], P["mpo"], P["itk"], P["mps"]
    args = (P["Ls"][c - 1], mps[c], mpo[c], P["Rs"][c + 1], itk[c])
    y_ref = oracle.contract_single(*args)
    y = native.contract_single(*args, precision=native.F64)
    assert _rel(y, y_ref) <= 1e-12


@pytest.mark.parametrize("direction", [1, 0])
@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES + HUB)
def test_env_f64_matches_oracle(gpu, oracle, model, L, N, D, seed, dense, direction):
    _say("env f64", direction, model, L, N, D, seed)
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk, mps = P["c"], P["mpo"], P["itk"], P["mps"]
    if direction > 0:
        BL, A, W, it = P["Ls"][c - 1], mps[c], mpo[c], itk[c]
    else:
        BL, A, W, it = P["Rs"][c + 2], mps[c + 1], mpo[c + 1], itk[c + 1]
    ref = oracle.add_layer(BL, A, W, it, direction)
    out = native.add_contraction_layer(BL, A, W, it, direction)
    assert out.keys() == ref.keys()
    assert _rel(out, ref) <= 1e-12


def test_env_mixed_variant_matches_oracle(gpu, oracle):
    """addContractionLayer with a distinct upper tensor Ap (src/MPSfunctions.hpp:499-555)."""
    _say("env Ap")
    P = problems.build(oracle, "tv", 16, 8, 30, 106)
    Q = problems.build(oracle, "tv", 16, 8, 30, 107)
    c = P["c"]
    BL, A, W, it = P["Ls"][c - 1], P["mps"][c], P["mpo"][c], P["itk"][c]
    Ap = Q["mps"][c]
    if Ap.keys() != A.keys():
        pytest.skip("profiles differ")
    ref = oracle.add_layer(BL, A, W, it, 1, Ap=Ap)
    out = native.add_contraction_layer(BL, A, W, it, 1, Ap=Ap)
    assert _rel(out, ref) <= 1e-12
    ref = oracle.add_layer(P["Rs"][c + 2], P["mps"][c + 1], P["mpo"][c + 1], it, 0, Ap=Q["mps"][c + 1])
    out = native.add_contraction_layer(P["Rs"][c + 2], P["mps"][c + 1], P["mpo"][c + 1], it, 0, Ap=Q["mps"][c + 1])
    assert _rel(out, ref) <= 1e-12


@pytest.mark.parametrize("prec,tol", [(native.F32, 2e-6), (native.BF16, 3e-2), (native.BF16X3, 1e-4),
                                      (native.F64 | native.PLAIN, 1e-12)])
@pytest.mark.parametrize("model,L,N,D,seed,dense", [CASES[2], CASES[3], CASES[4], ("xxz", 20, 0, 512, 106, True),
                                                     ("xxz", 24, 0, 720, 107, True)] + HUB)
def test_heff2_reduced_precision_bounds(gpu, oracle, model, L, N, D, seed, dense, prec, tol):
    """FP32 inputs/accumulate: relative error <= 2e-6; BF16 inputs with FP32
    accumulate: <= 3e-2 (bf16 has an 8-bit significand, unit roundoff 2^-9);
    BF16X3 (hi/lo split, three MFMAs): <= 1e-4; the FP64 plain-method schedule:
    <= 1e-12."""
    _say("heff2 reduced", prec, model, L, N, D, seed)
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    args = (P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    y_ref = oracle.contract_twosite(*args)
    y = native.contract_twosite(*args, precision=prec)
    assert y.keys() == y_ref.keys()
    assert _rel(y, y_ref) <= tol


def test_batched_apply_equals_single_applies(gpu, oracle):
    """nvec vectors in one grouped launch give the same bits as nvec launches."""
    torch = gpu
    _say("batched")
    P = problems.build(oracle, "tv", 20, 10, 60, 108)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    L, R, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(L), native.Table.from_blocks(R), native.Table.from_blocks(psi)
    yt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    vt = xt.union(yt)
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    dL = torch.from_numpy(Lt.pack(L)).cuda()
    dR = torch.from_numpy(Rt.pack(R)).cuda()
    plan.bind_env(dL, dR)
    nvec, stride = 7, vt.total + 128
    rng = np.random.default_rng(5)
    X = torch.zeros(nvec * stride, dtype=torch.float64, device="cuda")
    base = vt.pack(psi)
    for v in range(nvec):
        X[v * stride:v * stride + vt.total] = torch.from_numpy(base * (1 + v) + rng.standard_normal(vt.total) * (v > 0))
    Y = torch.zeros_like(X)
    plan.apply(X, Y, nvec, stride, stride)
    for v in range(nvec):
        y1 = torch.zeros(vt.total, dtype=torch.float64, device="cuda")
        plan.apply(X[v * stride:v * stride + vt.total].clone(), y1)
        assert torch.equal(y1, Y[v * stride:v * stride + vt.total])
    y0 = vt.unpack(Y[:vt.total].cpu().numpy())
    ref = oracle.contract_twosite(L, R, psi, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    num = sum(np.sum((y0[k] - ref[k]) ** 2) for k in ref.keys())
    den = sum(np.sum(ref[k] ** 2) for k in ref.keys())
    assert np.sqrt(num / den) <= 1e-12
    for k in y0.keys():  # blocks of the vector space that H_eff does not produce stay zero
        if k not in ref:
            assert not np.any(y0[k])


def test_dot_and_axpby(gpu):
    torch = gpu
    _say("dot/axpby")
    rng = np.random.default_rng(9)
    for n in (1, 1000, 1 << 20):
        x = rng.standard_normal(n)
        y = rng.standard_normal(n)
        dx, dy = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
        assert abs(native.dot(dx, dy) - x @ y) <= 1e-12 * (np.abs(x) @ np.abs(y))
        native.axpby(0.5, dx, -2.0, dy)
        torch.cuda.synchronize()
        assert np.max(np.abs(dy.cpu().numpy() - (0.5 * x - 2.0 * y))) < 1e-14


def test_full_size_dense_properties(gpu, oracle):
    """Dense (no-symmetry) Heisenberg bond at D=2048: linearity and the
    symmetry <x|H y> = <H x|y> of the real-symmetric H_eff, all on the GPU."""
    torch = gpu
    _say("full-size dense D=2048 properties")
    from mpstools_amd import bench_problem
    prob = bench_problem.dense_heisenberg_centre(D=2048, L=128, seed=20260003, build_envs="sweep")
    plan = prob["plan_f64"]
    n = prob["vt"].total
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    Hx, Hy, Hxy = torch.zeros_like(x), torch.zeros_like(x), torch.zeros_like(x)
    plan.apply(x, Hx)
    plan.apply(y, Hy)
    plan.apply(0.25 * x - 3.0 * y, Hxy)
    lin = (Hxy - (0.25 * Hx - 3.0 * Hy)).norm() / Hxy.norm()
    assert lin.item() <= 1e-13
    sym = abs(native.dot(Hy, x) - native.dot(y, Hx)) / (Hx.norm() * y.norm()).item()
    assert sym <= 1e-13


def _lanczos_setup(gpu, oracle, model, L, N, D, seed):
    torch = gpu
    P = problems.build(oracle, model, L, N, D, seed)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    yt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    vt = xt.union(yt)
    dL = torch.from_numpy(Lt.pack(Lb)).cuda()
    dR = torch.from_numpy(Rt.pack(Rb)).cuda()
    x0 = torch.from_numpy(vt.pack(psi)).cuda()
    return P, (Lt, Rt, vt, dL, dR, x0)


@pytest.mark.parametrize("model,L,N,D,seed", [("xxz", 12, 6, 24, 111), ("tv", 20, 10, 60, 112),
                                              ("hubbard", 10, 10, 48, 113)])
def test_lanczos_matches_oracle(gpu, oracle, model, L, N, D, seed):
    """Device Lanczos (same recurrence and stopping rules as TSLanzcosEngine)
    against the oracle's: 10 fixed steps give the same Ritz values; run to
    convergence, the same ground-state energy and state (up to sign)."""
    _say("lanczos", model, L, N, D, seed)
    P, (Lt, Rt, vt, dL, dR, x0) = _lanczos_setup(gpu, oracle, model, L, N, D, seed)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    args = (P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    plan.bind_env(dL, dR)
    r10 = oracle.lanczos_twosite(*args, itmax=10, delta=1e-12, testenergy=True)
    g10 = native.lanczos(plan, x0, 10, 1e-12, True)
    assert g10["it"] == r10["it"] == 10
    assert np.max(np.abs(g10["energies"] - r10["energies"])) <= 1e-10 * np.max(np.abs(r10["energies"]))
    r = oracle.lanczos_twosite(*args, itmax=300, delta=1e-10, testenergy=True)
    g = native.lanczos(plan, x0, 300, 1e-10, True)
    assert abs(g["energies"][0] - r["energies"][0]) <= 1e-9
    gs = vt.unpack(g["state"].cpu().numpy())
    ov = sum(float(np.sum(gs[k] * r["gs"][k])) for k in r["gs"].keys())
    assert abs(abs(ov) - 1.0) <= 1e-6


def test_lanczos_residual_mode_and_itmax_cap(gpu, oracle):
    _say("lanczos residual mode")
    P, (Lt, Rt, vt, dL, dR, x0) = _lanczos_setup(gpu, oracle, "xxz", 12, 6, 24, 113)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    plan.bind_env(dL, dR)
    g = native.lanczos(plan, x0, 400, 1e-8, False)
    r = oracle.lanczos_twosite(P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1],
                               itmax=400, delta=1e-8, testenergy=False)
    assert abs(g["energies"][0] - r["energies"][0]) <= 1e-9
    g3 = native.lanczos(plan, x0, 3, 1e-12, True)
    assert g3["it"] == 3 and len(g3["energies"]) == 3
