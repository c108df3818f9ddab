"""CPU checks of the host side of the HIP library: the C-ABI library loads and
exports every entry point of include/mpstools_amd.h, the planner reproduces the
reference's output key sets, and the grouped-GEMM schedule it hands to the
kernels, executed by a numpy interpreter, reproduces the oracle.  No GPU and
no kernel launch is involved."""
import ctypes
import os
import re

import numpy as np
import pytest

import plan_interp
import problems
from mpstools_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mpstools_amd.h")).read()
    names = set(re.findall(r"\b(mpsa_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 15
    lib = ctypes.CDLL(native.LIBPATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = [("xxz", 10, 5, 12, 201, False), ("asym", 10, 4, 16, 202, False), ("tv", 16, 8, 40, 203, False),
         ("xxz", 10, 0, 24, 204, True)]


@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES)
def test_heff2_schedule_reproduces_oracle(oracle, model, L, N, D, seed, dense):
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    ref = oracle.contract_twosite(Lb, Rb, psi, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    yt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    assert yt.keys() == ref.keys()
    plan = native.HeffPlan(Lt, Rt, xt, yt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    info = plan.info()
    y = np.zeros(yt.total)
    plan_interp.run_heff(plan.describe(), Lt.pack(Lb), Rt.pack(Rb), xt.pack(psi), y, info["workspace_bytes"] // 8)
    yb = yt.unpack(y)
    a = np.concatenate([yb[k].ravel() for k in ref.keys()])
    b = np.concatenate([ref[k].ravel() for k in ref.keys()])
    assert _rel(a, b) <= 1e-13
    assert info["flops_alg"] <= info["flops_ref"]


@pytest.mark.parametrize("direction", [1, 0])
@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES)
def test_env_schedule_reproduces_oracle(oracle, model, L, N, D, seed, dense, direction):
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk, mps = P["c"], P["mpo"], P["itk"], P["mps"]
    if direction > 0:
        BL, A, W, it = P["Ls"][c - 1], mps[c], mpo[c], itk[c]
    else:
        BL, A, W, it = P["Rs"][c + 2], mps[c + 1], mpo[c + 1], itk[c + 1]
    ref = oracle.add_layer(BL, A, W, it, direction)
    BLt, At = native.Table.from_blocks(BL), native.Table.from_blocks(A)
    ot = native.env_outkeys(BLt, At, W, it, direction)
    assert ot.keys() == ref.keys()
    plan = native.EnvPlan(BLt, At, W, it, direction, ot)
    out = np.zeros(ot.total)
    plan_interp.run_env(plan.describe(), BLt.pack(BL), At.pack(A), None, out, plan.info()["workspace_bytes"] // 8)
    ob = ot.unpack(out)
    a = np.concatenate([ob[k].ravel() for k in ref.keys()])
    b = np.concatenate([ref[k].ravel() for k in ref.keys()])
    assert _rel(a, b) <= 1e-13


def test_plan_errors_are_reported():
    from mpstools_amd import models
    from mpstools_amd.blocks import Blocks
    A = Blocks(1)
    t = native.Table.from_blocks(A)
    with pytest.raises(native.MpsaError, match="A is empty"):
        native.env_outkeys(t, t, models.xxz_mpo(4)[1], [0, 1], 1)
    B = Blocks(1)
    B[(0, 0, 0)] = np.ones((2, 3))
    L = Blocks(1)
    L[(0, 0, 0)] = np.ones((2, 2))
    with pytest.raises(native.MpsaError, match="dimensions"):
        native.env_outkeys(native.Table.from_blocks(L), native.Table.from_blocks(B), models.xxz_mpo(4)[1], [0, 0], 0)


def test_tridiagonal_eigensolver_matches_scipy():
    import scipy.linalg
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 40, 200):
        d, e = rng.standard_normal(n), rng.standard_normal(max(n - 1, 0))
        w, z = native.tridiag_eig(d, e)
        w2 = scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True) if n > 1 else d
        assert np.max(np.abs(w - w2)) < 1e-12 * max(1, np.max(np.abs(w2)))
        T = np.diag(d) + (np.diag(e, 1) + np.diag(e, -1) if n > 1 else 0)
        assert np.max(np.abs(T @ z - z * w)) < 1e-11
        assert np.max(np.abs(z.T @ z - np.eye(n))) < 1e-12
    # repeated and clustered eigenvalues
    d = np.array([1.0, 1.0, 1.0, 2.0, 2.0])
    e = np.array([0.0, 1e-9, 0.0, 0.0])
    w, z = native.tridiag_eig(d, e)
    assert np.max(np.abs(w - np.sort(scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True)))) < 1e-14


def test_plan_vector_lengths_and_union_layout(oracle):
    P = problems.build(oracle, "tv", 16, 8, 40, 205)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lt, Rt, xt = (native.Table.from_blocks(P["Ls"][c - 1]), native.Table.from_blocks(P["Rs"][c + 2]),
                  native.Table.from_blocks(P["psi"]))
    yt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    vt = xt.union(yt)
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]))
    x = np.zeros(1, np.int64)
    y = np.zeros(1, np.int64)
    s = np.zeros(1, np.int32)
    native._check(native.lib().mpsa_plan_vector_lengths(plan.h, native._ptr(x), native._ptr(y), native._ptr(s)))
    assert x[0] == y[0] == vt.total and s[0] == 1
    plan2 = native.HeffPlan(Lt, Rt, xt, yt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]))
    native._check(native.lib().mpsa_plan_vector_lengths(plan2.h, native._ptr(x), native._ptr(y), native._ptr(s)))
    assert s[0] == (1 if xt.keys() == yt.keys() else 0)


@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES)
def test_plain_schedule_reproduces_oracle_with_reference_flops(oracle, model, L, N, D, seed, dense):
    """MPSA_PLAIN: one product pair per tuple, the reference's F_ref flop count."""
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    ref = oracle.contract_twosite(Lb, Rb, psi, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    yt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    plan = native.HeffPlan(Lt, Rt, xt, yt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64 | native.PLAIN)
    info = plan.info()
    assert abs(info["flops_alg"] - info["flops_ref"]) <= 1e-9 * info["flops_ref"]
    assert len(plan.describe()["mix"]) == 0
    y = np.zeros(yt.total)
    plan_interp.run_heff(plan.describe(), Lt.pack(Lb), Rt.pack(Rb), xt.pack(psi), y, info["workspace_bytes"] // 8)
    yb = yt.unpack(y)
    a = np.concatenate([yb[k].ravel() for k in ref.keys()])
    b = np.concatenate([ref[k].ravel() for k in ref.keys()])
    assert _rel(a, b) <= 1e-13


def test_chain_environment_tables_match_oracle(oracle):
    """mpsa_chain_create plans every environment of the chain on the host; its
    key tables are the oracle's left/right environments (getLMin .. L[N-1],
    getRMax .. R[0])."""
    from mpstools_amd import models
    L = 10
    P = problems.build(oracle, "xxz", L, 5, 12, 3)
    mps, mpo, itk = P["mps"], P["mpo"], P["itk"]
    tabs = [native.Table.from_blocks(A) for A in mps]
    d = models.xxz_param_derivatives(L)
    ch = native.Chain(tabs, mpo, itk, [d["Jz"], d["Jxy"], d["hz"]], checkpoint=3)
    Ls = oracle.left_envs(mps, mpo, itk, L - 1)
    for i in range(-1, L):
        ref = Ls[i] if i >= 0 else oracle.lmin()
        assert ch.env_table(1, i).keys() == ref.keys(), i
    Rs = oracle.right_envs(mps, mpo, itk, 0)
    for i in range(0, L + 1):
        ref = Rs[i] if i < L else oracle.rmax(mps[-1])
        assert ch.env_table(0, i).keys() == ref.keys(), i
    # derivative tensors: W(theta) = W(0) + sum_p theta_p dW_p entrywise
    W = models.xxz_mpo(L, 0.7, 1.3, 0.2)
    W0 = models.xxz_mpo(L, 0.0, 0.0, 0.0)
    for i in range(L):
        dense = W0[i].dense() + 0.7 * d["Jz"][i].dense() + 1.3 * d["Jxy"][i].dense() + 0.2 * d["hz"][i].dense()
        assert np.allclose(dense, W[i].dense(), atol=1e-15)
