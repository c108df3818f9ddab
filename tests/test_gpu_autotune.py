"""Tile-class selection of H_eff plans on the GPU (mpsa_plan_set_tile_class,
mpsa_plan_autotune).

Bars: in FP64 every tile class (128x128, 64x64, 32x32 per workgroup, 16x16 and 32x32 per wave) of both grouped GEMM
stages gives the same bits (each output element is accumulated over k in the
same order whatever the tile) and the oracle's contraction (doContraction,
src/MPSfunctions.hpp:726-804) within 1e-12; in BF16X3 every admissible class, the 256x256 one included,
stays within the precision's bound (1e-4 relative to the oracle); the
autotuner reports a time for every admissible class, keeps the fastest per
stage, and the tuned plan then gives the bits of that class.
"""
import numpy as np
import pytest
import torch

import problems
from mpstools_amd import native

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _setup(gpu, oracle, model, L, N, D, seed, dense, prec):
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), prec)
    dL = torch.from_numpy(Lt.pack(Lb)).cuda()
    dR = torch.from_numpy(Rt.pack(Rb)).cuda()
    plan.bind_env(dL, dR)
    x = torch.from_numpy(vt.pack(psi)).cuda()
    ref = oracle.contract_twosite(Lb, Rb, psi, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    return plan, vt, x, ref, (dL, dR)


def _apply(plan, vt, x):
    y = torch.zeros(vt.total, dtype=torch.float64, device="cuda")
    plan.apply(x, y)
    torch.cuda.synchronize()
    return y


def _rel(vt, y, ref):
    yb = vt.unpack(y.cpu().numpy())
    num = sum(np.sum((yb[k] - ref[k]) ** 2) for k in ref.keys())
    den = sum(np.sum(ref[k] ** 2) for k in ref.keys())
    return float(np.sqrt(num / den))


def _admissible(plan, stage, prec):
    out = []
    for cls in range(5):
        try:
            plan.set_tile_class(stage, cls)
            out.append(cls)
        except native.MpsaError:
            pass
    return out


@pytest.mark.parametrize("model,L,N,D,seed,dense", [("tv", 20, 10, 60, 108, False), ("xxz", 20, 0, 256, 109, True)])
def test_fp64_tile_classes_give_identical_bits(gpu, oracle, model, L, N, D, seed, dense):
    _say("fp64 tile classes", model, L, N, D)
    plan, vt, x, ref, keep = _setup(gpu, oracle, model, L, N, D, seed, dense, native.F64)
    outs = {}
    for c1 in (0, 1, 2, 3, 4):
        for c2 in (0, 1, 2, 3, 4):
            plan.set_tile_class(1, c1)
            plan.set_tile_class(2, c2)
            outs[(c1, c2)] = _apply(plan, vt, x)
    y0 = outs[(0, 0)]
    assert all(torch.equal(y, y0) for y in outs.values())
    assert _rel(vt, y0, ref) <= 1e-12
    for bad in ((1, 5), (2, 6), (0, 0)):
        with pytest.raises(native.MpsaError):
            plan.set_tile_class(*bad)


@pytest.mark.parametrize("model,L,N,D,seed,dense", [("tv", 20, 10, 60, 108, False), ("xxz", 20, 0, 512, 106, True)])
def test_bf16x3_every_admissible_class_within_bound(gpu, oracle, model, L, N, D, seed, dense):
    _say("bf16x3 tile classes", model, L, N, D)
    plan, vt, x, ref, keep = _setup(gpu, oracle, model, L, N, D, seed, dense, native.BF16X3)
    a1, a2 = _admissible(plan, 1, native.BF16X3), _admissible(plan, 2, native.BF16X3)
    assert {0, 1, 2, 4} <= set(a1) and {0, 1, 2, 4} <= set(a2)
    if dense:  # unit segment scales: the 256x256 tile is admitted
        assert 3 in a1 and 3 in a2
    for c1 in a1:
        for c2 in a2:
            plan.set_tile_class(1, c1)
            plan.set_tile_class(2, c2)
            assert _rel(vt, _apply(plan, vt, x), ref) <= 1e-4, (c1, c2)


@pytest.mark.parametrize("prec", [native.F64, native.BF16X3])
def test_autotune_keeps_the_fastest_class(gpu, oracle, prec):
    _say("autotune", prec)
    plan, vt, x, ref, keep = _setup(gpu, oracle, "xxz", 20, 0, 512, 106, True, prec)
    ms = plan.autotune(nvec=1, reps=3)
    for stage in (1, 2):
        times = {c: t for (s, c), t in ms.items() if s == stage}
        assert set(times) == set(_admissible(plan, stage, prec)) and len(times) >= 3
        assert all(t > 0 for t in times.values())
    # _admissible moved the classes: tune again, then check the choice
    ms = plan.autotune(nvec=1, reps=3)
    si = plan.stage_info()
    for stage, key in ((1, "tile_class1"), (2, "tile_class2")):
        times = {c: t for (s, c), t in ms.items() if s == stage}
        assert times[si[key]] == min(times.values())
    y = _apply(plan, vt, x)
    plan.set_tile_class(1, si["tile_class1"])
    plan.set_tile_class(2, si["tile_class2"])
    assert torch.equal(_apply(plan, vt, x), y)
    assert _rel(vt, y, ref) <= (1e-12 if prec == native.F64 else 1e-4)
