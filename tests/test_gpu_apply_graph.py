"""H_eff applies replayed from the plan's own graph (mpsa_heff_apply on a
caller stream that is not being captured): the replayed apply equals the
directly launched one bit for bit (FP64 and BF16X3, one and several
vectors; the graph is recorded on the second apply in a row with the same
arguments), a change of vectors, vector count or environment (bind_env)
records anew instead of replaying stale arguments, and applies inside a
caller's capture (the batched Lanczos) still run directly."""
import numpy as np
import pytest

import problems
from mpstools_amd import native

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _setup(torch, oracle, prec):
    P = problems.build(oracle, "tv", 20, 10, 60, 131)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), prec)
    dL, dR = torch.from_numpy(Lt.pack(Lb)).cuda(), torch.from_numpy(Rt.pack(Rb)).cuda()
    plan.bind_env(dL, dR)
    return plan, vt, dL, dR


@pytest.mark.parametrize("prec", [native.F64, native.BF16X3])
def test_graph_apply_equals_direct(gpu, oracle, prec):
    torch = gpu
    _say("apply graph", prec)
    plan, vt, dL, dR = _setup(torch, oracle, prec)
    rng = np.random.default_rng(5)
    nvec, stride = 3, vt.total + 64
    x = torch.from_numpy(rng.standard_normal(nvec * stride)).cuda()
    y_ref = torch.zeros(nvec * stride, dtype=torch.float64, device="cuda")
    plan.apply(x, y_ref, nvec, stride, stride)  # null stream: direct launches
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    # the first apply on the stream launches directly, the second records the
    # graph, the next ones replay it
    with torch.cuda.stream(s):
        yb = torch.zeros(nvec * stride, dtype=torch.float64, device="cuda")
        for _ in range(4):
            plan.apply(x, yb, nvec, stride, stride)
            s.synchronize()
            for v in range(nvec):
                o = v * stride
                assert torch.equal(yb[o:o + vt.total], y_ref[o:o + vt.total])
        # one vector from the same buffers: direct, then recorded and replayed,
        # each equal to the direct one-vector apply on the null stream
        y1 = torch.zeros(stride, dtype=torch.float64, device="cuda")
        y1s = []
        for _ in range(3):
            y1.zero_()
            plan.apply(x, y1, 1, stride, stride)
            s.synchronize()
            y1s.append(y1.clone())
    y1_ref = torch.zeros(stride, dtype=torch.float64, device="cuda")
    plan.apply(x, y1_ref, 1, stride, stride)
    torch.cuda.synchronize()
    for y in y1s:
        assert torch.equal(y[:vt.total], y1_ref[:vt.total])
    # a new environment: the captured launches must not keep the old one
    with torch.cuda.stream(s):
        plan.bind_env(2.0 * dL, dR)
        yb2 = torch.zeros(nvec * stride, dtype=torch.float64, device="cuda")
        yb2s = []
        for _ in range(3):  # direct, recorded, replayed
            yb2.zero_()
            plan.apply(x, yb2, nvec, stride, stride)
            s.synchronize()
            yb2s.append(yb2.clone())
    # doubling L is exact in every precision (bf16 hi/lo splits scale by 2 too)
    for yb2 in yb2s:
        for v in range(nvec):
            o = v * stride
            assert torch.equal(yb2[o:o + vt.total], 2.0 * y_ref[o:o + vt.total])
    torch.cuda.synchronize()


def test_lanczos_on_a_stream_with_graph_applies(gpu, oracle):
    """The batched Lanczos captures its own graph around the applies (they
    run directly there) and runs its first step eagerly on its private stream,
    where the apply records its own graph: the result equals the host-driven
    Lanczos."""
    torch = gpu
    _say("lanczos with graph applies")
    plan, vt, dL, dR = _setup(torch, oracle, native.F64)
    rng = np.random.default_rng(9)
    x0 = torch.from_numpy(rng.standard_normal(vt.total)).cuda()
    a = native.lanczos_batched(plan, x0, 1, vt.total, 200, 1e-10, graph=True)
    b = native.lanczos(plan, x0, 200, 1e-10, True)
    assert int(a["it"][0]) == b["it"]
    assert abs(a["energies"][0][0] - b["energies"][0]) <= 1e-10 * max(1.0, abs(b["energies"][0]))
