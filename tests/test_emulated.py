"""The library's host logic on the CPU, against an emulation of the device
layer (tests/emu/emu.cpp): kernels computed on the host with every device
address they touch checked against the live allocations, never-written memory
NaN-filled, freed memory kept and poisoned.  Covered: the DMRG sweep
(mpsa_dmrg.cpp over the plan builder, the split and the Lanczos driver,
unchanged), the tile-class selection of H_eff plans, Hubbard (U(1) x U(1))
plans, complex128 plans (FP64 and FP32; the BF16X3 schedule), the split's
sector SVD paths, the Chebyshev moments driver and the C++ drop-in header.

Bars, as for the device sweep (DMRGengine::simulate, src/container.hpp:383-438):
untruncated sweeps reach the exact ground-state energy within 1e-9 (free
fermions; exact diagonalisation of the Sz sector); a truncated sweep from the
same right-normalised start gives the oracle's energy and truncated weight, its
key sets and a final state of the same energy (oracle/pydmrg.py); and the
emulation reports no out-of-bounds access, use after free or bad free.

CPU only.  The emulated library is built by tests/emu/Makefile with g++.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse.linalg

import dense
import problems
import pydmrg
from mpstools_amd import models, native
from test_oracle import _sector_H, right_normalised_mps

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_LIB = os.path.join(HERE, "emu", "_build", "libmpsa_emu.so")


@pytest.fixture(scope="module")
def emu():
    if not shutil.which("g++") or not os.path.exists("/opt/rocm/include/hip/hip_runtime.h"):
        pytest.skip("g++ or the HIP headers are not available")
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "emu")], check=True, capture_output=True)
    L = C.CDLL(EMU_LIB)
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.emu_violations.restype = C.c_long
    L.emu_message.restype = C.c_char_p
    return L


@pytest.fixture
def lib(emu, monkeypatch):
    """native's ABI wrappers bound to the emulated library for one test."""
    monkeypatch.setattr(native, "LIBPATH", EMU_LIB)
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setattr(native, "_stream", lambda: C.c_void_p(0))
    L = native.lib()
    v0 = emu.emu_violations()
    yield emu
    n = emu.emu_violations() - v0
    msgs = [emu.emu_message(i).decode() for i in range(min(emu.emu_violations(), 10)) if emu.emu_message(i)]
    assert n == 0, msgs
    assert L is native._lib


class DevBuf:
    """A device buffer of the emulated runtime, filled from a host array."""

    def __init__(self, L, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.L, self.n = L, a.size
        self.p = C.c_void_p()
        assert L.hipMalloc(C.byref(self.p), max(a.nbytes, 8)) == 0
        if a.nbytes:
            L.hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1)

    def data_ptr(self):
        return self.p.value

    def host(self):
        out = np.zeros(self.n)
        if self.n:
            self.L.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2)
        return out


def _dmrg(L, mps, mpo, itk):
    tabs = [native.Table.from_blocks(A) for A in mps]
    bufs = [DevBuf(L, t.pack(A)) for t, A in zip(tabs, mps)]
    return native.Dmrg(tabs, bufs, mpo, itk), bufs


def _download(L, d):
    out = []
    for i in range(d.N):
        nb, ln = np.zeros(1, np.int64), np.zeros(1, np.int64)
        native._check(L.mpsa_dmrg_site(d.h, i, native._ptr(nb), native._ptr(ln), None, None, None, None, None, None,
                                       None))
        n = int(nb[0])
        q, p = np.zeros(2 * n, np.uint64), np.zeros(n, np.uint32)
        r, c, o = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n, np.int64)
        buf = DevBuf(L, np.zeros(max(1, int(ln[0]))))
        native._check(L.mpsa_dmrg_site(d.h, i, native._ptr(nb), native._ptr(ln), native._ptr(q), native._ptr(p),
                                       native._ptr(r), native._ptr(c), native._ptr(o), buf.p, C.c_void_p(0)))
        out.append(native.Table(1, q.reshape(-1, 2), p.reshape(-1, 1), r, c, o).unpack(buf.host()))
    return out


@pytest.mark.parametrize("L,N,Jz,hz", [(10, 5, 0.0, 0.0), (8, 4, 1.0, 0.1)])
def test_emulated_sweeps_reach_exact_energy(lib, oracle, L, N, Jz, hz):
    mps, itk = right_normalised_mps(L, N, 64, 31)
    mpo = models.xxz_mpo(L, Jz, 1.0, hz)
    d, keep = _dmrg(lib, mps, mpo, itk)
    for _ in range(3):
        r = d.sweep(chimax=64, lanmax=300, delta=1e-12)
    if Jz == 0.0:
        eps = np.sort(np.cos(np.pi * np.arange(1, L + 1) / (L + 1)))
        e0 = eps[:N].sum()
    else:
        e0 = scipy.sparse.linalg.eigsh(_sector_H(dense.dense_H(mpo), L, N), k=1, which="SA")[0][0]
    assert abs(r["E"] - e0) < 1e-9
    assert r["tw"] < 1e-12
    assert abs(oracle.measure(_download(lib, d), mpo, itk) - r["E"]) < 1e-9
    del d


@pytest.mark.parametrize("model,L,N,chi", [("xxz", 12, 6, 8), ("tv", 14, 7, 12)])
def test_emulated_truncated_sweep_matches_oracle(lib, oracle, model, L, N, chi):
    mps, itk = right_normalised_mps(L, N, 24, 41)
    mpo = models.xxz_mpo(L, 1.0, 1.0, 0.0) if model == "xxz" else models.tv_mpo(L, 1.0, 2.0, 0.0)
    ref = pydmrg.Dmrg(mps, mpo, itk)
    rr = ref.sweep(chimax=chi, lanmax=300, delta=1e-12)
    d, keep = _dmrg(lib, mps, mpo, itk)
    r = d.sweep(chimax=chi, lanmax=300, delta=1e-12)
    assert abs(r["E"] - rr["E"]) < 1e-9
    assert abs(r["tw"] - rr["tw"]) < 1e-10
    got = _download(lib, d)
    assert [A.keys() for A in got] == [A.keys() for A in ref.mps]
    assert abs(oracle.measure(got, mpo, itk) - oracle.measure(ref.mps, mpo, itk)) < 1e-9
    del d


def test_emulation_reports_planted_mistakes(emu):
    """The emulation itself is pinned: an out-of-bounds copy, a use after free
    and a free of a foreign pointer are each reported."""
    v0 = emu.emu_violations()
    buf = DevBuf(emu, np.zeros(16))
    host = np.zeros(32)
    emu.hipMemcpy(host.ctypes.data, buf.p, 17 * 8, 2)  # one element past the end
    assert emu.emu_violations() == v0 + 1
    emu.hipFree.argtypes = [C.c_void_p]
    assert emu.hipFree(buf.p) == 0
    emu.hipMemcpy(host.ctypes.data, buf.p, 8, 2)  # read after free
    assert emu.emu_violations() == v0 + 2
    assert "use after free" in emu.emu_message(int(emu.emu_violations()) - 1).decode()
    assert emu.hipFree(C.c_void_p(host.ctypes.data)) != 0
    assert emu.emu_violations() == v0 + 3


def test_emulated_tile_classes_agree_and_autotune_keeps_ties(lib, oracle):
    """Every FP64 tile class of both GEMM stages (128, 64, 32 per workgroup,
    16 and 32 per wave) gives the same bits and the
    oracle's contraction (doContraction, src/MPSfunctions.hpp:726-804); bad
    classes are refused; the autotuner times every admissible class and, with
    the emulation's zero-duration events, keeps the plan's own classes."""
    P = problems.build(oracle, "tv", 20, 10, 60, 108)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    L, R, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(L), native.Table.from_blocks(R), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    dL, dR = DevBuf(lib, Lt.pack(L)), DevBuf(lib, Rt.pack(R))
    plan.bind_env(dL, dR)
    x = DevBuf(lib, vt.pack(psi))
    outs = {}
    for c1 in (0, 1, 2, 3, 4):
        for c2 in (0, 1, 2, 3, 4):
            plan.set_tile_class(1, c1)
            plan.set_tile_class(2, c2)
            assert (plan.stage_info()["tile_class1"], plan.stage_info()["tile_class2"]) == (c1, c2)
            y = DevBuf(lib, np.zeros(vt.total))
            plan.apply(x, y)
            outs[(c1, c2)] = y.host()
    ref = outs[(0, 0)]
    assert all(np.array_equal(v, ref) for v in outs.values())
    y_ref = oracle.contract_twosite(L, R, psi, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    y0 = vt.unpack(ref)
    num = sum(np.sum((y0[k] - y_ref[k]) ** 2) for k in y_ref.keys())
    den = sum(np.sum(y_ref[k] ** 2) for k in y_ref.keys())
    assert np.sqrt(num / den) <= 1e-12
    for bad in ((1, 5), (2, 6), (1, -1), (3, 0)):
        with pytest.raises(native.MpsaError):
            plan.set_tile_class(*bad)
    plan.set_tile_class(1, 1)
    plan.set_tile_class(2, 0)
    ms = plan.autotune(nvec=2, reps=1)
    assert set(ms) == {(s_, c_) for s_ in (1, 2) for c_ in range(5)}
    assert (plan.stage_info()["tile_class1"], plan.stage_info()["tile_class2"]) == (1, 0)
    y = DevBuf(lib, np.zeros(vt.total))
    plan.apply(x, y)
    assert np.array_equal(y.host(), ref)
    del plan


@pytest.mark.parametrize("V", [None, 0.7])
def test_emulated_hubbard_plans_match_oracle(lib, oracle, V):
    """Plans over U(1) x U(1) packed charges and d = 4 (FermiHubbardMPO): the
    two-site H_eff and both environment updates against the oracle, with every
    device access in bounds."""
    P = problems.build(oracle, "hubbard", 8, 8, 24, 33, U=4.0, V=V)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    L, R, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(L), native.Table.from_blocks(R), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    dL, dR = DevBuf(lib, Lt.pack(L)), DevBuf(lib, Rt.pack(R))
    plan.bind_env(dL, dR)
    x, y = DevBuf(lib, vt.pack(psi)), DevBuf(lib, np.zeros(vt.total))
    plan.apply(x, y)
    ref = oracle.contract_twosite(L, R, psi, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    y0 = vt.unpack(y.host())
    num = sum(np.sum((y0[k] - ref[k]) ** 2) for k in ref.keys())
    den = sum(np.sum(ref[k] ** 2) for k in ref.keys())
    assert np.sqrt(num / den) <= 1e-12
    del plan
    for direction, E, A, Wm in ((1, L, P["mps"][c], mpo[c]), (0, R, P["mps"][c + 1], mpo[c + 1])):
        Et, At = native.Table.from_blocks(E), native.Table.from_blocks(A)
        outt = native.env_outkeys(Et, At, Wm, itk[c], direction)
        ep = native.EnvPlan(Et, At, Wm, itk[c], direction, outt)
        out = DevBuf(lib, np.zeros(max(outt.total, 1)))
        ep.apply(DevBuf(lib, Et.pack(E)), DevBuf(lib, At.pack(A)), out)
        got = outt.unpack(out.host()[:outt.total] if outt.total else np.zeros(0))
        want = oracle.add_layer(E, A, Wm, itk[c], direction)
        assert got.keys() == want.keys()
        for k in want.keys():
            assert np.max(np.abs(got[k] - want[k])) <= 1e-12 * max(1.0, np.max(np.abs(want[k])))
        del ep


@pytest.mark.parametrize("model,L,N,D,seed", [("tv", 10, 5, 12, 41), ("hubbard", 6, 6, 12, 42)])
def test_emulated_complex_plans_match_oracle(lib, oracle, model, L, N, D, seed):
    """Complex128 plans (planar re/im storage, real MPO): two-site and one-site
    H_eff and both environment updates, plain and mixed, against the complex
    oracle (itself pinned against real contractions in test_complex.py), with
    every device access in bounds."""
    import complexcase as cc
    P = problems.build(oracle, model, L, N, D, seed)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lz, Rz = cc.with_imag(P["Ls"][c - 1], seed + 1), cc.with_imag(P["Rs"][c + 2], seed + 2)
    xz = cc.with_imag(P["psi"], seed + 3)
    Lt, Rt, xt = native.Table.from_blocks(Lz), native.Table.from_blocks(Rz), native.Table.from_blocks(xz)
    assert Lt.is_complex and Lt.length == 2 * Lt.total
    vt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]).as_complex()
    plan = native.HeffPlan(Lt, Rt, xt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    dL, dR = DevBuf(lib, Lt.pack(Lz)), DevBuf(lib, Rt.pack(Rz))
    plan.bind_env(dL, dR)
    x, y = DevBuf(lib, xt.pack(xz)), DevBuf(lib, np.zeros(vt.length))
    plan.apply(x, y)
    want = oracle.contract_twosite(Lz, Rz, xz, mpo[c], mpo[c + 1], itk[c], itk[c + 1], cplx=True)
    got = vt.unpack(y.host())
    assert cc.rel(got, {k: np.asarray(want[k]) for k in want.keys()}) <= 1e-12
    # FP32 complex plan: L and R converted as [re | im | -im] planes, every segment at unit scale
    p32 = native.HeffPlan(Lt, Rt, xt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F32)
    p32.bind_env(dL, dR)
    y32 = DevBuf(lib, np.zeros(vt.length))
    p32.apply(x, y32)
    assert cc.rel(vt.unpack(y32.host()), {k: np.asarray(want[k]) for k in want.keys()}) <= 1e-5
    del p32
    # BF16X3 complex plan (its GEMMs are not emulated): the schedule keeps unit
    # segment scales, so the 256x256 pipelined class stays admissible, and the
    # conversion of both planes at binding stays inside the buffers
    pb = native.HeffPlan(Lt, Rt, xt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.BF16X3)
    sched = native.describe(pb.h)
    assert np.all(sched["s1_segs"]["alpha"] == 1.0) and np.all(sched["s2_segs"]["alpha"] == 1.0)
    pb.set_tile_class(1, 3)
    pb.set_tile_class(2, 3)
    pb.bind_env(dL, dR)
    del pb
    with pytest.raises(native.MpsaError):  # every table complex, or none
        native.HeffPlan(native.Table.from_blocks(P["Ls"][c - 1]), Rt, xt, vt, (mpo[c], mpo[c + 1]),
                        (itk[c], itk[c + 1]), native.F64)
    del plan
    # one site
    R1 = cc.with_imag(P["Rs"][c + 1], seed + 4)
    Az = cc.with_imag(P["mps"][c], seed + 5)
    R1t, At = native.Table.from_blocks(R1), native.Table.from_blocks(Az)
    yt = native.heff1_outkeys(Lt, R1t, At, mpo[c], itk[c]).as_complex()
    p1 = native.HeffPlan(Lt, R1t, At, yt, mpo[c], itk[c], native.F64)
    p1.bind_env(dL, DevBuf(lib, R1t.pack(R1)))
    y1 = DevBuf(lib, np.zeros(yt.length))
    p1.apply(DevBuf(lib, At.pack(Az)), y1)
    want = oracle.contract_single(Lz, Az, mpo[c], R1, itk[c], cplx=True)
    got = yt.unpack(y1.host())
    assert cc.rel(got, {k: np.asarray(want[k]) for k in want.keys()}) <= 1e-12
    del p1
    # environments
    for direction, E, A in ((1, Lz, P["mps"][c]), (0, Rz, P["mps"][c + 1])):
        site = c if direction > 0 else c + 1
        Ez, Azz, Uz = E, cc.with_imag(A, seed + 6), cc.with_imag(A, seed + 7)
        for up in (None, Uz):
            Et, At2 = native.Table.from_blocks(Ez), native.Table.from_blocks(Azz)
            Ut = native.Table.from_blocks(up) if up is not None else None
            outt = native.env_outkeys(Et, At2, mpo[site], itk[site], direction, Ut).as_complex()
            ep = native.EnvPlan(Et, At2, mpo[site], itk[site], direction, outt, Ut)
            out = DevBuf(lib, np.zeros(max(outt.length, 1)))
            ep.apply(DevBuf(lib, Et.pack(Ez)), DevBuf(lib, At2.pack(Azz)), out,
                     DevBuf(lib, Ut.pack(up)) if up is not None else None)
            got = outt.unpack(out.host())
            want = oracle.add_layer(Ez, Azz, mpo[site], itk[site], direction, Ap=up, cplx=True)
            assert got.keys() == want.keys()
            assert cc.rel(got, {k: np.asarray(want[k]) for k in want.keys()}) <= 1e-12, (direction, up is None)
            del ep


@pytest.mark.parametrize("model,L,N,D,seed,dense,c,chimax", [("xxz", 12, 0, 48, 503, True, None, 10 ** 6),
                                                             ("xxz", 12, 0, 48, 503, True, None, 40),
                                                             ("xxz", 12, 0, 48, 503, True, None, 60),
                                                             ("tv", 16, 8, 80, 507, False, None, 10 ** 6),
                                                             ("xxz", 16, 0, 64, 508, True, 5, 10 ** 6),
                                                             ("xxz", 16, 0, 64, 508, True, 5, 50)])
def test_emulated_split_matches_oracle(lib, oracle, model, L, N, D, seed, dense, c, chimax):
    """The split's sector SVDs against the numpy restatement, with the
    density-matrix route built here for sectors with min(m, n) >= 64: the
    dense centre bond is one 96 x 96 sector of rank 48, the dense bond at
    sites 5, 6 of L = 16 one wide 64 x 128 sector.  Untruncated
    (chimax = 10^6) they go to gesvdj, which reveals the rank; count-limited
    (chimax = 40, 50, 60) to the eigendecomposition of M M^T, where
    chimax = 60 is above the rank (48) and needs the route's rank guard.  Bars: singular
    values within 1e-12 relative or 100 eps lambda_max (LAPACK on both sides
    here), the lambda-carrying tensor within 1e-10, isometries within 1e-12,
    the bare vectors weighted by lambda / lambda_max within 1e-10."""
    import pysplit
    P = problems.build(oracle, model, L, N, D, seed, c=c, dense=dense)
    c, itk, psi = P["c"], P["itk"], P["psi"]
    t = native.Table.from_blocks(psi)
    dpsi = DevBuf(lib, t.pack(psi))
    Lb = native.lib()
    itl, itr = native._itk(itk[c]), native._itk(itk[c + 1])
    h = C.c_void_p()
    native._check(Lb.mpsa_split(C.byref(t.s), dpsi.p, native._ptr(itl), len(itl), native._ptr(itr), len(itr), 1,
                                chimax, 0, 0.0, 1, C.byref(h), C.c_void_p(0)))
    tw, chi = np.zeros(1), np.zeros(1, np.int32)
    nl, nr, nlam, ll, rl = (np.zeros(1, np.int64) for _ in range(5))
    native._check(Lb.mpsa_split_info(h, native._ptr(tw), native._ptr(chi), native._ptr(nl), native._ptr(nr),
                                     native._ptr(nlam), native._ptr(ll), native._ptr(rl)))
    got = {}
    for which, n, ln, name in ((0, int(nl[0]), int(ll[0]), "left"), (1, int(nr[0]), int(rl[0]), "right")):
        q, p = np.zeros(2 * n, np.uint64), np.zeros(n, np.uint32)
        r, cc_, o = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n, np.int64)
        native._check(Lb.mpsa_split_table(h, which, native._ptr(q), native._ptr(p), native._ptr(r), native._ptr(cc_),
                                          native._ptr(o)))
        tb = native.Table(1, q.reshape(-1, 2), p.reshape(-1, 1), r, cc_, o)
        buf = DevBuf(lib, np.zeros(max(ln, 1)))
        native._check(Lb.mpsa_split_copy(h, which, buf.p, C.c_void_p(0)))
        got[name] = tb.unpack(buf.host()[:tb.total])
    k = int(nlam[0])
    lq, ls, vals = np.zeros(k, np.uint64), np.zeros(k, np.int64), np.zeros(max(1, int(chi[0])))
    native._check(Lb.mpsa_split_lambda(h, native._ptr(lq), native._ptr(ls), native._ptr(vals)))
    Lb.mpsa_split_destroy(h)
    left, lam, right, twr, chir = pysplit.split(psi, itk[c], itk[c + 1], chimax=chimax, direction=1)
    assert int(chi[0]) == chir
    assert abs(float(tw[0]) - twr) <= 1e-12
    if dense:  # one sector of 2 m(l) x 2 m(r), both >= 64: a preconditioned path ran
        assert 2 * int(t.rows.max()) >= 64 and 2 * int(t.cols.max()) >= 64
    o = 0
    lmax = max(float(np.max(v)) for v in lam.values())
    for i in range(k):
        want = np.asarray(lam[int(lq[i])])
        d = np.abs(vals[o:o + int(ls[i])] - want)
        assert np.all(d <= np.maximum(1e-12 * want, 100 * np.finfo(float).eps * lmax))
        o += int(ls[i])
    assert got["right"].keys() == right.keys()
    for key in right.keys():  # lambda rides on the right tensor
        assert np.max(np.abs(got["right"][key] - right[key]), initial=0.0) <= 1e-10, key
    for qc in lam:  # U^T U = 1 per sector
        X = np.vstack([got["left"][kk] for kk in sorted(got["left"].keys()) if kk[1] == qc])
        assert np.max(np.abs(X.T @ X - np.eye(X.shape[1]))) <= 1e-12
    # the bare left vectors, each weighted by lambda / lambda_max (fixed by the SVD where lambda matters)
    assert got["left"].keys() == left.keys()
    for key in left.keys():
        wgt = np.asarray(lam[key[1]]) / lmax
        assert np.max(np.abs((got["left"][key] - left[key]) * wgt[None, :]), initial=0.0) <= 1e-10, key


def test_emulated_chebyshev_moments_match_spectral_sum(lib, oracle):
    """mpsa_chebyshev_moments against the closed form
    mu_k = sum_n |<n|t_0>|^2 T_k((E_n - shift) / scale) from the dense H_eff's
    eigendecomposition, and the reconstructed doubled moments against the same
    closed form at 2k and 2k-1."""
    P = problems.build(oracle, "tv", 10, 5, 8, 61)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    H = problems.dense_heff(oracle, P, vt)
    assert np.max(np.abs(H - H.T)) <= 1e-12 * np.max(np.abs(H))
    E, U = np.linalg.eigh(H)
    shift, scale = 0.5 * (E[-1] + E[0]), 0.55 * (E[-1] - E[0])
    x0 = vt.pack(psi)
    w = (U.T @ x0) ** 2
    th = np.arccos(np.clip((E - shift) / scale, -1, 1))
    nmom = 40
    want = np.array([np.sum(w * np.cos(k * th)) for k in range(2 * nmom - 1)])
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    plan.bind_env(DevBuf(lib, Lt.pack(Lb)), DevBuf(lib, Rt.pack(Rb)))
    mu, rc = native.chebyshev_moments(plan, DevBuf(lib, x0), nmom, scale, shift)
    assert np.max(np.abs(mu - want[:nmom])) <= 1e-12 * want[0]
    assert np.max(np.abs(rc - want)) <= 1e-11 * want[0]
    with pytest.raises(native.MpsaError):
        native.chebyshev_moments(plan, DevBuf(lib, x0), 4, 0.0, shift)


def test_emulated_cpp_dropin(lib):
    """The C++ drop-in front end (include/mpstools_amd.hpp: contractTwosite,
    contractSingle, addContractionLayer with and without Ap, TSLanzcosEngine,
    real and complex128) against the oracle's templates, over the emulated
    device layer (tests/cpp/test_dropin.cpp; the same program runs on the GPU
    in tests/test_gpu_dropin.py)."""
    exe = os.path.join(HERE, "emu", "_build", "test_dropin_emu")
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ALL PASS" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    assert "complex128" in r.stdout
    assert "EMU VIOLATION" not in r.stderr, r.stderr[-2000:]


def test_emulated_vector_pool_stays_under_its_cap(lib, oracle):
    """Lanczos runs on bonds of three vector lengths with a small pool cap: the
    idle bytes the device vector pool keeps stay under the cap after every run
    (vectors of other lengths are freed first), and each run's lowest Ritz
    value is the dense H_eff's ground-state energy."""
    L = native.lib()
    L.mpsa_pool_info.argtypes = [C.c_void_p, C.c_void_p]
    L.mpsa_pool_set_cap.argtypes = [C.c_int64]
    native._check(L.mpsa_release_pools())
    cap = 48 * 1024
    native._check(L.mpsa_pool_set_cap(cap))
    try:
        for n, seed in [(8, 71), (10, 72), (12, 73), (8, 74)]:
            P = problems.build(oracle, "xxz", n, n // 2, 16, seed)
            c, mpo, itk = P["c"], P["mpo"], P["itk"]
            Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
            Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
            vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
            plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
            plan.bind_env(DevBuf(lib, Lt.pack(Lb)), DevBuf(lib, Rt.pack(Rb)))
            x0, st = DevBuf(lib, vt.pack(psi)), DevBuf(lib, np.zeros(vt.total))
            E = np.zeros(302)
            nit, nr = np.zeros(1, np.int32), np.zeros(1, np.int32)
            native._check(L.mpsa_lanczos(plan.h, x0.p, st.p, 300, 1e-12, 1, 0, native._ptr(E), 302, native._ptr(nit),
                                         native._ptr(nr), None))
            idle = np.zeros(1, np.int64)
            native._check(L.mpsa_pool_info(native._ptr(idle), None))
            assert 0 < idle[0] <= cap
            H = problems.dense_heff(oracle, P, vt)
            assert abs(E[0] - np.linalg.eigvalsh(H)[0]) <= 1e-9 * max(1.0, abs(E[0]))
            del plan
    finally:
        native._check(L.mpsa_pool_set_cap(8 << 30))


def test_emulated_autotune_on_a_non_default_stream(lib, oracle):
    """Tile-class tuning and applies issued on a stream of the caller's own
    (non-blocking, as torch's side streams are): the task-list rewrites and the
    workspace zeroing go on that stream; results equal the default-stream apply."""
    P = problems.build(oracle, "xxz", 12, 6, 20, 77)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    L = native.lib()
    s = C.c_void_p()
    assert lib.hipStreamCreateWithFlags(C.byref(s), 1) == 0
    ref = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    dL, dR, x = DevBuf(lib, Lt.pack(Lb)), DevBuf(lib, Rt.pack(Rb)), DevBuf(lib, vt.pack(psi))
    ref.bind_env(dL, dR)
    y0 = DevBuf(lib, np.zeros(vt.total))
    ref.apply(x, y0)
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    native._check(L.mpsa_plan_bind_env(plan.h, dL.p, dR.p, s))
    ms = np.zeros(2 * native.TILE_CLASSES)
    native._check(L.mpsa_plan_autotune(plan.h, 2, 1, native._ptr(ms), s))
    y1 = DevBuf(lib, np.zeros(vt.total))
    native._check(L.mpsa_heff_apply(plan.h, x.p, 0, y1.p, 0, 1, s))
    assert np.array_equal(y1.host(), y0.host())
    del plan, ref


def _emu_alloc(L):
    return (lambda a: DevBuf(L, a)), (lambda b: b.host())


@pytest.mark.parametrize("model,L,nvec,chimax,cplx", [("hubbard", 4, 2, 24, False), ("xxz", 6, 3, 15, False),
                                                      ("hubbard", 4, 2, 24, True)])
def test_emulated_chebyshev_engine_matches_oracle(lib, oracle, model, L, nvec, chimax, cplx):
    """The device Chebyshev engine (mpsa_cheb_*: applyMPO, computeOverlap,
    MPS sums, compressSingleSite, compress with truncation, the moment loop)
    on the CPU emulation, nvec vectors in one set of launches, against the
    oracle's restatement per vector: identical block structure, states and
    convergence measures within 1e-10, moments within 1e-10; untruncated
    moments equal the dense Chebyshev moments."""
    import chebcase
    from mpstools_amd import native
    sts, itk = chebcase.states(model, L, nvec, 40 + L, cplx=cplx)
    mpo = chebcase.scaled_mpo(model, L)
    al, th = _emu_alloc(lib)
    eng = native.Cheb(sts, mpo, itk, alloc=al, to_host=th)
    chebcase.check_steps(eng, sts, mpo, itk, chimax, 1e-10)
    eng2 = native.Cheb(sts, mpo, itk, alloc=al, to_host=th)
    chebcase.check_moments(eng2, sts, mpo, itk, 6, chimax, 1e-10)
    eng3 = native.Cheb(sts, mpo, itk, alloc=al, to_host=th)
    eng3.factor_log(100000)
    r = eng3.moments(6, 10000, 2)
    assert np.abs(r["moments"] - chebcase.dense_moments(sts, mpo, itk, 6)).max() <= 1e-10
    # the sector-shape log: this build sends sectors with k >= 3 to the
    # rocSOLVER path, the rest to the batched kernels
    recs = eng3.factor_records()
    assert recs and {r_["kind"] for r_ in recs} <= {"svd", "qr", "svd_rocsolver", "qr_rocsolver"}
    for r_ in recs:
        assert r_["sectors"] >= r_["k64"] >= r_["k128"] >= r_["k256"] >= 0
        assert 1 <= r_["max_k"] <= min(r_["max_m"], r_["max_n"])
        assert (r_["max_k"] >= 3) == r_["kind"].endswith("rocsolver")
    eng3.factor_log(0)
    assert eng3.factor_records() == []
    # every sector below the rocSOLVER threshold through the blocked SVD's
    # phases and rounds (pair table, per-sweep flags, finish)
    eng4 = native.Cheb(sts, mpo, itk, alloc=al, to_host=th)
    eng4.set_block_jacobi(1)
    r4 = eng4.moments(6, 10000, 2)
    assert np.abs(r4["moments"] - r["moments"]).max() <= 1e-12
    tm = eng4.timers()
    assert tm["heff_s"] > 0 and tm["env_s"] > 0 and tm["prepare_s"] > 0 and tm["plans"] > 0
    assert tm["syncs"] >= 1 and tm["allocs"] >= 0 and tm["alloc_s"] >= 0
    with pytest.raises(native.MpsaError):
        eng4.set_block_jacobi(-1)
    del eng, eng2, eng3, eng4


@pytest.mark.parametrize("cplx", [False, True])
def test_emulated_chebyshev_engine_lossless_invariants(lib, cplx):
    """The invariants the GPU test checks on large sectors, on the CPU
    emulation at small size (Hubbard L = 8, D = 24, 2 vectors): compress
    with chimax above every bond returns the state (conv <= 1e-10, <c|t> = 1),
    and compressSingleSite with chimax above w D gives the exact applyMPO
    state (|c - H t|^2 within 1e-10 of |H t|^2)."""
    import chebcase
    sts, itk = chebcase.states("hubbard", 8, 2, 77, D=24, cplx=cplx)
    mpo = models.rescale_mpo(models.hubbard_mpo(8, 1.0, 4.0), 4.0 * 7 + 4.0 * 8, 0.0)
    al, th = _emu_alloc(lib)
    e = native.Cheb(sts, mpo, itk, alloc=al, to_host=th)
    n0 = e.overlap(0, 0).real
    e.copy(1, 0)
    conv = e.compress(1, 10 ** 4, 2, 1e-14)
    assert np.abs(conv).max() <= 1e-10
    assert np.abs(e.overlap(1, 0) / n0 - 1.0).max() <= 1e-10
    e.apply_mpo(3, 0)
    e.compress_single_site(2, 0, 10 ** 5, 2, 1e-14)
    h2 = e.overlap(3, 3).real
    dist = e.overlap(2, 2).real + h2 - 2.0 * e.overlap(2, 3).real
    assert np.abs(dist / h2).max() <= 1e-10
    del e


def test_emulated_device_canonicalisation(lib, oracle):
    """A seeded random MPS with the sector profile's tables and entries from
    mpsa_fill_gauss (the device bits equal the host generator's), brought to
    right-normalised form by orthonormalize("rl") on the device: the state is
    unchanged up to its norm and sign, has unit norm, and every site after the
    first is right-normalised (B B^H = 1 per left charge)."""
    import chebcase
    import pycheb
    from mpstools_amd import native
    L = 6
    itk = [models.itokey_u1xu1()] * L
    tabs = native.random_mps_tables(L, itk, models.pack_charge(3, 3), 12)
    blocks = []
    for i, t in enumerate(tabs):
        b = DevBuf(lib, np.zeros(t.length))
        native.fill_gauss(b, 7 + i, t.length)
        data = b.host()
        assert np.array_equal(data, oracle.gauss(t.length, 7 + i))
        blocks.append(t.unpack(data))
    al, th = _emu_alloc(lib)
    eng = native.Cheb([blocks], models.hubbard_mpo(L), itk, alloc=al, to_host=th)
    v0 = pycheb.to_dense(pycheb.MPS(blocks, itk))
    eng.orthonormalize(0)
    got = eng.state(0, 0)
    v1 = pycheb.to_dense(pycheb.MPS(got, itk))
    assert abs(np.linalg.norm(v1) - 1.0) <= 1e-12
    assert np.abs(np.abs(np.vdot(v1, v0)) / np.linalg.norm(v0) - 1.0) <= 1e-12
    for s in range(1, L):
        secs = {}
        for k, blk in got[s].items():
            secs.setdefault(k[0], []).append(blk)
        for q, bl in secs.items():
            X = np.concatenate(bl, axis=1)
            assert np.abs(X @ X.T - np.eye(X.shape[0])).max() <= 1e-12
    del eng


def _emu_io(lib):
    return (lambda a: DevBuf(lib, a)), (lambda b: b.host())


def test_emulated_lanczos_projection(lib, oracle):
    """Projection against given states (one, two non-orthonormal, per-vector
    sets, degenerate sets rejected) against the dense spectrum and the oracle's
    projected run (tests/lzcase.py)."""
    import lzcase
    dev, host = _emu_io(lib)
    lzcase.check_projection(oracle, lzcase.case(oracle, "xxz", 10, 5, 12, 15), dev, host)


def test_emulated_lanczos_reorthogonalisation(lib, oracle):
    """Full FP64 reorthogonalisation: dim steps give the whole spectrum (the
    plain recurrence does not); short runs keep the oracle's Ritz values."""
    import lzcase
    dev, host = _emu_io(lib)
    lzcase.check_reortho(oracle, lzcase.case(oracle, "xxz", 10, 5, 12, 15), dev, host)


def test_emulated_lanczos_complex_projection(lib, oracle):
    import lzcase
    dev, host = _emu_io(lib)
    lzcase.check_complex_projection(oracle, lzcase.case(oracle, "tv", 12, 6, 16, 116), dev, host, 116)


@pytest.mark.parametrize("model,cplx", [("xxz", False), ("hubbard", False), ("hubbard", True)])
def test_emulated_twosite_product(lib, oracle, model, cplx):
    """BlockMatrix<2>(A_c, A_{c+1}) on the device layer (mpsa_twosite_product)
    against the host product of blocks.twosite: the same keys, shapes and
    values to rounding; an output block no pair produces is zeroed."""
    import complexcase as cc
    from mpstools_amd.blocks import twosite
    P = problems.build(oracle, model, 8 if model == "xxz" else 6, 4 if model == "xxz" else 6, 12, 41)
    c = P["c"]
    A1, A2 = P["mps"][c], P["mps"][c + 1]
    if cplx:
        A1, A2 = cc.with_imag(A1, 3), cc.with_imag(A2, 4)
    want = twosite(A1, A2)
    t1, t2 = native.Table.from_blocks(A1), native.Table.from_blocks(A2)
    to = native.twosite_outkeys(t1, t2)
    if cplx:
        to = to.as_complex()
    assert to.keys() == want.keys()
    d1, d2 = DevBuf(lib, t1.pack(A1)), DevBuf(lib, t2.pack(A2))
    dout = DevBuf(lib, np.full(to.length, 7.0))  # stale contents must not survive
    native.twosite_product(t1, d1, t2, d2, to, dout)
    got = to.unpack(dout.host())
    for k in want.keys():
        assert np.abs(np.asarray(got[k]) - np.asarray(want[k])).max() <= 1e-13 * max(1.0, np.abs(want[k]).max())
