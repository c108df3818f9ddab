"""Shared checks of the Lanczos options of mpsa_lanczos_batched_ex: projection
against given states (the reference's projected run, src/lanzcos.hpp:281-414)
and full FP64 reorthogonalisation.  Run on the CPU emulation of the device
layer (tests/test_emulated.py) and on the GPU (tests/test_gpu_lanczos_options.py);
buffers come from the `dev` / `host` callables of the caller.

What fixes the answers:
  * the dense H_eff of the bond (one oracle contraction per unit vector,
    restricted to the layout's non-gap coordinates) and its eigenpairs
    (E_0 < E_1 < ...);
  * projected against the exact ground state the run converges to E_1 and its
    Ritz vector is orthogonal to that state; against E_0's and E_1's states
    (given unnormalised and non-orthogonal) it converges to E_2;
  * the oracle's restatement of the projected run (same steps, same Ritz
    values to rounding);
  * with reorthogonalisation, as many steps as the space has dimensions give
    the whole spectrum as Ritz values (a basis that lost orthogonality gives
    repeated "ghost" values and misses others).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import complexcase as cc
import problems
from mpstools_amd import native


def case(oracle, model, L, N, D, seed):
    P = problems.build(oracle, model, L, N, D, seed)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    H = problems.dense_heff(oracle, P, vt)
    mask = vt.pack(vt.unpack(np.ones(vt.total))) != 0
    E, Um = np.linalg.eigh(H[np.ix_(mask, mask)])
    U = np.zeros((vt.total, len(E)))
    U[mask, :] = Um
    return dict(P=P, c=c, mpo=mpo, itk=itk, Lb=Lb, Rb=Rb, psi=psi, Lt=Lt, Rt=Rt, vt=vt, E=E, U=U, m=int(mask.sum()))


def plan_of(K, dev, Lb=None, Rb=None, vt=None):
    Lb = K["Lb"] if Lb is None else Lb
    Rb = K["Rb"] if Rb is None else Rb
    vt = K["vt"] if vt is None else vt
    c, mpo, itk = K["c"], K["mpo"], K["itk"]
    Lt, Rt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb)
    plan = native.HeffPlan(Lt, Rt, vt, vt, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), native.F64)
    keep = (dev(Lt.pack(Lb)), dev(Rt.pack(Rb)))
    plan.bind_env(*keep)
    plan._keep = keep
    return plan


def run(plan, dev, host, x0, itmax, delta, proj=None, nproj=0, proj_vstride=0, reortho=False, graph=True, target=0,
        stream=None):
    """mpsa_lanczos_batched_ex on the rows of x0 (nvec x n, host arrays);
    proj: host array of nproj (x nvec when proj_vstride) vectors of length n."""
    x0 = np.atleast_2d(np.asarray(x0, np.float64))
    nvec, n = x0.shape
    cap = itmax + 2
    E = np.zeros(nvec * cap)
    nit, nr = np.zeros(nvec, np.int32), np.zeros(nvec, np.int32)
    dx = dev(x0.ravel())
    ds = dev(np.zeros(nvec * n))
    dp = dev(np.asarray(proj, np.float64).ravel()) if proj is not None else None
    pstride = n * (nvec if proj_vstride else 1)
    native._check(native.lib().mpsa_lanczos_batched_ex(
        plan.h, C.c_void_p(dx.data_ptr()), n, nvec, C.c_void_p(ds.data_ptr()), itmax, float(delta), target,
        native._ptr(E), cap, native._ptr(nit), native._ptr(nr), int(graph),
        C.c_void_p(dp.data_ptr()) if dp is not None else None, nproj, pstride, n if proj_vstride else 0,
        int(reortho), stream))
    E = E.reshape(nvec, cap)
    return dict(it=nit.astype(int), energies=[E[v, :int(nr[v])].copy() for v in range(nvec)],
                states=host(ds).reshape(nvec, n))


def run_single(plan, dev, host, x0, itmax, delta, testenergy, proj=None, nproj=0, reortho=False, target=0,
               stream=None):
    """mpsa_lanczos_ex on one start vector (host array)."""
    x0 = np.asarray(x0, np.float64)
    n = x0.size
    cap = itmax + 2
    E = np.zeros(cap)
    nit, nr = np.zeros(1, np.int32), np.zeros(1, np.int32)
    dx, ds = dev(x0), dev(np.zeros(n))
    dp = dev(np.asarray(proj, np.float64).ravel()) if proj is not None else None
    native._check(native.lib().mpsa_lanczos_ex(
        plan.h, C.c_void_p(dx.data_ptr()), C.c_void_p(ds.data_ptr()), itmax, float(delta), int(testenergy), target,
        native._ptr(E), cap, native._ptr(nit), native._ptr(nr), C.c_void_p(dp.data_ptr()) if dp is not None else None,
        nproj, n, int(reortho), stream))
    return dict(it=int(nit[0]), energies=E[:int(nr[0])].copy(), state=host(ds))


def _close(a, b, tol):
    return abs(a - b) <= tol * max(1.0, abs(b))


def check_projection(oracle, K, dev, host, stream=None):
    E, U, vt, c = K["E"], K["U"], K["vt"], K["c"]
    plan = plan_of(K, dev)
    x0 = vt.pack(K["psi"])
    g0, g1 = U[:, 0], U[:, 1]
    # one state: the oracle's projected run, and E_1
    r = run(plan, dev, host, x0, 300, 1e-12, proj=g0, nproj=1, stream=stream)
    o = oracle.lanczos_twosite(K["Lb"], K["Rb"], K["psi"], K["mpo"][c], K["mpo"][c + 1], K["itk"][c],
                               K["itk"][c + 1], 300, 1e-12, True, state=vt.unpack(g0))
    assert r["it"][0] == o["it"]
    assert np.max(np.abs(r["energies"][0] - o["energies"])) <= 1e-9 * max(1.0, np.max(np.abs(o["energies"])))
    assert _close(r["energies"][0][0], E[1], 1e-9)
    s = r["states"][0]
    assert abs(np.dot(g0, s)) <= 1e-9 and abs(np.linalg.norm(s) - 1.0) <= 1e-9
    assert abs(abs(np.dot(U[:, 1], s)) - 1.0) <= 1e-6
    # two states given unnormalised and non-orthogonal: E_2
    r2 = run(plan, dev, host, x0, 300, 1e-12, proj=np.stack([2.0 * g0, g0 - 3.0 * g1]), nproj=2, stream=stream)
    assert _close(r2["energies"][0][0], E[2], 1e-9)
    s2 = r2["states"][0]
    assert abs(np.dot(g0, s2)) <= 1e-9 and abs(np.dot(g1, s2)) <= 1e-9
    # per-vector sets: vector 0 against g0 (E_1), vector 1 against g1 (E_0)
    r3 = run(plan, dev, host, np.stack([x0, x0]), 300, 1e-12, proj=np.stack([g0, g1]), nproj=1, proj_vstride=1,
             stream=stream)
    assert _close(r3["energies"][0][0], E[1], 1e-9) and _close(r3["energies"][1][0], E[0], 1e-9)
    # the residual test (host-driven loop) projected the same way
    r4 = run_single(plan, dev, host, x0, 300, 1e-9, False, proj=g0, nproj=1, stream=stream)
    o4 = oracle.lanczos_twosite(K["Lb"], K["Rb"], K["psi"], K["mpo"][c], K["mpo"][c + 1], K["itk"][c],
                                K["itk"][c + 1], 300, 1e-9, False, state=vt.unpack(g0))
    assert r4["it"] == o4["it"]
    assert np.max(np.abs(r4["energies"] - o4["energies"])) <= 1e-9 * max(1.0, np.max(np.abs(o4["energies"])))
    assert _close(r4["energies"][0], E[1], 1e-8) and abs(np.dot(g0, r4["state"])) <= 1e-9
    # a zero state and a dependent pair are errors
    for bad, nb in [(np.zeros_like(g0), 1), (np.stack([g0, -2.5 * g0]), 2)]:
        try:
            run(plan, dev, host, x0, 30, 1e-12, proj=bad, nproj=nb, stream=stream)
        except native.MpsaError as e:
            assert "linearly dependent" in str(e)
        else:
            raise AssertionError("a degenerate projection set was accepted")


def check_reortho(oracle, K, dev, host, stream=None):
    E, U, vt, c, m = K["E"], K["U"], K["vt"], K["c"], K["m"]
    plan = plan_of(K, dev)
    x0 = vt.pack(K["psi"])
    scale = np.max(np.abs(E))
    # dim steps, no stopping test (delta = 0): the whole spectrum, once each
    r = run(plan, dev, host, x0, m, 0.0, reortho=True, stream=stream)
    assert r["it"][0] == m and len(r["energies"][0]) == m
    assert np.max(np.abs(r["energies"][0] - E)) <= 1e-9 * scale
    assert abs(abs(np.dot(U[:, 0], r["states"][0])) - 1.0) <= 1e-9
    # the plain recurrence over as many steps does not (orthogonality is lost
    # once the extreme Ritz values converge): this case needs the option
    p = run(plan, dev, host, x0, m, 0.0, reortho=False, stream=stream)
    assert np.max(np.abs(p["energies"][0] - E)) > 1e-6 * scale
    # the host-driven loop (residual test) reorthogonalises the same way
    r = run_single(plan, dev, host, x0, m, 0.0, False, reortho=True, stream=stream)
    assert r["it"] == m and np.max(np.abs(r["energies"] - E)) <= 1e-9 * scale
    # a short run: the oracle's (plain) Ritz values to rounding
    r = run(plan, dev, host, x0, 12, 1e-12, reortho=True, stream=stream)
    o = oracle.lanczos_twosite(K["Lb"], K["Rb"], K["psi"], K["mpo"][c], K["mpo"][c + 1], K["itk"][c],
                               K["itk"][c + 1], 12, 1e-12, True)
    assert r["it"][0] == o["it"]
    assert np.max(np.abs(r["energies"][0] - o["energies"])) <= 1e-9 * scale


def check_complex_projection(oracle, K, dev, host, seed, stream=None):
    """A Hermitian complex H_eff (the real one in a gauged basis: the left
    environment built from A G with a unitary G per charge sector) has the real
    spectrum; its eigenvectors are the gauged real ones.  Projected against
    the gauged ground state (a complex vector) the run converges to E_1, and
    the Ritz vector has zero complex overlap with that state (projecting out
    the state alone, without i times it, would leave E_0 reachable)."""
    P, c, mpo, itk, E, U, vt = K["P"], K["c"], K["mpo"], K["itk"], K["E"], K["U"], K["vt"]
    mps = P["mps"]
    G = cc.bond_unitaries(mps[c - 1], seed)
    Lz = oracle.add_layer(cc.with_imag(P["Ls"][c - 2], 0, 0.0), cc.gauge_right(mps[c - 1], G), mpo[c - 1],
                          itk[c - 1], 1, cplx=True)
    Rz = cc.with_imag(K["Rb"], 0, 0.0)
    psiz = cc.with_imag(cc.gauge_left(K["psi"], G), seed + 1)
    Lt, Rt, xt = native.Table.from_blocks(Lz), native.Table.from_blocks(Rz), native.Table.from_blocks(psiz)
    vz = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    plan = plan_of(K, dev, Lz, Rz, vz)
    gz = cc.gauge_left(vt.unpack(U[:, 0]), G)
    g = vz.pack(gz)
    r = run(plan, dev, host, vz.pack(psiz), 300, 1e-12, proj=g, nproj=1, stream=stream)
    assert _close(r["energies"][0][0], E[1], 1e-9)
    s = r["states"][0]
    n = vz.total
    gc, sc = g[:n] + 1j * g[n:], s[:n] + 1j * s[n:]
    assert abs(np.vdot(gc, sc)) <= 1e-9
