"""Shared checks of the device Chebyshev engine (native.Cheb over the C ABI)
against the oracle's restatement (oracle/pycheb.py), run on the CPU emulation
of the device layer (tests/test_emulated.py) and on the GPU
(tests/test_gpu_cheb_engine.py).  Inputs only here; each check states its bar.  Cases are chosen so that no
charge sector of a compression start is rank deficient (see the GPU test's
docstring for why the reference's result is not reproducible there).
"""
from __future__ import annotations

import numpy as np

import dense
import pycheb
from mpstools_amd import models, synth


def states(model, L, nvec, seed, D=16, cplx=False):
    """nvec seeded, right-normalised MPS of the model's charge sector."""
    if model == "hubbard":
        itk = [models.itokey_u1xu1()] * L
        N = models.pack_charge(L // 2, L // 2)
    else:
        itk = [models.itokey_u1()] * L
        N = L // 2
    out = []
    for v in range(nvec):
        mps, _ = synth.random_mps(L, itk, N, D, seed + 17 * v)
        for i in range(L - 1, 0, -1):
            mps[i], mps[i - 1] = synth.right_normalize(mps[i], mps[i - 1])
        mps[0] = mps[0].scaled(1.0 / mps[0].norm())
        if cplx:
            rng = np.random.default_rng(seed + 17 * v + 1)
            for A in mps:
                for k in A.keys():
                    b = A[k]
                    A.dtype = np.dtype(np.complex128)
                    A.b[k] = np.asfortranarray(b + 0.3j * rng.standard_normal(b.shape))
        out.append(mps)
    return out, itk


def scaled_mpo(model, L):
    """The model's MPO rescaled so that its spectrum lies in [-0.95, 0.95]."""
    mpo = models.hubbard_mpo(L, 1.0, 2.0, 0.2, -0.1) if model == "hubbard" else models.xxz_mpo(L, 0.8, 1.0, 0.15)
    ev = np.linalg.eigvalsh(dense.dense_H(mpo))
    a, b = 0.5 * (ev[-1] - ev[0]) / 0.95, 0.5 * (ev[-1] + ev[0])
    return models.rescale_mpo(mpo, a, b)


def rel(x, y):
    x, y = np.asarray(x), np.asarray(y)
    return float(np.max(np.abs(x - y)) / max(1e-300, np.max(np.abs(y))))


def same_structure(got, ref):
    return all(g.keys() == r.keys() and all(g[k].shape == r[k].shape for k in g.keys()) for g, r in zip(got, ref))


def check_steps(eng, sts, mpo, itk, chimax, tol):
    """applyMPO, computeOverlap, the MPS sum, compressSingleSite and compress
    of every vector against the oracle: identical block structure, states (as
    dense vectors) and conv values within `tol` relative."""
    ref = [pycheb.MPS(s, itk) for s in sts]
    # applyMPO: exact structure, dense state
    eng.apply_mpo(2, 0)
    for v, M in enumerate(ref):
        t = pycheb.apply_mpo(mpo, M)
        got = eng.state(2, v)
        assert same_structure(got, t.A)
        assert rel(pycheb.to_dense(pycheb.MPS(got, itk)), pycheb.to_dense(t)) <= tol
    ov = eng.overlap(2, 0)
    for v, M in enumerate(ref):
        assert abs(ov[v] - pycheb.compute_overlap(pycheb.apply_mpo(mpo, M), M)) <= tol * max(1.0, abs(ov[v]))
    # compressSingleSite with truncation
    conv = eng.compress_single_site(3, 0, chimax, 3)
    comps = []
    for v, M in enumerate(ref):
        c, cv = pycheb.compress_single_site(M.copy(), mpo, chimax, 3)
        comps.append(c)
        got = eng.state(3, v)
        assert same_structure(got, c.A)
        assert rel(pycheb.to_dense(pycheb.MPS(got, itk)), pycheb.to_dense(c)) <= tol
        assert abs(conv[v] - cv) <= tol * max(1.0, abs(cv))
    # 2 comp - t0, then compress in place
    eng.axpby(4, 2.0, 3, -1.0, 0)
    conv2 = eng.compress(4, chimax, 3)
    for v, M in enumerate(ref):
        s = pycheb.add(pycheb.scale(comps[v], 2.0), pycheb.scale(M, -1.0))
        cc, cv2 = pycheb.compress(s.copy(), chimax, 3)
        got = eng.state(4, v)
        assert same_structure(got, cc.A)
        assert rel(pycheb.to_dense(pycheb.MPS(got, itk)), pycheb.to_dense(cc)) <= tol
        assert abs(conv2[v] - cv2) <= tol * max(1.0, abs(cv2)) + 1e-12


def check_moments(eng, sts, mpo, itk, nmom, chimax, tol):
    """simulateSSCompress moments, reconstructed moments and both convergence
    measures of every vector against the oracle within `tol` (relative to mu_0
    = 1 for the moments)."""
    r = eng.moments(nmom, chimax, 3)
    for v, s in enumerate(sts):
        ro = pycheb.chebyshev_moments(mpo, pycheb.MPS(s, itk), nmom, chimax, 3)
        assert np.abs(r["moments"][v] - ro["moments"]).max() <= tol
        assert np.abs(r["recons"][v] - ro["recons"]).max() <= tol * max(1.0, np.abs(ro["recons"]).max())
        assert np.abs(r["conv1"][v] - np.array(ro["conv1"])).max() <= tol
        assert np.abs(r["conv2"][v] - np.array(ro["conv2"])).max() <= tol
    return r


def dense_moments(sts, mpo, itk, nmom):
    """mu_n = <t_0|T_n(H)|t_0> from the dense three-term recurrence."""
    Hd = dense.dense_H(mpo)
    out = []
    for s in sts:
        t0 = pycheb.to_dense(pycheb.MPS(s, itk))
        ts = [t0, Hd @ t0]
        while len(ts) < nmom:
            ts.append(2 * Hd @ ts[-1] - ts[-2])
        out.append(np.array([np.vdot(t, t0) for t in ts[:nmom]]))
    return np.array(out)
