"""The oracle's restatement of the Chebyshev engine with MPS compression
(oracle/pycheb.py: applyMPO, computeOverlap, MPS sums, prepareSite, truncate,
compress, compressSingleSite, simulateSSCompress) pinned by what the
mathematics fixes, independently of the block-sparse code:

  * applyMPO is H |t> exactly: the dense vector of the result equals the
    dense Hamiltonian (tests/dense.py) times the dense vector of t, for the
    Fermi-Hubbard (U(1) x U(1), d = 4) and XXZ chains, real and complex;
  * computeOverlap(a, b) = <b|a> of the dense vectors; 2a - b likewise;
  * prepareSite leaves the state unchanged and makes the site an isometry;
  * untruncated compressSingleSite returns H |t>, untruncated compress the
    state itself;
  * with chimax below the bond dimension, truncate keeps the global top-chi
    Schmidt values of each bond (checked against the dense bipartition SVD
    on a one-bond cut) and compress returns a state of that bond dimension
    whose overlap with the input is at least the truncated one's;
  * the moments of an untruncated run are the Chebyshev moments
    mu_n = <t_0|T_n(H)|t_0> from the dense three-term recurrence, and the
    reconstructed doubled moments equal mu_{2n} and mu_{2n-1} of that same
    recurrence (a sign slip in t_n = 2 H t_{n-1} - t_{n-2} or a wrong index in
    the reconstruction fails them).
CPU only.
"""
import numpy as np
import pytest

import dense
import pycheb
from mpstools_amd import models, synth


def _mps(model, L, seed, D=16, cplx=False):
    if model == "hubbard":
        itk = [models.itokey_u1xu1()] * L
        N = models.pack_charge(L // 2, L // 2)
    else:
        itk = [models.itokey_u1()] * L
        N = L // 2
    mps, _ = synth.random_mps(L, itk, N, D, seed)
    for i in range(L - 1, 0, -1):
        mps[i], mps[i - 1] = synth.right_normalize(mps[i], mps[i - 1])
    mps[0] = mps[0].scaled(1.0 / mps[0].norm())
    M = pycheb.MPS(mps, itk)
    if cplx:
        rng = np.random.default_rng(seed + 1)
        for A in M.A:
            for k in A.keys():
                b = A[k]
                A.dtype = np.dtype(np.complex128)
                A.b[k] = np.asfortranarray(b + 0.3j * rng.standard_normal(b.shape))
    return M


def _mpo(model, L):
    if model == "hubbard":
        return models.hubbard_mpo(L, 1.0, 2.0, 0.2, -0.1)
    return models.xxz_mpo(L, 0.8, 1.0, 0.15)


CASES = [("hubbard", 4, 3, False), ("xxz", 6, 4, False), ("hubbard", 4, 5, True)]


@pytest.mark.parametrize("model,L,seed,cplx", CASES)
def test_apply_mpo_is_dense_h_times_t(model, L, seed, cplx):
    M = _mps(model, L, seed, cplx=cplx)
    mpo = _mpo(model, L)
    H = dense.dense_H(mpo)
    v = pycheb.to_dense(M)
    t = pycheb.apply_mpo(mpo, M)
    assert np.abs(pycheb.to_dense(t) - H @ v).max() <= 1e-13 * max(1.0, np.abs(H @ v).max())
    # the bond dimension is the MPO's w times the input's, per charge at most
    for b in range(L - 1):
        assert t.chi(b) <= mpo[b].Dr * M.chi(b)


@pytest.mark.parametrize("model,L,seed,cplx", CASES)
def test_overlap_and_sum_are_the_dense_ones(model, L, seed, cplx):
    a = _mps(model, L, seed, cplx=cplx)
    b = pycheb.apply_mpo(_mpo(model, L), _mps(model, L, seed + 7, cplx=cplx))
    va, vb = pycheb.to_dense(a), pycheb.to_dense(b)
    assert abs(pycheb.compute_overlap(a, b) - np.vdot(vb, va)) <= 1e-12 * np.linalg.norm(va) * np.linalg.norm(vb)
    s = pycheb.add(pycheb.scale(a, 2.0), pycheb.scale(b, -1.0))
    assert np.abs(pycheb.to_dense(s) - (2 * va - vb)).max() <= 1e-12 * np.abs(vb).max()
    # the sum of a state with one missing a charge sector is still the dense sum
    c = a.copy()
    k = c.A[1].keys()[0]
    del c.A[1].b[k]
    for kk in list(c.A[1].b):
        if kk[0] == k[0]:
            del c.A[1].b[kk]
    vc = pycheb.to_dense(c)
    s = pycheb.add(b, c)
    assert np.abs(pycheb.to_dense(s) - (vb + vc)).max() <= 1e-12 * np.abs(vb).max()


@pytest.mark.parametrize("model,L,seed,cplx", CASES)
def test_prepare_site_keeps_the_state_and_makes_isometries(model, L, seed, cplx):
    M = _mps(model, L, seed, cplx=cplx)
    v = pycheb.to_dense(M)
    for s in range(L - 1):
        pycheb.prepare_site(M, s, "l")
        assert np.abs(pycheb.to_dense(M) - v).max() <= 1e-13
        secs = {}
        for k, blk in M.A[s].items():
            secs.setdefault(k[1], []).append(blk)
        for q, bl in secs.items():
            X = np.concatenate(bl, axis=0)
            assert np.abs(X.conj().T @ X - np.eye(X.shape[1])).max() <= 1e-12
    for s in range(L - 1, 0, -1):
        pycheb.prepare_site(M, s, "r")
        assert np.abs(pycheb.to_dense(M) - v).max() <= 1e-13
        secs = {}
        for k, blk in M.A[s].items():
            secs.setdefault(k[0], []).append(blk)
        for q, bl in secs.items():
            X = np.concatenate(bl, axis=1)
            assert np.abs(X @ X.conj().T - np.eye(X.shape[0])).max() <= 1e-12


@pytest.mark.parametrize("model,L,seed,cplx", CASES)
def test_untruncated_compressions_are_exact(model, L, seed, cplx):
    M = _mps(model, L, seed, cplx=cplx)
    mpo = _mpo(model, L)
    v = pycheb.to_dense(M)
    Hv = dense.dense_H(mpo) @ v
    c, conv = pycheb.compress_single_site(M.copy(), mpo, chimax=10000, maxstep=3)
    assert np.abs(pycheb.to_dense(c) - Hv).max() <= 1e-12 * np.abs(Hv).max()
    # conv measures against the input state (the reference's definition)
    n1sq = np.vdot(Hv, Hv).real
    n2sq = pycheb.compute_overlap(c, c).real
    assert abs(conv - (n1sq + n2sq - 2 * abs(np.vdot(v, Hv))) / n1sq) <= 1e-10
    s = pycheb.add(pycheb.scale(c, 2.0), pycheb.scale(M, -1.0))
    cc, conv2 = pycheb.compress(s.copy(), chimax=10000, maxstep=3)
    assert np.abs(pycheb.to_dense(cc) - (2 * Hv - v)).max() <= 1e-12 * np.abs(Hv).max()
    assert abs(conv2) <= 1e-12


def test_truncate_keeps_the_global_top_chi_schmidt_values():
    """One-bond check: after the canonicalising sweeps the lambdas of bond 1
    are the singular values of the dense bipartition (sites 0-1 | 2-3) of the
    state, and truncate keeps the chi largest over all charge sectors."""
    M = _mps("hubbard", 4, 9)
    v = pycheb.to_dense(M)
    for s in range(3):
        pycheb.prepare_site(M, s, "l")
    for s in range(3, 0, -1):
        pycheb.prepare_site(M, s, "r")
    sv = np.linalg.svd(v.reshape(16, 16), compute_uv=False)
    lam = np.sort(np.concatenate([x for x in M.lam[1].values()]))[::-1]
    k = len(lam)
    assert np.abs(lam - sv[:k]).max() <= 1e-12 and np.all(sv[k:] <= 1e-12)
    chi = 5
    kept = pycheb.lambda_truncate(M.lam[1], chi)
    vals = np.sort(np.concatenate([x for x in kept.values()]))[::-1]
    assert np.abs(vals - sv[:chi]).max() <= 1e-12
    T = M.copy()
    pycheb.truncate(T, chi)
    assert T.chi(1) == chi


def test_truncated_compress_has_the_bond_dimension_and_fidelity():
    M = _mps("hubbard", 4, 11)
    mpo = _mpo("hubbard", 4)
    t = pycheb.apply_mpo(mpo, M)
    v = pycheb.to_dense(t)
    for chi in (4, 8):
        c, conv = pycheb.compress(t.copy(), chimax=chi, maxstep=4)
        assert max(c.chi(b) for b in range(3)) <= chi
        vc = pycheb.to_dense(c)
        fid = abs(np.vdot(vc, v)) / (np.linalg.norm(vc) * np.linalg.norm(v))
        # the variational state is at least as close as the truncated start
        T = t.copy()
        for s in range(3):
            pycheb.prepare_site(T, s, "l")
        for s in range(2, 0, -1):
            pycheb.prepare_site(T, s, "r")
        pycheb.truncate(T, chi)
        vt = pycheb.to_dense(T)
        fid0 = abs(np.vdot(vt, v)) / (np.linalg.norm(vt) * np.linalg.norm(v))
        assert fid >= fid0 - 1e-10
        assert abs(conv - (np.vdot(v, v).real + np.vdot(vc, vc).real - 2 * abs(np.vdot(v, vc))) / np.vdot(v, v).real) <= 1e-9


@pytest.mark.parametrize("model,L,seed,cplx", CASES)
def test_untruncated_moments_are_the_dense_chebyshev_moments(model, L, seed, cplx):
    M = _mps(model, L, seed, cplx=cplx)
    mpo = _mpo(model, L)
    H = dense.dense_H(mpo)
    ev = np.linalg.eigvalsh(H)
    a, b = 1.05 * 0.5 * (ev[-1] - ev[0]), 0.5 * (ev[-1] + ev[0])
    Ht = models.rescale_mpo(mpo, a, b)
    Hd = (H - b * np.eye(len(H))) / a
    nmom = 7
    r = pycheb.chebyshev_moments(Ht, M, nmom, chimax=10000, ncmax=3)
    t0 = pycheb.to_dense(M)
    ts = [t0, Hd @ t0]
    for _ in range(2 * nmom):
        ts.append(2 * Hd @ ts[-1] - ts[-2])
    mu = np.array([np.vdot(t, t0) for t in ts])
    assert np.abs(r["moments"] - mu[:nmom]).max() <= 1e-11
    rec = r["recons"]
    assert np.abs(rec[0] - mu[0]) <= 1e-11 and np.abs(rec[1] - mu[1]) <= 1e-11 and np.abs(rec[2] - mu[2]) <= 1e-11
    for n in range(2, nmom):
        assert np.abs(rec[2 * n - 1] - mu[2 * n - 1]) <= 1e-10
        assert np.abs(rec[2 * n] - mu[2 * n]) <= 1e-10
    # |mu_n| <= mu_0 for a spectrum inside [-1, 1]
    assert np.all(np.abs(r["moments"]) <= abs(r["moments"][0]) + 1e-12)
