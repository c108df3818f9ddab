"""Pins of the CPU oracle (the parity checker) against mathematics that does not
go through it: dense brute force, closed forms, scipy eigensolvers.

The reference ships no tests, fixtures or golden vectors for this path, so the
oracle is pinned by: (1) the dense projected Hamiltonian P^T H^T P built from
the MPO and the MPS without any block-sparse code; (2) a deliberately
non-symmetric MPO, which fails if the in/out convention is transposed;
(3) free-fermion energies of the XX chain and exact diagonalisation; (4)
scipy's tridiagonal eigensolver.  Every tolerance is stated in the assert.
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse.linalg

import dense
import problems
from mpstools_amd import models
from mpstools_amd.blocks import Blocks, twosite


def _dense_vec(B, dims_l, dims_r, d=2):
    return dense.twosite_dense(B, dims_l, dims_r, d).ravel()


@pytest.mark.parametrize("model,L,N,D,seed", [("xxz", 8, 4, 6, 1), ("asym", 8, 3, 8, 2), ("tv", 10, 5, 10, 3),
                                              ("xxz", 9, 0, 16, 4)])
def test_heff2_matches_dense_projection(oracle, model, L, N, D, seed):
    dense_flag = (N == 0)
    P = problems.build(oracle, model, L, N, D, seed, dense=dense_flag)
    c, mps, mpo, itk, dims = P["c"], P["mps"], P["mpo"], P["itk"], P["dims"]
    y = oracle.contract_twosite(P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    H, _ = dense.heff_dense(mpo, mps, dims, c, 2)
    x = _dense_vec(P["psi"], dims[c], dims[c + 2])
    yd = H @ x
    yo = _dense_vec(y, dims[c], dims[c + 2])
    assert np.max(np.abs(yo - yd)) <= 1e-13 * max(1.0, np.max(np.abs(yd)))
    # block structure: every produced key conserves charge l + q(s1) + q(s2) = r
    for (l, r, s1, s2) in y.keys():
        assert (l + itk[c][s1] + itk[c + 1][s2]) % (1 << 64) == r


def test_asym_mpo_separates_transpose(oracle):
    """The non-symmetric MPO gives H != H^T; the oracle must match H^T (the
    reference reads psi at the operator's out index and writes at in)."""
    P = problems.build(oracle, "asym", 8, 4, 8, 11)
    c, mps, mpo, itk, dims = P["c"], P["mps"], P["mpo"], P["itk"], P["dims"]
    y = oracle.contract_twosite(P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    PL = dense.left_basis(mps, dims, c, 2)
    PR = dense.right_basis(mps, dims, c + 2, 2)
    Pm = np.kron(np.kron(PL, np.eye(4)), PR.T)
    H = dense.dense_H(mpo)
    x = _dense_vec(P["psi"], dims[c], dims[c + 2])
    yo = _dense_vec(y, dims[c], dims[c + 2])
    y_T = Pm.T @ H.T @ Pm @ x
    y_H = Pm.T @ H @ Pm @ x
    assert np.max(np.abs(yo - y_T)) < 1e-13
    assert np.max(np.abs(yo - y_H)) > 1e-3  # the two conventions differ visibly


@pytest.mark.parametrize("model,L,N,D,seed", [("xxz", 8, 4, 8, 5), ("asym", 8, 4, 6, 6), ("tv", 10, 4, 12, 7)])
def test_heff1_matches_dense_projection(oracle, model, L, N, D, seed):
    P = problems.build(oracle, model, L, N, D, seed)
    c, mps, mpo, itk, dims = P["c"], P["mps"], P["mpo"], P["itk"], P["dims"]
    A = mps[c]
    y = oracle.contract_single(P["Ls"][c - 1], A, mpo[c], P["Rs"][c + 1], itk[c])
    PL = dense.left_basis(mps, dims, c, 2)
    PR = dense.right_basis(mps, dims, c + 1, 2)
    Pm = np.kron(np.kron(PL, np.eye(2)), PR.T)
    H = dense.dense_H(mpo)
    x = dense.site_dense(A, dims[c], dims[c + 1], 2).ravel()
    yd = Pm.T @ H.T @ Pm @ x
    yo = dense.site_dense(y, dims[c], dims[c + 1], 2).ravel()
    assert np.max(np.abs(yo - yd)) <= 1e-13 * max(1.0, np.max(np.abs(yd)))


@pytest.mark.parametrize("model,L,N,D,seed", [("xxz", 8, 4, 8, 8), ("asym", 7, 3, 8, 9), ("tv", 8, 4, 6, 10)])
def test_measure_and_environments_match_dense_energy(oracle, model, L, N, D, seed):
    """MPO::measure (left environments) and the right-environment stack both give
    <psi|H^T|psi> of the dense state."""
    P = problems.build(oracle, model, L, N, D, seed, c=2)
    mps, mpo, itk, dims = P["mps"], P["mpo"], P["itk"], P["dims"]
    psi_full = dense.left_basis(mps, dims, L, 2)[:, 0] if len(dims[L]) == 1 else None
    H = dense.dense_H(mpo)
    E_d = psi_full @ H.T @ psi_full
    E_o = oracle.measure(mps, mpo, itk)
    assert abs(E_o - E_d) <= 1e-12 * max(1.0, abs(E_d))
    Rs = oracle.right_envs(mps, mpo, itk, 0)
    E_r = sum(np.trace(v) for k, v in Rs[0].items() if k[0] == k[1])
    assert abs(E_r - E_d) <= 1e-12 * max(1.0, abs(E_d))


def test_env_left_block_matches_dense(oracle):
    """L[c-1] blocks against the dense left-block operators: for each aux w,
    L_w = P_L^T O_w^T P_L with O_w the partial MPO product over sites < c."""
    P = problems.build(oracle, "asym", 7, 3, 8, 12, c=3)
    c, mps, mpo, dims = P["c"], P["mps"], P["mpo"], P["dims"]
    PL = dense.left_basis(mps, dims, c, 2)
    T = mpo[0].dense()[0]  # (w, o, i)
    ops = [T[w] for w in range(T.shape[0])]
    for s in range(1, c):
        Wd = mpo[s].dense()
        ops = [sum(np.kron(ops[wl], Wd[wl, wr]) for wl in range(Wd.shape[0])) for wr in range(Wd.shape[1])]
    off = dense.sector_offsets(dims[c])[0]
    Lc = P["Ls"][c - 1]
    for (a, b, w), blk in Lc.items():
        full = PL.T @ ops[w].T @ PL
        oa, ma = off[a]
        ob, mb = off[b]
        assert np.max(np.abs(full[oa:oa + ma, ob:ob + mb] - blk)) < 1e-13


def _sector_H(H, L, N):
    idx = [i for i in range(2 ** L) if bin(i).count("1") == N]
    return H[np.ix_(idx, idx)]


@pytest.mark.parametrize("L,N", [(10, 5), (8, 4), (12, 6)])
def test_lanczos_xx_chain_free_fermions(oracle, L, N):
    """XX chain (Jz=0): E0 = sum of the N lowest eps_k = Jxy cos(pi k/(L+1)).
    At half filling with full bond dimension the centre two-site problem spans
    the whole charge sector, so its ground state is the chain's."""
    P = problems.build(oracle, "xxz", L, N, 64, 13, Jz=0.0, Jxy=1.0)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    r = oracle.lanczos_twosite(P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1],
                               itmax=200, delta=1e-12, testenergy=True)
    eps = np.sort(np.cos(np.pi * np.arange(1, L + 1) / (L + 1)))
    assert abs(r["energies"][0] - eps[:N].sum()) < 1e-10


def test_lanczos_xxz_matches_exact_diagonalisation(oracle):
    L, N = 8, 4
    P = problems.build(oracle, "xxz", L, N, 64, 14, Jz=1.0, Jxy=1.0, hz=0.1)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    r = oracle.lanczos_twosite(P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1],
                               itmax=150, delta=1e-12, testenergy=True)
    Hs = _sector_H(dense.dense_H(mpo), L, N)
    e0 = scipy.sparse.linalg.eigsh(Hs, k=1, which="SA")[0][0]
    assert abs(r["energies"][0] - e0) < 1e-10
    # ground state from getState is an eigenvector: residual of the dense problem
    gs = r["gs"]
    assert abs(gs.norm() - 1.0) < 1e-8


def test_lanczos_off_half_filling_matches_projected_problem(oracle):
    """Away from half filling the Schmidt-bounded bases do not span the sector;
    the Lanczos energy is then the lowest eigenvalue of the projected problem,
    an upper bound of the free-fermion energy."""
    L, N = 10, 3
    P = problems.build(oracle, "xxz", L, N, 64, 13, Jz=0.0, Jxy=1.0)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    r = oracle.lanczos_twosite(P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1],
                               itmax=200, delta=1e-12, testenergy=True)
    H, _ = dense.heff_dense(mpo, P["mps"], P["dims"], c, 2)
    mask = Blocks(2)
    for k, v in P["psi"].items():
        mask[k] = np.ones_like(v)
    support = _dense_vec(mask, P["dims"][c], P["dims"][c + 2]) > 0  # charge-conserving (l, s1, s2, r)
    w = np.linalg.eigvalsh(H[np.ix_(support, support)])
    assert abs(r["energies"][0] - w[0]) < 1e-10
    eps = np.sort(np.cos(np.pi * np.arange(1, L + 1) / (L + 1)))
    assert r["energies"][0] >= eps[:N].sum() - 1e-12


def test_lanczos_residual_criterion_and_itmax(oracle):
    """testenergy=False uses the Ritz-vector residual; itmax caps the steps."""
    P = problems.build(oracle, "xxz", 10, 5, 12, 15)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    args = (P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    r10 = oracle.lanczos_twosite(*args, itmax=10, delta=1e-12, testenergy=True)
    assert r10["it"] == 10 and len(r10["energies"]) == 10
    r = oracle.lanczos_twosite(*args, itmax=300, delta=1e-9, testenergy=False)
    H, _ = dense.heff_dense(mpo, P["mps"], P["dims"], c, 2)
    w = np.linalg.eigvalsh(0.5 * (H + H.T))
    assert abs(r["energies"][0] - w[0]) < 1e-8


def test_lanczos_projected_run_reaches_the_first_excited_state(oracle):
    """The projected run (src/lanzcos.hpp:281-414) against the exact ground
    state of the dense H_eff converges to the second eigenvalue, its Ritz
    vector is that eigenvector and orthogonal to the ground state, for the
    energy test and the residual test; given the state scaled by 3, the same.
    A run that skipped the projection of the start vector or of a recurrence
    vector would find E_0 again (rounding-level components grow)."""
    import lzcase
    K = lzcase.case(oracle, "xxz", 10, 5, 12, 15)
    E, U, vt, c = K["E"], K["U"], K["vt"], K["c"]
    args = (K["Lb"], K["Rb"], K["psi"], K["mpo"][c], K["mpo"][c + 1], K["itk"][c], K["itk"][c + 1])
    g0 = U[:, 0]
    assert E[1] - E[0] > 1e-3 and E[2] - E[1] > 1e-3
    plain = oracle.lanczos_twosite(*args, itmax=300, delta=1e-12, testenergy=True)
    assert abs(plain["energies"][0] - E[0]) <= 1e-9
    for testenergy, tol in [(True, 1e-9), (False, 1e-8)]:
        for f in (1.0, 3.0):
            r = oracle.lanczos_twosite(*args, itmax=300, delta=1e-12 if testenergy else 1e-9, testenergy=testenergy,
                                       state=vt.unpack(f * g0))
            assert abs(r["energies"][0] - E[1]) <= tol
            s = vt.pack(r["gs"])
            assert abs(np.dot(g0, s)) <= 1e-8
            assert abs(abs(np.dot(U[:, 1], s)) / np.linalg.norm(s) - 1.0) <= 1e-6


def test_tridiag_matches_scipy(oracle):
    rng = np.random.default_rng(0)
    d, e = rng.standard_normal(30), rng.standard_normal(29)
    w, z = oracle.tridiag(d, e)
    w2, z2 = scipy.linalg.eigh_tridiagonal(d, e)
    assert np.max(np.abs(w - w2)) < 1e-12
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.max(np.abs(T @ z - z * w)) < 1e-12


def test_twosite_and_dot(oracle):
    P = problems.build(oracle, "xxz", 8, 4, 8, 16)
    c, mps = P["c"], P["mps"]
    ts = twosite(mps[c], mps[c + 1])
    h1 = oracle.BM.from_blocks(mps[c])
    h2 = oracle.BM.from_blocks(mps[c + 1])
    out = oracle.BM(2)
    oracle._check(oracle.lib().orc_twosite(out.h, h1.h, h2.h))
    o = out.to_blocks()
    assert o.keys() == ts.keys()
    for k in ts.keys():
        assert np.max(np.abs(o[k] - ts[k])) < 1e-14
    assert abs(oracle.dot(ts, ts) - ts.dot(ts)) < 1e-12


def test_add_layer_rejects_empty_tensor(oracle):
    P = problems.build(oracle, "xxz", 6, 3, 4, 17)
    with pytest.raises(RuntimeError, match="A is empty"):
        oracle.add_layer(P["Ls"][P["c"] - 1], Blocks(1), P["mpo"][P["c"]], P["itk"][0], 1)


def test_models_match_dense_spin_chain(oracle):
    """HeisenbergMPO reproduces H = sum Jz SzSz + Jxy/2 (S+S- + S-S+) - B Sz."""
    L = 5
    mpo = models.heisenberg_mpo([0.9, 1.1, 0.7, 1.3], [1.0, 0.5, 0.8, 1.2], [0.1, -0.2, 0.3, 0.0, 0.5])
    H = dense.dense_H(mpo)
    sz = np.diag([-0.5, 0.5])
    sp = np.array([[0, 0], [1, 0.]])
    sm = sp.T

    def op(o, i):
        m = np.eye(1)
        for s in range(L):
            m = np.kron(m, o if s == i else np.eye(2))
        return m
    Jz, Jxy, B = [0.9, 1.1, 0.7, 1.3], [1.0, 0.5, 0.8, 1.2], [0.1, -0.2, 0.3, 0.0, 0.5]
    Href = sum(Jz[i] * op(sz, i) @ op(sz, i + 1) + Jxy[i] / 2 * (op(sp, i) @ op(sm, i + 1) + op(sm, i) @ op(sp, i + 1))
               for i in range(L - 1)) - sum(B[i] * op(sz, i) for i in range(L))
    assert np.max(np.abs(H - Href)) < 1e-14


# ---------------------------------------------------------------- split oracle
def _dense_twosite(psi, itk_l, itk_r):
    """psi as one dense matrix over row labels (l, s1) and column labels (r, s2)."""
    rows, cols = {}, {}
    for (l, r, s1, s2) in psi.keys():
        rows[(l, s1)] = psi[(l, r, s1, s2)].shape[0]
        cols[(r, s2)] = psi[(l, r, s1, s2)].shape[1]
    ro, o = {}, 0
    for k in sorted(rows):
        ro[k] = o
        o += rows[k]
    co, p = {}, 0
    for k in sorted(cols):
        co[k] = p
        p += cols[k]
    M = np.zeros((o, p))
    for (l, r, s1, s2) in psi.keys():
        b = psi[(l, r, s1, s2)]
        M[ro[(l, s1)]:ro[(l, s1)] + b.shape[0], co[(r, s2)]:co[(r, s2)] + b.shape[1]] = b
    return M


@pytest.mark.parametrize("model,L,N,D,seed", [("xxz", 10, 5, 12, 401), ("tv", 12, 6, 20, 402)])
def test_split_oracle_exact_decomposition(oracle, model, L, N, D, seed):
    import pysplit
    P = problems.build(oracle, model, L, N, D, seed)
    c, itk, psi = P["c"], P["itk"], P["psi"]
    left, lam, right, tw, chi = pysplit.split(psi, itk[c], itk[c + 1], chimax=10 ** 6, normalize=False, direction=1)
    # singular values: the union over sectors is the spectrum of the dense matrix
    sv = np.sort(np.concatenate([v for v in lam.values()]))[::-1]
    ref = np.linalg.svd(_dense_twosite(psi, itk[c], itk[c + 1]), compute_uv=False)
    ref = ref[ref > 1e-15]
    assert len(sv) == len(ref) == chi
    assert np.allclose(sv, ref, rtol=1e-12, atol=1e-14)
    assert abs(tw - (1.0 - np.sum(ref ** 2))) <= 1e-12
    # reconstruction psi(l, r, s1, s2) = U[(l, qc, s1)] (lambda V')[(qc, r, s2)]
    for (l, r, s1, s2) in psi.keys():
        qc = (l + itk[c][s1]) % (1 << 64)
        got = left[(l, qc, s1)] @ right[(qc, r, s2)]
        assert np.allclose(got, psi[(l, r, s1, s2)], atol=1e-12)
    # left isometry per sector, FixPhase sign rule on V' = lambda^-1 (lambda V')
    for qc, v in lam.items():
        us = [left[k] for k in left.keys() if k[1] == qc]
        G = sum(u.T @ u for u in us)
        assert np.allclose(G, np.eye(len(v)), atol=1e-12)
        vt = np.concatenate([right[k] for k in right.keys() if k[0] == qc], axis=1) / v[:, None]
        for i in range(vt.shape[0]):
            j = int(np.argmax(np.abs(vt[i])))
            assert vt[i, j] > 0


def test_split_oracle_truncation_rules(oracle):
    import pysplit
    P = problems.build(oracle, "xxz", 12, 6, 24, 403)
    c, itk, psi = P["c"], P["itk"], P["psi"]
    full = np.linalg.svd(_dense_twosite(psi, itk[c], itk[c + 1]), compute_uv=False)
    for chi in (1, 5, 17):
        left, lam, right, tw, cnt = pysplit.split(psi, itk[c], itk[c + 1], chimax=chi, direction=0)
        kept = np.sort(np.concatenate([v for v in lam.values()]))[::-1]
        assert cnt == chi == len(kept)
        assert abs(np.sum(kept ** 2) - 1.0) <= 1e-12  # normalised
        top = full[:chi]
        assert np.allclose(kept, top / np.linalg.norm(top), rtol=1e-12)
        assert abs(tw - (1.0 - np.sum(top ** 2))) <= 1e-12
        # direction 0: lambda sits in the left tensor, the right one is row-orthonormal
        for qc in lam:
            vt = np.concatenate([right[k] for k in right.keys() if k[0] == qc], axis=1)
            assert np.allclose(vt @ vt.T, np.eye(vt.shape[0]), atol=1e-12)
    # weight mode: stop once the remaining weight drops below tr_weight * norm
    nrm = np.sqrt(np.sum(full ** 2))
    for w in (1e-1, 1e-3):
        _, lam, _, tw, cnt = pysplit.split(psi, itk[c], itk[c + 1], chimax=1000, do_trunc=True, tr_weight=w)
        rem = nrm - np.cumsum(full ** 2) / nrm
        expect = int(np.argmax(rem <= nrm * w)) + 1
        assert cnt == expect


# ---------------------------------------------------------------- DMRG oracle
def right_normalised_mps(L, N, D, seed, Jz=1.0):
    """Seeded random MPS, right-normalised by LQ sweeps, unit norm."""
    from mpstools_amd import models, synth
    itk = [models.itokey_u1()] * L
    mps, dims = synth.random_mps(L, itk, N, D, seed)
    for i in range(L - 1, 0, -1):
        mps[i], mps[i - 1] = synth.right_normalize(mps[i], mps[i - 1])
    mps[0] = mps[0].scaled(1.0 / mps[0].norm())
    return mps, itk


@pytest.mark.parametrize("L,N,Jz,hz", [(10, 5, 0.0, 0.0), (8, 4, 1.0, 0.1)])
def test_dmrg_oracle_sweeps_reach_exact_energy(oracle, L, N, Jz, hz):
    """The restated DMRG sweep with untruncated bonds converges to the exact
    ground state: free fermions for the XX chain, exact diagonalisation of the
    Sz sector for XXZ with a field."""
    import pydmrg
    from mpstools_amd import models
    mps, itk = right_normalised_mps(L, N, 64, 31)
    mpo = models.xxz_mpo(L, Jz, 1.0, hz)
    d = pydmrg.Dmrg(mps, mpo, itk)
    for _ in range(3):
        r = d.sweep(chimax=64, lanmax=300, delta=1e-12)
    if Jz == 0.0:
        eps = np.sort(np.cos(np.pi * np.arange(1, L + 1) / (L + 1)))
        e0 = eps[:N].sum()
    else:
        Hs = _sector_H(dense.dense_H(mpo), L, N)
        e0 = scipy.sparse.linalg.eigsh(Hs, k=1, which="SA")[0][0]
    assert abs(r["E"] - e0) < 1e-9
    assert r["tw"] < 1e-12
    # the final state is normalised and its energy (measure) is the DMRG energy
    assert abs(oracle.measure(d.mps, mpo, itk) - r["E"]) < 1e-9
