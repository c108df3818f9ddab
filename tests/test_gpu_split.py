"""Device split of a two-site tensor (mpsa_split) against the numpy
restatement of split + FixPhase + truncate + dropDangling (oracle/pysplit.py),
on the same seeded two-site tensors.

Bar: identical left/right key tables and block shapes, kept count and
truncated weight; singular values within 1e-12 relative or 100 eps lambda_max
absolute (Weyl: a backward-stable SVD moves every singular value by at most
the size of its backward error, so small singular values are fixed only
absolutely); the tensor that
carries lambda within 1e-10; the other (the bare singular vectors) within
1e-10 after weighting each vector by lambda / lambda_max, and orthonormal
within 1e-12.  A singular vector is determined only to about
eps * sigma_max / gap by any backward-stable SVD; the untruncated splits keep
every sigma above 1e-15, and the rank-deficient dense sector below has
clusters of such sigmas whose vectors no two algorithms agree on, while the
state (vectors weighted by lambda) and the isometry are fixed.  FixPhase fixes
the sign of every well-separated vector.
"""
import numpy as np
import pytest

import problems
import pysplit
from mpstools_amd import native
from mpstools_amd.blocks import twosite

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _cmp_blocks(t, buf, ref, tol):
    got = t.unpack(buf.cpu().numpy())
    assert got.keys() == ref.keys()
    for k in ref.keys():
        assert got[k].shape == ref[k].shape, k
        assert np.max(np.abs(got[k] - ref[k]), initial=0.0) <= tol, k


def _cmp_lambda(lam_got, lam_ref):
    lmax = max(float(np.max(v)) for v in lam_ref.values())
    atol = 100 * np.finfo(float).eps * lmax
    for qc in lam_ref:
        d = np.abs(np.asarray(lam_got[qc]) - np.asarray(lam_ref[qc]))
        assert np.all(d <= np.maximum(1e-12 * np.asarray(lam_ref[qc]), atol)), (qc, float(np.max(d)), atol)


def _cmp_weighted(tb, ref, lam_got, lam_ref, side, tol):
    """Bare singular vectors against the oracle's, each weighted by its
    lambda / lambda_max: columns of the left blocks (l, qc, s1), rows of the
    right blocks (qc, r, s2)."""
    t, buf = tb
    got = t.unpack(buf.cpu().numpy())
    assert got.keys() == ref.keys()
    lmax = max(float(np.max(v)) for v in lam_ref.values())
    for k in ref.keys():
        qc = k[1] if side == "left" else k[0]
        w = np.asarray(lam_ref[qc]) / lmax
        # the lambda bar of _cmp_lambda (1e-12 relative or 100 eps lambda_max), in lambda_max units
        assert np.all(np.abs(np.asarray(lam_got[qc]) / lmax - w) <= np.maximum(1e-12 * w, 100 * np.finfo(float).eps))
        d = got[k] - ref[k]
        d = d * w[None, :] if side == "left" else d * w[:, None]
        assert np.max(np.abs(d), initial=0.0) <= tol, k


def _check_isometry(tb, lam, side, tol):
    """U^T U = 1 per sector (left), V' V'^T = 1 per sector (right)."""
    t, buf = tb
    got = t.unpack(buf.cpu().numpy())
    for qc in lam:
        blocks = [got[k] for k in sorted(got.keys()) if (k[1] if side == "left" else k[0]) == qc]
        X = np.vstack(blocks) if side == "left" else np.hstack(blocks)
        G = X.T @ X if side == "left" else X @ X.T
        assert np.max(np.abs(G - np.eye(G.shape[0]))) <= tol, qc


# the dense centre bonds are one (2D) x (2D) sector (96 x 96, 256 x 256); the
# dense bond at sites 5, 6 of L = 16 is one wide 64 x 128 sector; all of them
# below the preconditioning threshold (1024 columns): gesvdj directly
CASES = [("xxz", 10, 5, 12, 501, False, None), ("tv", 14, 7, 30, 502, False, None),
         ("xxz", 12, 0, 48, 503, True, None), ("xxz", 16, 0, 128, 505, True, None), ("xxz", 16, 0, 64, 508, True, 5)]


@pytest.mark.parametrize("model,L,N,D,seed,dense,c", CASES)
@pytest.mark.parametrize("direction,chimax,do_trunc,tw", [(1, 10 ** 6, False, 0.0), (0, 7, False, 0.0),
                                                          (1, 1000, True, 1e-2)])
def test_split_matches_oracle(gpu, oracle, model, L, N, D, seed, dense, c, direction, chimax, do_trunc, tw):
    torch = gpu
    _say("split", model, L, N, D, seed, "c", c, "dir", direction, "chimax", chimax, "trunc", do_trunc)
    P = problems.build(oracle, model, L, N, D, seed, c=c, dense=dense)
    c, itk, psi = P["c"], P["itk"], P["psi"]
    t = native.Table.from_blocks(psi)
    dpsi = torch.from_numpy(t.pack(psi)).cuda()
    out = native.split(t, dpsi, itk[c], itk[c + 1], direction=direction, chimax=chimax, do_trunc=do_trunc,
                       tr_weight=tw)
    left, lam, right, twr, chi = pysplit.split(psi, itk[c], itk[c + 1], chimax=chimax, do_trunc=do_trunc,
                                               tr_weight=tw, direction=direction)
    assert out["chi"] == chi
    assert abs(out["tw"] - twr) <= 1e-12
    assert sorted(out["lam"]) == sorted(lam)
    _cmp_lambda(out["lam"], lam)
    # lambda rides on the right tensor for direction > 0, on the left one otherwise
    bare, carrier = ("left", "right") if direction > 0 else ("right", "left")
    _cmp_blocks(*out[carrier], {"left": left, "right": right}[carrier], 1e-10)
    _cmp_weighted(out[bare], {"left": left, "right": right}[bare], out["lam"], lam, bare, 1e-10)
    _check_isometry(out[bare], out["lam"], bare, 1e-12)


@pytest.mark.parametrize("model,L,N,D,seed,dense,chi", [("xxz", 12, 0, 48, 504, True, 23),
                                                      ("tv", 14, 7, 30, 505, False, 20)])
def test_split_of_a_ground_state_truncated(gpu, oracle, model, L, N, D, seed, dense, chi):
    """The split a sweep performs: the bond's Lanczos ground state (the
    oracle's), whose Schmidt values decay, truncated to chi.  The spectrum
    comes in SU(2) multiplets split only by the Lanczos residual, so the
    vectors inside a multiplet are not determined (any rotation of them is an
    SVD); chi sits at a gap of the spectrum (23: 59 %, 20: 39 %) so that the
    kept subspace is.  The bars are the determined quantities: kept count, truncated
    weight, lambda, the truncated two-site tensor rebuilt from the two site
    tensors (BlockMatrix<2>, src/BlockTensor.hpp:285-301) within 1e-12, the
    left isometry, and the right tensor's Gram matrix diag(lambda^2)."""
    torch = gpu
    _say("split ground state", model, L, N, D, chi)
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    r = oracle.lanczos_twosite(P["Ls"][c - 1], P["Rs"][c + 2], P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1],
                               itmax=300, delta=1e-12, testenergy=True)
    gs = r["gs"]
    t = native.Table.from_blocks(gs)
    out = native.split(t, torch.from_numpy(t.pack(gs)).cuda(), itk[c], itk[c + 1], direction=1, chimax=chi)
    left, lam, right, twr, kept = pysplit.split(gs, itk[c], itk[c + 1], chimax=chi, direction=1)
    assert out["chi"] == kept == chi
    assert abs(out["tw"] - twr) <= 1e-12
    assert sorted(out["lam"]) == sorted(lam)
    _cmp_lambda(out["lam"], lam)
    gl = out["left"][0].unpack(out["left"][1].cpu().numpy())
    gr = out["right"][0].unpack(out["right"][1].cpu().numpy())
    assert gl.keys() == left.keys() and gr.keys() == right.keys()
    got2, ref2 = twosite(gl, gr), twosite(left, right)
    assert got2.keys() == ref2.keys()
    assert max(np.max(np.abs(got2[k] - ref2[k])) for k in ref2.keys()) <= 1e-12
    _check_isometry(out["left"], out["lam"], "left", 1e-12)
    for qc, v in out["lam"].items():
        X = np.hstack([gr[k] for k in sorted(gr.keys()) if k[0] == qc])
        assert np.max(np.abs(X @ X.T - np.diag(np.asarray(v) ** 2))) <= 1e-12, qc



def test_split_preconditioned_sector_within_tolerance(gpu, oracle):
    """A 1024 x 1024 sector (the dense D = 512 bond), untruncated, so gesvdj
    decides the rank.  Its accuracy is that of gesvdj's stopping test, an
    absolute bound on the off-diagonal of M^T M, not rounding level, so the
    bars are: lambda within 1e-9 lambda_max, the lambda-carrying tensor and
    the lambda-weighted bare vectors within 1e-9, both isometries within 1e-12,
    and the rebuilt psi within 1e-9 of its norm."""
    torch = gpu
    _say("split preconditioned 1024 x 1024")
    P = problems.build(oracle, "xxz", 22, 0, 512, 509, dense=True)
    c, itk, psi = P["c"], P["itk"], P["psi"]
    t = native.Table.from_blocks(psi)
    assert 2 * int(t.cols.max()) >= 1024
    dpsi = torch.from_numpy(t.pack(psi)).cuda()
    out = native.split(t, dpsi, itk[c], itk[c + 1], direction=1, chimax=10 ** 6)
    left, lam, right, twr, chi = pysplit.split(psi, itk[c], itk[c + 1], chimax=10 ** 6, direction=1)
    assert out["chi"] == chi
    lmax = max(float(np.max(v)) for v in lam.values())
    for qc in lam:
        assert np.max(np.abs(np.asarray(out["lam"][qc]) - np.asarray(lam[qc]))) <= 1e-9 * lmax
    _cmp_blocks(*out["right"], right, 1e-9)
    _cmp_weighted(out["left"], left, out["lam"], lam, "left", 1e-9)
    _check_isometry(out["left"], out["lam"], "left", 1e-12)
    gl = out["left"][0].unpack(out["left"][1].cpu().numpy())
    gr = out["right"][0].unpack(out["right"][1].cpu().numpy())
    rebuilt = twosite(gl, gr)
    num = sum(float(np.sum((rebuilt[k] - psi[k]) ** 2)) for k in psi.keys())
    den = sum(float(np.sum(psi[k] ** 2)) for k in psi.keys())
    assert np.sqrt(num / den) <= 1e-9  # psi is normalised, so lambda is not rescaled


@pytest.mark.parametrize("chimax,kept", [(600, 512), (300, 300)])
@pytest.mark.parametrize("direction", [1, 0])
def test_split_density_route_count_limited(gpu, oracle, direction, chimax, kept):
    """The same 1024 x 1024 sector (rank 512: smallest nonzero singular
    value 3.2e-4, the rest LAPACK noise below 3e-17) split count-limited, so
    it takes the density-matrix route: syevd of M M^T (lambda right) or
    M^T M (lambda left) and P = M^T E or M E.  chimax = 600 needs the
    route's zero level (1e-10 sigma_max: the directions outside the range come
    out near 3e-14 there), chimax = 300 cuts inside the spectrum.  Bars: kept count equal, truncated
    weight within 1e-10, lambda within 1e-9 lambda_max, the lambda-carrying
    tensor and the lambda-weighted bare vectors within 1e-9, the bare
    isometry within 1e-12, and the truncated state within 1e-9 of the
    oracle's."""
    torch = gpu
    _say("split density route 1024 x 1024, chimax", chimax, "direction", direction)
    P = problems.build(oracle, "xxz", 22, 0, 512, 509, dense=True)
    c, itk, psi = P["c"], P["itk"], P["psi"]
    t = native.Table.from_blocks(psi)
    dpsi = torch.from_numpy(t.pack(psi)).cuda()
    out = native.split(t, dpsi, itk[c], itk[c + 1], direction=direction, chimax=chimax)
    left, lam, right, twr, chi = pysplit.split(psi, itk[c], itk[c + 1], chimax=chimax, direction=direction)
    assert out["chi"] == chi == kept
    assert abs(out["tw"] - twr) <= 1e-10
    lmax = max(float(np.max(v)) for v in lam.values())
    for qc in lam:
        assert np.max(np.abs(np.asarray(out["lam"][qc]) - np.asarray(lam[qc]))) <= 1e-9 * lmax
    if direction > 0:
        _cmp_blocks(*out["right"], right, 1e-9)
        _cmp_weighted(out["left"], left, out["lam"], lam, "left", 1e-9)
        _check_isometry(out["left"], out["lam"], "left", 1e-12)
    else:
        _cmp_blocks(*out["left"], left, 1e-9)
        _cmp_weighted(out["right"], right, out["lam"], lam, "right", 1e-9)
        _check_isometry(out["right"], out["lam"], "right", 1e-12)
    gl = out["left"][0].unpack(out["left"][1].cpu().numpy())
    gr = out["right"][0].unpack(out["right"][1].cpu().numpy())
    got, want = twosite(gl, gr), twosite(left, right)
    assert got.keys() == want.keys()
    assert max(np.max(np.abs(got[k] - want[k])) for k in want.keys()) <= 1e-9
