"""The device Chebyshev engine (mpsa_cheb_*) on the GPU: applyMPO,
computeOverlap, MPS sums, compressSingleSite and compress with truncation,
and the moment loop of simulateSSCompress (src/container.hpp:2343-2470),
for several start vectors in one set of grouped launches, against the
oracle's restatement (oracle/pycheb.py).

Bars: identical block structure; states (as dense vectors), convergence
measures and moments within 1e-10 relative in FP64 (the sector SVD is the
batched Jacobi kernel here and LAPACK in the oracle, so agreement is to
rounding); untruncated moments equal the dense Chebyshev moments within
1e-10; with the one-site contraction in FP32 (complex64) the moments agree
with FP64 within 1e-5.  A mid-size chain (Hubbard L = 16, D = 48, 4 vectors)
checks |mu_n| <= mu_0 (spectrum in [-1, 1]) and the bond dimensions.

Parity is unpinned where the compression start (applyMPO of the truncated
state, canonicalised) has charge sectors with singular values at rounding
level: prepareSite keeps all min(m, n) directions of a sector
(src/MPSfunctions.hpp:60-75), so the variational space then contains basis
vectors fixed only by how the SVD completes a rank-deficient basis, and the
reference's own result depends on its LAPACK build (seen on XXZ L = 8 at
chimax = 20: singular values 2e-36 and 5e-19 at two bonds, moments differing
by 1e-5 between two LAPACK-based runs).  The cases here are well conditioned.
"""
import numpy as np
import pytest

import chebcase
from mpstools_amd import models, native

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _engine(sts, mpo, itk, big, **kw):
    e = native.Cheb(sts, mpo, itk, **kw)
    if big == "blocked":
        e.set_block_jacobi(2)
    elif big:
        e.set_big_sector(big)
    return e


@pytest.mark.parametrize("big", [0, 3, "blocked"])
@pytest.mark.parametrize("model,L,nvec,chimax,cplx", [("hubbard", 4, 2, 24, False), ("xxz", 6, 3, 15, False),
                                                      ("hubbard", 4, 2, 24, True), ("xxz", 8, 2, 40, True)])
def test_cheb_engine_steps_and_moments_match_oracle(gpu, oracle, model, L, nvec, chimax, cplx, big):
    """big = 3 sends every sector of at least 3 columns to the rocSOLVER path
    (gesvd / geqrf+ungqr / gelqf+unglq) that the engine uses for large sectors;
    "blocked" sends every SVD sector of at least 2 columns to the blocked
    Jacobi SVD (several workgroups per sector) that the engine uses for
    sectors of 32 to 512 columns (QR and LQ: 32 or more)."""
    _say("cheb engine", model, L, nvec, chimax, cplx, big)
    sts, itk = chebcase.states(model, L, nvec, 40 + L, cplx=cplx)
    mpo = chebcase.scaled_mpo(model, L)
    chebcase.check_steps(_engine(sts, mpo, itk, big), sts, mpo, itk, chimax, 1e-10)
    chebcase.check_moments(_engine(sts, mpo, itk, big), sts, mpo, itk, 6, chimax, 1e-10)
    r = _engine(sts, mpo, itk, big).moments(6, 10000, 2)
    assert np.abs(r["moments"] - chebcase.dense_moments(sts, mpo, itk, 6)).max() <= 1e-10


def test_cheb_engine_fp32_one_site_contraction(gpu, oracle):
    _say("cheb engine fp32")
    sts, itk = chebcase.states("hubbard", 4, 2, 71, cplx=True)
    mpo = chebcase.scaled_mpo("hubbard", 4)
    r64 = native.Cheb(sts, mpo, itk).moments(6, 24, 3)
    r32 = native.Cheb(sts, mpo, itk, h1_precision=native.F32).moments(6, 24, 3)
    assert np.abs(r32["moments"] - r64["moments"]).max() <= 1e-5


def test_cheb_engine_mid_size_properties(gpu):
    _say("cheb engine mid size")
    L = 16
    sts, itk = chebcase.states("hubbard", L, 4, 91, D=48)
    raw = models.hubbard_mpo(L, 1.0, 4.0)
    # |H| <= 2 t (L - 1) * 2 + U L bounds the spectrum; rescale into [-1, 1]
    bound = 4.0 * (L - 1) + 4.0 * L
    mpo = models.rescale_mpo(raw, bound, 0.0)
    eng = native.Cheb(sts, mpo, itk)
    eng.factor_log(100000)
    r = eng.moments(5, 96, 2)
    recs = eng.factor_records()  # sector shapes: all in the batched kernels at the default threshold
    assert recs and {x["kind"] for x in recs} <= {"svd", "qr"}
    assert all(1 <= x["max_k"] <= min(x["max_m"], x["max_n"]) for x in recs)
    mu = r["moments"]
    assert np.all(np.abs(mu) <= np.abs(mu[:, :1]) * (1 + 1e-3))
    # mu_2 = 2 <t1|t1> - mu_0 is reconstructed exactly from the compressed t1
    assert np.all(np.isfinite(r["recons"]))
    # the last compressed state's bond dimension is at most chimax
    for v in range(4):
        st = eng.state(1, v)
        for s in range(L - 1):
            dims = {}
            for k, b in st[s].items():
                dims[k[1]] = b.shape[1]
            assert sum(dims.values()) <= 96


@pytest.mark.parametrize("cplx,block", [(True, None), (False, None), (True, 16), (False, 0)])
def test_cheb_engine_large_sector_invariants(gpu, cplx, block):
    """Sector matrices of up to a few hundred rows (Hubbard L = 24, D = 512,
    U(1) x U(1), 2 device-generated start vectors) through the batched
    Householder QR and the Jacobi SVD on R^H: the start vectors have unit
    norm after the device orthonormalisation; compress with chimax above
    every bond (canonical form by SVD and QR, then sweeps) returns the same
    state (conv <= 1e-10, <c|t> = 1 within 1e-10); compressSingleSite with
    chimax above w D gives the exact applyMPO state (|c - H t|^2 within 1e-10
    of |H t|^2).  block: the blocked Jacobi SVD's threshold (None: the
    default 32, 0: every sector in the one-workgroup kernel)."""
    _say("cheb engine large sectors", cplx, block)
    L, D = 24, 512
    itk = [models.itokey_u1xu1()] * L
    N = models.pack_charge(L // 2, L // 2)
    mpo = models.rescale_mpo(models.hubbard_mpo(L, 1.0, 4.0), 4.0 * (L - 1) + 4.0 * L, 0.0)
    e = native.ChebRandom(L, itk, N, D, 2, mpo, 424242, cplx=cplx)
    if block is not None:
        e.set_block_jacobi(block)
    e.factor_log(100000)
    assert np.abs(e.overlap(0, 0) - 1.0).max() <= 1e-12
    e.copy(1, 0)
    conv = e.compress(1, 8 * D, 2, 1e-14)
    assert np.abs(conv).max() <= 1e-10
    assert np.abs(e.overlap(1, 0) - 1.0).max() <= 1e-10
    assert np.abs(e.overlap(1, 1) - 1.0).max() <= 1e-10
    # compressSingleSite(H, t) with chimax above w D against the exact applyMPO
    e.apply_mpo(3, 0)
    e.compress_single_site(2, 0, 8 * 6 * D, 2, 1e-14)
    h2 = e.overlap(3, 3).real
    dist = e.overlap(2, 2).real + h2 - 2.0 * e.overlap(2, 3).real
    assert np.abs(dist / h2).max() <= 1e-10
    recs = e.factor_records()
    assert {x["kind"] for x in recs} == {"svd", "qr"}
    assert max(x["max_k"] for x in recs) >= 32 and max(x["max_m"] for x in recs) >= 100
