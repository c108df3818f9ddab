"""The device two-site product BlockMatrix<2>(A_c, A_{c+1})
(src/BlockTensor.hpp:285-301; mpsa_twosite_product) on the GPU against the
host product of blocks.twosite (numpy, pinned in tests/test_oracle.py):
identical key sets and shapes, values within 1e-13 relative (FP64 MFMA
against numpy's BLAS), U(1), U(1) x U(1), complex128, and a dense D = 512
pair (no symmetry, 4 blocks of 512 x 512)."""
import numpy as np
import pytest

import complexcase as cc
import problems
from mpstools_amd import native
from mpstools_amd.blocks import Blocks, twosite

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _check(A1, A2):
    want = twosite(A1, A2)
    got = native.twosite(A1, A2)
    assert got.keys() == want.keys()
    for k in want.keys():
        w = np.asarray(want[k])
        assert np.asarray(got[k]).shape == w.shape
        assert np.abs(np.asarray(got[k]) - w).max() <= 1e-13 * max(1.0, np.abs(w).max())


@pytest.mark.parametrize("model,L,N,D,seed,cplx", [("xxz", 16, 8, 48, 51, False), ("hubbard", 10, 10, 40, 52, False),
                                                   ("tv", 20, 10, 60, 53, True)])
def test_twosite_product_matches_host(gpu, oracle, model, L, N, D, seed, cplx):
    _say("twosite", model, L, N, D, cplx)
    P = problems.build(oracle, model, L, N, D, seed)
    c = P["c"]
    A1, A2 = P["mps"][c], P["mps"][c + 1]
    if cplx:
        A1, A2 = cc.with_imag(A1, seed), cc.with_imag(A2, seed + 1)
    _check(A1, A2)


def test_twosite_product_dense_512(gpu):
    _say("twosite dense 512")
    rng = np.random.default_rng(54)
    D = 512
    A1, A2 = Blocks(1), Blocks(1)
    for s in range(2):
        A1[(0, 0, s)] = np.asfortranarray(rng.standard_normal((D, D)))
        A2[(0, 0, s)] = np.asfortranarray(rng.standard_normal((D, D)))
    _check(A1, A2)
