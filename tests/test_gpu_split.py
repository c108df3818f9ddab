"""Device split of a two-site tensor (mpsa_split) against the numpy
restatement of split + FixPhase + truncate + dropDangling (oracle/pysplit.py),
on the same seeded two-site tensors.

Bar: identical left/right key tables and block shapes, kept count and
truncated weight; singular values within 1e-12 relative; the site tensors
within 1e-10 (the random spectra here are non-degenerate, so FixPhase fixes
every singular vector).
"""
import numpy as np
import pytest

import problems
import pysplit
from mpstools_amd import native

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _cmp_blocks(t, buf, ref, tol):
    got = t.unpack(buf.cpu().numpy())
    assert got.keys() == ref.keys()
    for k in ref.keys():
        assert got[k].shape == ref[k].shape, k
        assert np.max(np.abs(got[k] - ref[k]), initial=0.0) <= tol, k


CASES = [("xxz", 10, 5, 12, 501, False), ("tv", 14, 7, 30, 502, False), ("xxz", 12, 0, 48, 503, True)]


@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES)
@pytest.mark.parametrize("direction,chimax,do_trunc,tw", [(1, 10 ** 6, False, 0.0), (0, 7, False, 0.0),
                                                          (1, 1000, True, 1e-2)])
def test_split_matches_oracle(gpu, oracle, model, L, N, D, seed, dense, direction, chimax, do_trunc, tw):
    torch = gpu
    _say("split", model, L, N, D, seed, "dir", direction, "chimax", chimax, "trunc", do_trunc)
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
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
    for qc in lam:
        assert np.allclose(out["lam"][qc], lam[qc], rtol=1e-12, atol=1e-15)
    _cmp_blocks(*out["left"], left, 1e-10)
    _cmp_blocks(*out["right"], right, 1e-10)
