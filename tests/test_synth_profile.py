"""The seeded generator's sector profiles (mpstools_amd.synth.bond_dims): the
default profile apportions D over all charge sectors of a bond by weight; the
clamped profile of the t-V workload with about 1700 states keeps the 24
heaviest sectors of every bond with dimensions in [20, 120].  Both are
consistent (no sector larger than what its neighbours' sectors and the site's
charges can reach), which the MPS builders rely on."""
import numpy as np

from mpstools_amd import models, synth


def _consistent(dims, itk):
    L = len(itk)
    for l in range(1, L + 1):
        for r, m in dims[l].items():
            assert m <= sum(dims[l - 1].get(synth.qsub(r, qs), 0) for qs in itk[l - 1])
    for l in range(L - 1, -1, -1):
        for q, m in dims[l].items():
            assert m <= sum(dims[l + 1].get(synth.qadd(q, qs), 0) for qs in itk[l])


def test_clamped_profile_of_the_tv_workload():
    L, N = 100, 50
    itk = [models.itokey_u1()] * L
    dims = synth.bond_dims(L, itk, N, 8000, (20, 120), 24)
    _consistent(dims, itk)
    centre = dims[L // 2]
    assert len(centre) == 24 and sum(centre.values()) == 1698
    assert min(centre.values()) == 20 and max(centre.values()) == 120
    # the kept sectors are the 24 charges nearest half filling (binomial weights)
    qs = sorted(centre)
    assert qs == list(range(qs[0], qs[0] + 24)) and abs((qs[0] + qs[-1]) / 2 - N / 2) <= 0.5
    # near the ends the exact state counts cap the dimensions
    assert sum(dims[1].values()) == 2 and sum(dims[L - 1].values()) == 2
    for l in range(L + 1):
        assert all(1 <= m <= 120 for m in dims[l].values())


def test_default_profile_is_unchanged_by_the_clamp_arguments():
    L, N = 32, 16
    itk = [models.itokey_u1()] * L
    a = synth.bond_dims(L, itk, N, 64)
    b = synth.bond_dims(L, itk, N, 64, None, None)
    assert a == b
    _consistent(a, itk)
    assert max(sum(d.values()) for d in a) == 64
    mps1, _ = synth.random_mps(L, itk, N, 64, 7)
    mps2, _ = synth.random_mps(L, itk, N, 64, 7)
    for s1, s2 in zip(mps1, mps2):
        assert s1.keys() == s2.keys()
        for k in s1.keys():
            assert np.array_equal(s1[k], s2[k])
