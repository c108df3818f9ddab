"""The Lanczos options of mpsa_lanczos_batched_ex on the GPU: projection
against given states (the reference's projected run, src/lanzcos.hpp:281-414)
and full FP64 reorthogonalisation, through the C ABI.

Small bonds: the checks of tests/lzcase.py against the dense spectrum and the
oracle's projected run (E_1 and E_2 within 1e-9, Ritz vectors orthogonal to
the given states within 1e-9, the whole spectrum from dim reorthogonalised
steps within 1e-9 of the largest eigenvalue, complex plans projected in the
complex inner product).  At the size of the XXZ L = 32, D = 64 bond: the
projected run from the device's own ground state against the oracle's
projected run from the same state (lowest Ritz value within 1e-9), graph
replay and step-by-step launches alike.
"""
import numpy as np
import pytest
import torch

import lzcase
import problems
from mpstools_amd import native

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_lanczos_projection_small(gpu, oracle):
    _say("lanczos projection xxz 10")
    lzcase.check_projection(oracle, lzcase.case(oracle, "xxz", 10, 5, 12, 15), _dev, _host, native._stream())


def test_lanczos_reorthogonalisation_small(gpu, oracle):
    _say("lanczos reortho xxz 10")
    lzcase.check_reortho(oracle, lzcase.case(oracle, "xxz", 10, 5, 12, 15), _dev, _host, native._stream())


def test_lanczos_complex_projection_small(gpu, oracle):
    _say("lanczos complex projection tv 12")
    lzcase.check_complex_projection(oracle, lzcase.case(oracle, "tv", 12, 6, 16, 116), _dev, _host, 116,
                                    native._stream())


def test_lanczos_projection_config1_size(gpu, oracle):
    _say("lanczos projection xxz 32 D64")
    P = problems.build(oracle, "xxz", 32, 16, 64, 5)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lb, Rb, psi = P["Ls"][c - 1], P["Rs"][c + 2], P["psi"]
    Lt, Rt, xt = native.Table.from_blocks(Lb), native.Table.from_blocks(Rb), native.Table.from_blocks(psi)
    vt = xt.union(native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1]))
    K = dict(c=c, mpo=mpo, itk=itk, Lb=Lb, Rb=Rb, vt=vt)
    plan = lzcase.plan_of(K, _dev)
    st = native._stream()
    x0 = vt.pack(psi)
    g = lzcase.run(plan, _dev, _host, x0, 300, 1e-12, stream=st)
    gs = g["states"][0]
    args = (Lb, Rb, psi, mpo[c], mpo[c + 1], itk[c], itk[c + 1], 300, 1e-12, True)
    o = oracle.lanczos_twosite(*args, state=vt.unpack(gs), nthreads=8)
    for graph in (True, False):
        r = lzcase.run(plan, _dev, _host, x0, 300, 1e-12, proj=gs, nproj=1, graph=graph, stream=st)
        e = r["energies"][0]
        assert abs(e[0] - o["energies"][0]) <= 1e-9 * max(1.0, abs(o["energies"][0]))
        assert e[0] > g["energies"][0][0] + 1e-6
        s = r["states"][0]
        assert abs(np.dot(gs, s)) <= 1e-9 and abs(np.linalg.norm(s) - 1.0) <= 1e-9
    # reorthogonalised: the same lowest Ritz value as the plain run
    r = lzcase.run(plan, _dev, _host, x0, 300, 1e-12, reortho=True, stream=st)
    assert abs(r["energies"][0][0] - g["energies"][0][0]) <= 1e-9 * max(1.0, abs(g["energies"][0][0]))
