"""Python restatement of the reference two-site DMRG sweep on the oracle.

TEST INFRASTRUCTURE ONLY (like the rest of oracle/): imported by tests/.

Follows DMRGengine::doStep and the sweep loops of DMRGengine::simulate
(src/container.hpp:293-337, 383-438): for each bond the two-site tensor
BlockMatrix<2>(A[s], A[s+1]) (src/BlockTensor.hpp:285-301), the Lanczos of
TSLanzcosEngine (oracle lanczos_twosite), getState(0), split (pysplit), lambda
into the site ahead (container.hpp:321-328) and addContractionLayer for the
environment of the sweep direction.
"""
from __future__ import annotations

import pyoracle
import pysplit


def _twosite(A1, A2):
    from mpstools_amd.blocks import twosite
    return twosite(A1, A2)


class Dmrg:
    def __init__(self, mps, mpo, itk):
        self.mps = list(mps)
        self.mpo, self.itk = mpo, itk
        N = len(mps)
        self.L = {-1: pyoracle.lmin()}
        self.R = {N: pyoracle.rmax(self.mps[-1])}
        for i in range(N - 1, 1, -1):
            self.R[i] = pyoracle.add_layer(self.R[i + 1], self.mps[i], mpo[i], itk[i], 0)

    def step(self, s, d, chimax, do_trunc, tr_weight, lanmax, delta):
        mps, mpo, itk = self.mps, self.mpo, self.itk
        psi = _twosite(mps[s], mps[s + 1])
        r = pyoracle.lanczos_twosite(self.L[s - 1], self.R[s + 2], psi, mpo[s], mpo[s + 1], itk[s], itk[s + 1],
                                     itmax=lanmax, delta=delta, testenergy=True)
        left, lam, right, tw, chi = pysplit.split(r["gs"], itk[s], itk[s + 1], chimax=chimax, do_trunc=do_trunc,
                                                  tr_weight=tr_weight, direction=1 if d > 0 else 0)
        mps[s], mps[s + 1] = left, right
        if d > 0:
            self.L[s] = pyoracle.add_layer(self.L[s - 1], mps[s], mpo[s], itk[s], 1)
        else:
            self.R[s + 1] = pyoracle.add_layer(self.R[s + 2], mps[s + 1], mpo[s + 1], itk[s + 1], 0)
        return r["energies"][0], tw, r["it"]

    def sweep(self, chimax, do_trunc=False, tr_weight=0.0, lanmax=500, delta=1e-12):
        N = len(self.mps)
        out = None
        tws = []
        for s in range(0, N - 2):
            out = self.step(s, 1, chimax, do_trunc, tr_weight, lanmax, delta)
            tws.append(out[1])
        for s in range(N - 2, 0, -1):
            out = self.step(s, -1, chimax, do_trunc, tr_weight, lanmax, delta)
            tws.append(out[1])
        return dict(E=float(out[0]), tw=max(tws))
