"""Seeded synthetic MPS inputs for the DMRG hot path.

Sector profile: for bond l (after l sites) of an L-site chain with total
charge N, the charges allowed are those reachable from the left boundary
(charge 0) and able to reach N on the right.  count_left[l][n] is the number
of left basis states with charge n, count_right[l][n] the number of right
states completing it to N; the exact Schmidt bound is
min(count_left, count_right) and the weight count_left * count_right.  If the
exact bounds fit in D they are the dimensions; otherwise D is apportioned in
proportion to the weights (largest remainder), capped at the exact bounds,
and sectors that get 0 are dropped.  A final pass makes every site tensor
reshapable to an isometry from either side.

Block entries are i.i.d. N(0,1) from numpy's PCG64 with the given seed, filled
in key order; QR sweeps bring the chain into mixed canonical form around a
two-site centre.  The reference's own initialisers are time-seeded
(src/BlockTensor.hpp:394-404), so the seed scheme is this package's.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

from .blocks import Blocks, qadd, qsub, twosite


def _counts_left(L, itk):
    cl = [{0: 1}]
    for i in range(L):
        nxt = {}
        for q, c in cl[-1].items():
            for qs in itk[i]:
                r = qadd(q, qs)
                nxt[r] = nxt.get(r, 0) + c
        cl.append(nxt)
    return cl


def _counts_right(L, itk, Ntot):
    cr = [None] * (L + 1)
    cr[L] = {Ntot: 1}
    for i in range(L - 1, -1, -1):
        nxt = {}
        for r, c in cr[i + 1].items():
            for qs in itk[i]:
                q = qsub(r, qs)
                nxt[q] = nxt.get(q, 0) + c
        cr[i] = nxt
    return cr


def _apportion(D, weights):
    tot = sum(weights.values())
    quota = {k: Fraction(D * w, tot) for k, w in weights.items()}
    base = {k: int(q) for k, q in quota.items()}
    rest = D - sum(base.values())
    order = sorted(weights, key=lambda k: (-(quota[k] - base[k]), k))
    for k in order[:rest]:
        base[k] += 1
    return base


def bond_dims(L, itk, Ntot, D, clamp=None, nsec=None):
    """Per-bond {charge: dim} for bonds 0..L (see module docstring).  With
    clamp = (lo, hi) and nsec, a bond keeps its nsec heaviest sectors (weight
    cl * cr, ties by charge), apportions D over them by weight and clamps each
    dimension into [lo, hi] (and to the exact count), so sector sizes are
    bounded below as well as above."""
    cl, cr = _counts_left(L, itk), _counts_right(L, itk, Ntot)
    dims = []
    for l in range(L + 1):
        qs = sorted(set(cl[l]) & set(cr[l]))
        mmax = {q: min(cl[l][q], cr[l][q]) for q in qs}
        if clamp is not None:
            w = {q: cl[l][q] * cr[l][q] for q in qs}
            keep = sorted(qs, key=lambda q: (-w[q], q))[:nsec or len(qs)]
            a = _apportion(D, {q: w[q] for q in keep}) if keep else {}
            d = {q: min(max(a[q], clamp[0]), clamp[1], mmax[q]) for q in keep}
        elif sum(mmax.values()) <= D:
            d = dict(mmax)
        else:
            w = {q: cl[l][q] * cr[l][q] for q in qs}
            a = _apportion(D, w)
            d = {q: min(a[q], mmax[q]) for q in qs}
        dims.append({q: m for q, m in d.items() if m > 0})
    # consistency: m_l(r) <= sum_s m_{l-1}(r - q_s) and <= sum_s m_{l+1}(r + q_s)
    changed = True
    while changed:
        changed = False
        for l in range(1, L + 1):
            for r in list(dims[l]):
                cap = sum(dims[l - 1].get(qsub(r, qs), 0) for qs in itk[l - 1])
                if dims[l][r] > cap:
                    dims[l][r] = cap
                    changed = True
                if dims[l][r] == 0:
                    del dims[l][r]
        for l in range(L - 1, -1, -1):
            for q in list(dims[l]):
                cap = sum(dims[l + 1].get(qadd(q, qs), 0) for qs in itk[l])
                if dims[l][q] > cap:
                    dims[l][q] = cap
                    changed = True
                if dims[l][q] == 0:
                    del dims[l][q]
    return dims


def random_site(rng, dl, dr, itk_site, dtype=np.float64):
    A = Blocks(1, dtype)
    for lq in sorted(dl):
        for s, qs in enumerate(itk_site):
            r = qadd(lq, qs)
            if r in dr:
                shape = (dl[lq], dr[r])
                if np.issubdtype(np.dtype(dtype), np.complexfloating):
                    A[(lq, r, s)] = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
                else:
                    A[(lq, r, s)] = rng.standard_normal(shape)
    return A


def left_normalize(A: Blocks, Anext: Blocks | None):
    """QR per right charge: stacked [A(l,r,s)] over keys -> isometry; R moves into Anext."""
    out = Blocks(1, A.dtype)
    byr = {}
    for k in A.keys():
        byr.setdefault(k[1], []).append(k)
    carry = {}
    for r, ks in byr.items():
        M = np.vstack([A[k] for k in ks])
        Q, R = np.linalg.qr(M)
        row = 0
        for k in ks:
            m = A[k].shape[0]
            out[k] = Q[row:row + m, :]
            row += m
        carry[r] = R
    if Anext is not None:
        nxt = Blocks(1, Anext.dtype)
        for k, v in Anext.items():
            nxt[k] = carry[k[0]] @ v
        return out, nxt
    return out, carry


def right_normalize(A: Blocks, Aprev: Blocks | None):
    """LQ per left charge: [A(l,r,s)] side by side -> co-isometry; L moves into Aprev."""
    out = Blocks(1, A.dtype)
    byl = {}
    for k in A.keys():
        byl.setdefault(k[0], []).append(k)
    carry = {}
    for lq, ks in byl.items():
        M = np.hstack([A[k] for k in ks])
        Q, R = np.linalg.qr(M.conj().T)
        Lm, Qh = R.conj().T, Q.conj().T
        col = 0
        for k in ks:
            m = A[k].shape[1]
            out[k] = Qh[:, col:col + m]
            col += m
        carry[lq] = Lm
    if Aprev is not None:
        prv = Blocks(1, Aprev.dtype)
        for k, v in Aprev.items():
            prv[k] = v @ carry[k[1]]
        return out, prv
    return out, carry


def random_mps(L, itk, Ntot, D, seed, dtype=np.float64, clamp=None, nsec=None):
    """Seeded random MPS (list of Blocks) with the sector profile above."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dims = bond_dims(L, itk, Ntot, D, clamp, nsec)
    return [random_site(rng, dims[i], dims[i + 1], itk[i], dtype) for i in range(L)], dims


def mixed_canonical(mps, c):
    """Left-normalise sites < c, right-normalise sites > c+1 (in place on a copy)."""
    mps = [A.copy() for A in mps]
    for i in range(c):
        mps[i], mps[i + 1] = left_normalize(mps[i], mps[i + 1])
    for i in range(len(mps) - 1, c + 1, -1):
        mps[i], mps[i - 1] = right_normalize(mps[i], mps[i - 1])
    return mps


def centre_problem(L, itk, Ntot, D, seed, c=None, dtype=np.float64):
    """Random mixed-canonical MPS and the normalised two-site psi at sites (c, c+1)."""
    if c is None:
        c = L // 2 - 1
    mps, dims = random_mps(L, itk, Ntot, D, seed, dtype)
    mps = mixed_canonical(mps, c)
    psi = twosite(mps[c], mps[c + 1])
    nrm = psi.norm()
    psi = psi.scaled(1.0 / nrm)
    return mps, psi, c, dims
