"""Dense brute-force references (numpy) for the block-sparse hot path.

Everything here is independent of the block-sparse code: the MPO becomes a
dense Hamiltonian H[out_config, in_config] = prod_i W_i[w_i, w_{i+1}, out_i, in_i]
summed over aux paths, and the MPS becomes dense bases of the bond spaces,
with charge sectors concatenated in ascending charge order.
"""
from __future__ import annotations

import numpy as np


def dense_H(mpo):
    """Full operator H[out, in] over d**L configurations (site 0 most significant)."""
    W0 = mpo[0].dense()  # (1, w, d, d)
    T = W0[0]            # (w, o, i)
    T = np.moveaxis(T, 0, -1)  # (o, i, w)
    dim = W0.shape[2]
    T = T.reshape(dim, dim, -1)
    for W in mpo[1:]:
        Wd = W.dense()   # (wl, wr, o, i)
        T = np.einsum("abw,wxcd->acbdx", T, Wd)
        n_o, d, n_i = T.shape[0], T.shape[1], T.shape[2]
        T = T.reshape(T.shape[0] * T.shape[1], T.shape[2] * T.shape[3], T.shape[4])
    assert T.shape[2] == 1
    return T[:, :, 0]


def sector_offsets(dims_at_bond):
    """{charge: (offset, dim)} with sectors in ascending charge order."""
    off = 0
    out = {}
    for q in sorted(dims_at_bond):
        out[q] = (off, dims_at_bond[q])
        off += dims_at_bond[q]
    return out, off


def site_dense(A, dl, dr, d):
    """Block site tensor -> dense T[left, s, right]."""
    ol, nl = sector_offsets(dl)
    orr, nr = sector_offsets(dr)
    T = np.zeros((nl, d, nr), dtype=A.dtype)
    for (lq, rq, s), blk in A.items():
        a, _ = ol[lq]
        b, _ = orr[rq]
        T[a:a + blk.shape[0], s, b:b + blk.shape[1]] = blk
    return T


def left_basis(mps, dims, c, d):
    """P_L[(s_0..s_{c-1}), bond_c] from sites 0..c-1."""
    P = np.ones((1, 1))
    for i in range(c):
        T = site_dense(mps[i], dims[i], dims[i + 1], d)
        P = np.einsum("xa,asb->xsb", P, T).reshape(P.shape[0] * d, T.shape[2])
    return P


def right_basis(mps, dims, c2, d):
    """P_R[bond_{c2}, (s_{c2}..s_{L-1})] from sites c2..L-1."""
    L = len(mps)
    P = np.ones((1, 1))
    for i in range(L - 1, c2 - 1, -1):
        T = site_dense(mps[i], dims[i], dims[i + 1], d)
        P = np.einsum("asb,bx->asx", T, P).reshape(T.shape[0], d * P.shape[1])
    return P


def twosite_dense(psi, dl, dr, d):
    """Two-site Blocks (l, r, s1, s2) -> dense psi[l, s1, s2, r]."""
    ol, nl = sector_offsets(dl)
    orr, nr = sector_offsets(dr)
    X = np.zeros((nl, d, d, nr), dtype=psi.dtype)
    for (lq, rq, s1, s2), blk in psi.items():
        a, _ = ol[lq]
        b, _ = orr[rq]
        X[a:a + blk.shape[0], s1, s2, b:b + blk.shape[1]] = blk
    return X


def heff_dense(mpo, mps, dims, c, d):
    """P^T H^T P for the two-site problem at sites (c, c+1): the reference applies
    sum_(out,in) v(out,in) psi[out] -> y[in], i.e. the transpose of H[out,in]."""
    PL = left_basis(mps, dims, c, d)
    PR = right_basis(mps, dims, c + 2, d)
    nl, nr = PL.shape[1], PR.shape[0]
    Id = np.eye(d * d)
    P = np.kron(np.kron(PL, Id), PR.T)  # (d^L, nl*d*d*nr)
    H = dense_H(mpo)
    return P.T @ H.T @ P, (nl, nr)
