"""Complex128 cases made from the real problems, and the complex contractions
written as sums of real ones.

Every block of a real problem gets a seeded imaginary part of its shape.  The
MPO values are real, so the contractions are linear in each block argument
(the environment update conjugates its upper tensor: A_up^H E A,
src/MPSfunctions.hpp:335-341).  Splitting every argument into real and
imaginary parts therefore gives the complex result as a sum of real
contractions with coefficients i^a i^b (+-i)^c, which the real oracle
(pinned by the golden vectors) computes without any complex code.
"""
import numpy as np

from mpstools_amd.blocks import Blocks


def with_imag(B: Blocks, seed: int, scale: float = 0.7) -> Blocks:
    rng = np.random.default_rng(seed)
    Z = Blocks(B.ns, np.complex128)
    for k in B.keys():
        v = np.asarray(B[k])
        Z[k] = v + 1j * scale * rng.standard_normal(v.shape)
    return Z


def parts(Z: Blocks):
    re, im = Blocks(Z.ns), Blocks(Z.ns)
    for k in Z.keys():
        v = np.asarray(Z[k])
        re[k] = v.real
        im[k] = v.imag
    return re, im


def _add(acc: dict, B: Blocks, coef: complex):
    for k in B.keys():
        v = coef * np.asarray(B[k], dtype=np.complex128)
        acc[k] = acc[k] + v if k in acc else v


def heff2_by_parts(oracle, L, R, psi, Wl, Wr, itl, itr) -> dict:
    Lp, Rp, Xp = parts(L), parts(R), parts(psi)
    acc = {}
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                _add(acc, oracle.contract_twosite(Lp[a], Rp[b], Xp[c], Wl, Wr, itl, itr), 1j ** (a + b + c))
    return acc


def heff1_by_parts(oracle, L, A, W, R, itk) -> dict:
    Lp, Ap, Rp = parts(L), parts(A), parts(R)
    acc = {}
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                _add(acc, oracle.contract_single(Lp[a], Ap[b], W, Rp[c], itk), 1j ** (a + b + c))
    return acc


def env_by_parts(oracle, E, A, W, itk, direction, Aup) -> dict:
    """A_up enters conjugated: its imaginary part carries -i."""
    Ep, Ap, Up = parts(E), parts(A), parts(Aup)
    acc = {}
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                _add(acc, oracle.add_layer(Ep[a], Ap[b], W, itk, direction, Ap=Up[c]),
                     (1j ** (a + b)) * ((-1j) ** c))
    return acc


def rel(got, want: dict) -> float:
    """Relative Frobenius distance over want's keys (got may be Blocks or a
    dict); the key sets must agree."""
    gk = set(got.keys())
    assert gk == set(want.keys()), (len(gk), len(want))
    num = sum(np.sum(np.abs(np.asarray(got[k]) - want[k]) ** 2) for k in want)
    den = sum(np.sum(np.abs(want[k]) ** 2) for k in want)
    return float(np.sqrt(num / max(den, 1e-300)))


def bond_unitaries(A_left: Blocks, seed: int) -> dict:
    """A random unitary per charge sector of the right bond of A_left (QR of a
    complex Gaussian matrix)."""
    dims = {}
    for k in A_left.keys():
        dims[k[1]] = np.asarray(A_left[k]).shape[1]
    rng = np.random.default_rng(seed)
    G = {}
    for q in sorted(dims):
        m = dims[q]
        Q, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
        G[q] = Q
    return G


def gauge_right(A: Blocks, G: dict) -> Blocks:
    """A -> A G on the right bond."""
    Z = Blocks(A.ns, np.complex128)
    for k in A.keys():
        Z[k] = np.asarray(A[k]) @ G[k[1]]
    return Z


def gauge_left(B: Blocks, G: dict) -> Blocks:
    """B -> G^H B on the left bond (MPS tensors and two-site states alike)."""
    Z = Blocks(B.ns, np.complex128)
    for k in B.keys():
        Z[k] = G[k[0]].conj().T @ np.asarray(B[k])
    return Z
