"""Complex128 contractions in the oracle (the checker of the device's complex
plans), pinned on the CPU.

The oracle's complex path (the same templates as its real one, with the
reference's conj() copies, src/MPSfunctions.hpp:335-341) must equal the sum of
real-oracle contractions over the real and imaginary parts of every argument
(tests/complexcase.py): with a real MPO each contraction is linear in its
block arguments and the environment update conjugates its upper tensor.  A
dropped conjugation, a conjugation on the wrong operand or a lost imaginary
cross term fails these at O(1).  Bar: 1e-13 relative (FP64 sums of a few
products).
"""
import numpy as np
import pytest

import complexcase as cc
import problems

CASES = [("xxz", 8, 4, 8, 21, False), ("asym", 8, 4, 8, 22, False), ("tv", 10, 5, 10, 23, False),
         ("hubbard", 6, 6, 12, 24, False)]


@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES)
def test_complex_heff2_is_the_sum_of_real_parts(oracle, model, L, N, D, seed, dense):
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lz, Rz = cc.with_imag(P["Ls"][c - 1], seed + 1), cc.with_imag(P["Rs"][c + 2], seed + 2)
    xz = cc.with_imag(P["psi"], seed + 3)
    got = oracle.contract_twosite(Lz, Rz, xz, mpo[c], mpo[c + 1], itk[c], itk[c + 1], cplx=True)
    want = cc.heff2_by_parts(oracle, Lz, Rz, xz, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    assert cc.rel(got, want) <= 1e-13
    # the imaginary parts matter: the real parts alone give a different result
    yr = oracle.contract_twosite(*[cc.parts(b)[0] for b in (Lz, Rz, xz)], mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    assert cc.rel({k: np.asarray(yr[k], np.complex128) for k in yr.keys()}, want) > 1e-2


@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES[:2])
def test_complex_heff1_is_the_sum_of_real_parts(oracle, model, L, N, D, seed, dense):
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    Lz, Rz = cc.with_imag(P["Ls"][c - 1], seed + 4), cc.with_imag(P["Rs"][c + 1], seed + 5)
    Az = cc.with_imag(P["mps"][c], seed + 6)
    got = oracle.contract_single(Lz, Az, mpo[c], Rz, itk[c], cplx=True)
    assert cc.rel(got, cc.heff1_by_parts(oracle, Lz, Az, mpo[c], Rz, itk[c])) <= 1e-13


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("model,L,N,D,seed,dense", CASES[:1] + CASES[3:])
def test_complex_env_conjugates_the_upper_tensor(oracle, model, L, N, D, seed, dense, mixed):
    P = problems.build(oracle, model, L, N, D, seed, dense=dense)
    c, mpo, itk = P["c"], P["mpo"], P["itk"]
    for direction, E, A in ((1, P["Ls"][c - 1], P["mps"][c]), (0, P["Rs"][c + 2], P["mps"][c + 1])):
        site = c if direction > 0 else c + 1
        Ez, Az = cc.with_imag(E, seed + 7), cc.with_imag(A, seed + 8)
        Uz = cc.with_imag(A, seed + 9) if mixed else Az
        got = oracle.add_layer(Ez, Az, mpo[site], itk[site], direction, Ap=Uz if mixed else None, cplx=True)
        want = cc.env_by_parts(oracle, Ez, Az, mpo[site], itk[site], direction, Uz)
        assert cc.rel(got, want) <= 1e-13, direction
        # without the conjugation (+i on the upper tensor) the sum differs at O(1)
        Up = cc.parts(Uz)
        unconj = {}
        for a in (0, 1):
            for b in (0, 1):
                for k in (0, 1):
                    cc._add(unconj, oracle.add_layer(cc.parts(Ez)[a], cc.parts(Az)[b], mpo[site], itk[site],
                                                     direction, Ap=Up[k]), 1j ** (a + b + k))
        assert cc.rel(unconj, want) > 1e-2


@pytest.mark.parametrize("model,L,N,D,seed", [("tv", 12, 6, 16, 31), ("hubbard", 8, 8, 16, 32)])
def test_complex_oracle_is_gauge_covariant(oracle, model, L, N, D, seed):
    """A random unitary G per sector on the bond left of the centre
    (A -> A G, psi -> G^H psi): the complex environment is G^H L G and
    H_eff psi becomes G^H (H_eff psi), with the real oracle's L and y."""
    P = problems.build(oracle, model, L, N, D, seed)
    c, mps, mpo, itk = P["c"], P["mps"], P["mpo"], P["itk"]
    G = cc.bond_unitaries(mps[c - 1], seed)
    Lg = oracle.add_layer(cc.with_imag(P["Ls"][c - 2], 0, 0.0), cc.gauge_right(mps[c - 1], G), mpo[c - 1],
                          itk[c - 1], 1, cplx=True)
    L0 = P["Ls"][c - 1]
    assert cc.rel(Lg, {k: G[k[0]].conj().T @ np.asarray(L0[k]) @ G[k[1]] for k in L0.keys()}) <= 1e-13
    R0 = P["Rs"][c + 2]
    y = oracle.contract_twosite(L0, R0, P["psi"], mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    yg = oracle.contract_twosite(Lg, cc.with_imag(R0, 0, 0.0), cc.gauge_left(P["psi"], G), mpo[c], mpo[c + 1],
                                 itk[c], itk[c + 1], cplx=True)
    w = cc.gauge_left(y, G)
    assert cc.rel(yg, {k: np.asarray(w[k]) for k in w.keys()}) <= 1e-13
