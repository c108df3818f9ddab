"""MPO tensors of the models the hot path is exercised with.

`MPOMat` mirrors ``MPOMat<1,T>`` (src/MPOTypes.hpp:839-899): a map from the
aux pair ``ukey<2>(wl, wr)`` to a sparse local operator, itself a map from
``OpKeyType<1>(out, in)`` to a value.  Iteration order is the reference's map
order: ukey lexicographic (src/keyTypes.hpp:471-483), operator entries by
``(in, out)`` (src/keyTypes.hpp:313-343 with OpKeyType storing in in [0]).

The constructors follow src/Hamiltonians.hpp line by line (same aux layout,
same operator products, same ``squeeze(1e-16)``), so the W tensors fed to the
contraction are the reference's.
"""
from __future__ import annotations

import numpy as np


class LocalOp:
    """SparseLocalOperator<T> (src/MPOTypes.hpp:418-530); entries {(out,in): v}."""

    def __init__(self, d, entries=None):
        self.d = d
        self.e: dict[tuple[int, int], complex] = dict(entries or {})

    def items(self):
        # map order of OpKeyType: (in, out)
        return sorted(self.e.items(), key=lambda kv: (kv[0][1], kv[0][0]))

    def __mul__(self, other):
        if isinstance(other, LocalOp):
            # SparseOperator::operator*= (src/MPOTypes.hpp:218-243): matrix
            # product, key (i.out, j.in) whenever i.in == j.out.
            out = LocalOp(self.d)
            for (o1, i1), v1 in self.items():
                for (o2, i2), v2 in other.items():
                    if i1 == o2:
                        k = (o1, i2)
                        out.e[k] = out.e.get(k, 0.0) + v1 * v2
            return out
        return LocalOp(self.d, {k: v * other for k, v in self.e.items()})

    __rmul__ = __mul__

    def __add__(self, other):
        # SparseOperator addition: entries of both, summed on common keys
        out = LocalOp(self.d, self.e)
        for k, v in other.e.items():
            out.e[k] = out.e.get(k, 0.0) + v
        return out

    def dense(self):
        m = np.zeros((self.d, self.d), dtype=np.result_type(*[type(v) for v in self.e.values()], float)
                     if self.e else float)
        for (o, i), v in self.e.items():
            m[o, i] += v
        return m


class MPOMat:
    """MPOMat<1,T>: {(wl,wr): LocalOp}, with Dl/Dr."""

    def __init__(self, Dl=0, Dr=0, d=2):
        self.Dl, self.Dr, self.d = Dl, Dr, d
        self.m: dict[tuple[int, int], LocalOp] = {}

    def __setitem__(self, k, op: LocalOp):
        self.m[tuple(k)] = LocalOp(op.d, op.e)

    def squeeze(self, thresh):
        # MPOMat::squeeze (src/MPOTypes.hpp:891-899): drops small entries, keeps keys.
        for op in self.m.values():
            op.e = {k: v for k, v in op.e.items() if abs(v) >= thresh}

    def entries(self):
        """(wl, wr, out, in, value) in the reference's iteration order."""
        out = []
        for (wl, wr) in sorted(self.m):
            for (o, i), v in self.m[(wl, wr)].items():
                out.append((wl, wr, o, i, v))
        return out

    def arrays(self, complex_=False):
        e = self.entries()
        wl = np.array([x[0] for x in e], dtype=np.uint32)
        wr = np.array([x[1] for x in e], dtype=np.uint32)
        o = np.array([x[2] for x in e], dtype=np.uint32)
        i = np.array([x[3] for x in e], dtype=np.uint32)
        if complex_:
            v = np.array([complex(x[4]) for x in e], dtype=np.complex128)
        else:
            v = np.array([float(np.real(x[4])) for x in e], dtype=np.float64)
        return wl, wr, o, i, v

    def dense(self):
        """W[wl, wr, out, in] as a dense array."""
        W = np.zeros((self.Dl, self.Dr, self.d, self.d),
                     dtype=np.complex128 if any(isinstance(x[4], complex) for x in self.entries()) else np.float64)
        for wl, wr, o, i, v in self.entries():
            W[wl, wr, o, i] += v
        return W


def _spin_ops():
    Sz = LocalOp(2, {(0, 0): -0.5, (1, 1): 0.5})
    Sp = LocalOp(2, {(1, 0): 1.0})
    Sm = LocalOp(2, {(0, 1): 1.0})
    Id = LocalOp(2, {(0, 0): 1.0, (1, 1): 1.0})
    return Sz, Sp, Sm, Id


def heisenberg_mpo(Jz, Jxy, B):
    """HeisenbergMPO (src/Hamiltonians.hpp:86-133).  Jz, Jxy have N-1 entries, B has N."""
    N = len(B)
    Sz, Sp, Sm, Id = _spin_ops()
    mpo = []
    M = MPOMat(1, 5)
    M[(0, 0)] = -B[0] * Sz
    M[(0, 1)] = Jxy[0] / 2. * Sm
    M[(0, 2)] = Jxy[0] / 2. * Sp
    M[(0, 3)] = Jz[0] * Sz
    M[(0, 4)] = Id
    M.squeeze(1e-16)
    mpo.append(M)
    for site in range(1, N - 1):
        Mb = MPOMat(5, 5)
        Mb[(0, 0)] = Id
        Mb[(1, 0)] = Sp
        Mb[(2, 0)] = Sm
        Mb[(3, 0)] = Sz
        Mb[(4, 0)] = -B[site] * Sz
        Mb[(4, 1)] = Jxy[site] / 2. * Sm
        Mb[(4, 2)] = Jxy[site] / 2. * Sp
        Mb[(4, 3)] = Jz[site] * Sz
        Mb[(4, 4)] = Id
        Mb.squeeze(1e-16)
        mpo.append(Mb)
    M = MPOMat(5, 1)
    M[(0, 0)] = Id
    M[(1, 0)] = Sp
    M[(2, 0)] = Sm
    M[(3, 0)] = Sz
    M[(4, 0)] = -B[N - 1] * Sz
    M.squeeze(1e-16)
    mpo.append(M)
    return mpo


def xxz_mpo(N, Jz=1.0, Jxy=1.0, hz=0.0):
    return heisenberg_mpo([Jz] * (N - 1), [Jxy] * (N - 1), [hz] * N)


def mpo_difference(A, B, thresh=0.0):
    """Entrywise A - B of two MPO chains with the same aux layout: the
    derivative tensors dW/dtheta of an MPO affine in theta, as
    mpo(theta = e_p) - mpo(theta = 0).  Entries that cancel are dropped."""
    out = []
    for Wa, Wb in zip(A, B):
        M = MPOMat(Wa.Dl, Wa.Dr, Wa.d)
        keys = set(Wa.m) | set(Wb.m)
        for k in sorted(keys):
            ea = Wa.m[k].e if k in Wa.m else {}
            eb = Wb.m[k].e if k in Wb.m else {}
            e = {}
            for oi in set(ea) | set(eb):
                v = ea.get(oi, 0.0) - eb.get(oi, 0.0)
                if abs(v) > thresh:
                    e[oi] = v
            if e:
                M[k] = LocalOp(Wa.d, e)
        out.append(M)
    return out


def xxz_param_derivatives(N):
    """dW[i]/dtheta for theta = (Jz, Jxy, hz) of xxz_mpo (uniform couplings)."""
    base = xxz_mpo(N, 0.0, 0.0, 0.0)
    return {"Jz": mpo_difference(xxz_mpo(N, 1.0, 0.0, 0.0), base),
            "Jxy": mpo_difference(xxz_mpo(N, 0.0, 1.0, 0.0), base),
            "hz": mpo_difference(xxz_mpo(N, 0.0, 0.0, 1.0), base)}


def spinless_fermions_mpo(V, t, mu):
    """SpinlessFermionsMPO (src/Hamiltonians.hpp:521-569), Jordan-Wigner sign p."""
    N = len(mu)
    n = LocalOp(2, {(0, 0): 0.0, (1, 1): 1.0})
    cd = LocalOp(2, {(1, 0): 1.0})
    c = LocalOp(2, {(0, 1): 1.0})
    Id = LocalOp(2, {(0, 0): 1.0, (1, 1): 1.0})
    p = LocalOp(2, {(0, 0): 1.0, (1, 1): -1.0})
    mpo = []
    M = MPOMat(1, 5)
    M[(0, 0)] = -mu[0] * n
    M[(0, 1)] = t[0] * (c * p)
    M[(0, 2)] = -t[0] * (cd * p)
    M[(0, 3)] = V[0] * n
    M[(0, 4)] = Id
    M.squeeze(1e-16)
    mpo.append(M)
    for site in range(1, N - 1):
        Mb = MPOMat(5, 5)
        Mb[(0, 0)] = Id
        Mb[(1, 0)] = cd
        Mb[(2, 0)] = c
        Mb[(3, 0)] = n
        Mb[(4, 0)] = -mu[site] * n
        Mb[(4, 1)] = t[site] * (c * p)
        Mb[(4, 2)] = -t[site] * (cd * p)
        Mb[(4, 3)] = V[site] * n
        Mb[(4, 4)] = Id
        Mb.squeeze(1e-16)
        mpo.append(Mb)
    M = MPOMat(5, 1)
    M[(0, 0)] = Id
    M[(1, 0)] = cd
    M[(2, 0)] = c
    M[(3, 0)] = n
    M[(4, 0)] = -mu[N - 1] * n
    M.squeeze(1e-16)
    mpo.append(M)
    return mpo


def tv_mpo(N, t=1.0, V=2.0, mu=0.0):
    return spinless_fermions_mpo([V] * (N - 1), [t] * (N - 1), [mu] * N)


def itokey_u1(d=2):
    """i_to_key for spin-1/2 / spinless fermions: state s -> charge s (N_up / N)."""
    return [s for s in range(d)]


def itokey_dense(d=2):
    """No symmetry: every local state carries charge 0 (dense blocks)."""
    return [0] * d


def pack_charge(q1, q2):
    """AbKeyType(q1, q2, 2): the two U(1) charges packed as (q2 << 16) + q1
    (src/keyTypes.hpp:25-30); sums and differences act on the packed integer."""
    return (int(q2) << 16) + int(q1)


def itokey_u1xu1():
    """i_to_key of the Hubbard site (main/TO_solver_new.cpp:407-415): states
    0 (empty), 1 (down), 2 (up), 3 (up and down) carry (N_up, N_down) =
    (0,0), (0,1), (1,0), (1,1)."""
    return [pack_charge(0, 0), pack_charge(0, 1), pack_charge(1, 0), pack_charge(1, 1)]


def _hubbard_ops():
    # src/Hamiltonians.hpp:829-843: the down fermion is created after the up
    # one inside a site (|3> = c+_up c+_dn |0>), p is the site's fermion parity
    nu = LocalOp(4, {(2, 2): 1.0, (3, 3): 1.0})
    nd = LocalOp(4, {(1, 1): 1.0, (3, 3): 1.0})
    cuD = LocalOp(4, {(2, 0): 1.0, (3, 1): 1.0})
    cu = LocalOp(4, {(0, 2): 1.0, (1, 3): 1.0})
    cdD = LocalOp(4, {(1, 0): 1.0, (3, 2): -1.0})
    cd = LocalOp(4, {(0, 1): 1.0, (2, 3): -1.0})
    Id = LocalOp(4, {(k, k): 1.0 for k in range(4)})
    p = LocalOp(4, {(0, 0): 1.0, (1, 1): -1.0, (2, 2): -1.0, (3, 3): 1.0})
    return nu, nd, cuD, cu, cdD, cd, Id, p


def fermi_hubbard_mpo(hop, U, ch_up, ch_down, V=None):
    """FermiHubbardMPO (src/Hamiltonians.hpp:823-889; with the doublon-doublon
    term V, :891-962), w = 6 (7 with V), d = 4:
      H = sum_i -hop_i sum_s (c+_{i s} c_{i+1 s} + h.c.)
          + sum_i (ch_up_i n_{i up} + ch_down_i n_{i dn} + U_i n_{i up} n_{i dn})
          [+ sum_i V_i (n_up n_dn)_i (n_up n_dn)_{i+1}]
    with the reference's aux layout, operator products and squeeze(1e-15)."""
    N = len(ch_up)
    assert len(ch_down) == N
    nu, nd, cuD, cu, cdD, cd, Id, p = _hubbard_ops()
    n = nu * nd

    def onsite(s):
        return ch_up[s] * nu + ch_down[s] * nd + U[s] * (nu * nd)

    w = 6 if V is None else 7
    last = w - 1
    mpo = []
    M = MPOMat(1, w, 4)
    M[(0, 0)] = onsite(0)
    M[(0, 1)] = -hop[0] * (cuD * p)
    M[(0, 2)] = -hop[0] * (cdD * p)
    M[(0, 3)] = hop[0] * (cu * p)
    M[(0, 4)] = hop[0] * (cd * p)
    if V is not None:
        M[(0, 5)] = V[0] * n
    M[(0, last)] = Id
    M.squeeze(1e-15)
    mpo.append(M)
    for site in range(1, N - 1):
        Mb = MPOMat(w, w, 4)
        Mb[(0, 0)] = Id
        Mb[(1, 0)] = cu
        Mb[(2, 0)] = cd
        Mb[(3, 0)] = cuD
        Mb[(4, 0)] = cdD
        if V is not None:
            Mb[(5, 0)] = n
        Mb[(last, 0)] = onsite(site)
        Mb[(last, 1)] = -hop[site] * (cuD * p)
        Mb[(last, 2)] = -hop[site] * (cdD * p)
        Mb[(last, 3)] = hop[site] * (cu * p)
        Mb[(last, 4)] = hop[site] * (cd * p)
        if V is not None:
            Mb[(last, 5)] = V[site] * n
        Mb[(last, last)] = Id
        Mb.squeeze(1e-15)
        mpo.append(Mb)
    M = MPOMat(w, 1, 4)
    M[(0, 0)] = Id
    M[(1, 0)] = cu
    M[(2, 0)] = cd
    M[(3, 0)] = cuD
    M[(4, 0)] = cdD
    if V is not None:
        M[(5, 0)] = n
    M[(last, 0)] = onsite(N - 1)
    M.squeeze(1e-15)
    mpo.append(M)
    return mpo


def hubbard_mpo(N, t=1.0, U=4.0, ch_up=0.0, ch_down=0.0, V=None):
    """Uniform Fermi-Hubbard chain of N sites (optional uniform V)."""
    return fermi_hubbard_mpo([t] * (N - 1), [U] * N, [ch_up] * N, [ch_down] * N,
                             None if V is None else [V] * (N - 1))


def rescale_mpo(mpo, scale, shift):
    """The MPO of (H - shift) / scale, for the Chebyshev recurrence (the
    reference rescales its Hamiltonian's couplings before simulateSSCompress,
    src/container.hpp:2343).  For MPOs in the layout of the constructors above
    (aux channel Dl-1 of a site is the 'term not started' channel, 0 the 'term
    done' one): every entry leaving the start channel other than its
    pass-through identity begins exactly one term, so those entries are
    divided by `scale`; -shift / (scale N) identity is added to every site's
    on-site entry (start -> done)."""
    N = len(mpo)
    out = []
    for site, W in enumerate(mpo):
        M = MPOMat(W.Dl, W.Dr, W.d)
        start = W.Dl - 1 if site > 0 else 0
        for k, op in W.m.items():
            wl, wr = k
            if wl == start and not (site < N - 1 and wr == W.Dr - 1):
                M[k] = op * (1.0 / scale)
            else:
                M[k] = op
        onsite = (start, 0)
        base = M.m.get(onsite, LocalOp(W.d))
        M[onsite] = base + LocalOp(W.d, {(i, i): -shift / (scale * N) for i in range(W.d)})
        M.squeeze(1e-16)
        out.append(M)
    return out
