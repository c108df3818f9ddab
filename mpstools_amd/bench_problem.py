"""Device-resident synthetic workload: the centre bond of a dense (no symmetry)
spin-1/2 Heisenberg chain, J = 1, L sites, MPS bond dimension D.

Inputs are a seeded Gaussian MPS (torch's CUDA generator; random input
generation is plumbing), left-normalised (CholeskyQR2) on sites < c and
right-normalised on sites > c+1.  The environments L[c-1] and R[c+2] are then
built layer by layer with this package's environment kernels
(addContractionLayer on the GPU), exactly as DMRGengine builds them
(src/MPSfunctions.hpp:2025-2052).  Every local state carries charge 0, so all
blocks are dense: L and R hold one D x D block per MPO channel and psi four
D x D blocks (s1, s2).
"""
from __future__ import annotations

import sys

import numpy as np

from . import models, native
from .blocks import Blocks


def _dims(L, D, d=2):
    return [int(min(d ** min(l, L - l), D)) for l in range(L + 1)]


def _site_table(dl, dr, d=2):
    q = np.zeros((d, 2), np.uint64)
    p = np.arange(d, dtype=np.uint32).reshape(d, 1)
    return native.Table.from_keys(1, q, p, [dl] * d, [dr] * d)


def _lmin_table():
    return native.Table.from_keys(1, np.zeros((1, 2), np.uint64), np.zeros((1, 1), np.uint32), [1], [1])


def _isometry(torch, G):
    """Q with orthonormal columns spanning G (m x n, m >= n): CholeskyQR2
    (Q = G R^-1 with R^T R = G^T G, applied twice).  A few large GEMM/Cholesky
    calls instead of a Householder sweep's many small kernels."""
    Q = G
    for _ in range(2):
        R = torch.linalg.cholesky(Q.T @ Q, upper=True)
        Q = torch.linalg.solve_triangular(R, Q, upper=True, left=False)
    return Q


def _add_plan(torch, out, pr, W, itk, seed):
    """out["plan_" + pr]: the H_eff plan of the bond at precision pr.  "c128"
    is the complex128 plan (FP64, planar re/im; "c128_<prec>" the same at a
    reduced GEMM precision) over the same keys, bound to
    complex environments made of the real ones plus seeded imaginary parts of
    the same magnitude, with a complex start vector in out["psi_c128"]."""
    Lt, Rt, vt, dL, dR = out["Lt"], out["Rt"], out["vt"], out["dL"], out["dR"]
    if not pr.startswith("c128"):
        plan = native.HeffPlan(Lt, Rt, vt, vt, W, itk, native.PREC[pr])
        plan.bind_env(dL, dR)
        out["plan_" + pr] = plan
        return
    g = torch.Generator(device=dL.device)
    g.manual_seed(seed + 128)

    def cplx(v):
        im = torch.randn(v.numel(), dtype=torch.float64, device=v.device, generator=g) * v.abs().mean()
        return torch.cat([v, im])

    if "dL_c128" not in out:
        out["dL_c128"], out["dR_c128"] = cplx(dL), cplx(dR)
        x = cplx(out["psi"])
        out["psi_c128"] = x / x.norm()
    out["psi_" + pr] = out["psi_c128"]
    vz = vt.as_complex()
    sub = pr[len("c128"):].lstrip("_") or "f64"  # c128 (FP64), c128_f32, c128_bf16x3, ...
    plan = native.HeffPlan(Lt.as_complex(), Rt.as_complex(), vz, vz, W, itk, native.PREC[sub])
    plan.bind_env(out["dL_c128"], out["dR_c128"])
    out["plan_" + pr] = plan


def dense_heisenberg_centre(D=2048, L=128, seed=20260003, build_envs="sweep", precisions=("f64",), verbose=False,
                            keep_right=0):
    """keep_right = n also keeps, for a measured run of n bond steps to the
    right of the centre, the right-normalised site tensors A[c+2 .. c+1+n]
    (out["sites"][s] = (Table, device buffer)) and the right environments
    R[c+3 .. c+2+n] (out["R"][s] = (Table, device buffer))."""
    import torch
    dev = "cuda"
    mpo = models.xxz_mpo(L, 1.0, 1.0, 0.0)
    itk = models.itokey_dense()
    c = L // 2 - 1
    dims = _dims(L, D)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    plans = {}
    kept_sites, kept_R = {}, {}

    def env_step(Et, dE, s, direction):
        dl, dr = dims[s], dims[s + 1]
        At = _site_table(dl, dr)
        key = (s if s in (0, L - 1) else -1, dl, dr, direction, tuple(Et.keys()))
        if key not in plans:
            outt = native.env_outkeys(Et, At, mpo[s], itk, direction)
            plans[key] = (native.EnvPlan(Et, At, mpo[s], itk, direction, outt), outt)
        plan, outt = plans[key]
        # random site tensor -> isometry
        if direction > 0:
            G = torch.randn(2 * dl, dr, dtype=torch.float64, device=dev, generator=g)
            Q = _isometry(torch, G)
            blocks = [Q[i * dl:(i + 1) * dl, :] for i in range(2)]
        else:
            G = torch.randn(dl, 2 * dr, dtype=torch.float64, device=dev, generator=g)
            Qh = _isometry(torch, G.T.contiguous()).T
            blocks = [Qh[:, i * dr:(i + 1) * dr] for i in range(2)]
        dA = torch.zeros(max(At.total, 1), dtype=torch.float64, device=dev)
        for i in range(2):
            o = int(At.offsets[i])
            dA[o:o + dl * dr] = blocks[i].T.contiguous().reshape(-1)  # column-major
        dout = torch.zeros(max(outt.total, 1), dtype=torch.float64, device=dev)
        plan.apply(dE, dA, dout)
        if direction <= 0 and c + 2 <= s <= c + 1 + keep_right:
            kept_sites[s] = (At, dA)
        return outt, dout

    if build_envs == "sweep":
        Et, dE = _lmin_table(), torch.ones(1, dtype=torch.float64, device=dev)
        for s in range(0, c):
            Et, dE = env_step(Et, dE, s, 1)
            if verbose and s % 16 == 0:
                print(f"[bench_problem] left env site {s}", file=sys.stderr, flush=True)
        Lt, dL = Et, dE
        # R[L] = identity on the last bond (getRMax), aux 0
        Et, dE = _lmin_table(), torch.ones(1, dtype=torch.float64, device=dev)
        for s in range(L - 1, c + 1, -1):
            if c + 3 <= s + 1 <= c + 2 + keep_right:
                kept_R[s + 1] = (Et, dE)  # R[s + 1], before site s is added
            Et, dE = env_step(Et, dE, s, 0)
            if verbose and s % 16 == 0:
                print(f"[bench_problem] right env site {s}", file=sys.stderr, flush=True)
        Rt, dR = Et, dE
    else:
        raise ValueError(build_envs)
    plans.clear()

    # centre two-site psi = A_c A_{c+1}, normalised
    dl, dm, dr = dims[c], dims[c + 1], dims[c + 2]
    A1 = torch.randn(2, dl, dm, dtype=torch.float64, device=dev, generator=g)
    A2 = torch.randn(2, dm, dr, dtype=torch.float64, device=dev, generator=g)
    q = np.zeros((4, 2), np.uint64)
    p = np.array([[s1, s2] for s1 in range(2) for s2 in range(2)], np.uint32)
    xt = native.Table.from_keys(2, q, p, [dl] * 4, [dr] * 4)
    yt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk, itk)
    vt = xt.union(yt)
    psi = torch.zeros(vt.total, dtype=torch.float64, device=dev)
    for i, k in enumerate(vt.keys()):
        s1, s2 = k[2], k[3]
        blk = A1[s1] @ A2[s2]
        o = int(vt.offsets[i])
        psi[o:o + dl * dr] = blk.T.contiguous().reshape(-1)
    psi /= psi.norm()
    out = dict(mpo=mpo, itk=itk, c=c, dims=dims, Lt=Lt, Rt=Rt, dL=dL, dR=dR, xt=xt, yt=yt, vt=vt, psi=psi, L=L, D=D,
               sites=kept_sites, R=kept_R)
    for pr in precisions:
        _add_plan(torch, out, pr, (mpo[c], mpo[c + 1]), (itk, itk), seed)
    return out


def u1_centre(model="tv", L=100, N=50, D=512, seed=20260002, precisions=("f64",), verbose=False, clamp=None,
              nsec=None):
    """Centre bond of a U(1)-symmetric chain: t-V (t = 1, V = 2, mu = 0), XXZ
    (Jz = Jxy = 1, hz = 0) or Fermi-Hubbard (t = 1, U = 4, U(1) x U(1)), N
    particles (up spins; electrons for Hubbard, half of them up), bond
    dimension D split over charge sectors by the package's profile
    (synth.bond_dims; with clamp = (lo, hi) and nsec: the nsec heaviest sectors
    of every bond, D apportioned over them, dimensions clamped into [lo, hi]).

    The MPS is this package's seeded generator (synth.random_mps, host numpy),
    brought to mixed-canonical form around (c, c+1) on the host; the
    environments L[c-1] and R[c+2] are built on the GPU layer by layer with the
    environment plans (addContractionLayer, src/MPSfunctions.hpp:442-493), from
    getLMin / getRMax (:1917-1926); psi is BlockMatrix<2>(A_c, A_{c+1})
    (src/BlockTensor.hpp:285-301), normalised."""
    import torch
    from . import synth
    from .blocks import twosite
    dev = "cuda"
    if model == "tv":
        mpo = models.tv_mpo(L, 1.0, 2.0, 0.0)
    elif model == "xxz":
        mpo = models.xxz_mpo(L, 1.0, 1.0, 0.0)
    elif model == "hubbard":  # t = 1, U = 4; N electrons, N // 2 of them up
        mpo = models.hubbard_mpo(L, 1.0, 4.0)
    else:
        raise ValueError(model)
    if model == "hubbard":
        itk = [models.itokey_u1xu1()] * L
        N = models.pack_charge(N // 2, N - N // 2)
    else:
        itk = [models.itokey_u1()] * L
    c = L // 2 - 1
    mps, dims = synth.random_mps(L, itk, N, D, seed, clamp=clamp, nsec=nsec)
    mps = synth.mixed_canonical(mps, c)

    def run(Et, dE, s, direction):
        At = native.Table.from_blocks(mps[s])
        outt = native.env_outkeys(Et, At, mpo[s], itk[s], direction)
        plan = native.EnvPlan(Et, At, mpo[s], itk[s], direction, outt)
        dA = torch.from_numpy(At.pack(mps[s])).to(dev)
        dout = torch.zeros(max(outt.total, 1), dtype=torch.float64, device=dev)
        plan.apply(dE, dA, dout)
        return outt, dout

    Et, dE = _lmin_table(), torch.ones(1, dtype=torch.float64, device=dev)
    for s in range(0, c):
        Et, dE = run(Et, dE, s, 1)
        if verbose and s % 16 == 0:
            print(f"[bench_problem] left env site {s}", file=sys.stderr, flush=True)
    Lt, dL = Et, dE
    # getRMax: identity on the last bond's charge (N), aux 0
    Et = native.Table.from_keys(1, np.array([[N, N]], np.uint64), np.zeros((1, 1), np.uint32), [1], [1])
    dE = torch.ones(1, dtype=torch.float64, device=dev)
    for s in range(L - 1, c + 1, -1):
        Et, dE = run(Et, dE, s, 0)
        if verbose and s % 16 == 0:
            print(f"[bench_problem] right env site {s}", file=sys.stderr, flush=True)
    Rt, dR = Et, dE
    psi_b = twosite(mps[c], mps[c + 1])
    psi_b = psi_b.scaled(1.0 / psi_b.norm())
    xt = native.Table.from_blocks(psi_b)
    yt = native.heff2_outkeys(Lt, Rt, xt, mpo[c], mpo[c + 1], itk[c], itk[c + 1])
    vt = xt.union(yt)
    psi = torch.from_numpy(vt.pack(psi_b)).to(dev)
    out = dict(mpo=mpo, itk=itk[c], c=c, dims=dims, Lt=Lt, Rt=Rt, dL=dL, dR=dR, xt=xt, yt=yt, vt=vt, psi=psi, L=L,
               D=D, centre_D=int(sum(dims[c + 1].values())), psi_blocks=len(xt.keys()), psi_block_dims=(int(min(xt.rows.min(), xt.cols.min())),
                                                             int(max(xt.rows.max(), xt.cols.max()))))
    for pr in precisions:
        _add_plan(torch, out, pr, (mpo[c], mpo[c + 1]), (itk[c], itk[c + 1]), seed)
    return out
