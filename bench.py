"""Benchmark of the DMRG hot path on one MI355X.

Workload: the centre bond of a dense (no symmetry) spin-1/2 Heisenberg chain,
J = 1, L = 128, MPO bond w = 5, d = 2, MPS bond D = 2048, seeded Gaussian MPS
(left/right normalised by CholeskyQR2), environments built on the GPU with this
package's environment kernels.  A "step" is one two-site H_eff . psi over the
four D x D psi blocks, inputs resident in HBM.

Reported: the mixed-precision grouped matvec (default BF16X3: every operand
split into hi + lo bf16, fp32 MFMA accumulation) as `value` in TFLOP/s of
algorithmic flops (F_alg = 64 D^3 for this MPO), the FP64, FP64 plain-method,
FP32 and BF16 matvecs beside it, the roofline of the dominant GEMM stage, the
Lanczos energy of the bond in mixed precision with FP64 refinement against
all-FP64 (|dE|/L), and a CPU baseline: the oracle's restatement of the
reference tuple loop (one dgemm pair per tuple, MKL) timed on this host: one
warm-up, then the median of 5 matvecs on all cores, and a one-thread figure.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6, "f64plain": 78.6, "f32": 157.3, "bf16": 2500.0, "bf16x3": 2500.0,
        "c128": 78.6, "c128_f32": 157.3, "c128_bf16x3": 2500.0}  # MI355X dense TFLOP/s (complex: its GEMM path)
MFMA_FLOPS_PER_ALG = {"bf16x3": 3.0, "c128_bf16x3": 3.0}  # MFMA work per algorithmic flop


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def time_apply(torch, plan, x, y, steps, warmup, nvec=1, stride=0):
    """Wall time per (batched) matvec over `steps` applies after `warmup`,
    bracketed by device synchronisation; the same applies carry HIP events on
    the plan's stream, giving the average time of each stage inside the timed
    region."""
    for _ in range(warmup):
        plan.apply(x, y, nvec, stride, stride)
    torch.cuda.synchronize()
    plan.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.apply(x, y, nvec, stride, stride)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    st, n = plan.timing_summary()
    plan.set_timing(False)
    assert n == steps
    return ms, st


def time_apply_stream(torch, plan, x, y, steps, warmup):
    """Wall time per untimed single-vector apply on a stream of its own,
    where the plan replays its captured graph of launches."""
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        for _ in range(warmup):
            plan.apply(x, y)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.apply(x, y)
        s.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
    torch.cuda.synchronize()
    return ms


TILE = {0: "256x128 stream, 4 waves", 1: "64x64", 2: "128x128", 3: "256x256 pipelined, 4 waves", 4: "32x32"}
KERNEL = {0: "gemm_bf16_stream_kernel", 1: "gemm_bf16_kernel", 2: "gemm_bf16_kernel", 3: "gemm_bf16_pipe_kernel",
          4: "gemm_bf16_kernel"}
# U(1) bonds: (model, L, N, D, seed) -- the t-V chain at half filling, the XXZ
# chain of the gradient workload, the small XXZ chain
WORKLOADS = {"dense": None, "tv512": ("tv", 100, 50, 512, 20260002), "xxz1024": ("xxz", 64, 32, 1024, 20260004),
             "xxz64": ("xxz", 32, 16, 64, 20260001), "hub1024": ("hubbard", 80, 80, 1024, 20260005),
             "tv1700": ("tv", 100, 50, 8000, 20260006)}
# sector profiles other than the default: tv1700 keeps the 24 heaviest charge
# sectors of every bond, apportions 8000 over them by weight and clamps each
# into [20, 120] (centre bond: 24 sectors, sum 1698)
PROFILES = {"tv1700": {"clamp": (20, 120), "nsec": 24}}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "gemm_bf16_pipe_dense_D2048_hbm_traffic.txt")


def profiled_traffic(prec, stage):
    """HBM bytes per launch of one GEMM stage, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated; see the
    file's header).  None when the profile is absent."""
    try:
        for line in open(TRAFFIC_PROFILE):
            f = line.rstrip("\n").split("\t")
            if len(f) >= 5 and f[0] == prec and f[1] == f"stage{stage}":
                return int(float(f[4]))
    except OSError:
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _median_time(fn, reps=5):
    """1 warm-up call, then the median wall time of `reps` calls."""
    import statistics
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def cpu_baseline(prob, workload, flops_alg, flops_ref):
    """The oracle (restated reference tuple loop, one BLAS dgemm pair per
    tuple, OpenMP over left charges) on this host: 1 warm-up, then the median
    of 5 matvecs, on all cores and on one thread.  On the dense bond every
    tuple is a D x D x D product pair, so the one-thread figure is the median
    of 5 single-thread D^3 dgemm calls times the matvec's F_ref / (2 D^3)
    calls (a full one-thread matvec takes minutes)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    import numpy as np
    if not os.path.exists(pyoracle.LIB):
        pyoracle.build()
    L = prob["Lt"].unpack(prob["dL"].cpu().numpy())
    R = prob["Rt"].unpack(prob["dR"].cpu().numpy())
    psi = prob["vt"].unpack(prob["psi"].cpu().numpy())
    c, mpo, itk = prob["c"], prob["mpo"], prob["itk"]
    h = pyoracle.Handles(L, R, psi, mpo[c], mpo[c + 1], itk, itk)
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    pyoracle.set_blas_threads(cores)
    t_all = _median_time(lambda: h.matvec(cores))
    pyoracle.set_blas_threads(1)
    if workload == "dense":
        D = int(round((flops_alg / 64.0) ** (1.0 / 3.0)))
        rng = np.random.default_rng(7)
        A = np.asfortranarray(rng.standard_normal((D, D)))
        B = np.asfortranarray(rng.standard_normal((D, D)))
        calls = flops_ref / (2.0 * D ** 3)
        t_one = calls * _median_time(lambda: pyoracle.dgemm(A, B))
        one_how = (f"median of 5 one-thread {D}^3 dgemm calls x {calls:.0f} calls (F_ref / 2D^3) per matvec")
    else:
        t_one = _median_time(lambda: h.matvec(1))
        one_how = "median of 5 one-thread matvecs"
    pyoracle.set_blas_threads(cores)
    return {"t_all": t_all, "t_one": t_one, "cores": cores, "blas": pyoracle.blas_name(), "cpu": cpu_model(),
            "one_how": one_how}


def u1_bonds(torch, steps, warmup):
    """The U(1) bonds of the other workloads beside the headline, each with
    its GPU matvec (FP64 and BF16X3, tuned tiles) and the CPU baseline on the
    same bond (oracle tuple loop: 1 warm-up, median of 5, all cores and one
    thread):
      xxz64  : XXZ L = 32, D = 64 -- one matvec and the 10-step Lanczos of
               the plain-method workload, both timed on GPU and CPU;
      tv512  : t-V L = 100, D = 512 -- also 64 vectors per batched matvec;
      tv1700 : t-V L = 100, 24 sectors per bond clamped to [20, 120]
               (centre sum 1698) -- also 64 vectors per batched matvec;
      hub1024: Fermi-Hubbard L = 80, D = 1024 (U(1) x U(1)) -- the bond of
               the Chebyshev workload, real FP64."""
    from mpstools_amd import bench_problem, native
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    out = {}
    for name, nvecs in (("xxz64", (1,)), ("tv512", (1, 64)), ("tv1700", (1, 64)), ("hub1024", (1, 16))):
        model, L, N, D, seed = WORKLOADS[name]
        t0 = time.perf_counter()
        prob = bench_problem.u1_centre(model, L, N, D, seed, precisions=("f64", "bf16x3"), **PROFILES.get(name, {}))
        row = {"psi_blocks": prob["psi_blocks"], "psi_block_dims": prob["psi_block_dims"],
               "centre_D": prob["centre_D"]}
        for pr in ("f64", "bf16x3"):
            plan = prob["plan_" + pr]
            info = plan.info()
            for nv in nvecs:
                plan.autotune(nvec=nv, reps=3)
                v = prob["psi"]
                x = v.repeat(nv) if nv > 1 else v
                ms, st = time_apply(torch, plan, x, torch.zeros_like(x), steps, warmup, nv,
                                    v.numel() if nv > 1 else 0)
                row[f"{pr}_nvec{nv}"] = {"ms": round(ms, 4), "tflops": round(nv * info["flops_alg"] / ms / 1e9, 3)}
                if nv == 1:
                    # the same apply untimed on its own stream (graph replay), beside its event-timed stages
                    row[f"{pr}_nvec1"]["ms_graph_replay"] = round(
                        time_apply_stream(torch, plan, x, torch.zeros_like(x), steps, warmup), 4)
                    row[f"{pr}_nvec1"]["stages_ms_events"] = round(float(sum(st)), 4)
            row["flops_alg"] = info["flops_alg"]
            row["flops_ref"] = info["flops_ref"]
        cb = cpu_baseline(prob, name, row["flops_alg"], row["flops_ref"])
        row["cpu"] = {"ms_all_cores": round(cb["t_all"] * 1e3, 3), "ms_one_thread": round(cb["t_one"] * 1e3, 3),
                      "cores": cb["cores"], "blas": cb["blas"], "kind": "port",
                      "sample": "oracle tuple loop on the same bond, 1 warm-up then the median of 5"}
        row["gpu_f64_vs_cpu_all_cores"] = round(cb["t_all"] * 1e3 / row["f64_nvec1"]["ms"], 1)
        if name == "xxz64":
            # the plain-method workload: one matvec plus a 10-step Lanczos
            plan = prob["plan_f64"]
            torch.cuda.synchronize()
            t_gpu = _median_time(lambda: native.lanczos(plan, prob["psi"], 10, 1e-12, True, want_state=False))
            Lb = prob["Lt"].unpack(prob["dL"].cpu().numpy())
            Rb = prob["Rt"].unpack(prob["dR"].cpu().numpy())
            psi = prob["vt"].unpack(prob["psi"].cpu().numpy())
            c, mpo, itk = prob["c"], prob["mpo"], prob["itk"]
            t_cpu = _median_time(lambda: pyoracle.lanczos_twosite(Lb, Rb, psi, mpo[c], mpo[c + 1], itk, itk, 10, 1e-12,
                                                                  True, nthreads=cb["cores"]))
            row["lanczos10"] = {"gpu_ms": round(t_gpu * 1e3, 3), "cpu_ms_all_cores": round(t_cpu * 1e3, 3)}
        row["build_s"] = round(time.perf_counter() - t0, 1)
        log(name, json.dumps(row))
        out[name] = row
        del prob
        native.lib().mpsa_release_pools()
    return out


def chebyshev(torch, args):
    """The Chebyshev workload: spinful Fermi-Hubbard chain (t = 1, U = 4,
    U(1) x U(1) charges, d = 4), L = 80, half filling, MPO rescaled into
    [-1, 1], nvec seeded start vectors (device generator, canonicalised on
    the device) of bond dimension D in complex128, moments of
    simulateSSCompress (src/container.hpp:2343-2470) with chimax = D and 2
    compression sweeps.  A moment n >= 2 is one compressSingleSite of H t_{n-1},
    the sum 2 H t_{n-1} - t_{n-2}, one compress of it and the overlaps; its wall
    time (moment 2 of a 3-moment run) is the figure, measured with the
    one-site contraction in FP64 (complex128) and in FP32 (complex64).  The CPU
    baseline is the oracle's numpy restatement (oracle/pycheb.py) on a bounded
    sample: one start vector, real, D = 32, the same chain."""
    from mpstools_amd import models, native
    L, D, nv = 80, args.cheb_D, args.cheb_nvec
    itk = [models.itokey_u1xu1()] * L
    N = models.pack_charge(L // 2, L // 2)
    mpo = models.rescale_mpo(models.hubbard_mpo(L, 1.0, 4.0), 4.0 * (L - 1) + 4.0 * L, 0.0)
    out = {"workload": f"Fermi-Hubbard t=1 U=4, L={L}, d=4, U(1)xU(1), MPO w={mpo[1].Dr}, D={D}, {nv} start "
                       f"vectors, complex128 states, chimax=D, 2 compression sweeps"}
    legs = {"complex128": native.F64, "complex64_contraction": native.F32}
    for name in args.cheb_legs.split(","):
        prec = legs[name]
        t0 = time.perf_counter()
        e = native.ChebRandom(L, itk, N, D, nv, mpo, args.seed + 2, cplx=True, h1_precision=prec)
        if args.cheb_block_min is not None:
            e.set_block_jacobi(args.cheb_block_min)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        r = e.moments(3, D, 2)
        st = e.stats()
        host_timers = e.timers()
        out[name] = {"s_per_moment": round(float(r["moment_s"][2]), 3),
                     "moment_s": [round(float(x), 3) for x in r["moment_s"]], "setup_s": round(setup, 2),
                     "mu": [[round(float(z.real), 10), round(float(z.imag), 10)] for z in r["moments"][0]],
                     "host_plan_s": round(st["host_plan_s"], 2), "host_keys_s": round(st["host_keys_s"], 2),
                     "device_bytes": st["pool_bytes"], "host_timers": host_timers}
        log("chebyshev", name, json.dumps(out[name]))
        del e
        native.lib().mpsa_release_pools()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pycheb
        from mpstools_amd import synth
        Dc = 32
        mps, _ = synth.random_mps(L, itk, N, Dc, 7)
        for i in range(L - 1, 0, -1):
            mps[i], mps[i - 1] = synth.right_normalize(mps[i], mps[i - 1])
        mps[0] = mps[0].scaled(1.0 / mps[0].norm())
        t0 = time.perf_counter()
        pycheb.chebyshev_moments(mpo, pycheb.MPS(mps, itk), 3, Dc, 2)
        tc = time.perf_counter() - t0
        e = native.ChebRandom(L, itk, N, Dc, 1, mpo, args.seed + 3, cplx=False)
        r = e.moments(3, Dc, 2)
        out["cpu_baseline"] = {"s_3_moments": round(tc, 3), "kind": "port",
                               "sample": f"oracle numpy restatement, 1 real start vector, D={Dc}, 3 moments",
                               "gpu_s_3_moments_same_sample": round(float(r["s"]), 3)}
        del e
    return out


def lanczos_compare(torch, prob, mixed, args):
    """Bond ground state, the reference's plain Lanczos (TSLanzcosEngine rules)
    run to convergence (|dE| < 1e-10 between tests) in three ways:
      fp64   : every matvec in FP64;
      mixed  : matvecs in `mixed` precision, recurrence/tridiagonal solve in FP64;
      refined: the mixed Ritz vector followed by a short FP64 Lanczos (at most
               `refine_it` FP64 matvecs), the final FP64 refinement step.
    |dE|/L is reported for mixed and refined against fp64."""
    from mpstools_amd import native
    L = prob["L"]
    psi = prob["psi"]
    it_max, delta = args.lanczos_itmax, 1e-10
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g64 = native.lanczos(prob["plan_f64"], psi, it_max, delta, True)
    t64 = time.perf_counter() - t0
    log(f"lanczos f64: {g64['it']} it, E={g64['energies'][0]:.12f}, {t64:.2f}s")
    t0 = time.perf_counter()
    gm = native.lanczos(prob["plan_" + mixed], psi, it_max, delta, True)
    tm = time.perf_counter() - t0
    t0 = time.perf_counter()
    gr = native.lanczos(prob["plan_f64"], gm["state"], args.refine_it, delta, True, want_state=False)
    tr = time.perf_counter() - t0
    e64, emix, eref = g64["energies"][0], gm["energies"][0], gr["energies"][0]
    log(f"lanczos {mixed}: {gm['it']} it E={emix:.12f} {tm:.2f}s; fp64 refinement {gr['it']} it "
        f"E={eref:.12f} {tr:.2f}s")
    native.lib().mpsa_release_pools()
    return {"E_fp64": e64, "it_fp64": g64["it"], "s_fp64": round(t64, 4),
            "E_mixed": emix, "it_mixed": gm["it"], "s_mixed": round(tm, 4),
            "E_refined": eref, "it_refine_fp64": gr["it"], "s_refine": round(tr, 4),
            "dE_per_site_mixed": abs(emix - e64) / L, "dE_per_site_refined": abs(eref - e64) / L,
            "meets_1e-9": bool(abs(eref - e64) / L <= 1e-9),
            "solve_speedup": round(t64 / (tm + tr), 3)}


def sweep_segment(torch, prob, mixed, args, nbonds):
    """nbonds consecutive full-D bond steps of a left-to-right two-site sweep
    (DMRGengine::doStep, src/container.hpp:293-337), each timed on the
    device from its parts:
      lanczos : the mixed-precision Lanczos from the bond's two-site state,
                then at most refine_it FP64 steps from its Ritz vector;
      split   : split of the refined ground state keeping chi = D
                (src/MPSfunctions.hpp:1142-1278), lambda into the right tensor;
      env     : L[b] = addContractionLayer(L[b-1], A_b) (:442-493);
      twosite : the next bond's state BlockMatrix<2>(lambda V_{b+1}, A_{b+2})
                (src/BlockTensor.hpp:285-301) -- the reference's sweep makes
                it the same way;
    plus building the bond's H_eff plans.  A sweep is 2 (L - 2) bond steps;
    the measured steps are bulk bonds, so sweep_s = 2 (L - 2) x the mean
    measured step is an upper estimate (bonds near the ends are smaller)."""
    from mpstools_amd import native
    c, mpo, itk, D = prob["c"], prob["mpo"], prob["itk"], prob["D"]
    Lt, dL = prob["Lt"], prob["dL"]
    vt, psi = prob["vt"], prob["psi"]
    steps = []
    for b in range(c, c + nbonds):
        rec = {"bond": b}
        t0 = time.perf_counter()
        Rt, dR = (prob["Rt"], prob["dR"]) if b == c else prob["R"][b + 2]
        if b == c:
            pm, p64 = prob["plan_" + mixed], prob["plan_f64"]
        else:
            pm = native.HeffPlan(Lt, Rt, vt, vt, (mpo[b], mpo[b + 1]), (itk, itk), native.PREC[mixed])
            pm.bind_env(dL, dR)
            p64 = native.HeffPlan(Lt, Rt, vt, vt, (mpo[b], mpo[b + 1]), (itk, itk), native.F64)
            p64.bind_env(dL, dR)
        torch.cuda.synchronize()
        rec["plans_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        gm = native.lanczos(pm, psi, args.lanczos_itmax, 1e-10, True)
        gr = native.lanczos(p64, gm["state"], args.refine_it, 1e-10, True)
        torch.cuda.synchronize()
        rec["lanczos_s"] = time.perf_counter() - t0
        rec["E"] = gr["energies"][0]
        rec["it"] = gm["it"] + gr["it"]
        t0 = time.perf_counter()
        sp = native.split(vt, gr["state"], itk, itk, direction=1, chimax=D)
        torch.cuda.synchronize()
        rec["split_s"] = time.perf_counter() - t0
        rec["chi"], rec["tw"] = sp["chi"], sp["tw"]
        At, dA = sp["left"]
        Bt, dB = sp["right"]
        t0 = time.perf_counter()
        outt = native.env_outkeys(Lt, At, mpo[b], itk, 1)
        ep = native.EnvPlan(Lt, At, mpo[b], itk, 1, outt)
        dLn = torch.zeros(max(outt.total, 1), dtype=torch.float64, device="cuda")
        ep.apply(dL, dA, dLn)
        torch.cuda.synchronize()
        rec["env_s"] = time.perf_counter() - t0
        Lt, dL = outt, dLn
        t0 = time.perf_counter()
        if b + 1 < c + nbonds:
            Ct, dC = prob["sites"][b + 2]
            xt = native.twosite_outkeys(Bt, Ct)
            dx = torch.zeros(max(xt.total, 1), dtype=torch.float64, device="cuda")
            native.twosite_product(Bt, dB, Ct, dC, xt, dx)
            yt = native.heff2_outkeys(Lt, prob["R"][b + 3][0], xt, mpo[b + 1], mpo[b + 2], itk, itk)
            vt = xt.union(yt)
            psi = torch.zeros(max(vt.total, 1), dtype=torch.float64, device="cuda")
            for k_, (o_new, o_old, n_) in enumerate(_block_moves(xt, vt)):
                psi[o_new:o_new + n_] = dx[o_old:o_old + n_]
        torch.cuda.synchronize()
        rec["twosite_s"] = time.perf_counter() - t0
        rec["step_s"] = sum(rec[k] for k in ("plans_s", "lanczos_s", "split_s", "env_s", "twosite_s"))
        log(f"bond {b}: E={rec['E']:.10f} it={rec['it']} chi={rec['chi']} " +
            " ".join(f"{k}={rec[k]:.3f}" for k in ("plans_s", "lanczos_s", "split_s", "env_s", "twosite_s", "step_s")))
        steps.append({k: (round(v, 4) if isinstance(v, float) and k != "E" and k != "tw" else v)
                      for k, v in rec.items()})
        del pm, p64
        native.lib().mpsa_release_pools()
    mean = sum(r["step_s"] for r in steps) / len(steps)
    L = prob["L"]
    return {"bond_steps_measured": len(steps), "steps": steps, "mean_bond_step_s": round(mean, 4),
            "bond_steps_per_sweep": 2 * (L - 2), "sweep_s": round(2 * (L - 2) * mean, 1),
            "sweep_s_kind": f"measured {len(steps)} consecutive full-D bulk bond steps (mixed {mixed} Lanczos + "
                            f"FP64 refinement, split chi = D, env update, two-site product, plan builds), "
                            f"extrapolated to 2 (L - 2) = {2 * (L - 2)} steps (upper estimate: end bonds are "
                            f"smaller)"}


def _block_moves(src: "Table", dst: "Table"):
    """(dst offset, src offset, elements) of every block of src inside dst."""
    where = {k: i for i, k in enumerate(src.keys())}
    out = []
    for i, k in enumerate(dst.keys()):
        if k in where:
            j = where[k]
            out.append((int(dst.offsets[i]), int(src.offsets[j]), int(src.rows[j] * src.cols[j])))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--D", type=int, default=2048)
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--seed", type=int, default=20260003)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precisions", default="f64,f64plain,f32,bf16,bf16x3,c128,c128_bf16x3",
                    help="plans to time; c128 and c128_<prec> are complex128 plans of the same bond")
    ap.add_argument("--mixed", default="bf16x3", help="precision of the mixed-precision matvec")
    ap.add_argument("--lanczos-itmax", type=int, default=700)
    ap.add_argument("--refine-it", type=int, default=10)
    ap.add_argument("--no-lanczos", action="store_true")
    ap.add_argument("--workload", default="dense", choices=sorted(WORKLOADS),
                    help="dense: the D=2048 dense bond (the headline); the others are U(1) bonds")
    ap.add_argument("--nvec", type=int, default=1, help="vectors per batched apply")
    ap.add_argument("--sweep-bonds", type=int, default=3,
                    help="full-D bond steps of a sweep measured on the dense bond (0 skips)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the plans' heuristic tile classes instead of autotuning them")
    ap.add_argument("--no-u1", action="store_true", help="skip the U(1) bonds of the other workloads")
    ap.add_argument("--cheb-D", type=int, default=1024, help="bond dimension of the Chebyshev workload (0 skips)")
    ap.add_argument("--cheb-nvec", type=int, default=16, help="start vectors of the Chebyshev workload")
    ap.add_argument("--cheb-block-min", type=int, default=None,
                    help="sectors of at least this many columns are factorised by several workgroups "
                         "(engine default 32; 0: one workgroup per sector)")
    ap.add_argument("--cheb-legs", default="complex128,complex64_contraction",
                    help="Chebyshev legs to run: complex128 and/or complex64_contraction")
    ap.add_argument("--calib", action="store_true",
                    help="run a known-size axpby (2 GiB read, 1 GiB written) for HBM counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; multi-GPU sharding is described in "
                                   "DESIGN.md but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from mpstools_amd import bench_problem
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    precs = args.precisions.split(",")
    t0 = time.perf_counter()
    if args.workload == "dense":
        log(f"building dense Heisenberg centre problem L={args.L} D={args.D}")
        prob = bench_problem.dense_heisenberg_centre(D=args.D, L=args.L, seed=args.seed, precisions=precs,
                                                     verbose=True, keep_right=args.sweep_bonds)
        workload = (f"dense Heisenberg J=1 centre bond, L={args.L}, w=5, d=2, D={args.D}, seeded Gaussian MPS, "
                    f"QR-normalised, GPU-built environments")
    else:
        model, L, N, D, seed = WORKLOADS[args.workload]
        log(f"building U(1) {model} centre problem L={L} N={N} D={D}")
        prob = bench_problem.u1_centre(model, L, N, D, seed, precisions=precs, verbose=True,
                                       **PROFILES.get(args.workload, {}))
        name = {"tv": "U(1) t-V t=1 V=2", "xxz": "U(1) XXZ Jz=Jxy=1",
                "hubbard": "U(1)xU(1) Fermi-Hubbard t=1 U=4 (real FP64, N_up=N_dn)"}[model]
        dlabel = f"D={D}" if args.workload not in PROFILES else (
            f"sum D={prob['centre_D']} (24 heaviest sectors per bond, dims clamped to [20, 120])")
        workload = (f"{name} centre bond, L={L}, N={N}, w={6 if model == 'hubbard' else 5}, "
                    f"d={4 if model == 'hubbard' else 2}, "
                    f"{dlabel} over charge sectors ({prob['psi_blocks']} psi blocks, dims "
                    f"{prob['psi_block_dims'][0]}-{prob['psi_block_dims'][1]}), seeded MPS, GPU-built environments")
    torch.cuda.synchronize()
    log(f"problem built in {time.perf_counter() - t0:.1f}s")
    nvec = max(1, args.nvec)

    def vectors(pr):
        """Start vector(s) of precision pr's plan (complex plans: planar re/im)."""
        v = prob.get("psi_" + pr, prob["psi"])
        x_ = v.repeat(nvec) if nvec > 1 else v
        return x_, torch.zeros_like(x_), (v.numel() if nvec > 1 else 0)
    if args.calib:
        from mpstools_amd import native
        n = (1 << 30) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.ones(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        native.axpby(1.0, a, 1.0, b)
        torch.cuda.synchronize()
        log("calibration axpby: 2 GiB read, 1 GiB written")
        del a, b
    res = {}
    for pr in precs:
        plan = prob["plan_" + pr]
        info = plan.info()
        tuning = None
        if not args.no_tune:
            # time every admissible tile class of both GEMM stages, keep the fastest
            t = plan.autotune(nvec=nvec, reps=3)
            tuning = {"stage%d" % s_: {str(c): round(ms_, 4) for (s2, c), ms_ in sorted(t.items()) if s2 == s_}
                      for s_ in (1, 2)}
        si = plan.stage_info()
        x, y, stride = vectors(pr)
        ms, st = time_apply(torch, plan, x, y, args.steps, args.warmup, nvec, stride)
        res[pr] = dict(ms=ms, tflops=nvec * info["flops_alg"] / ms / 1e9, flops_alg=info["flops_alg"],
                       flops_ref=info["flops_ref"], stage_ms=st, stage_flops=[si["flops1"], si["flops2"]],
                       tile_class=[si["tile_class1"], si["tile_class2"]], tuning=tuning)
        log(pr, f"{ms:.3f} ms/matvec", f"{res[pr]['tflops']:.1f} TFLOP/s", "stages(ms)", [round(s, 3) for s in st])
    mixed = args.mixed if args.mixed in res else precs[-1]
    r = res[mixed]
    # dominant kernel of the mixed matvec: the slower GEMM stage
    k = 1 if r["stage_ms"][1] >= r["stage_ms"][3] else 3
    stage = 1 if k == 1 else 2
    fl = nvec * r["stage_flops"][stage - 1]
    achieved = fl / r["stage_ms"][k] / 1e9
    mf = MFMA_FLOPS_PER_ALG.get(mixed, 1.0)
    roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK[mixed], "unit": "TFLOP/s",
            "frac": round(achieved / PEAK[mixed], 4),
            "mfma_issue_frac": round(achieved * mf / PEAK[mixed], 4),
            "traffic": profiled_traffic(mixed, stage) if args.workload == "dense" and nvec == 1 else None,
            "kernel": "%s %s, stage %d (%s)" % (KERNEL.get(r["tile_class"][stage - 1], "gemm_bf16_kernel"),
                                                 TILE.get(r["tile_class"][stage - 1], "?"), stage, mixed),
            "stage_ms": round(r["stage_ms"][k], 4), "stage_flops_alg": fl,
            "note": "achieved = algorithmic flops of the stage / its average HIP-event duration over the timed "
                    "steps, frac = achieved / peak; mfma_issue_frac counts the MFMA work the kernel issues "
                    "(%.0fx the algorithmic flops for %s); traffic = HBM bytes per launch from "
                    "profiles/gemm_bf16_pipe_dense_D2048_hbm_traffic.txt" % (mf, mixed)}
    lan = None
    if not args.no_lanczos and "f64" in res:
        lan = lanczos_compare(torch, prob, mixed, args)
    sweep = None
    if args.sweep_bonds > 0 and args.workload == "dense" and "f64" in res:
        sweep = sweep_segment(torch, prob, mixed, args, args.sweep_bonds)
    u1 = None
    if args.workload == "dense" and not args.no_u1:
        u1 = u1_bonds(torch, args.steps, args.warmup)
    cheb = None
    if args.workload == "dense" and args.cheb_D > 0:
        cheb = chebyshev(torch, args)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline: oracle contractTwosite on host cores (1 warm-up, median of 5; all cores and 1 thread)")
        cb = cpu_baseline(prob, args.workload, r["flops_alg"], r["flops_ref"])
        cpu = {"value": round(r["flops_alg"] / cb["t_all"] / 1e12, 5), "unit": "TFLOP/s", "cores": cb["cores"],
               "kind": "port",
               "sample": f"two-site matvec of the same bond ({args.workload}), the reference tuple loop with one "
                         f"{cb['blas']} dgemm pair per tuple: 1 warm-up then the median of 5, "
                         f"{cb['t_all'] * 1e3:.1f} ms on {cb['cores']} cores; value = F_alg / t",
               "cpu_model": cb["cpu"], "blas": cb["blas"],
               "ms_all_cores": round(cb["t_all"] * 1e3, 3), "ms_one_thread": round(cb["t_one"] * 1e3, 3),
               "value_one_thread": round(r["flops_alg"] / cb["t_one"] / 1e12, 5),
               "one_thread_sample": cb["one_how"],
               "gpu_speedup_vs_all_cores": round(cb["t_all"] * 1e3 / r["ms"], 1)}
    out = {
        "metric": "H_eff\u00b7\u03c8 TFLOP/s + DMRG sweep s at D=2048; mixed vs FP64 at |\u0394E/L|\u22641e-9",
        "value": round(r["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3 (split bf16 in, fp32 acc)" if mixed == "bf16x3" else mixed,
        "data": "synthetic",
        "config": {"workload": workload, "global_batch": nvec, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "by_precision": {pr: {"ms": round(v["ms"], 4), "tflops": round(v["tflops"], 3),
                              "frac_of_peak": round(v["tflops"] / PEAK[pr], 4),
                              "stage_ms": [round(s, 4) for s in v["stage_ms"]],
                              "tile_class": v["tile_class"],
                              "tile_tuning_ms": v["tuning"]} for pr, v in res.items()},
        "flops_alg": r["flops_alg"],
        "flops_ref": r["flops_ref"],
    }
    if "f64" in res:
        out["speedup_mixed_vs_f64"] = round(res["f64"]["ms"] / r["ms"], 3)
    if "f64plain" in res:
        out["speedup_mixed_vs_f64_plain_method"] = round(res["f64plain"]["ms"] / r["ms"], 3)
    if lan is not None:
        out["lanczos"] = lan
    if sweep is not None:
        out["dmrg_sweep"] = sweep
    if u1 is not None:
        out["u1_bonds"] = u1
    if cheb is not None:
        out["chebyshev"] = cheb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
