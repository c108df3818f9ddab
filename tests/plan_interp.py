"""Numpy interpreter of a plan's device schedule (mpsa_plan_describe).

Test infrastructure: it executes the exact task lists the GPU kernels receive
(stage-1 products, W mix, zero fill, stage-2 segmented products) with numpy
matmuls, so the planner's regrouping of the reference tuple loop can be checked
against the oracle on the CPU, independently of the kernels.  FP64 layouts
only (column-major blocks, ld = rows).
"""
from __future__ import annotations

import numpy as np


def _op(buf, off, ld, rows, cols, trans):
    """op(X) of shape rows x cols where X is stored column-major at off with ld."""
    assert off >= 0 and ld >= 1
    last = off + (cols - 1 if not trans else rows - 1) * ld + (rows if not trans else cols)
    assert last <= len(buf), ("operand read past the buffer", off, ld, rows, cols, len(buf))
    if not trans:  # X is rows x cols
        out = np.zeros((rows, cols))
        for j in range(cols):
            out[:, j] = buf[off + j * ld: off + j * ld + rows]
        return out
    # X is cols x rows; op = X^T
    out = np.zeros((cols, rows))
    for j in range(rows):
        out[:, j] = buf[off + j * ld: off + j * ld + cols]
    return out.T


def run_stage(tasks, segs, tr, A, B, C, c_vs=0):
    ta, tb = bool(tr & 2), bool(tr & 1)
    for t in tasks:
        m, n = int(t["m"]), int(t["n"])
        acc = np.zeros((m, n))
        for s in segs[int(t["seg0"]):int(t["seg0"]) + int(t["nseg"])]:
            k = int(s["k"])
            if k <= 0:
                continue
            a = _op(A, int(s["a_off"]), int(s["lda"]), m, k, ta)
            b = _op(B, int(s["b_off"]), int(s["ldb"]), k, n, tb)
            acc += float(s["alpha"]) * (a @ b)
        assert 0 <= int(t["c_off"]) and int(t["c_off"]) + m * n <= len(C), "output block past the buffer"
        C[int(t["c_off"]):int(t["c_off"]) + m * n] = acc.ravel(order="F")


def run_heff(desc, L, R, x, y, ws_len):
    ws = np.zeros(ws_len)
    run_stage(desc["s1_tasks"], desc["s1_segs"], desc["s1_tasks_extra"], L, x, ws)
    for mt in desc["mix"]:
        d, ln = int(mt["dst_off"]), int(mt["len"])
        acc = np.zeros(ln)
        for tm in desc["terms"][int(mt["t0"]):int(mt["t0"]) + int(mt["nt"])]:
            acc += float(tm["coef"]) * ws[int(tm["src_off"]):int(tm["src_off"]) + ln]
        ws[d:d + ln] = acc
    for z in desc["zero"]:
        assert 0 <= int(z["off"]) and int(z["off"]) + int(z["len"]) <= len(y), "zero range past the output"
        y[int(z["off"]):int(z["off"]) + int(z["len"])] = 0.0
    run_stage(desc["s2_tasks"], desc["s2_segs"], desc["s2_tasks_extra"], ws, R, y)


def run_env(desc, BL, A, Ap, out, ws_len):
    ws = np.zeros(ws_len)
    run_stage(desc["s1_tasks"], desc["s1_segs"], desc["s1_tasks_extra"], BL, A, ws)
    for mt in desc["mix"]:
        d, ln = int(mt["dst_off"]), int(mt["len"])
        acc = np.zeros(ln)
        for tm in desc["terms"][int(mt["t0"]):int(mt["t0"]) + int(mt["nt"])]:
            acc += float(tm["coef"]) * ws[int(tm["src_off"]):int(tm["src_off"]) + ln]
        ws[d:d + ln] = acc
    for z in desc["zero"]:
        assert 0 <= int(z["off"]) and int(z["off"]) + int(z["len"]) <= len(out), "zero range past the output"
        out[int(z["off"]):int(z["off"]) + int(z["len"])] = 0.0
    run_stage(desc["s2_tasks"], desc["s2_segs"], desc["s2_tasks_extra"], Ap if Ap is not None else A, ws, out)
