"""ctypes binding of the gfx950 library (libmpstools_amd.so, include/mpstools_amd.h).

This module is the Python mirror of the reference's contraction interface for
the hot path: `contract_twosite`, `contract_single` and `add_contraction_layer`
take the same arguments as the reference's contractTwosite / contractSingle /
addContractionLayer (src/MPSfunctions.hpp:949, 857, 442) -- environments,
tensor, MPO matrices, i_to_key -- as host `Blocks` and return host `Blocks`.
The plan classes keep everything device-resident for repeated application
(Lanczos).  Device memory comes from torch (plumbing only); every contraction
runs in the HIP library, and importing this module fails loudly when the
library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .blocks import Blocks

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmpstools_amd.so")

F64, F32, BF16, BF16X3 = 0, 1, 2, 3
TILE_CLASSES = 5  # tile classes per GEMM stage in mpsa_plan_autotune's table
PLAIN = 0x100
PREC = {"f64": F64, "fp64": F64, "f32": F32, "fp32": F32, "bf16": BF16, "bf16x3": BF16X3, "f64plain": F64 | PLAIN}


class MpsaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mpsa error {code}: {msg}")
        self.code = code


class _Blocks(C.Structure):
    _fields_ = [("n", C.c_int64), ("ns", C.c_int32), ("is_complex", C.c_int32), ("q", C.c_void_p),
                ("p", C.c_void_p), ("rows", C.c_void_p), ("cols", C.c_void_p), ("offset", C.c_void_p)]


class _MPO(C.Structure):
    _fields_ = [("n", C.c_int64), ("Dl", C.c_uint32), ("Dr", C.c_uint32), ("is_complex", C.c_int32),
                ("wl", C.c_void_p), ("wr", C.c_void_p), ("out", C.c_void_p), ("in_", C.c_void_p),
                ("val", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"HIP library not built: {LIBPATH} (run __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        P = C.c_void_p
        L.mpsa_last_error.restype = C.c_char_p
        L.mpsa_version.restype = C.c_char_p
        L.mpsa_heff2_outkeys.argtypes = [P, P, P, P, P, P, C.c_int, P, C.c_int, C.c_int64, P, P, P, P, P]
        L.mpsa_twosite_outkeys.argtypes = [P, P, C.c_int64, P, P, P, P, P]
        L.mpsa_twosite_product.argtypes = [P, P, P, P, P, P, P]
        L.mpsa_heff1_outkeys.argtypes = [P, P, P, P, P, C.c_int, C.c_int64, P, P, P, P, P]
        L.mpsa_env_outkeys.argtypes = [P, P, P, P, P, C.c_int, C.c_int, C.c_int64, P, P, P, P, P]
        L.mpsa_heff2_plan.argtypes = [P, P, P, P, P, P, P, P, C.c_int, P, C.c_int, C.c_int]
        L.mpsa_heff1_plan.argtypes = [P, P, P, P, P, P, P, C.c_int, C.c_int]
        L.mpsa_env_plan.argtypes = [P, P, P, P, P, P, C.c_int, C.c_int, P, C.c_int]
        L.mpsa_plan_bind_env.argtypes = [P, P, P, P]
        L.mpsa_heff_apply.argtypes = [P, P, C.c_int64, P, C.c_int64, C.c_int, P]
        L.mpsa_env_apply.argtypes = [P, P, P, P, P, P]
        L.mpsa_plan_info.argtypes = [P, P, P, P, P, P]
        L.mpsa_plan_destroy.argtypes = [P]
        L.mpsa_plan_stage_info.argtypes = [P, P, P, P, P]
        L.mpsa_pool_info.argtypes = [P, P]
        L.mpsa_pool_set_cap.argtypes = [C.c_int64]
        L.mpsa_plan_set_timing.argtypes = [P, C.c_int]
        L.mpsa_plan_describe.argtypes = [P, C.c_int, P, C.c_int64, P, P]
        L.mpsa_plan_last_times.argtypes = [P, P]
        L.mpsa_plan_timing_summary.argtypes = [P, P, P]
        L.mpsa_plan_vector_lengths.argtypes = [P, P, P, P]
        L.mpsa_lanczos.argtypes = [P, P, P, C.c_int, C.c_double, C.c_int, C.c_int, P, C.c_int, P, P, P]
        L.mpsa_lanczos_batched.argtypes = [P, P, C.c_int64, C.c_int, P, C.c_int, C.c_double, C.c_int, P, C.c_int,
                                            P, P, C.c_int, P]
        L.mpsa_lanczos_ex.argtypes = [P, P, P, C.c_int, C.c_double, C.c_int, C.c_int, P, C.c_int, P, P, P, C.c_int,
                                       C.c_int64, C.c_int, P]
        L.mpsa_lanczos_batched_ex.argtypes = [P, P, C.c_int64, C.c_int, P, C.c_int, C.c_double, C.c_int, P, C.c_int,
                                               P, P, C.c_int, P, C.c_int, C.c_int64, C.c_int64, C.c_int, P]
        L.mpsa_chebyshev_moments.argtypes = [P, P, C.c_int, C.c_double, C.c_double, P, P, P]
        L.mpsa_chain_create.argtypes = [P, C.c_int, P, P, P, P, C.c_int, P, C.c_int]
        L.mpsa_energy_forward.argtypes = [P, P, P, P]
        L.mpsa_energy_backward.argtypes = [P, P, P, P, P]
        L.mpsa_chain_env_table.argtypes = [P, C.c_int, C.c_int, C.c_int64, P, P, P, P, P, P]
        L.mpsa_chain_destroy.argtypes = [P]
        L.mpsa_split.argtypes = [P, P, P, C.c_int, P, C.c_int, C.c_int, C.c_uint, C.c_int, C.c_double, C.c_int, P, P]
        L.mpsa_split_info.argtypes = [P, P, P, P, P, P, P, P]
        L.mpsa_split_table.argtypes = [P, C.c_int, P, P, P, P, P]
        L.mpsa_split_data.argtypes = [P, C.c_int]
        L.mpsa_split_data.restype = P
        L.mpsa_split_copy.argtypes = [P, C.c_int, P, P]
        L.mpsa_split_lambda.argtypes = [P, P, P, P]
        L.mpsa_split_destroy.argtypes = [P]
        L.mpsa_dmrg_create.argtypes = [P, C.c_int, P, P, P, P, P, P]
        L.mpsa_dmrg_sweep.argtypes = [P, C.c_int, C.c_int, C.c_uint, C.c_int, C.c_double, C.c_int, C.c_double, P, P,
                                      P, P]
        L.mpsa_dmrg_site.argtypes = [P, C.c_int, P, P, P, P, P, P, P, P, P]
        L.mpsa_dmrg_destroy.argtypes = [P]
        L.mpsa_tridiag_eig.argtypes = [C.c_int, P, P, P]
        L.mpsa_cheb_create.argtypes = [P, C.c_int, C.c_int, P, P, P, P, P, C.c_int, P]
        L.mpsa_cheb_copy.argtypes = [P, C.c_int, C.c_int]
        L.mpsa_cheb_orthonormalize.argtypes = [P, C.c_int]
        L.mpsa_fill_gauss.argtypes = [P, C.c_int64, C.c_uint64, P]
        L.mpsa_cheb_apply_mpo.argtypes = [P, C.c_int, C.c_int]
        L.mpsa_cheb_overlap.argtypes = [P, C.c_int, C.c_int, P]
        L.mpsa_cheb_axpby.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int]
        L.mpsa_cheb_compress_single_site.argtypes = [P, C.c_int, C.c_int, C.c_uint, C.c_int, C.c_double, P]
        L.mpsa_cheb_compress.argtypes = [P, C.c_int, C.c_uint, C.c_int, C.c_double, P]
        L.mpsa_cheb_moments.argtypes = [P, C.c_int, C.c_uint, C.c_int, C.c_double, P, P, P, P, P]
        L.mpsa_cheb_site.argtypes = [P, C.c_int, C.c_int, C.c_int, P, P, P, P, P, P, P, P]
        L.mpsa_cheb_stats.argtypes = [P, P, P]
        L.mpsa_cheb_profile.argtypes = [P, P]
        L.mpsa_cheb_moment_times.argtypes = [P, P, C.c_int, P]
        L.mpsa_cheb_set_big_sector.argtypes = [P, C.c_int]
        L.mpsa_cheb_set_block_jacobi.argtypes = [P, C.c_int]
        L.mpsa_cheb_timers.argtypes = [P, C.POINTER(C.c_double), C.c_int]
        L.mpsa_cheb_factor_log.argtypes = [P, C.c_int64]
        L.mpsa_cheb_factor_records.argtypes = [P, P, C.c_int64, P]
        L.mpsa_cheb_destroy.argtypes = [P]
        L.mpsa_dot.argtypes = [P, P, C.c_int64, C.c_int, P, P]
        L.mpsa_axpby.argtypes = [P, P, P, P, C.c_int64, C.c_int, P]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MpsaError(rc, lib().mpsa_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Table:
    """A key table (mpsa_blocks) with its arrays kept alive."""

    def __init__(self, ns, q, p, rows, cols, offsets, is_complex=0):
        self.ns = ns
        self.is_complex = int(is_complex)
        self.q = np.ascontiguousarray(q, dtype=np.uint64).reshape(-1, 2)
        self.p = np.ascontiguousarray(p, dtype=np.uint32).reshape(-1, ns)
        self.rows = np.ascontiguousarray(rows, dtype=np.int64)
        self.cols = np.ascontiguousarray(cols, dtype=np.int64)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.n = len(self.rows)
        self.total = int(max([0] + [int(o + r * c) for o, r, c in zip(self.offsets, self.rows, self.cols)]))
        self.s = _Blocks(self.n, ns, self.is_complex, self.q.ctypes.data, self.p.ctypes.data, self.rows.ctypes.data,
                         self.cols.ctypes.data, self.offsets.ctypes.data)

    @staticmethod
    def from_blocks(B: Blocks, align=64):
        t = B.table(align)
        return Table(B.ns, t["q"], t["p"], t["rows"], t["cols"], t["offsets"], B.dtype.kind == "c")

    def as_complex(self) -> "Table":
        """The same keys and offsets as a complex128 table (planar layout: the
        real parts, then the imaginary parts one extent further on)."""
        return Table(self.ns, self.q, self.p, self.rows, self.cols, self.offsets, 1)

    @property
    def length(self):
        """Doubles of a packed buffer: the extent, twice for complex tables."""
        return self.total * (2 if self.is_complex else 1)

    @staticmethod
    def from_keys(ns, q, p, rows, cols, align=64):
        offs = np.zeros(len(rows), np.int64)
        off = 0
        for i in range(len(rows)):
            off = (off + align - 1) // align * align
            offs[i] = off
            off += int(rows[i]) * int(cols[i])
        return Table(ns, q, p, rows, cols, offs)

    def as_dict(self):
        return dict(q=self.q, p=self.p, rows=self.rows, cols=self.cols, offsets=self.offsets, total=self.total)

    def keys(self):
        return [(int(self.q[i, 0]), int(self.q[i, 1])) + tuple(int(x) for x in self.p[i]) for i in range(self.n)]

    def union(self, other: "Table", align=64) -> "Table":
        """Table over the union of both key sets (key order)."""
        d = {}
        for t in (self, other):
            for i, k in enumerate(t.keys()):
                d[k] = (int(t.rows[i]), int(t.cols[i]))
        ks = sorted(d)
        q = np.array([[k[0], k[1]] for k in ks], np.uint64).reshape(-1, 2)
        p = np.array([k[2:] for k in ks], np.uint32).reshape(-1, self.ns)
        t = Table.from_keys(self.ns, q, p, [d[k][0] for k in ks], [d[k][1] for k in ks], align)
        return t.as_complex() if self.is_complex else t

    def pack(self, B: Blocks, dtype=np.float64):
        if self.is_complex:
            z = B.pack(table=self.as_dict()).astype(np.complex128)
            return np.concatenate([z.real, z.imag]).astype(dtype)
        return B.pack(table=self.as_dict()).astype(dtype)

    def unpack(self, buf, keep=None) -> Blocks:
        buf = np.asarray(buf)
        if self.is_complex:
            buf = buf[:self.total] + 1j * buf[self.total:2 * self.total]
        return Blocks.unpack(self.ns, self.as_dict(), buf, keep=keep)


class MPOArg:
    def __init__(self, W):
        self.wl, self.wr, self.o, self.i, self.v = W.arrays()
        self.s = _MPO(len(self.wl), W.Dl, W.Dr, 0, self.wl.ctypes.data, self.wr.ctypes.data, self.o.ctypes.data,
                      self.i.ctypes.data, self.v.ctypes.data)


def _itk(itk):
    return np.ascontiguousarray(np.array(itk, dtype=np.uint64))


def _outkeys(fn, args, ns):
    n = np.zeros(1, np.int64)
    _check(fn(*args, 0, _ptr(n), None, None, None, None))
    m = int(n[0])
    q = np.zeros((max(m, 1), 2), np.uint64)
    p = np.zeros((max(m, 1), ns), np.uint32)
    rows = np.zeros(max(m, 1), np.int64)
    cols = np.zeros(max(m, 1), np.int64)
    if m:
        _check(fn(*args, m, _ptr(n), _ptr(q), _ptr(p), _ptr(rows), _ptr(cols)))
    return q[:m], p[:m], rows[:m], cols[:m]


def heff2_outkeys(Lt: Table, Rt: Table, xt: Table, Wl, Wr, itl, itr, align=64) -> Table:
    a, b = _itk(itl), _itk(itr)
    ml, mr = MPOArg(Wl), MPOArg(Wr)
    q, p, r, c = _outkeys(lib().mpsa_heff2_outkeys,
                          [C.byref(Lt.s), C.byref(Rt.s), C.byref(xt.s), C.byref(ml.s), C.byref(mr.s), _ptr(a),
                           len(a), _ptr(b), len(b)], 2)
    return Table.from_keys(2, q, p, r, c, align)


def twosite_outkeys(t1: Table, t2: Table, align=64) -> Table:
    """Keys of BlockMatrix<2>(B1, B2) (src/BlockTensor.hpp:285-301)."""
    q, p, r, c = _outkeys(lib().mpsa_twosite_outkeys, [C.byref(t1.s), C.byref(t2.s)], 2)
    return Table.from_keys(2, q, p, r, c, align)


def twosite_product(t1: Table, d1, t2: Table, d2, tout: Table, dout):
    """dout = BlockMatrix<2>(B1, B2) on the device (mpsa_twosite_product)."""
    _check(lib().mpsa_twosite_product(C.byref(t1.s), _dptr(d1), C.byref(t2.s), _dptr(d2), C.byref(tout.s),
                                      _dptr(dout), _stream()))


def heff1_outkeys(Lt: Table, Rt: Table, xt: Table, W, itk, align=64) -> Table:
    a = _itk(itk)
    m = MPOArg(W)
    q, p, r, c = _outkeys(lib().mpsa_heff1_outkeys,
                          [C.byref(Lt.s), C.byref(Rt.s), C.byref(xt.s), C.byref(m.s), _ptr(a), len(a)], 1)
    return Table.from_keys(1, q, p, r, c, align)


def env_outkeys(BLt: Table, At: Table, W, itk, direction, Apt: Table | None = None, align=64) -> Table:
    a = _itk(itk)
    m = MPOArg(W)
    q, p, r, c = _outkeys(lib().mpsa_env_outkeys,
                          [C.byref(BLt.s), C.byref(At.s), C.byref(Apt.s) if Apt is not None else None,
                           C.byref(m.s), _ptr(a), len(a), int(direction)], 1)
    return Table.from_keys(1, q, p, r, c, align)


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class HeffPlan:
    """Grouped-GEMM plan of y = H_eff x (two-site or one-site)."""

    def __init__(self, Lt: Table, Rt: Table, xt: Table, yt: Table, W, itk, precision=F64):
        self.Lt, self.Rt, self.xt, self.yt = Lt, Rt, xt, yt
        self.h = C.c_void_p()
        self._L = lib()
        self.precision = PREC.get(precision, precision) if isinstance(precision, str) else precision
        if xt.ns == 2:
            Wl, Wr = W
            itl, itr = itk
            a, b = _itk(itl), _itk(itr)
            ml, mr = MPOArg(Wl), MPOArg(Wr)
            _check(lib().mpsa_heff2_plan(C.byref(self.h), C.byref(Lt.s), C.byref(Rt.s), C.byref(xt.s),
                                         C.byref(yt.s), C.byref(ml.s), C.byref(mr.s), _ptr(a), len(a), _ptr(b),
                                         len(b), self.precision))
        else:
            a = _itk(itk)
            m = MPOArg(W)
            _check(lib().mpsa_heff1_plan(C.byref(self.h), C.byref(Lt.s), C.byref(Rt.s), C.byref(xt.s),
                                         C.byref(yt.s), C.byref(m.s), _ptr(a), len(a), self.precision))
        self._env = None

    def __del__(self):
        # destroyed by the library that created it
        L = getattr(self, "_L", None)
        if getattr(self, "h", None) and self.h.value and L is not None:
            L.mpsa_plan_destroy(self.h)
            self.h = None

    def info(self):
        fa, fr = np.zeros(1), np.zeros(1)
        g1, g2, ws = np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64)
        _check(lib().mpsa_plan_info(self.h, _ptr(fa), _ptr(fr), _ptr(g1), _ptr(g2), _ptr(ws)))
        return dict(flops_alg=float(fa[0]), flops_ref=float(fr[0]), n_gemm1=int(g1[0]), n_gemm2=int(g2[0]),
                    workspace_bytes=int(ws[0]))

    def stage_info(self):
        f1, f2 = np.zeros(1), np.zeros(1)
        c1, c2 = np.zeros(1, np.int32), np.zeros(1, np.int32)
        _check(lib().mpsa_plan_stage_info(self.h, _ptr(f1), _ptr(f2), _ptr(c1), _ptr(c2)))
        return dict(flops1=float(f1[0]), flops2=float(f2[0]), tile_class1=int(c1[0]), tile_class2=int(c2[0]))

    def describe(self):
        """The plan's device schedule as numpy record arrays (see mpsa_plan_describe)."""
        return describe(self.h)

    def set_timing(self, on=True):
        _check(lib().mpsa_plan_set_timing(self.h, int(on)))

    def set_tile_class(self, stage, tile_class):
        """Launch GEMM stage 1 or 2 with the given tile class from now on."""
        _check(lib().mpsa_plan_set_tile_class(self.h, int(stage), int(tile_class)))

    def autotune(self, nvec=1, reps=3):
        """Time every admissible tile class of both stages (environments bound)
        and keep the fastest; returns {(stage, class): ms per launch}."""
        ms = np.zeros(2 * TILE_CLASSES)
        _check(lib().mpsa_plan_autotune(self.h, int(nvec), int(reps), _ptr(ms), _stream()))
        return {(st + 1, c): float(ms[st * TILE_CLASSES + c]) for st in range(2) for c in range(TILE_CLASSES)
                if ms[st * TILE_CLASSES + c] >= 0}

    def last_times(self):
        """ms of [convert x, stage-1 GEMM, mix + zero fill, stage-2 GEMM] of the last apply."""
        ms = np.zeros(4, np.float32)
        _check(lib().mpsa_plan_last_times(self.h, _ptr(ms)))
        return [float(x) for x in ms]

    def timing_summary(self):
        """Average [convert, stage 1, mix + zero, stage 2] ms over the applies
        timed since timing was enabled (HIP events on the apply stream)."""
        ms = np.zeros(4, np.float32)
        n = np.zeros(1, np.int32)
        _check(lib().mpsa_plan_timing_summary(self.h, _ptr(ms), _ptr(n)))
        return [float(x) for x in ms], int(n[0])

    def bind_env(self, dL, dR):
        self._env = (dL, dR)
        _check(lib().mpsa_plan_bind_env(self.h, _dptr(dL), _dptr(dR), _stream()))

    def apply(self, dx, dy, nvec=1, x_stride=0, y_stride=0):
        _check(lib().mpsa_heff_apply(self.h, _dptr(dx), int(x_stride), _dptr(dy), int(y_stride), int(nvec),
                                     _stream()))


_REC = {
    0: np.dtype([("c_off", "<i8"), ("m", "<i4"), ("n", "<i4"), ("seg0", "<i4"), ("nseg", "<i4"),
                 ("tile0", "<i4"), ("tiles_m", "<i4")]),
    1: np.dtype([("a_off", "<i8"), ("b_off", "<i8"), ("k", "<i4"), ("lda", "<i4"), ("ldb", "<i4"),
                 ("pad", "<i4"), ("alpha", "<f8")]),
    4: np.dtype([("dst_off", "<i8"), ("len", "<i8"), ("t0", "<i4"), ("nt", "<i4"), ("chunk0", "<i4"),
                 ("pad", "<i4")]),
    5: np.dtype([("src_off", "<i8"), ("coef", "<f8")]),
    6: np.dtype([("off", "<i8"), ("len", "<i8")]),
}
_REC[2], _REC[3] = _REC[0], _REC[1]


def describe(h):
    out = {}
    names = ["s1_tasks", "s1_segs", "s2_tasks", "s2_segs", "mix", "terms", "zero"]
    for w in range(7):
        n = np.zeros(1, np.int64)
        ex = np.zeros(1, np.int32)
        _check(lib().mpsa_plan_describe(h, w, None, 0, _ptr(n), _ptr(ex)))
        arr = np.zeros(int(n[0]), dtype=_REC[w])
        if n[0]:
            _check(lib().mpsa_plan_describe(h, w, _ptr(arr), arr.nbytes, _ptr(n), _ptr(ex)))
        out[names[w]] = arr
        out[names[w] + "_extra"] = int(ex[0])
    return out


class EnvPlan:
    """Plan of one environment update (addContractionLayer)."""

    def __init__(self, BLt: Table, At: Table, W, itk, direction, outt: Table, Apt: Table | None = None):
        self.BLt, self.At, self.Apt, self.outt = BLt, At, Apt, outt
        self.h = C.c_void_p()
        self._L = lib()
        a = _itk(itk)
        m = MPOArg(W)
        _check(lib().mpsa_env_plan(C.byref(self.h), C.byref(BLt.s), C.byref(At.s),
                                   C.byref(Apt.s) if Apt is not None else None, C.byref(m.s), _ptr(a), len(a),
                                   int(direction), C.byref(outt.s), F64))

    def __del__(self):
        # destroyed by the library that created it
        L = getattr(self, "_L", None)
        if getattr(self, "h", None) and self.h.value and L is not None:
            L.mpsa_plan_destroy(self.h)
            self.h = None

    def info(self):
        fa, fr = np.zeros(1), np.zeros(1)
        g1, g2, ws = np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64)
        _check(lib().mpsa_plan_info(self.h, _ptr(fa), _ptr(fr), _ptr(g1), _ptr(g2), _ptr(ws)))
        return dict(flops_alg=float(fa[0]), flops_ref=float(fr[0]), n_gemm1=int(g1[0]), n_gemm2=int(g2[0]),
                    workspace_bytes=int(ws[0]))

    def describe(self):
        return describe(self.h)

    def apply(self, dBL, dA, dout, dAp=None):
        _check(lib().mpsa_env_apply(self.h, _dptr(dBL), _dptr(dA), _dptr(dAp), _dptr(dout), _stream()))


def lanczos(plan, dpsi0, itmax, delta, testenergy=True, target=0, want_state=True, proj=None, nproj=0,
            proj_stride=0, reortho=False):
    """TSLanzcosEngine::simulate + getState on the device (src/lanzcos.hpp:165-278);
    with nproj > 0 the projected run against the states in `proj` (state p at
    offset p * proj_stride, src/lanzcos.hpp:281-414); reortho adds full FP64
    reorthogonalisation."""
    import torch
    cap = itmax + 2
    E = np.zeros(cap)
    nit = np.zeros(1, np.int32)
    nr = np.zeros(1, np.int32)
    state = torch.zeros_like(dpsi0) if want_state else None
    _check(lib().mpsa_lanczos_ex(plan.h, _dptr(dpsi0), _dptr(state), int(itmax), float(delta), int(testenergy),
                                 int(target), _ptr(E), cap, _ptr(nit), _ptr(nr), _dptr(proj), int(nproj),
                                 int(proj_stride), int(bool(reortho)), _stream()))
    return dict(it=int(nit[0]), energies=E[:int(nr[0])].copy(), state=state)


def chebyshev_moments(plan, dt0, nmom, scale, shift=0.0, recons=True):
    """Chebyshev moments mu_k = <t_0|T_k((H_eff - shift)/scale) t_0> on one bond
    (the recurrence of ChebychevEngine::simulateSSCompress,
    src/container.hpp:2343-2470, with the exact H_eff).  Returns (moments,
    reconstructed doubled moments or None)."""
    mu = np.zeros(nmom)
    rc = np.zeros(2 * nmom - 1) if recons else None
    _check(lib().mpsa_chebyshev_moments(plan.h, _dptr(dt0), int(nmom), float(scale), float(shift), _ptr(mu),
                                        _ptr(rc) if recons else None, _stream()))
    return mu, rc


def lanczos_batched(plan, dpsi0, nvec, stride, itmax, delta, target=0, want_state=True, graph=True, proj=None,
                    nproj=0, proj_stride=0, proj_vstride=0, reortho=False):
    """nvec independent Lanczos runs (energy test), one batched matvec per step;
    with graph=True each block of 5 steps is replayed from a hipGraph.  With
    nproj > 0 every run is projected against the states in `proj` (state p at
    offset p * proj_stride, per-vector sets at v * proj_vstride when nonzero;
    the reference's projected run, src/lanzcos.hpp:281-414); reortho adds full
    FP64 reorthogonalisation against each run's Krylov basis."""
    import torch
    cap = itmax + 2
    E = np.zeros(nvec * cap)
    nit = np.zeros(nvec, np.int32)
    nr = np.zeros(nvec, np.int32)
    state = torch.zeros(nvec * stride, dtype=torch.float64, device=dpsi0.device) if want_state else None
    _check(lib().mpsa_lanczos_batched_ex(plan.h, _dptr(dpsi0), int(stride), int(nvec), _dptr(state), int(itmax),
                                         float(delta), int(target), _ptr(E), cap, _ptr(nit), _ptr(nr),
                                         int(bool(graph)), _dptr(proj), int(nproj), int(proj_stride),
                                         int(proj_vstride), int(bool(reortho)), _stream()))
    E = E.reshape(nvec, cap)
    return dict(it=nit.astype(int), energies=[E[v, :int(nr[v])].copy() for v in range(nvec)], state=state)


class Chain:
    """A whole MPS chain under an MPO: energy (MPO::measure) and its adjoint
    (dE/dA[i] for every site, dE/dtheta for the given derivative MPOs)."""

    def __init__(self, site_tables, mpo, itk, dmpo=(), checkpoint=8):
        self.N = len(site_tables)
        self.tables = list(site_tables)
        self._mpo = [MPOArg(W) for W in mpo]
        self._dmpo = [MPOArg(W) for dW in dmpo for W in dW]
        self.nparam = len(dmpo)
        sites = (_Blocks * self.N)(*[t.s for t in self.tables])
        mpos = (_MPO * self.N)(*[m.s for m in self._mpo])
        dm = (_MPO * max(1, len(self._dmpo)))(*[m.s for m in self._dmpo])
        self._keep = (sites, mpos, dm)
        self._itk = np.ascontiguousarray(np.concatenate([_itk(x) for x in itk]), dtype=np.uint64)
        self._d = np.array([len(x) for x in itk], dtype=np.int32)
        self._L = lib()
        h = C.c_void_p()
        _check(lib().mpsa_chain_create(C.byref(h), self.N, C.cast(sites, C.c_void_p), C.cast(mpos, C.c_void_p),
                                       _ptr(self._itk), _ptr(self._d), self.nparam,
                                       C.cast(dm, C.c_void_p) if self._dmpo else None, int(checkpoint)))
        self.h = h

    def __del__(self):
        L = getattr(self, "_L", None)
        if getattr(self, "h", None) and L is not None:
            L.mpsa_chain_destroy(self.h)
            self.h = None

    def _ptrs(self, bufs):
        return (C.c_void_p * self.N)(*[_dptr(b) for b in bufs])

    def forward(self, dA):
        E = np.zeros(1)
        arr = self._ptrs(dA)
        _check(lib().mpsa_energy_forward(self.h, C.cast(arr, C.c_void_p), _ptr(E), _stream()))
        return float(E[0])

    def backward(self, dA):
        import torch
        grads = [torch.zeros_like(b) for b in dA]
        dth = np.zeros(max(1, self.nparam))
        a, g = self._ptrs(dA), self._ptrs(grads)
        _check(lib().mpsa_energy_backward(self.h, C.cast(a, C.c_void_p), C.cast(g, C.c_void_p), _ptr(dth),
                                          _stream()))
        return grads, dth[:self.nparam].copy()

    def env_table(self, side, index):
        n = C.c_int64()
        _check(lib().mpsa_chain_env_table(self.h, side, index, 0, C.byref(n), None, None, None, None, None))
        k = n.value
        q, p = np.zeros(2 * k, np.uint64), np.zeros(k, np.uint32)
        r, c, o = np.zeros(k, np.int64), np.zeros(k, np.int64), np.zeros(k, np.int64)
        _check(lib().mpsa_chain_env_table(self.h, side, index, k, C.byref(n), _ptr(q), _ptr(p), _ptr(r), _ptr(c),
                                          _ptr(o)))
        return Table(1, q.reshape(-1, 2), p.reshape(-1, 1), r, c, o)


def split(psi_t: Table, dpsi, itk_l, itk_r, direction=1, chimax=0, do_trunc=False, tr_weight=0.0, normalize=True):
    """split of a two-site tensor on the device (src/MPSfunctions.hpp:1142-1278).
    Returns dict(left=(Table, device tensor), right=(Table, device tensor),
    lam={qc: values}, tw, chi); lambda is multiplied into the right tensor for
    direction > 0, into the left one otherwise."""
    import torch
    L = lib()
    itl, itr = _itk(itk_l), _itk(itk_r)
    h = C.c_void_p()
    _check(L.mpsa_split(C.byref(psi_t.s), _dptr(dpsi), _ptr(itl), len(itl), _ptr(itr), len(itr), int(direction),
                        int(chimax), int(bool(do_trunc)), float(tr_weight), int(bool(normalize)), C.byref(h),
                        _stream()))
    try:
        tw, chi = np.zeros(1), np.zeros(1, np.int32)
        nl, nr, nlam, ll, rl = (np.zeros(1, np.int64) for _ in range(5))
        _check(L.mpsa_split_info(h, _ptr(tw), _ptr(chi), _ptr(nl), _ptr(nr), _ptr(nlam), _ptr(ll), _ptr(rl)))
        out = dict(tw=float(tw[0]), chi=int(chi[0]))
        for which, n, ln, name in ((0, int(nl[0]), int(ll[0]), "left"), (1, int(nr[0]), int(rl[0]), "right")):
            q, p = np.zeros(2 * n, np.uint64), np.zeros(n, np.uint32)
            r, c, o = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n, np.int64)
            if n:
                _check(L.mpsa_split_table(h, which, _ptr(q), _ptr(p), _ptr(r), _ptr(c), _ptr(o)))
            t = Table(1, q.reshape(-1, 2), p.reshape(-1, 1), r, c, o)
            buf = torch.zeros(max(ln, 1), dtype=torch.float64, device=dpsi.device)
            if ln:
                _check(L.mpsa_split_copy(h, which, _dptr(buf), _stream()))
            out[name] = (t, buf)
        k = int(nlam[0])
        lq, ls = np.zeros(k, np.uint64), np.zeros(k, np.int64)
        vals = np.zeros(max(1, out["chi"]))
        if k:
            _check(L.mpsa_split_lambda(h, _ptr(lq), _ptr(ls), _ptr(vals)))
        lam, o = {}, 0
        for i in range(k):
            lam[int(lq[i])] = vals[o:o + int(ls[i])].copy()
            o += int(ls[i])
        out["lam"] = lam
        return out
    finally:
        L.mpsa_split_destroy(h)


class Dmrg:
    """Device-resident two-site DMRG (DMRGengine::doStep / simulate sweeps).
    site_tables/dA: a right-normalised MPS (ns = 1 tables, device FP64)."""

    def __init__(self, site_tables, dA, mpo, itk):
        self.N = len(site_tables)
        self._tabs = list(site_tables)
        self._mpo = [MPOArg(W) for W in mpo]
        sites = (_Blocks * self.N)(*[t.s for t in self._tabs])
        mpos = (_MPO * self.N)(*[m.s for m in self._mpo])
        ptrs = (C.c_void_p * self.N)(*[_dptr(b) for b in dA])
        itk_all = np.ascontiguousarray(np.concatenate([_itk(x) for x in itk]), dtype=np.uint64)
        d = np.array([len(x) for x in itk], dtype=np.int32)
        self._L = lib()
        h = C.c_void_p()
        _check(lib().mpsa_dmrg_create(C.byref(h), self.N, C.cast(sites, C.c_void_p), C.cast(ptrs, C.c_void_p),
                                      C.cast(mpos, C.c_void_p), _ptr(itk_all), _ptr(d), _stream()))
        self.h = h

    def __del__(self):
        L = getattr(self, "_L", None)
        if getattr(self, "h", None) and L is not None:
            L.mpsa_dmrg_destroy(self.h)
            self.h = None

    def sweep(self, chimax, do_trunc=False, tr_weight=0.0, lanmax=500, delta=1e-12, precision=F64, refine_it=10):
        E, tw, sec = np.zeros(1), np.zeros(1), np.zeros(1)
        it = np.zeros(1, np.int32)
        _check(lib().mpsa_dmrg_sweep(self.h, int(precision), int(refine_it), int(chimax), int(bool(do_trunc)),
                                     float(tr_weight), int(lanmax), float(delta), _ptr(E), _ptr(tw), _ptr(it),
                                     _ptr(sec)))
        return dict(E=float(E[0]), tw=float(tw[0]), it=int(it[0]), s=float(sec[0]))

    def site(self, i):
        """(Table, device tensor) of site tensor i."""
        import torch
        nb, ln = np.zeros(1, np.int64), np.zeros(1, np.int64)
        _check(lib().mpsa_dmrg_site(self.h, int(i), _ptr(nb), _ptr(ln), None, None, None, None, None, None, None))
        n = int(nb[0])
        q, p = np.zeros(2 * n, np.uint64), np.zeros(n, np.uint32)
        r, c, o = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n, np.int64)
        buf = torch.zeros(max(1, int(ln[0])), dtype=torch.float64, device="cuda")
        _check(lib().mpsa_dmrg_site(self.h, int(i), _ptr(nb), _ptr(ln), _ptr(q), _ptr(p), _ptr(r), _ptr(c), _ptr(o),
                                    _dptr(buf), _stream()))
        return Table(1, q.reshape(-1, 2), p.reshape(-1, 1), r, c, o), buf


class Cheb:
    """Chebyshev moments with variational MPS compression on the device
    (ChebychevEngine::simulateSSCompress, src/container.hpp:2343-2470, and
    applyMPO / compressSingleSite / compress / computeOverlap of
    src/MPSfunctions.hpp) for nvec start vectors in one set of grouped
    launches.  `states`: per vector a list of site Blocks (ns = 1; all real or
    all complex); `alloc(np_array)` returns a device buffer with data_ptr()
    (default: a torch tensor on the GPU) and `to_host(buf)` reads one back."""

    def __init__(self, states, mpo, itk, h1_precision=F64, alloc=None, to_host=None):
        import ctypes as C
        if alloc is None:
            import torch
            alloc = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
            to_host = lambda b: b.cpu().numpy()  # noqa: E731
        self._alloc, self._to_host = alloc, to_host
        self.nvec, self.N = len(states), len(states[0])
        self.cplx = any(B.dtype.kind == "c" for st in states for B in st)
        tabs, bufs = [], []
        for st in states:
            for B in st:
                if self.cplx and B.dtype.kind != "c":
                    B = Blocks.from_table(B.ns, *[B.table()[k] for k in ("q", "p", "rows", "cols")],
                                          B.pack().astype(np.complex128), np.complex128)
                t = Table.from_blocks(B)
                tabs.append(t)
                bufs.append(alloc(t.pack(B)))
        self._tabs, self._bufs = tabs, bufs
        self._mpo = [MPOArg(W) for W in mpo]
        self._itk = [list(x) for x in itk]
        sites = (_Blocks * len(tabs))(*[t.s for t in tabs])
        mpos = (_MPO * self.N)(*[m.s for m in self._mpo])
        ptrs = (C.c_void_p * len(bufs))(*[_dptr(b) for b in bufs])
        itk_all = np.ascontiguousarray(np.concatenate([_itk(x) for x in itk]), dtype=np.uint64)
        d = np.array([len(x) for x in itk], dtype=np.int32)
        self._L = lib()
        h = C.c_void_p()
        _check(lib().mpsa_cheb_create(C.byref(h), self.N, self.nvec, C.cast(sites, C.c_void_p),
                                      C.cast(ptrs, C.c_void_p), C.cast(mpos, C.c_void_p), _ptr(itk_all), _ptr(d),
                                      int(h1_precision), _stream()))
        self.h = h

    def __del__(self):
        L = getattr(self, "_L", None)
        if getattr(self, "h", None) and L is not None:
            L.mpsa_cheb_destroy(self.h)
            self.h = None

    def _cv(self, a):
        return a[0::2] + 1j * a[1::2] if self.cplx else a[0::2]

    def copy(self, dst, src):
        _check(lib().mpsa_cheb_copy(self.h, dst, src))

    def apply_mpo(self, dst, src):
        _check(lib().mpsa_cheb_apply_mpo(self.h, dst, src))

    def orthonormalize(self, slot=0):
        """orthonormalize(mps, "rl") on the device (right-normalised, unit norm)."""
        _check(lib().mpsa_cheb_orthonormalize(self.h, slot))

    def overlap(self, a, b):
        """computeOverlap(a, b) = <b|a> per vector."""
        out = np.zeros(2 * self.nvec)
        _check(lib().mpsa_cheb_overlap(self.h, a, b, _ptr(out)))
        return self._cv(out)

    def axpby(self, dst, xa, a, xb, b):
        _check(lib().mpsa_cheb_axpby(self.h, dst, float(xa), a, float(xb), b))

    def compress_single_site(self, dst, src, chimax, maxstep=5, delta=1e-10):
        conv = np.zeros(self.nvec)
        _check(lib().mpsa_cheb_compress_single_site(self.h, dst, src, int(chimax), int(maxstep), float(delta),
                                                    _ptr(conv)))
        return conv

    def compress(self, slot, chimax, maxstep=5, delta=1e-10):
        conv = np.zeros(self.nvec)
        _check(lib().mpsa_cheb_compress(self.h, slot, int(chimax), int(maxstep), float(delta), _ptr(conv)))
        return conv

    def moments(self, nmom, chimax, ncmax=5, delta=1e-10):
        mu = np.zeros(2 * self.nvec * nmom)
        rc = np.zeros(2 * self.nvec * (2 * nmom - 1))
        c1, c2, sec = np.zeros(self.nvec * nmom), np.zeros(self.nvec * nmom), np.zeros(1)
        _check(lib().mpsa_cheb_moments(self.h, int(nmom), int(chimax), int(ncmax), float(delta), _ptr(mu), _ptr(rc),
                                       _ptr(c1), _ptr(c2), _ptr(sec)))
        lap, n = np.zeros(nmom), np.zeros(1, np.int32)
        _check(lib().mpsa_cheb_moment_times(self.h, _ptr(lap), nmom, _ptr(n)))
        return dict(moments=self._cv(mu).reshape(self.nvec, nmom), recons=self._cv(rc).reshape(self.nvec, -1),
                    conv1=c1.reshape(self.nvec, nmom), conv2=c2.reshape(self.nvec, nmom), s=float(sec[0]),
                    moment_s=lap[:int(n[0])].copy())

    def set_big_sector(self, k):
        """Sectors with min(m, n) >= k go to rocSOLVER (default 1024)."""
        _check(lib().mpsa_cheb_set_big_sector(self.h, int(k)))

    def timers(self):
        """Host-side seconds and counts of the engine since its creation."""
        out = (C.c_double * 9)()
        _check(lib().mpsa_cheb_timers(self.h, out, 9))
        k = ("heff_s", "env_s", "prepare_s", "sync_s", "alloc_s", "allocs", "plan_s", "syncs", "plans")
        return {n: (int(v) if n in ("allocs", "syncs", "plans") else round(v, 4)) for n, v in zip(k, out)}

    def set_block_jacobi(self, r_min):
        """Sectors with min(m, n) >= r_min and max(m, n) <= 1024 are
        factorised by several workgroups each (Householder panels, Q
        formation, and for SVD steps with min(m, n) <= 512 the block Jacobi
        rounds); 0: one workgroup per sector; default 32."""
        _check(lib().mpsa_cheb_set_block_jacobi(self.h, int(r_min)))

    def factor_log(self, cap):
        """Record the sector shapes of the next cap canonicalisation launches
        (0 stops and clears)."""
        _check(lib().mpsa_cheb_factor_log(self.h, int(cap)))

    def factor_records(self):
        """Logged launches as dicts: kind ('svd', 'qr', 'svd_rocsolver',
        'qr_rocsolver'), sectors, max_m, max_n, max_k, k64, k128, k256."""
        n = np.zeros(1, np.int64)
        _check(lib().mpsa_cheb_factor_records(self.h, None, 0, _ptr(n)))
        out = np.zeros(8 * max(int(n[0]), 1), np.int64)
        _check(lib().mpsa_cheb_factor_records(self.h, _ptr(out), int(n[0]), _ptr(n)))
        kinds = ("svd", "qr", "svd_rocsolver", "qr_rocsolver")
        names = ("sectors", "max_m", "max_n", "max_k", "k64", "k128", "k256")
        return [dict(kind=kinds[int(r[0])], **{k: int(v) for k, v in zip(names, r[1:])})
                for r in out[:8 * int(n[0])].reshape(-1, 8)]

    def site(self, slot, vec, site):
        nb, ln = np.zeros(1, np.int64), np.zeros(1, np.int64)
        _check(lib().mpsa_cheb_site(self.h, slot, vec, site, _ptr(nb), _ptr(ln), None, None, None, None, None, None))
        n = int(nb[0])
        q, p = np.zeros(2 * max(n, 1), np.uint64), np.zeros(max(n, 1), np.uint32)
        r, c, o = np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1), np.int64)
        buf = self._alloc(np.zeros(max(1, int(ln[0]))))
        _check(lib().mpsa_cheb_site(self.h, slot, vec, site, _ptr(nb), _ptr(ln), _ptr(q), _ptr(p), _ptr(r), _ptr(c),
                                    _ptr(o), _dptr(buf)))
        t = Table(1, q[:2 * n].reshape(-1, 2), p[:n].reshape(-1, 1), r[:n], c[:n], o[:n], int(self.cplx))
        return t.unpack(self._to_host(buf))

    def state(self, slot, vec):
        return [self.site(slot, vec, i) for i in range(self.N)]

    def stats(self):
        a, b = np.zeros(1, np.int64), np.zeros(1, np.int64)
        _check(lib().mpsa_cheb_stats(self.h, _ptr(a), _ptr(b)))
        t = np.zeros(5)
        _check(lib().mpsa_cheb_profile(self.h, _ptr(t)))
        return dict(plans_built=int(a[0]), pool_bytes=int(b[0]), host_plan_s=float(t[0]), host_keys_s=float(t[1]),
                    host_wait_s=float(t[2]), waits=int(t[3]), plans_cached=int(t[4]))


def fill_gauss(buf, seed, n=None):
    """Seeded N(0, 1) entries on the device (mpsa_fill_gauss)."""
    n = buf.numel() if n is None else n
    _check(lib().mpsa_fill_gauss(_dptr(buf), int(n), int(seed) & ((1 << 64) - 1), _stream()))


def random_mps_tables(L, itk, Ntot, D, cplx=False):
    """Block tables of a seeded random MPS with the sector profile of
    synth.bond_dims (no data): one Table per site."""
    from .synth import bond_dims
    from .blocks import qadd
    dims = bond_dims(L, itk, Ntot, D)
    tabs = []
    for i in range(L):
        q, p, r, c = [], [], [], []
        for a in sorted(dims[i]):
            for s, qs in enumerate(itk[i]):
                b = qadd(a, qs)
                if b in dims[i + 1]:
                    q.append((a, b))
                    p.append(s)
                    r.append(dims[i][a])
                    c.append(dims[i + 1][b])
        order = sorted(range(len(q)), key=lambda k: (q[k][0], q[k][1], p[k]))
        t = Table.from_keys(1, np.array([q[k] for k in order], np.uint64).reshape(-1, 2),
                            np.array([p[k] for k in order], np.uint32).reshape(-1, 1),
                            [r[k] for k in order], [c[k] for k in order])
        tabs.append(t.as_complex() if cplx else t)
    return tabs


class ChebRandom(Cheb):
    """A Cheb engine whose nvec start vectors are seeded random MPS generated
    on the device: block tables from the sector profile, entries from
    mpsa_fill_gauss (seed + 1000 * vector + site; both planes for complex),
    then orthonormalize("rl") on the device."""

    def __init__(self, L, itk, Ntot, D, nvec, mpo, seed, cplx=False, h1_precision=F64):
        import ctypes as C
        import torch
        self._alloc = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        self._to_host = lambda b: b.cpu().numpy()  # noqa: E731
        self.nvec, self.N, self.cplx = nvec, L, bool(cplx)
        base = random_mps_tables(L, itk, Ntot, D, cplx)
        tabs, bufs = [], []
        for v in range(nvec):
            for i, t in enumerate(base):
                b = torch.empty(max(1, t.length), dtype=torch.float64, device="cuda")
                fill_gauss(b, seed + 1000 * v + i, t.length)
                tabs.append(t)
                bufs.append(b)
        self._tabs, self._bufs = tabs, bufs
        self._mpo = [MPOArg(W) for W in mpo]
        sites = (_Blocks * len(tabs))(*[t.s for t in tabs])
        mpos = (_MPO * L)(*[m.s for m in self._mpo])
        ptrs = (C.c_void_p * len(bufs))(*[_dptr(b) for b in bufs])
        itk_all = np.ascontiguousarray(np.concatenate([_itk(x) for x in itk]), dtype=np.uint64)
        d = np.array([len(x) for x in itk], dtype=np.int32)
        self._L = lib()
        h = C.c_void_p()
        _check(lib().mpsa_cheb_create(C.byref(h), L, nvec, C.cast(sites, C.c_void_p), C.cast(ptrs, C.c_void_p),
                                      C.cast(mpos, C.c_void_p), _ptr(itk_all), _ptr(d), int(h1_precision), _stream()))
        self.h = h
        self._bufs = None  # the engine holds its own copies
        self.orthonormalize(0)


def tridiag_eig(d, e, vectors=True):
    """Host FP64 tridiagonal eigensolver of the library (replaces dstev)."""
    n = len(d)
    dd = np.ascontiguousarray(np.array(d, dtype=np.float64))
    ee = np.ascontiguousarray(np.array(e, dtype=np.float64))
    z = np.zeros(n * n) if vectors else None
    _check(lib().mpsa_tridiag_eig(n, _ptr(dd), _ptr(ee) if n > 1 else None, _ptr(z) if vectors else None))
    return dd, (z.reshape((n, n), order="F") if vectors else None)


def dot(dx, dy, n=None, cplx=False):
    """<y|x> on device (fixed-order FP64 reduction)."""
    out = np.zeros(2)
    n = dx.numel() if n is None else n
    _check(lib().mpsa_dot(_dptr(dx), _dptr(dy), int(n), int(cplx), _ptr(out), _stream()))
    return complex(out[0], out[1]) if cplx else float(out[0])


def axpby(a, dx, b, dy, n=None):
    """y = a x + b y on device (real)."""
    av = np.array([float(a), 0.0])
    bv = np.array([float(b), 0.0])
    n = dx.numel() if n is None else n
    _check(lib().mpsa_axpby(_ptr(av), _dptr(dx), _ptr(bv), _dptr(dy), int(n), 0, _stream()))


# ---------------------------------------------------------------- drop-in calls
def _to_dev(t: Table, B: Blocks):
    import torch
    return torch.from_numpy(t.pack(B)).cuda()


def _promote(*bs):
    """Complex128 copies of all tables when any of them is complex (a complex
    plan takes complex tables only), else the tables themselves."""
    if not any(b is not None and b.dtype.kind == "c" for b in bs):
        return bs
    out = []
    for b in bs:
        if b is None or b.dtype.kind == "c":
            out.append(b)
            continue
        z = Blocks(b.ns, np.complex128)
        for k, v in b.b.items():
            z[k] = v
        out.append(z)
    return tuple(out)


def _out_table(t: Table, cplx: bool) -> Table:
    return t.as_complex() if cplx else t


def contract_twosite(L: Blocks, R: Blocks, psi: Blocks, Wl, Wr, itl, itr, precision=F64) -> Blocks:
    """contractTwosite (src/MPSfunctions.hpp:949-981): returns H_eff psi with the
    reference's output key set.  Complex128 blocks (with the real MPO) run a
    complex plan."""
    import torch
    L, R, psi = _promote(L, R, psi)
    Lt, Rt, xt = Table.from_blocks(L), Table.from_blocks(R), Table.from_blocks(psi)
    yt = _out_table(heff2_outkeys(Lt, Rt, xt, Wl, Wr, itl, itr), xt.is_complex)
    plan = HeffPlan(Lt, Rt, xt, yt, (Wl, Wr), (itl, itr), precision)
    dL, dR, dx = _to_dev(Lt, L), _to_dev(Rt, R), _to_dev(xt, psi)
    dy = torch.zeros(max(yt.length, 1), dtype=torch.float64, device="cuda")
    plan.bind_env(dL, dR)
    plan.apply(dx, dy)
    torch.cuda.synchronize()
    return yt.unpack(dy.cpu().numpy())


def twosite(B1: Blocks, B2: Blocks) -> Blocks:
    """BlockMatrix<2>(B1, B2) (src/BlockTensor.hpp:285-301) computed on the
    device; complex when either input is complex."""
    import torch
    B1, B2 = _promote(B1, B2)
    t1, t2 = Table.from_blocks(B1), Table.from_blocks(B2)
    tout = _out_table(twosite_outkeys(t1, t2), t1.is_complex)
    d1, d2 = _to_dev(t1, B1), _to_dev(t2, B2)
    dout = torch.zeros(max(tout.length, 1), dtype=torch.float64, device="cuda")
    twosite_product(t1, d1, t2, d2, tout, dout)
    torch.cuda.synchronize()
    return tout.unpack(dout.cpu().numpy())


def contract_single(L: Blocks, A: Blocks, W, R: Blocks, itk, precision=F64) -> Blocks:
    """contractSingle (src/MPSfunctions.hpp:857-885)."""
    import torch
    L, A, R = _promote(L, A, R)
    Lt, Rt, xt = Table.from_blocks(L), Table.from_blocks(R), Table.from_blocks(A)
    yt = _out_table(heff1_outkeys(Lt, Rt, xt, W, itk), xt.is_complex)
    plan = HeffPlan(Lt, Rt, xt, yt, W, itk, precision)
    dL, dR, dx = _to_dev(Lt, L), _to_dev(Rt, R), _to_dev(xt, A)
    dy = torch.zeros(max(yt.length, 1), dtype=torch.float64, device="cuda")
    plan.bind_env(dL, dR)
    plan.apply(dx, dy)
    torch.cuda.synchronize()
    return yt.unpack(dy.cpu().numpy())


def add_contraction_layer(BL: Blocks, A: Blocks, W, itk, direction, Ap: Blocks | None = None) -> Blocks:
    """addContractionLayer (src/MPSfunctions.hpp:442-555)."""
    import torch
    if len(A) == 0:
        raise MpsaError(1, "Error in addContractionLayer: A is empty.")
    BL, A, Ap = _promote(BL, A, Ap)
    BLt, At = Table.from_blocks(BL), Table.from_blocks(A)
    Apt = Table.from_blocks(Ap) if Ap is not None else None
    outt = _out_table(env_outkeys(BLt, At, W, itk, direction, Apt), At.is_complex)
    plan = EnvPlan(BLt, At, W, itk, direction, outt, Apt)
    dBL, dA = _to_dev(BLt, BL), _to_dev(At, A)
    dAp = _to_dev(Apt, Ap) if Ap is not None else None
    dout = torch.zeros(max(outt.length, 1), dtype=torch.float64, device="cuda")
    plan.apply(dBL, dA, dout, dAp)
    torch.cuda.synchronize()
    return outt.unpack(dout.cpu().numpy())
