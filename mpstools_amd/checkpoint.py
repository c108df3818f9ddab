"""Versioned checkpoint files of block-sparse tensors (mpsa_ckpt_* of the C
ABI), in place of the reference's Boost archives (MPSBase::save/load,
src/MPSTypes.hpp:377-426; Container::write/read, src/container.hpp:87-146).

    save(path, [A0, A1, ...], meta="...")        host Blocks (an MPS, envs)
    tensors, meta = load(path)                   back as host Blocks
    save_device(path, [(table, buf), ...], meta)  device tensors (torch)
    pairs, meta = load_device(path)              back as (Table, device buf)

All tensors of one file are real FP64 or all complex128.  A file that is
truncated, corrupt or of another format version raises MpsaError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .blocks import Blocks

_bound = False


def _lib():
    global _bound
    L = native.lib()
    if not _bound:
        P = C.c_void_p
        L.mpsa_ckpt_write.argtypes = [C.c_char_p, C.c_int, P, P, C.c_int, C.c_char_p, P]
        L.mpsa_ckpt_open.argtypes = [C.c_char_p, P, P, P, P]
        L.mpsa_ckpt_meta.argtypes = [P, P, C.c_int64]
        L.mpsa_ckpt_table.argtypes = [P, C.c_int, P, P, P, P, P, P, P, P]
        L.mpsa_ckpt_read.argtypes = [P, C.c_int, P, C.c_int, P]
        L.mpsa_ckpt_close.argtypes = [P]
        _bound = True
    return L


def _write(path, tables, ptrs, on_device, meta, stream):
    arr = (native._Blocks * len(tables))(*[t.s for t in tables])
    pp = (C.c_void_p * len(tables))(*ptrs)
    native._check(_lib().mpsa_ckpt_write(str(path).encode(), len(tables), C.cast(arr, C.c_void_p),
                                         C.cast(pp, C.c_void_p), int(on_device), meta.encode(), stream))


def save(path, tensors, meta=""):
    """Write host Blocks (ns = 1 or 2), all real or all complex128."""
    tabs = [native.Table.from_blocks(B) for B in tensors]
    bufs = [np.ascontiguousarray(t.pack(B)) for t, B in zip(tabs, tensors)]
    _write(path, tabs, [b.ctypes.data for b in bufs], 0, meta, None)


def save_device(path, pairs, meta=""):
    """Write (Table, device buffer) pairs."""
    _write(path, [t for t, _ in pairs], [b.data_ptr() for _, b in pairs], 1, meta, native._stream())


class _Open:
    def __init__(self, path):
        self.h = C.c_void_p()
        nt, cx, ml = np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1, np.int64)
        native._check(_lib().mpsa_ckpt_open(str(path).encode(), C.byref(self.h), native._ptr(nt), native._ptr(cx),
                                            native._ptr(ml)))
        self.n, self.cplx = int(nt[0]), int(cx[0])
        m = C.create_string_buffer(int(ml[0]) + 1)
        native._check(_lib().mpsa_ckpt_meta(self.h, m, int(ml[0])))
        self.meta = m.raw[:int(ml[0])].decode()

    def table(self, i):
        ns, nb, ln = np.zeros(1, np.int32), np.zeros(1, np.int64), np.zeros(1, np.int64)
        L = _lib()
        native._check(L.mpsa_ckpt_table(self.h, i, native._ptr(ns), native._ptr(nb), native._ptr(ln), None, None,
                                        None, None, None))
        n, s = int(nb[0]), int(ns[0])
        q, p = np.zeros((max(n, 1), 2), np.uint64), np.zeros((max(n, 1), s), np.uint32)
        r, c, o = (np.zeros(max(n, 1), np.int64) for _ in range(3))
        native._check(L.mpsa_ckpt_table(self.h, i, None, native._ptr(nb), None, native._ptr(q), native._ptr(p),
                                        native._ptr(r), native._ptr(c), native._ptr(o)))
        t = native.Table(s, q[:n], p[:n], r[:n], c[:n], o[:n], self.cplx)
        return t, int(ln[0])

    def close(self):
        if self.h:
            _lib().mpsa_ckpt_close(self.h)
            self.h = C.c_void_p()


def load(path):
    """(list of host Blocks, metadata text)."""
    f = _Open(path)
    try:
        out = []
        for i in range(f.n):
            t, ln = f.table(i)
            buf = np.zeros(max(ln, 1))
            native._check(_lib().mpsa_ckpt_read(f.h, i, native._ptr(buf), 0, None))
            out.append(t.unpack(buf[:max(ln, 0)] if ln else buf[:0]))
        return out, f.meta
    finally:
        f.close()


def load_device(path):
    """(list of (Table, device buffer), metadata text)."""
    import torch
    f = _Open(path)
    try:
        out = []
        for i in range(f.n):
            t, ln = f.table(i)
            d = torch.zeros(max(ln, 1), dtype=torch.float64, device="cuda")
            native._check(_lib().mpsa_ckpt_read(f.h, i, native._dptr(d), 1, native._stream()))
            out.append((t, d))
        return out, f.meta
    finally:
        f.close()


__all__ = ["save", "load", "save_device", "load_device", "Blocks"]
