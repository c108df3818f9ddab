"""Host-side block-sparse tensors with the MPSTools key semantics.

`Blocks` mirrors ``BlockMatrix<Ns,AbKeyType,T>`` (src/BlockTensor.hpp:279-379):
a map from a key ``(q_left, q_right, p_0, ..., p_{Ns-1})`` to a dense
column-major block.  Keys sort like ``TensorKeyType::operator<``
(src/keyTypes.hpp:313-343): charges first, then physical indices.

Charges are AbKeyType packed words (src/keyTypes.hpp:25-30): up to two U(1)
charges as ``(q2 << 16) + q1`` in an unsigned 64-bit word; ``+``/``-`` wrap
modulo 2**64 exactly as the reference's ``unsigned long`` arithmetic
(src/keyTypes.hpp:46-65), so an impossible charge becomes a key that no block
has.
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1


def pack_qn(q1: int, q2: int = 0) -> int:
    """AbKeyType(q1,q2,2) packing (src/keyTypes.hpp:25-30)."""
    return ((q2 << 16) + q1) & MASK64


def qadd(a: int, b: int) -> int:
    return (a + b) & MASK64


def qsub(a: int, b: int) -> int:
    return (a - b) & MASK64


class Blocks:
    """Ordered block map key -> 2-D numpy array (Fortran/column-major order)."""

    def __init__(self, ns: int, dtype=np.float64):
        self.ns = ns
        self.dtype = np.dtype(dtype)
        self.b: dict[tuple, np.ndarray] = {}

    # -- map interface ---------------------------------------------------
    def __setitem__(self, key, val):
        key = tuple(int(k) for k in key)
        assert len(key) == 2 + self.ns, key
        self.b[key] = np.asfortranarray(val, dtype=self.dtype)

    def __getitem__(self, key):
        return self.b[tuple(key)]

    def __contains__(self, key):
        return tuple(key) in self.b

    def __len__(self):
        return len(self.b)

    def keys(self):
        return sorted(self.b.keys())

    def items(self):
        return [(k, self.b[k]) for k in self.keys()]

    def copy(self):
        o = Blocks(self.ns, self.dtype)
        o.b = {k: v.copy(order="F") for k, v in self.b.items()}
        return o

    def nelem(self) -> int:
        return int(sum(v.size for v in self.b.values()))

    # -- algebra (BlockTensor.hpp:443-544) -------------------------------
    def dot(self, other: "Blocks"):
        """X.dot(Y) = sum conj(Y).X over matching keys (BlockTensor.hpp:520-544)."""
        s = 0.0
        for k, v in self.b.items():
            if k in other.b:
                s = s + np.vdot(other.b[k].ravel(order="F"), v.ravel(order="F"))
        return s

    def norm(self) -> float:
        return float(np.sqrt(abs(self.dot(self))))

    def axpy(self, a, other: "Blocks"):
        for k, v in other.b.items():
            if k in self.b:
                self.b[k] = self.b[k] + a * v
            else:
                self.b[k] = np.asfortranarray(a * v)
        return self

    def scaled(self, a):
        o = Blocks(self.ns, self.dtype)
        o.b = {k: np.asfortranarray(a * v) for k, v in self.b.items()}
        return o

    # -- flat tables ------------------------------------------------------
    def table(self, align: int = 1):
        """Flat key table in key order: q (n,2) uint64, p (n,ns) uint32,
        rows/cols/offsets (n,) int64 (element offsets; each block starts at a
        multiple of `align` elements), and the total element count."""
        keys = self.keys()
        n = len(keys)
        q = np.zeros((n, 2), dtype=np.uint64)
        p = np.zeros((n, self.ns), dtype=np.uint32)
        rows = np.zeros(n, dtype=np.int64)
        cols = np.zeros(n, dtype=np.int64)
        offs = np.zeros(n, dtype=np.int64)
        off = 0
        for i, k in enumerate(keys):
            q[i, 0], q[i, 1] = k[0], k[1]
            p[i, :] = k[2:]
            r, c = self.b[k].shape
            rows[i], cols[i] = r, c
            off = (off + align - 1) // align * align
            offs[i] = off
            off += r * c
        return dict(q=q, p=p, rows=rows, cols=cols, offsets=offs, total=off)

    def pack(self, align: int = 1, table=None) -> np.ndarray:
        """Concatenate blocks column-major at the table's offsets."""
        t = table if table is not None else self.table(align)
        buf = np.zeros(int(t["total"]), dtype=self.dtype)
        for i in range(len(t["rows"])):
            k = (int(t["q"][i, 0]), int(t["q"][i, 1])) + tuple(int(x) for x in t["p"][i])
            o, r, c = int(t["offsets"][i]), int(t["rows"][i]), int(t["cols"][i])
            if k in self.b:
                buf[o:o + r * c] = self.b[k].ravel(order="F")
        return buf

    @staticmethod
    def unpack(ns, t, buf, dtype=None, keep=None) -> "Blocks":
        """Inverse of pack; `keep` (bool per block) selects which blocks exist."""
        dtype = buf.dtype if dtype is None else dtype
        o = Blocks(ns, dtype)
        for i in range(len(t["rows"])):
            if keep is not None and not keep[i]:
                continue
            k = (int(t["q"][i, 0]), int(t["q"][i, 1])) + tuple(int(x) for x in t["p"][i])
            off, r, c = int(t["offsets"][i]), int(t["rows"][i]), int(t["cols"][i])
            o.b[k] = np.asfortranarray(np.asarray(buf[off:off + r * c]).reshape((r, c), order="F"))
        return o

    @staticmethod
    def from_table(ns, q, p, rows, cols, data, dtype=np.float64) -> "Blocks":
        o = Blocks(ns, dtype)
        off = 0
        for i in range(len(rows)):
            r, c = int(rows[i]), int(cols[i])
            k = (int(q[i][0]), int(q[i][1])) + tuple(int(x) for x in np.atleast_1d(p[i]))
            o.b[k] = np.asfortranarray(np.asarray(data[off:off + r * c]).reshape((r, c), order="F"))
            off += r * c
        return o


def twosite(A1: Blocks, A2: Blocks) -> Blocks:
    """BlockMatrix<2>(B1,B2): product of blocks with matching middle charge
    keyed (l, r, s1, s2) (src/BlockTensor.hpp:285-301)."""
    out = Blocks(2, np.result_type(A1.dtype, A2.dtype))
    for k1, a in A1.items():
        for k2, b in A2.items():
            if k1[1] == k2[0]:
                out[(k1[0], k2[1], k1[2], k2[2])] = a @ b
    return out
