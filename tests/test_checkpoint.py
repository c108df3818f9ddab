"""The versioned checkpoint format (mpsa_ckpt_*, mpstools_amd/checkpoint.py),
in place of the reference's Boost archives (MPSBase::save/load,
src/MPSTypes.hpp:377-426; Container::write/read, src/container.hpp:87-146).

CPU tests through the library itself with host buffers: an MPS (U(1) x U(1)
Hubbard, real and complex) and a two-site state round-trip bit for bit with
their key tables and metadata; a flipped byte in the data, a flipped byte in
a table, a truncated file, another format version and a non-checkpoint file
are refused with a message.  The GPU test writes device tensors and reads
them back to the device."""
import os

import numpy as np
import pytest

import complexcase as cc
from mpstools_amd import checkpoint, models, native, synth
from mpstools_amd.blocks import twosite


def _mps(L=8, D=12, seed=5):
    itk = [models.itokey_u1xu1()] * L
    mps, _ = synth.random_mps(L, itk, models.pack_charge(L // 2, L // 2), D, seed)
    return mps


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a.keys():
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape and x.dtype == y.dtype and np.array_equal(x, y)


@pytest.mark.parametrize("cplx", [False, True])
def test_checkpoint_round_trip_is_exact(tmp_path, cplx):
    mps = _mps()
    if cplx:
        mps = [cc.with_imag(A, 30 + i) for i, A in enumerate(mps)]
    path = tmp_path / "state.mpsa"
    checkpoint.save(path, mps, meta='{"model": "hubbard", "L": 8}')
    got, meta = checkpoint.load(path)
    assert meta == '{"model": "hubbard", "L": 8}'
    assert len(got) == len(mps)
    for a, b in zip(got, mps):
        _same(a, b)
    # a two-site tensor (ns = 2)
    psi = twosite(mps[3], mps[4])
    checkpoint.save(path, [psi])
    (back,), meta = checkpoint.load(path)
    assert meta == ""
    _same(back, psi)


def _corrupt(path, offset_from, delta=1):
    b = bytearray(path.read_bytes())
    i = offset_from if offset_from >= 0 else len(b) + offset_from
    b[i] = (b[i] + delta) % 256
    path.write_bytes(bytes(b))


def test_checkpoint_refuses_damaged_files(tmp_path):
    mps = _mps(L=6, D=8, seed=9)
    path = tmp_path / "s.mpsa"
    checkpoint.save(path, mps, meta="x")
    good = path.read_bytes()
    cases = {
        "data byte": lambda: _corrupt(path, len(good) // 2),
        "table byte": lambda: _corrupt(path, 8 + 4 + 4 + 8 + 8 + 1 + 4 + 4 + 8 + 3),
        "trailer": lambda: _corrupt(path, -3),
        "truncated": lambda: path.write_bytes(good[:-20]),
        "trailing bytes": lambda: path.write_bytes(good + b"\0"),
        "version": lambda: path.write_bytes(good[:8] + (2).to_bytes(4, "little") + good[12:]),
        "magic": lambda: path.write_bytes(b"NOTACKPT" + good[8:]),
    }
    for name, damage in cases.items():
        path.write_bytes(good)
        damage()
        with pytest.raises(native.MpsaError) as e:
            checkpoint.load(path)
        assert "checkpoint_open" in str(e.value), name
    path.write_bytes(good)
    got, _ = checkpoint.load(path)
    assert len(got) == len(mps)
    with pytest.raises(native.MpsaError):
        checkpoint.load(tmp_path / "missing.mpsa")


@pytest.mark.gpu
def test_checkpoint_device_round_trip(gpu, tmp_path):
    torch = gpu
    print("[case] checkpoint device round trip", flush=True)
    mps = [cc.with_imag(A, 50 + i) for i, A in enumerate(_mps(L=10, D=24, seed=11))]
    tabs = [native.Table.from_blocks(A) for A in mps]
    bufs = [torch.from_numpy(t.pack(A)).cuda() for t, A in zip(tabs, mps)]
    path = tmp_path / "dev.mpsa"
    checkpoint.save_device(path, list(zip(tabs, bufs)), meta="device")
    pairs, meta = checkpoint.load_device(path)
    assert meta == "device"
    for (t, d), t0, b0 in zip(pairs, tabs, bufs):
        assert t.keys() == t0.keys()
        assert torch.equal(d[:b0.numel()], b0)
    # the host reader sees the same tensors
    got, _ = checkpoint.load(path)
    for a, b in zip(got, mps):
        _same(a, b)
    os.remove(path)
