"""Catalog I/O (host code of libbaorec_hip.so, no GPU): the drivers' ASCII
catalogs and .npy outputs.

The reader must turn every decimal into the nearest Float32, as CSV.jl's
Float32 columns do (test/benchmarks_cpu.jl:22-23); the golden fixture's
expected bits come from exact rational rounding (tests/golden/make_golden.py),
not from a float parser.  The .npy writer must give numpy the (n, ncols)
matrix NPZ.jl writes for hcat(pos', nz) (examples/lightcone.jl:125-127).
"""
import json
import os
import struct

import numpy as np
import pytest

import baorec_hip as bh

F32 = np.float32
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.asarray(a, F32).view(np.uint32)


def test_reader_matches_exact_rounding_golden():
    cols = bh.read_catalog(os.path.join(GOLD, "catalog.txt"))
    want = np.array(json.load(open(os.path.join(GOLD, "catalog.json")))["rows_f32_bits"], np.uint32)
    got = np.stack([bits(c) for c in cols], axis=1)
    np.testing.assert_array_equal(got, want)


def test_reader_column_selection_and_counts(tmp_path):
    p = tmp_path / "c.txt"
    p.write_text("1 2 3 4 5 6\n7 8 9 10 11 12\n")
    ra, z = bh.read_catalog(str(p), columns=(0, 3))
    np.testing.assert_array_equal(ra, [1, 7])
    np.testing.assert_array_equal(z, [4, 10])
    import ctypes
    lib = bh.load()
    n, nc = ctypes.c_int64(), ctypes.c_int32()
    assert lib.baorec_catalog_rows(os.fsencode(str(p)), ctypes.byref(n), ctypes.byref(nc)) == 0
    assert (n.value, nc.value) == (2, 6)


def test_reader_large_catalog_multithreaded_is_bit_exact(tmp_path):
    """200k rows (~14 MB, several parsing threads): shortest round-trip
    spellings of float32 values come back bit-exact, in order."""
    rng = np.random.default_rng(41)
    n = 200_000
    src = [rng.uniform(0, 360, n).astype(F32), rng.uniform(-90, 90, n).astype(F32),
           rng.uniform(0.6, 1.1, n).astype(F32), np.ones(n, F32),
           (10 ** rng.uniform(-5, -3, n)).astype(F32)]
    p = tmp_path / "big.txt"
    with open(p, "w") as fh:
        fh.write("\n".join(" ".join(np.format_float_positional(c[i], unique=True)
                                    for c in src) for i in range(n)) + "\n")
    assert os.path.getsize(p) > 8 << 20
    got = bh.read_catalog(str(p))
    for g, s in zip(got, src):
        np.testing.assert_array_equal(bits(g), bits(s))


def test_reader_reports_malformed_line(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("1 2 3 4 5\n1 2 x 4 5\n")
    with pytest.raises(bh.BaorecError, match="line 2"):
        bh.read_catalog(str(p))
    p.write_text("1 2 3 4 5\n\n1 2 3\n")
    with pytest.raises(bh.BaorecError, match="line 3"):
        bh.read_catalog(str(p))
    with pytest.raises(bh.BaorecError, match="cannot open"):
        bh.read_catalog(str(tmp_path / "missing.txt"))


def test_reader_empty_and_comment_only(tmp_path):
    p = tmp_path / "e.txt"
    p.write_text("")
    assert all(len(c) == 0 for c in bh.read_catalog(str(p)))
    p.write_text("# nothing\n\n   \n")
    assert all(len(c) == 0 for c in bh.read_catalog(str(p)))


def test_npy_writer_layout(tmp_path):
    rng = np.random.default_rng(42)
    cols = [rng.standard_normal(1001).astype(F32) for _ in range(4)]
    p = tmp_path / "out.npy"
    bh.write_npy(str(p), *cols)
    a = np.load(p, allow_pickle=False)
    assert a.shape == (1001, 4) and a.dtype == F32 and a.flags.f_contiguous
    np.testing.assert_array_equal(a, np.stack(cols, axis=1))
    raw = p.read_bytes()
    hlen = struct.unpack("<H", raw[8:10])[0]
    assert raw[:6] == b"\x93NUMPY" and (10 + hlen) % 64 == 0
    # columns are contiguous after the header, as a Julia Matrix{Float32}
    np.testing.assert_array_equal(np.frombuffer(raw[10 + hlen:], F32), np.concatenate(cols))
    bh.write_npy(str(p), np.zeros(0, F32), np.zeros(0, F32))
    assert np.load(p).shape == (0, 2)
