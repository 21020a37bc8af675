"""C-ABI checks that need no GPU: the library loads, exports every function
include/baorec.h declares, reports a missing device as an error code (never
an abort), and its host-only entry points agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import baorec_hip as bh
import baorec_oracle as orc
from baorec_hip import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F32 = np.float32


def declared_functions():
    src = open(os.path.join(ROOT, "include", "baorec.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(baorec_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes mirror binds every declared function
    assert set(names) <= set(L.SIGNATURES), set(names) - set(L.SIGNATURES)


def test_params_layout_and_defaults():
    lib = L.load()
    p = L.Params()
    lib.baorec_params_default(ctypes.byref(p), 2.2, 0.75, 15.0)
    assert p.n_iter == 3 and p.jacobi_niterations == 5 and p.vcycle_niterations == 6
    assert p.jacobi_damping_factor == pytest.approx(0.4)
    assert p.beta == F32(F32(0.75) / F32(2.2))
    assert p.box_pad == 500.0 and p.ran_min == pytest.approx(0.01) and p.mutate_inputs == 1
    assert p.has_los == 0 and p.mas == 0 and p.mg_schedule == 0


def test_version_and_errors_without_gpu():
    lib = L.load()
    assert b"gfx950" in lib.baorec_version()
    assert lib.baorec_create(None, 0) == L.ERR_INVALID_ARGUMENT
    assert lib.baorec_destroy(None) == L.ERR_INVALID_ARGUMENT
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = lib.baorec_create(ctypes.byref(h), 0)
    assert rc == L.ERR_NO_DEVICE
    assert b"device" in lib.baorec_last_error(None)
    with pytest.raises(bh.BaorecError) as e:
        bh.Context(0)
    assert e.value.code == L.ERR_NO_DEVICE


def test_null_context_calls_fail_cleanly():
    lib = L.load()
    p = L.Params()
    g = (ctypes.c_int32 * 3)(16, 16, 16)
    assert lib.baorec_run_iterative(None, ctypes.byref(p), g, 0, None, None, None, None, 0,
                                    None, None, None, None, 0) == L.ERR_INVALID_ARGUMENT
    assert lib.baorec_result_info(None, g, ctypes.byref(ctypes.c_int32()), None) == L.ERR_INVALID_ARGUMENT
    assert lib.baorec_profile(None, 1) == L.ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("grid,box", [((16, 32, 8), (1000.0, 500.0, 250.0)), ((64, 64, 64), (2000.0,) * 3)])
def test_k_vec_matches_oracle(grid, box):
    got = bh.k_vec(grid, box)
    ref = orc.k_vec(grid, box)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_x_vec_matches_oracle():
    grid, box, bmin = (16, 8, 12), (1980.0, 1980.0, 1980.0), (-90.25, 3.5, 12.0)
    for a, b in zip(bh.x_vec(grid, box, bmin), orc.x_vec(grid, box, bmin)):
        np.testing.assert_array_equal(a, b)


def test_k_vec_rejects_bad_arguments():
    lib = L.load()
    out = np.zeros(64, F32)
    p = out.ctypes.data_as(ctypes.c_void_p)
    bad = (ctypes.c_int32 * 3)(0, 8, 8)
    box = np.ones(3, F32)
    assert lib.baorec_k_vec(bad, box.ctypes.data_as(ctypes.c_void_p), p, p, p) == L.ERR_INVALID_ARGUMENT
