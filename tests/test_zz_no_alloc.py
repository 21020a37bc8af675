"""The allocation-free sky conversion (baorec_sky_to_cartesian_noalloc)
against baorec_sky_to_cartesian on device pointers, through the C ABI."""
import ctypes

import numpy as np
import pytest

import baorec_hip as bh
from baorec_hip import _lib as L
from baorec_hip.devmem import DeviceArray

pytestmark = pytest.mark.gpu

F32 = np.float32


def sky_catalog(n, seed):
    rng = np.random.default_rng(seed)
    ra = (rng.random(n) * 90.0).astype(F32)
    dec = (rng.random(n) * 90.0).astype(F32)
    red = (0.05 + rng.random(n) * 1.5).astype(F32)
    return ra, dec, red


def convert(fn_name, ctx, cosmo, ra, dec, red):
    n = len(ra)
    din = [DeviceArray.from_numpy(a) for a in (ra, dec, red)]
    dout = [DeviceArray((n,), F32) for _ in range(3)]
    s = cosmo.struct()
    fn = getattr(ctx.lib, fn_name)
    ptrs = [ctypes.c_void_p(a.ptr) for a in din] + [n] + [ctypes.c_void_p(a.ptr) for a in dout]
    if fn_name == "baorec_sky_to_cartesian":
        rc = fn(ctx.handle, ctypes.byref(s), L.MEM_DEVICE, *ptrs)
    else:
        rc = fn(ctx.handle, ctypes.byref(s), *ptrs)
    ctx.check(rc)
    return tuple(a.numpy() for a in dout)


def test_noalloc_matches_allocating_entry_point_twice():
    """Two calls in a row, the second with a smaller catalog: each equals
    baorec_sky_to_cartesian bit for bit (same table and kernels)."""
    ctx = bh.Context(0)
    cosmo = bh.Cosmology()
    for n, seed in ((200_000, 1), (3_000, 2)):
        ra, dec, red = sky_catalog(n, seed)
        got = convert("baorec_sky_to_cartesian_noalloc", ctx, cosmo, ra, dec, red)
        ref = convert("baorec_sky_to_cartesian", ctx, cosmo, ra, dec, red)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
        assert np.all(np.isfinite(got[0]))
    ctx.close()


def test_noalloc_new_cosmology_and_errors():
    """A second cosmology rebuilds the table; a redshift outside the table is
    BoundsError and the next call starts from a cleared flag; a table longer
    than the scratch array is refused."""
    ctx = bh.Context(0)
    ra, dec, red = sky_catalog(5_000, 3)
    desi = bh.Cosmology.desi()
    got = convert("baorec_sky_to_cartesian_noalloc", ctx, desi, ra, dec, red)
    ref = convert("baorec_sky_to_cartesian", ctx, desi, ra, dec, red)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    bad = red.copy()
    bad[7] = F32(5.0)
    with pytest.raises(bh.BoundsError):
        convert("baorec_sky_to_cartesian_noalloc", ctx, desi, ra, dec, bad)
    got = convert("baorec_sky_to_cartesian_noalloc", ctx, desi, ra, dec, red)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    big = bh.Cosmology(z_tab_num=131073)
    with pytest.raises(bh.BaorecError, match="scratch limit"):
        convert("baorec_sky_to_cartesian_noalloc", ctx, big, ra, dec, red)
    ctx.close()
