"""Batched reconstruction (baorec_run_batch) against the single-catalog path
and the CPU oracle.

A batch runs, per catalog, the same kernels as run!(IterativeRecon, ...) +
reconstructed_positions(...; field = :sum) with every launch covering all
catalogs of a group (a scatter group for the sort and the mass assignment,
an FFT subgroup for the rest); the per-catalog FP64 sums and the
fixed-point mass assignment are deterministic, so the batch must reproduce
the single-catalog outputs bit for bit.  One catalog per case is also checked
against the oracle (1e-4 Mpc/h RMS on the shifted positions).
"""
import numpy as np
import pytest

import baorec_hip as bh
import baorec_oracle as orc
from baorec_hip import synth

pytestmark = pytest.mark.gpu

F32 = np.float32


def rms3(a, b):
    return float(np.sqrt(np.mean([np.mean((np.asarray(p, np.float64) - np.asarray(q, np.float64)) ** 2)
                                  for p, q in zip(a, b)])))


def box_kw(los=(0.0, 0.0, 1.0), **extra):
    return dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=[0.0] * 3,
                los=None if los is None else list(los), **extra)


def single_path(kw, grid, x, y, z, w, off):
    """run! + reconstructed_positions(field = :sum) per catalog; returns the
    positions and the wrapped inputs."""
    xs, ys, zs = x.copy(), y.copy(), z.copy()
    out = [np.empty_like(x) for _ in range(3)]
    for c in range(len(off) - 1):
        s = slice(off[c], off[c + 1])
        if off[c + 1] == off[c]:
            continue
        cx, cy, cz = xs[s].copy(), ys[s].copy(), zs[s].copy()
        rec = bh.IterativeRecon(**kw)
        bh.run_(rec, grid, cx, cy, cz, w[s], copy=False)
        p = bh.reconstructed_positions(rec, cx, cy, cz, field="sum")
        for k in range(3):
            out[k][s] = p[k]
        xs[s], ys[s], zs[s] = cx, cy, cz
    return out, (xs, ys, zs)


@pytest.mark.parametrize("los,mas", [((0.0, 0.0, 1.0), "cic"), (None, "cic"), ((0.0, 0.0, 1.0), "tsc")])
def test_batch_matches_single_catalog_path(ctx, los, mas):
    grid = (64, 64, 64)
    sizes = np.array([3000, 0, 50000, 12345, 120000], np.int64)
    x, y, z, w, off = synth.batch_catalogs(len(sizes), mesh=64, seed0=300, sizes=sizes)
    # a few particles on and beyond the upper face (one-sided wrap)
    x[5] = F32(1000.0)
    y[60000] = F32(1003.5)
    kw = box_kw(los=los, mas=mas)
    exp, wrapped = single_path(kw, grid, x, y, z, w, off)
    bx, by, bz = x.copy(), y.copy(), z.copy()
    got = bh.run_batch(bh.IterativeRecon(**kw), grid, bx, by, bz, w, off)
    for k in range(3):
        np.testing.assert_array_equal(got[k], exp[k])
    # the inputs are wrapped in place as run! wraps them
    for a, b in zip((bx, by, bz), wrapped):
        np.testing.assert_array_equal(a, b)


def test_batch_catalog_matches_oracle(ctx):
    grid = (64, 64, 64)
    sizes = np.array([40000, 80000], np.int64)
    x, y, z, w, off = synth.batch_catalogs(2, mesh=64, seed0=310, sizes=sizes)
    kw = box_kw()
    got = bh.run_batch(bh.IterativeRecon(**kw), grid, x.copy(), y.copy(), z.copy(), w, off)
    for c in range(2):
        s = slice(off[c], off[c + 1])
        o = orc.IterativeRecon(**kw)
        ox, oy, oz = x[s].copy(), y[s].copy(), z[s].copy()
        orc.run_iterative(o, grid, ox, oy, oz, w[s])
        exp = orc.reconstructed_positions(o, ox, oy, oz, kind="sum")
        err = rms3([g[s] for g in got], exp)
        print(f"catalog {c}: RMS vs oracle {err:.2e} Mpc/h")
        assert err < 1e-4


def test_batch_several_groups_device_memory(ctx):
    """At 256^3 the 20 catalogs scatter as two groups of 10 and run the FFT
    pipeline in subgroups of 2 (2 GiB of grids and spectra) over each
    group's density grids; device-resident inputs."""
    from baorec_hip.devmem import DeviceArray
    grid = (256, 256, 256)
    rng = np.random.default_rng(5)
    sizes = rng.integers(2000, 20000, 20).astype(np.int64)
    x, y, z, w, off = synth.batch_catalogs(20, mesh=64, seed0=400, sizes=sizes)
    kw = box_kw()
    dev = [DeviceArray.from_numpy(a.copy()) for a in (x, y, z, w)]
    got = bh.run_batch(bh.IterativeRecon(**kw), grid, *dev, off)
    got = [g.numpy() for g in got]
    # both of the first FFT subgroup, a middle one, the last of the first
    # scatter group, the first and last of the second
    for c in (0, 1, 5, 9, 10, 19):
        s = slice(off[c], off[c + 1])
        cx, cy, cz = x[s].copy(), y[s].copy(), z[s].copy()
        rec = bh.IterativeRecon(**kw)
        bh.run_(rec, grid, cx, cy, cz, w[s], copy=False)
        exp = bh.reconstructed_positions(rec, cx, cy, cz, field="sum")
        for k in range(3):
            np.testing.assert_array_equal(got[k][s], exp[k])
    assert all(np.isfinite(g).all() for g in got)


def test_batch_out_of_box_raises(ctx):
    grid = (32, 32, 32)
    sizes = np.array([1000, 2000], np.int64)
    x, y, z, w, off = synth.batch_catalogs(2, mesh=32, seed0=500, sizes=sizes)
    x[1500] = F32(-5.0)   # below box_min: not wrapped, index -1
    with pytest.raises(bh.BoundsError):
        bh.run_batch(bh.IterativeRecon(**box_kw()), grid, x, y, z, w, off)


def test_batch_rejects_bad_offsets(ctx):
    grid = (32, 32, 32)
    x, y, z, w, off = synth.batch_catalogs(3, mesh=32, seed0=510, sizes=np.array([100, 200, 300]))
    bad = off.copy()
    bad[1], bad[2] = bad[2], bad[1]
    with pytest.raises(bh.BaorecError):
        bh.run_batch(bh.IterativeRecon(**box_kw()), grid, x, y, z, w, bad)
    with pytest.raises(ValueError):
        bh.run_batch(bh.IterativeRecon(**box_kw()), grid, x, y, z, w, off[:-1])
