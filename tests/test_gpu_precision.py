"""Precision of every result against the FP64 oracle.

Each case runs the HIP path (FP32 bulk arithmetic, FP64 global reductions)
and the oracle twice: in the reference's FP32 operation order and in FP64.
The HIP error against FP64 must be of the size of the FP32 algorithm's own
error against FP64 (at most 2x it, or within the parity tolerance), so
that the GPU adds no error beyond FP32 rounding.  Set BAOREC_PRECISION_OUT
to a path to collect the table rows (markdown).
"""
import os

import numpy as np
import pytest

import baorec_hip as bh
import baorec_oracle as orc
from baorec_hip import synth

pytestmark = pytest.mark.gpu

F32, F64 = np.float32, np.float64


def rel_l2(a, b):
    a, b = np.asarray(a, F64), np.asarray(b, F64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def rms3(a, b):
    return float(np.sqrt(np.mean([np.mean((np.asarray(p, F64) - np.asarray(q, F64)) ** 2) for p, q in zip(a, b)])))


def report(case, quantity, metric, e_hip, e_f32, bound):
    row = f"| {case} | {quantity} | {metric} | {e_hip:.2e} | {e_f32:.2e} | {bound} |"
    print(row)
    out = os.environ.get("BAOREC_PRECISION_OUT")
    if out:
        with open(out, "a") as f:
            f.write(row + "\n")


def check(e_hip, e_f32, tol):
    assert e_hip <= max(2.0 * e_f32, tol), (e_hip, e_f32, tol)


BOX = dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=[0.0] * 3)


@pytest.fixture(scope="module")
def box1m():
    """The 128^3 box configuration: 1,000,000 lognormal tracers, L = 1000."""
    return synth.box_catalog(1_000_000, L=1000.0, mesh=256, seed=1)


def test_precision_iterative_box(ctx, box1m):
    x, y, z, w = box1m
    g = (128, 128, 128)
    kw = dict(BOX, los=[0.0, 0.0, 1.0])
    case = "iterative, box 128³, 1M, LOS z"
    ds = bh.setup_overdensity_(np.zeros((128,) * 3, F32), bh.IterativeRecon(**kw), x.copy(), y.copy(), z.copy(), w)
    ref = {dt: orc.setup_overdensity(np.zeros((128,) * 3, dt), orc.IterativeRecon(**kw), x.copy(), y.copy(),
                                     z.copy(), w) for dt in (F32, F64)}
    e, e32 = rel_l2(ds, ref[F64]), rel_l2(ref[F32], ref[F64])
    report(case, "δ_s (setup_overdensity!)", "rel. L2", e, e32, "1e-5")
    check(e, e32, 1e-5)
    rec = bh.IterativeRecon(**kw)
    hx, hy, hz = x.copy(), y.copy(), z.copy()
    d = bh.run_(rec, g, hx, hy, hz, w)
    o = {dt: orc.IterativeRecon(**kw) for dt in (F32, F64)}
    dr = {dt: orc.run_iterative(o[dt], g, x.copy(), y.copy(), z.copy(), w, dtype=dt) for dt in (F32, F64)}
    e, e32 = rel_l2(d, dr[F64]), rel_l2(dr[F32], dr[F64])
    report(case, "δ_r (run!, 3 iterations)", "rel. L2", e, e32, "1e-5")
    check(e, e32, 1e-5)
    for kind in ("disp", "sum"):
        got = bh.read_shifts(rec, hx, hy, hz, field=kind)
        ex = {dt: orc.read_shifts(o[dt], hx, hy, hz, kind=kind) for dt in (F32, F64)}
        e, e32 = rms3(got, ex[F64]), rms3(ex[F32], ex[F64])
        report(case, f"shifts :{kind}", "RMS Mpc/h", e, e32, "1e-4")
        check(e, e32, 1e-4)
    got = bh.reconstructed_positions(rec, hx, hy, hz, field="sum")
    ex = {dt: orc.reconstructed_positions(o[dt], hx, hy, hz, kind="sum") for dt in (F32, F64)}
    e, e32 = rms3(got, ex[F64]), rms3(ex[F32], ex[F64])
    report(case, "positions (:sum)", "RMS Mpc/h", e, e32, "1e-4")
    check(e, e32, 1e-4)


def test_precision_iterative_lightcone(ctx):
    g = (128, 128, 128)
    x, y, z, w = synth.lightcone_catalog(200000, r_min=500.0, r_max=900.0, mesh=128, seed=11)
    rx, ry, rz, rw = synth.lightcone_randoms(2000000, r_min=500.0, r_max=900.0, seed=12)
    kw = dict(bias=2.2, f=0.75, smoothing_radius=15.0, los=None)
    case = "iterative, lightcone 128³, 200k + 2M, local LOS"
    rec = bh.IterativeRecon(**kw, box_pad=200.0)
    d = bh.run_(rec, g, x, y, z, w, rx, ry, rz, rw)
    o = {dt: orc.IterativeRecon(**kw) for dt in (F32, F64)}
    dr = {dt: orc.run_iterative(o[dt], g, x, y, z, w, rx, ry, rz, rw, box_pad=200.0, dtype=dt) for dt in (F32, F64)}
    e, e32 = rel_l2(d, dr[F64]), rel_l2(dr[F32], dr[F64])
    report(case, "δ_r (run!, 3 iterations)", "rel. L2", e, e32, "≤ 2× FP32")
    check(e, e32, 1e-5)
    got = bh.reconstructed_positions(rec, x, y, z, field="sum")
    ex = {dt: orc.reconstructed_positions(o[dt], x, y, z, kind="sum") for dt in (F32, F64)}
    e, e32 = rms3(got, ex[F64]), rms3(ex[F32], ex[F64])
    report(case, "positions (:sum)", "RMS Mpc/h", e, e32, "≤ 2× FP32")
    check(e, e32, 1e-4)
    got = bh.reconstructed_positions(rec, rx, ry, rz, field="disp")
    ex = {dt: orc.reconstructed_positions(o[dt], rx, ry, rz, kind="disp") for dt in (F32, F64)}
    e, e32 = rms3(got, ex[F64]), rms3(ex[F32], ex[F64])
    report(case, "randoms positions (:disp)", "RMS Mpc/h", e, e32, "≤ 2× FP32")
    check(e, e32, 1e-4)


def test_precision_multigrid_box(ctx, box1m):
    x, y, z, w = box1m
    g = (128, 128, 128)
    kw = dict(BOX, los=None)
    case = "multigrid (reference FMG), box 128³, 1M, local LOS"
    rec = bh.MultigridRecon(**kw)
    hx, hy, hz = x.copy(), y.copy(), z.copy()
    phi = bh.run_(rec, g, hx, hy, hz, w)
    o = {dt: orc.MultigridRecon(**kw) for dt in (F32, F64)}
    pr = {dt: orc.run_multigrid(o[dt], g, x.copy(), y.copy(), z.copy(), w, dtype=dt) for dt in (F32, F64)}
    e, e32 = rel_l2(phi, pr[F64]), rel_l2(pr[F32], pr[F64])
    report(case, "φ (run!)", "rel. L2", e, e32, "≤ 2× FP32")
    check(e, e32, 1e-5)
    got = bh.reconstructed_positions(rec, hx, hy, hz, field="sum")
    ex = {dt: orc.reconstructed_positions(o[dt], hx, hy, hz, kind="sum") for dt in (F32, F64)}
    e, e32 = rms3(got, ex[F64]), rms3(ex[F32], ex[F64])
    report(case, "positions (:sum)", "RMS Mpc/h", e, e32, "1e-4")
    check(e, e32, 1e-4)


def test_precision_batch(ctx):
    g = (128, 128, 128)
    sizes = np.array([150000, 400000], np.int64)
    x, y, z, w, off = synth.batch_catalogs(2, mesh=128, seed0=700, sizes=sizes)
    kw = dict(BOX, los=[0.0, 0.0, 1.0])
    got = bh.run_batch(bh.IterativeRecon(**kw), g, x.copy(), y.copy(), z.copy(), w, off)
    e_all, e32_all = [], []
    for c in range(2):
        s = slice(off[c], off[c + 1])
        ex = {}
        for dt in (F32, F64):
            o = orc.IterativeRecon(**kw)
            cx, cy, cz = x[s].copy(), y[s].copy(), z[s].copy()
            orc.run_iterative(o, g, cx, cy, cz, w[s], dtype=dt)
            ex[dt] = orc.reconstructed_positions(o, cx, cy, cz, kind="sum")
        e_all.append(rms3([a[s] for a in got], ex[F64]))
        e32_all.append(rms3(ex[F32], ex[F64]))
    e, e32 = max(e_all), max(e32_all)
    report("batch, 2 boxes 128³ (150k, 400k), LOS z", "positions (:sum), worst catalog", "RMS Mpc/h", e, e32, "1e-4")
    check(e, e32, 1e-4)


def test_precision_power_spectrum(ctx):
    n, L = 128, 1000.0
    x, y, z, w = synth.box_catalog(200000, L=L, mesh=128, seed=97)
    got = bh.power_multipoles(x, y, z, w, (n, n, n), L, dk=0.01, nbins=40, ctx=ctx)
    ref = orc.power_multipoles(x, y, z, w, (n, n, n), (L, L, L), dk=0.01, nbins=40)
    scale = np.abs(ref["p0"]) + L ** 3 / len(x)
    for q in ("p0", "p2", "p4"):
        e = float(np.max(np.abs(got[q] - ref[q]) / scale))
        report("power spectrum, box 128³, 200k, TSC + interlacing", f"P_{q[1]}", "max rel. (to P_0 + shot)", e,
               0.0, "1e-4 (oracle grids FP32, FFT FP64)")
        assert e < 1e-4


def test_precision_sky(ctx):
    c, o = bh.Cosmology(z_tab_num=2001), orc.Cosmology(z_tab_num=2001)
    # at the table knots (between them the reference interpolates linearly,
    # and so does the library)
    zq = np.linspace(0.0, 3.0, 2001)[1::20]
    r = bh.comoving_distance(zq, c, ctx=ctx)
    r64 = orc.comoving_distance(o, zq)
    e = float(np.max(np.abs(r - r64) / r64))
    report("comoving distance, 0 < z ≤ 3", "χ(z) at the table knots vs adaptive quadrature", "max rel.", e, 0.0,
           "1e-9")
    assert e < 1e-9
    rng = np.random.default_rng(3)
    ra = rng.uniform(0, 90, 100000).astype(F32)
    dec = rng.uniform(0, 90, 100000).astype(F32)
    red = rng.uniform(0.1, 1.5, 100000).astype(F32)
    got = bh.sky_to_cartesian(ra, dec, red, c, ctx=ctx)
    # the driver's algorithm in FP64: the same distance table and linear
    # interpolation, angles and trigonometry in FP64
    zk, rk = orc.distance_table(o)
    hh = float(F32(F32(F32(0.67) * F32(100)) / F32(100)))
    dist = orc._interp(zk, rk, red.astype(F64)) * hh
    rr, dd = np.deg2rad(ra.astype(F64)), np.deg2rad(dec.astype(F64))
    ex = (dist * np.cos(dd) * np.cos(rr), dist * np.cos(dd) * np.sin(rr), dist * np.sin(dd))
    ref32 = orc.sky_to_cartesian(ra, dec, red, o, (zk, rk))
    e, e32 = rms3(got, ex), rms3(ref32, ex)
    report("sky → Cartesian, 100k galaxies", "positions", "RMS Mpc/h", e, e32, "≤ 2× FP32")
    check(e, e32, 1e-4)
