"""The reference drivers' sequence end to end on the GPU, and the
configurations 3 and 4 at their own sizes.

examples/lightcone.jl:22-127 reads ASCII catalogs (ra dec d z nz), keeps
0.8 < z < 1, converts sky -> Cartesian, weights w = 1 / (1 + nz P0), runs the
reconstruction with randoms, shifts data (:sum) and randoms (:sum, :disp),
converts back to the sky and writes .npy files of hcat(sky', nz).  The test
runs that sequence through the library (catalog reader, sky conversion,
run!, read-outs, .npy writer) and through the oracle, and compares them."""
import os

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


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def sky_catalog(n, seed, cosmo):
    """(ra, dec, d, z, nz) rows of a cap catalog around 0.7 < z < 1.1 (some
    outside the kept redshift range), with a smooth n(z) column."""
    rng = np.random.default_rng(seed)
    ra = (rng.random(n) * 60.0 + 10.0).astype(F32)
    dec = (rng.random(n) * 50.0 + 5.0).astype(F32)
    z = (0.7 + rng.random(n) * 0.4).astype(F32)
    table = orc.distance_table(orc.Cosmology(z_tab_max=cosmo.z_tab_max, z_tab_num=cosmo.z_tab_num))
    d = np.interp(z.astype(np.float64), *table).astype(F32)
    nz = (3e-4 * np.exp(-((z - 0.9) / 0.2) ** 2)).astype(F32)
    return np.stack([ra, dec, d, z, nz], axis=1)


def write_ascii(path, rows):
    with open(path, "w") as fh:
        fh.write("# ra dec d z nz\n")
        for r in rows:
            fh.write(" ".join(repr(float(v)) for v in r) + "\n")


def test_lightcone_driver_sequence(ctx, tmp_path):
    P0 = F32(5e3)
    cosmo = bh.Cosmology(z_tab_max=10.0)
    ocosmo = orc.Cosmology(z_tab_max=10.0)
    for name, n, seed in (("data", 30000, 1), ("rand", 150000, 2)):
        write_ascii(tmp_path / f"{name}.txt", sky_catalog(n, seed, cosmo))
    out = {}
    for impl in ("hip", "oracle"):
        cats = {}
        for name in ("data", "rand"):
            if impl == "hip":
                cols = bh.read_catalog(str(tmp_path / f"{name}.txt"), columns=(0, 1, 2, 3, 4))
            else:
                a = np.loadtxt(tmp_path / f"{name}.txt", dtype=np.float64)
                cols = [a[:, k].astype(F32) for k in range(5)]
            ra, dec, _, z, nz = cols
            keep = (z > 0.8) & (z < 1.0)
            ra, dec, z, nz = ra[keep], dec[keep], z[keep], nz[keep]
            if impl == "hip":
                pos = bh.sky_to_cartesian(ra, dec, z, cosmo, ctx=ctx)
            else:
                pos = orc.sky_to_cartesian(ra, dec, z, ocosmo)
            w = (F32(1) / (F32(1) + nz * P0)).astype(F32)
            cats[name] = ([np.ascontiguousarray(p, dtype=F32) for p in pos], w, nz)
        (dp, dw, dnz), (rp, rw, rnz) = cats["data"], cats["rand"]
        kw = dict(bias=2.2, f=0.757, smoothing_radius=15.0, los=None)
        n = 64
        if impl == "hip":
            rec = bh.IterativeRecon(**kw, box_pad=100.0, ctx=ctx)
            bh.run_(rec, (n, n, n), *[p.copy() for p in dp], dw, *rp, rw)
            new = bh.reconstructed_positions(rec, *dp, field="sum")
            sym = bh.reconstructed_positions(rec, *rp, field="sum")
            iso = bh.reconstructed_positions(rec, *rp, field="disp")
            sky = [bh.cartesian_to_sky(*q, cosmo, ctx=ctx) for q in (new, sym, iso)]
            for tag, s, nzc in zip(("rec", "sym", "iso"), sky, (dnz, rnz, rnz)):
                bh.write_npy(str(tmp_path / f"{tag}.npy"), *s, nzc)
            arrs = [np.load(tmp_path / f"{tag}.npy") for tag in ("rec", "sym", "iso")]
        else:
            rec = orc.IterativeRecon(**kw)
            orc.run_iterative(rec, (n, n, n), *[p.copy() for p in dp], dw, *rp, rw, box_pad=100.0)
            new = orc.reconstructed_positions(rec, *dp, kind="sum")
            sym = orc.reconstructed_positions(rec, *rp, kind="sum")
            iso = orc.reconstructed_positions(rec, *rp, kind="disp")
            sky = [orc.cartesian_to_sky(*q, ocosmo) for q in (new, sym, iso)]
            arrs = [np.stack([*s, nzc], axis=1) for s, nzc in zip(sky, (dnz, rnz, rnz))]
        out[impl] = (new, sym, iso, arrs)
    for k in range(3):
        assert rms3(out["hip"][k], out["oracle"][k]) < 1e-4      # Mpc/h
    for a, b in zip(out["hip"][3], out["oracle"][3]):
        assert a.shape == b.shape and a.dtype == np.float32
        np.testing.assert_array_equal(a[:, 3], b[:, 3])                           # nz column
        assert np.max(np.abs(a[:, 1] - b[:, 1])) < 1e-4                          # Dec, degrees
        dra = np.abs(a[:, 0] - b[:, 0])
        assert np.max(np.minimum(dra, 360.0 - dra)) < 1e-4                         # RA, degrees
        assert np.max(np.abs(a[:, 2] - b[:, 2])) < 1e-6                           # redshift


def test_config3_multigrid_512_box_catalog(ctx):
    """Configuration 3: RBGS multigrid (2+2 sweeps, FP64 levels <= 8^3) to
    relative residual 1e-6 on the configuration-1 catalog (1M lognormal
    tracers, L = 1000 Mpc/h) at 512^3, local LOS: reaches the tolerance, the
    positions are finite and bitwise reproducible, and they agree with the
    256^3 reconstruction of the same catalog to a tenth of the displacement
    (the two grids resolve the smoothed field alike)."""
    x, y, z, w = synth.box_catalog(1_000_000, L=1000.0, mesh=256, seed=1)
    kw = dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=[0.0] * 3, los=None,
              schedule="rbgs", tolerance=1e-6, rb_sweeps=2, fp64_coarse_max=8, ctx=ctx)
    pos = {}
    for n, reps in ((512, 2), (256, 1)):
        runs = []
        for _ in range(reps):
            rec = bh.MultigridRecon(**kw)
            bh.run_(rec, (n, n, n), x.copy(), y.copy(), z.copy(), w, copy=False)
            assert rec.final_residual <= 1e-6, (n, rec.final_residual)
            runs.append(bh.reconstructed_positions(rec, x, y, z, field="sum"))
            print(f"{n}^3: {rec.vcycles_done} V-cycles to {rec.final_residual:.2e}, shift {rec.rhs_shift:.3e}")
        for a, b in zip(runs[0], runs[-1]):
            np.testing.assert_array_equal(a, b)
        assert all(np.all(np.isfinite(a)) for a in runs[0])
        pos[n] = runs[0]
    disp = rms3(pos[512], (x, y, z))
    diff = rms3(pos[512], pos[256])
    print(f"rms displacement {disp:.3f} Mpc/h; 512^3 vs 256^3 positions {diff:.3f} Mpc/h")
    assert 1.0 < disp < 30.0 and diff < 0.1 * disp


def test_config3_multigrid_128_matches_oracle_solution(ctx):
    """The configuration-3 solve (on a smaller catalog at 128^3) against the
    FP64 oracle: the
    oracle's own V-cycles on the same right-hand side, run to rounding with
    the constant shift of a local LOS, give the same potential."""
    x, y, z, w = synth.box_catalog(200_000, L=1000.0, mesh=128, seed=11)
    n = 128
    kw = dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=[0.0] * 3, los=None)
    rec = bh.MultigridRecon(**kw, schedule="rbgs", tolerance=1e-6, ctx=ctx)
    f = np.zeros((n, n, n), F32)
    bh.setup_overdensity_(f, rec, x.copy(), y.copy(), z.copy(), w)
    phi = bh.fmg(f, rec)
    assert rec.final_residual <= 1e-6
    o = orc.MultigridRecon(**kw)
    f64 = f.astype(np.float64)
    v = np.zeros_like(f64)
    ff = f64 - f64.mean()
    for _ in range(20):
        orc.mg_vcycle(v, ff, o.box_size, o.box_min, o.beta, 0.4, 5, los=None)
        r = orc.mg_residual(v, ff, o.box_size, o.box_min, o.beta, los=None)
        ff -= r.mean()
    res = np.linalg.norm(orc.mg_residual(v, ff, o.box_size, o.box_min, o.beta, los=None)) / np.linalg.norm(f64)
    err = rel_l2(phi - phi.mean(), v - v.mean())
    shift = f64.mean() - ff.mean()
    print(f"oracle residual {res:.1e}; potential rel L2 {err:.2e}; shift {rec.rhs_shift:.6e} vs {shift:.6e}")
    assert res < 1e-8 and err < 1e-5
    assert abs(rec.rhs_shift - shift) < 1e-6 * float(np.std(f64))


def test_config4_batch_64_catalogs_256(ctx):
    """Configuration 4: 64 lognormal box catalogs (N log-uniform in
    [2e5, 5e6], slopes in [-2, -1]) at 256^3 in one batched call: every
    output finite, three catalogs bit for bit the single-catalog run! +
    reconstructed_positions, and the smallest against the oracle."""
    x, y, z, w, off = synth.batch_catalogs(64, mesh=128)
    kw = dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=[0.0] * 3,
              los=[0.0, 0.0, 1.0])
    n = 256
    bx, by, bz = x.copy(), y.copy(), z.copy()
    out = bh.run_batch(bh.IterativeRecon(**kw, ctx=ctx), (n, n, n), bx, by, bz, w, off)
    assert all(np.all(np.isfinite(a)) for a in out)
    sizes = np.diff(off)
    small = int(np.argmin(sizes))
    for c in (small, int(np.argmax(sizes)), 31):
        s = slice(off[c], off[c + 1])
        rec = bh.IterativeRecon(**kw, ctx=ctx)
        cx, cy, cz = x[s].copy(), y[s].copy(), z[s].copy()
        bh.run_(rec, (n, n, n), cx, cy, cz, w[s], copy=False)
        single = bh.reconstructed_positions(rec, cx, cy, cz, field="sum")
        for a, b in zip(single, out):
            np.testing.assert_array_equal(a, b[s])
    s = slice(off[small], off[small + 1])
    o = orc.IterativeRecon(**kw)
    ox, oy, oz = x[s].copy(), y[s].copy(), z[s].copy()
    orc.run_iterative(o, (n, n, n), ox, oy, oz, w[s])
    exp = orc.reconstructed_positions(o, ox, oy, oz, kind="sum")
    err = rms3([a[s] for a in out], exp)
    print(f"catalog {small} ({sizes[small]} tracers) against the oracle: {err:.2e} Mpc/h")
    assert err < 1e-4
