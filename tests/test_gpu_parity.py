"""GPU parity: every HIP entry point against the CPU oracle on seeded inputs.

Tolerances: density grids within 1e-5 relative L2 (FP32);
displacements and shifted positions within 1e-4 Mpc/h RMS.  Primitives that
are pure FP32 arithmetic in the reference's order (scatter, gather, Jacobi,
residual, restriction, prolongation) are held to tighter bounds.
"""
import numpy as np
import pytest

import baorec_hip as bh
import baorec_oracle as orc
from baorec_hip import synth

pytestmark = pytest.mark.gpu

F32 = np.float32


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def rms3(a, b):
    return float(np.sqrt(np.mean([np.mean((np.asarray(p, np.float64) - np.asarray(q, np.float64)) ** 2)
                                  for p, q in zip(a, b)])))


def randoms_mask(dran, thr, margin=3.0):
    """Cells whose smoothed randoms are clear of the normalisation threshold
    thr (src/recon.jl:81-89): above margin * thr, or below thr by more than
    1e-4 (there delta is 0 in every implementation).  In cells with
    thr < rho_r <= margin * thr the division by a small rho_r amplifies the
    FFT rounding of the smoothed randoms, so two FP32 implementations of the
    reference algorithm differ there by more than 1e-5 (each is that far from
    FP64)."""
    return (dran > margin * thr) | (dran < thr * (1 - 1e-4))


def assert_density_parity(gpu, ref32, ref64, mask, tol=1e-5, what=""):
    """Density parity: relative L2 against the FP32 oracle below tol on the
    masked cells, no fallback.  The error of both against the FP64 oracle is
    printed and checked separately (HIP no worse than 1.5x the FP32
    reference algorithm's own error, on all cells)."""
    e32 = rel_l2(gpu[mask], ref32[mask])
    e_gpu, e_ref = rel_l2(gpu, ref64), rel_l2(ref32, ref64)
    print(f"{what}: rel L2 vs FP32 oracle {e32:.2e} on {int(mask.sum())} of {mask.size} cells; "
          f"all cells vs FP64: HIP {e_gpu:.2e}, FP32 reference {e_ref:.2e}")
    assert e32 < tol, e32
    assert e_gpu <= 1.5 * e_ref + 1e-7, (e_gpu, e_ref)


def assert_fp64_precision(gpu, ref32, ref64, what=""):
    """Precision of a result the parity bounds do not cover directly (the
    multigrid potential of a thresholded lightcone density): its error against
    the FP64 oracle is no worse than 1.5x the FP32 reference algorithm's."""
    e_gpu, e_ref = rel_l2(gpu, ref64), rel_l2(ref32, ref64)
    print(f"{what}: vs FP64: HIP {e_gpu:.2e}, FP32 reference {e_ref:.2e}; vs FP32 oracle {rel_l2(gpu, ref32):.2e}")
    assert e_gpu <= 1.5 * e_ref + 1e-7, (e_gpu, e_ref)


def box_kw(L=1000.0, los=(0.0, 0.0, 1.0)):
    return dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[L] * 3, box_min=[0.0] * 3,
                los=None if los is None else list(los))


# ------------------------------------------------------------------ FFT

@pytest.mark.parametrize("shape", [(16, 16, 16), (16, 32, 64), (64, 16, 32), (128, 128, 128), (32, 256, 16),
                                   (16, 512, 32), (64, 32, 512), (32, 512, 512), (16, 1024, 16)])
def test_rfft3d_matches_numpy(ctx, shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(1)
    a = rng.standard_normal((nz, ny, nx)).astype(F32)
    got = bh.rfft3d(a, ctx)
    ref = np.fft.rfftn(a.astype(np.float64))
    assert rel_l2(got, ref) < 2e-6
    back = bh.irfft3d(got, (nx, ny, nz), ctx)
    assert rel_l2(back, a) < 2e-6


def test_irfft3d_ignores_selfconjugate_imaginary_parts(ctx):
    nx = ny = nz = 16
    rng = np.random.default_rng(2)
    spec = (rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1)))
    spec = spec.astype(np.complex64)
    got = bh.irfft3d(spec, (nx, ny, nz), ctx)
    ref = np.fft.irfftn(spec.astype(np.complex128), s=(nz, ny, nx), axes=(0, 1, 2))
    assert rel_l2(got, ref) < 2e-6


# ------------------------------------------------------------------ mass assignment

@pytest.mark.parametrize("grid", [(32, 32, 32), (64, 32, 16), (40, 24, 20)])
def test_cic_matches_oracle(ctx, grid):
    nx, ny, nz = grid
    rng = np.random.default_rng(3)
    n = 20000
    L = 1000.0
    x, y, z = (rng.random(n).astype(F32) * F32(L) for _ in range(3))
    # a few particles exactly on the upper face and beyond it (one-sided wrap)
    x[:5] = F32(L)
    y[5:10] = F32(L + 3.0)
    w = (rng.random(n) + 0.5).astype(F32)
    rec = bh.IterativeRecon(**box_kw(L))
    rho = np.zeros((nz, ny, nx), F32)
    gx, gy, gz = x.copy(), y.copy(), z.copy()
    bh.cic_(rho, gx, gy, gz, w, rec, wrap=True)
    ref = np.zeros((nz, ny, nx), F32)
    ox, oy, oz = x.copy(), y.copy(), z.copy()
    orc.cic(ref, ox, oy, oz, w, rec.box_size, rec.box_min, wrap=True)
    assert rel_l2(rho, ref) < 1e-6
    assert abs(rho.sum(dtype=np.float64) - w.sum(dtype=np.float64)) < 1e-4 * w.sum()
    # input mutation identical to the reference
    np.testing.assert_array_equal(gx, ox)
    np.testing.assert_array_equal(gy, oy)


def test_cic_nowrap_out_of_box_raises(ctx):
    rec = bh.IterativeRecon(**box_kw(1000.0))
    rho = np.zeros((16, 16, 16), F32)
    x = np.array([999.9], F32)  # last cell: the +1 corner is outside without wrap
    y = np.array([10.0], F32)
    z = np.array([10.0], F32)
    with pytest.raises(bh.BoundsError):
        bh.cic_(rho, x, y, z, np.ones(1, F32), rec, wrap=False)


def test_tsc_mass_and_symmetry(ctx):
    rec = bh.IterativeRecon(**box_kw(32.0), mas="tsc")
    rho = np.zeros((32, 32, 32), F32)
    rng = np.random.default_rng(4)
    n = 5000
    x, y, z = (rng.random(n).astype(F32) * F32(32.0) for _ in range(3))
    bh.cic_(rho, x, y, z, np.ones(n, F32), rec, wrap=True)
    assert abs(rho.sum(dtype=np.float64) - n) < 1e-3
    # single particle on a node: weights 1/8, 3/4, 1/8 per axis
    rho = np.zeros((32, 32, 32), F32)
    p = np.array([5.0], F32)
    bh.cic_(rho, p.copy(), p.copy(), p.copy(), np.ones(1, F32), rec, wrap=True)
    w1 = np.array([0.125, 0.75, 0.125])
    ref = np.zeros((32, 32, 32))
    ref[4:7, 4:7, 4:7] = w1[:, None, None] * w1[None, :, None] * w1[None, None, :]
    np.testing.assert_allclose(rho, ref, atol=1e-7)


def test_read_cic_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    nx, ny, nz = 32, 16, 24
    field = rng.standard_normal((nz, ny, nx)).astype(F32)
    n = 10000
    L = 500.0
    x, y, z = (rng.random(n).astype(F32) * F32(L) for _ in range(3))
    rec = bh.IterativeRecon(**box_kw(L))
    out = np.zeros(n, F32)
    bh.read_cic_(out, field, x, y, z, rec)
    ref = orc.read_cic(field, x, y, z, rec.box_size, rec.box_min, gpu_formula=True)
    assert np.max(np.abs(out - ref)) < 1e-5


@pytest.mark.parametrize("n", [100001, 4099, 7])
def test_setup_box_exact(ctx, n):
    """Extremes planted at the first and last particles and near the start
    (where the block and thread ranges of the reduction begin and end) must
    all be found, for lengths that are not multiples of the block size."""
    rng = np.random.default_rng(6)
    x, y, z = (rng.standard_normal(n).astype(F32) * 300 + 50 for _ in range(3))
    x[-1], y[0], z[-1] = 5000.0, -5000.0, -4000.0
    x[1], z[2] = -3000.0, 3500.0
    bs, bm = bh.setup_box(x, y, z, 500.0, ctx)
    rs, rm = orc.setup_box(x, y, z, 500.0)
    np.testing.assert_array_equal(bs, rs)
    np.testing.assert_array_equal(bm, rm)


# ------------------------------------------------------------------ smoothing / overdensity

@pytest.mark.parametrize("shape", [(32, 32, 32), (512, 16, 32), (16, 512, 64), (256, 32, 16), (16, 32, 1024)])
def test_smooth_matches_oracle(ctx, shape):
    """(nz, ny, nx): 512-long y and z axes take the two-columns-per-thread
    strided pass (fwd * Gaussian * inv in z, plain y)."""
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape).astype(F32)
    rec = bh.IterativeRecon(**box_kw(300.0))
    got = bh.smooth_(a.copy(), rec)
    ref = orc.smooth(a.copy(), 15.0, rec.box_size)
    assert rel_l2(got, ref) < 1e-5


def test_setup_overdensity_box(ctx):
    x, y, z, w = synth.box_catalog(50000, L=1000.0, mesh=64, seed=11)
    rec = bh.IterativeRecon(**box_kw())
    d = np.zeros((64, 64, 64), F32)
    bh.setup_overdensity_(d, rec, x.copy(), y.copy(), z.copy(), w)
    o = orc.IterativeRecon(**box_kw())
    ref = orc.setup_overdensity(np.zeros((64, 64, 64), F32), o, x.copy(), y.copy(), z.copy(), w)
    assert rel_l2(d, ref) < 1e-5


@pytest.mark.parametrize("g", [(64, 64, 64), (16, 16, 1024)])
def test_setup_overdensity_randoms(ctx, g):
    """(nz, ny, nx); 1024-point x lines run the largest smoothing x pass."""
    x, y, z, w = synth.lightcone_catalog(40000, r_min=300.0, r_max=600.0, mesh=64, seed=12)
    rx, ry, rz, rw = synth.lightcone_randoms(200000, r_min=300.0, r_max=600.0, seed=13)
    bs, bm = orc.setup_box(rx, ry, rz, 200.0)
    kw = dict(bias=2.2, f=0.75, smoothing_radius=15.0, box_size=list(bs), box_min=list(bm), los=None)
    rec = bh.IterativeRecon(**kw)
    d = np.zeros(g, F32)
    bh.setup_overdensity_(d, rec, x, y, z, w, rx, ry, rz, rw)
    ref = orc.setup_overdensity_randoms(np.zeros(g, F32), orc.IterativeRecon(**kw), x, y, z, w,
                                        rx, ry, rz, rw)
    ref64, dran, thr = orc.setup_overdensity_randoms(np.zeros(g, np.float64), orc.IterativeRecon(**kw),
                                                     x, y, z, w, rx, ry, rz, rw, return_aux=True)
    assert_density_parity(d, ref, ref64, randoms_mask(dran, thr), what="setup_overdensity randoms")


# ------------------------------------------------------------------ iterative recon end to end

@pytest.mark.parametrize("los", [(0.0, 0.0, 1.0), (1.0, 0.0, 0.0), None])
def test_run_iterative_box(ctx, los):
    n = 64
    x, y, z, w = synth.box_catalog(100000, L=1000.0, mesh=64, seed=21)
    kw = box_kw(los=los)
    rec = bh.IterativeRecon(**kw)
    gx, gy, gz = x.copy(), y.copy(), z.copy()
    d = bh.run_(rec, (n, n, n), gx, gy, gz, w)
    o = orc.IterativeRecon(**kw)
    ox, oy, oz = x.copy(), y.copy(), z.copy()
    ref = orc.run_iterative(o, (n, n, n), ox, oy, oz, w)
    assert rel_l2(d, ref) < 1e-5
    for kind in ("disp", "rsd", "sum"):
        got = bh.read_shifts(rec, gx, gy, gz, field=kind)
        exp = orc.read_shifts(o, ox, oy, oz, kind=kind)
        assert rms3(got, exp) < 1e-4, kind
    got = bh.reconstructed_positions(rec, gx, gy, gz, field="sum")
    exp = orc.reconstructed_positions(o, ox, oy, oz, kind="sum")
    assert rms3(got, exp) < 1e-4


@pytest.mark.parametrize("los,g", [((0.0, 0.0, 1.0), (16, 512, 512)), (None, (16, 512, 512)),
                                   ((0.0, 0.0, 1.0), (32, 256, 256)), (None, (1024, 16, 16)),
                                   ((0.0, 0.0, 1.0), (1024, 16, 16)), (None, (512, 32, 16))])
def test_run_iterative_box_long_axes(ctx, los, g):
    """y and z passes of length 256 and 512: the two-columns-per-thread
    kernel for single-output passes, the one-column kernel for multi-output
    ones; x lines of 512 and 1024 points (the x passes' LDS tables, the
    local-LOS iteration pass at its largest line)."""
    x, y, z, w = synth.box_catalog(100000, L=1000.0, mesh=64, seed=23)
    kw = box_kw(los=los)
    rec = bh.IterativeRecon(**kw)
    gx, gy, gz = x.copy(), y.copy(), z.copy()
    d = bh.run_(rec, g, gx, gy, gz, w)
    o = orc.IterativeRecon(**kw)
    ox, oy, oz = x.copy(), y.copy(), z.copy()
    ref = orc.run_iterative(o, g, ox, oy, oz, w)
    assert rel_l2(d, ref) < 1e-5
    got = bh.reconstructed_positions(rec, gx, gy, gz, field="sum")
    exp = orc.reconstructed_positions(o, ox, oy, oz, kind="sum")
    assert rms3(got, exp) < 1e-4


def noncubic_case(seed):
    """A catalog in a (1000, 800, 600) Mpc/h box on a (64, 32, 128) grid:
    per-axis cell sizes, k and x tables."""
    x, y, z, w = synth.box_catalog(100000, L=1000.0, mesh=64, seed=seed)
    y = (y * np.float32(0.8)).astype(F32)
    z = (z * np.float32(0.6)).astype(F32)
    kw = dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0, 800.0, 600.0], box_min=[0.0] * 3)
    return (x, y, z, w), kw, (64, 32, 128)


@pytest.mark.parametrize("los", [(0.0, 0.6, 0.8), None])
def test_run_iterative_noncubic(ctx, los):
    (x, y, z, w), kw, g = noncubic_case(25)
    kw["los"] = None if los is None else list(los)
    rec = bh.IterativeRecon(**kw)
    gx, gy, gz = x.copy(), y.copy(), z.copy()
    d = bh.run_(rec, g, gx, gy, gz, w)
    o = orc.IterativeRecon(**kw)
    ox, oy, oz = x.copy(), y.copy(), z.copy()
    ref = orc.run_iterative(o, g, ox, oy, oz, w)
    assert d.shape == g[::-1]
    assert rel_l2(d, ref) < 1e-5
    got = bh.reconstructed_positions(rec, gx, gy, gz, field="sum")
    exp = orc.reconstructed_positions(o, ox, oy, oz, kind="sum")
    assert rms3(got, exp) < 1e-4


def test_run_multigrid_noncubic(ctx):
    (x, y, z, w), kw, g = noncubic_case(26)
    kw["los"] = None
    rec = bh.MultigridRecon(**kw)
    gx, gy, gz = x.copy(), y.copy(), z.copy()
    phi = bh.run_(rec, g, gx, gy, gz, w)
    o = orc.MultigridRecon(**kw)
    ox, oy, oz = x.copy(), y.copy(), z.copy()
    ref = orc.run_multigrid(o, g, ox, oy, oz, w)
    assert rel_l2(phi, ref) < 1e-5
    got = bh.reconstructed_positions(rec, gx, gy, gz, field="sum")
    exp = orc.reconstructed_positions(o, ox, oy, oz, kind="sum")
    assert rms3(got, exp) < 1e-4


def lightcone_case(seed_d, seed_r):
    """Lightcone test catalogs whose smoothed randoms stay >= 1e-4 (relative)
    away from the normalisation threshold in every cell, so that the
    thresholded overdensity is not decided by summation order."""
    x, y, z, w = synth.lightcone_catalog(50000, r_min=500.0, r_max=900.0, mesh=64, seed=seed_d)
    rx, ry, rz, rw = synth.lightcone_randoms(250000, r_min=500.0, r_max=900.0, seed=seed_r)
    bs, bm = orc.setup_box(rx, ry, rz, 200.0)
    o = orc.IterativeRecon(bias=2.2, f=0.75, smoothing_radius=15.0, los=None, box_size=bs, box_min=bm)
    _, dran, thr = orc.setup_overdensity_randoms(np.zeros((64, 64, 64)), o, x, y, z, w, rx, ry, rz, rw,
                                                 return_aux=True)
    assert np.abs(dran / thr - 1).min() > 1e-4
    return (x, y, z, w), (rx, ry, rz, rw)


def test_run_iterative_lightcone(ctx):
    n = 64
    (x, y, z, w), (rx, ry, rz, rw) = lightcone_case(71, 72)
    kw = dict(bias=2.2, f=0.75, smoothing_radius=15.0, los=None)
    rec = bh.IterativeRecon(**kw, box_pad=200.0)
    d = bh.run_(rec, (n, n, n), x, y, z, w, rx, ry, rz, rw)
    o = orc.IterativeRecon(**kw)
    ref = orc.run_iterative(o, (n, n, n), x, y, z, w, rx, ry, rz, rw, box_pad=200.0)
    ref64 = orc.run_iterative(orc.IterativeRecon(**kw), (n, n, n), x, y, z, w, rx, ry, rz, rw, box_pad=200.0,
                              dtype=np.float64)
    np.testing.assert_array_equal(rec.box_size, o.box_size)
    np.testing.assert_array_equal(rec.box_min, o.box_min)
    _, dran, thr = orc.setup_overdensity_randoms(np.zeros((n, n, n)), orc.IterativeRecon(**kw, box_size=o.box_size,
                                                 box_min=o.box_min), x, y, z, w, rx, ry, rz, rw, return_aux=True)
    assert_density_parity(d, ref, ref64, randoms_mask(dran, thr), what="iterative lightcone delta")
    got = bh.reconstructed_positions(rec, x, y, z, field="sum")
    exp = orc.reconstructed_positions(o, x, y, z, kind="sum")
    assert rms3(got, exp) < 1e-4
    # randoms read-outs reuse the cached displacement grids
    got = bh.reconstructed_positions(rec, rx, ry, rz, field="disp")
    exp = orc.reconstructed_positions(o, rx, ry, rz, kind="disp")
    assert rms3(got, exp) < 1e-4


def test_read_shifts_explicit_field(ctx):
    rng = np.random.default_rng(41)
    field = rng.standard_normal((32, 32, 32)).astype(F32)
    kw = box_kw(500.0)
    x, y, z = (rng.random(5000).astype(F32) * F32(500.0) for _ in range(3))
    for cls, ocls in ((bh.IterativeRecon, orc.IterativeRecon), (bh.MultigridRecon, orc.MultigridRecon)):
        rec = cls(**kw)
        got = bh.read_shifts(rec, x, y, z, field, field="sum")
        exp = orc.read_shifts(ocls(**kw), x, y, z, field=field, kind="sum")
        scale = max(np.abs(exp[0]).max(), 1e-3)
        assert rms3(got, exp) < 1e-5 * scale + 1e-6


# ------------------------------------------------------------------ multigrid

def mg_rec(los=None, L=1000.0):
    return dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[L] * 3, box_min=[-L / 2, -L / 2, -L / 2],
                los=None if los is None else list(los))


@pytest.mark.parametrize("shape", [(16, 32, 32), (32, 16, 128)])   # (nz, ny, nx); the second is
@pytest.mark.parametrize("los", [None, (0.0, 0.0, 1.0)])            # a multiple of the plane-marching tile
def test_mg_primitives(ctx, los, shape):
    rng = np.random.default_rng(51)
    kw = mg_rec(los)
    rec = bh.MultigridRecon(**kw)
    o = orc.MultigridRecon(**kw)
    v = rng.standard_normal(shape).astype(F32)
    f = rng.standard_normal(shape).astype(F32)
    got = bh.jacobi_(v.copy(), f, rec)
    exp = orc.mg_jacobi(v.copy(), f, o.box_size, o.box_min, o.beta, F32(0.4), 5, los=o.los)
    assert np.max(np.abs(got - exp)) < 1e-6 * np.abs(exp).max()
    r = np.zeros_like(v)
    bh.residual_(r, v, f, rec)
    exp = orc.mg_residual(v, f, o.box_size, o.box_min, o.beta, los=o.los)
    assert np.max(np.abs(r - exp)) < 1e-6 * np.abs(exp).max()
    c = np.zeros(tuple(a // 2 for a in shape), F32)
    bh.reduce_(c, v, ctx)
    np.testing.assert_allclose(c, orc.mg_reduce(v), rtol=0, atol=1e-6)
    fine = np.zeros(shape, F32)
    bh.prolong_(fine, c, add=False, ctx=ctx)
    np.testing.assert_array_equal(fine, orc.mg_prolong(c))
    fine2 = v.copy()
    bh.prolong_(fine2, c, add=True, ctx=ctx)
    np.testing.assert_allclose(fine2, v + orc.mg_prolong(c), rtol=0, atol=1e-6)


@pytest.mark.parametrize("n", [64, 128, 256])
def test_mg_residual_local_los_against_analytic_operator(ctx, n):
    """The HIP residual with a local LOS, applied to a Gaussian blob, against
    the continuum operator lap + beta [(rhat.grad)^2 + (2/|x|) rhat.grad]
    evaluated in closed form: within 2 % of the FP64 oracle's own
    truncation error (which falls 4x per halving of the cell) plus twice the
    FP32 oracle's rounding of the stencil, so the GPU's
    first-derivative term, cross terms and position-dependent coefficients
    are pinned without reference to the oracle's code."""
    from test_oracle import BLOB, gaussian_blob_operator
    phi, lphi = gaussian_blob_operator(n, BLOB["L"], BLOB["box_min"], BLOB["beta"], BLOB["centre"], BLOB["s"])
    rec = bh.MultigridRecon(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[BLOB["L"]] * 3,
                            box_min=list(BLOB["box_min"]), los=None, beta=BLOB["beta"])
    v = phi.astype(F32)
    r = np.zeros_like(v)
    bh.residual_(r, v, np.zeros_like(v), rec)
    e_gpu = rel_l2(r.astype(np.float64), lphi)
    r64 = orc.mg_residual(phi, np.zeros_like(phi), [BLOB["L"]] * 3, BLOB["box_min"], BLOB["beta"], los=None)
    e64 = rel_l2(r64, lphi)
    r32 = orc.mg_residual(v, np.zeros_like(v), [BLOB["L"]] * 3, BLOB["box_min"], F32(BLOB["beta"]), los=None)
    rnd = rel_l2(r32.astype(np.float64), r64)   # FP32 rounding of the stencil, grows as 1/h^2
    print(f"{n}^3: HIP {e_gpu:.4e}, FP64 oracle {e64:.4e} against the continuum operator; "
          f"FP32 rounding {rnd:.1e}")
    assert abs(e_gpu - e64) < 0.02 * e64 + 2 * rnd


@pytest.mark.parametrize("los", [None, (0.0, 0.0, 1.0)])
def test_mg_vcycle_and_fmg(ctx, los):
    rng = np.random.default_rng(52)
    kw = mg_rec(los)
    rec = bh.MultigridRecon(**kw)
    o = orc.MultigridRecon(**kw)
    f = rng.standard_normal((32, 32, 32)).astype(F32)
    v = rng.standard_normal((32, 32, 32)).astype(F32)
    got = bh.vcycle_(v.copy(), f, rec)
    exp = orc.mg_vcycle(v.copy(), f, o.box_size, o.box_min, o.beta, F32(0.4), 5, los=o.los)
    assert rel_l2(got, exp) < 1e-5
    got = bh.fmg(f, rec)
    exp = orc.fmg(f, o.box_size, o.box_min, o.beta, F32(0.4), 5, 6, los=o.los)
    assert rel_l2(got, exp) < 1e-5


def test_run_multigrid_box(ctx):
    n = 64
    x, y, z, w = synth.box_catalog(100000, L=1000.0, mesh=64, seed=61)
    kw = box_kw(los=None)
    kw["box_min"] = [0.0, 0.0, 0.0]
    rec = bh.MultigridRecon(**kw)
    gx, gy, gz = x.copy(), y.copy(), z.copy()
    phi = bh.run_(rec, (n, n, n), gx, gy, gz, w)
    o = orc.MultigridRecon(**kw)
    ox, oy, oz = x.copy(), y.copy(), z.copy()
    ref = orc.run_multigrid(o, (n, n, n), ox, oy, oz, w)
    assert rel_l2(phi, ref) < 1e-5
    got = bh.reconstructed_positions(rec, gx, gy, gz, field="sum")
    exp = orc.reconstructed_positions(o, ox, oy, oz, kind="sum")
    assert rms3(got, exp) < 1e-4


def test_run_multigrid_lightcone(ctx):
    n = 64
    (x, y, z, w), (rx, ry, rz, rw) = lightcone_case(71, 72)
    kw = dict(bias=2.2, f=0.75, smoothing_radius=15.0, los=None)
    rec = bh.MultigridRecon(**kw, box_pad=200.0)
    phi = bh.run_(rec, (n, n, n), x, y, z, w, rx, ry, rz, rw)
    o = orc.MultigridRecon(**kw)
    ref = orc.run_multigrid(o, (n, n, n), x, y, z, w, rx, ry, rz, rw, box_pad=200.0)
    ref64 = orc.run_multigrid(orc.MultigridRecon(**kw), (n, n, n), x, y, z, w, rx, ry, rz, rw, box_pad=200.0,
                              dtype=np.float64)
    assert_fp64_precision(phi, ref, ref64, what="multigrid lightcone phi")
    got = bh.reconstructed_positions(rec, x, y, z, field="sum")
    exp = orc.reconstructed_positions(o, x, y, z, kind="sum")
    assert rms3(got, exp) < 1e-4


def test_device_pointer_path(ctx):
    from baorec_hip.devmem import DeviceArray
    n = 32
    x, y, z, w = synth.box_catalog(20000, L=1000.0, mesh=32, seed=81)
    kw = box_kw()
    rec = bh.IterativeRecon(**kw)
    hx, hy, hz = x.copy(), y.copy(), z.copy()
    dh = bh.run_(rec, (n, n, n), hx, hy, hz, w)
    ph = bh.reconstructed_positions(rec, hx, hy, hz, field="sum")
    dev = [DeviceArray.from_numpy(a.copy()) for a in (x, y, z, w)]
    rec2 = bh.IterativeRecon(**kw)
    dd = bh.run_(rec2, (n, n, n), *dev)
    pd = bh.reconstructed_positions(rec2, dev[0], dev[1], dev[2], field="sum")
    assert rel_l2(dd, dh) < 1e-6
    assert rms3([t.numpy() for t in pd], ph) < 1e-5
    np.testing.assert_array_equal(dev[0].numpy(), hx)


def test_no_result_error(ctx):
    c = bh.Context(0)
    rec = bh.IterativeRecon(**box_kw(), ctx=c)
    with pytest.raises(bh.BaorecError):
        bh.reconstructed_positions(rec, np.ones(3, F32), np.ones(3, F32), np.ones(3, F32))
    c.close()


def test_stale_result_of_an_earlier_recon(ctx):
    """Two recons on one context, the second run! on a larger grid: the first
    recon's read-outs still use its own result (its host copy), as the
    reference reads each recon's result_cache, and a device-only cache of a
    replaced result refuses to copy instead of copying the newer grid."""
    c = bh.Context(0)
    x, y, z, w = synth.box_catalog(30000, L=1000.0, mesh=32, seed=81)
    rec1 = bh.IterativeRecon(**box_kw(), ctx=c)
    d1 = bh.run_(rec1, (32, 32, 32), x.copy(), y.copy(), z.copy(), w)
    p1 = bh.reconstructed_positions(rec1, x, y, z, field="sum")
    rec2 = bh.IterativeRecon(**box_kw(), ctx=c)
    lazy = bh.run_(rec2, (64, 64, 64), x.copy(), y.copy(), z.copy(), w, copy=False)
    assert not rec1.result_cache.is_current() and lazy.is_current()
    np.testing.assert_array_equal(rec1.result_cache.numpy(), d1)
    p1b = bh.reconstructed_positions(rec1, x, y, z, field="sum")
    assert rms3(p1b, p1) < 1e-6
    bh.run_(rec1, (16, 16, 16), x.copy(), y.copy(), z.copy(), w)
    with pytest.raises(bh.BaorecError, match="stale"):
        lazy.numpy()
    with pytest.raises(bh.BaorecError, match="stale"):
        bh.reconstructed_positions(rec2, x, y, z, field="sum")
    c.close()


# ------------------------------------------------------------------ RBGS schedule (extension)

def fixed_los_symbol(n, L, beta, los):
    """Fourier symbol of the discrete multigrid operator for a fixed LOS
    (constant coefficients, so the FFT diagonalises it exactly)."""
    cell = np.array(L, np.float64) / np.array(n)
    p = np.array(los, dtype=F32).astype(np.float64) / cell
    g = beta / np.sum(cell ** 2 * p ** 2)
    gp = 1 / cell ** 2 + g * p ** 2
    th = [2 * np.pi * np.fft.fftfreq(n[a]) for a in range(3)]
    T = [th[0][None, None, :], th[1][None, :, None], th[2][:, None, None]]
    A = sum(gp[a] * (2 - 2 * np.cos(T[a])) for a in range(3))
    for i, j in ((0, 1), (0, 2), (1, 2)):
        A = A + 2 * g * p[i] * p[j] * np.sin(T[i]) * np.sin(T[j])
    return A


@pytest.mark.parametrize("fp64_coarse_max", [8, 0])
def test_mg_rbgs_fixed_los_matches_exact_inverse(ctx, fp64_coarse_max):
    """Multicolour Gauss-Seidel V-cycles (2+2 sweeps) to relative residual
    1e-6 reproduce the exact inverse of the discrete operator.  Tolerance:
    rel L2 3e-5 on the zero-mean potential (a 1e-6 residual bounds the error
    by the operator's condition number over the modes it leaves)."""
    n = (64, 64, 64)
    kw = mg_rec((0.0, 0.6, 0.8))
    rng = np.random.default_rng(91)
    f = rng.standard_normal(n[::-1]).astype(F32)
    rec = bh.MultigridRecon(**kw, schedule="rbgs", tolerance=1e-6, fp64_coarse_max=fp64_coarse_max)
    phi = bh.fmg(f, rec)
    assert rec.final_residual <= 1e-6 and 0 < rec.vcycles_done < rec.max_vcycles
    o = orc.MultigridRecon(**kw)
    A = fixed_los_symbol(n, kw["box_size"], o.beta, kw["los"])
    fk = np.fft.fftn(f.astype(np.float64) - f.astype(np.float64).mean())
    exact = np.fft.ifftn(np.where(A > 1e-12, fk / np.where(A > 1e-12, A, 1), 0)).real
    got = phi.astype(np.float64) - phi.mean(dtype=np.float64)
    err = rel_l2(got, exact)
    print(f"rbgs fixed LOS: {rec.vcycles_done} V-cycles, residual {rec.final_residual:.2e}, rel L2 {err:.2e}")
    assert err < 3e-5


def test_mg_rbgs_fixed_los_refined_to_1e6_at_256(ctx):
    """256^3 with a smooth right-hand side (a smoothed lognormal density):
    the FP32 residual of A v alone is rounding-limited near 1e-4 of ||f||
    here, so reaching 1e-6 needs the FP64-refined solution.  It must also
    equal the exact inverse of the discrete operator (FFT of its symbol)."""
    n = (256, 256, 256)
    kw = mg_rec((0.0, 0.0, 1.0))
    x, y, z, w = synth.box_catalog(400000, L=1000.0, box_min=-500.0, mesh=128, seed=95)
    rec = bh.MultigridRecon(**kw, schedule="rbgs", tolerance=1e-6)
    f = bh.setup_overdensity_(np.zeros(n[::-1], F32), rec, x.copy(), y.copy(), z.copy(), w)
    phi = bh.fmg(f, rec)
    print(f"rbgs refined 256^3: {rec.vcycles_done} V-cycles, residual {rec.final_residual:.2e}")
    assert rec.final_residual <= 1e-6 and 0 < rec.vcycles_done < 20
    o = orc.MultigridRecon(**kw)
    A = fixed_los_symbol(n, kw["box_size"], o.beta, kw["los"])
    fk = np.fft.fftn(f.astype(np.float64) - f.astype(np.float64).mean())
    exact = np.fft.ifftn(np.where(A > 1e-12, fk / np.where(A > 1e-12, A, 1), 0)).real
    got = phi.astype(np.float64) - phi.mean(dtype=np.float64)
    err = rel_l2(got, exact)
    print(f"rel L2 to the exact inverse {err:.2e}")
    assert err < 1e-5


def test_mg_rbgs_spectral_warm_start_fixed_los(ctx):
    """With a fixed LOS the spectral seed is the exact inverse of the discrete
    operator up to FP32 rounding: the solve starts near the tolerance, needs
    far fewer V-cycles than from v = 0, and reaches the same solution."""
    n = (256, 256, 256)
    kw = mg_rec((0.0, 0.0, 1.0))
    x, y, z, w = synth.box_catalog(400000, L=1000.0, box_min=-500.0, mesh=128, seed=95)
    cold = bh.MultigridRecon(**kw, schedule="rbgs", tolerance=1e-6)
    f = bh.setup_overdensity_(np.zeros(n[::-1], F32), cold, x.copy(), y.copy(), z.copy(), w)
    phi_cold = bh.fmg(f, cold)
    warm = bh.MultigridRecon(**kw, schedule="rbgs", tolerance=1e-6, warm_start="spectral")
    phi_warm = bh.fmg(f, warm)
    print(f"cold: {cold.vcycles_done} V-cycles to {cold.final_residual:.2e}; warm: seed "
          f"{warm.seed_residual:.2e}, {warm.vcycles_done} V-cycles to {warm.final_residual:.2e}")
    assert abs(cold.seed_residual - 1.0) < 1e-5
    assert warm.seed_residual < 1e-3
    assert warm.final_residual <= 1e-6 and warm.vcycles_done <= cold.vcycles_done // 2
    assert rel_l2(phi_warm, phi_cold) < 1e-5


@pytest.mark.parametrize("los", [None, (0.0, 0.6, 0.8)])
def test_mg_rbgs_spectral_warm_start_matches_cold(ctx, los):
    """Local LOS (the seed freezes the box-centre LOS) and a fixed oblique one:
    the warm-started solve stops at the same solution, in no more V-cycles."""
    n = 64
    kw = mg_rec(los)
    kw["box_min"] = [200.0, 300.0, 100.0]
    rng = np.random.default_rng(97)
    f = bh.smooth_(rng.standard_normal((n, n, n)).astype(F32), bh.IterativeRecon(**{**kw, "los": None}))
    out = {}
    for ws in ("none", "spectral"):
        rec = bh.MultigridRecon(**kw, schedule="rbgs", tolerance=1e-6, warm_start=ws)
        out[ws] = (bh.fmg(f, rec), rec)
        print(f"{ws}: seed {rec.seed_residual:.2e}, {rec.vcycles_done} V-cycles to {rec.final_residual:.2e}")
    (pc, rc), (pw, rw) = out["none"], out["spectral"]
    assert rw.seed_residual < rc.seed_residual
    assert rw.vcycles_done <= rc.vcycles_done
    assert rw.final_residual <= max(1e-6, 1.5 * rc.final_residual)
    assert rel_l2(pw - pw.mean(), pc - pc.mean()) < 1e-3


@pytest.mark.parametrize("n,box_min", [(32, (200.0, 300.0, 100.0)), (32, (-500.0, -500.0, -500.0)),
                                        (64, (200.0, 300.0, 100.0))])
def test_mg_rbgs_local_los_manufactured_solution(ctx, n, box_min):
    """Local (radial) LOS, observer outside and inside the box: f = A v_true
    with A the oracle's operator (FP64), so f lies in A's range; the RBGS
    V-cycles must reach relative residual 1e-6 and return v_true up to a
    constant.  At 64^3 the finest level runs the plane-marching kernels."""
    from scipy.ndimage import gaussian_filter
    kw = dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=list(box_min), los=None)
    o = orc.MultigridRecon(**kw)
    rng = np.random.default_rng(93)
    vt = gaussian_filter(rng.standard_normal((n, n, n)), 1.5, mode="wrap")
    f = -orc.mg_residual(vt, np.zeros_like(vt), o.box_size, o.box_min, o.beta, los=None)
    rec = bh.MultigridRecon(**kw, schedule="rbgs", tolerance=1e-6)
    phi = bh.fmg(f.astype(F32), rec)
    assert rec.final_residual <= 1e-6, (rec.final_residual, rec.vcycles_done)
    err = rel_l2(phi.astype(np.float64) - phi.mean(dtype=np.float64), vt - vt.mean())
    print(f"rbgs local LOS: {rec.vcycles_done} V-cycles, residual {rec.final_residual:.2e}, rel L2 {err:.2e}")
    assert err < 3e-5


@pytest.mark.parametrize("box_min", [(200.0, 300.0, 100.0), (-500.0, -500.0, -500.0)])
def test_mg_rbgs_local_los_random_f_reaches_tolerance(ctx, box_min):
    """Random f with a local LOS is not in the range of the non-symmetric
    operator; the RBGS solve subtracts the constant alpha that makes it
    consistent and reaches 1e-6.  alpha must be (w . f) / (w . 1), w the left
    null vector of the oracle's operator (sparse LU), and v must solve the
    FP64 oracle's system for f - alpha to FP32 rounding."""
    from test_oracle import left_null_vector, operator_matrix
    n = 32
    kw = dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=list(box_min), los=None)
    o = orc.MultigridRecon(**kw)
    rng = np.random.default_rng(92)
    f = rng.standard_normal((n, n, n)).astype(F32)
    rec = bh.MultigridRecon(**kw, schedule="rbgs", tolerance=1e-6)
    v = bh.fmg(f, rec)
    M = operator_matrix(n, o.box_size[0], list(o.box_min), float(o.beta))
    w = left_null_vector(M)
    f64 = f.astype(np.float64)
    alpha = float(w @ f64.ravel()) / float(w.sum())
    res = np.linalg.norm(M @ v.astype(np.float64).ravel() - (f64.ravel() - alpha)) / np.linalg.norm(f64 - f64.mean())
    floor = np.linalg.norm(M @ v.astype(np.float64).ravel() - (f64.ravel() - f64.mean()))
    print(f"rbgs local LOS: {rec.vcycles_done} V-cycles to {rec.final_residual:.2e}; shift {rec.rhs_shift:.6e} "
          f"vs (w.f)/(w.1) {alpha:.6e} (mean f {f64.mean():.3e}); FP64 residual of f - alpha {res:.2e}")
    assert rec.final_residual <= 1e-6 and rec.vcycles_done < 40
    assert abs(rec.rhs_shift - alpha) < 1e-6 * f64.std()
    assert res < 1e-5
    assert floor > 10 * res   # the plain mean would not have been consistent


# ------------------------------------------------------------------ edge cases

def test_empty_catalog_gives_nan_overdensity(ctx):
    """No tracers: mean density 0, so ((rho/mean) - 1)/b is NaN everywhere,
    as in the reference (src/recon.jl:54-55)."""
    n = 16
    e = np.zeros(0, F32)
    rec = bh.IterativeRecon(**box_kw())
    delta = bh.setup_overdensity_(np.zeros((n, n, n), F32), rec, e, e.copy(), e.copy(), e.copy())
    o = orc.IterativeRecon(**box_kw())
    with np.errstate(invalid="ignore", divide="ignore"):
        ref = orc.setup_overdensity(np.zeros((n, n, n), F32), o, e, e.copy(), e.copy(), e.copy())
    assert np.isnan(ref).all() and np.isnan(delta).all()


# ------------------------------------------------------------------ full size (512^3)

def test_full_size_fft_roundtrip_and_parseval(ctx):
    """512^3: irfft(rfft(x)) == x to FP32 rounding, and Parseval."""
    n = 512
    rng = np.random.default_rng(93)
    x = rng.standard_normal((n, n, n), dtype=F32)
    X = bh.rfft3d(x, ctx=ctx)
    back = bh.irfft3d(X, (n, n, n), ctx=ctx)
    assert rel_l2(back, x) < 2e-6
    w = np.full(X.shape[-1], 2.0)
    w[0] = w[-1] = 1.0                                    # DC and Nyquist planes counted once
    p_spec = float(np.sum(w * np.sum(np.abs(X.astype(np.complex128)) ** 2, axis=(0, 1))))
    p_real = float(np.sum(x.astype(np.float64) ** 2)) * n ** 3
    assert abs(p_spec / p_real - 1) < 1e-5


def test_full_size_lightcone_mass_and_determinism(ctx):
    """The 512^3 lightcone workload (2M tracers, 20M randoms): CIC conserves
    mass on the full grid, and the whole reconstruction is bitwise
    reproducible (fixed-point scatter sums, ordered FP64 reductions)."""
    n = 512
    data = synth.lightcone_catalog(2_000_000, r_min=1000.0, r_max=1800.0, mesh=256, seed=2)
    rand = synth.lightcone_randoms(20_000_000, r_min=1000.0, r_max=1800.0, seed=3)
    ext = max(float(np.max(a) - np.min(a)) for a in rand[:3])
    kw = dict(bias=2.2, f=0.75, smoothing_radius=15.0, los=None)
    outs = []
    for _ in range(2):
        rec = bh.IterativeRecon(**kw, box_pad=0.1 * ext, ctx=ctx)
        bh.run_(rec, (n, n, n), *data, *rand, copy=False)
        outs.append(bh.reconstructed_positions(rec, *data[:3], field="sum"))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
    shift = np.sqrt(np.mean([(np.asarray(p, np.float64) - q) ** 2 for p, q in zip(outs[0], data[:3])]))
    assert np.all(np.isfinite(outs[0])) and 0.5 < shift < 50.0
    rho = np.zeros((n, n, n), F32)
    bh.cic_(rho, *(a.copy() for a in data), rec, wrap=False)
    assert abs(float(np.sum(rho, dtype=np.float64)) / float(np.sum(data[3], dtype=np.float64)) - 1) < 1e-6


# ------------------------------------------------------------------ sky <-> Cartesian

def ulp_err(got, ref):
    """Largest |got - ref| in units of the FP32 spacing at ref."""
    ref = np.asarray(ref, F32)
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref.astype(np.float64)) /
                        np.spacing(np.abs(ref)).astype(np.float64)))


@pytest.mark.parametrize("num", [301, 100000])
def test_comoving_distance_table_matches_quadrature(ctx, num):
    """The GPU table (FP64 Gauss-Legendre per interval + ordered prefix sum)
    against adaptive quadrature at the reference's knots, incl. the default
    100000-knot table."""
    c = bh.Cosmology(z_tab_num=num)
    o = orc.Cosmology(z_tab_num=num)
    zk = np.linspace(0.0, 3.0, num)
    idx = np.unique(np.concatenate([[0, 1, num - 1], np.random.default_rng(94).integers(0, num, 60)]))
    got = bh.comoving_distance(zk[idx], c, ctx=ctx)
    ref = orc.comoving_distance(o, zk[idx])
    assert got[0] == 0.0
    np.testing.assert_allclose(got[1:], ref[1:], rtol=1e-9)


def test_comoving_distance_einstein_de_sitter_gpu(ctx):
    h = 0.5
    c = bh.Cosmology(omega_b=0.0625, omega_c=0.1875, h=h, T_cmb=0.0, z_tab_num=3001)
    z = np.linspace(0.0, 3.0, 3001)[::100]
    exact = 2 * orc.C_KMS / (100 * h) * (1 - 1 / np.sqrt(1 + z))
    np.testing.assert_allclose(bh.comoving_distance(z, c, ctx=ctx), exact, rtol=1e-9, atol=1e-9)


def sky_case(n=200000, seed=95):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0, 360, n).astype(F32), rng.uniform(-90, 90, n).astype(F32),
            rng.uniform(0.0, 3.0, n).astype(F32))


def test_sky_to_cartesian_matches_oracle(ctx):
    """Within 1 FP32 ulp of the oracle (same Float32 rounding points; the
    tables differ by quadrature error ~1e-12)."""
    c, o = bh.Cosmology(z_tab_num=2001), orc.Cosmology(z_tab_num=2001)
    table = orc.distance_table(o)
    ra, dec, red = sky_case()
    got = bh.sky_to_cartesian(ra, dec, red, c, ctx=ctx)
    ref = orc.sky_to_cartesian(ra, dec, red, o, table)
    for g, r in zip(got, ref):
        assert ulp_err(g, r) <= 1.0
    assert rms3(got, ref) < 1e-4


def test_cartesian_to_sky_matches_oracle(ctx):
    c, o = bh.Cosmology(z_tab_num=2001), orc.Cosmology(z_tab_num=2001)
    table = orc.distance_table(o)
    ra, dec, red = sky_case(seed=96)
    x, y, z = orc.sky_to_cartesian(ra, dec, red, o, table)
    got = bh.cartesian_to_sky(x, y, z, c, ctx=ctx)
    ref = orc.cartesian_to_sky(x, y, z, o, table)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g.astype(np.float64) - r)) <= np.max(np.spacing(np.abs(r))) + 1e-12
    assert np.all(got[0] <= 0) and np.all(got[0] > -360)


def test_sky_outside_table_raises(ctx):
    c = bh.Cosmology(z_tab_num=1001)
    with pytest.raises(bh.BoundsError):
        bh.sky_to_cartesian(np.zeros(3, F32), np.zeros(3, F32), np.array([0.5, 3.5, 1.0], F32), c, ctx=ctx)
    with pytest.raises(bh.BoundsError):
        bh.cartesian_to_sky(np.array([1e5], F32), np.zeros(1, F32), np.zeros(1, F32), c, ctx=ctx)


# ------------------------------------------------------------------ power spectrum

def test_power_multipoles_plane_wave_gpu(ctx):
    """The lattice plane wave of the oracle's closed-form test, on the GPU."""
    n, L, A = 16, 160.0, 0.25
    g = np.arange(n, dtype=F32) * F32(L / n)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    x, y, z = X.ravel().copy(), Y.ravel().copy(), Z.ravel().copy()
    w = (1 + A * np.cos(2 * np.pi / L * x.astype(np.float64))).astype(F32)
    kf = 2 * np.pi / L
    got = bh.power_multipoles(x, y, z, w, (n, n, n), L, kmin=0.8 * kf, dk=0.3 * kf, nbins=1, ctx=ctx)
    ref = orc.power_multipoles(x, y, z, w, (n, n, n), (L, L, L), kmin=0.8 * kf, dk=0.3 * kf, nbins=1)
    assert got["nmodes"][0] == 6
    for q in ("p0", "p2", "p4"):
        np.testing.assert_allclose(got[q], ref[q], rtol=2e-5)


@pytest.mark.parametrize("interlace", [True, False])
def test_power_multipoles_matches_oracle(ctx, interlace):
    """Clustered box catalog, 128^3: every bin below Nyquist within 1e-4
    relative of the oracle (FP32 grid FFT against FP64)."""
    n, L = 128, 1000.0
    x, y, z, w = synth.box_catalog(200000, L=L, mesh=128, seed=97)
    got = bh.power_multipoles(x, y, z, w, (n, n, n), L, interlace=interlace, dk=0.01, nbins=40, ctx=ctx)
    ref = orc.power_multipoles(x, y, z, w, (n, n, n), (L, L, L), interlace=interlace, dk=0.01, nbins=40)
    np.testing.assert_array_equal(got["nmodes"], ref["nmodes"])
    np.testing.assert_allclose(got["k"], ref["k"], rtol=1e-12)
    scale = np.abs(ref["p0"]) + L ** 3 / len(x)
    for q in ("p0", "p2", "p4"):
        assert np.max(np.abs(got[q] - ref[q]) / scale) < 1e-4, q


def test_power_multipoles_full_size_poisson(ctx):
    """512^3, 2M unclustered points: P0 - shot noise within 5 sigma of zero
    below 0.7 k_Nyquist (the TSC-interlaced estimator is unbiased there)."""
    n, L, N = 512, 1000.0, 2_000_000
    rng = np.random.default_rng(98)
    x, y, z = (rng.uniform(0, L, N).astype(F32) for _ in range(3))
    out = bh.power_multipoles(x, y, z, None, (n, n, n), L, dk=0.02, nbins=50, ctx=ctx)
    shot = L ** 3 / N
    ok = (out["nmodes"] > 100) & (out["k"] < 0.7 * np.pi * n / L)
    sig = shot / np.sqrt(out["nmodes"][ok])
    assert ok.sum() > 30
    assert np.all(np.abs(out["p0"][ok]) < 5 * sig + 0.01 * shot)
