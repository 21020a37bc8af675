"""The multigrid smoothers and fused level passes on the GPU: the one-launch 4-colour sweep
(k_mg_gs4) against the per-colour launches (bit for bit) and against the
oracle's multicolour Gauss-Seidel (FP64, within FP32 rounding); the
temporally blocked damped Jacobi against single sweeps (bit for bit)."""
import numpy as np
import pytest

import baorec_hip as bh
import baorec_oracle as orc

pytestmark = pytest.mark.gpu

F32 = np.float32


def recon(shape, los, ctx):
    nz, ny, nx = shape
    return bh.MultigridRecon(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[4.0 * nx, 4.0 * ny, 4.0 * nz],
                             box_min=[200.0, -300.0, 500.0], los=los, ctx=ctx)


def fields(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape).astype(F32), rng.standard_normal(shape).astype(F32)


@pytest.mark.parametrize("shape,los,sweeps", [
    ((8, 16, 64), None, 1),
    ((24, 32, 64), [0.0, 0.0, 1.0], 2),
    ((40, 48, 192), None, 3),
    ((256, 256, 256), None, 2),
    ((512, 512, 512), None, 1),
])
def test_fused_sweep_bitwise_per_colour(ctx, shape, los, sweeps):
    """Shapes: one workgroup column, halos wrapping onto the tile itself
    (64 x 16), non-cubic sides, and the solver's 256^3 and 512^3 levels."""
    v, f = fields(shape, sum(shape))
    rec = recon(shape, los, ctx)
    a = bh.gauss_seidel_(v.copy(), f, rec, sweeps=sweeps, fused=True)
    b = bh.gauss_seidel_(v.copy(), f, rec, sweeps=sweeps, fused=False)
    assert np.all(np.isfinite(a))
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("los", [None, [0.3, -0.4, 0.866]])
def test_fused_sweep_matches_oracle(ctx, los):
    shape = (32, 32, 128)
    v, f = fields(shape, 5)
    rec = recon(shape, los, ctx)
    got = bh.gauss_seidel_(v.copy(), f, rec, sweeps=2, fused=True)
    exp = orc.mg_gauss_seidel(v.astype(np.float64), f.astype(np.float64), rec.box_size, rec.box_min, rec.beta,
                              los=los, sweeps=2)
    err = float(np.linalg.norm(got - exp) / np.linalg.norm(exp))
    print(f"los {los}: rel L2 vs FP64 oracle {err:.2e}, max abs {np.max(np.abs(got - exp)):.2e}")
    assert err < 2e-6


def test_fused_sweep_rejects_unsupported_shape(ctx):
    shape = (8, 16, 48)   # x side not a multiple of 64
    v, f = fields(shape, 1)
    with pytest.raises(bh.BaorecError):
        bh.gauss_seidel_(v, f, recon(shape, None, ctx), fused=True)
    bh.gauss_seidel_(v, f, recon(shape, None, ctx), fused=False)


@pytest.mark.parametrize("shape,los,sweeps", [
    ((128, 128, 128), None, 5),
    ((128, 128, 128), [0.0, 0.0, 1.0], 3),
    ((32, 256, 256), None, 4),
    ((64, 128, 256), [0.6, 0.0, 0.8], 2),
])
def test_blocked_jacobi_bitwise_single_sweeps(ctx, shape, los, sweeps):
    """jacobi! with 2-5 sweeps on grids of >= 128^3 cells runs 3 or 2 sweeps
    per launch (temporal blocking); one sweep per call runs the plain
    kernel.  Both give the same bits."""
    v, f = fields(shape, 7 + sweeps)
    rec = recon(shape, los, ctx)
    rec.jacobi_niterations = sweeps
    got = bh.jacobi_(v.copy(), f, rec)
    rec.jacobi_niterations = 1
    exp = v.copy()
    for _ in range(sweeps):
        bh.jacobi_(exp, f, rec)
    assert np.all(np.isfinite(got))
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("shape,los", [((128, 128, 128), None), ((64, 64, 512), [0.0, 0.0, 1.0]),
                                       ((32, 256, 256), [0.6, 0.0, 0.8])])
def test_zmarching_residual_and_prolongation_match_oracle(ctx, shape, los):
    """On grids of >= 128^3 cells the FP32 residual runs as a z-marching
    pass and the prolongation two cells per thread; both keep the reference's
    FP32 operation order (residual!, prolong!: src/multigrid.jl:378-395,
    398-518)."""
    v, f = fields(shape, 3)
    rec = recon(shape, los, ctx)
    o = orc.MultigridRecon(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=rec.box_size, box_min=rec.box_min,
                           los=los)
    r = np.zeros_like(v)
    bh.residual_(r, v, f, rec)
    exp = orc.mg_residual(v, f, o.box_size, o.box_min, o.beta, los=o.los)
    print("residual max |diff| / max |r|:", np.max(np.abs(r - exp)) / np.abs(exp).max())
    assert np.max(np.abs(r - exp)) < 1e-6 * np.abs(exp).max()
    c = v[::2, ::2, ::2].copy()
    fine = np.zeros(shape, F32)
    bh.prolong_(fine, c, add=False, ctx=ctx)
    np.testing.assert_array_equal(fine, orc.mg_prolong(c))
    fine2 = v.copy()
    bh.prolong_(fine2, c, add=True, ctx=ctx)
    np.testing.assert_allclose(fine2, v + orc.mg_prolong(c), rtol=0, atol=1e-6)


@pytest.mark.parametrize("shape,los", [
    ((2, 16, 64), None),
    ((24, 32, 64), [0.0, 0.0, 1.0]),
    ((40, 48, 192), None),
    ((128, 128, 128), None),
    ((64, 64, 512), [0.6, 0.0, 0.8]),
    ((512, 512, 512), None),
])
def test_fused_residual_restriction_bitwise(ctx, shape, los):
    """The V-cycle's step down, reduce!(residual!(v, f)), as one pass: the
    same bits as the two kernels, on one-plane-pair grids, halos wrapping
    onto the tile itself, non-cubic sides and the solver's 128^3 and 512^3
    levels."""
    v, f = fields(shape, 11 + shape[0])
    rec = recon(shape, los, ctx)
    half = tuple(s // 2 for s in shape)
    a = bh.residual_restrict_(np.full(half, np.nan, F32), v, f, rec, fused=True)
    b = bh.residual_restrict_(np.full(half, np.nan, F32), v, f, rec, fused=False)
    assert np.all(np.isfinite(a))
    np.testing.assert_array_equal(a, b)
    if shape[0] * shape[1] * shape[2] <= 64 * 64 * 512:
        o = orc.MultigridRecon(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=rec.box_size,
                               box_min=rec.box_min, los=los)
        exp = orc.mg_reduce(orc.mg_residual(v, f, o.box_size, o.box_min, o.beta, los=o.los))
        assert np.max(np.abs(a - exp)) < 1e-6 * np.abs(exp).max()


def test_fused_residual_restriction_rejects_unsupported_shape(ctx):
    shape = (8, 16, 96)   # x side not a multiple of 64
    v, f = fields(shape, 2)
    c = np.zeros((4, 8, 48), F32)
    with pytest.raises(bh.BaorecError):
        bh.residual_restrict_(c, v, f, recon(shape, None, ctx), fused=True)
    bh.residual_restrict_(c, v, f, recon(shape, None, ctx), fused=False)
