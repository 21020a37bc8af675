"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something its own code does not
restate: exact-arithmetic golden vectors (tests/golden/), closed forms
(plane waves), invariants (mass conservation, adjointness), a brute-force DFT
on a tiny grid, and, for the multigrid operator, its exact Fourier symbol.
A planted-mistake test shows the checks would catch a plausible slip.
"""
import json
import os

import numpy as np
import pytest

import baorec_oracle as orc

F32, F64 = np.float32, np.float64
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


# ------------------------------------------------------------------ CIC

@pytest.mark.parametrize("dt", [F32, F64])
def test_cic_golden(dt):
    for case in load("cic.json"):
        n, L, b = case["n"], case["L"], case["b"]
        rho = np.zeros((n[2], n[1], n[0]), dt)
        p = [np.array([v], F32) for v in case["p"]]
        orc.cic(rho, p[0], p[1], p[2], np.array([case["w"]], F32), L, b, wrap=case["wrap"])
        ref = np.zeros_like(rho)
        for i, j, k, v in case["cells"]:
            ref[k, j, i] += v
        np.testing.assert_array_equal(rho, ref)


def test_cic_mass_conservation_and_wrap_mutation():
    rng = np.random.default_rng(0)
    n = 5000
    x, y, z = (rng.random(n).astype(F32) * F32(100.0) for _ in range(3))
    x[:10] = F32(100.0) + F32(7.0)  # beyond the box: wrapped once
    w = rng.random(n).astype(F32)
    rho = np.zeros((16, 8, 12), F64)
    xx = x.copy()
    orc.cic(rho, xx, y.copy(), z.copy(), w, [100.0] * 3, [0.0] * 3, wrap=True)
    assert abs(rho.sum() - w.astype(F64).sum()) < 1e-9 * w.sum()
    np.testing.assert_array_equal(xx[:10], np.full(10, F32(7.0)))
    np.testing.assert_array_equal(xx[10:], x[10:])


def test_cic_nowrap_bounds_error():
    rho = np.zeros((8, 8, 8), F32)
    with pytest.raises(IndexError):
        orc.cic(rho, np.array([7.9], F32), np.array([1.0], F32), np.array([1.0], F32), np.ones(1, F32),
                [8.0] * 3, [0.0] * 3, wrap=False)


# ------------------------------------------------------------------ gather

def test_read_cic_golden():
    g = load("read_cic.json")
    field = np.array(g["field"], F32)
    for pt in g["points"]:
        p = [np.array([v], F32) for v in pt["p"]]
        for fml in (True, False):
            got = orc.read_cic(field, *p, g["L"], g["b"], gpu_formula=fml)
            assert got[0] == pytest.approx(pt["value"], abs=1e-6)


def test_read_cic_linear_and_constant():
    n = 16
    L = 32.0
    ix = np.arange(n, dtype=F64)
    field = (2.0 * ix[None, None, :] - 1.0 * ix[None, :, None] + 0.5 * ix[:, None, None] + 3.0)
    rng = np.random.default_rng(1)
    # interior points (no wrap of the upper corner): linear reproduced exactly
    p = [rng.uniform(0, L * (n - 1) / n, 200).astype(F32) for _ in range(3)]
    got = orc.read_cic(field, *p, [L] * 3, [0.0] * 3)
    u = [q.astype(F64) * n / L for q in p]
    np.testing.assert_allclose(got, 2 * u[0] - u[1] + 0.5 * u[2] + 3.0, atol=1e-9)
    np.testing.assert_allclose(orc.read_cic(np.full((n, n, n), 4.25), *p, [L] * 3, [0.0] * 3), 4.25)


def test_scatter_gather_adjoint():
    """<gather(F), w> == <F, scatter(w)> (same nodes, same weights)."""
    rng = np.random.default_rng(2)
    n, L = 12, 24.0
    x, y, z = (rng.random(3000).astype(F32) * F32(L) for _ in range(3))
    w = rng.random(3000).astype(F32)
    F = rng.standard_normal((n, n, n))
    rho = np.zeros((n, n, n), F64)
    orc.cic(rho, x.copy(), y.copy(), z.copy(), w, [L] * 3, [0.0] * 3, wrap=True)
    lhs = np.dot(orc.read_cic(F, x, y, z, [L] * 3, [0.0] * 3), w.astype(F64))
    rhs = np.sum(F * rho)
    assert lhs == pytest.approx(rhs, rel=1e-5)


# ------------------------------------------------------------------ grid vectors

def test_k_vec_and_x_vec_conventions():
    kx, ky, kz = orc.k_vec((8, 8, 6), [1000.0, 1000.0, 600.0])
    dk = F32(F32(2 * np.pi * 8 / 1000.0) / F32(8))
    np.testing.assert_array_equal(kx, np.arange(5, dtype=F32) * dk)          # rfftfreq: 0..n/2
    np.testing.assert_array_equal(ky, np.array([0, 1, 2, 3, -4, -3, -2, -1], F32) * dk)  # Nyquist negative
    assert kz.shape == (6,)
    xs = orc.x_vec((4, 4, 4), [8.0] * 3, [-4.0, 0.0, 1.0])
    np.testing.assert_array_equal(xs[0], np.array([-3, -1, 1, 3], F32))
    np.testing.assert_array_equal(xs[2], np.array([2, 4, 6, 8], F32))


def test_setup_box():
    x = np.array([1.0, 5.0], F32)
    y = np.array([-2.0, 0.0], F32)
    z = np.array([3.0, 3.5], F32)
    bs, bm = orc.setup_box(x, y, z, 10.0)
    np.testing.assert_array_equal(bm, np.array([-4.0, -7.0, -2.0], F32))
    np.testing.assert_array_equal(bs, np.full(3, 14.0, F32))


# ------------------------------------------------------------------ smoothing / FFT steps

def plane_wave(n, L, m, phase=0.0):
    x = np.arange(n) * (L / n)
    kv = [2 * np.pi * mi / L for mi in m]
    arg = kv[0] * x[None, None, :] + kv[1] * x[None, :, None] + kv[2] * x[:, None, None] + phase
    return np.cos(arg), np.array(kv)


def test_smoothing_plane_wave():
    n, L, R = 16, 100.0, 7.0
    f, k = plane_wave(n, L, (2, -3, 1))
    got = orc.smooth(f.copy(), R, [L] * 3)
    np.testing.assert_allclose(got, np.exp(-0.5 * R * R * (k @ k)) * f, atol=1e-12)


def nodal_catalog(n, L, w, m=None):
    """One FP32 particle exactly on every grid node (CIC then deposits all of
    its weight into that node's cell), weights w[z, y, x] (exact in FP32)."""
    c = L / n
    g = np.arange(n if m is None else m, dtype=F32) * F32(c)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    assert np.all(w.astype(F32) == w) and w.shape == x.shape
    return x.ravel().copy(), y.ravel().copy(), z.ravel().copy(), w.ravel().astype(F32)


def test_setup_overdensity_box_plane_wave():
    """rho = 1 + eps cos(k.x) on the nodes: mean 1, so delta must be the
    smoothed mode eps exp(-R^2 k^2 / 2) cos(k.x) / b (src/recon.jl:42-57).
    k.x = pi (i + j) / 2, so the weights are exact in FP32."""
    n, L, R, b, eps = 16, 160.0, 12.0, 1.7, 0.25
    f, k = plane_wave(n, L, (4, 4, 0))
    f = np.round(f)
    x, y, z, w = nodal_catalog(n, L, 1 + eps * f)
    rec = orc.IterativeRecon(bias=b, f=0.5, smoothing_radius=R, box_size=[L] * 3, box_min=[0.0] * 3)
    delta = orc.setup_overdensity(np.zeros((n, n, n)), rec, x, y, z, w)
    np.testing.assert_allclose(delta, eps * np.exp(-0.5 * R * R * (k @ k)) * f / b, atol=1e-12)


def test_setup_overdensity_randoms_exact():
    """R = 0 (smoothing is the identity): alpha = S_d/S_r, cells whose randoms
    density is at most 0.01 S_r/N_r give 0, others (d - alpha r)/(b alpha r)
    (src/recon.jl:60-91); expectations in exact rational arithmetic.  The
    particles sit on the nodes 0..n-2 (no wrap: a particle on the last node
    would index past the grid), so the last plane of every axis stays empty."""
    from fractions import Fraction as Fr
    n, L, b = 8, 80.0, 2.0
    m = n - 1
    N = m ** 3
    wr = np.ones((m, m, m)); wr[1, 2, 3] = 1 / 256         # below the threshold
    wd = np.ones((m, m, m)); wd[4, 5, 6] = 3.0
    rec = orc.IterativeRecon(bias=b, f=0.5, smoothing_radius=0.0, box_size=[L] * 3, box_min=[0.0] * 3)
    delta = orc.setup_overdensity_randoms(np.zeros((n, n, n)), rec, *nodal_catalog(n, L, wd, m),
                                          *nodal_catalog(n, L, wr, m))
    alpha = Fr(N + 2) / (Fr(N) - 1 + Fr(1, 256))
    expect = np.zeros((n, n, n))
    expect[:m, :m, :m] = float((1 - alpha) / (b * alpha))
    expect[4, 5, 6] = float((3 - alpha) / (b * alpha))
    expect[1, 2, 3] = 0.0
    np.testing.assert_allclose(delta, expect, atol=1e-12)


def test_iterate_beta_zero_is_identity():
    rng = np.random.default_rng(3)
    ds = rng.standard_normal((8, 8, 8))
    kv = orc.k_vec((8, 8, 8), [50.0] * 3, F64)
    xv = orc.x_vec((8, 8, 8), [50.0] * 3, [10.0] * 3, F64)
    for los, x in ((None, xv), ((0.0, 0.0, 1.0), None)):
        dr = ds.copy()
        orc.iterate(dr, ds, kv, 1, 0.0, los=los, xv=x)
        np.testing.assert_allclose(dr, ds, atol=1e-14)


def test_iterate_fixed_los_plane_wave_recursion():
    """delta^(n) = delta_s - fac_n mu^2 delta^(n-1), mu = k_z/k, fac_1 = b/(1+b)."""
    n, L, beta = 16, 200.0, 0.4
    ds, k = plane_wave(n, L, (1, 2, 3))
    mu2 = k[2] ** 2 / (k @ k)
    kv = orc.k_vec((n, n, n), [L] * 3, F64)
    dr = ds.copy()
    expect = ds.copy()
    for it in (1, 2, 3):
        orc.iterate(dr, ds, kv, it, beta, los=(0.0, 0.0, 1.0))
        fac = beta / (1 + beta) if it == 1 else beta
        expect = ds - fac * mu2 * expect
        np.testing.assert_allclose(dr, expect, atol=1e-12)


def dft_matrix(n, inverse=False):
    j = np.arange(n)
    s = 1 if inverse else -1
    return np.exp(s * 2j * np.pi * np.outer(j, j) / n)


def test_iterate_local_los_brute_force():
    """Local-LOS update against explicit DFT matrices (no FFT library).  All
    sides are odd: with no self-paired Nyquist bins the k_i k_j spectra stay
    Hermitian, so the full-spectrum brute force equals the reference's
    R2C/C2R exactly (at a Nyquist bin the two conventions legitimately differ)."""
    n = (7, 5, 9)
    L = (70.0, 50.0, 45.0)
    beta = 0.35
    bmin = [5.0, -20.0, 13.0]
    rng = np.random.default_rng(4)
    ds = rng.standard_normal((n[2], n[1], n[0]))
    W = [dft_matrix(m) for m in n]
    Wi = [dft_matrix(m, True) for m in n]
    dk = np.einsum("ai,bj,ck,kji->cba", W[0], W[1], W[2], ds)  # axes (z,y,x), full spectrum
    m = [np.fft.fftfreq(n[a], 1.0 / n[a]) * 2 * np.pi / L[a] for a in range(3)]
    K = [m[0][None, None, :], m[1][None, :, None], m[2][:, None, None]]
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = np.where(k2 > 0, dk / k2, 0)
    xv = orc.x_vec(n, L, bmin, F64)
    X = [xv[0][None, None, :], xv[1][None, :, None], xv[2][:, None, None]]
    x2 = X[0] ** 2 + X[1] ** 2 + X[2] ** 2
    expect = ds.copy()
    for i in range(3):
        for j in range(i, 3):
            D = np.einsum("ai,bj,ck,kji->cba", Wi[0], Wi[1], Wi[2], K[i] * K[j] * phi).real / np.prod(n)
            fac = (2.0 if i != j else 1.0) * beta / (1 + beta)
            expect = expect - fac * D * X[i] * X[j] / x2
    dr = ds.copy()
    orc.iterate(dr, ds, orc.k_vec(n, L, F64), 1, beta, los=None, xv=xv)
    np.testing.assert_allclose(dr, expect, atol=1e-12)


def test_displacement_plane_wave():
    """delta = A cos(k.x)  ->  Psi = -A k sin(k.x)/k^2 (iterative: Psi_k = i k delta_k / k^2)."""
    n, L, A = 16, 160.0, 0.7
    d, k = plane_wave(n, L, (1, -2, 2))
    rec = orc.IterativeRecon(bias=1.0, f=0.5, smoothing_radius=0.0, box_size=[L] * 3, box_min=[0.0] * 3)
    grids = orc.displacement_grids(A * d, rec)
    s, _ = plane_wave(n, L, (1, -2, 2), phase=-np.pi / 2)  # sin(k.x)
    for i in range(3):
        np.testing.assert_allclose(grids[i], -A * k[i] / (k @ k) * s, atol=1e-12)
    # multigrid read-out: Psi_k = i k phi_k  ->  Psi = -A k sin(k.x) for phi = A cos(k.x)
    mg = orc.MultigridRecon(bias=1.0, f=0.5, smoothing_radius=0.0, box_size=[L] * 3, box_min=[0.0] * 3)
    grids = orc.displacement_grids(A * d, mg)
    for i in range(3):
        np.testing.assert_allclose(grids[i], -A * k[i] * s, atol=1e-12)


def test_rsd_shift_properties():
    n, L = 8, 80.0
    rng = np.random.default_rng(5)
    field = rng.standard_normal((n, n, n))
    x, y, z = (rng.uniform(1.0, L - 1.0, 300).astype(F32) for _ in range(3))
    base = dict(bias=2.0, smoothing_radius=10.0, box_size=[L] * 3, box_min=[0.0] * 3)
    r0 = orc.IterativeRecon(f=0.0, **base)
    np.testing.assert_array_equal(np.array(orc.read_shifts(r0, x, y, z, field, "sum")),
                                  np.array(orc.read_shifts(r0, x, y, z, field, "disp")))
    r1 = orc.IterativeRecon(f=0.7, los=(0.0, 0.0, 1.0), **base)
    disp = orc.read_shifts(r1, x, y, z, field, "disp")
    rsd = orc.read_shifts(r1, x, y, z, field, "rsd")
    np.testing.assert_allclose(rsd[2], 0.7 * disp[2], rtol=1e-12)   # along LOS: f * Psi
    np.testing.assert_array_equal(rsd[0], 0.0)                       # perpendicular: 0
    r2 = orc.IterativeRecon(f=0.7, los=None, **base)
    disp = orc.read_shifts(r2, x, y, z, field, "disp")
    rsd = orc.read_shifts(r2, x, y, z, field, "rsd")
    rhat = np.stack([x, y, z]).astype(F64)
    rhat /= np.linalg.norm(rhat, axis=0)
    dot = np.sum(np.stack(disp) * rhat, axis=0)
    np.testing.assert_allclose(np.stack(rsd), 0.7 * dot * rhat, atol=1e-12)


# ------------------------------------------------------------------ multigrid

def test_restrict_prolong_golden():
    g = load("restrict.json")
    np.testing.assert_array_equal(orc.mg_reduce(np.array(g["fine"], F64)), np.array(g["restricted"]))
    np.testing.assert_array_equal(orc.mg_reduce(np.array(g["fine"], F32)), np.array(g["restricted"], F32))
    np.testing.assert_array_equal(orc.mg_prolong(np.array(g["coarse"], F64)), np.array(g["prolonged"]))


def test_restrict_prolong_constants():
    np.testing.assert_array_equal(orc.mg_reduce(np.full((8, 8, 8), 2.5)), np.full((4, 4, 4), 2.5))
    np.testing.assert_array_equal(orc.mg_prolong(np.full((4, 2, 4), -1.25)), np.full((8, 4, 8), -1.25))


def mg_symbol(n, L, beta, los):
    """Fourier symbol of A v = D v - N - X for a fixed LOS (src/multigrid.jl:213-326)."""
    cell = np.array(L) / np.array(n)
    p = np.array(los, dtype=F32).astype(F64) / cell   # the oracle takes the LOS in FP32
    g = beta / np.sum(cell ** 2 * p ** 2)
    gp = 1 / cell ** 2 + g * p ** 2
    th = [2 * np.pi * np.fft.fftfreq(n[a]) for a in range(3)]
    T = [th[0][None, None, :], th[1][None, :, None], th[2][:, None, None]]
    A = sum(gp[a] * (2 - 2 * np.cos(T[a])) for a in range(3))
    for i, j in ((0, 1), (0, 2), (1, 2)):
        A = A + 2 * g * p[i] * p[j] * np.sin(T[i]) * np.sin(T[j])
    return A


@pytest.mark.parametrize("los", [(0.0, 0.0, 1.0), (0.6, 0.0, 0.8)])
def test_mg_residual_matches_operator_symbol(los):
    n, L, beta = (16, 8, 12), (32.0, 16.0, 24.0), 0.4
    rng = np.random.default_rng(6)
    v = rng.standard_normal((n[2], n[1], n[0]))
    r = orc.mg_residual(v, np.zeros_like(v), L, [0.0] * 3, beta, los=los)   # r = -A v
    Av = np.fft.ifftn(np.fft.fftn(v) * mg_symbol(n, L, beta, los)).real
    np.testing.assert_allclose(-r, Av, atol=1e-11)


@pytest.mark.parametrize("los", [(0.0, 0.0, 1.0), (0.6, 0.0, 0.8)])
def test_mg_jacobi_plane_wave_amplification(los):
    """With f = 0 a damped Jacobi sweep multiplies each Fourier mode by
    1 - omega lambda(k) / D, D the stencil's diagonal 2 sum_a (1/h_a^2 + g p_a^2)
    (src/multigrid.jl:14-135)."""
    n, L, beta, omega, niter = (16, 8, 12), (32.0, 16.0, 24.0), 0.4, 0.4, 3
    cell = np.array(L) / np.array(n)
    p = np.array(los, dtype=F32).astype(F64) / cell
    g = beta / np.sum(cell ** 2 * p ** 2)
    D = 2 * np.sum(1 / cell ** 2 + g * p ** 2)
    rng = np.random.default_rng(9)
    v = rng.standard_normal((n[2], n[1], n[0]))
    amp = (1 - omega * mg_symbol(n, L, beta, los) / D) ** niter
    expect = np.fft.ifftn(np.fft.fftn(v) * amp).real
    got = orc.mg_jacobi(v.copy(), np.zeros_like(v), L, [0.0] * 3, beta, omega, niter, los=los)
    np.testing.assert_allclose(got, expect, atol=1e-11)


def test_mg_vcycles_converge_to_fft_solution():
    n, L, beta, los = (16, 16, 16), (64.0,) * 3, 0.4, (0.0, 0.0, 1.0)
    rng = np.random.default_rng(7)
    f = rng.standard_normal((16, 16, 16))
    f -= f.mean()
    A = mg_symbol(n, L, beta, los)
    fk = np.fft.fftn(f)
    phik = np.where(A > 1e-14, fk / np.where(A > 1e-14, A, 1), 0)
    phi = np.fft.ifftn(phik).real
    v = np.zeros_like(f)
    res = []
    for _ in range(25):
        orc.mg_vcycle(v, f, L, [0.0] * 3, beta, 0.4, 5, los=los)
        res.append(np.linalg.norm(orc.mg_residual(v, f, L, [0.0] * 3, beta, los=los)))
    assert res[1] / res[0] < 0.5 and res[5] / res[4] < 0.5           # contraction per cycle
    np.testing.assert_allclose(v - v.mean(), phi, atol=1e-8 * np.abs(phi).max())


def test_mg_fmg_reduces_residual_local_los():
    n, L = 32, 500.0
    rng = np.random.default_rng(8)
    f = rng.standard_normal((n, n, n))
    f -= f.mean()
    r0 = np.linalg.norm(f)
    v = orc.fmg(f, [L] * 3, [200.0, 300.0, 100.0], 0.375, 0.4, 5, 6, los=None)
    r = np.linalg.norm(orc.mg_residual(v, f, [L] * 3, [200.0, 300.0, 100.0], 0.375, los=None))
    assert r / r0 < 1e-2


def gaussian_blob_operator(n, L, box_min, beta, centre, s):
    """phi = exp(-|x - c|^2 / 2 s^2) at the cell centres b + (i + 1/2) L/n and
    the continuum operator the stencil of src/multigrid.jl:213-326
    discretises with a local LOS:  r = f - A v with f = 0 gives
    L phi = lap phi + beta [(rhat.grad)^2 phi + (2/|x|) rhat.grad phi],
    evaluated from the closed-form derivatives of the Gaussian."""
    h = L / n
    xs = [box_min[a] + (np.arange(n) + 0.5) * h for a in range(3)]
    X, Y, Z = xs[0][None, None, :], xs[1][None, :, None], xs[2][:, None, None]
    dx, dy, dz = X - centre[0], Y - centre[1], Z - centre[2]
    phi = np.exp(-(dx ** 2 + dy ** 2 + dz ** 2) / (2 * s * s))
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    dr = (dx * X + dy * Y + dz * Z) / r                   # (x - c) . rhat
    lap = ((dx ** 2 + dy ** 2 + dz ** 2) / s ** 4 - 3 / s ** 2) * phi
    rr = (dr ** 2 / s ** 4 - 1 / s ** 2) * phi           # (rhat.grad)^2 phi
    rg = -dr / s ** 2 * phi                               # rhat.grad phi
    return phi, lap + beta * (rr + 2 / r * rg)


BLOB = dict(L=600.0, box_min=(50.0, -300.0, 100.0), beta=0.4, centre=(350.0, 0.0, 400.0), s=50.0)


def blob_operator_errors(sizes, dtype=F64):
    errs = []
    for n in sizes:
        phi, Lphi = gaussian_blob_operator(n, BLOB["L"], BLOB["box_min"], BLOB["beta"], BLOB["centre"],
                                           BLOB["s"])
        r = orc.mg_residual(phi.astype(dtype), np.zeros(phi.shape, dtype), [BLOB["L"]] * 3, BLOB["box_min"],
                            BLOB["beta"], los=None)
        errs.append(np.linalg.norm(r - Lphi) / np.linalg.norm(Lphi))
    return errs


def test_mg_operator_local_los_second_order_against_analytic():
    """The local-LOS operator (position-dependent p = x/cell, cross terms and
    the first-derivative term sG, src/multigrid.jl:213-326) converges to the
    continuum operator at second order: error ratio 4 per halving of the cell,
    on a Gaussian blob whose tails vanish at the periodic faces."""
    e = blob_operator_errors((32, 64, 128))
    print("operator rel. errors", e)
    assert 3.6 < e[0] / e[1] < 4.4 and 3.6 < e[1] / e[2] < 4.4
    assert e[2] < 5e-3
    # the FP32 instantiation (the GPU's comparison target) is the same operator
    e32 = blob_operator_errors((64,), F32)[0]
    assert abs(e32 - e[1]) < 0.02 * e[1]


@pytest.mark.parametrize("plant", [1, 2])
def test_mg_operator_analytic_check_catches_planted(plant):
    """A flipped sign of sG (1) or p measured from box_min instead of the
    observer (2) breaks the second-order convergence above."""
    with orc.planted(plant):
        e = blob_operator_errors((32, 64, 128))
    print(f"plant {plant}: operator rel. errors", e)
    assert not (3.6 < e[0] / e[1] < 4.4 and 3.6 < e[1] / e[2] < 4.4 and e[2] < 5e-3)


def smooth_field(n, L, R, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n, n, n))
    k = 2 * np.pi * np.fft.fftfreq(n, L / n)
    k2 = k[None, None, :] ** 2 + k[None, :, None] ** 2 + k[:, None, None] ** 2
    f = np.fft.ifftn(np.fft.fftn(f) * np.exp(-0.5 * k2 * R * R)).real
    return f - f.mean()


def fmg_fixed_los_errors(ncycles_list):
    """fmg (src/multigrid.jl:722-752) against the exact inverse of the same
    discrete operator (its Fourier symbol, fixed LOS), for a smooth f."""
    n, L, beta, los = 64, 1000.0, 0.4, (0.0, 0.6, 0.8)
    f = smooth_field(n, L, 40.0, 3)
    A = mg_symbol((n,) * 3, (L,) * 3, beta, los)
    fk = np.fft.fftn(f)
    phi = np.fft.ifftn(np.where(A > 1e-14, fk / np.where(A > 1e-14, A, 1), 0)).real
    out = []
    for nc in ncycles_list:
        v = orc.fmg(f, [L] * 3, [0.0] * 3, beta, 0.4, 5, nc, los=los)
        out.append(np.linalg.norm(v - v.mean() - phi) / np.linalg.norm(phi))
    return out


def test_fmg_fixed_los_full_multigrid_accuracy():
    """Full multigrid's defining property: one V-cycle per level, started
    from the prolonged coarse solution, already lands near the discrete
    solution (1e-2 here, where one V-cycle from v = 0 leaves 0.23); the
    reference's 6 V-cycles per level reach 1e-5."""
    e1, e6 = fmg_fixed_los_errors((1, 6))
    print(f"fmg vs exact inverse: 1 cycle {e1:.2e}, 6 cycles {e6:.2e}")
    assert e1 < 0.03 and e6 < 3e-5


@pytest.mark.parametrize("plant", [3, 4])
def test_fmg_accuracy_catches_planted(plant):
    """A wrong prolongation index map (3) or a level started from v = 0
    instead of the prolonged coarse solution (4) fails the FMG accuracy."""
    with orc.planted(plant):
        e1, e6 = fmg_fixed_los_errors((1, 6))
    print(f"plant {plant}: fmg 1 cycle {e1:.2e}, 6 cycles {e6:.2e}")
    assert not (e1 < 0.03 and e6 < 3e-5)


def test_fmg_local_los_recovers_manufactured_solution():
    """Local LOS, observer outside the box: f = A v_true (A the oracle's
    operator, pinned against the continuum above), so f lies in A's range;
    the reference schedule (6 V-cycles per level, 5+5 damped Jacobi) returns
    v_true up to a constant within 1e-4, one cycle per level within 0.1."""
    from scipy.ndimage import gaussian_filter
    n, L, beta, bm = 64, 1000.0, 0.4, [200.0, 300.0, 100.0]
    rng = np.random.default_rng(3)
    vt = gaussian_filter(rng.standard_normal((n, n, n)), 3.0, mode="wrap")
    vt -= vt.mean()
    f = -orc.mg_residual(vt, np.zeros_like(vt), [L] * 3, bm, beta, los=None)
    errs = []
    for nc in (1, 6):
        v = orc.fmg(f, [L] * 3, bm, beta, 0.4, 5, nc, los=None)
        errs.append(np.linalg.norm(v - v.mean() - vt) / np.linalg.norm(vt))
    print("local LOS fmg errors", errs)
    assert errs[0] < 0.1 and errs[1] < 1e-4


def operator_matrix(n, L, box_min, beta, los=None):
    """The oracle's multigrid operator M (r = f - M v, src/multigrid.jl:213-326)
    on an n^3 periodic grid as a sparse matrix, read off by applying it to
    the 64 indicator fields of the cells (i mod 4, j mod 4, k mod 4): the
    19-point stencil reaches one cell of each such class per row."""
    import scipy.sparse as sp
    N = n ** 3
    idx = np.arange(N).reshape(n, n, n)
    Z, Y, X = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rows, cols, vals = [], [], []
    for cz in range(4):
        for cy in range(4):
            for cx in range(4):
                v = np.zeros((n, n, n))
                v[cz::4, cy::4, cx::4] = 1.0
                Mv = -orc.mg_residual(v, np.zeros_like(v), [L] * 3, box_min, beta, los=los)
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            zz, yy, xx = (Z + dz) % n, (Y + dy) % n, (X + dx) % n
                            m = (zz % 4 == cz) & (yy % 4 == cy) & (xx % 4 == cx)
                            rows.append(idx[m])
                            cols.append(idx[zz[m], yy[m], xx[m]])
                            vals.append(Mv[m])
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    M.eliminate_zeros()
    return M


def left_null_vector(M):
    """w with M^T w = 0, mean 1 (one equation replaced by w_0 = 1; sparse LU)."""
    import scipy.sparse.linalg as spl
    N = M.shape[0]
    MT = M.T.tolil()
    MT[0, :] = 0
    MT[0, 0] = 1.0
    b = np.zeros(N)
    b[0] = 1.0
    w = spl.spsolve(MT.tocsc(), b)
    return w / w.mean()


def test_local_los_operator_left_null_vector():
    """With a local LOS the operator is not symmetric: it annihilates constants
    (zero row sums), but its left null vector w is not constant, so f is in
    its range only when w . f = 0.  The reference's own V-cycles on an
    arbitrary zero-mean f stall at a residual floor; on f - (w.f)/(w.1) they
    converge to rounding.  Moving the residual's mean into the shift of f
    after every V-cycle reaches that same shift (w.f)/(w.1) without w."""
    n, L, bm, beta = 16, 1000.0, [200.0, 300.0, 100.0], 0.375
    M = operator_matrix(n, L, bm, beta)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((n, n, n))
    np.testing.assert_allclose(M @ v.ravel(), -orc.mg_residual(v, np.zeros_like(v), [L] * 3, bm, beta).ravel(),
                               atol=1e-12)
    assert np.abs(M @ np.ones(n ** 3)).max() < 1e-12
    w = left_null_vector(M)
    assert np.abs(M.T @ w).max() < 1e-12 and np.sqrt(np.mean((w - 1) ** 2)) > 1e-2
    f = rng.standard_normal((n, n, n))
    f -= f.mean()
    alpha = float(w @ f.ravel()) / float(w.sum())

    def cycles(ff, shift_each_cycle):
        ff = ff.copy()
        vv = np.zeros_like(ff)
        shift, res = 0.0, []
        for _ in range(24):
            orc.mg_vcycle(vv, ff, [L] * 3, bm, beta, 0.4, 5, los=None)
            r = orc.mg_residual(vv, ff, [L] * 3, bm, beta, los=None)
            res.append(np.linalg.norm(r) / np.linalg.norm(f))
            if shift_each_cycle:
                ff -= r.mean()
                shift += r.mean()
        return res, shift

    stalled, _ = cycles(f, False)
    projected, _ = cycles(f - alpha, False)
    shifted, shift = cycles(f, True)
    print(f"alpha {alpha:.3e}: floor {stalled[-1]:.1e}, projected {projected[-1]:.1e}, "
          f"shifted {shifted[-1]:.1e} (shift {shift:.6e})")
    assert stalled[-1] > 1e-5 and projected[-1] < 1e-12 and shifted[-1] < 1e-12
    assert abs(shift - alpha) < 1e-12


def textbook_gauss_seidel(M, v, f, order):
    """Forward Gauss-Seidel x <- (D + L)^-1 (f - U x) on the system permuted to
    the cell order `order` (D, L, U: diagonal, strictly lower and upper parts)."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve_triangular
    Mp = M[order][:, order].tocsr()
    x = v.ravel()[order]
    rhs = f.ravel()[order] - sp.triu(Mp, k=1) @ x
    x = spsolve_triangular(sp.tril(Mp, k=0).tocsr(), rhs, lower=True)
    out = np.empty_like(x)
    out[order] = x
    return out.reshape(v.shape)


@pytest.mark.parametrize("los", [None, [0.0, 0.0, 1.0]])
def test_mg_gauss_seidel_is_textbook_colour_ordered_gauss_seidel(los):
    """One sweep of the oracle's multicolour Gauss-Seidel equals forward
    Gauss-Seidel on the operator matrix with the cells ordered by colour (any
    order inside a colour: no two cells of a colour are coupled)."""
    n, L, bm, beta = 8, 1000.0, [200.0, 300.0, 100.0], 0.375
    M = operator_matrix(n, L, bm, beta, los=los)
    col = orc.mg_colors((n, n, n)).ravel()
    for c in range(4):   # the colouring decouples each colour
        sel = np.flatnonzero(col == c)
        blk = M[sel][:, sel]
        assert abs(blk - sp_diag(blk)).max() == 0.0
    rng = np.random.default_rng(4)
    v, f = rng.standard_normal((n, n, n)), rng.standard_normal((n, n, n))
    got = orc.mg_gauss_seidel(v.copy(), f, [L] * 3, bm, beta, los=los)
    order = np.concatenate([rng.permutation(np.flatnonzero(col == c)) for c in range(4)])
    np.testing.assert_allclose(got, textbook_gauss_seidel(M, v, f, order), rtol=1e-12, atol=1e-12)


def sp_diag(m):
    import scipy.sparse as sp
    return sp.diags(m.diagonal())


def test_mg_gauss_seidel_check_catches_planted_colouring(monkeypatch):
    """A 2-colour (red-black) ordering leaves edge neighbours in one colour
    and fails the textbook comparison."""
    monkeypatch.setattr(orc, "mg_colors", lambda shape: np.indices(shape).sum(axis=0) % 2 * 1)
    with pytest.raises(AssertionError):
        test_mg_gauss_seidel_is_textbook_colour_ordered_gauss_seidel(None)


def centre_los_inverse(n, L, bmin, beta):
    """The warm start's spectral preconditioner: the exact inverse of the
    operator's Fourier symbol with the LOS fixed at the box centre (no
    first-derivative term), zero on the k = 0 mode."""
    d = np.full(3, bmin + L / 2)
    d /= np.linalg.norm(d)
    S = mg_symbol((n,) * 3, (L,) * 3, beta, d)
    S[0, 0, 0] = np.inf
    return lambda r: np.fft.ifftn(np.fft.fftn(r) / S).real


def test_spectral_richardson_keeps_contracting_in_fp64():
    """Richardson steps v += M^-1 (f - A v) with the centre-LOS spectral
    inverse M^-1 on the local-LOS operator of a lightcone box (an octant shell
    r in [1000, 1800] Mpc/h, 64^3 cells of 31 Mpc/h): in FP64 every step cuts
    the relative residual (mean removed, as the consistency shift does) by
    more than 4x, below 3e-6 in six steps.  The 1e-4 floor of the FP32 steps
    at 512^3 is the rounding of the FP32 residual, not the preconditioner."""
    n, L, bmin, beta = 64, 1980.0, -90.0, 0.757 / 2.2
    rng = np.random.default_rng(3)
    c = bmin + (np.arange(n) + 0.5) * (L / n)
    r = np.sqrt(c[None, None, :] ** 2 + c[None, :, None] ** 2 + c[:, None, None] ** 2)
    f = orc.smooth(rng.standard_normal((n, n, n)), 60.0, [L] * 3) * ((r > 1000) & (r < 1800))
    f -= f.mean()
    kw = dict(box_size=[L] * 3, box_min=[bmin] * 3, beta=beta, los=None)
    Minv = centre_los_inverse(n, L, bmin, beta)

    def rel(v):
        res = orc.mg_residual(v, f, **kw)
        return np.linalg.norm(res - res.mean()) / np.linalg.norm(f)

    v = Minv(f)
    prev = rel(v)
    for step in range(6):
        res = orc.mg_residual(v, f, **kw)
        v = v + Minv(res - res.mean())
        cur = rel(v)
        print(f"step {step + 1}: {cur:.2e}")
        assert cur < prev / 4, (step, prev, cur)
        prev = cur
    assert prev < 3e-6


# ------------------------------------------------------------------ planted mistake

def test_planted_sign_error_is_caught(monkeypatch):
    """A sign flip in the displacement multiplier must fail the plane-wave check."""
    real = orc.displacement_grids

    def flipped(field, recon):
        return tuple(-g for g in real(field, recon))

    monkeypatch.setattr(orc, "displacement_grids", flipped)
    with pytest.raises(AssertionError):
        test_displacement_plane_wave()


# ------------------------------------------------------------------ cosmology / sky conversion

def test_cosmology_radiation_and_neutrino_densities():
    """Photon density of a T = 2.725 K CMB: omega_gamma = 2.47e-5; neutrinos
    Neff (7/8)(4/11)^(4/3) times that; the budget closes (src/cosmo.jl:36-58)."""
    c = orc.Cosmology()
    d = orc.cosmo_derived(c)
    h2 = float(np.float32(c.h) ** 2)
    assert abs(float(d["omega_g0"]) * h2 / 2.4730e-5 - 1) < 2e-3
    assert abs(float(d["omega_nu0"]) / float(d["omega_g0"]) - 3.044 * 7 / 8 * (4 / 11) ** (4 / 3)) < 1e-6
    tot = sum(float(d[k]) for k in ("omega_nu0", "omega_g0", "omega_b0", "omega_c0", "omega_k0", "omega_l0"))
    assert abs(tot - 1) < 1e-6


def test_comoving_distance_einstein_de_sitter():
    """No radiation (T_cmb = 0), omega_b + omega_c = h^2: E = (1+z)^1.5 and
    chi = 2 c / H0 (1 - 1 / sqrt(1 + z))."""
    h = 0.5
    c = orc.Cosmology(omega_b=0.0625, omega_c=0.1875, h=h, T_cmb=0.0)
    d = orc.cosmo_derived(c)
    assert abs(float(d["omega_l0"])) < 1e-6
    z = np.array([0.1, 0.5, 1.0, 2.0, 3.0])
    exact = 2 * orc.C_KMS / (100 * h) * (1 - 1 / np.sqrt(1 + z))
    np.testing.assert_allclose(orc.comoving_distance(c, z), exact, rtol=2e-6)


@pytest.mark.parametrize("zmin,zmax,num", [(0.0, 3.0, 100000), (0.1, 10.0, 2001), (0.0, 1.0, 2)])
def test_distance_table_matches_adaptive_quadrature(zmin, zmax, num):
    """The interval-by-interval table against an adaptive quadrature of the
    whole integral at knots spread over the table."""
    c = orc.Cosmology(z_tab_min=zmin, z_tab_max=zmax, z_tab_num=num)
    zk, dk = orc.distance_table(c)
    assert len(zk) == num and zk[0] == zmin and zk[-1] == zmax
    idx = np.unique(np.linspace(0, num - 1, 40).astype(int))
    np.testing.assert_allclose(dk[idx], orc.comoving_distance(c, zk[idx]), rtol=1e-10, atol=1e-9)
    assert orc.distance_table(c) is orc.distance_table(c)


def test_sky_cartesian_round_trip_and_geometry():
    """|pos| = chi(z) h; cartesian_to_sky(sky_to_cartesian(.)) returns Dec,
    z and RA as rem(RA - 360, 360) (examples/lightcone.jl:30-80)."""
    c = orc.Cosmology(z_tab_num=401)
    table = orc.distance_table(c)
    rng = np.random.default_rng(31)
    ra = rng.uniform(0, 360, 2000).astype(F32)
    dec = rng.uniform(-89, 89, 2000).astype(F32)
    red = rng.uniform(0.05, 2.9, 2000).astype(F32)
    x, y, z = orc.sky_to_cartesian(ra, dec, red, c, table)
    r = np.sqrt(x.astype(F64) ** 2 + y.astype(F64) ** 2 + z.astype(F64) ** 2)
    np.testing.assert_allclose(r, np.interp(red.astype(F64), *table) * float(np.float32(c.h)), rtol=1e-6)
    ra2, dec2, red2 = orc.cartesian_to_sky(x, y, z, c, table)
    assert np.all(ra2 <= 0) and np.all(ra2 > -360)
    lon = np.where(ra > 180, ra - 360, ra)          # atan2 range
    np.testing.assert_allclose(ra2, np.fmod(lon - 360, 360), atol=2e-3)
    np.testing.assert_allclose(dec2, dec, atol=2e-3)
    np.testing.assert_allclose(red2, red, atol=2e-5)
    with pytest.raises(IndexError):
        orc.sky_to_cartesian(ra[:1], dec[:1], np.array([3.5], F32), c, table)


# ------------------------------------------------------------------ power spectrum

def test_tsc_weights_closed_form():
    """One particle: three weights per axis summing to 1; on a node
    (1/8, 3/4, 1/8) around it; mass is conserved with periodic wrap."""
    g = orc.tsc_grid(np.array([2.0], F32), np.array([0.0], F32), np.array([7.75], F32), np.array([2.0], F32),
                     (8, 8, 8), (8.0, 8.0, 8.0), (0.0, 0.0, 0.0))
    assert abs(float(g.sum(dtype=F64)) - 2.0) < 1e-6
    wx = g.sum(axis=(0, 1)) / 2.0
    np.testing.assert_allclose(wx[[1, 2, 3]], [0.125, 0.75, 0.125], atol=1e-7)
    wz = g.sum(axis=(1, 2)) / 2.0      # z = 7.75: d = -0.25 around node 8 == 0 (wrapped)
    np.testing.assert_allclose(wz[[7, 0, 1]], [0.5 * 0.75 ** 2, 0.75 - 0.0625, 0.5 * 0.25 ** 2], atol=1e-7)


def test_power_multipoles_plane_wave_closed_form():
    """A particle lattice (one per cell, on the nodes) with weights
    1 + A cos(k0 x), k0 = 2 pi / L along x: the mode amplitude through TSC
    is T1 = prod_a (3/4 + cos(k_a H)/4) on the grid and T2 = prod_a cos(k_a H/2)
    on the interlaced grid, so P(k0) = V (A/2)^2 ((T1 + T2)/2)^2 / W(k0)^2 with
    W = prod sinc^3(k_a H/2).  The bin [0.8, 1.1) k_f holds exactly the six
    unit modes, two of them with power, all with mu = 0 for a z line of sight:
    P0 = P/3 (plus the shot noise the estimator subtracts), P2 = -5P/6,
    P4 = 9P/8."""
    n, L, A = 16, 160.0, 0.25
    H = L / n
    g = np.arange(n, dtype=F32) * F32(H)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    k0 = 2 * np.pi / L
    w = (1 + A * np.cos(k0 * x.astype(F64))).astype(F32)
    kf = 2 * np.pi / L
    out = orc.power_multipoles(x, y, z, w, (n, n, n), (L, L, L), kmin=0.8 * kf, dk=0.3 * kf, nbins=1)
    assert out["nmodes"][0] == 6
    a = k0 * H
    T1 = 0.75 + 0.25 * np.cos(a)          # y, z factors are 1 at k_y = k_z = 0
    T2 = np.cos(a / 2)
    Wk = (np.sin(a / 2) / (a / 2)) ** 3
    V = L ** 3
    Wt = float(np.sum(w, dtype=F64))
    amp = float(np.sum(w.astype(F64) * np.cos(k0 * x)))  # = N A / 2 for the lattice
    P = V * (amp / Wt) ** 2 * ((T1 + T2) / 2) ** 2 / Wk ** 2
    shot = V * float(np.sum(w.astype(F64) ** 2)) / Wt ** 2
    np.testing.assert_allclose(out["p0"][0] + shot, P / 3, rtol=1e-6)
    np.testing.assert_allclose(out["p2"][0], -5 * P / 6, rtol=1e-6)
    np.testing.assert_allclose(out["p4"][0], 9 * P / 8, rtol=1e-6)


def test_power_multipoles_poisson_is_shot_noise():
    """Unclustered points: P0 - shot noise consistent with zero, P2 and P4
    consistent with zero (5 sigma of the Gaussian mode scatter), below
    0.7 k_Nyquist (beyond it the window correction of aliased shot noise is
    not flat)."""
    rng = np.random.default_rng(43)
    n, L, N = 32, 500.0, 40000
    x, y, z = (rng.uniform(0, L, N).astype(F32) for _ in range(3))
    out = orc.power_multipoles(x, y, z, None, (n, n, n), (L, L, L), dk=0.05, nbins=6)
    shot = L ** 3 / N
    sig = shot * np.sqrt(1.0 / out["nmodes"])
    ok = (out["nmodes"] > 20) & (out["k"] < 0.7 * np.pi * n / L)
    assert ok.sum() >= 3
    assert np.all(np.abs(out["p0"][ok]) < 5 * sig[ok] + 0.03 * shot)
    assert np.all(np.abs(out["p2"][ok]) < 5 * np.sqrt(5) * sig[ok] + 0.05 * shot)
    assert np.all(np.abs(out["p4"][ok]) < 5 * 3 * sig[ok] + 0.05 * shot)
