"""Seeded synthetic catalogs for the reconstruction workloads.

The reference benchmarks real mocks that are not available offline
(test/benchmarks_gpu.jl:19-20), so every workload here uses catalogs drawn from
a lognormal density field:

* a Gaussian random field with power-law spectrum
  P(k) = amp * (k / 0.1)**slope  for k in [kmin, kmax], 0 elsewhere,
  built on a generator mesh by colouring white noise;
* variance-clipped: if var(G) > 1 the field is rescaled to unit variance;
* rho = exp(G - var(G)/2), times a selection function;
* exactly N tracers, drawn multinomially over mesh cells, uniform inside a cell.

Box catalogs fill [box_min, box_min + L)^3.  Lightcone catalogs fill the shell
r in [r_min, r_max] over RA, Dec in [0, 90] degrees around an observer at the
origin with n(r) proportional to r**-0.5; their randoms follow the same n(r)
without clustering.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def gaussian_field(n: int, L: float, rng, amp=2e4, slope=-1.5, kmin=0.005, kmax=0.5):
    """Real Gaussian field on an n^3 mesh of side L with the power-law P(k)."""
    w = rng.standard_normal((n, n, n), dtype=np.float64)
    wk = np.fft.rfftn(w)
    kf = 2 * np.pi / L
    kx = np.fft.rfftfreq(n, 1.0 / n) * kf
    ky = np.fft.fftfreq(n, 1.0 / n) * kf
    k = np.sqrt(kx[None, None, :] ** 2 + ky[None, :, None] ** 2 + ky[:, None, None] ** 2)
    with np.errstate(divide="ignore"):
        pk = np.where((k >= kmin) & (k <= kmax), amp * (k / 0.1) ** slope, 0.0)
    vcell = (L / n) ** 3
    wk *= np.sqrt(pk / vcell)
    return np.fft.irfftn(wk, s=(n, n, n), axes=(0, 1, 2))


def lognormal(g):
    var = g.var()
    if var > 1.0:
        g = g / np.sqrt(var)
        var = 1.0
    return np.exp(g - var / 2)


def _sample_cells(weights, n_part, rng):
    p = weights.ravel().astype(np.float64)
    p /= p.sum()
    counts = rng.multinomial(n_part, p)
    return np.repeat(np.arange(p.size), counts)


def box_catalog(n_part: int, L: float = 1000.0, box_min=0.0, mesh: int = 256, seed: int = 1,
                slope: float = -1.5, amp: float = 2e4):
    """Periodic-box tracers: (x, y, z, w) float32 arrays, w = 1."""
    rng = np.random.default_rng(seed)
    rho = lognormal(gaussian_field(mesh, L, rng, amp=amp, slope=slope))
    cells = _sample_cells(rho, n_part, rng)
    iz, rem = np.divmod(cells, mesh * mesh)
    iy, ix = np.divmod(rem, mesh)
    h = L / mesh
    pos = []
    for idx in (ix, iy, iz):
        p = box_min + (idx + rng.random(idx.size)) * h
        p = np.minimum(p, box_min + L * (1 - 1e-7))
        pos.append(p.astype(F32))
    perm = rng.permutation(n_part)
    x, y, z = (a[perm] for a in pos)
    return x, y, z, np.ones(n_part, dtype=F32)


def _octant_dirs(n, rng):
    ra = rng.random(n) * (np.pi / 2)
    sdec = rng.random(n)
    cdec = np.sqrt(1 - sdec ** 2)
    return cdec * np.cos(ra), cdec * np.sin(ra), sdec


def lightcone_randoms(n_part: int, r_min=1000.0, r_max=1800.0, seed: int = 3):
    """Unclustered randoms with n(r) ~ r^-0.5, i.e. dN/dr ~ r^1.5."""
    rng = np.random.default_rng(seed)
    a, b = r_min ** 2.5, r_max ** 2.5
    r = (a + rng.random(n_part) * (b - a)) ** (1 / 2.5)
    ux, uy, uz = _octant_dirs(n_part, rng)
    return ((r * ux).astype(F32), (r * uy).astype(F32), (r * uz).astype(F32),
            np.ones(n_part, dtype=F32))


def lightcone_catalog(n_part: int, r_min=1000.0, r_max=1800.0, mesh: int = 256, seed: int = 2,
                      slope: float = -1.5, amp: float = 2e4):
    """Clustered tracers in the octant shell with n(r) ~ r^-0.5."""
    rng = np.random.default_rng(seed)
    L = r_max * 1.0
    h = L / mesh
    rho = lognormal(gaussian_field(mesh, L, rng, amp=amp, slope=slope))
    c = (np.arange(mesh) + 0.5) * h
    r = np.sqrt(c[None, None, :] ** 2 + c[None, :, None] ** 2 + c[:, None, None] ** 2)
    sel = np.where((r >= r_min) & (r <= r_max), r ** -0.5, 0.0)
    cells = _sample_cells(rho * sel, n_part, rng)
    iz, rem = np.divmod(cells, mesh * mesh)
    iy, ix = np.divmod(rem, mesh)
    pos = [((idx + rng.random(idx.size)) * h).astype(F32) for idx in (ix, iy, iz)]
    perm = rng.permutation(n_part)
    x, y, z = (p[perm] for p in pos)
    return x, y, z, np.ones(n_part, dtype=F32)


def batch_catalogs(ncat: int, n_min: float = 2e5, n_max: float = 5e6, L: float = 1000.0, mesh: int = 256,
                   seed0: int = 100, sizes=None):
    """A batch of periodic-box catalogs laid end to end: catalog c has its own
    seed (seed0 + c), a power-law slope drawn uniformly in [-2, -1] and a size
    drawn log-uniformly in [n_min, n_max] (or sizes[c]).  Returns
    (x, y, z, w, offsets) with catalog c at [offsets[c], offsets[c+1])."""
    rng = np.random.default_rng(seed0)
    if sizes is None:
        sizes = np.exp(rng.uniform(np.log(n_min), np.log(n_max), ncat)).astype(np.int64)
    slopes = rng.uniform(-2.0, -1.0, ncat)
    cols = [[], [], [], []]
    for c in range(ncat):
        if sizes[c] == 0:
            continue
        for k, a in enumerate(box_catalog(int(sizes[c]), L=L, mesh=mesh, seed=seed0 + c, slope=float(slopes[c]))):
            cols[k].append(a)
    offsets = np.zeros(ncat + 1, np.int64)
    offsets[1:] = np.cumsum(sizes)
    out = [np.concatenate(a) if a else np.zeros(0, F32) for a in cols]
    return (*out, offsets)
