"""CPU oracle for the BAOrec.jl reconstruction hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module: as the checker the HIP
path is compared against, or as the timed CPU restatement.  The product path
(``baorec.jl_amd``) never imports it.

It restates the reference algorithm step by step (file:line of BAOrec.jl given
per function).  Grids are numpy arrays of shape ``(nz, ny, nx)`` in C order,
i.e. the same bytes as the reference's column-major ``(nx, ny, nz)`` Julia
arrays (x fastest).  ``dtype=np.float32`` keeps the reference's FP32 operation
order; ``dtype=np.float64`` is the same algorithm in FP64 (used to state the
error of the FP32 paths).  FFTs use scipy's pocketfft (FFTW is not installed);
loops that are not FFTs (CIC scatter/gather, multigrid) live in the C library
``_build/liboracle.so`` built from ``baorec_oracle.c``.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess
from dataclasses import dataclass, field as dfield
from typing import Optional, Sequence, Tuple

import numpy as np
import scipy.fft as sfft

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

F32 = np.float32
F64 = np.float64


def build() -> str:
    """Compile the C part of the oracle in-tree (``oracle/_build``)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "_build", "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


@contextlib.contextmanager
def planted(k: int):
    """Within the block, the C functions come from ``liboracle_plant<k>.so``:
    the oracle with deliberate mistake k (listed in oracle_impl.h).  Tests use
    it to show that a check fails on a plausible mistake."""
    global _LIB
    path = os.path.join(_HERE, "_build", f"liboracle_plant{int(k)}.so")
    if not os.path.exists(path):
        build()
    saved = lib()
    _LIB = ctypes.CDLL(path)
    try:
        yield
    finally:
        _LIB = saved


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=F32).reshape(3))


def _sfx(dtype):
    return "_f32" if np.dtype(dtype) == np.float32 else "_f64"


def _workers():
    return int(os.environ.get("BAOREC_ORACLE_WORKERS", os.cpu_count() or 1))


# --------------------------------------------------------------- parameters

@dataclass
class IterativeRecon:
    """Mirror of ``IterativeRecon`` (src/recon.jl:2-16)."""
    bias: float
    f: float
    smoothing_radius: float
    box_size: Optional[Sequence[float]] = None
    box_min: Optional[Sequence[float]] = None
    los: Optional[Sequence[float]] = None
    n_iter: int = 3
    beta: Optional[float] = None
    result_cache: Optional[np.ndarray] = None

    def __post_init__(self):
        if self.beta is None:
            self.beta = float(F32(F32(self.f) / F32(self.bias)))


@dataclass
class MultigridRecon:
    """Mirror of ``MultigridRecon`` (src/recon.jl:18-33)."""
    bias: float
    f: float
    smoothing_radius: float
    box_size: Optional[Sequence[float]] = None
    box_min: Optional[Sequence[float]] = None
    los: Optional[Sequence[float]] = None
    jacobi_damping_factor: float = 0.4
    jacobi_niterations: int = 5
    vcycle_niterations: int = 6
    beta: Optional[float] = None
    result_cache: Optional[np.ndarray] = None

    def __post_init__(self):
        if self.beta is None:
            self.beta = float(F32(F32(self.f) / F32(self.bias)))


# ------------------------------------------------------------- grid vectors

def k_vec(dims: Sequence[int], box_size, dtype=F32) -> Tuple[np.ndarray, ...]:
    """src/utils.jl:3-10.  sample_rate = fl32(2*pi*n/L) (computed in Float64),
    step = fl32(sample_rate/n), k[i] = fl32(i*step); x uses rfftfreq (0..n/2),
    y and z use fftfreq (Nyquist negative)."""
    out = []
    for a, n in enumerate(dims):
        L = F32(box_size[a])
        if dtype == F32:
            fs = F32(2 * np.pi * n / float(L))
            mult = F32(fs / F32(n))
        else:
            mult = 2 * np.pi / float(L)
        if a == 0:
            idx = np.arange(n // 2 + 1)
        else:
            idx = np.arange(n)
            idx = np.where(idx < (n + 1) // 2, idx, idx - n)
        out.append((idx.astype(dtype) * dtype(mult)).astype(dtype))
    return tuple(out)


def x_vec(dims: Sequence[int], box_size, box_min, dtype=F32) -> Tuple[np.ndarray, ...]:
    """src/utils.jl:21-25: cell centres b + (i+1/2)*fl32(L/n), rounded to T."""
    out = []
    fn = getattr(lib(), "oracle_xvec" + _sfx(dtype))
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float]
    for a, n in enumerate(dims):
        v = np.zeros(n, dtype=dtype)
        fn(_p(v), int(n), float(box_size[a]), float(box_min[a]))
        out.append(v)
    return tuple(out)


def setup_box(x, y, z, box_pad):
    """src/utils.jl:100-109: cubic box, min - pad/2 .. max + pad/2 (Float32)."""
    pad = F32(box_pad)
    bmin = np.array([np.min(x), np.min(y), np.min(z)], dtype=F32) - pad / F32(2)
    bmax = np.array([np.max(x), np.max(y), np.max(z)], dtype=F32) + pad / F32(2)
    L = np.max((bmax - bmin).astype(F32))
    return np.full(3, L, dtype=F32), bmin.astype(F32)


# -------------------------------------------------------------- mass assign

def cic(rho: np.ndarray, x, y, z, w, box_size, box_min, wrap=True) -> np.ndarray:
    """cic!, src/mas.jl:1-52 (serial; mutates x,y,z in place when wrap)."""
    nz, ny, nx = rho.shape
    fn = getattr(lib(), "oracle_cic" + _sfx(rho.dtype))
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + \
        [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    for a in (x, y, z, w):
        assert a.dtype == F32 and a.flags.c_contiguous
    assert rho.flags.c_contiguous
    bs, bm = _f3(box_size), _f3(box_min)
    rc = fn(_p(rho), nx, ny, nz, _p(x), _p(y), _p(z), _p(w), len(x), _p(bs), _p(bm), int(bool(wrap)))
    if rc != 0:
        raise IndexError("cic!: particle outside the grid (reference BoundsError)")
    return rho


def read_cic(field: np.ndarray, x, y, z, box_size, box_min, gpu_formula=True) -> np.ndarray:
    """read_cic!, src/mas.jl:218-269 (CPU) / :271-318 (GPU formula)."""
    nz, ny, nx = field.shape
    out = np.zeros(len(x), dtype=field.dtype)
    fn = getattr(lib(), "oracle_read_cic" + _sfx(field.dtype))
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    x, y, z = (np.ascontiguousarray(a, dtype=F32) for a in (x, y, z))
    field = np.ascontiguousarray(field)
    bs, bm = _f3(box_size), _f3(box_min)
    rc = fn(_p(out), _p(field), nx, ny, nz, _p(x), _p(y), _p(z), len(x), _p(bs), _p(bm), int(gpu_formula))
    if rc != 0:
        raise IndexError("read_cic!: particle outside the grid")
    return out


# -------------------------------------------------------------------- FFTs

def _rfft(a):
    return sfft.rfftn(a, workers=_workers())


def _irfft(ak, shape):
    return sfft.irfftn(ak, s=shape, workers=_workers())


def _k2(kv, dtype):
    kx, ky, kz = kv
    return (kx[None, None, :] ** 2 + ky[None, :, None] ** 2).astype(dtype) + kz[:, None, None] ** 2


def smooth(field: np.ndarray, R, box_size) -> np.ndarray:
    """smooth!, src/utils.jl:45-56: F_k *= exp(-0.5 R^2 k^2) with the exponent
    in Float64 (the -0.5 literal promotes), product rounded back to T."""
    dt = field.dtype.type
    kv = k_vec(field.shape[::-1], box_size, dt)
    fk = _rfft(field)
    k2 = _k2(kv, dt)
    R2 = dt(R) * dt(R)
    fac = np.exp(-0.5 * np.float64(R2) * k2.astype(F64))
    if dt == F32:
        fk = (fk.astype(np.complex128) * fac).astype(np.complex64)
    else:
        fk = fk * fac
    field[...] = _irfft(fk, field.shape)
    return field


# -------------------------------------------------------------- overdensity

def setup_overdensity(delta: np.ndarray, recon, x, y, z, w, wrap=True) -> np.ndarray:
    """setup_overdensity!, src/recon.jl:42-57 (periodic box, no randoms)."""
    dt = delta.dtype.type
    cic(delta, x, y, z, w, recon.box_size, recon.box_min, wrap=wrap)
    smooth(delta, recon.smoothing_radius, recon.box_size)
    mean = dt(np.mean(delta, dtype=dt))
    delta[...] = ((delta / mean) - dt(1)) / dt(recon.bias)
    return delta


def setup_overdensity_randoms(delta: np.ndarray, recon, x, y, z, w, rx, ry, rz, rw,
                              ran_min=0.01, return_aux=False):
    """setup_overdensity!, src/recon.jl:60-91 (data minus alpha*randoms).
    return_aux=True also returns the smoothed randoms grid and the threshold
    (cells with rho_r within rounding of the threshold are ambiguous: the
    reference's own result there depends on its summation order)."""
    dt = delta.dtype.type
    dran = np.zeros_like(delta)
    cic(delta, x, y, z, w, recon.box_size, recon.box_min, wrap=False)
    cic(dran, rx, ry, rz, rw, recon.box_size, recon.box_min, wrap=False)
    smooth(delta, recon.smoothing_radius, recon.box_size)
    smooth(dran, recon.smoothing_radius, recon.box_size)
    sd = dt(np.sum(delta, dtype=dt))
    sr = dt(np.sum(dran, dtype=dt))
    alpha = dt(sd / sr)
    thr = np.float64(ran_min) * np.float64(sr) / len(rx)
    delta -= alpha * dran
    den = (dt(recon.bias) * alpha) * dran
    with np.errstate(divide="ignore", invalid="ignore"):
        delta[...] = np.where(dran > thr, delta / den, dt(0))
    if return_aux:
        return delta, dran, thr
    return delta


# ------------------------------------------------------------------ iterate

def iterate(delta_r, delta_s, kv, it, beta, los=None, xv=None):
    """iterate!, src/iterative.jl:5-67 (CPU semantics: axes with los_i == 0 are
    skipped, x^2 == 0 cells are set to 0)."""
    dt = delta_r.dtype.type
    shape = delta_r.shape
    dk = _rfft(delta_r)
    k2 = _k2(kv, dt)
    k2 = np.where(k2 == 0, dt(1), k2)
    dk = (dk / k2).astype(dk.dtype)
    dk[0, 0, 0] = 0
    delta_r[...] = delta_s
    beta = dt(beta)
    K = (kv[0][None, None, :], kv[1][None, :, None], kv[2][:, None, None])
    if los is None:
        X = (xv[0][None, None, :], xv[1][None, :, None], xv[2][:, None, None])
        x2 = (X[0] ** 2 + X[1] ** 2).astype(dt) + X[2] ** 2
        for i in range(3):
            for j in range(i, 3):
                fac = dt((1.0 + float(dt(i != j))) * float(beta))
                if it == 1:
                    fac = dt(fac / (dt(1) + beta))
                d = _irfft(((K[i] * K[j]).astype(dt) * dk).astype(dk.dtype), shape).astype(dt)
                with np.errstate(divide="ignore", invalid="ignore"):
                    upd = (((fac * d) * X[i]) * X[j]) / x2
                    delta_r[...] = np.where(x2 > 0, delta_r - upd, dt(0))
    else:
        los = [dt(v) for v in los]
        for i in range(3):
            if los[i] == 0:
                continue
            fac = beta
            if it == 1:
                fac = dt(fac / (dt(1) + beta))
            d = _irfft(((K[i] ** 2).astype(dt) * los[i] * dk).astype(dk.dtype), shape).astype(dt)
            delta_r -= fac * d
    return delta_r


def reconstructed_overdensity(delta, recon, x, y, z, w, rx=None, ry=None, rz=None, rw=None):
    """reconstructed_overdensity!, src/recon.jl:93-132."""
    dt = delta.dtype.type
    if rx is None:
        setup_overdensity(delta, recon, x, y, z, w)
    else:
        setup_overdensity_randoms(delta, recon, x, y, z, w, rx, ry, rz, rw)
    ds = delta.copy()
    dims = delta.shape[::-1]
    kv = k_vec(dims, recon.box_size, dt)
    xv = x_vec(dims, recon.box_size, recon.box_min, dt) if recon.los is None else None
    for it in range(1, recon.n_iter + 1):
        iterate(delta, ds, kv, it, recon.beta, los=recon.los, xv=xv)
    return delta


def run_iterative(recon: IterativeRecon, grid_size, x, y, z, w, rx=None, ry=None, rz=None,
                  rw=None, dtype=F32, box_pad=500.0):
    """run!(::IterativeRecon), src/recon.jl:134-180."""
    nx, ny, nz = grid_size
    delta = np.zeros((nz, ny, nx), dtype=dtype)
    if rx is not None:
        recon.box_size, recon.box_min = setup_box(rx, ry, rz, box_pad)
    reconstructed_overdensity(delta, recon, x, y, z, w, rx, ry, rz, rw)
    recon.result_cache = delta
    return delta


# ---------------------------------------------------------------- multigrid

def _mgfn(name, dtype):
    return getattr(lib(), "oracle_mg_" + name + _sfx(dtype))


def _los_args(los):
    if los is None:
        return 0, _f3([0, 0, 0])
    return 1, _f3(los)


def fmg(f: np.ndarray, box_size, box_min, beta, omega, niter, ncycles, los=None) -> np.ndarray:
    """fmg, src/multigrid.jl:722-752 (returns a fresh v)."""
    dt = f.dtype
    nz, ny, nx = f.shape
    v = np.zeros_like(f)
    fn = _mgfn("fmg", dt)
    R = ctypes.c_float if dt == F32 else ctypes.c_double
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 3 + \
        [ctypes.c_void_p, ctypes.c_void_p, R, R, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    hl, lv = _los_args(los)
    bs, bm = _f3(box_size), _f3(box_min)
    fn(_p(v), _p(np.ascontiguousarray(f)), nx, ny, nz, _p(bs), _p(bm), float(beta), float(omega),
       int(niter), int(ncycles), hl, _p(lv))
    return v


def mg_jacobi(v, f, box_size, box_min, beta, omega, niter, los=None):
    """jacobi!, src/multigrid.jl:14-135 (in place on v)."""
    dt = v.dtype
    nz, ny, nx = v.shape
    fn = _mgfn("jacobi", dt)
    R = ctypes.c_float if dt == F32 else ctypes.c_double
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 3 + \
        [ctypes.c_void_p, ctypes.c_void_p, R, R, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    hl, lv = _los_args(los)
    bs, bm = _f3(box_size), _f3(box_min)
    fn(_p(v), _p(np.ascontiguousarray(f)), nx, ny, nz, _p(bs), _p(bm), float(beta), float(omega),
       int(niter), hl, _p(lv))
    return v


def mg_residual(v, f, box_size, box_min, beta, los=None):
    """residual!, src/multigrid.jl:213-326."""
    dt = v.dtype
    nz, ny, nx = v.shape
    r = np.zeros_like(v)
    fn = _mgfn("residual", dt)
    R = ctypes.c_float if dt == F32 else ctypes.c_double
    fn.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 3 + \
        [ctypes.c_void_p, ctypes.c_void_p, R, ctypes.c_int, ctypes.c_void_p]
    hl, lv = _los_args(los)
    bs, bm = _f3(box_size), _f3(box_min)
    fn(_p(r), _p(np.ascontiguousarray(v)), _p(np.ascontiguousarray(f)), nx, ny, nz, _p(bs), _p(bm),
       float(beta), hl, _p(lv))
    return r


def mg_colors(shape):
    """Colour ((ix+iy)&1) + 2((iy+iz)&1) of every cell of an (nz, ny, nx) grid:
    the RBGS schedule's smoother order (an extension; the reference smooths
    with jacobi!, src/multigrid.jl:193-210)."""
    nz, ny, nx = shape
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((ix + iy) & 1) + 2 * ((iy + iz) & 1)


def mg_gauss_seidel(v, f, box_size, box_min, beta, los=None, sweeps=1):
    """Multicolour Gauss-Seidel sweeps (in place on v) stated through the
    residual: v_i <- (f_i + sum_j a_ij v_j) / d_i = v_i + r_i / d_i with
    r = f - A v of residual! (src/multigrid.jl:213-326), one colour at a time
    (no two cells of a colour are stencil neighbours, so a colour's updates
    are independent).  d_i, the operator's diagonal, is -r_i of the colour's
    indicator with f = 0: its neighbours carry no weight there."""
    col = mg_colors(v.shape)
    zero = np.zeros_like(v)
    diag = [-mg_residual((col == c).astype(v.dtype), zero, box_size, box_min, beta, los) for c in range(4)]
    for _ in range(sweeps):
        for c in range(4):
            m = col == c
            r = mg_residual(v, f, box_size, box_min, beta, los)
            v[m] += r[m] / diag[c][m]
    return v


def mg_prolong(v2h: np.ndarray) -> np.ndarray:
    """prolong!, src/multigrid.jl:398-453 (fresh fine array)."""
    cz, cy, cx = v2h.shape
    v1h = np.zeros((2 * cz, 2 * cy, 2 * cx), dtype=v2h.dtype)
    fn = _mgfn("prolong", v2h.dtype)
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    fn(_p(v1h), _p(np.ascontiguousarray(v2h)), cx, cy, cz)
    return v1h


def mg_reduce(v1h: np.ndarray) -> np.ndarray:
    """reduce!, src/multigrid.jl:520-584."""
    fz, fy, fx = v1h.shape
    v2h = np.zeros((fz // 2, fy // 2, fx // 2), dtype=v1h.dtype)
    fn = _mgfn("reduce", v1h.dtype)
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    fn(_p(v2h), _p(np.ascontiguousarray(v1h)), fx, fy, fz)
    return v2h


def mg_vcycle(v, f, box_size, box_min, beta, omega, niter, los=None):
    """vcycle!, src/multigrid.jl:654-687 (in place on v)."""
    dt = v.dtype
    nz, ny, nx = v.shape
    fn = _mgfn("vcycle", dt)
    R = ctypes.c_float if dt == F32 else ctypes.c_double
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_int] * 3 + \
        [ctypes.c_void_p, ctypes.c_void_p, R, R, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    hl, lv = _los_args(los)
    bs, bm = _f3(box_size), _f3(box_min)
    fn(_p(v), _p(np.ascontiguousarray(f)), nx, ny, nz, _p(bs), _p(bm), float(beta), float(omega),
       int(niter), hl, _p(lv))
    return v


def reconstructed_potential(phi, recon: MultigridRecon, x, y, z, w, rx=None, ry=None, rz=None, rw=None):
    """reconstructed_potential!, src/recon.jl:184-212."""
    delta = np.zeros_like(phi)
    if rx is None:
        setup_overdensity(delta, recon, x, y, z, w)
    else:
        setup_overdensity_randoms(delta, recon, x, y, z, w, rx, ry, rz, rw)
    phi[...] = fmg(delta, recon.box_size, recon.box_min, recon.beta, recon.jacobi_damping_factor,
                   recon.jacobi_niterations, recon.vcycle_niterations, los=recon.los)
    return phi


def run_multigrid(recon: MultigridRecon, grid_size, x, y, z, w, rx=None, ry=None, rz=None,
                  rw=None, dtype=F32, box_pad=500.0):
    """run!(::MultigridRecon), src/recon.jl:215-261."""
    nx, ny, nz = grid_size
    phi = np.zeros((nz, ny, nx), dtype=dtype)
    if rx is not None:
        recon.box_size, recon.box_min = setup_box(rx, ry, rz, box_pad)
    reconstructed_potential(phi, recon, x, y, z, w, rx, ry, rz, rw)
    recon.result_cache = phi
    return phi


# ------------------------------------------------------------ displacements

def displacement_grids(field: np.ndarray, recon) -> Tuple[np.ndarray, ...]:
    """The three real-space displacement grids of compute_displacements:
    iterative (src/iterative.jl:253-275): Psi_k = i k_i delta_k / k^2 (0 at k=0);
    multigrid (src/multigrid.jl:755-773): Psi_k = i k_i phi_k."""
    dt = field.dtype.type
    kv = k_vec(field.shape[::-1], recon.box_size, dt)
    fk = _rfft(field)
    K = (kv[0][None, None, :], kv[1][None, :, None], kv[2][:, None, None])
    out = []
    if isinstance(recon, MultigridRecon):
        for i in range(3):
            pk = (1j * K[i]).astype(fk.dtype) * fk
            out.append(_irfft(pk, field.shape).astype(dt))
    else:
        k2 = _k2(kv, dt)
        for i in range(3):
            with np.errstate(divide="ignore", invalid="ignore"):
                pk = np.where(k2 > 0, ((1j * K[i]).astype(fk.dtype) * fk) / k2, 0).astype(fk.dtype)
            pk[0, 0, 0] = 0
            out.append(_irfft(pk, field.shape).astype(dt))
    return tuple(out)


def compute_displacements(field, x, y, z, recon, gpu_formula=True):
    grids = displacement_grids(field, recon)
    return tuple(read_cic(g, x, y, z, recon.box_size, recon.box_min, gpu_formula) for g in grids)


def read_shifts(recon, x, y, z, field=None, kind="disp", gpu_formula=True):
    """read_shifts, src/recon.jl:264-306 (:disp, :rsd or :sum)."""
    if field is None:
        field = recon.result_cache
    dt = field.dtype.type
    disp = compute_displacements(field, x, y, z, recon, gpu_formula)
    if kind == "disp":
        return disp
    f = dt(recon.f)
    if recon.los is None:
        xx, yy, zz = (np.asarray(a, dtype=dt) for a in (x, y, z))
        dist = np.sqrt((xx * xx + yy * yy) + zz * zz).astype(dt)
        l = (xx / dist, yy / dist, zz / dist)
    else:
        l = tuple(np.full(len(x), dt(v), dtype=dt) for v in recon.los)
    dot = ((disp[0] * l[0]) + (disp[1] * l[1])) + (disp[2] * l[2])
    rsd = tuple(((f * dot) * l[j]).astype(dt) for j in range(3))
    if kind == "rsd":
        return rsd
    if kind == "sum":
        return tuple((disp[j] + rsd[j]).astype(dt) for j in range(3))
    raise ValueError(kind)


def reconstructed_positions(recon, x, y, z, field=None, kind="disp", gpu_formula=True):
    """reconstructed_positions, src/recon.jl:366-388: pos - shift."""
    sh = read_shifts(recon, x, y, z, field, kind, gpu_formula)
    dt = sh[0].dtype.type
    return tuple((np.asarray(p, dtype=dt) - s).astype(dt) for p, s in zip((x, y, z), sh))


# ------------------------------------------------------------- cosmology

@dataclass
class Cosmology:
    """Mirror of ``BAOrec.Cosmology`` (src/cosmo.jl:22-72), Float32 fields."""
    omega_b: float = 0.0225
    omega_c: float = 0.12
    h: float = 0.67
    Neff: float = 3.044
    Omega_k0: float = 0.0
    T_cmb: float = 2.725
    w0: float = -1.0
    wa: float = 0.0
    z_tab_min: float = 0.0
    z_tab_max: float = 3.0
    z_tab_num: int = 100000


C_KMS = 299792.458   # src/cosmo.jl:12


def cosmo_derived(c: Cosmology) -> dict:
    """Derived densities as the Julia constructor computes them
    (src/cosmo.jl:36-58): Float32 fields, photon and neutrino densities in
    Float64 rounded to Float32, Omega_Lambda closing the budget in Float32."""
    f = np.float32
    G, hbar, mpc = 6.67408e-11, 6.582119514e-16, 3.085677581491367e22
    evc2, kb = 1.782661907e-36, 8.6173303e-5
    h = f(c.h)
    h2 = f(h * h)
    rho_g0 = f(3 * 100 ** 2 / (8 * np.pi * G) * hbar ** 3 * (1.0 / (1e-3 * mpc)) ** 2 * (1.0 / evc2)
               * (C_KMS * 1e3) ** 3)
    omega_g0 = f((np.pi ** 2 / 15) * (2.725 * kb) ** 4 / float(rho_g0) * (float(f(c.T_cmb)) / 2.725) ** 4
                 / float(h2))
    omega_nu0 = f(float(f(c.Neff)) * (7 / 8 * (4 / 11) ** (4 / 3)) * float(omega_g0))
    omega_b0 = f(f(c.omega_b) / h2)
    omega_c0 = f(f(c.omega_c) / h2)
    omega_k0 = f(c.Omega_k0)
    omega_l0 = f(f(f(f(f(f(1) - omega_k0) - omega_b0) - omega_c0) - omega_nu0) - omega_g0)
    return dict(omega_nu0=omega_nu0, omega_g0=omega_g0, omega_b0=omega_b0, omega_c0=omega_c0,
                omega_k0=omega_k0, omega_l0=omega_l0, w0=f(c.w0), wa=f(c.wa), h=h, h100=f(h * f(100)))


def inv_hubble(c: Cosmology, z):
    """1/H(z) (src/cosmo.jl:74-85), FP64 with the Float32 parameters."""
    d = cosmo_derived(c)
    z = np.asarray(z, dtype=np.float64)
    zp = 1.0 + z
    a = 1.0 / zp
    with np.errstate(divide="ignore", invalid="ignore"):
        fde = -3 * (1 + float(d["w0"])) + 3 * float(d["wa"]) * ((a - 1) / np.log(a) - 1)
        ol = float(d["omega_l0"]) * zp ** (-fde)
    ol = np.where(zp == 1.0, float(d["omega_l0"]), ol)
    e = np.sqrt(float(d["omega_nu0"]) * zp ** 4 + float(d["omega_g0"]) * zp ** 4 + float(d["omega_b0"]) * zp ** 3
                + float(d["omega_c0"]) * zp ** 3 + float(d["omega_k0"]) * zp ** 2 + ol)
    return 1.0 / (float(d["h100"]) * e)


def comoving_distance(c: Cosmology, z) -> np.ndarray:
    """comoving_distance(c, z) (src/cosmo.jl:85-87): c * quadgk(1/H, 0, z,
    rtol = 1e-8); scipy's adaptive quadrature at the same relative tolerance."""
    from scipy.integrate import quad
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    return np.array([C_KMS * quad(lambda t: float(inv_hubble(c, t)), 0.0, float(zi), epsabs=0.0,
                                  epsrel=1e-10, limit=200)[0] for zi in z])


_TABLES = {}


def distance_table(c: Cosmology):
    """Knots and distances of comoving_distance_interp (src/cosmo.jl:88-93):
    the integral of 1/H from 0 to the first knot by comoving_distance, then
    knot interval by knot interval with an 8-point Gauss-Legendre rule,
    summed in order (1e5 adaptive quadratures would take minutes).  Cached
    per cosmology; the arrays are read-only."""
    import dataclasses
    key = dataclasses.astuple(c)
    if key not in _TABLES:
        zk = np.linspace(c.z_tab_min, c.z_tab_max, c.z_tab_num)
        xg, wg = np.polynomial.legendre.leggauss(8)
        mid, half = 0.5 * (zk[1:] + zk[:-1]), 0.5 * (zk[1:] - zk[:-1])
        seg = C_KMS * half * (inv_hubble(c, mid[:, None] + half[:, None] * xg[None, :]) @ wg)
        dist = comoving_distance(c, zk[:1])[0] + np.concatenate([[0.0], np.cumsum(seg)])
        zk.setflags(write=False)
        dist.setflags(write=False)
        _TABLES[key] = (zk, dist)
    return _TABLES[key]


def _interp(xk, yk, x):
    x = np.asarray(x, dtype=np.float64)
    if np.any(~(x >= xk[0]) | ~(x <= xk[-1])):
        raise IndexError("outside the interpolation table (reference BoundsError)")
    return np.interp(x, xk, yk)


def sky_to_cartesian(ra, dec, red, c: Cosmology, table=None):
    """sky_to_cartesian of examples/lightcone.jl:30-48."""
    f = np.float32
    zk, rk = table if table is not None else distance_table(c)
    dist = _interp(zk, rk, np.asarray(red, f).astype(np.float64))
    hh = f(f(f(c.h) * f(100)) / f(100))
    pif = f(np.pi)
    rar = (np.asarray(ra, f) * pif) / f(180)
    der = (np.asarray(dec, f) * pif) / f(180)
    cd = np.cos(der.astype(np.float64)).astype(f).astype(np.float64)
    sd = np.sin(der.astype(np.float64)).astype(f).astype(np.float64)
    cr = np.cos(rar.astype(np.float64)).astype(f).astype(np.float64)
    sr = np.sin(rar.astype(np.float64)).astype(f).astype(np.float64)
    hd = float(hh)
    return (((dist * cd) * cr * hd).astype(f), ((dist * cd) * sr * hd).astype(f), ((dist * sd) * hd).astype(f))


def cartesian_to_sky(x, y, z, c: Cosmology, table=None):
    """cartesian_to_sky of examples/lightcone.jl:50-80 (RA = rem(lon - 360, 360))."""
    f = np.float32
    zk, rk = table if table is not None else distance_table(c)
    x, y, z = (np.asarray(a, f) for a in (x, y, z))
    r = np.sqrt((x * x + y * y) + z * z) / f(c.h)
    red = _interp(rk, zk, r.astype(np.float64)).astype(f)
    s = np.sqrt(x * x + y * y)
    pif = f(np.pi)
    lon = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(f)
    lat = np.arctan2(z.astype(np.float64), s.astype(np.float64)).astype(f)
    lon = (lon * f(180)) / pif
    lat = (lat * f(180)) / pif
    lon = np.fmod(lon - f(360), f(360))
    return lon, lat, red


# ------------------------------------------------------------- power spectrum

def tsc_grid(x, y, z, w, grid, box_size, box_min):
    """TSC assignment with periodic wrap on the CIC nodes (grid coordinate
    g = (p - b) n / L, node round(g), weights 0.5(0.5-d)^2, 0.75-d^2,
    0.5(0.5+d)^2), in the FP32 steps the GPU kernel uses, summed in FP64.
    TSC has no counterpart in the reference; the steps follow the CIC of
    src/mas.jl:62-64 for the grid coordinate."""
    f = np.float32
    nx, ny, nz = grid
    idx, wts = [], []
    for p, n, L, b in zip((x, y, z), grid, box_size, box_min):
        u = ((np.asarray(p, f) - f(b)) * f(n)) / f(L) + f(1)
        g = u - f(1)
        c = np.floor(g + f(0.5))
        d = g - c
        ww = [f(0.5) * (f(0.5) - d) * (f(0.5) - d), f(0.75) - d * d, f(0.5) * (f(0.5) + d) * (f(0.5) + d)]
        ci = c.astype(np.int64)
        idx.append([np.mod(ci - 1 + k, n) for k in range(3)])
        wts.append(ww)
    wv = np.asarray(w, f)
    rho = np.zeros(nx * ny * nz)
    for k in range(3):
        for j in range(3):
            for i in range(3):
                val = ((wv * wts[0][i]) * wts[1][j]) * wts[2][k]
                flat = idx[0][i] + nx * (idx[1][j] + ny * idx[2][k])
                rho += np.bincount(flat, weights=val.astype(np.float64), minlength=nx * ny * nz)
    return rho.reshape(nz, ny, nx).astype(f)


def power_multipoles(x, y, z, w, grid, box_size, box_min=(0.0, 0.0, 0.0), los=(0.0, 0.0, 1.0),
                     interlace=True, kmin=0.0, dk=0.005, nbins=300):
    """FFT multipole estimator of a periodic box with TSC and interlacing
    (the pypowspec configuration of test_helpers/powspec_auto.conf):
    delta_k = (F1 + F2 e^{i k.H/2}) / 2 / prod sinc^3(k_a H_a / 2), P_l =
    (2l+1) V/W^2 <|delta_k|^2 L_l(mu)> over the modes of each |k| bin (the
    full k-space: Hermitian weights on the half spectrum), P_0 minus the
    shot noise V sum w^2 / W^2.  pypowspec is not available here; this
    restates the standard estimator and is pinned by closed forms."""
    nx, ny, nz = grid
    box_size = [float(np.float32(v)) for v in box_size]
    w = np.ones(len(x), np.float32) if w is None else np.asarray(w, np.float32)
    F1 = np.fft.rfftn(tsc_grid(x, y, z, w, grid, box_size, box_min).astype(np.float64))
    H = [L / n for L, n in zip(box_size, grid)]
    kf = [2 * np.pi / L for L in box_size]
    fx = np.arange(nx // 2 + 1)
    fy = np.where(np.arange(ny) < ny // 2, np.arange(ny), np.arange(ny) - ny)
    fz = np.where(np.arange(nz) < nz // 2, np.arange(nz), np.arange(nz) - nz)
    KX, KY, KZ = fx[None, None, :] * kf[0], fy[None, :, None] * kf[1], fz[:, None, None] * kf[2]
    d = F1
    if interlace:
        bm2 = [float(np.float32(np.float32(b) - np.float32(0.5) * np.float32(np.float32(L) / np.float32(n))))
               for b, L, n in zip(box_min, box_size, grid)]
        F2 = np.fft.rfftn(tsc_grid(x, y, z, w, grid, box_size, bm2).astype(np.float64))
        ph = np.exp(0.5j * (KX * H[0] + KY * H[1] + KZ * H[2]))
        d = 0.5 * (F1 + F2 * ph)

    def sinc(a):
        return np.where(a == 0, 1.0, np.sin(a) / np.where(a == 0, 1.0, a))

    win = sinc(0.5 * KX * H[0]) ** 3 * sinc(0.5 * KY * H[1]) ** 3 * sinc(0.5 * KZ * H[2]) ** 3
    p = np.abs(d) ** 2 / win ** 2
    k = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    lv = np.asarray(los, np.float64)
    lv = lv / np.linalg.norm(lv)
    with np.errstate(invalid="ignore", divide="ignore"):
        mu = (KX * lv[0] + KY * lv[1] + KZ * lv[2]) / k
    mu2 = mu * mu
    l2 = 1.5 * mu2 - 0.5
    l4 = (35 * mu2 * mu2 - 30 * mu2 + 3) / 8
    herm = np.where((fx == 0) | (fx == nx // 2), 1.0, 2.0)[None, None, :] * np.ones_like(k)
    b = np.floor((k - kmin) / dk)
    sel = (k > 0) & (b >= 0) & (b < nbins)
    bi = b[sel].astype(np.int64)
    hw = herm[sel]
    cnt = np.bincount(bi, weights=hw, minlength=nbins)
    sk = np.bincount(bi, weights=hw * k[sel], minlength=nbins)
    s0 = np.bincount(bi, weights=hw * p[sel], minlength=nbins)
    s2 = np.bincount(bi, weights=hw * p[sel] * l2[sel], minlength=nbins)
    s4 = np.bincount(bi, weights=hw * p[sel] * l4[sel], minlength=nbins)
    V = float(np.prod(box_size))
    W = float(np.sum(w, dtype=np.float64))
    W2 = float(np.sum(w.astype(np.float64) ** 2))
    norm, shot = V / W ** 2, V * W2 / W ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        out = dict(k=sk / cnt, nmodes=cnt, p0=norm * s0 / cnt - shot, p2=5 * norm * s2 / cnt,
                   p4=9 * norm * s4 / cnt)
    return out
