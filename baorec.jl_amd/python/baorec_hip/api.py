"""Python mirror of BAOrec.jl's reconstruction API over libbaorec_hip.

Names follow the Julia package (``!`` becomes a trailing ``_``):

    IterativeRecon, MultigridRecon          src/recon.jl:2-33
    run_(recon, grid, x,y,z,w[, rx..rw])    run!, src/recon.jl:134-261
    setup_overdensity_, reconstructed_overdensity_, reconstructed_potential_
    read_shifts, reconstructed_positions    src/recon.jl:264-388
    k_vec, x_vec, setup_box                 src/utils.jl:3-25,100-109
    cic_, read_cic_, smooth_                src/mas.jl, src/utils.jl
    jacobi_, residual_, reduce_, prolong_, vcycle_, fmg   src/multigrid.jl

Particle arrays are float32 numpy arrays (host) or CUDA/HIP torch tensors
(device, used in place).  Grids are numpy arrays of shape (nz, ny, nx), i.e.
the bytes of Julia's column-major (nx, ny, nz) arrays.  Like the reference,
the periodic-box path (no randoms) wraps and overwrites input positions.
Every call runs on the GPU; errors raise BaorecError.
"""
from __future__ import annotations

import ctypes
import os
import threading
from dataclasses import dataclass, field as dfield
from typing import Optional, Sequence

import numpy as np

from . import _lib as L

F32 = np.float32


class BaorecError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"baorec error {code}: {msg}")
        self.code = code


class BoundsError(BaorecError, IndexError):
    """A particle indexes outside the grid (the reference throws BoundsError)."""


class Context:
    """One GPU context (stream + cached buffers + result_cache)."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = ctypes.c_void_p()
        rc = self.lib.baorec_create(ctypes.byref(h), int(device))
        if rc != L.OK:
            raise BaorecError(rc, self.lib.baorec_last_error(None).decode())
        self.handle = h
        self.device = device

    def check(self, rc: int):
        if rc != L.OK:
            msg = self.lib.baorec_last_error(self.handle).decode()
            if rc == L.ERR_OUT_OF_BOX:
                raise BoundsError(rc, msg)
            raise BaorecError(rc, msg)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.baorec_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_tls = threading.local()


def default_context() -> Context:
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context(0)
        _tls.ctx = ctx
    return ctx


# ------------------------------------------------------------------ recon types

@dataclass
class _Recon:
    bias: float
    f: float
    smoothing_radius: float
    box_size: Optional[Sequence[float]] = None
    box_min: Optional[Sequence[float]] = None
    los: Optional[Sequence[float]] = None
    beta: Optional[float] = None
    # extensions (defaults reproduce the reference)
    mas: str = "cic"
    box_pad: float = 500.0
    ran_min: float = 0.01
    mutate_inputs: bool = True
    result_cache: object = None
    ctx: Optional[Context] = dfield(default=None, repr=False)

    def __post_init__(self):
        if self.beta is None:
            self.beta = float(F32(F32(self.f) / F32(self.bias)))

    def context(self) -> Context:
        if self.ctx is None:
            self.ctx = default_context()
        return self.ctx

    def params(self) -> L.Params:
        p = L.Params()
        self.context().lib.baorec_params_default(ctypes.byref(p), float(self.bias), float(self.f),
                                                 float(self.smoothing_radius))
        p.beta = float(self.beta)
        if self.box_size is not None:
            for a in range(3):
                p.box_size[a] = float(self.box_size[a])
        if self.box_min is not None:
            for a in range(3):
                p.box_min[a] = float(self.box_min[a])
        if self.los is not None:
            p.has_los = 1
            for a in range(3):
                p.los[a] = float(self.los[a])
        p.mas = L.MAS[self.mas]
        p.box_pad = float(self.box_pad)
        p.ran_min = float(self.ran_min)
        p.mutate_inputs = int(bool(self.mutate_inputs))
        return p

    def _take_box(self, p: L.Params):
        self.box_size = np.array(list(p.box_size), dtype=F32)
        self.box_min = np.array(list(p.box_min), dtype=F32)


@dataclass
class IterativeRecon(_Recon):
    """IterativeRecon, src/recon.jl:2-16."""
    n_iter: int = 3

    kind = L.ITERATIVE

    def params(self):
        p = super().params()
        p.n_iter = int(self.n_iter)
        return p


@dataclass
class MultigridRecon(_Recon):
    """MultigridRecon, src/recon.jl:18-33 (+ RBGS schedule extension)."""
    jacobi_damping_factor: float = 0.4
    jacobi_niterations: int = 5
    vcycle_niterations: int = 6
    schedule: str = "reference"
    tolerance: float = 1e-6
    max_vcycles: int = 200
    rb_sweeps: int = 2
    fp64_coarse_max: int = 8
    warm_start: str = "none"          # RBGS start: "none" (v = 0), "spectral", "learned" or "refined"
    seed_steps: int = 3               # spectral Richardson steps of the learned and refined starts
    vcycles_done: int = 0
    final_residual: float = float("nan")
    final_change: float = float("nan")
    seed_residual: float = float("nan")
    rhs_shift: float = float("nan")

    kind = L.MULTIGRID

    def params(self):
        p = super().params()
        p.jacobi_damping_factor = float(self.jacobi_damping_factor)
        p.jacobi_niterations = int(self.jacobi_niterations)
        p.vcycle_niterations = int(self.vcycle_niterations)
        p.mg_schedule = L.MG_SCHEDULE[self.schedule]
        p.tolerance = float(self.tolerance)
        p.max_vcycles = int(self.max_vcycles)
        p.rb_sweeps = int(self.rb_sweeps)
        p.fp64_coarse_max = int(self.fp64_coarse_max)
        p.mg_warm_start = {"none": 0, "spectral": 1, "learned": 2, "refined": 3}[self.warm_start]
        p.mg_seed_steps = int(self.seed_steps)
        return p


@dataclass
class ResultCache:
    """recon.result_cache (src/recon.jl:153, 180, 234, 261).  The field lives
    on the GPU inside the context, which keeps only the result of its last
    run!; `generation` identifies this recon's.  `host` is this recon's own
    host copy when run_ made one: read-outs of a result that a later run! on
    the same context replaced use it as an explicit field, as the reference
    reads each recon's own result_cache."""
    ctx: Context
    grid: tuple
    kind: int
    generation: int
    host: Optional[np.ndarray] = dfield(default=None, repr=False)

    def is_current(self) -> bool:
        g = (ctypes.c_int32 * 3)()
        k = ctypes.c_int32()
        gen = ctypes.c_uint64()
        rc = self.ctx.lib.baorec_result_info(self.ctx.handle, g, ctypes.byref(k), ctypes.byref(gen))
        return rc == L.OK and gen.value == self.generation

    def numpy(self) -> np.ndarray:
        if self.host is not None:
            return self.host.copy()
        out = np.empty(self.grid[::-1], dtype=F32)
        self.ctx.check(self.ctx.lib.baorec_result_copy(self.ctx.handle, L.MEM_HOST, _ptr(out),
                                                       ctypes.c_uint64(self.generation)))
        return out


# ------------------------------------------------------------------ arrays

def _is_device(a) -> bool:
    return bool(getattr(a, "is_cuda", False))


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a):
    if a is None:
        return None
    if _is_device(a):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _mem(*arrays) -> int:
    dev = [_is_device(a) for a in arrays if a is not None]
    if any(dev) and not all(dev):
        raise ValueError("mix of host and device arrays")
    return L.MEM_DEVICE if dev and dev[0] else L.MEM_HOST


def _f32(a, writable=False):
    if a is None:
        return None
    if _is_device(a):
        if _is_torch(a):
            import torch
            if a.dtype != torch.float32 or not a.is_contiguous():
                raise TypeError("device arrays must be contiguous float32 tensors")
        elif a.dtype != F32:
            raise TypeError("device arrays must be float32")
        return a
    if not (isinstance(a, np.ndarray) and a.dtype == F32 and a.flags.c_contiguous):
        if writable:
            raise TypeError("positions must be contiguous float32 numpy arrays (they may be wrapped in place)")
        a = np.ascontiguousarray(a, dtype=F32)
    return a


def _grid(g) -> ctypes.Array:
    g = tuple(int(v) for v in g)
    if len(g) != 3:
        raise ValueError("grid_size must have 3 entries")
    return (ctypes.c_int32 * 3)(*g)


def _cat(x, y, z, w, rx=None, ry=None, rz=None, rw=None):
    x, y, z = (_f32(a, writable=True) for a in (x, y, z))
    w = _f32(w)
    n = len(x)
    if not (len(y) == n and len(z) == n and len(w) == n):
        raise ValueError("x, y, z, w lengths differ")
    if rx is not None:
        rx, ry, rz, rw = (_f32(a) for a in (rx, ry, rz, rw))
        nr = len(rx)
        if not (len(ry) == nr and len(rz) == nr and len(rw) == nr):
            raise ValueError("randoms lengths differ")
    else:
        nr = 0
    return x, y, z, w, rx, ry, rz, rw, n, nr


def _new_grid(grid, like):
    nx, ny, nz = grid
    if _is_device(like):
        return _dev_empty((nz, ny, nx), like)
    return np.zeros((nz, ny, nx), dtype=F32)


def _dev_empty(shape, like):
    if _is_torch(like):
        import torch
        return torch.empty(shape, dtype=torch.float32, device=like.device)
    from .devmem import DeviceArray
    return DeviceArray(shape, F32, like.device)


# ------------------------------------------------------------------ run!

def run_(recon: _Recon, grid_size, x, y, z, w, rx=None, ry=None, rz=None, rw=None, copy=True):
    """run!(recon, grid_size, x,y,z,w[, rx,ry,rz,rw]), src/recon.jl:134-261.
    Returns the result grid (delta for iterative, phi for multigrid) as a host
    array when copy=True; the result also stays cached on the GPU."""
    ctx = recon.context()
    x, y, z, w, rx, ry, rz, rw, n, nr = _cat(x, y, z, w, rx, ry, rz, rw)
    mem = _mem(x, y, z, w, rx, ry, rz, rw)
    p = recon.params()
    fn = ctx.lib.baorec_run_iterative if recon.kind == L.ITERATIVE else ctx.lib.baorec_run_multigrid
    ctx.check(fn(ctx.handle, ctypes.byref(p), _grid(grid_size), mem, _ptr(x), _ptr(y), _ptr(z), _ptr(w),
                 n, _ptr(rx), _ptr(ry), _ptr(rz), _ptr(rw), nr))
    if nr > 0:
        recon._take_box(p)
    if isinstance(recon, MultigridRecon):
        recon.vcycles_done = p.vcycles_done
        recon.final_residual = p.final_residual
        recon.final_change = p.final_change
        recon.seed_residual = p.seed_residual
        recon.rhs_shift = p.rhs_shift
    g = (ctypes.c_int32 * 3)()
    k = ctypes.c_int32()
    gen = ctypes.c_uint64()
    ctx.check(ctx.lib.baorec_result_info(ctx.handle, g, ctypes.byref(k), ctypes.byref(gen)))
    cache = ResultCache(ctx, tuple(int(v) for v in grid_size), recon.kind, int(gen.value))
    recon.result_cache = cache
    if not copy:
        return cache
    cache.host = cache.numpy()
    return cache.host.copy()


def _grid_call(name, recon, delta, x, y, z, w, rx, ry, rz, rw, pass_p_mut=False):
    ctx = recon.context()
    x, y, z, w, rx, ry, rz, rw, n, nr = _cat(x, y, z, w, rx, ry, rz, rw)
    mem = _mem(delta, x, y, z, w, rx, ry, rz, rw)
    grid = tuple(delta.shape[::-1])
    p = recon.params()
    ctx.check(getattr(ctx.lib, name)(ctx.handle, ctypes.byref(p), _grid(grid), mem, _ptr(delta), _ptr(x),
                                     _ptr(y), _ptr(z), _ptr(w), n, _ptr(rx), _ptr(ry), _ptr(rz), _ptr(rw), nr))
    if isinstance(recon, MultigridRecon):
        recon.vcycles_done = p.vcycles_done
        recon.final_residual = p.final_residual
        recon.final_change = p.final_change
        recon.seed_residual = p.seed_residual
        recon.rhs_shift = p.rhs_shift
    return delta


def setup_overdensity_(delta, recon, x, y, z, w, rx=None, ry=None, rz=None, rw=None):
    """setup_overdensity!, src/recon.jl:42-91 (delta is overwritten)."""
    return _grid_call("baorec_setup_overdensity", recon, delta, x, y, z, w, rx, ry, rz, rw)


def reconstructed_overdensity_(delta, recon: IterativeRecon, x, y, z, w, rx=None, ry=None, rz=None, rw=None):
    """reconstructed_overdensity!, src/recon.jl:93-132."""
    return _grid_call("baorec_reconstructed_overdensity", recon, delta, x, y, z, w, rx, ry, rz, rw)


def reconstructed_potential_(phi, recon: MultigridRecon, x, y, z, w, rx=None, ry=None, rz=None, rw=None):
    """reconstructed_potential!, src/recon.jl:184-212."""
    return _grid_call("baorec_reconstructed_potential", recon, phi, x, y, z, w, rx, ry, rz, rw)


# ------------------------------------------------------------------ shifts

def _shifts(name, recon, x, y, z, field3d, field):
    ctx = recon.context()
    x, y, z = (_f32(a) for a in (x, y, z))
    n = len(x)
    mem = _mem(x, y, z, field3d)
    if _is_device(x):
        outs = [_dev_empty((n,), x) for _ in range(3)]
    else:
        outs = [np.empty(n, dtype=F32) for _ in range(3)]
    if field3d is None and isinstance(recon.result_cache, ResultCache):
        field3d = recon.result_cache
    if isinstance(field3d, ResultCache) and not field3d.is_current():
        # a later run! on this context replaced the cached result: read this
        # recon's own copy, or refuse rather than read another recon's field
        if field3d.host is None or _is_device(x):
            raise BaorecError(L.ERR_NO_RESULT, "stale result: a later run! on this context replaced "
                                               "the cached result of this recon")
        field3d = field3d.host
    if field3d is None or isinstance(field3d, ResultCache):
        fptr, grid = None, (ctypes.c_int32 * 3)(0, 0, 0)
    else:
        field3d = _f32(field3d)
        fptr, grid = _ptr(field3d), _grid(field3d.shape[::-1])
    p = recon.params()
    ctx.check(getattr(ctx.lib, name)(ctx.handle, ctypes.byref(p), mem, _ptr(x), _ptr(y), _ptr(z), n, fptr,
                                     grid, recon.kind, L.FIELD[field], *(_ptr(o) for o in outs)))
    return tuple(outs)


def read_shifts(recon, x, y, z, field3d=None, field="disp"):
    """read_shifts(recon, x,y,z[, field3d]; field), src/recon.jl:264-364."""
    return _shifts("baorec_read_shifts", recon, x, y, z, field3d, field)


def reconstructed_positions(recon, x, y, z, field3d=None, field="disp"):
    """reconstructed_positions, src/recon.jl:366-388."""
    return _shifts("baorec_reconstructed_positions", recon, x, y, z, field3d, field)


def run_batch(recon: IterativeRecon, grid_size, x, y, z, w, offsets):
    """Extension (no reference counterpart): for every catalog c of a batch,
    run!(recon, grid_size, x,y,z,w) on particles [offsets[c], offsets[c+1])
    followed by reconstructed_positions(recon, x,y,z; field=:sum), all
    catalogs in one launch sequence per group.  Periodic box (recon.box_size
    and recon.box_min set), shared grid and parameters.  Positions are wrapped
    in place as run! does.  Returns the reconstructed positions."""
    if recon.kind != L.ITERATIVE:
        raise ValueError("run_batch reconstructs with IterativeRecon")
    ctx = recon.context()
    x, y, z, w, _, _, _, _, n, _ = _cat(x, y, z, w)
    mem = _mem(x, y, z, w)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    if off.ndim != 1 or len(off) < 1 or off[0] != 0 or off[-1] != n:
        raise ValueError("offsets must run from 0 to the number of particles")
    if _is_device(x):
        outs = [_dev_empty((n,), x) for _ in range(3)]
    else:
        outs = [np.empty(n, dtype=F32) for _ in range(3)]
    p = recon.params()
    ctx.check(ctx.lib.baorec_run_batch(ctx.handle, ctypes.byref(p), _grid(grid_size), mem, len(off) - 1,
                                       off.ctypes.data_as(ctypes.c_void_p), _ptr(x), _ptr(y), _ptr(z), _ptr(w),
                                       *(_ptr(o) for o in outs)))
    return tuple(outs)


def run_positions_graph(recon: IterativeRecon, grid_size, x, y, z, w, out=None):
    """Extension (no reference counterpart): run!(recon, grid_size, x,y,z,w)
    followed by reconstructed_positions(recon, x,y,z; field=:sum), replayed
    as a HIP graph when called again with the same device arrays, sizes and
    parameters (the first call runs eagerly and captures).  Device arrays
    only; periodic box (recon.box_size and recon.box_min set).  Returns the
    reconstructed positions (the arrays of `out` when given)."""
    if recon.kind != L.ITERATIVE:
        raise ValueError("run_positions_graph reconstructs with IterativeRecon")
    if not all(_is_device(a) for a in (x, y, z, w)):
        raise ValueError("run_positions_graph takes device arrays")
    ctx = recon.context()
    n = int(x.shape[0])
    outs = list(out) if out is not None else [_dev_empty((n,), x) for _ in range(3)]
    p = recon.params()
    ctx.check(ctx.lib.baorec_run_positions_graph(ctx.handle, ctypes.byref(p), _grid(grid_size), _ptr(x), _ptr(y),
                                                 _ptr(z), _ptr(w), n, *(_ptr(o) for o in outs)))
    return tuple(outs)


def autotune(recon: _Recon, grid_size, x, y, z, w):
    """Extension (no reference counterpart): times the library's scatter
    tile shapes and strided-pass kernels on this catalog and keeps the
    fastest in the recon's context (results are bitwise the same for every
    choice).  Needs recon.box_size / box_min.  Returns the tuned choices."""
    ctx = recon.context()
    x, y, z, w = (_f32(a) for a in (x, y, z, w))
    n = len(x)
    p = recon.params()
    ctx.check(ctx.lib.baorec_autotune(ctx.handle, ctypes.byref(p), _grid(grid_size), _mem(x, y, z, w), _ptr(x),
                                      _ptr(y), _ptr(z), _ptr(w), n))
    return tune_report(ctx)


def tune_report(ctx: Optional[Context] = None) -> dict:
    """The context's tuned choices (as baorec_tune_save writes them)."""
    import json
    ctx = ctx or default_context()
    need = ctx.lib.baorec_tune_report(ctx.handle, None, 0)
    buf = ctypes.create_string_buffer(int(need))
    ctx.lib.baorec_tune_report(ctx.handle, buf, need)
    return json.loads(buf.value.decode())


def tune_save(path, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    ctx.check(ctx.lib.baorec_tune_save(ctx.handle, os.fsencode(path)))


def tune_load(path, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    ctx.check(ctx.lib.baorec_tune_load(ctx.handle, os.fsencode(path)))


# ------------------------------------------------------------------ utils

def k_vec(grid, box_size):
    """k_vec, src/utils.jl:3-10 (host)."""
    lib = L.load()
    nx, ny, nz = (int(v) for v in grid)
    out = [np.empty(nx // 2 + 1, F32), np.empty(ny, F32), np.empty(nz, F32)]
    bs = np.asarray(box_size, dtype=F32)
    rc = lib.baorec_k_vec(_grid(grid), _ptr(bs), *(_ptr(o) for o in out))
    if rc:
        raise BaorecError(rc, "k_vec")
    return tuple(out)


def x_vec(grid, box_size, box_min):
    """x_vec, src/utils.jl:21-25 (host)."""
    lib = L.load()
    out = [np.empty(int(n), F32) for n in grid]
    bs = np.asarray(box_size, dtype=F32)
    bm = np.asarray(box_min, dtype=F32)
    rc = lib.baorec_x_vec(_grid(grid), _ptr(bs), _ptr(bm), *(_ptr(o) for o in out))
    if rc:
        raise BaorecError(rc, "x_vec")
    return tuple(out)


def setup_box(x, y, z, box_pad, ctx: Optional[Context] = None):
    """setup_box, src/utils.jl:100-109 (GPU reduction)."""
    ctx = ctx or default_context()
    x, y, z = (_f32(a) for a in (x, y, z))
    bs = np.zeros(3, F32)
    bm = np.zeros(3, F32)
    ctx.check(ctx.lib.baorec_setup_box(ctx.handle, _mem(x, y, z), _ptr(x), _ptr(y), _ptr(z), len(x),
                                       float(box_pad), _ptr(bs), _ptr(bm)))
    return bs, bm


def cic_(rho, x, y, z, w, recon_or_box, wrap=True):
    """cic!(rho, x,y,z,w, box_size, box_min; wrap), src/mas.jl:1-107 (adds into rho)."""
    recon = recon_or_box
    ctx = recon.context()
    x, y, z = (_f32(a, writable=True) for a in (x, y, z))
    w = _f32(w)
    p = recon.params()
    ctx.check(ctx.lib.baorec_cic(ctx.handle, ctypes.byref(p), _grid(rho.shape[::-1]), _mem(rho, x, y, z, w),
                                 _ptr(rho), _ptr(x), _ptr(y), _ptr(z), _ptr(w), len(x), int(bool(wrap))))
    return rho


def read_cic_(out, field, x, y, z, recon):
    """read_cic!(out, field, x,y,z, box_size, box_min), src/mas.jl:271-327."""
    ctx = recon.context()
    x, y, z = (_f32(a) for a in (x, y, z))
    field = _f32(field)
    p = recon.params()
    ctx.check(ctx.lib.baorec_read_cic(ctx.handle, ctypes.byref(p), _grid(field.shape[::-1]),
                                      _mem(out, field, x, y, z), _ptr(out), _ptr(field), _ptr(x), _ptr(y),
                                      _ptr(z), len(x)))
    return out


def smooth_(field, recon):
    """smooth!(field, R, box_size, plan), src/utils.jl:85-96 (in place)."""
    ctx = recon.context()
    p = recon.params()
    ctx.check(ctx.lib.baorec_smooth(ctx.handle, ctypes.byref(p), _grid(field.shape[::-1]), _mem(field),
                                    _ptr(field)))
    return field


def rfft3d(field, ctx: Optional[Context] = None):
    """fft_plan * field (R2C over all three axes, x halved)."""
    ctx = ctx or default_context()
    field = _f32(field)
    nz, ny, nx = field.shape
    out = np.empty((nz, ny, nx // 2 + 1), dtype=np.complex64)
    ctx.check(ctx.lib.baorec_rfft3d(ctx.handle, _grid((nx, ny, nz)), L.MEM_HOST, _ptr(field), _ptr(out)))
    return out


def irfft3d(spec, grid, ctx: Optional[Context] = None):
    """ldiv!(field, fft_plan, spec): normalised C2R."""
    ctx = ctx or default_context()
    spec = np.ascontiguousarray(spec, dtype=np.complex64)
    nx, ny, nz = grid
    out = np.empty((nz, ny, nx), dtype=F32)
    ctx.check(ctx.lib.baorec_irfft3d(ctx.handle, _grid(grid), L.MEM_HOST, _ptr(spec), _ptr(out)))
    return out


# ------------------------------------------------------------------ multigrid

def jacobi_(v, f, recon: MultigridRecon):
    """jacobi!, src/multigrid.jl:193-210 (in place on v)."""
    ctx = recon.context()
    p = recon.params()
    ctx.check(ctx.lib.baorec_mg_jacobi(ctx.handle, ctypes.byref(p), _grid(v.shape[::-1]), _mem(v, f), _ptr(v),
                                       _ptr(_f32(f))))
    return v


def gauss_seidel_(v, f, recon: MultigridRecon, sweeps=1, fused=True):
    """The RBGS schedule's smoother (extension): `sweeps` 4-colour
    Gauss-Seidel sweeps in place on v; fused=True runs each sweep as one
    launch (bit for bit the per-colour launches of fused=False)."""
    ctx = recon.context()
    p = recon.params()
    ctx.check(ctx.lib.baorec_mg_gauss_seidel(ctx.handle, ctypes.byref(p), _grid(v.shape[::-1]), _mem(v, f),
                                             _ptr(v), _ptr(_f32(f)), int(sweeps), int(bool(fused))))
    return v


def residual_(r, v, f, recon: MultigridRecon):
    """residual!, src/multigrid.jl:378-395."""
    ctx = recon.context()
    p = recon.params()
    ctx.check(ctx.lib.baorec_mg_residual(ctx.handle, ctypes.byref(p), _grid(v.shape[::-1]), _mem(r, v, f),
                                         _ptr(r), _ptr(_f32(v)), _ptr(_f32(f))))
    return r


def reduce_(coarse, fine, ctx: Optional[Context] = None):
    """reduce!, src/multigrid.jl:641-651."""
    ctx = ctx or default_context()
    ctx.check(ctx.lib.baorec_mg_reduce(ctx.handle, _grid(fine.shape[::-1]), _mem(coarse, fine), _ptr(coarse),
                                       _ptr(_f32(fine))))
    return coarse


def residual_restrict_(coarse, v, f, recon: MultigridRecon, fused=True):
    """reduce!(residual!(v, f)), the V-cycle's step down (src/multigrid.jl:
    700-704), into coarse (half the sides of v).  fused=True runs one pass
    that does not store the residual (sides multiples of 64, 16 and 2), bit
    for bit the two-kernel path of fused=False."""
    ctx = recon.context()
    p = recon.params()
    ctx.check(ctx.lib.baorec_mg_residual_restrict(ctx.handle, ctypes.byref(p), _grid(v.shape[::-1]),
                                                  _mem(coarse, v, f), _ptr(coarse), _ptr(_f32(v)), _ptr(_f32(f)),
                                                  int(bool(fused))))
    return coarse


def prolong_(fine, coarse, add=False, ctx: Optional[Context] = None):
    """prolong!, src/multigrid.jl:511-518 (add=True adds into fine)."""
    ctx = ctx or default_context()
    ctx.check(ctx.lib.baorec_mg_prolong(ctx.handle, _grid(coarse.shape[::-1]), _mem(fine, coarse), _ptr(fine),
                                        _ptr(_f32(coarse)), int(bool(add))))
    return fine


def vcycle_(v, f, recon: MultigridRecon):
    """vcycle!, src/multigrid.jl:689-719 (in place on v)."""
    ctx = recon.context()
    p = recon.params()
    ctx.check(ctx.lib.baorec_mg_vcycle(ctx.handle, ctypes.byref(p), _grid(v.shape[::-1]), _mem(v, f), _ptr(v),
                                       _ptr(_f32(f))))
    return v


def fmg(f, recon: MultigridRecon, v=None):
    """fmg, src/multigrid.jl:722-752 (returns v)."""
    ctx = recon.context()
    if v is None:
        v = _new_grid(f.shape[::-1], f)
    p = recon.params()
    ctx.check(ctx.lib.baorec_mg_fmg(ctx.handle, ctypes.byref(p), _grid(f.shape[::-1]), _mem(v, f), _ptr(v),
                                    _ptr(_f32(f))))
    recon.vcycles_done = p.vcycles_done
    recon.final_residual = p.final_residual
    recon.final_change = p.final_change
    recon.seed_residual = p.seed_residual
    recon.rhs_shift = p.rhs_shift
    return v


# ------------------------------------------------------------------ sky <-> Cartesian

@dataclass
class Cosmology:
    """BAOrec.Cosmology (src/cosmo.jl:22-72): the Float32 parameters; the
    library derives the densities as the Julia constructor does.  The
    comoving-distance table spans [z_tab_min, z_tab_max] with z_tab_num knots."""
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

    @classmethod
    def desi(cls, **kw):
        """DESICosmology() (src/cosmo.jl:60-62)."""
        c = cls(omega_b=0.02237, omega_c=0.1200, h=0.6736, Neff=float(F32(F32(2.0328) + F32(1.0))))
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def struct(self) -> L.Cosmology:
        s = L.Cosmology()
        for name, _ in L.Cosmology._fields_:
            setattr(s, name, getattr(self, name))
        return s


def comoving_distance(z, cosmo: Cosmology, ctx: Optional[Context] = None):
    """comoving_distance_interp(cosmo)(z) (src/cosmo.jl:86-93), Mpc, FP64."""
    ctx = ctx or default_context()
    z = np.ascontiguousarray(z, dtype=np.float64)
    r = np.empty_like(z)
    c = cosmo.struct()
    ctx.check(ctx.lib.baorec_comoving_distance(ctx.handle, ctypes.byref(c), L.MEM_HOST, _ptr(z), z.size, _ptr(r)))
    return r


def _sky(name, a, b, c3, cosmo, ctx):
    ctx = ctx or default_context()
    a, b, c3 = (_f32(v) for v in (a, b, c3))
    n = len(a)
    if not (len(b) == n and len(c3) == n):
        raise ValueError("array lengths differ")
    mem = _mem(a, b, c3)
    out = [np.empty(n, F32) if mem == L.MEM_HOST else _dev_empty((n,), a) for _ in range(3)]
    s = cosmo.struct()
    ctx.check(getattr(ctx.lib, name)(ctx.handle, ctypes.byref(s), mem, _ptr(a), _ptr(b), _ptr(c3), n,
                                     _ptr(out[0]), _ptr(out[1]), _ptr(out[2])))
    return tuple(out)


def sky_to_cartesian(ra, dec, red, cosmo: Cosmology, ctx: Optional[Context] = None):
    """sky_to_cartesian of the reference's drivers (examples/lightcone.jl:30-48):
    RA, Dec in degrees and redshift -> (x, y, z) in Mpc/h."""
    return _sky("baorec_sky_to_cartesian", ra, dec, red, cosmo, ctx)


def cartesian_to_sky(x, y, z, cosmo: Cosmology, ctx: Optional[Context] = None):
    """cartesian_to_sky (examples/lightcone.jl:50-80): (RA, Dec, redshift), with
    RA = rem(lon - 360, 360) as the driver computes it."""
    return _sky("baorec_cartesian_to_sky", x, y, z, cosmo, ctx)


# ------------------------------------------------------------------ learned warm start

def warmstart_nparams() -> int:
    """Number of FP32 parameters of the warm-start network."""
    return int(L.load().baorec_warmstart_nparams())


def warmstart_init(seed: int = 0, h_train: float = 1.0, ctx: Optional[Context] = None):
    """A fresh warm-start network in the context (He-normal weights); h_train
    is the cell size (Mpc/h) of the grids it is trained on."""
    ctx = ctx or default_context()
    ctx.check(ctx.lib.baorec_warmstart_init(ctx.handle, ctypes.c_uint64(int(seed)), float(h_train)))


def warmstart_save(path, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    ctx.check(ctx.lib.baorec_warmstart_save(ctx.handle, os.fsencode(path)))


def warmstart_load(path, ctx: Optional[Context] = None):
    """Loads a network written by warmstart_save into the context; RBGS solves
    with warm_start="learned" then start from its seed."""
    ctx = ctx or default_context()
    ctx.check(ctx.lib.baorec_warmstart_load(ctx.handle, os.fsencode(path)))


def warmstart_params(which: str = "params", ctx: Optional[Context] = None) -> np.ndarray:
    """The network's parameters (which="params") or the gradients of the last
    step / eval (which="grads"), FP32."""
    ctx = ctx or default_context()
    out = np.empty(warmstart_nparams(), F32)
    ctx.check(ctx.lib.baorec_warmstart_params(ctx.handle, {"params": 0, "grads": 2}[which], _ptr(out), out.size))
    return out


def warmstart_set_params(params, ctx: Optional[Context] = None):
    ctx = ctx or default_context()
    a = np.ascontiguousarray(params, dtype=F32)
    ctx.check(ctx.lib.baorec_warmstart_params(ctx.handle, 1, _ptr(a), a.size))


def warmstart_step(recon: MultigridRecon, f, phi, lr: float = 1e-3) -> float:
    """One training step on a batch of (f, phi) pairs, arrays of shape
    (nbatch, nz, ny, nx) (or one grid); phi is the converged local-LOS
    potential of f.  lr = 0 computes the loss and gradients only.  Returns the
    loss."""
    ctx = recon.context()
    f = _f32(f)
    phi = _f32(phi)
    shape = f.shape if f.ndim == 4 else (1,) + tuple(f.shape)
    if tuple(phi.shape if phi.ndim == 4 else (1,) + tuple(phi.shape)) != tuple(shape):
        raise ValueError("f and phi shapes differ")
    loss = ctypes.c_double()
    p = recon.params()
    ctx.check(ctx.lib.baorec_warmstart_step(ctx.handle, ctypes.byref(p), _grid(shape[1:][::-1]), _mem(f, phi),
                                            _ptr(f), _ptr(phi), shape[0], float(lr), ctypes.byref(loss)))
    return loss.value


def warmstart_eval(features, target, ctx: Optional[Context] = None):
    """The network on explicit features (uint16 BF16 bit patterns of shape
    (nbatch, nz, ny, nx, 8)) against a target (nbatch, nz, ny, nx): returns
    (output, loss); the parameter gradients are then warmstart_params("grads")."""
    ctx = ctx or default_context()
    X = np.ascontiguousarray(features, dtype=np.uint16)
    t = np.ascontiguousarray(target, dtype=F32)
    if X.ndim != 5 or X.shape[-1] != 8 or X.shape[:4] != t.shape:
        raise ValueError("features must be (nbatch, nz, ny, nx, 8) and match the target")
    out = np.empty(t.shape, F32)
    loss = ctypes.c_double()
    ctx.check(ctx.lib.baorec_warmstart_eval(ctx.handle, _grid(t.shape[1:][::-1]), L.MEM_HOST,
                                            X.ctypes.data_as(ctypes.c_void_p), _ptr(t), t.shape[0], _ptr(out),
                                            ctypes.byref(loss)))
    return out, loss.value


def warmstart_tensor(kind: str, index: int, shape, ctx: Optional[Context] = None) -> np.ndarray:
    """A BF16 tensor of the last warmstart_eval / warmstart_step as uint16 bit
    patterns of the given shape: kind "act" (layer output), "grad" (its
    gradient after the ReLU mask), "pool", "pool_grad" or "input"."""
    ctx = ctx or default_context()
    out = np.empty(shape, np.uint16)
    k = {"act": 0, "grad": 1, "pool": 2, "pool_grad": 3, "input": 4}[kind]
    ctx.check(ctx.lib.baorec_warmstart_tensor(ctx.handle, k, int(index), out.ctypes.data_as(ctypes.c_void_p),
                                              out.size))
    return out


def warmstart_predict(recon: MultigridRecon, f):
    """The learned seed (spectral seed + the network's correction) of a
    multigrid solve of f on its grid."""
    ctx = recon.context()
    f = _f32(f)
    out = _new_grid(f.shape[::-1], f)
    p = recon.params()
    ctx.check(ctx.lib.baorec_warmstart_predict(ctx.handle, ctypes.byref(p), _grid(f.shape[::-1]), _mem(f, out),
                                               _ptr(f), _ptr(out)))
    return out


# ------------------------------------------------------------------ catalog I/O

def _io_check(lib, rc):
    if rc != L.OK:
        raise BaorecError(rc, lib.baorec_io_last_error().decode())


def read_catalog(path, columns=(0, 1, 2, 3, 4)):
    """An ASCII catalog as the reference's drivers read it (CSV.File(path;
    delim=" ", types=Float32), test/benchmarks_cpu.jl:22-23): the selected
    0-based columns as float32 arrays."""
    lib = L.load()
    bpath = os.fsencode(path)
    n = ctypes.c_int64()
    _io_check(lib, lib.baorec_catalog_rows(bpath, ctypes.byref(n), None))
    cols = [int(c) for c in columns]
    out = [np.empty(n.value, F32) for _ in cols]
    ptrs = (ctypes.c_void_p * len(cols))(*[o.ctypes.data for o in out])
    sel = (ctypes.c_int32 * len(cols))(*cols)
    got = ctypes.c_int64()
    _io_check(lib, lib.baorec_catalog_read(bpath, len(cols), sel, ptrs, n.value, ctypes.byref(got)))
    return tuple(o[:got.value] for o in out)


def write_npy(path, *columns):
    """npzwrite(path, hcat(columns...)) of the drivers (examples/lightcone.jl:
    125-127): an (n, ncols) float32 .npy in Fortran order."""
    lib = L.load()
    cols = [np.ascontiguousarray(c, dtype=F32) for c in columns]
    n = len(cols[0])
    if any(len(c) != n for c in cols):
        raise ValueError("column lengths differ")
    ptrs = (ctypes.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    _io_check(lib, lib.baorec_write_npy(os.fsencode(path), len(cols), ptrs, n))


# ------------------------------------------------------------------ power spectrum

def power_multipoles(x, y, z, w=None, grid=(512, 512, 512), box_size=1000.0, box_min=0.0, los=(0.0, 0.0, 1.0),
                     mas="tsc", interlace=True, kmin=0.0, dk=0.005, nbins=300, ctx: Optional[Context] = None):
    """P_0, P_2, P_4 of a periodic box catalog (TSC, interlacing; the
    configuration of the reference's validation, test_helpers/
    powspec_auto.conf).  Returns a dict of per-bin arrays k (mean |k| of the
    modes), nmodes, p0 (shot noise subtracted), p2, p4."""
    ctx = ctx or default_context()
    x, y, z = (_f32(a) for a in (x, y, z))
    w = _f32(w)
    n = len(x)
    if not (len(y) == n and len(z) == n and (w is None or len(w) == n)):
        raise ValueError("array lengths differ")
    p = L.PkParams()
    ctx.lib.baorec_pk_params_default(ctypes.byref(p))
    bs = np.broadcast_to(np.asarray(box_size, np.float64), (3,))
    bm = np.broadcast_to(np.asarray(box_min, np.float64), (3,))
    for a in range(3):
        p.box_size[a] = bs[a]
        p.box_min[a] = bm[a]
        p.los[a] = los[a]
    p.mas = L.MAS[mas]
    p.interlace = int(bool(interlace))
    p.kmin, p.dk, p.nbins = float(kmin), float(dk), int(nbins)
    out = {k: np.empty(nbins) for k in ("k", "nmodes", "p0", "p2", "p4")}
    ctx.check(ctx.lib.baorec_power_multipoles(ctx.handle, ctypes.byref(p), _grid(grid), _mem(x, y, z, w), _ptr(x),
                                              _ptr(y), _ptr(z), _ptr(w), n, *(_ptr(out[k]) for k in
                                                                             ("k", "nmodes", "p0", "p2", "p4"))))
    return out
