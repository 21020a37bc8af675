"""Training of the learned warm start (host driver; the network, its training
step and the solver are the library's kernels).

The network predicts, from the right-hand side f of the local-LOS multigrid
solve, the correction that turns the spectral seed (the exact inverse of the
operator with the LOS fixed at the box centre) into the converged local-LOS
potential.  Its training pairs come from the plain solver itself:

* lightcone realizations: tracers in the shell r in [1000, 1800] Mpc/h over a
  90 x 90 degree cap with n(r) ~ r^-0.5, from lognormal fields with power-law
  slopes drawn uniformly in [-2, -1] and sizes drawn log-uniformly in
  [5e5, 3e6] (each with its own seed), and one shared set of unclustered
  randoms, so every realization has the same box (setup_box with a 10 % pad);
* f = setup_overdensity!(randoms) at grid^3 (bias 2.2, f 0.75, smoothing
  15 Mpc/h), phi = the RBGS solve of f to relative residual 1e-6 (cold start).

Training: Adam on batches of these pairs, loss = mean square of the
finite-difference gradient of (network output - target), i.e. of the
displacement the seed implies.  Run as a module on a GPU host:

    python -m baorec_hip.warmstart --realizations 2000 --grid 128 --out net.bin
"""
from __future__ import annotations

import argparse
import ctypes
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

from . import _lib as L
from . import api as bh
from . import synth
from .devmem import D2D, DeviceArray, hip

F32 = np.float32
SHELL = dict(r_min=1000.0, r_max=1800.0)
RECON = dict(bias=2.2, f=0.75, smoothing_radius=15.0)


def _realization(args):
    seed, n, slope, mesh = args
    return synth.lightcone_catalog(int(n), mesh=mesh, seed=int(seed), slope=float(slope), **SHELL)


def realization_plan(count, seed0, n_min=5e5, n_max=3e6):
    """(seed, size, slope) of each realization: sizes log-uniform, slopes
    uniform in [-2, -1]."""
    rng = np.random.default_rng(seed0)
    sizes = np.exp(rng.uniform(np.log(n_min), np.log(n_max), count)).astype(np.int64)
    slopes = rng.uniform(-2.0, -1.0, count)
    return [(seed0 + 1 + i, sizes[i], slopes[i]) for i in range(count)]


class TrainingSet:
    """(f, phi) pairs of grid^3 cells on the device: f[i] and phi[i] are
    consecutive grids of two device arrays, so a batch of consecutive
    realizations is one contiguous block."""

    def __init__(self, count, grid, ctx, pool, seed0=10000, n_randoms=5_000_000, mesh=128, log=print):
        self.count, self.grid, self.ctx = count, grid, ctx
        rand = synth.lightcone_randoms(n_randoms, seed=seed0, **SHELL)
        ext = max(float(np.max(a) - np.min(a)) for a in rand[:3])
        self.box_size, self.box_min = bh.setup_box(*rand[:3], 0.1 * ext, ctx=ctx)
        self.randoms = [DeviceArray.from_numpy(a, ctx.device) for a in rand]
        n3 = grid ** 3
        self.F = DeviceArray((count, grid, grid, grid), F32, ctx.device)
        self.PHI = DeviceArray((count, grid, grid, grid), F32, ctx.device)
        self.vcycles = np.zeros(count, np.int32)
        plan = [(s, n, sl, mesh) for s, n, sl in realization_plan(count, seed0)]
        t0 = time.time()
        last = t0
        for i, cat in enumerate(pool.imap(_realization, plan, chunksize=2)):
            self._add(i, cat)
            if time.time() - last > 30 or i == count - 1:
                last = time.time()
                log(f"training set: {i + 1}/{count} realizations, {last - t0:.0f} s "
                    f"(mean V-cycles {self.vcycles[:i + 1].mean():.1f})")
        self.nbytes = 2 * count * n3 * 4

    def recon(self, **kw):
        return bh.MultigridRecon(**RECON, box_size=list(self.box_size), box_min=list(self.box_min), los=None,
                                 schedule="rbgs", tolerance=1e-6, ctx=self.ctx, **kw)

    def _add(self, i, cat):
        g, n3 = self.grid, self.grid ** 3
        data = [DeviceArray.from_numpy(a, self.ctx.device) for a in cat]
        f = DeviceArray((g, g, g), F32, self.ctx.device)
        rec = self.recon()
        bh.setup_overdensity_(f, rec, *data, *self.randoms)
        phi = bh.fmg(f, rec)
        if not rec.final_residual <= 1e-6:
            raise RuntimeError(f"realization {i}: the plain solve stopped at {rec.final_residual:.2e}")
        self.vcycles[i] = rec.vcycles_done
        h = hip()
        for dst, src in ((self.F, f), (self.PHI, phi)):
            rc = h.hipMemcpy(ctypes.c_void_p(dst.ptr + i * n3 * 4), ctypes.c_void_p(src.ptr), n3 * 4, D2D)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy failed: {rc}")

    def step(self, start, nbatch, lr):
        """One training step on realizations [start, start + nbatch)."""
        ctx = self.ctx
        n3 = self.grid ** 3
        p = self.recon().params()
        loss = ctypes.c_double()
        ctx.check(ctx.lib.baorec_warmstart_step(ctx.handle, ctypes.byref(p), bh._grid((self.grid,) * 3),
                                                L.MEM_DEVICE, ctypes.c_void_p(self.F.ptr + start * n3 * 4),
                                                ctypes.c_void_p(self.PHI.ptr + start * n3 * 4), nbatch, float(lr),
                                                ctypes.byref(loss)))
        return loss.value



def _slice(arr, i, g):
    n3 = g ** 3
    out = np.empty((g, g, g), F32)
    rc = hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(arr.ptr + i * n3 * 4), n3 * 4, 2)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    return out


def train(ts: TrainingSet, epochs=6, batch=4, lr=1e-3, lr_min=2e-5, seed=0, holdout=0, log=print):
    """Adam with a cosine learning-rate decay over batches of consecutive
    realizations in random order; the last `holdout` realizations are kept
    out.  Returns the per-epoch mean losses (and held-out losses)."""
    bh.warmstart_init(seed=seed, h_train=float(ts.box_size[0]) / ts.grid, ctx=ts.ctx)
    ntrain = ts.count - holdout
    nb = ntrain // batch
    total = epochs * nb
    rng = np.random.default_rng(seed)
    hist = []
    k = 0
    t0 = time.time()
    last = t0
    for ep in range(epochs):
        losses = []
        for b in rng.permutation(nb):
            rate = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * k / max(1, total - 1)))
            losses.append(ts.step(int(b) * batch, batch, rate))
            k += 1
            if time.time() - last > 30:
                last = time.time()
                log(f"epoch {ep} step {k}/{total}: loss {np.mean(losses[-20:]):.4e}, lr {rate:.2e}, "
                    f"{last - t0:.0f} s")
        held = [ts.step(i, 1, 0.0) for i in range(ntrain, ts.count)] if holdout else []
        hist.append((float(np.mean(losses)), float(np.mean(held)) if held else float("nan")))
        log(f"epoch {ep}: train loss {hist[-1][0]:.4e}, held-out loss {hist[-1][1]:.4e}, {time.time() - t0:.0f} s")
    return hist


def compare_seeds(ts: TrainingSet, indices, log=print):
    """Plain (cold), spectral and learned starts of the RBGS solve on the
    given realizations: V-cycles and seed residuals."""
    out = []
    for i in indices:
        f = _slice(ts.F, i, ts.grid)
        row = {}
        for ws in ("none", "spectral", "refined", "learned"):
            rec = ts.recon(warm_start=ws)
            bh.fmg(f, rec)
            row[ws] = (rec.vcycles_done, rec.seed_residual, rec.final_residual)
        out.append(row)
        log(f"realization {i}: " + ", ".join(f"{k} seed {v[1]:.2e} -> {v[0]} V-cycles" for k, v in row.items()))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--realizations", type=int, default=2000)
    ap.add_argument("--holdout", type=int, default=50)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--randoms", type=int, default=5_000_000)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    # the catalog workers start before this process touches the GPU
    workers = max(1, min(14, (os.cpu_count() or 2) - 2))
    pool = mp.get_context("fork").Pool(workers)
    ctx = bh.default_context()
    t0 = time.time()
    ts = TrainingSet(a.realizations, a.grid, ctx, pool, n_randoms=a.randoms, log=log)
    pool.close()
    pool.join()
    log(f"training set ready: {a.realizations} realizations at {a.grid}^3, box {ts.box_size[0]:.1f} Mpc/h, "
        f"{time.time() - t0:.0f} s")
    hist = train(ts, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, holdout=a.holdout, log=log)
    bh.warmstart_save(a.out, ctx=ctx)
    log(f"saved {a.out}")
    rows = compare_seeds(ts, range(a.realizations - min(a.holdout, 10), a.realizations), log=log)
    for ws in ("none", "spectral", "refined", "learned"):
        log(f"held-out mean: {ws}: {np.mean([r[ws][0] for r in rows]):.2f} V-cycles, "
            f"seed residual {np.mean([r[ws][1] for r in rows]):.3e}")
    return hist


if __name__ == "__main__":
    sys.exit(0 if main() is not None else 1)
