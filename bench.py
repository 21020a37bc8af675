#!/usr/bin/env python3
"""Reconstructions per second of the BAOrec.jl hot path on one MI355X.

One step = one reconstruction as the reference's GPU benchmark times it
(test/benchmarks_gpu.jl:114-126, 140-151): run!(recon, 512^3, data, randoms)
followed by reconstructed_positions(recon, data; field=:sum), with the
catalogs already resident in HBM.

Catalogs of the 512^3 lightcone configuration: 2,000,000 lognormal tracers in
the shell r in [1000, 1800] Mpc/h over a 90x90 deg cap with n(r) ~ r^-0.5,
20,000,000 unclustered randoms in the same geometry, box padded 10 %, local
radial LOS, CIC, FP32, R = 15 Mpc/h, b = 2.2, f = 0.75.  Synthetic data (the
reference's mocks are not available).

Default workload lc512_warm: the multigrid reconstruction of these catalogs
(RBGS V-cycles to relative residual 1e-6) started from the learned seed is the
headline value; the same reconstruction from the plain cold start, from the
spectral seed and from the refined spectral seed is timed beside it, and the
iterative reconstruction (3 iterations) of configurations 1 (128^3 box, 1M
tracers, LOS z) and 2 (these catalogs) is timed briefly.

Prints one JSON line (see the "roofline" and "cpu_baseline" fields).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "baorec.jl_amd", "python"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(args):
    from baorec_hip import synth
    t0 = time.perf_counter()
    if args.workload in ("lightcone512", "mg_lc512"):
        data = synth.lightcone_catalog(args.ndata, r_min=1000.0, r_max=1800.0, mesh=256, seed=2)
        rand = synth.lightcone_randoms(args.nrand, r_min=1000.0, r_max=1800.0, seed=3)
        ext = max(float(np.max(a) - np.min(a)) for a in rand[:3])
        cfg = dict(grid=(512, 512, 512), bias=2.2, f=0.75, R=15.0, los=None, pad=0.1 * ext)
        if args.workload == "mg_lc512":
            # the same catalogs through the multigrid solver: RBGS to 1e-6
            # (or its attainable floor), FP64 levels <= 8^3
            cfg.update(recon="multigrid", schedule="rbgs")
    elif args.workload == "box128":
        data = synth.box_catalog(1_000_000, L=1000.0, mesh=256, seed=1)
        rand = None
        cfg = dict(grid=(128, 128, 128), bias=2.0, f=0.8, R=15.0, los=(0.0, 0.0, 1.0), pad=0.0)
    elif args.workload in ("mg512", "mg512_ref", "mg512_zlos"):
        # multigrid on the 1M-tracer periodic-box catalog at 512^3, local LOS
        # from the box corner (LOS z for mg512_zlos); RBGS to relative
        # residual 1e-6 with FP64 levels <= 8^3, or the reference's FMG
        # schedule
        data = synth.box_catalog(1_000_000, L=1000.0, mesh=256, seed=1)
        rand = None
        cfg = dict(grid=(512, 512, 512), bias=2.0, f=0.8, R=15.0,
                   los=(0.0, 0.0, 1.0) if args.workload == "mg512_zlos" else None, pad=0.0, recon="multigrid",
                   schedule="reference" if args.workload == "mg512_ref" else "rbgs")
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    log(f"synthetic catalogs in {time.perf_counter() - t0:.1f} s")
    return data, rand, cfg


WARM_STARTS = {"none": 0, "spectral": 1, "learned": 2, "refined": 3}
NET = os.path.join(ROOT, "baorec.jl_amd", "weights", "unet_lc128.bin")


def gpu_bench(args, data, rand, cfg, warm="none", steps=None, host_leg=True, ctx=None):
    import baorec_hip as bh
    from baorec_hip import _lib as L
    from baorec_hip.devmem import DeviceArray, hip

    steps = args.steps if steps is None else steps
    ctx = ctx or bh.Context(0)
    lib = ctx.lib
    p = L.Params()
    lib.baorec_params_default(ctypes.byref(p), cfg["bias"], cfg["f"], cfg["R"])
    p.box_pad = cfg["pad"]
    multigrid = cfg.get("recon") == "multigrid"
    if multigrid:
        p.mg_schedule = L.MG_SCHEDULE[cfg["schedule"]]
        p.tolerance = 1e-6
        p.rb_sweeps = 2
        p.fp64_coarse_max = 8
        p.mg_warm_start = WARM_STARTS[warm]
        p.mg_seed_steps = args.seed_steps
        if warm == "learned":
            ctx.check(lib.baorec_warmstart_load(ctx.handle, os.fsencode(args.net)))
    if cfg["los"] is not None:
        p.has_los = 1
        for a in range(3):
            p.los[a] = cfg["los"][a]
    else:
        p.has_los = 0
    if rand is None:
        for a in range(3):
            p.box_size[a] = 1000.0
            p.box_min[a] = 0.0
    grid = (ctypes.c_int32 * 3)(*cfg["grid"])
    d = [DeviceArray.from_numpy(a) for a in data]
    r = [DeviceArray.from_numpy(a) for a in rand] if rand is not None else [None] * 4
    n = len(data[0])
    nr = len(rand[0]) if rand is not None else 0
    out = [DeviceArray((n,)) for _ in range(3)]
    # the periodic-box path wraps positions in place; wrapping is idempotent,
    # so repeated steps see the same inputs
    vp = lambda a: None if a is None else ctypes.c_void_p(a.ptr)  # noqa: E731

    run = lib.baorec_run_multigrid if multigrid else lib.baorec_run_iterative
    kind = L.MULTIGRID if multigrid else L.ITERATIVE
    tuned = None
    if args.autotune:
        # tune the scatter of each catalog (its size class) with the box run!
        # would set up from the randoms, and the FFT passes
        q = L.Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(p))
        if rand is not None:
            bs, bm = np.zeros(3, np.float32), np.zeros(3, np.float32)
            ctx.check(lib.baorec_setup_box(ctx.handle, L.MEM_DEVICE, vp(r[0]), vp(r[1]), vp(r[2]), nr,
                                           float(cfg["pad"]), bs.ctypes.data_as(ctypes.c_void_p),
                                           bm.ctypes.data_as(ctypes.c_void_p)))
            for a in range(3):
                q.box_size[a], q.box_min[a] = float(bs[a]), float(bm[a])
        t0 = time.perf_counter()
        for src, nsrc in ((d, n), (r, nr)):
            if nsrc:
                ctx.check(lib.baorec_autotune(ctx.handle, ctypes.byref(q), grid, L.MEM_DEVICE, vp(src[0]),
                                              vp(src[1]), vp(src[2]), vp(src[3]), nsrc))
        tuned = bh.tune_report(ctx)
        log(f"autotuned in {time.perf_counter() - t0:.1f} s: {tuned}")

    def step():
        rc = run(ctx.handle, ctypes.byref(p), grid, L.MEM_DEVICE, vp(d[0]), vp(d[1]),
                 vp(d[2]), vp(d[3]), n, vp(r[0]), vp(r[1]), vp(r[2]), vp(r[3]), nr)
        ctx.check(rc)
        rc = lib.baorec_reconstructed_positions(ctx.handle, ctypes.byref(p), L.MEM_DEVICE, vp(d[0]), vp(d[1]),
                                                vp(d[2]), n, None, grid, kind, L.FIELD["sum"],
                                                vp(out[0]), vp(out[1]), vp(out[2]))
        ctx.check(rc)

    h = hip()
    for i in range(args.warmup):
        step()
        log(f"warmup step {i + 1}/{args.warmup} done")
    h.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step()
    h.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    log(f"{cfg.get('name', args.workload)} ({warm}): timed {steps} steps in {dt:.3f} s")
    if not host_leg:
        pos = [o.numpy() for o in out]
        mg = {"vcycles": p.vcycles_done, "seed_residual": p.seed_residual, "final_residual": p.final_residual,
              "rhs_shift": p.rhs_shift} if multigrid else None
        return dict(dt=dt, steps=steps, pos=pos, mg=mg)

    # the same step with host-resident catalogs and outputs (PCIe transfers
    # included); reported beside the device-resident value, never as it
    hp = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    hd = [np.array(a, np.float32) for a in data]
    hr = [np.ascontiguousarray(a) for a in rand] if rand is not None else [None] * 4
    hout = [np.empty(n, np.float32) for _ in range(3)]

    def step_host():
        rc = run(ctx.handle, ctypes.byref(p), grid, L.MEM_HOST, hp(hd[0]), hp(hd[1]),
                 hp(hd[2]), hp(hd[3]), n, hp(hr[0]), hp(hr[1]), hp(hr[2]), hp(hr[3]), nr)
        ctx.check(rc)
        rc = lib.baorec_reconstructed_positions(ctx.handle, ctypes.byref(p), L.MEM_HOST, hp(hd[0]), hp(hd[1]),
                                                hp(hd[2]), n, None, grid, kind, L.FIELD["sum"],
                                                hp(hout[0]), hp(hout[1]), hp(hout[2]))
        ctx.check(rc)

    step_host()
    nh = 3
    t0 = time.perf_counter()
    for _ in range(nh):
        step_host()
    dt_host = (time.perf_counter() - t0) / nh
    log(f"host-resident step {dt_host * 1e3:.2f} ms")

    # one profiled step: per-kernel HIP-event durations with algorithmic bytes
    lib.baorec_profile(ctx.handle, 1)
    step()
    buf = ctypes.create_string_buffer(1 << 16)
    lib.baorec_profile_report(ctx.handle, buf, len(buf))
    lib.baorec_profile(ctx.handle, 0)
    prof = json.loads(buf.value.decode())
    pos = [o.numpy() for o in out]
    finite = all(np.isfinite(a).all() for a in pos)
    mg = {"schedule": cfg["schedule"], "warm_start": warm,
          "vcycles": p.vcycles_done, "seed_residual": p.seed_residual, "final_residual": p.final_residual,
          "final_change": p.final_change, "rhs_shift": p.rhs_shift} if multigrid else None
    return dict(dt=dt, steps=steps, prof=prof, finite=finite, box=(list(p.box_size), list(p.box_min)),
                dt_host=dt_host, mg=mg, tuned=tuned, pos=pos)


def kernel_symbol(label, grid, rand):
    """Kernel symbol (as rocprofv3 names it) behind one of the library's
    profiler labels; None for labels that cover several kernels."""
    nx, ny, nz = grid
    h = nx // 2
    return {"fft_y": f"k_strided<{ny}>", "fft_z": f"k_strided<{nz}>", "fft_y_pair": f"k_strided2<{ny}>",
            "fft_z_pair": f"k_strided2<{nz}>", "fft_y_iter": f"k_y_iter<{ny}>",
            "fft_x_r2c": f"k_x_r2c<{h}>", "fft_x_c2r": f"k_x_c2r<{h}, 0>",
            "fft_x_c2r_od": f"k_x_c2r<{h}, {2 if rand else 1}>", "fft_x_c2r_iter": f"k_x_c2r<{h}, 3>",
            "fft_x_c2r_disp": f"k_x_c2r<{h}, 5>", "scatter_accum": "k_scatter_accum<0>", "zero": "k_zero4",
            "gather": "k_gather", "mg_jacobi": "k_mg_jacobi", "mg_residual": "k_mg_residual<float>",
            "mg_reduce": "k_mg_reduce<float, float>", "mg_prolong": "k_mg_prolong<float, float>",
            "mg_mcgs": "k_mg_mcgs<float>", "mg_mcgs64": "k_mg_mcgs<double>", "mg_gs4": "k_mg_gs4",
            "mg_refine": "k_mg_refine", "mg_residual_fp64": "k_mg_residual_mixed", "mg_jacobi_tb": "k_mg_jacobi_tb",
            "mg_correct": "k_add_correction<double>",
            "mg_coarse": "k_mg_coarse<float>", "mg_coarse64": "k_mg_coarse<double>"}.get(label)


def measured_traffic(workload):
    """HBM bytes per dispatch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (already scaled by the copy
    calibration); {} when there is none."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return {}
    txt = open(path).read()
    m = re.search(r"```json\n(.*?)\n```", txt, re.S)
    return json.loads(m.group(1))["kernels"] if m else {}


def roofline(prof, grid, workload, rand=True):
    """Roofline entry of the dominant kernel (largest total time over the
    profiled reconstruction): algorithmic bytes per launch / average HIP-event
    launch duration, against the 8 TB/s HBM peak."""
    by_sym = {}
    for label, st in prof.items():
        sym = kernel_symbol(label, grid, rand)
        if sym is None:
            continue
        a = by_sym.setdefault(sym, {"ms": 0.0, "bytes": 0.0, "count": 0, "labels": []})
        a["ms"] += st["ms"]
        a["bytes"] += st["bytes"]
        a["count"] += st["count"]
        a["labels"].append(label)
    sym, st = max(by_sym.items(), key=lambda kv: kv[1]["ms"])
    avg_s = st["ms"] * 1e-3 / st["count"]
    per_launch = st["bytes"] / st["count"]
    achieved = per_launch / avg_s / 1e9
    t = measured_traffic(workload).get(sym)
    traffic = None if t is None else t["read_bytes"] + t["write_bytes"]
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": sym,
            "launches": st["count"], "avg_launch_ms": round(avg_s * 1e3, 4),
            "algorithmic_bytes_per_launch": round(per_launch), "profiler_labels": st["labels"]}


def graph_bench(args, data, cfg, steps, ctx):
    """Configuration 1 through baorec_run_positions_graph: run! and
    reconstructed_positions(:sum) replayed as one HIP graph per step
    (device-resident catalog; the first call captures).  Returns seconds."""
    from baorec_hip import _lib as L
    from baorec_hip.devmem import DeviceArray, hip
    lib = ctx.lib
    p = L.Params()
    lib.baorec_params_default(ctypes.byref(p), cfg["bias"], cfg["f"], cfg["R"])
    p.has_los = 1
    for a in range(3):
        p.los[a] = cfg["los"][a]
        p.box_size[a], p.box_min[a] = 1000.0, 0.0
    grid = (ctypes.c_int32 * 3)(*cfg["grid"])
    d = [DeviceArray.from_numpy(a) for a in data]
    n = len(data[0])
    out = [DeviceArray((n,)) for _ in range(3)]
    vp = lambda a: ctypes.c_void_p(a.ptr)  # noqa: E731

    def step():
        ctx.check(lib.baorec_run_positions_graph(ctx.handle, ctypes.byref(p), grid, *(vp(a) for a in d), n,
                                                 *(vp(o) for o in out)))

    h = hip()
    for _ in range(max(args.warmup, 2)):
        step()
    h.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    h.hipDeviceSynchronize()
    return time.perf_counter() - t0


def cpu_baseline(args, data, rand, cfg):
    """FP32 CPU restatement (oracle) of the reference path, one whole
    reconstruction at the full size: setup (CIC + smoothing + normalisation),
    every iteration, and the shifted-position read-out, each timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import baorec_oracle as orc
    workers = int(os.environ.get("BAOREC_ORACLE_WORKERS", os.cpu_count() or 1))
    os.environ["BAOREC_ORACLE_WORKERS"] = str(workers)
    nx, ny, nz = cfg["grid"]
    x, y, z, w = (a.copy() for a in data)
    rec = orc.IterativeRecon(bias=cfg["bias"], f=cfg["f"], smoothing_radius=cfg["R"], los=cfg["los"])
    t0 = time.perf_counter()
    if rand is not None:
        rec.box_size, rec.box_min = orc.setup_box(rand[0], rand[1], rand[2], cfg["pad"])
        delta = np.zeros((nz, ny, nx), np.float32)
        orc.setup_overdensity_randoms(delta, rec, x, y, z, w, *rand)
    else:
        rec.box_size, rec.box_min = [1000.0] * 3, [0.0] * 3
        delta = np.zeros((nz, ny, nx), np.float32)
        orc.setup_overdensity(delta, rec, x, y, z, w)
    t_setup = time.perf_counter() - t0
    ds = delta.copy()
    kv = orc.k_vec((nx, ny, nz), rec.box_size)
    xv = orc.x_vec((nx, ny, nz), rec.box_size, rec.box_min) if rec.los is None else None
    t_iters = []
    for it in range(1, rec.n_iter + 1):
        t0 = time.perf_counter()
        orc.iterate(delta, ds, kv, it, rec.beta, los=rec.los, xv=xv)
        t_iters.append(time.perf_counter() - t0)
    rec.result_cache = delta
    t0 = time.perf_counter()
    orc.reconstructed_positions(rec, x, y, z, kind="sum")
    t_read = time.perf_counter() - t0
    total = t_setup + sum(t_iters) + t_read
    its = " + ".join(f"{t:.1f}" for t in t_iters)
    log(f"cpu oracle: setup {t_setup:.1f} s, iterations {its} s, read-out {t_read:.1f} s")
    return dict(value=1.0 / total, unit="recons/s", cores=workers, kind="port",
                sample=(f"one whole reconstruction by the FP32 numpy/C oracle at the full {nx}^3 workload: "
                        f"setup {t_setup:.1f} s + {rec.n_iter} iterations {its} s + read-out {t_read:.1f} s"))


def cpu_baseline_multigrid(args, data, rand, cfg):
    """The reference's own multigrid reconstruction on the host: the oracle's
    FP32 restatement of run!(::MultigridRecon) (setup_box, the randoms'
    overdensity, fmg with its fixed schedule of 6 V-cycles per level and 5 + 5
    damped Jacobi sweeps, src/recon.jl:184-261, src/multigrid.jl:722-752) and
    reconstructed_positions(:sum), all timed at the full 512^3 workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import baorec_oracle as orc
    workers = int(os.environ.get("BAOREC_ORACLE_WORKERS", os.cpu_count() or 1))
    os.environ["BAOREC_ORACLE_WORKERS"] = str(workers)
    x, y, z, w = (a.copy() for a in data)
    rec = orc.MultigridRecon(bias=cfg["bias"], f=cfg["f"], smoothing_radius=cfg["R"], los=cfg["los"])
    t0 = time.perf_counter()
    if rand is not None:
        orc.run_multigrid(rec, cfg["grid"], x, y, z, w, *rand, box_pad=cfg["pad"])
    else:
        rec.box_size, rec.box_min = [1000.0] * 3, [0.0] * 3
        orc.run_multigrid(rec, cfg["grid"], x, y, z, w)
    t_run = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.reconstructed_positions(rec, x, y, z, kind="sum")
    t_read = time.perf_counter() - t0
    total = t_run + t_read
    log(f"cpu oracle (multigrid): run! {t_run:.1f} s, read-out {t_read:.1f} s")
    return dict(value=1.0 / total, unit="recons/s", cores=workers, kind="port",
                sample=(f"one whole multigrid reconstruction by the FP32 numpy/C oracle at the full "
                        f"{cfg['grid'][0]}^3 workload ({len(data[0])} tracers"
                        f"{'' if rand is None else f', {len(rand[0])} randoms'}) with the reference's own FMG "
                        f"schedule: run! {t_run:.1f} s + read-out {t_read:.1f} s"))


def readout_bench(args):
    """The drivers' read-outs after a 512^3 lightcone reconstruction
    (examples/lightcone.jl:110-111): reconstructed_positions of the 20M
    randoms with field :sum and with field :disp.  One reconstruction runs
    first; a step is the two read-outs, from the displacement grid the first
    read-out caches (the reference recomputes it per call,
    src/recon.jl:271)."""
    import baorec_hip as bh
    from baorec_hip import _lib as L
    from baorec_hip.devmem import DeviceArray, hip
    args.workload = "lightcone512"
    data, rand, cfg = make_workload(args)
    args.workload = "readout512"
    ctx = bh.Context(0)
    lib = ctx.lib
    p = L.Params()
    lib.baorec_params_default(ctypes.byref(p), cfg["bias"], cfg["f"], cfg["R"])
    p.box_pad = cfg["pad"]
    p.has_los = 0
    grid = (ctypes.c_int32 * 3)(*cfg["grid"])
    d = [DeviceArray.from_numpy(a) for a in data]
    r = [DeviceArray.from_numpy(a) for a in rand]
    n, nr = len(data[0]), len(rand[0])
    out = [DeviceArray((nr,)) for _ in range(3)]
    vp = lambda a: ctypes.c_void_p(a.ptr)  # noqa: E731
    ctx.check(lib.baorec_run_iterative(ctx.handle, ctypes.byref(p), grid, L.MEM_DEVICE, *(vp(a) for a in d), n,
                                       *(vp(a) for a in r), nr))

    def step():
        for field in ("sum", "disp"):
            ctx.check(lib.baorec_reconstructed_positions(ctx.handle, ctypes.byref(p), L.MEM_DEVICE, vp(r[0]),
                                                         vp(r[1]), vp(r[2]), nr, None, grid, L.ITERATIVE,
                                                         L.FIELD[field], *(vp(o) for o in out)))

    h = hip()
    for _ in range(max(args.warmup, 1)):   # the first read-out builds the displacement grid
        step()
    h.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    h.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    lib.baorec_profile(ctx.handle, 1)
    step()
    buf = ctypes.create_string_buffer(1 << 16)
    lib.baorec_profile_report(ctx.handle, buf, len(buf))
    lib.baorec_profile(ctx.handle, 0)
    prof = json.loads(buf.value.decode())
    st = prof["gather"]
    achieved = st["bytes"] / st["count"] / (st["ms"] * 1e-3 / st["count"]) / 1e9
    finite = all(np.isfinite(o.numpy()).all() for o in out)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import baorec_oracle as orc
        workers = int(os.environ.get("BAOREC_ORACLE_WORKERS", os.cpu_count() or 1))
        os.environ["BAOREC_ORACLE_WORKERS"] = str(workers)
        nx, ny, nz = cfg["grid"]
        m = 2_000_000
        rec = orc.IterativeRecon(bias=cfg["bias"], f=cfg["f"], smoothing_radius=cfg["R"], los=None)
        rec.box_size, rec.box_min = orc.setup_box(rand[0], rand[1], rand[2], cfg["pad"])
        # any field of the grid's size: the read-out cost does not depend on it
        rec.result_cache = np.random.default_rng(0).standard_normal((nz, ny, nx)).astype(np.float32)
        t0 = time.perf_counter()
        orc.reconstructed_positions(rec, rand[0][:m].copy(), rand[1][:m].copy(), rand[2][:m].copy(), kind="sum")
        t1 = time.perf_counter()
        cpu = {"value": round(m / (t1 - t0), 1), "unit": "particles/s", "cores": workers, "kind": "port",
               "sample": f"FP32 numpy/C oracle: one :sum read-out of {m} randoms at {nx}^3, recomputing the "
                         f"displacement grid as the reference does per call ({t1 - t0:.1f} s)"}
    line = {
        "metric": "randoms shifted/s, the drivers' two read-outs (:sum, :disp) after a 512\u00b3 lightcone recon",
        "value": round(2 * nr * args.steps / dt, 1), "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "readout512", "grid": list(cfg["grid"]), "n_randoms": nr, "fields": ["sum", "disp"],
                   "displacement_grid": "cached from the first read-out"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_gather",
                     "avg_launch_ms": round(st["ms"] / st["count"], 4)},
        "cpu_baseline": cpu,
        "kernels": {k: {"count": v["count"], "ms": round(v["ms"], 3),
                        "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in prof.items()},
        "outputs_finite": finite,
    }
    print(json.dumps(line), flush=True)


def sky_bench(args):
    """Sky <-> Cartesian conversion of the 22M galaxies of the lightcone
    workload (2M tracers + 20M randoms): one step = sky_to_cartesian then
    cartesian_to_sky (the drivers' step before and after the reconstruction),
    device-resident arrays."""
    import baorec_hip as bh
    from baorec_hip import _lib as L
    from baorec_hip.devmem import DeviceArray, hip
    n = args.ndata + args.nrand
    rng = np.random.default_rng(4)
    ra = rng.uniform(0, 90, n).astype(np.float32)
    dec = np.degrees(np.arcsin(rng.uniform(0, 1, n))).astype(np.float32)
    red = rng.uniform(0.3, 0.8, n).astype(np.float32)
    ctx = bh.Context(0)
    lib = ctx.lib
    cos = bh.Cosmology().struct()
    din = [DeviceArray.from_numpy(a) for a in (ra, dec, red)]
    mid = [DeviceArray((n,)) for _ in range(3)]
    out = [DeviceArray((n,)) for _ in range(3)]
    vp = lambda a: ctypes.c_void_p(a.ptr)  # noqa: E731

    def step():
        ctx.check(lib.baorec_sky_to_cartesian(ctx.handle, ctypes.byref(cos), L.MEM_DEVICE, *(vp(a) for a in din),
                                              n, *(vp(a) for a in mid)))
        ctx.check(lib.baorec_cartesian_to_sky(ctx.handle, ctypes.byref(cos), L.MEM_DEVICE, *(vp(a) for a in mid),
                                              n, *(vp(a) for a in out)))

    h = hip()
    for _ in range(args.warmup):
        step()
    h.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    h.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    lib.baorec_profile(ctx.handle, 1)
    step()
    buf = ctypes.create_string_buffer(1 << 16)
    lib.baorec_profile_report(ctx.handle, buf, len(buf))
    prof = json.loads(buf.value.decode())
    name, st = max(prof.items(), key=lambda kv: kv[1]["ms"])
    achieved = st["bytes"] / st["count"] / (st["ms"] * 1e-3 / st["count"]) / 1e9
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import baorec_oracle as orc
        m = 2_000_000
        o = orc.Cosmology()
        t0 = time.perf_counter()
        table = orc.distance_table(orc.Cosmology(z_tab_num=2001))
        t1 = time.perf_counter()
        xyz = orc.sky_to_cartesian(ra[:m], dec[:m], red[:m], o, table)
        orc.cartesian_to_sky(*xyz, o, table)
        t2 = time.perf_counter()
        cpu = {"value": round(m / (t2 - t1), 1), "unit": "galaxies/s", "cores": 1, "kind": "port",
               "sample": f"numpy oracle on {m} galaxies ({t2 - t1:.1f} s; table of 2001 knots by quadrature "
                         f"in {t1 - t0:.1f} s not counted)"}
    line = {
        "metric": "galaxies/s, sky->Cartesian->sky (22M galaxies)", "value": round(n * args.steps / dt, 1),
        "unit": "galaxies/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64 table", "data": "synthetic",
        "config": {"workload": "sky22m", "n": n, "z_tab_num": 100000},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name,
                     "avg_launch_ms": round(st["ms"] / st["count"], 4)},
        "cpu_baseline": cpu,
        "kernels": {k: {"count": v["count"], "ms": round(v["ms"], 3),
                        "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in prof.items()},
    }
    print(json.dumps(line), flush=True)


def catalog_bench(args):
    """Catalog I/O of the drivers (host code): parse a 2M-row ASCII catalog
    in the reference's format (ra dec z w nz) and write the (n, 4) .npy
    output.  The file repeats a 200k-row block of distinct galaxies ten times."""
    import tempfile
    import baorec_hip as bh
    rng = np.random.default_rng(5)
    m, rep = 200_000, 10
    cols = [rng.uniform(0, 360, m), np.degrees(np.arcsin(rng.uniform(-1, 1, m))), rng.uniform(0.6, 1.1, m),
            np.ones(m), 10 ** rng.uniform(-5, -3, m)]
    cols = [c.astype(np.float32) for c in cols]
    block = "\n".join(" ".join(f"{c[i]:.8g}" for c in cols) for i in range(m)) + "\n"
    d = tempfile.mkdtemp(dir=os.environ.get("TMPDIR", "/tmp"))
    path = os.path.join(d, "catalog.txt")
    with open(path, "w") as fh:
        for _ in range(rep):
            fh.write(block)
    size = os.path.getsize(path)
    bh.read_catalog(path)   # warm the page cache
    times = []
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        got = bh.read_catalog(path)
        times.append(time.perf_counter() - t0)
    t_read = min(times)
    t0 = time.perf_counter()
    bh.write_npy(os.path.join(d, "out.npy"), *got[:4])
    t_write = time.perf_counter() - t0
    cpu = None
    if not args.no_cpu:
        sample = os.path.join(d, "sample.txt")
        with open(sample, "w") as fh:
            fh.write(block)
        t0 = time.perf_counter()
        np.loadtxt(sample, dtype=np.float32)
        t1 = time.perf_counter()
        cpu = {"value": round(m / (t1 - t0), 1), "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"numpy.loadtxt (the oracle's text reader) on the {m}-row block: {t1 - t0:.2f} s"}
    n = m * rep
    line = {"metric": "rows/s, ASCII catalog parse (2M galaxies, 5 columns)", "value": round(n / t_read, 1),
            "unit": "rows/s", "n_gpus": 0, "steps": len(times), "warmup": 1, "ms_per_step": round(t_read * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "catalog2m", "rows": n, "bytes": size, "threads": "min(16, cores)"},
            "roofline": None, "parse_MBps": round(size / t_read / 1e6, 1),
            "npy_write": {"rows": n, "ms": round(t_write * 1e3, 2)}, "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    import shutil
    shutil.rmtree(d, ignore_errors=True)


def fft_bench(args):
    """The 512^3 real transforms against the vendor library: the library's
    baorec_rfft3d / baorec_irfft3d (device-resident, one launch sequence
    each) beside rocFFT through torch.fft.rfftn / irfftn on the same device,
    each timed over the same number of calls.  Algorithmic bytes per 3D
    R2C: the real grid read once and the half spectrum written once (and the
    reverse for C2R)."""
    import torch
    import baorec_hip as bh
    from baorec_hip import _lib as L
    from baorec_hip.devmem import DeviceArray, hip
    n = 512
    ctx = bh.Context(0)
    lib = ctx.lib
    grid = (ctypes.c_int32 * 3)(n, n, n)
    rng = np.random.default_rng(1)
    host = rng.standard_normal((n, n, n), dtype=np.float32)
    x = DeviceArray.from_numpy(host)
    spec = DeviceArray((n * n * (n // 2 + 1) * 2,))
    back = DeviceArray((n * n * n,))
    h = hip()

    def ours(fwd):
        if fwd:
            ctx.check(lib.baorec_rfft3d(ctx.handle, grid, L.MEM_DEVICE, ctypes.c_void_p(x.ptr),
                                        ctypes.c_void_p(spec.ptr)))
        else:
            ctx.check(lib.baorec_irfft3d(ctx.handle, grid, L.MEM_DEVICE, ctypes.c_void_p(spec.ptr),
                                         ctypes.c_void_p(back.ptr)))

    def timed(fn, reps):
        for _ in range(args.warmup):
            fn()
        h.hipDeviceSynchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        h.hipDeviceSynchronize()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    res = {"ours_r2c_ms": timed(lambda: ours(True), args.steps),
           "ours_c2r_ms": timed(lambda: ours(False), args.steps)}
    err = float(np.max(np.abs(back.numpy().reshape(n, n, n) - host)))
    xt = torch.from_numpy(host).cuda()
    holder = {}

    def t_r2c():
        holder["k"] = torch.fft.rfftn(xt)

    def t_c2r():
        holder["x"] = torch.fft.irfftn(holder["k"], s=(n, n, n))

    res["rocfft_r2c_ms"] = timed(t_r2c, args.steps)
    res["rocfft_c2r_ms"] = timed(t_c2r, args.steps)
    err_t = float((holder["x"] - xt).abs().max().item())
    # Gaussian smoothing (smooth!, src/utils.jl:85-96) of a 1000 Mpc/h box
    # at R = 15: the library's fused passes against rfftn * G, irfftn
    R, Lb = 15.0, 1000.0
    p = L.Params()
    lib.baorec_params_default(ctypes.byref(p), 2.0, 0.8, R)
    for a in range(3):
        p.box_size[a], p.box_min[a] = Lb, 0.0
    field = DeviceArray.from_numpy(host)
    res["ours_smooth_ms"] = timed(lambda: ctx.check(lib.baorec_smooth(ctx.handle, ctypes.byref(p), grid, L.MEM_DEVICE,
                                                                      ctypes.c_void_p(field.ptr))), args.steps)
    k = torch.fft.fftfreq(n, d=Lb / n, device="cuda") * 2 * np.pi
    kr = torch.fft.rfftfreq(n, d=Lb / n, device="cuda") * 2 * np.pi
    gauss = torch.exp(-0.5 * R * R * (k[:, None, None] ** 2 + k[None, :, None] ** 2 + kr[None, None, :] ** 2))

    def t_smooth():
        holder["s"] = torch.fft.irfftn(torch.fft.rfftn(xt) * gauss, s=(n, n, n))

    res["rocfft_smooth_ms"] = timed(t_smooth, args.steps)
    nbytes = n ** 3 * 4 + n * n * (n // 2 + 1) * 8
    line = {"metric": "3D R2C + C2R ms, 512^3 FP32 (library vs rocFFT)",
            "value": round(res["ours_r2c_ms"] + res["ours_c2r_ms"], 4), "unit": "ms", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "fft512", "grid": [n, n, n]},
            "library": {k: round(v, 4) for k, v in res.items()},
            "GBps_algorithmic": {k: round(nbytes / (v * 1e-3) / 1e9, 1) for k, v in res.items() if "smooth" not in k},
            "round_trip_max_abs_err": {"library": err, "rocfft": err_t},
            "vendor": "rocFFT via torch " + torch.__version__}
    print(json.dumps(line), flush=True)


def pk_bench(args):
    """Power-spectrum multipoles of a 1M-tracer periodic box at 512^3 with
    TSC and interlacing, dk = 0.005 (the configuration of the reference's
    validation runs, test_helpers/powspec_auto.conf); one step = one call
    with device-resident positions."""
    import baorec_hip as bh
    from baorec_hip import _lib as L
    from baorec_hip import synth
    from baorec_hip.devmem import DeviceArray, hip
    data = synth.box_catalog(1_000_000, L=1000.0, mesh=256, seed=1)
    ctx = bh.Context(0)
    lib = ctx.lib
    n = len(data[0])
    d = [DeviceArray.from_numpy(a) for a in data]
    p = L.PkParams()
    lib.baorec_pk_params_default(ctypes.byref(p))
    nb = p.nbins
    outs = [np.empty(nb) for _ in range(5)]
    grid = (ctypes.c_int32 * 3)(512, 512, 512)
    vp = lambda a: ctypes.c_void_p(a.ptr)  # noqa: E731
    hp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731

    def step():
        ctx.check(lib.baorec_power_multipoles(ctx.handle, ctypes.byref(p), grid, L.MEM_DEVICE, *(vp(a) for a in d),
                                              n, *(hp(o) for o in outs)))

    h = hip()
    for _ in range(args.warmup):
        step()
    h.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dt = time.perf_counter() - t0
    lib.baorec_profile(ctx.handle, 1)
    step()
    buf = ctypes.create_string_buffer(1 << 16)
    lib.baorec_profile_report(ctx.handle, buf, len(buf))
    prof = json.loads(buf.value.decode())
    name, st = max(prof.items(), key=lambda kv: kv[1]["ms"])
    achieved = st["bytes"] / (st["ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import baorec_oracle as orc
        t0 = time.perf_counter()
        orc.power_multipoles(*data, (512, 512, 512), (1000.0,) * 3, dk=0.005, nbins=nb)
        t1 = time.perf_counter()
        cpu = {"value": round(1.0 / (t1 - t0), 4), "unit": "spectra/s", "cores": 1, "kind": "port",
               "sample": f"numpy oracle (TSC by bincount, FP64 rfftn) at the full size: {t1 - t0:.1f} s"}
    line = {"metric": "power spectra/s (P0, P2, P4; 512^3, 1M tracers, TSC + interlacing)",
            "value": round(args.steps / dt, 3), "unit": "spectra/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 grids, f64 sums", "data": "synthetic",
            "config": {"workload": "pk512", "grid": [512, 512, 512], "n": n, "dk": 0.005, "nbins": nb},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel_family": name,
                         "ms": round(st["ms"], 4)},
            "cpu_baseline": cpu,
            "kernels": {k: {"count": v["count"], "ms": round(v["ms"], 3),
                            "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in prof.items()},
            "p0_at_0p1": float(np.interp(0.1, outs[0], outs[2]))}
    print(json.dumps(line), flush=True)


def _box_catalog_job(job):
    from baorec_hip import synth
    n, seed, slope = job
    return synth.box_catalog(n, L=1000.0, mesh=128, seed=seed, slope=slope)


def batch_bench(args):
    """Batched plain reconstruction: 64 periodic-box catalogs (L = 1000 Mpc/h,
    256^3 grids) of sizes drawn log-uniformly in [2e5, 5e6], each a lognormal
    field with its own seed (100..163) and power-law slope drawn uniformly in
    [-2, -1]; iterative FFT, LOS along z, CIC, FP32, 3 iterations.  One step =
    baorec_run_batch over all 64 catalogs (run! + reconstructed_positions with
    field = :sum per catalog), inputs resident in HBM."""
    from concurrent.futures import ProcessPoolExecutor
    ncat = args.ncat
    rng = np.random.default_rng(100)
    sizes = np.exp(rng.uniform(np.log(2e5), np.log(5e6), ncat)).astype(np.int64)
    slopes = rng.uniform(-2.0, -1.0, ncat)
    t0 = time.perf_counter()
    jobs = [(int(sizes[c]), 100 + c, float(slopes[c])) for c in range(ncat)]
    if args.gen_workers > 1:
        # forked workers, started before anything touches the GPU
        with ProcessPoolExecutor(max_workers=args.gen_workers) as ex:
            cats = list(ex.map(_box_catalog_job, jobs))
    else:
        cats = [_box_catalog_job(j) for j in jobs]
    data = [np.concatenate([c[k] for c in cats]) for k in range(4)]
    off = np.zeros(ncat + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    n = int(off[-1])
    log(f"{ncat} catalogs, {n} particles, generated in {time.perf_counter() - t0:.1f} s")

    import baorec_hip as bh
    from baorec_hip import _lib as L
    from baorec_hip.devmem import DeviceArray, hip
    ctx = bh.Context(0)
    lib = ctx.lib
    p = L.Params()
    cfg = dict(grid=(256, 256, 256), bias=2.0, f=0.8, R=15.0, los=(0.0, 0.0, 1.0), pad=0.0)
    lib.baorec_params_default(ctypes.byref(p), cfg["bias"], cfg["f"], cfg["R"])
    p.has_los = 1
    for a in range(3):
        p.los[a] = cfg["los"][a]
        p.box_size[a] = 1000.0
        p.box_min[a] = 0.0
    grid = (ctypes.c_int32 * 3)(*cfg["grid"])
    d = [DeviceArray.from_numpy(a) for a in data]
    out = [DeviceArray((n,)) for _ in range(3)]
    vp = lambda a: ctypes.c_void_p(a.ptr)  # noqa: E731
    offp = off.ctypes.data_as(ctypes.c_void_p)

    def step():
        ctx.check(lib.baorec_run_batch(ctx.handle, ctypes.byref(p), grid, L.MEM_DEVICE, ncat, offp,
                                       *(vp(a) for a in d), *(vp(o) for o in out)))

    tuned = None
    if args.autotune:
        # a batch looks its scatter tile up by the mean catalog size: tune on
        # that many particles (the first catalogs' particles)
        ctx.check(lib.baorec_autotune(ctx.handle, ctypes.byref(p), grid, L.MEM_DEVICE, *(vp(a) for a in d),
                                      int(off[-1]) // ncat))
        tuned = bh.tune_report(ctx)
        log(f"autotuned: {tuned}")

    h = hip()
    for i in range(args.warmup):
        step()
        log(f"warmup step {i + 1}/{args.warmup} done")
    h.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    h.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {dt:.3f} s")
    lib.baorec_profile(ctx.handle, 1)
    step()
    buf = ctypes.create_string_buffer(1 << 16)
    lib.baorec_profile_report(ctx.handle, buf, len(buf))
    lib.baorec_profile(ctx.handle, 0)
    prof = json.loads(buf.value.decode())
    roof = roofline(prof, cfg["grid"], "batch64", rand=False)
    finite = all(np.isfinite(o.numpy()).all() for o in out)
    total_ms = sum(v["ms"] for v in prof.values())
    total_bytes = sum(v["bytes"] for v in prof.values())
    cpu = None
    if not args.no_cpu:
        c0 = [a[off[0]:off[1]] for a in data]
        cpu = cpu_baseline(args, c0, None, cfg)
        cpu["unit"] = "catalogs/s"
        cpu["sample"] = f"catalog 0 of the batch (N = {int(sizes[0])}): " + cpu["sample"]
    line = {"metric": "catalogs/s, batched plain recon (64 periodic boxes, 256\u00b3, N ~ logU[2e5, 5e6])",
            "value": round(ncat * args.steps / dt, 3), "unit": "catalogs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "batch64", "grid": [256, 256, 256], "ncat": ncat, "n_total": n,
                       "n_min": int(sizes.min()), "n_max": int(sizes.max()), "los": "fixed (z)", "n_iter": 3},
            "roofline": roof, "cpu_baseline": cpu,
            "kernels": {k: {"count": v["count"], "ms": round(v["ms"], 3),
                            "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in prof.items()},
            "profiled_step": {"kernel_ms": round(total_ms, 3), "algorithmic_GB": round(total_bytes / 1e9, 3)},
            "outputs_finite": finite, "autotuned": tuned}
    print(json.dumps(line), flush=True)


def copy_calibration(nbytes=1 << 30, reps=5):
    """Plain device-to-device copies of a known size: the reference point for
    the HBM byte counters (FETCH_SIZE / WRITE_SIZE) of the profiled kernels."""
    from baorec_hip.devmem import DeviceArray, hip, D2D
    h = hip()
    a = DeviceArray((nbytes // 4,))
    b = DeviceArray((nbytes // 4,))
    for _ in range(reps):
        rc = h.hipMemcpy(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr), nbytes, D2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
    h.hipDeviceSynchronize()
    print(json.dumps({"calibration_copy_bytes": nbytes, "copies": reps}), flush=True)


def warm_bench(args):
    """The headline: the 512^3 lightcone multigrid reconstruction (RBGS to
    relative residual 1e-6, local LOS) started from the learned seed, against
    the plain cold start and the two non-learned seeds, each a full timed
    run of run! + reconstructed_positions(:sum) on the same device-resident
    catalogs; plus the iterative forms of configurations 1 and 2."""
    import baorec_hip as bh
    if not os.path.exists(args.net):
        raise SystemExit(f"warm-start network {args.net} not found")
    args.workload = "mg_lc512"
    data, rand, cfg = make_workload(args)
    args.workload = "lc512_warm"
    cfg["name"] = "mg_lc512"
    ctx = bh.Context(0)
    forms = {}
    for warm in ("none", "spectral", "refined"):
        forms[warm] = gpu_bench(args, data, rand, cfg, warm=warm, host_leg=False, ctx=ctx)
    g = gpu_bench(args, data, rand, cfg, warm="learned", ctx=ctx)
    forms["learned"] = g
    ref = forms["none"]["pos"]
    comp = {}
    for k, v in forms.items():
        rms = float(np.sqrt(np.mean([np.mean((a.astype(np.float64) - b) ** 2) for a, b in zip(v["pos"], ref)])))
        comp[k] = {"recons_per_s": round(v["steps"] / v["dt"], 3), "ms_per_step": round(v["dt"] / v["steps"] * 1e3, 3),
                   "vcycles": v["mg"]["vcycles"], "seed_residual": v["mg"]["seed_residual"],
                   "final_residual": v["mg"]["final_residual"], "positions_rms_vs_plain_mpch": rms}
    # configurations 1 and 2 (iterative), timed briefly on the same context
    others = {}
    for wl, nsteps in (("lightcone512", 40), ("box128", 200)):
        args.workload = wl
        d2, r2, c2 = (data, rand, dict(grid=(512, 512, 512), bias=2.2, f=0.75, R=15.0, los=None,
                                       pad=cfg["pad"])) if wl == "lightcone512" else make_workload(args)
        c2["name"] = wl
        o = gpu_bench(args, d2, r2, c2, steps=nsteps, host_leg=False, ctx=ctx)
        others[wl] = {"recons_per_s": round(nsteps / o["dt"], 2), "ms_per_step": round(o["dt"] / nsteps * 1e3, 3),
                      "steps": nsteps}
        if wl == "box128":   # the same step replayed as a HIP graph
            dt = graph_bench(args, d2, c2, nsteps, ctx)
            others["box128_graph"] = {"recons_per_s": round(nsteps / dt, 2), "ms_per_step": round(dt / nsteps * 1e3, 3),
                                      "steps": nsteps, "entry": "baorec_run_positions_graph"}
    args.workload = "lc512_warm"
    prof = g["prof"]
    roof = roofline(prof, cfg["grid"], "mg_lc512", True)
    cpu = None if args.no_cpu else cpu_baseline_multigrid(args, data, rand, cfg)
    total_ms = sum(v["ms"] for v in prof.values())
    total_bytes = sum(v["bytes"] for v in prof.values())
    plain = comp["none"]
    line = {
        "metric": "recons/s at 512\u00b3 grid, 2M particles: learned warm-start vs plain at equal residual",
        "value": round(args.steps / g["dt"], 4),
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(g["dt"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 solve, bf16 network",
        "data": "synthetic",
        "config": {"workload": "mg_lc512, learned warm start", "grid": list(cfg["grid"]),
                   "n_data": len(data[0]), "n_randoms": len(rand[0]), "recon": "multigrid (RBGS, 2+2 sweeps, "
                   "FP64 levels <= 8^3)", "tolerance": 1e-6, "los": "local", "box_pad": round(cfg["pad"], 2),
                   "network": os.path.relpath(args.net, ROOT), **g["mg"]},
        "warm_vs_plain": {"speedup": round(comp["learned"]["recons_per_s"] / plain["recons_per_s"], 3),
                          "vcycle_ratio": round(plain["vcycles"] / max(comp["learned"]["vcycles"], 1), 3),
                          "forms": comp},
        "other_configs": others,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": {k: {"count": v["count"], "ms": round(v["ms"], 3),
                        "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in prof.items()},
        "profiled_step": {"kernel_ms": round(total_ms, 3), "algorithmic_GB": round(total_bytes / 1e9, 3)},
        "outputs_finite": g["finite"],
        "pcie_inclusive": {"value": round(1.0 / g["dt_host"], 4), "unit": "recons/s",
                           "note": "catalogs in host memory, copied in and out by the library each step"},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 120 for lc512_warm: at least 3 s per form; 10 otherwise)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="lc512_warm",
                    help="lc512_warm (default: the 512^3 lightcone multigrid recon, learned warm start against "
                         "the plain cold start), lightcone512, mg_lc512, box128, mg512, mg512_zlos, mg512_ref, "
                         "batch64, readout512, sky22m, catalog2m, pk512, fft512")
    ap.add_argument("--ndata", type=int, default=2_000_000)
    ap.add_argument("--nrand", type=int, default=20_000_000)
    ap.add_argument("--ncat", type=int, default=64, help="catalogs of the batch64 workload")
    ap.add_argument("--gen-workers", type=int, default=8,
                    help="processes generating the batch64 catalogs (1 under a profiler)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--warm-start", default="none", choices=sorted(WARM_STARTS),
                    help="multigrid workloads: RBGS start (cold, spectral seed, learned seed, refined spectral)")
    ap.add_argument("--net", default=NET, help="warm-start network file (learned starts)")
    ap.add_argument("--seed-steps", type=int, default=3,
                    help="spectral Richardson steps of the learned and refined starts")
    ap.add_argument("--autotune", action="store_true",
                    help="run baorec_autotune on the workload's catalog before timing")
    ap.add_argument("--graph", action="store_true",
                    help="box128: time baorec_run_positions_graph (run! + read-out replayed as a HIP graph)")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run 1 GiB device-to-device copies (byte-counter calibration)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 120 if args.workload == "lc512_warm" else 10
    if args.calibrate:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    if args.workload == "sky22m":
        sky_bench(args)
        return
    if args.workload == "catalog2m":
        catalog_bench(args)
        return
    if args.workload == "pk512":
        pk_bench(args)
        return
    if args.workload == "fft512":
        fft_bench(args)
        return
    if args.workload == "batch64":
        batch_bench(args)
        return
    if args.workload == "readout512":
        readout_bench(args)
        return
    if args.workload == "lc512_warm":
        warm_bench(args)
        return
    data, rand, cfg = make_workload(args)
    if args.graph:
        if args.workload != "box128":
            raise SystemExit("--graph times the box128 workload")
        import baorec_hip as bh
        dt = graph_bench(args, data, cfg, args.steps, bh.Context(0))
        print(json.dumps({"metric": "recons/s, box128 (run! + reconstructed_positions as a HIP graph)",
                          "value": round(args.steps / dt, 2), "unit": "recons/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
                          "higher_is_better": True, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": "box128", "grid": list(cfg["grid"]), "n_data": len(data[0])}}),
              flush=True)
        return
    g = gpu_bench(args, data, rand, cfg, warm=args.warm_start)
    prof = g["prof"]
    roof = roofline(prof, cfg["grid"], args.workload, rand is not None)
    total_ms = sum(v["ms"] for v in prof.values())
    total_bytes = sum(v["bytes"] for v in prof.values())
    cpu = None
    if not args.no_cpu:
        if cfg.get("recon") == "multigrid":
            cpu = cpu_baseline_multigrid(args, data, rand, cfg)
        else:
            cpu = cpu_baseline(args, data, rand, cfg)
    ms = g["dt"] / args.steps * 1e3
    nx, ny, nz = cfg["grid"]
    multigrid = cfg.get("recon") == "multigrid"
    metric = ("recons/s at 512\u00b3 grid, 2M particles: learned warm-start vs plain at equal residual"
              if args.workload == "lightcone512" else f"recons/s, {args.workload}")
    line = {
        "metric": metric,
        "value": round(args.steps / g["dt"], 4),
        "unit": "recons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload,
                   "form": "plain (cold start)" if args.warm_start == "none" else f"{args.warm_start} warm start",
                   "grid": [nx, ny, nz],
                   "n_data": len(data[0]),
                   "n_randoms": 0 if rand is None else len(rand[0]),
                   "recon": "multigrid" if multigrid else "iterative",
                   "los": "local" if cfg["los"] is None else "fixed", "box_pad": round(cfg["pad"], 2),
                   **(g["mg"] if multigrid else {"n_iter": 3})},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": {k: {"count": v["count"], "ms": round(v["ms"], 3),
                        "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in prof.items()},
        "profiled_step": {"kernel_ms": round(total_ms, 3), "algorithmic_GB": round(total_bytes / 1e9, 3)},
        "outputs_finite": g["finite"],
        "autotuned": g["tuned"],
        "pcie_inclusive": {"value": round(1.0 / g["dt_host"], 4), "unit": "recons/s",
                           "note": "catalogs in host memory, copied in and out by the library each step"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
