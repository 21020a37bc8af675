"""run! + reconstructed_positions replayed as a HIP graph: every replay
against the eager calls (bit for bit), with eager work and synchronisations
between the replays, a recapture after the library's buffers are
reallocated, and new arguments captured separately."""
import ctypes

import numpy as np
import pytest

import baorec_hip as bh
from baorec_hip import synth
from baorec_hip.devmem import H2D, DeviceArray, hip

pytestmark = pytest.mark.gpu

KW = dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=[0.0] * 3, los=[0.0, 0.0, 1.0])


def device_catalog(n, seed):
    return [DeviceArray.from_numpy(a) for a in synth.box_catalog(n, L=1000.0, mesh=64, seed=seed)]


def eager(ctx, grid, cat):
    rec = bh.IterativeRecon(**KW, ctx=ctx)
    bh.run_(rec, grid, *cat, copy=False)
    return [a.numpy() for a in bh.reconstructed_positions(rec, *cat[:3], field="sum")]


def test_graph_replays_match_eager(ctx):
    grid = (128, 128, 128)
    cat = device_catalog(300_000, 5)
    other = device_catalog(100_000, 6)
    exp = eager(ctx, grid, cat)
    rec = bh.IterativeRecon(**KW, ctx=ctx)
    outs = [DeviceArray((300_000,)) for _ in range(3)]
    for i in range(4):
        got = bh.run_positions_graph(rec, grid, *cat, out=outs)
        for a, b in zip(got, exp):
            np.testing.assert_array_equal(a.numpy(), b)
        eager(ctx, (64, 64, 64), other)   # eager work of another size in between
    # a run that grows the library's buffers invalidates the graph: recaptured
    eager(ctx, (256, 256, 256), other)
    for i in range(2):
        got = bh.run_positions_graph(rec, grid, *cat, out=outs)
        for a, b in zip(got, exp):
            np.testing.assert_array_equal(a.numpy(), b)
    # other arguments (another grid) get their own graph
    exp64 = eager(ctx, (64, 64, 64), cat)
    for i in range(2):
        got = bh.run_positions_graph(rec, (64, 64, 64), *cat, out=outs)
        for a, b in zip(got, exp64):
            np.testing.assert_array_equal(a.numpy(), b)
    got = bh.run_positions_graph(rec, grid, *cat, out=outs)
    for a, b in zip(got, exp):
        np.testing.assert_array_equal(a.numpy(), b)


def test_graph_reports_out_of_box(ctx):
    """A particle outside the box raises on a replay as on the eager call;
    with the catalog restored the same graph runs clean again."""
    grid = (64, 64, 64)
    x, y, z, w = synth.box_catalog(50_000, L=1000.0, mesh=64, seed=7)
    cat = [DeviceArray.from_numpy(a) for a in (x, y, z, w)]
    rec = bh.IterativeRecon(**KW, ctx=ctx)
    outs = [DeviceArray((50_000,)) for _ in range(3)]
    good = [a.numpy() for a in bh.run_positions_graph(rec, grid, *cat, out=outs)]

    def upload(dst, a):
        assert hip().hipMemcpy(ctypes.c_void_p(dst.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    bad = x.copy()
    bad[123] = np.float32(5000.0)
    upload(cat[0], bad)
    with pytest.raises(bh.BaorecError):
        bh.run_positions_graph(rec, grid, *cat, out=outs)   # replay of the captured graph
    upload(cat[0], x)
    got = bh.run_positions_graph(rec, grid, *cat, out=outs)
    for a, b in zip(got, good):
        np.testing.assert_array_equal(a.numpy(), b)
