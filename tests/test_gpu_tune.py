"""Autotuning (baorec_autotune): the tuned scatter tiles and strided-pass
kernels must leave every result bitwise unchanged, and tuned tables must
survive a save / load round trip."""
import json

import numpy as np
import pytest

import baorec_hip as bh
from baorec_hip import synth

pytestmark = pytest.mark.gpu


def box_kw(ctx, los=(0.0, 0.0, 1.0)):
    return dict(bias=2.0, f=0.8, smoothing_radius=15.0, box_size=[1000.0] * 3, box_min=[0.0] * 3,
                los=None if los is None else list(los), ctx=ctx)


@pytest.mark.parametrize("g,los", [((128, 128, 128), (0.0, 0.0, 1.0)), ((64, 256, 256), None)])
def test_tuned_results_are_bitwise_unchanged(g, los):
    x, y, z, w = synth.box_catalog(300000, L=1000.0, mesh=64, seed=41)
    plain, tuned = bh.Context(0), bh.Context(0)
    rep = bh.autotune(bh.IterativeRecon(**box_kw(tuned, los)), g, x, y, z, w)
    print(json.dumps(rep))
    assert len(rep["scatter"]) == 1
    e = rep["scatter"][0]
    assert e["grid"] == list(g) and e["log2n"] == int(np.floor(np.log2(len(x))))
    tx, ty, tz = e["tile"]
    assert tx * ty * tz <= 32 * 16 * 16
    assert {(s["axis"], s["n"]) for s in rep["strided"]} == {(a, n) for a, n in ((1, g[1]), (2, g[2])) if n >= 256}
    out = []
    for ctx in (plain, tuned):
        rec = bh.IterativeRecon(**box_kw(ctx, los))
        cx, cy, cz = x.copy(), y.copy(), z.copy()
        d = bh.run_(rec, g, cx, cy, cz, w)
        p = bh.reconstructed_positions(rec, cx, cy, cz, field="sum")
        out.append((d, p))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)


def test_tune_save_load_roundtrip(tmp_path):
    x, y, z, w = synth.box_catalog(100000, L=1000.0, mesh=64, seed=42)
    a, b = bh.Context(0), bh.Context(0)
    rep = bh.autotune(bh.IterativeRecon(**box_kw(a)), (64, 64, 256), x, y, z, w)
    path = tmp_path / "tune.json"
    bh.tune_save(path, a)
    assert json.loads(path.read_text()) == rep
    bh.tune_load(path, b)
    assert bh.tune_report(b) == rep


def test_tune_load_rejects_unusable_entries(tmp_path):
    ctx = bh.Context(0)
    bad = tmp_path / "bad.json"
    bad.write_text('{"scatter": [\n  {"grid": [64, 64, 64], "tsc": 0, "log2n": 10, "tile": [64, 64, 64]}],\n'
                   ' "strided": []}\n')
    with pytest.raises(bh.BaorecError):
        bh.tune_load(bad, ctx)
    assert bh.tune_report(ctx) == {"scatter": [], "strided": []}


def test_failed_tuning_leaves_no_forced_choice():
    """A catalog outside the box: BoundsError, nothing tuned, and later calls
    use the library's own choices."""
    x, y, z, w = synth.box_catalog(50000, L=1000.0, mesh=64, seed=43)
    ctx = bh.Context(0)
    bad = x.copy()
    bad[:100] = -50.0      # below box_min: outside the grid
    with pytest.raises(bh.BoundsError):
        bh.autotune(bh.IterativeRecon(**box_kw(ctx)), (64, 64, 64), bad, y, z, w)
    assert bh.tune_report(ctx) == {"scatter": [], "strided": []}
    plain = bh.Context(0)
    out = []
    for c in (plain, ctx):
        rec = bh.IterativeRecon(**box_kw(c))
        out.append(bh.run_(rec, (64, 64, 64), x.copy(), y.copy(), z.copy(), w))
    np.testing.assert_array_equal(out[0], out[1])
