"""The learned warm start on the GPU: the BF16 matrix-core U-Net against its
FP64 CPU oracle (output, loss and every parameter gradient), a training
step that lowers the loss, and the learned seed inside the RBGS solve."""
import os

import numpy as np
import pytest

import baorec_hip as bh
import unet_oracle as U
from baorec_hip import synth

pytestmark = pytest.mark.gpu

F32 = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def features(shape, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal(shape + (8,))
    X[..., 6] = (X[..., 6] > 0).astype(float)
    X[..., 7] = 0.0
    return U.bf16_bits(X)


def assert_bf16_close(gpu_bits, exact, what, max_frac=0.02, inner=0.0):
    """A BF16 tensor the GPU computed against the exact value of the same
    operation on the same inputs: every element within one BF16 step
    (2^-7 relative) plus FP32 accumulation noise, and almost all of them equal
    to the exact value rounded to BF16 (the GPU accumulates in FP32, so a
    sum within FP32 rounding of a BF16 tie can round the other way).  inner:
    magnitude of an intermediate BF16 value the result was formed from (one
    more rounding step of that size is allowed)."""
    got = U.from_bits(gpu_bits)
    exact = np.asarray(exact, np.float64).reshape(got.shape)
    scale = float(np.sqrt(np.mean(exact ** 2))) + 1e-30
    err = np.abs(got - exact)
    bad = err > 2.0 ** -7 * (np.abs(exact) + np.abs(inner)) + 1e-5 * scale
    frac = float(np.mean(got != U.bf16(exact)))
    print(f"{what}: max err / rms {err.max() / scale:.1e}, {frac:.1e} of the entries off the rounded value")
    assert not bad.any(), (what, int(bad.sum()))
    assert frac < max_frac, (what, frac)


@pytest.mark.parametrize("shape", [(1, 16, 16, 16), (2, 16, 32, 16)])
def test_unet_layers_match_oracle_on_their_inputs(ctx, shape):
    """Each kernel of the network against the oracle's operation applied to
    the GPU's own inputs of that layer (shapes (batch, nz, ny, nx)): every
    convolution and pooling of the forward pass, the head, the loss
    gradient, each layer's weight and bias gradient (FP32, 1e-5) and each
    input gradient through the ReLU masks, the concatenation split and the
    pooling adjoint."""
    B, Z, Y, Xn = shape
    p = U.init_params(5).astype(F32)
    Xb = features(shape, 11)
    T = np.random.default_rng(12).standard_normal(shape).astype(F32)
    bh.warmstart_init(seed=1, h_train=1.0, ctx=ctx)
    bh.warmstart_set_params(p, ctx=ctx)
    out, loss = bh.warmstart_eval(Xb, T, ctx=ctx)
    grads = bh.warmstart_params("grads", ctx=ctx).astype(np.float64)
    W, b, hw, hb = U.unpack(p.astype(np.float64))
    Wq = [U.bf16(w) for w in W]

    def dims(level):
        return (B, Z >> level, Y >> level, Xn >> level)

    def tensor(kind, i, level, c):
        return bh.warmstart_tensor(kind, i, dims(level) + (c,), ctx=ctx)

    np.testing.assert_array_equal(bh.warmstart_tensor("input", 0, shape + (8,), ctx=ctx), Xb)
    A = [U.from_bits(tensor("act", i, L[3], L[2])) for i, L in enumerate(U.LAYERS)]
    P = [U.from_bits(tensor("pool", l, l + 1, 16 << l)) for l in range(3)]
    dA = [U.from_bits(tensor("grad", i, L[3], L[2])) for i, L in enumerate(U.LAYERS)]
    dP = [U.from_bits(tensor("pool_grad", l, l + 1, 16 << l)) for l in range(3)]
    Xf = U.from_bits(Xb)

    def inp(i):
        c1, c2, co, lvl, i1, i2 = U.LAYERS[i]
        a1 = Xf if i1 == -1 else (P[-i1 - 2] if i1 <= -2 else A[i1])
        return a1 if i2 is None else np.concatenate([a1, U.up(A[i2])], axis=-1)

    # forward
    for i in range(len(U.LAYERS)):
        assert_bf16_close(U.bf16_bits(A[i]), np.maximum(U.conv(inp(i), Wq[i], b[i]), 0.0), f"conv {i}")
    for l, sk in U.POOLED_FROM.items():
        assert_bf16_close(U.bf16_bits(P[l]), U.pool(A[sk]), f"pool {l}")
    head = A[13] @ hw + hb
    assert rel_l2(out, head) < 1e-6
    L_ex, dout = U.loss_and_grad_out(out.astype(np.float64), T.astype(np.float64))
    assert abs(loss - L_ex) < 1e-5 * L_ex
    assert_bf16_close(U.bf16_bits(dA[13]), (dout[..., None] * hw) * (A[13] > 0), "head backward")
    offs, ho, hbo = U.offsets()
    assert rel_l2(grads[ho:ho + 16], (A[13] * dout[..., None]).reshape(-1, 16).sum(axis=0)) < 1e-5
    # backward, layer by layer on the GPU's masked output gradients
    for i in range(len(U.LAYERS) - 1, -1, -1):
        c1, c2, co, lvl, i1, i2 = U.LAYERS[i]
        dx, dw, db = U.conv_bwd(inp(i), Wq[i], dA[i])
        wo, bo = offs[i]
        ew, eb = rel_l2(grads[wo:bo], dw.ravel()), rel_l2(grads[bo:bo + co], db)
        print(f"layer {i}: weight gradient {ew:.1e}, bias gradient {eb:.1e}")
        assert ew < 1e-5 and eb < 1e-5
        if i == 0:
            continue
        dxq = U.bf16(dx)
        if i2 is not None:   # decoder: skip part (plus the pooling adjoint) and the coarse part
            pl = {1: 0, 3: 1, 5: 2}[i1]
            exp_skip = (U.bf16(dxq[..., :c1]) + U.pool_bwd(dP[pl])) * (A[i1] > 0)
            assert_bf16_close(U.bf16_bits(dA[i1]), exp_skip, f"input gradient {i} (skip)",
                              inner=dxq[..., :c1] * (A[i1] > 0))
            assert_bf16_close(U.bf16_bits(dA[i2]), U.down_sum(dxq[..., c1:]) * (A[i2] > 0),
                              f"input gradient {i} (coarse)", inner=U.down_sum(np.abs(dxq[..., c1:])))
        elif i1 <= -2:
            assert_bf16_close(U.bf16_bits(dP[-i1 - 2]), dx, f"input gradient {i} (pooled)")
        else:
            assert_bf16_close(U.bf16_bits(dA[i1]), dx * (A[i1] > 0), f"input gradient {i}")


def test_unet_gradients_within_bf16_rounding(ctx):
    """The whole network's output and gradients against the exact FP64
    network: per layer no further from it than 1.5x (plus 0.02) what the
    same BF16 roundings alone cost (the oracle with quantize=True)."""
    shape = (1, 16, 16, 16)
    p = U.init_params(5).astype(F32)
    Xb = features(shape, 11)
    T = np.random.default_rng(12).standard_normal(shape).astype(F32)
    bh.warmstart_init(seed=1, h_train=1.0, ctx=ctx)
    bh.warmstart_set_params(p, ctx=ctx)
    out, loss = bh.warmstart_eval(Xb, T, ctx=ctx)
    g = bh.warmstart_params("grads", ctx=ctx).astype(np.float64)
    Xf = U.from_bits(Xb)
    oq, lq, gq = U.forward_backward(p.astype(np.float64), Xf, T.astype(np.float64), quantize=True)
    oe, le, ge = U.forward_backward(p.astype(np.float64), Xf, T.astype(np.float64), quantize=False)
    print(f"output vs FP64: GPU {rel_l2(out, oe):.2e}, BF16 oracle {rel_l2(oq, oe):.2e}")
    assert rel_l2(out, oe) < 1.5 * rel_l2(oq, oe) + 0.01
    offs, hw, hb = U.offsets()
    for i, (wo, bo) in enumerate(offs):
        e_gpu, e_q = rel_l2(g[wo:bo], ge[wo:bo]), rel_l2(gq[wo:bo], ge[wo:bo])
        print(f"layer {i:2d}: weight gradient vs FP64: GPU {e_gpu:.2e}, BF16 oracle {e_q:.2e}")
        assert e_gpu < 1.5 * e_q + 0.02


def smooth(n, seed, R=2.0):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n) * 2 * np.pi
    k2 = k[None, None, :] ** 2 + k[None, :, None] ** 2 + k[:, None, None] ** 2
    return np.fft.ifftn(np.fft.fftn(f) * np.exp(-0.5 * k2 * R * R)).real.astype(F32)


def lc_recon(n=32, **kw):
    """A local-LOS multigrid recon on a lightcone-like box (observer outside)."""
    return bh.MultigridRecon(bias=2.2, f=0.75, smoothing_radius=15.0, box_size=[800.0] * 3,
                             box_min=[600.0, 300.0, 200.0], los=None, schedule="rbgs", tolerance=1e-6, **kw)


def test_training_steps_lower_the_loss(ctx):
    """Adam steps on one (f, converged phi) pair at 32^3 lower the loss."""
    n = 32
    f = smooth(n, 3)
    rec = lc_recon(n)
    phi = bh.fmg(f, rec)
    assert rec.final_residual <= 1e-6
    bh.warmstart_init(seed=2, h_train=800.0 / n, ctx=ctx)
    losses = [bh.warmstart_step(rec, f, phi, lr=1e-3) for _ in range(40)]
    print("loss", losses[0], "->", losses[-1])
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.5 * losses[0]
    # gradients only: the parameters do not move
    p0 = bh.warmstart_params(ctx=ctx)
    l1 = bh.warmstart_step(rec, f, phi, lr=0.0)
    l2 = bh.warmstart_step(rec, f, phi, lr=0.0)
    assert l1 == l2
    np.testing.assert_array_equal(bh.warmstart_params(ctx=ctx), p0)


def test_learned_seed_solve_reaches_tolerance(ctx, tmp_path):
    """warm_start="learned" with a network (here untrained) starts from
    spectral seed + correction and still runs to the plain tolerance, to the
    same potential as the cold start; the network file round-trips."""
    n = 32
    f = smooth(n, 4)
    bh.warmstart_init(seed=3, h_train=800.0 / n, ctx=ctx)
    path = tmp_path / "net.bin"
    bh.warmstart_save(str(path), ctx=ctx)
    p0 = bh.warmstart_params(ctx=ctx)
    bh.warmstart_init(seed=4, h_train=1.0, ctx=ctx)
    bh.warmstart_load(str(path), ctx=ctx)
    np.testing.assert_array_equal(bh.warmstart_params(ctx=ctx), p0)
    cold = lc_recon(n, ctx=ctx)
    pc = bh.fmg(f, cold)
    warm = lc_recon(n, ctx=ctx, warm_start="learned")
    pw = bh.fmg(f, warm)
    seed = bh.warmstart_predict(warm, f)
    print(f"cold {cold.vcycles_done} V-cycles; learned seed {warm.seed_residual:.2e}, {warm.vcycles_done} V-cycles")
    assert warm.final_residual <= 1e-6 and np.all(np.isfinite(seed))
    assert rel_l2(pw - pw.mean(), pc - pc.mean()) < 1e-4
    assert abs(warm.rhs_shift - cold.rhs_shift) < 1e-6 * float(np.std(f))


def test_learned_seed_without_network_is_an_error():
    c = bh.Context(0)
    rec = lc_recon(32, ctx=c, warm_start="learned")
    with pytest.raises(bh.BaorecError, match="network"):
        bh.fmg(smooth(32, 5), rec)
    c.close()


NET = os.path.join(ROOT, "baorec.jl_amd", "weights", "unet_lc128.bin")


@pytest.mark.skipif(not os.path.exists(NET), reason="no trained network in the tree")
def test_trained_network_saves_vcycles_on_held_out_lightcones(ctx):
    """The committed network (trained on 1950 lightcone realizations at
    128^3 by baorec_hip.warmstart) on three realizations it never saw: with
    the same number of spectral Richardson steps, the learned start needs
    fewer V-cycles than the refined spectral start without the network, and
    every start converges to the cold start's potential."""
    from baorec_hip import warmstart as W
    bh.warmstart_load(NET, ctx=ctx)
    rand = synth.lightcone_randoms(5_000_000, seed=10000, **W.SHELL)
    ext = max(float(np.max(a) - np.min(a)) for a in rand[:3])
    bs, bm = bh.setup_box(*rand[:3], 0.1 * ext, ctx=ctx)
    n = 128
    saved = []
    for seed, size, slope in W.realization_plan(3, seed0=90000):
        cat = synth.lightcone_catalog(int(size), mesh=128, seed=int(seed), slope=float(slope), **W.SHELL)
        kw = dict(**W.RECON, box_size=list(bs), box_min=list(bm), los=None, schedule="rbgs", tolerance=1e-6,
                  ctx=ctx)
        f = np.zeros((n, n, n), F32)
        bh.setup_overdensity_(f, bh.MultigridRecon(**kw), *cat, *rand)
        out = {}
        for ws, k in (("none", 1), ("refined", 1), ("learned", 1), ("refined", 3), ("learned", 3)):
            rec = bh.MultigridRecon(**kw, warm_start=ws, seed_steps=k)
            phi = bh.fmg(f, rec)
            assert rec.final_residual <= 1e-6
            out[(ws, k)] = (rec.vcycles_done, rec.seed_residual, phi)
        print(f"realization {seed}: " + ", ".join(f"{w}/{k}: seed {v[1]:.2e}, {v[0]} V-cycles"
                                                  for (w, k), v in out.items()))
        p0 = out[("none", 1)][2]
        for v in out.values():
            assert rel_l2(v[2] - v[2].mean(), p0 - p0.mean()) < 1e-4
        saved.append(out[("refined", 1)][0] - out[("learned", 1)][0])
        assert out[("learned", 3)][0] <= out[("refined", 3)][0] and out[("learned", 1)][0] <= out[("refined", 1)][0]
        assert 2 * out[("learned", 3)][0] <= out[("none", 1)][0]
    assert np.mean(saved) > 0
