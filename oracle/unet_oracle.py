"""CPU oracle of the learned warm start's network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: tests/ compare the HIP U-Net (baorec.jl_amd/csrc/
unet.hip, reached through baorec_warmstart_eval) against this restatement;
the product path never imports it.  The reference has no such network
(BAOrec.jl's multigrid starts every level from zero, src/multigrid.jl:722-752),
so the oracle restates the network the library defines:

* tensors are channels last, shape (batch, nz, ny, nx, channels), periodic;
* 14 convolutions with 3^3 kernels, W[t][ci][co] with tap t = 9 (dz+1) +
  3 (dy+1) + (dx+1) reading in[z+dz, y+dy, x+dx], plus a bias and a ReLU;
  levels of 16, 32, 64, 128 channels, 2x2x2 average pooling on the way down,
  nearest upsampling of the coarser features concatenated after the skip
  input on the way up, and a 1x1 head (16 -> 1, FP32 output);
* the parameter vector holds, per layer, W (27 cin co) then b (co), and at
  the end the head's 16 weights and its bias;
* loss = (1 / 3N) sum_v sum_a (d(v + e_a) - d(v))^2, d = out - target.

``quantize=True`` rounds to BF16 (nearest even) wherever the GPU stores BF16:
the input features, the convolution weights, every activation and every
activation gradient; accumulation stays FP64 (the GPU accumulates in FP32).
``quantize=False`` is the exact FP64 network whose gradient the finite
difference checks pin.
"""
from __future__ import annotations

import numpy as np

LAYERS = [  # (c1, c2, co, level, in1, in2) as in unet.hip kLayers
    (8, 0, 16, 0, -1, None), (16, 0, 16, 0, 0, None),
    (16, 0, 32, 1, -2, None), (32, 0, 32, 1, 2, None),
    (32, 0, 64, 2, -3, None), (64, 0, 64, 2, 4, None),
    (64, 0, 128, 3, -4, None), (128, 0, 128, 3, 6, None),
    (64, 128, 64, 2, 5, 7), (64, 0, 64, 2, 8, None),
    (32, 64, 32, 1, 3, 9), (32, 0, 32, 1, 10, None),
    (16, 32, 16, 0, 1, 11), (16, 0, 16, 0, 12, None),
]
POOLED_FROM = {0: 1, 1: 3, 2: 5}          # pooled tensor l is the pool of layer output


def offsets():
    out = []
    o = 0
    for c1, c2, co, *_ in LAYERS:
        out.append((o, o + 27 * (c1 + c2) * co))
        o += 27 * (c1 + c2) * co + co
    return out, o, o + 16


def nparams():
    _, hw, hb = offsets()
    return hb + 1


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """BF16 bit patterns (uint16) of a float array, nearest even."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def conv(x, w, b):
    """out[v] = b + sum_t x[v + off_t] @ W[t] (periodic); x (B, Z, Y, X, cin)."""
    out = np.zeros(x.shape[:4] + (w.shape[2],))
    for t, (dz, dy, dx) in enumerate(taps()):
        out += np.roll(x, shift=(-dz, -dy, -dx), axis=(1, 2, 3)) @ w[t]
    return out + b


def conv_bwd(x, w, dout):
    """(dx, dw, db) of conv for the output gradient dout."""
    dx = np.zeros(x.shape)
    dw = np.zeros(w.shape)
    cin = x.shape[-1]
    d2 = dout.reshape(-1, dout.shape[-1])
    for t, (dz, dy, dx_) in enumerate(taps()):
        dx += np.roll(dout, shift=(dz, dy, dx_), axis=(1, 2, 3)) @ w[t].T
        dw[t] = np.roll(x, shift=(-dz, -dy, -dx_), axis=(1, 2, 3)).reshape(-1, cin).T @ d2
    return dx, dw, d2.sum(axis=0)


def pool(x):
    B, Z, Y, X, C = x.shape
    return x.reshape(B, Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(dc):
    return up(dc) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, axis=1), 2, axis=2), 2, axis=3)


def down_sum(x):
    B, Z, Y, X, C = x.shape
    return x.reshape(B, Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(2, 4, 6))


def unpack(params):
    offs, hw, hb = offsets()
    W, b = [], []
    for (c1, c2, co, *_), (wo, bo) in zip(LAYERS, offs):
        W.append(params[wo:bo].reshape(27, c1 + c2, co))
        b.append(params[bo:bo + co])
    return W, b, params[hw:hw + 16], params[hb]


def loss_and_grad_out(out, target):
    d = out - target
    n = d.size
    L = 0.0
    g = np.zeros_like(d)
    for ax in (1, 2, 3):
        diff = np.roll(d, -1, axis=ax) - d
        L += np.sum(diff ** 2)
        g += 2.0 * (np.roll(diff, 1, axis=ax) - diff)    # d/dd(v) of sum diff^2
    scale = 1.0 / (3.0 * n)
    return L * scale, g * scale


def forward_backward(params, X, target, quantize=True):
    """Network output, loss and the gradient of the loss w.r.t. the parameter
    vector.  X: (B, Z, Y, X, 8) features, target: (B, Z, Y, X)."""
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    W, b, hw, hb = unpack(np.asarray(params, np.float64))
    Wq = [q(w) for w in W]
    X = q(X)
    A = [None] * len(LAYERS)
    P = [None] * 3

    def inp(i):
        c1, c2, co, lvl, i1, i2 = LAYERS[i]
        a1 = X if i1 == -1 else (P[-i1 - 2] if i1 <= -2 else A[i1])
        return a1 if i2 is None else np.concatenate([a1, up(A[i2])], axis=-1)

    for i in range(len(LAYERS)):
        A[i] = q(np.maximum(conv(inp(i), Wq[i], b[i]), 0.0))
        if i in (1, 3, 5):
            P[(i - 1) // 2] = q(pool(A[i]))
    out = (A[13] @ hw + hb).astype(np.float32).astype(np.float64) if quantize else A[13] @ hw + hb
    L, dout = loss_and_grad_out(out, target)

    grads = np.zeros(nparams())
    offs, ho, hbo = offsets()
    grads[ho:ho + 16] = (A[13] * dout[..., None]).reshape(-1, 16).sum(axis=0)
    grads[hbo] = dout.sum()
    dA = [None] * len(LAYERS)
    dP = [None] * 3
    dA[13] = q(dout[..., None] * hw)
    for i in range(len(LAYERS) - 1, -1, -1):
        c1, c2, co, lvl, i1, i2 = LAYERS[i]
        g = dA[i] * (A[i] > 0)
        dx, dw, db = conv_bwd(inp(i), Wq[i], g)
        wo, bo = offs[i]
        grads[wo:bo] = dw.ravel()
        grads[bo:bo + co] = db
        if i == 0:
            break
        dx = q(dx)
        if i2 is not None:
            dA[i1] = dx[..., :c1]
            dA[i2] = q(down_sum(dx[..., c1:]))
        elif i1 <= -2:
            pl = -i1 - 2
            dP[pl] = dx
            sk = POOLED_FROM[pl]
            dA[sk] = q(dA[sk] + pool_bwd(dP[pl]))
        else:
            dA[i1] = dx
    return out, L, grads


def init_params(seed=0):
    """He-normal convolution weights, zero biases, a small head (the library's
    own initialisation uses another generator; tests set parameters
    explicitly)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(nparams())
    offs, hw, hb = offsets()
    for (c1, c2, co, *_), (wo, bo) in zip(LAYERS, offs):
        p[wo:bo] = rng.standard_normal(bo - wo) * np.sqrt(2.0 / (27 * (c1 + c2)))
        p[bo:bo + co] = 0.05 * rng.standard_normal(co)
    p[hw:hw + 16] = 0.3 * rng.standard_normal(16)
    p[hb] = 0.01
    return p
