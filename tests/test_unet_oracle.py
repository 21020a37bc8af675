"""Pins of the warm-start network's CPU oracle (oracle/unet_oracle.py): its
convolution against a direct sum, the adjoint pairs of pooling and
upsampling, and its gradient against finite differences of its own loss,
with a planted mistake that the gradient check must catch."""
import numpy as np
import pytest

import unet_oracle as U


def test_conv_matches_direct_sum():
    """out[z, y, x] = b + sum over the 27 taps of in[z+dz, y+dy, x+dx] W[t]
    with periodic wrap, t = 9 (dz+1) + 3 (dy+1) + (dx+1)."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 4, 3, 5, 2))
    w = rng.standard_normal((27, 2, 3))
    b = rng.standard_normal(3)
    got = U.conv(x, w, b)
    exp = np.zeros((1, 4, 3, 5, 3))
    for z in range(4):
        for y in range(3):
            for xx in range(5):
                s = b.copy()
                for t in range(27):
                    dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                    s += x[0, (z + dz) % 4, (y + dy) % 3, (xx + dx) % 5] @ w[t]
                exp[0, z, y, xx] = s
    np.testing.assert_allclose(got, exp, atol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((2, 4, 4, 4, 3))
    c = rng.standard_normal((2, 2, 2, 2, 3))
    assert abs(np.sum(U.pool(a) * c) - np.sum(a * U.pool_bwd(c))) < 1e-12
    assert abs(np.sum(U.up(c) * a) - np.sum(c * U.down_sum(a))) < 1e-12


def test_conv_backward_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 4, 4, 3))
    w = rng.standard_normal((27, 3, 2))
    d = rng.standard_normal((1, 4, 4, 4, 2))
    dx, dw, db = U.conv_bwd(x, w, d)
    # <conv(x), d> is linear in x and in w
    assert abs(np.sum((U.conv(x, w, 0.0)) * d) - np.sum(x * dx)) < 1e-10
    assert abs(np.sum((U.conv(x, w, 0.0)) * d) - np.sum(w * dw)) < 1e-10
    np.testing.assert_allclose(db, d.reshape(-1, 2).sum(axis=0))


def fd_case(seed=0, n=16):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed + 1)
    X = rng.standard_normal((1, n, n, n, 8))
    X[..., 7] = 0.0
    T = rng.standard_normal((1, n, n, n))
    return p, X, T


def fd_errors(p, X, T, limit=None):
    """Relative errors of the oracle's gradient against central differences
    of its loss: single parameters in every layer (weights, biases, head)."""
    _, _, g = U.forward_backward(p, X, T, quantize=False)
    offs, hw, hb = U.offsets()
    rng = np.random.default_rng(7)
    idx = [hw + 3]
    for wo, bo in offs:
        idx += list(rng.integers(wo, bo, 2)) + [bo]
    idx = idx[:limit]
    errs = []
    for i in idx:
        best = np.inf
        for h in (1e-6, 1e-7):   # two steps: one of them may straddle a ReLU kink
            e = np.zeros(p.size)
            e[i] = h
            lp = U.forward_backward(p + e, X, T, quantize=False)[1]
            lm = U.forward_backward(p - e, X, T, quantize=False)[1]
            fd = (lp - lm) / (2 * h)
            best = min(best, abs(fd - g[i]) / (max(abs(fd), abs(g[i])) + 1e-4))
        errs.append(best)
    return errs


def test_oracle_gradient_matches_finite_differences():
    errs = fd_errors(*fd_case(n=8))
    print("max rel. error", max(errs))
    assert max(errs) < 1e-4


def test_gradient_check_catches_planted_mistake(monkeypatch):
    """The loss gradient with its sign flipped (the mistake a hand-derived
    adjoint of the gradient loss invites) fails the finite-difference check."""
    real = U.loss_and_grad_out

    def flipped(out, target):
        L, g = real(out, target)
        return L, -g

    monkeypatch.setattr(U, "loss_and_grad_out", flipped)
    assert max(fd_errors(*fd_case(n=8), limit=4)) > 0.5


def test_quantized_network_close_to_exact():
    """The BF16 rounding the GPU applies changes the output and gradients by
    a BF16-sized amount only."""
    p, X, T = fd_case(3)
    X = U.bf16(X)
    o1, l1, g1 = U.forward_backward(p, X, T, quantize=False)
    o2, l2, g2 = U.forward_backward(p, X, T, quantize=True)
    assert np.linalg.norm(o2 - o1) / np.linalg.norm(o1) < 2e-2
    assert np.linalg.norm(g2 - g1) / np.linalg.norm(g1) < 5e-2
