"""Pins the warm-start network oracle (oracle/warmstart.py)."""
import numpy as np

from oracle import warmstart as ws


def test_bf16_round_known_bits():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.140625, 0.0], np.float32)
    r = ws.bf16_round(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1.0
    assert r[0] == 1.0 and r[1] == 1.0
    # 1.005859375 = 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert r[2] == np.float32(1.0078125)
    assert r[3] == np.float32(-3.140625) and r[4] == 0.0
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 5, 6))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = ws.conv3x3(x, w, b)
    ref = np.zeros_like(y)
    for bb in range(2):
        for o in range(4):
            for i in range(5):
                for j in range(6):
                    s = b[o]
                    for c in range(3):
                        for ky in range(3):
                            for kx in range(3):
                                ii, jj = i + ky - 1, j + kx - 1
                                if 0 <= ii < 5 and 0 <= jj < 6:
                                    s += w[o, c, ky, kx] * x[bb, c, ii, jj]
                    ref[bb, o, i, j] = s
    assert np.allclose(y, ref, atol=1e-12)


def test_gradient_finite_differences():
    """Central differences in float64 of the unquantized network.  With every
    other weight fixed the output is piecewise linear in one weight, so the
    loss is piecewise quadratic and the central difference is exact (to
    rounding) unless the step crosses a ReLU kink, which h = 1e-7 makes
    vanishingly unlikely."""
    rng = np.random.default_rng(1)
    ch = [2, 4, 4, 2]
    theta = ws.init_params(3, ch).astype(np.float64) * 3.0
    theta += 0.05 * rng.standard_normal(theta.size)          # nonzero biases
    x = rng.standard_normal((2, 2, 6, 5))
    t = rng.standard_normal((2, 2, 6, 5))
    loss, g = ws.loss_and_grad(theta, x, t, ch, quantize=False, bf16_acts=False)
    h = 1e-7
    for i in rng.choice(theta.size, 60, replace=False):
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += h
        tm[i] -= h
        lp = np.mean((ws.forward(tp, x, ch, quantize=False, bf16_acts=False) - t) ** 2)
        lm = np.mean((ws.forward(tm, x, ch, quantize=False, bf16_acts=False) - t) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-1, abs(fd)), (i, fd, g[i])


def test_quantized_gradient_is_gradient_at_rounded_weights():
    rng = np.random.default_rng(2)
    ch = [2, 4, 2]
    theta = ws.init_params(5, ch) + np.float32(1e-3) * rng.standard_normal(
        ws.num_params(ch)).astype(np.float32)
    x = rng.standard_normal((1, 2, 5, 7))
    t = rng.standard_normal((1, 2, 5, 7))
    lq, gq = ws.loss_and_grad(theta, x, t, ch, bf16_acts=False)
    lr, gr = ws.loss_and_grad(ws.bf16_round(theta).astype(np.float64), x, t, ch, quantize=False,
                              bf16_acts=False)
    assert lq == lr and np.array_equal(gq, gr)
    assert not np.array_equal(ws.bf16_round(theta), theta)


def test_adam_first_step_is_signed_lr():
    """At t = 1 Adam moves every weight by lr * g / (|g| + eps) = ~lr * sign(g)."""
    theta = np.zeros(5, np.float64)
    g = np.array([1.0, -2.0, 0.5, -1e-3, 3.0])
    th, m, v = ws.adam_step(theta, g, np.zeros(5), np.zeros(5), 1, 0.01)
    assert np.allclose(th, -0.01 * np.sign(g), rtol=1e-4)


def test_device_rounding_points_are_small_perturbations():
    """bf16 activations and gradients change the loss and the gradient by
    about the bf16 unit roundoff (2^-9), not more."""
    rng = np.random.default_rng(4)
    theta = ws.init_params(7)
    theta[ws.num_params() - 2:] = 0.1                     # nonzero last bias
    x = rng.standard_normal((2, 2, 12, 10))
    t = x + 0.3 * rng.standard_normal(x.shape)
    l0, g0 = ws.loss_and_grad(theta, x, t, bf16_acts=False)
    l1, g1 = ws.loss_and_grad(theta, x, t, bf16_acts=True)
    assert abs(l1 - l0) <= 2e-2 * l0
    assert np.linalg.norm(g1 - g0) <= 2e-2 * np.linalg.norm(g0)
    assert np.linalg.norm(g1 - g0) > 0


def test_apply_net_scale_equivariance():
    """s * net(x / s): scaling the input image by any c > 0 scales the output
    by c exactly when c is a power of two (bf16 rounding commutes with it)."""
    rng = np.random.default_rng(5)
    theta = ws.init_params(8)
    theta[-2:] = 0.2
    x = rng.standard_normal((1, 8, 8)) + 1j * rng.standard_normal((1, 8, 8))
    y1 = ws.apply_net(theta, x)
    y2 = ws.apply_net(theta, 8.0 * x)
    assert np.allclose(y2, 8.0 * y1, rtol=1e-12, atol=0)
    # and the skip: zero last layer -> identity
    th0 = theta.copy()
    th0[ws.num_params() - 2 * 32 * 9 - 2:] = 0
    assert np.allclose(ws.apply_net(th0, x), x, rtol=1e-6)
