"""CPU oracle of the learned warm-start network (test infrastructure only).

The reference has no learned component; this module defines the network the
device kernels implement, in plain NumPy (float64 arithmetic on the same
bf16-rounded weights), so the HIP forward pass, gradients and optimizer step
can be checked against it:

* 6 layers of 3x3 "same" convolutions (zero padding), 2 -> 32 -> 32 -> 32 ->
  32 -> 32 -> 2 channels, bias, ReLU after every layer but the last;
* residual skip: output = input + conv6(...);
* loss = mean over all elements of (output - target)^2;
* complex images enter as s * net(x / s) with s = sqrt(mean |x|^2) per image
  (the network is then equivariant to the scale of the data);
* Adam (Kingma & Ba 2015) on float32 master weights; the forward pass uses
  the weights rounded to bfloat16 (round to nearest even).

Activations are NCHW: [batch, channel, H, W].

Pinned by ``tests/test_oracle_warmstart.py``: gradients against central
finite differences (which fail for a dropped term, a flipped kernel, a
transposed channel index or a wrong ReLU mask), the convolution against a
direct quadruple loop, and bf16 rounding against exact bit patterns; the
rounding points of the device semantics (bf16_acts) are checked to be small
perturbations of the exact gradient.
"""
import numpy as np

CHANNELS = [2, 32, 32, 32, 32, 32, 2]


def bf16_round(w):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(w, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def param_shapes(channels=CHANNELS):
    out = []
    for l in range(len(channels) - 1):
        out.append((channels[l + 1], channels[l], 3, 3))
        out.append((channels[l + 1],))
    return out


def num_params(channels=CHANNELS):
    return int(sum(np.prod(s) for s in param_shapes(channels)))


def unflatten(theta, channels=CHANNELS):
    """Flat parameter vector -> [W1, b1, W2, b2, ...] (weights [Cout, Cin, 3, 3])."""
    out, o = [], 0
    for s in param_shapes(channels):
        n = int(np.prod(s))
        out.append(np.asarray(theta[o:o + n]).reshape(s))
        o += n
    return out


def init_params(seed, channels=CHANNELS):
    """He-normal weights, zero biases; the last layer is scaled by 0.1 so the
    untrained network is close to the identity (residual skip)."""
    rng = np.random.default_rng(seed)
    parts = []
    L = len(channels) - 1
    for l in range(L):
        cin, cout = channels[l], channels[l + 1]
        w = rng.standard_normal((cout, cin, 3, 3)) * np.sqrt(2.0 / (cin * 9))
        if l == L - 1:
            w *= 0.1
        parts += [w.reshape(-1), np.zeros(cout)]
    return np.concatenate(parts).astype(np.float32)


def conv3x3(x, w, b):
    """'same' 3x3 convolution (cross-correlation), x [B,Cin,H,W], w [Cout,Cin,3,3]."""
    B, Cin, H, W = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    y = np.zeros((B, w.shape[0], H, W))
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, :, ky:ky + H, kx:kx + W]                # [B, Cin, H, W]
            y += np.einsum("bchw,oc->bohw", patch, w[:, :, ky, kx])
    return y + b[None, :, None, None]


def conv3x3_grad(x, w, gy):
    """Gradients of conv3x3 w.r.t. x, w and b given gy = dL/dy."""
    B, Cin, H, W = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w, dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            patch = xp[:, :, ky:ky + H, kx:kx + W]
            gw[:, :, ky, kx] = np.einsum("bohw,bchw->oc", gy, patch)
            gxp[:, :, ky:ky + H, kx:kx + W] += np.einsum("bohw,oc->bchw", gy, w[:, :, ky, kx])
    gb = gy.sum(axis=(0, 2, 3))
    return gxp[:, :, 1:-1, 1:-1], gw, gb


def forward(theta, x, channels=CHANNELS, keep=False, quantize=True, bf16_acts=True):
    """Network output for input x [B, 2, H, W]; with keep, also the
    activations needed by the backward pass.

    quantize:  weights and biases rounded to bf16 (else used as given);
    bf16_acts: the network input and every hidden activation rounded to bf16
               (the operands the device feeds to its bf16 matrix cores).
    The residual skip adds the unrounded input.  Both flags on is the
    function the device computes (up to FP32 accumulation order).
    """
    p = unflatten(bf16_round(theta) if quantize else np.asarray(theta, np.float64), channels)
    rnd = (lambda a: bf16_round(a).astype(np.float64)) if bf16_acts else (lambda a: a)
    L = len(channels) - 1
    x = np.asarray(x, dtype=np.float64)
    acts = [rnd(x)]
    h = acts[0]
    for l in range(L):
        z = conv3x3(h, p[2 * l].astype(np.float64), p[2 * l + 1].astype(np.float64))
        h = rnd(np.maximum(z, 0.0)) if l < L - 1 else z
        acts.append(h)
    out = x + h
    return (out, acts) if keep else out


def loss_and_grad(theta, x, target, channels=CHANNELS, quantize=True, bf16_acts=True):
    """Mean squared error and its gradient w.r.t. the (bf16-rounded) weights.

    With bf16_acts the output gradient and every back-propagated gradient
    are rounded to bf16 before they meet a weight or an activation (the
    device's rounding points); the ReLU mask is taken from the stored
    (rounded) activation."""
    p = unflatten(bf16_round(theta) if quantize else np.asarray(theta, np.float64), channels)
    rnd = (lambda a: bf16_round(a).astype(np.float64)) if bf16_acts else (lambda a: a)
    out, acts = forward(theta, x, channels, keep=True, quantize=quantize, bf16_acts=bf16_acts)
    n = out.size
    diff = out - target
    loss = float(np.sum(diff ** 2) / n)
    g = rnd(2.0 * diff / n)                 # dL/dout = dL/d(last conv) (skip adds identity)
    L = len(channels) - 1
    grads = [None] * (2 * L)
    for l in range(L - 1, -1, -1):
        gx, gw, gb = conv3x3_grad(acts[l], p[2 * l].astype(np.float64), g)
        grads[2 * l], grads[2 * l + 1] = gw, gb
        if l > 0:
            g = rnd(gx * (acts[l] > 0))     # ReLU of the layer that produced acts[l]
    return loss, np.concatenate([q.reshape(-1) for q in grads])


def image_scale(ximg):
    """Per-image scale s = sqrt(mean |x|^2) in float32 (1 for a zero image);
    ximg is complex [B, H, W]."""
    s = np.sqrt(np.mean(np.abs(ximg.astype(np.complex128)) ** 2, axis=(1, 2))).astype(np.float32)
    s[~(s > 0)] = 1.0
    return s


def to_channels(ximg):
    """Complex [B, H, W] -> real [B, 2, H, W] (re, im)."""
    return np.stack([ximg.real, ximg.imag], axis=1).astype(np.float64)


def apply_net(theta, ximg, channels=CHANNELS, quantize=True, bf16_acts=True):
    """Warm-start image for complex images ximg [B, H, W]: s * net(x / s)."""
    s = image_scale(ximg).astype(np.float64)[:, None, None, None]
    xs = to_channels(ximg) / s
    o = forward(theta, xs, channels, quantize=quantize, bf16_acts=bf16_acts) * s
    return o[:, 0] + 1j * o[:, 1]


def net_loss_and_grad(theta, ximg, timg, channels=CHANNELS, quantize=True, bf16_acts=True):
    """Training loss on complex image pairs: mean((x/s + z - t/s)^2), s per image."""
    s = image_scale(ximg).astype(np.float64)[:, None, None, None]
    return loss_and_grad(theta, to_channels(ximg) / s, to_channels(timg) / s, channels,
                         quantize=quantize, bf16_acts=bf16_acts)


def adam_step(theta, grad, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (t is the 1-based step count); returns new (theta, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    theta = theta - lr * mh / (np.sqrt(vh) + eps)
    return theta, m, v
