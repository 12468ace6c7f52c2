"""CPU oracle of the warm-start network (test infrastructure only: the tests
check the gfx950 kernels of libmpd_nn.so against it; nothing in the product
imports it).

A float64 numpy restatement of the BF16 conv encoder-decoder and of its
backward pass: 3x3 convolutions (stride 1, zero padding 1, weights
[Cout][3][3][Cin], NHWC tensors), SiLU, 2x2 average pooling, nearest x2
upsampling, channel concatenation, mean squared error and Adam.  The network
has no counterpart in the reference (it replaces the initial condition of
src/main.cpp:117-148 by a prediction); these functions are pinned by
independent checks in tests/test_surrogate_oracle.py: scipy's correlate for
the convolution, finite differences for every backward function and for the
whole network, and closed forms for pooling and upsampling.

`bf16_round` reproduces the GPU's round-to-nearest-even to bfloat16, so the
oracle can be fed the same bf16 operands as the kernels.
"""
import numpy as np


def bf16_round(x):
    """Round float values to the nearest bfloat16 (ties to even), as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return out.astype(np.float64)


def bf16_bits(x):
    """bfloat16 bit patterns (uint16) of float values (round to nearest even)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def bits_to_f64(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _patches(x):
    """(B, H, W, 9, C): the 3x3 neighbourhood of every pixel, zero padded;
    tap = 3 ky + kx reads pixel (y + ky - 1, x + kx - 1)."""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C), dtype=np.float64)
    xp[:, 1:-1, 1:-1, :] = x
    return np.stack([xp[:, ky:ky + H, kx:kx + W, :] for ky in range(3) for kx in range(3)], axis=3)


def conv_fwd(x, w, b=None):
    """z[b, y, x, co] = sum_{ky, kx, ci} w[co, ky, kx, ci] x[b, y+ky-1, x+kx-1, ci] + b[co]."""
    B, H, W, C = x.shape
    Co = w.shape[0]
    z = _patches(x).reshape(B * H * W, 9 * C) @ np.asarray(w, np.float64).reshape(Co, 9 * C).T
    if b is not None:
        z = z + b
    return z.reshape(B, H, W, Co)


def conv_bwd_data(dz, w, shape):
    B, H, W, C = shape
    Co = w.shape[0]
    dxp = np.zeros((B, H + 2, W + 2, C))
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + H, kx:kx + W, :] += dz @ np.asarray(w[:, ky, kx, :], np.float64)
    return dxp[:, 1:-1, 1:-1, :]


def conv_bwd_weight(x, dz):
    B, H, W, C = x.shape
    Co = dz.shape[-1]
    P = _patches(x).reshape(B * H * W, 9 * C)
    return (dz.reshape(-1, Co).T @ P).reshape(Co, 3, 3, C)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def pool_fwd(x):
    B, H, W, C = x.shape
    return 0.25 * x.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def pool_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def up_fwd(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(dy):
    B, H, W, C = dy.shape
    return dy.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def mse(y, t):
    d = y - t
    return float(np.mean(d * d)), 2.0 * d / d.size


def adam(w, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return w - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def unet_layers(cin, cout, base=32, levels=4):
    """(Cin, Cout) of every conv in execution order: per encoder level two
    convs, per decoder level two (the first after the skip concatenation),
    then the output head."""
    ch = [base * 2 ** i for i in range(levels)]
    layers = []
    c = cin
    for i in range(levels):
        layers += [(c, ch[i]), (ch[i], ch[i])]
        c = ch[i]
    for i in range(levels - 2, -1, -1):
        layers += [(c + ch[i], ch[i]), (ch[i], ch[i])]
        c = ch[i]
    layers.append((c, cout))
    return layers


def unet_forward(x, params, levels=4, bf16=False):
    """Forward pass.  params: list of (w, b) per conv (unet_layers order).
    With bf16, activations and operands are rounded where the GPU rounds
    them (conv inputs, weights, activations after SiLU / pooling)."""
    rnd = bf16_round if bf16 else (lambda a: a)
    tape = []
    k = 0

    def conv(a, act):
        nonlocal k
        w, b = params[k]
        k += 1
        wq = rnd(w)
        z = conv_fwd(a, wq, b)
        out = rnd(silu(z)) if act else z
        tape.append((a, wq, z, act))
        return out

    a = rnd(x)
    skips = []
    for i in range(levels):
        a = conv(a, True)
        a = conv(a, True)
        if i < levels - 1:
            skips.append(a)
            a = rnd(pool_fwd(a))
    for i in range(levels - 2, -1, -1):
        a = np.concatenate([up_fwd(a), skips[i]], axis=-1)
        a = conv(a, True)
        a = conv(a, True)
    y = conv(a, False)
    return y, tape


def unet_backward(dy, tape, levels=4, bf16=False):
    """Gradients (dw, db) per conv, in unet_layers order, from dL/dy."""
    rnd = bf16_round if bf16 else (lambda a: a)
    grads = [None] * len(tape)
    k = len(tape) - 1

    def conv_back(da):
        nonlocal k
        a_in, wq, z, act = tape[k]
        dz = da * silu_grad(z) if act else da
        dzq = rnd(dz)
        grads[k] = (conv_bwd_weight(a_in, dzq), dz.reshape(-1, dz.shape[-1]).sum(axis=0))
        dx = conv_bwd_data(dzq, wq, a_in.shape)
        k -= 1
        return dx

    d = conv_back(dy)
    ch = [tape[2 * i][2].shape[-1] for i in range(levels)]
    dskips = [None] * (levels - 1)
    for i in range(levels - 1):
        d = conv_back(d)
        d = conv_back(d)
        cu = d.shape[-1] - ch[i]
        dskips[i] = d[..., cu:]
        d = up_bwd(d[..., :cu])
    d = conv_back(d)
    d = conv_back(d)
    for i in range(levels - 2, -1, -1):
        d = pool_bwd(d) + dskips[i]
        d = conv_back(d)
        d = conv_back(d)
    assert k == -1
    return grads
