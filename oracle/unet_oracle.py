"""CPU FP64 oracle of the learned preconditioner's network (U-Net).

TEST INFRASTRUCTURE ONLY.  The reference has no learned preconditioner; this
module states the network of the learned flexible PCG in plain NumPy double
precision, as the checker of the GPU convolutions, their backward passes and
the optimiser.  Only ``tests/`` import it.

Network (4 levels, 3x3 "same" convolutions with zero padding, ReLU after every
convolution except the last, 2x2 average pooling, nearest upsampling, the
upsampled tensor concatenated before the skip tensor):

    x(1) -> 16 -> 16 | pool -> 32 -> 32 | pool -> 64 -> 64 | pool -> 64 -> 64
    up, cat(64, 64) -> 64 -> 64 | up, cat(64, 32) -> 32 -> 32
    up, cat(32, 16) -> 16 -> 16 | -> 1 (linear)

Tensors are HWC.  Parameters form one flat vector, layer by layer: weights
[tap][Cout][Cin] (tap = 3 * (dh + 1) + (dw + 1)), then the Cout biases.

``bf16=True`` rounds what the GPU stores in bfloat16 (weights of the 16+
channel layers, activations, the masked gradients) to nearest-even, so the GPU
can be compared against it at accumulation-order accuracy.

Parity pins (tests/test_unet_oracle.py): the convolution against a brute-force
loop, pooling/upsampling against their adjoint identities, and the whole
backward pass against central finite differences.
"""

from __future__ import annotations

import numpy as np

# type, cin1, cin2, cout, level;  type 0 = 1->C input layer, 1 = matrix-core
# convolution, 2 = C->1 output layer
LAYERS = [
    (0, 1, 0, 16, 0), (1, 16, 0, 16, 0), (1, 16, 0, 32, 1), (1, 32, 0, 32, 1),
    (1, 32, 0, 64, 2), (1, 64, 0, 64, 2), (1, 64, 0, 64, 3), (1, 64, 0, 64, 3),
    (1, 64, 64, 64, 2), (1, 64, 0, 64, 2), (1, 64, 32, 32, 1), (1, 32, 0, 32, 1),
    (1, 32, 16, 16, 0), (1, 16, 0, 16, 0), (2, 16, 0, 1, 0),
]


def offsets():
    """[(w_off, b_off)] per layer and the total parameter count."""
    out, off = [], 0
    for _t, c1, c2, co, _l in LAYERS:
        w = off
        off += 9 * (c1 + c2) * co
        out.append((w, off))
        off += co
    return out, off


def param_count() -> int:
    return offsets()[1]


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (std sqrt(2 / fan_in)), zero biases; the linear output
    layer uses std sqrt(1 / fan_in)."""
    rng = np.random.default_rng(seed)
    offs, n = offsets()
    w = np.zeros(n, dtype=np.float64)
    for (t, c1, c2, co, _l), (wo, bo) in zip(LAYERS, offs):
        fan_in = 9 * (c1 + c2)
        std = np.sqrt((1.0 if t == 2 else 2.0) / fan_in)
        w[wo:bo] = rng.standard_normal(bo - wo) * std
    return w.astype(np.float32)


def bf16(a):
    """Round to bfloat16 (nearest-even) and return as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(xp, t, H, W):
    dh, dw = t // 3 - 1, t % 3 - 1
    return xp[1 + dh:1 + dh + H, 1 + dw:1 + dw + W, :]


def conv3x3(x, Wt, b):
    """y[h, w, co] = b[co] + sum_t sum_ci Wt[t, co, ci] x[h + dh_t, w + dw_t, ci]."""
    H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    y = np.zeros((H, W, Wt.shape[1])) + (0.0 if b is None else b)
    for t in range(9):
        y += _shift(xp, t, H, W) @ Wt[t].T
    return y


def conv3x3_bwd(x, Wt, g):
    """Gradients of conv3x3 with upstream gradient g: (dx, dWt, db)."""
    H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(Wt)
    g2 = g.reshape(H * W, -1)
    for t in range(9):
        dh, dw = t // 3 - 1, t % 3 - 1
        dW[t] = g2.T @ _shift(xp, t, H, W).reshape(H * W, -1)
        dxp[1 + dh:1 + dh + H, 1 + dw:1 + dw + W, :] += g @ Wt[t]
    return dxp[1:-1, 1:-1, :], dW, g2.sum(axis=0)


def pool2(x):
    return 0.25 * (x[0::2, 0::2] + x[0::2, 1::2] + x[1::2, 0::2] + x[1::2, 1::2])


def pool2_bwd(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=0), 2, axis=1)


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up2_bwd(g):
    return g[0::2, 0::2] + g[0::2, 1::2] + g[1::2, 0::2] + g[1::2, 1::2]


# inputs of each layer by activation name; "x" = network input
_IO = [("x", None, "a0a"), ("a0a", None, "a0b"), ("p1", None, "a1a"), ("a1a", None, "a1b"),
       ("p2", None, "a2a"), ("a2a", None, "a2b"), ("p3", None, "a3a"), ("a3a", None, "a3b"),
       ("u2", "a2b", "b2a"), ("b2a", None, "b2b"), ("u1", "a1b", "b1a"), ("b1a", None, "b1b"),
       ("u0", "a0b", "b0a"), ("b0a", None, "b0b"), ("b0b", None, "y")]
_POOL = {2: ("a0b", "p1"), 4: ("a1b", "p2"), 6: ("a2b", "p3")}
_UP = {8: ("a3b", "u2"), 10: ("b2b", "u1"), 12: ("b1b", "u0")}


def _weights(params, bf):
    offs, _ = offsets()
    out = []
    for (t, c1, c2, co, _l), (wo, bo) in zip(LAYERS, offs):
        Wt = np.asarray(params[wo:bo], dtype=np.float64).reshape(9, co, c1 + c2)
        if bf and t == 1:
            Wt = bf16(Wt)
        out.append((Wt, np.asarray(params[bo:bo + co], dtype=np.float64)))
    return out


def forward(params, x, bf16_emulate: bool = False):
    """Network output y (H, W) for input x (H, W); returns (y, cache)."""
    rnd = bf16 if bf16_emulate else (lambda a: a)
    ws = _weights(params, bf16_emulate)
    A = {"x": np.asarray(x, dtype=np.float64)[:, :, None]}
    for l, ((t, c1, c2, co, _lv), (Wt, b)) in enumerate(zip(LAYERS, ws)):
        if l in _POOL:
            s, d = _POOL[l]
            A[d] = rnd(pool2(A[s]))
        if l in _UP:
            s, d = _UP[l]
            A[d] = up2(A[s])
        i1, i2, o = _IO[l]
        xin = A[i1] if i2 is None else np.concatenate([A[i1], A[i2]], axis=2)
        y = conv3x3(xin, Wt, b)
        A[o] = y if t == 2 else rnd(np.maximum(y, 0.0))
    return A["y"][:, :, 0], A


def backward(params, cache, dy, bf16_emulate: bool = False):
    """Gradient of sum(dy * y) with respect to the flat parameter vector."""
    rnd = bf16 if bf16_emulate else (lambda a: a)
    ws = _weights(params, bf16_emulate)
    offs, n = offsets()
    grad = np.zeros(n)
    D = {k: np.zeros_like(v) for k, v in cache.items()}
    D["y"] = np.asarray(dy, dtype=np.float64)[:, :, None]
    for l in range(len(LAYERS) - 1, -1, -1):
        t, c1, c2, co, _lv = LAYERS[l]
        Wt, _b = ws[l]
        i1, i2, o = _IO[l]
        g = D[o] if t == 2 else rnd(D[o] * (cache[o] > 0.0))
        xin = cache[i1] if i2 is None else np.concatenate([cache[i1], cache[i2]], axis=2)
        dx, dW, db = conv3x3_bwd(xin, Wt, g)
        wo, bo = offs[l]
        grad[wo:bo] = dW.ravel()
        grad[bo:bo + co] = db
        if i1 != "x":
            D[i1] += dx[:, :, :c1]
            if i2 is not None:
                D[i2] += dx[:, :, c1:]
        if l in _POOL:
            s, d = _POOL[l]
            D[s] += pool2_bwd(D[d])
        if l in _UP:
            s, d = _UP[l]
            D[s] += up2_bwd(D[d])
    return grad


def loss_grad(params, x, target, mask=None, tnorm2=None, bf16_emulate=False):
    """loss = sum m (y - t)^2 / tnorm2 and its parameter gradient."""
    y, cache = forward(params, x, bf16_emulate)
    m = np.ones_like(y) if mask is None else np.asarray(mask, dtype=np.float64)
    tn = float(np.sum(m * np.asarray(target) ** 2)) if tnorm2 is None else tnorm2
    d = m * (y - target)
    loss = float(np.sum(d * d) / tn)
    return loss, backward(params, cache, 2.0 * d / tn, bf16_emulate)


def adam(w, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in place; returns w."""
    m1 *= b1
    m1 += (1 - b1) * g
    m2 *= b2
    m2 += (1 - b2) * g * g
    w -= lr * (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps)
    return w
