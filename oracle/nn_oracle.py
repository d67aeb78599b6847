"""CPU oracle of the learned signed-distance network (test infrastructure only).

Only tests/, `__graft_entry__.smoke()` and bench.py's CPU-baseline leg use this
module, as the checker; the product never imports it.

The network has no reference counterpart (the reference only has exact queries).
This restates the device computation in numpy FP64:

    features   u = (p - centre) * scale (FP32), then xyz, sin/cos(pi 2^k u_d)
               ordered [x, y, z, (sin u_x, cos u_x, sin u_y, cos u_y, sin u_z,
               cos u_z) for k = 0..5], zero-padded to 48 columns
    forward    H1 = relu(X W1^T + b1), H2, H3 likewise, y = b4 + H3 . w4
    loss       mean over points of (y - label)^2
    backward   straight-through: gradients of the FP32 network evaluated at the
               rounded activations

With `bf16=True` every value the device stores in BF16 (features, W1..W3,
activations, back-propagated layer gradients) is rounded to nearest-even BF16 at
the same point as on the device; with `bf16=False` the computation is plain FP64
and `gradient` is the exact gradient of `loss`, which the tests pin against
central finite differences.
"""

import numpy as np

IN, K1, H, OCTAVES = 39, 48, 64, 6
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + H
PARAMS = OFF_B4 + 1
BSTRIDE = H * K1 + 2 * H * H
FSTRIDE = 264
FB1, FB2, FB3, FW4, FB4, FCENTER, FSCALE, FMARGIN = 0, 64, 128, 192, 256, 257, 260, 261


def bf16_bits(x):
    """FP32 -> BF16 bit patterns, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(x):
    """Round to BF16 and return the value as FP64."""
    return from_bf16(bf16_bits(np.asarray(x, np.float32))).astype(np.float64)


def unpack(params):
    p = np.asarray(params, np.float64)
    return dict(
        W1=p[OFF_W1:OFF_B1].reshape(H, IN), b1=p[OFF_B1:OFF_W2],
        W2=p[OFF_W2:OFF_B2].reshape(H, H), b2=p[OFF_B2:OFF_W3],
        W3=p[OFF_W3:OFF_B3].reshape(H, H), b3=p[OFF_B3:OFF_W4],
        w4=p[OFF_W4:OFF_B4], b4=p[OFF_B4],
    )


def features(points, centre, scale, rounded=True):
    """Feature rows (N, 39).  `rounded`: FP32 argument arithmetic + BF16 outputs."""
    if rounded:
        u = (np.asarray(points, np.float32)[:, :3] - np.asarray(centre, np.float32)) * np.float32(scale)
        cols = [u.astype(np.float64)]
        freq = np.float32(3.14159265358979)
        for _ in range(OCTAVES):
            a = (freq * u).astype(np.float64)
            cols.append(np.stack([np.sin(a), np.cos(a)], axis=2).reshape(len(u), 6))
            freq = np.float32(freq * 2)
        return bf16(np.concatenate(cols, axis=1).astype(np.float32))
    u = (np.asarray(points, np.float64)[:, :3] - np.asarray(centre, np.float64)) * float(scale)
    cols = [u]
    for k in range(OCTAVES):
        a = np.pi * 2.0**k * u
        cols.append(np.stack([np.sin(a), np.cos(a)], axis=2).reshape(len(u), 6))
    return np.concatenate(cols, axis=1)


def forward(params, X, bf16_mode=True):
    """Network output per feature row; returns (y, cache)."""
    q = unpack(params)
    r = bf16 if bf16_mode else (lambda v: np.asarray(v, np.float64))
    W1, W2, W3 = r(q["W1"]), r(q["W2"]), r(q["W3"])
    X = np.asarray(X, np.float64)[:, :IN]
    H1 = r(np.maximum(X @ W1.T + q["b1"], 0))
    H2 = r(np.maximum(H1 @ W2.T + q["b2"], 0))
    H3 = r(np.maximum(H2 @ W3.T + q["b3"], 0))
    y = q["b4"] + H3 @ q["w4"]
    return y, dict(X=X, H1=H1, H2=H2, H3=H3, W2=W2, W3=W3, q=q, r=r)


def loss(params, X, labels, bf16_mode=False):
    y, _ = forward(params, X, bf16_mode)
    return float(np.mean((y - labels) ** 2))


def gradient(params, X, labels, bf16_mode=True, count=None):
    """Gradient of the mean squared error (mean over `count` points, default len(X))."""
    y, c = forward(params, X, bf16_mode)
    r, q = c["r"], c["q"]
    n = len(X) if count is None else count
    dy = 2.0 * (y - labels) / n
    g = np.zeros(PARAMS)
    H3, H2, H1, X = c["H3"], c["H2"], c["H1"], c["X"]
    g[OFF_B4] = dy.sum()
    g[OFF_W4:OFF_B4] = H3.T @ dy
    g3 = (H3 > 0) * (dy[:, None] * q["w4"][None, :])
    g[OFF_B3:OFF_W4] = g3.sum(0)
    g3 = r(g3)
    g[OFF_W3:OFF_B3] = (g3.T @ H2).reshape(-1)
    g2 = (g3 @ c["W3"]) * (H2 > 0)
    g[OFF_B2:OFF_W3] = g2.sum(0)
    g2 = r(g2)
    g[OFF_W2:OFF_B2] = (g2.T @ H1).reshape(-1)
    g1 = (g2 @ c["W2"]) * (H1 > 0)
    g[OFF_B1:OFF_W2] = g1.sum(0)
    g1 = r(g1)
    g[OFF_W1:OFF_B1] = (g1.T @ X).reshape(-1)
    return g


def pack(params, centre=(0, 0, 0), scale=1.0, margin=np.inf):
    """Device block layout: (BF16 weight block as uint16, FP32 block)."""
    q = unpack(params)
    wb = np.zeros(BSTRIDE, np.uint16)
    W1 = np.zeros((H, K1), np.float32)
    W1[:, :IN] = q["W1"]
    wb[: H * K1] = bf16_bits(W1.reshape(-1))
    wb[H * K1 : H * K1 + H * H] = bf16_bits(q["W2"].astype(np.float32).reshape(-1))
    wb[H * K1 + H * H :] = bf16_bits(q["W3"].astype(np.float32).reshape(-1))
    fb = np.zeros(FSTRIDE, np.float32)
    fb[FB1:FB1 + H] = q["b1"]
    fb[FB2:FB2 + H] = q["b2"]
    fb[FB3:FB3 + H] = q["b3"]
    fb[FW4:FW4 + H] = q["w4"]
    fb[FB4] = q["b4"]
    fb[FCENTER:FCENTER + 3] = centre
    fb[FSCALE] = scale
    fb[FMARGIN] = margin
    return wb, fb
