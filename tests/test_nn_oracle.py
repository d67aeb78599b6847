"""Pins of the learned-cull network oracle (CPU only).

The network has no reference counterpart, so the oracle is pinned by mathematics:
central finite differences of the FP64 loss (catches a dropped term, a wrong
transpose or a wrong activation mask in the backward pass), closed-form feature
values, and BF16 rounding against torch's own float->bfloat16 conversion.
"""

import numpy as np
import pytest
import torch

from oracle import nn_oracle as nn


def _params(rng, scale=1.0):
    p = np.zeros(nn.PARAMS)
    p[nn.OFF_W1:nn.OFF_B1] = rng.normal(0, scale * np.sqrt(2 / nn.IN), nn.H * nn.IN)
    p[nn.OFF_W2:nn.OFF_B2] = rng.normal(0, scale * np.sqrt(2 / nn.H), nn.H * nn.H)
    p[nn.OFF_W3:nn.OFF_B3] = rng.normal(0, scale * np.sqrt(2 / nn.H), nn.H * nn.H)
    p[nn.OFF_W4:nn.OFF_B4] = rng.normal(0, np.sqrt(1 / nn.H), nn.H)
    for lo, hi in ((nn.OFF_B1, nn.OFF_W2), (nn.OFF_B2, nn.OFF_W3), (nn.OFF_B3, nn.OFF_W4)):
        p[lo:hi] = rng.normal(0, 0.1, hi - lo)
    p[nn.OFF_B4] = 0.05
    return p


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(11)
    p = _params(rng)
    X = nn.features(rng.uniform(-1, 1, (96, 3)), (0.1, -0.2, 0.05), 0.9, rounded=False)
    X = np.pad(X, ((0, 0), (0, nn.K1 - nn.IN)))
    labels = rng.uniform(-1, 1, 96)
    g = nn.gradient(p, X, labels, bf16_mode=False)
    blocks = {
        "W1": (nn.OFF_W1, nn.OFF_B1), "b1": (nn.OFF_B1, nn.OFF_W2), "W2": (nn.OFF_W2, nn.OFF_B2),
        "b2": (nn.OFF_B2, nn.OFF_W3), "W3": (nn.OFF_W3, nn.OFF_B3), "b3": (nn.OFF_B3, nn.OFF_W4),
        "w4": (nn.OFF_W4, nn.OFF_B4), "b4": (nn.OFF_B4, nn.PARAMS),
    }
    h = 1e-6
    for name, (lo, hi) in blocks.items():
        for k in rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False):
            a, b = p.copy(), p.copy()
            a[k] += h
            b[k] -= h
            fd = (nn.loss(a, X, labels) - nn.loss(b, X, labels)) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (name, k, fd, g[k])


def test_gradient_vanishes_at_exact_fit_and_is_linear_in_residual():
    rng = np.random.default_rng(3)
    p = _params(rng)
    X = np.pad(nn.features(rng.uniform(-1, 1, (64, 3)), (0, 0, 0), 1.0, rounded=False), ((0, 0), (0, nn.K1 - nn.IN)))
    y, _ = nn.forward(p, X, bf16_mode=False)
    assert np.abs(nn.gradient(p, X, y, bf16_mode=False)).max() < 1e-14
    g1 = nn.gradient(p, X, y - 0.1, bf16_mode=False)
    g2 = nn.gradient(p, X, y - 0.2, bf16_mode=False)
    np.testing.assert_allclose(g2, 2 * g1, rtol=1e-10, atol=1e-14)


def test_features_closed_form_layout():
    pts = np.array([[0.5, 0.0, -0.25]])
    X = nn.features(pts, (0, 0, 0), 1.0, rounded=False)[0]
    assert X.shape == (nn.IN,)
    np.testing.assert_allclose(X[:3], [0.5, 0.0, -0.25])
    # k = 0: sin/cos(pi u) for x, y, z in turn
    np.testing.assert_allclose(X[3:9], [1, 0, 0, 1, -np.sqrt(0.5), np.sqrt(0.5)], atol=1e-15)
    # k = 1: sin/cos(2 pi u)
    np.testing.assert_allclose(X[9:15], [0, -1, 0, 1, -1, 0], atol=1e-15)
    # centre and scale act before the encoding
    Y = nn.features(pts * 2 + 1, (1, 1, 1), 0.5, rounded=False)[0]
    np.testing.assert_allclose(X, Y, atol=1e-15)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    x = np.concatenate([
        rng.normal(0, 1, 20000).astype(np.float32),
        rng.normal(0, 1e-30, 2000).astype(np.float32),
        # exact ties: low half = 0x8000, even and odd upper halves
        (np.arange(2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
        np.array([0.0, -0.0, 1.0, -1.0, 3.0e38], np.float32),
    ])
    x = x[np.isfinite(x)]
    mine = nn.bf16_bits(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(mine, ref)


def test_pack_layout_round_trips():
    rng = np.random.default_rng(2)
    p = _params(rng).astype(np.float32)
    wb, fb = nn.pack(p, (1, 2, 3), 0.5, 0.25)
    W1 = nn.from_bf16(wb[: nn.H * nn.K1]).reshape(nn.H, nn.K1)
    assert np.all(W1[:, nn.IN:] == 0)
    np.testing.assert_allclose(W1[:, : nn.IN], p[nn.OFF_W1:nn.OFF_B1].reshape(nn.H, nn.IN), rtol=2**-8)
    W3 = nn.from_bf16(wb[nn.H * nn.K1 + nn.H * nn.H:]).reshape(nn.H, nn.H)
    np.testing.assert_allclose(W3, p[nn.OFF_W3:nn.OFF_B3].reshape(nn.H, nn.H), rtol=2**-8)
    assert fb[nn.FB4] == p[nn.OFF_B4] and list(fb[nn.FCENTER:nn.FCENTER + 3]) == [1, 2, 3]
    assert fb[nn.FSCALE] == 0.5 and fb[nn.FMARGIN] == 0.25


def test_rounded_forward_tracks_exact_forward():
    """BF16 emulation stays within a few BF16 ulps of the FP64 network."""
    rng = np.random.default_rng(8)
    p = _params(rng)
    pts = rng.uniform(-1, 1, (256, 3))
    Xr = np.pad(nn.features(pts, (0, 0, 0), 1.0), ((0, 0), (0, nn.K1 - nn.IN)))
    Xe = np.pad(nn.features(pts, (0, 0, 0), 1.0, rounded=False), ((0, 0), (0, nn.K1 - nn.IN)))
    yr, _ = nn.forward(p, Xr, True)
    ye, _ = nn.forward(p, Xe, False)
    assert np.abs(yr - ye).max() < 0.05 * np.abs(ye).max()
    with pytest.raises(AssertionError):
        np.testing.assert_array_equal(yr, ye)
