"""Learned signed-distance cull on the GPU: MFMA layout, network kernels vs the
CPU network oracle, and the cull itself (audit misses, unchanged contacts, fewer
exact queries)."""

import ctypes as ct

import numpy as np
import pytest

from oracle import nn_oracle as nn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(pkg):
    from sand_atlas_dem_amd.hip import HipBackend

    return HipBackend(0)


def _p(t):
    return ct.c_void_p(t.data_ptr())


def _random_master(rng, templates):
    from sand_atlas_dem_amd import learned

    master = learned.initial_parameters(templates, rng)
    for t in range(templates):
        m = master[t]
        for lo, hi in ((nn.OFF_B1, nn.OFF_W2), (nn.OFF_B2, nn.OFF_W3), (nn.OFF_B3, nn.OFF_W4)):
            m[lo:hi] = rng.normal(0, 0.1, hi - lo)
        m[nn.OFF_B4] = 0.03
    return master


@pytest.mark.parametrize("k", [16, 32])
def test_mfma_operand_layout(hip, k):
    """D = A B for one 16x16xk BF16 tile (k = 16, and gfx950's k = 32), exact on small
    integers: pins the per-lane operand layouts the network's layers use."""
    torch = hip.torch
    rng = np.random.default_rng(1)
    A = rng.integers(-8, 9, (16, k)).astype(np.float32)
    B = rng.integers(-8, 9, (k, 16)).astype(np.float32)
    a = torch.from_numpy(nn.bf16_bits(A).view(np.int16)).to(hip.torch_device)
    b = torch.from_numpy(nn.bf16_bits(B).view(np.int16)).to(hip.torch_device)
    d = torch.zeros((16, 16), dtype=torch.float32, device=hip.torch_device)
    hip._check(hip.lib.sad_hip_mfma_probe(_p(a), _p(b), _p(d), k, hip.stream), "mfma probe")
    assert np.array_equal(d.cpu().numpy(), A @ B)


def test_pack_matches_oracle_layout(hip):
    torch = hip.torch
    rng = np.random.default_rng(2)
    T = 3
    master = _random_master(rng, T)
    m = torch.from_numpy(master).to(hip.torch_device)
    wb = torch.zeros((T, nn.BSTRIDE), dtype=torch.int16, device=hip.torch_device)
    fb = torch.zeros((T, nn.FSTRIDE), dtype=torch.float32, device=hip.torch_device)
    hip._check(hip.lib.sad_hip_nn_pack(_p(m), _p(wb), _p(fb), T, hip.stream), "pack")
    wbh, fbh = wb.cpu().numpy().view(np.uint16), fb.cpu().numpy()
    for t in range(T):
        ow, of = nn.pack(master[t, : nn.PARAMS], (0, 0, 0), 0.0, 0.0)
        assert np.array_equal(wbh[t], ow)
        assert np.array_equal(fbh[t, : nn.FCENTER], of[: nn.FCENTER])


def _setup(hip, rng, T, tiles):
    torch = hip.torch
    master = _random_master(rng, T)
    pts = rng.uniform(-1.2, 1.2, (T * tiles * 64, 4)).astype(np.float32)
    labels = rng.uniform(-1, 1, T * tiles * 64).astype(np.float32)
    centre = rng.uniform(-0.1, 0.1, (T, 3)).astype(np.float32)
    scale = np.float32(0.9)
    dev = hip.torch_device
    m = torch.from_numpy(master).to(dev)
    wb = torch.zeros((T, nn.BSTRIDE), dtype=torch.int16, device=dev)
    fbh = np.zeros((T, nn.FSTRIDE), np.float32)
    fbh[:, nn.FCENTER:nn.FCENTER + 3] = centre
    fbh[:, nn.FSCALE] = scale
    fb = torch.from_numpy(fbh).to(dev)
    hip._check(hip.lib.sad_hip_nn_pack(_p(m), _p(wb), _p(fb), T, hip.stream), "pack")
    return master, pts, labels, centre, scale, m, wb, fb


def test_inference_matches_oracle(hip):
    torch = hip.torch
    rng = np.random.default_rng(3)
    T, tiles = 3, 8
    master, pts, _, centre, scale, _, wb, fb = _setup(hip, rng, T, tiles)
    dev = hip.torch_device
    tile_t = torch.from_numpy(np.repeat(np.arange(T, dtype=np.uint32), tiles).view(np.int32)).to(dev)
    p = torch.from_numpy(pts).to(dev)
    out = torch.zeros(len(pts), dtype=torch.float32, device=dev)
    hip._check(hip.lib.sad_hip_nn_infer(_p(p), _p(tile_t), T * tiles, _p(wb), _p(fb), _p(out), hip.stream), "infer")
    got = out.cpu().numpy().reshape(T, -1)
    for t in range(T):
        X = nn.features(pts[t * tiles * 64:(t + 1) * tiles * 64], centre[t], scale)
        want, _ = nn.forward(master[t, : nn.PARAMS], X, bf16_mode=True)
        err = np.abs(got[t] - want)
        # FP32 accumulation order and hardware sin/cos can move a BF16 activation
        # by one ulp (2^-8 relative): bound the worst case and the typical case.
        assert err.max() <= 2e-2 * np.abs(want).max(), err.max()
        assert np.median(err) <= 1e-3 * np.abs(want).max(), np.median(err)


def test_gradient_matches_oracle(hip):
    torch = hip.torch
    rng = np.random.default_rng(4)
    T, tiles, slices = 2, 6, 3
    master, pts, labels, centre, scale, _, wb, fb = _setup(hip, rng, T, tiles)
    dev = hip.torch_device
    n = tiles * 64
    p = torch.from_numpy(pts).to(dev)
    lab = torch.from_numpy(labels).to(dev)
    tstart = torch.from_numpy((np.arange(T, dtype=np.uint32) * n).view(np.int32)).to(dev)
    tcount = torch.from_numpy(np.full(T, n, np.uint32).view(np.int32)).to(dev)
    partial = torch.zeros((T, slices, 11008), dtype=torch.float32, device=dev)
    loss = torch.zeros(T, dtype=torch.float32, device=dev)
    grad = torch.zeros((T, 11008), dtype=torch.float32, device=dev)
    zeros = torch.zeros((T, 11008), dtype=torch.float32, device=dev)
    lib = hip.lib
    hip._check(lib.sad_hip_nn_grad(_p(p), _p(lab), _p(tstart), _p(tcount), T, slices, _p(wb), _p(fb), _p(partial),
                                   _p(loss), 1.0 / n, hip.stream), "grad")
    hip._check(lib.sad_hip_nn_adam(_p(zeros), _p(zeros), _p(zeros), _p(partial), _p(grad), slices, T, 0.0, 0.9,
                                   0.999, 1e-8, 1, hip.stream), "sum")
    g = grad.cpu().numpy()[:, : nn.PARAMS]
    lo = loss.cpu().numpy()
    blocks = [(nn.OFF_W1, nn.OFF_B1), (nn.OFF_B1, nn.OFF_W2), (nn.OFF_W2, nn.OFF_B2), (nn.OFF_B2, nn.OFF_W3),
              (nn.OFF_W3, nn.OFF_B3), (nn.OFF_B3, nn.OFF_W4), (nn.OFF_W4, nn.OFF_B4), (nn.OFF_B4, nn.PARAMS)]
    for t in range(T):
        X = nn.features(pts[t * n:(t + 1) * n], centre[t], scale)
        want = nn.gradient(master[t, : nn.PARAMS], X, labels[t * n:(t + 1) * n], bf16_mode=True)
        y, _ = nn.forward(master[t, : nn.PARAMS], X, True)
        assert abs(lo[t] - np.sum((y - labels[t * n:(t + 1) * n]) ** 2)) <= 1e-3 * lo[t]
        for a, b in blocks:
            rel = np.linalg.norm(g[t, a:b] - want[a:b]) / max(np.linalg.norm(want[a:b]), 1e-30)
            assert rel < 1e-2, (t, a, rel)


def test_sampler_points_pass_hull_test_and_labels_are_exact_distances(pkg, hip):
    """Every sampled point passes the hull-group test (64 redraws suffice here) and
    its label is the exact query's signed distance / label scale, clamped."""
    from sand_atlas_dem_amd import assembly, learned

    w = assembly.config_workload("concave", n=200, templates=2, nodes=256)
    e = w.engine(hip)
    lc = learned.LearnedCull(e, points=4096, batch=1024, validation=1024, seed=3)
    T, per = lc.templates, lc.points + lc.validation
    pts = lc.train_pts.cpu().numpy().reshape(T, per, 4)
    labels = lc.train_labels.cpu().numpy().reshape(T, per)
    b = e.bindings()
    for t in range(T):
        inside = hip.probe("meshCullProbe", [per, t], b, pts[t], np.zeros(per, np.uint32))
        assert inside.mean() > 0.999
        d = hip.probe("meshProbe", [per, t], b, pts[t], np.zeros((per, 4), np.float32))[:, 3]
        want = np.clip(d / lc.label_scale[t], -1, 1).astype(np.float32)
        assert np.array_equal(labels[t], want)
        near = np.abs(d[0::2]) < 1.0  # even points: within one grain diameter of the surface
        assert near.mean() > 0.99
        assert (labels[t] < 0).mean() > 0.1 and (labels[t] > 0).mean() > 0.1


def test_learned_cull_gain_on_concave_grains(pkg):
    """On grains whose hull bounds are loose the trained network removes exact
    queries; the audit finds no culled contact and the culled run is identical to
    the plain run (contacts, forces, torques, poses)."""
    from sand_atlas_dem_amd import assembly, learned
    from sand_atlas_dem_amd.hip import HipBackend

    # Exact-query counts compare the network's cull with the plain walk, so the
    # temporal skip (which removes far nodes for both) stays off here.
    m = HipBackend(0, temporal_skip=False)
    w = assembly.config_workload("concave", n=400, templates=4, nodes=512)
    e = w.engine(m)
    e.initialize()
    lc = learned.LearnedCull(e, points=1 << 20, batch=16384, validation=16384, seed=1)
    lc.train(iterations=4000, lr=3e-3, log_every=0)
    lc.calibrate()
    assert lc.validation_report["validation_misses"] == 0
    state, step = e.state(), e.step

    def run(mode):
        e.restore(state, step)
        if mode == "plain":
            lc.detach()
        else:
            assert lc.attach(audit=(mode == "audit"))
        before = m.stats()
        e.advance(3, rate=w.rate)
        after = m.stats()
        lc.detach()
        return e.state(), {k: after[k] - before[k] for k in after}

    plain, d = run("plain")
    queried = d["queried_nodes"]
    audit, d = run("audit")
    assert d["nn_audited"] == d["nn_culled"] > 0
    assert d["nn_misses"] == 0
    culled, d = run("culled")
    queued = d["queried_nodes"]
    exact = queued - d["nn_culled"]  # queued nodes that still took the exact query
    assert queued == queried
    assert exact <= 0.9 * queried, (exact, queried)
    for other in (audit, culled):
        assert other["contacts"] == plain["contacts"]
        for name in ("meta", "force", "torque", "position", "rotation", "slip", "normalArea"):
            assert np.array_equal(other[name], plain[name]), name
