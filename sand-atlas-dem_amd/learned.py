"""Learned signed-distance cull for the mesh contact collection (no reference counterpart).

The reference decides every node->triangle query exactly: after the cluster,
rejection-plane and hull-group culls (`mesh.py`, `solver.slang:319-343`) each
surviving surface node walks the target template's BVH.  Most of those nodes are
outside the target surface.  This module trains, per template, a small MLP that
predicts the signed distance of a point in the template frame and lets the
collection stage skip the exact query of a node whose prediction exceeds a
calibrated margin:

    features  xyz and sin/cos(2^k pi u), k = 0..5, of u = (x - centre) * scale
    network   39 (padded to 48) -> 64 -> 64 -> 64 -> 1, ReLU, BF16 weights and
              activations, FP32 accumulation (v_mfma_f32_16x16x16_bf16)
    target    signed distance clamped to +-clamp * R, divided by clamp * R
              (R = template bounding radius), so outputs live in [-1, 1]
    margin    3 x the largest |prediction - target| over a held-out validation set

Training runs on the GPU through the library's own kernels (`sad_hip_nn_grad`:
forward + backward with MFMA, `sad_hip_nn_adam`, `sad_hip_nn_pack`); the labels
come from the library's exact mesh distance query, i.e. from the very function
the cull stands in for.  `sad_hip_nn_sample` draws the points on the GPU, half
within one grain diameter of the surface and half uniform in the bounding box,
redrawing each until it passes the hull-group test the collection applies first
(only such points can reach the exact query).

A cull changes results only if it drops a node that would have been a contact.
`attach(audit=True)` keeps every node, queries it exactly and counts culled
nodes that were contacts ("misses"); a run with zero misses proves the culled
run of the same steps gives identical contacts.
"""

import math
import time

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
STRIDE = 11008  # FP32 master / gradient / Adam stride per template
BSTRIDE = 2 * H * H + H * K1  # BF16 weight block per template
FSTRIDE = 264  # FP32 block: biases, output layer, centre, scale, margin, near bound
FB1, FB2, FB3, FW4, FB4, FCENTER, FSCALE, FMARGIN, FNEAR, FLABEL = 0, 64, 128, 192, 256, 257, 260, 261, 262, 263


def initial_parameters(templates, rng):
    """He-normal hidden layers, 1/sqrt(64) output layer, zero biases."""
    master = np.zeros((templates, STRIDE), np.float32)
    for t in range(templates):
        m = master[t]
        m[OFF_W1:OFF_B1] = rng.normal(0, math.sqrt(2 / IN), H * IN)
        m[OFF_W2:OFF_B2] = rng.normal(0, math.sqrt(2 / H), H * H)
        m[OFF_W3:OFF_B3] = rng.normal(0, math.sqrt(2 / H), H * H)
        m[OFF_W4:OFF_B4] = rng.normal(0, math.sqrt(1 / H), H)
    return master


class LearnedCull:
    """Per-template SDF networks for an `Engine` on the hip backend.

    engine      a built Engine whose backend is `HipBackend`
    points      training points per template (a multiple of `batch`)
    batch       points per template per Adam step (a multiple of 64)
    validation  held-out points per template for the margin
    clamp       label clamp as a fraction of the template bounding radius
    factor      margin = factor x max validation error
    """

    def __init__(self, engine, *, points=1 << 20, batch=16384, validation=16384, clamp=0.2, seed=0, factor=3.0):
        self.engine = engine
        self.m = engine.m
        if getattr(self.m, "name", None) != "hip":
            raise ValueError("LearnedCull runs on the hip backend")
        self.torch = self.m.torch
        self.dev = self.m.torch_device
        self.templates = len(engine.shapes)
        self.batch = 64 * math.ceil(min(batch, points) / 64)
        self.points = self.batch * max(1, points // self.batch)
        self.validation = 64 * math.ceil(validation / 64)
        self.clamp = float(clamp)
        self.factor = float(factor)
        self.rng = np.random.default_rng(seed)
        self.history = []
        self.timings = {}
        t0 = time.perf_counter()
        self._make_data()
        self.timings["data_s"] = time.perf_counter() - t0
        torch = self.torch
        T = self.templates
        self.master = torch.from_numpy(initial_parameters(T, self.rng)).to(self.dev)
        self.m1 = torch.zeros_like(self.master)
        self.m2 = torch.zeros_like(self.master)
        self.grad = torch.zeros_like(self.master)
        self.weights = torch.zeros((T, BSTRIDE), dtype=torch.int16, device=self.dev)
        blocks = np.zeros((T, FSTRIDE), np.float32)
        blocks[:, FCENTER:FCENTER + 3] = self.centre
        blocks[:, FSCALE] = self.scale
        blocks[:, FMARGIN] = np.inf  # no cull before calibration
        self.blocks = torch.from_numpy(blocks).to(self.dev)
        self.slices = max(1, math.ceil(512 / T))
        self.partial = torch.zeros((T, self.slices, STRIDE), dtype=torch.float32, device=self.dev)
        self.loss = torch.zeros(T, dtype=torch.float32, device=self.dev)
        self.step = 0
        self._pack()

    # ---- data
    def _make_data(self):
        """Sample and label every template's points on the GPU (`sad_hip_nn_sample`)."""
        import ctypes as ct

        from .engine import config_values
        from .hip import _table

        e, torch = self.engine, self.torch
        bindings = e.bindings()
        unit = e.length_unit
        T = self.templates
        self.centre = np.zeros((T, 3), np.float32)
        self.scale = np.zeros(T, np.float32)
        self.label_scale = np.zeros(T, np.float32)
        tris, cdfs, starts, counts = [], [], [], []
        box = np.zeros((T, 2, 4), np.float32)
        first = 0
        for t, shape in enumerate(e.shapes):
            vertices = shape.vertices / unit
            lo, hi = vertices.min(0), vertices.max(0)
            centre = 0.5 * (lo + hi)
            self.centre[t] = centre
            self.scale[t] = 1.0 / (0.55 * float((hi - lo).max()))
            radius = float(np.linalg.norm(vertices - centre, axis=1).max())
            self.label_scale[t] = self.clamp * radius
            tri = vertices[shape.faces]
            area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
            cdf = np.cumsum(area) / area.sum()
            cdf[-1] = 1.0
            tris.append(np.pad(tri, ((0, 0), (0, 0), (0, 1))).reshape(-1, 4))
            cdfs.append(cdf)
            starts.append(first)
            counts.append(len(tri))
            first += len(tri)
            box[t, 0, :3], box[t, 0, 3] = lo, shape.equivalent_diameter / unit
            box[t, 1, :3], box[t, 1, 3] = hi, self.label_scale[t]
        dev = self.dev
        tri_d = torch.from_numpy(np.concatenate(tris).astype(np.float32)).to(dev)
        cdf_d = torch.from_numpy(np.concatenate(cdfs).astype(np.float32)).to(dev)
        start_d = torch.from_numpy(np.array(starts, np.uint32).view(np.int32)).to(dev)
        count_d = torch.from_numpy(np.array(counts, np.uint32).view(np.int32)).to(dev)
        box_d = torch.from_numpy(box).to(dev)
        per = self.points + self.validation
        pts = torch.zeros((T * per, 4), dtype=torch.float32, device=dev)
        labels = torch.zeros(T * per, dtype=torch.float32, device=dev)
        seed = int(self.rng.integers(0, 2**63))
        u = (ct.c_uint32 * 4)(T, per, (seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF)
        ptrs, sizes = _table(bindings)
        values = config_values(bindings["c"])
        self._check(self.m.lib.sad_hip_nn_sample(
            u, ptrs, sizes, values.ctypes.data_as(ct.POINTER(ct.c_double)), self._ptr(tri_d), self._ptr(cdf_d),
            self._ptr(start_d), self._ptr(count_d), self._ptr(box_d), self._ptr(pts), self._ptr(labels),
            self.m.stream), "nn_sample")
        self.train_pts, self.train_labels = pts, labels  # training rows: [t * per, t * per + points)
        self.valid_pts = pts.view(T, per, 4)[:, self.points:].reshape(-1, 4).contiguous()
        self.valid_labels = labels.view(T, per)[:, self.points:].cpu().numpy()
        batches = self.points // self.batch
        rows = (np.arange(T, dtype=np.int64)[None, :] * per + np.arange(batches, dtype=np.int64)[:, None] * self.batch)
        if rows.max() + self.batch > 2**32:
            raise ValueError("training set exceeds 32-bit point indices")
        self.tstarts = torch.from_numpy(rows.astype(np.uint32).view(np.int32)).to(dev)
        self.tcount = torch.from_numpy(np.full(T, self.batch, np.uint32).view(np.int32)).to(dev)
        self.valid_tiles = torch.from_numpy(
            np.repeat(np.arange(T, dtype=np.uint32), self.validation // 64).view(np.int32)).to(dev)

    # ---- native calls
    def _ptr(self, x):
        import ctypes as ct

        return ct.c_void_p(x.data_ptr())

    def _check(self, status, what):
        self.m._check(status, what)

    def _pack(self):
        self._check(self.m.lib.sad_hip_nn_pack(self._ptr(self.master), self._ptr(self.weights), self._ptr(self.blocks),
                                               self.templates, self.m.stream), "nn_pack")

    def _grad(self, batch_index):
        self.loss.zero_()
        self._check(self.m.lib.sad_hip_nn_grad(
            self._ptr(self.train_pts), self._ptr(self.train_labels), self._ptr(self.tstarts[batch_index]),
            self._ptr(self.tcount), self.templates, self.slices, self._ptr(self.weights), self._ptr(self.blocks),
            self._ptr(self.partial), self._ptr(self.loss), 1.0 / self.batch, self.m.stream), "nn_grad")

    def gradient(self, batch_index=0):
        """Mean-squared-error gradient of every template on one training batch (into `self.grad`)."""
        self._grad(batch_index)
        self._adam(0.0)
        return self.grad

    def _adam(self, lr):
        self._check(self.m.lib.sad_hip_nn_adam(
            self._ptr(self.master), self._ptr(self.m1), self._ptr(self.m2), self._ptr(self.partial),
            self._ptr(self.grad), self.slices, self.templates, float(lr), 0.9, 0.999, 1e-8, max(1, self.step),
            self.m.stream), "nn_adam")

    def train(self, iterations=2000, lr=2e-3, lr_final=5e-5, log_every=250):
        """Minibatch Adam with cosine decay; returns per-template MSE of the last batch."""
        torch = self.torch
        torch.cuda.synchronize(self.dev)
        t0 = time.perf_counter()
        batches = self.tstarts.shape[0]
        for it in range(iterations):
            self._grad(it % batches)
            self.step += 1
            rate = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * it / max(1, iterations - 1)))
            self._adam(rate)
            self._pack()
            if log_every and (it % log_every == 0 or it == iterations - 1):
                mse = (self.loss / self.batch).cpu().numpy()
                self.history.append((it, float(mse.mean()), float(mse.max())))
        torch.cuda.synchronize(self.dev)
        self.timings["train_s"] = time.perf_counter() - t0
        self.timings["iterations"] = iterations
        return (self.loss / self.batch).cpu().numpy()

    def predict(self, points_tensor, tile_template):
        """Network outputs for 64-point tiles (device tensors)."""
        out = self.torch.zeros(points_tensor.shape[0], dtype=self.torch.float32, device=self.dev)
        self._check(self.m.lib.sad_hip_nn_infer(
            self._ptr(points_tensor), self._ptr(tile_template), int(tile_template.numel()), self._ptr(self.weights),
            self._ptr(self.blocks), self._ptr(out), self.m.stream), "nn_infer")
        return out

    def calibrate(self):
        """Margin per template = factor x max |prediction - target| on the validation set."""
        pred = self.predict(self.valid_pts, self.valid_tiles).cpu().numpy().reshape(self.templates, -1)
        err = np.abs(pred - self.valid_labels)
        self.max_error = err.max(1)
        self.margin = (self.factor * self.max_error).astype(np.float32)
        # A margin at or above the label clamp can never cull (targets saturate at 1):
        # such templates skip the network in the collection stage.
        self.enabled = self.margin < 1.0
        device_margin = np.where(self.enabled, self.margin, np.float32(np.inf)).astype(np.float32)
        self.blocks[:, FMARGIN] = self.torch.from_numpy(device_margin).to(self.dev)
        # Within its validated error the network predicts at most d / label_scale +
        # max_error for a node at signed distance d, so it drops none with
        # d <= (margin - max_error) x label_scale (template units): the collection
        # queries such nodes without evaluating the network.
        self.near = np.where(self.enabled, (self.margin - self.max_error) * self.label_scale, 0.0).astype(np.float32)
        self.blocks[:, FNEAR] = self.torch.from_numpy(self.near).to(self.dev)
        self.blocks[:, FLABEL] = self.torch.from_numpy(self.label_scale.astype(np.float32)).to(self.dev)
        inside = self.valid_labels <= 0
        would_cull = pred > np.where(self.margin < 1.0, self.margin, np.inf)[:, None]
        self.validation_report = {
            "max_error_mean": float(self.max_error.mean()),
            "max_error_max": float(self.max_error.max()),
            "margin_mean": float(self.margin.mean()),
            "cull_fraction_outside": float(would_cull[~inside].mean()) if (~inside).any() else 0.0,
            "validation_misses": int((would_cull & inside).sum()),
            "templates_enabled": int(self.enabled.sum()),
            "templates": int(self.templates),
        }
        return self.margin

    def attach(self, audit=False):
        """Enable the cull in the engine's collection stage."""
        if not hasattr(self, "margin"):
            raise RuntimeError("calibrate() before attach()")
        self.torch.cuda.synchronize(self.dev)
        if not self.enabled.any():  # nothing can be culled: keep the plain collection kernel
            self.m.configure_cull(False)
            return False
        self.m.configure_cull(True, audit, self.weights, self.blocks, self.templates)
        return True

    def detach(self):
        self.torch.cuda.synchronize(self.dev)
        self.m.configure_cull(False)

    def parameters(self):
        return self.master.cpu().numpy()

    def packed(self):
        """(BF16 weight blocks as uint16, FP32 blocks) on the host."""
        return self.weights.cpu().numpy().view(np.uint16), self.blocks.cpu().numpy()
