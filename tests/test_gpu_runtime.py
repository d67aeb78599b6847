"""Runtime paths that must not change results: hipGraph-captured steps, every
collection launch shape the tuner may pick, and the tuner itself."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("position", "positionRemainder", "velocity", "velocityRemainder", "rotation", "rotationRemainder",
         "momentum", "force", "torque", "accumulated", "cell", "slip", "normalArea", "meta")


@pytest.fixture(scope="module")
def dense(pkg):
    from sand_atlas_dem_amd import assembly

    return assembly.config_workload("dense", n=300, templates=8)


def _equal(a, b):
    assert a["contacts"] == b["contacts"]
    for name in STATE:
        assert np.array_equal(a[name], b[name], equal_nan=True), name


def _run(workload, backend, chunks=(2, 3)):
    e = workload.engine(backend)
    e.initialize()
    for k in chunks:
        e.advance(k, rate=workload.rate)
    return e


def test_graph_captured_steps_match_eager(dense):
    from sand_atlas_dem_amd.hip import HipBackend

    eager = _run(dense, HipBackend(0, graph=False)).state()
    g = HipBackend(0, graph=True)
    captured = _run(dense, g, chunks=(2, 3, 1))
    ref = _run(dense, HipBackend(0, graph=False), chunks=(2, 3, 1)).state()
    _equal(captured.state(), ref)
    # replays after eager work on the same stream stay exact
    captured.advance(2, rate=dense.rate)
    ref_e = _run(dense, HipBackend(0, graph=False), chunks=(2, 3, 1, 2)).state()
    _equal(captured.state(), ref_e)
    assert eager["contacts"] > 0


@pytest.mark.parametrize("setting", [
    {"collect_blocks_per_cu": 2, "collect_pop": 4, "collect_queue": 128, "skip_limit_micro": 2000,
     "collect_leaf_batch": 1, "collect_batch_tasks": 1, "collect_batch_fill": 1},
    {"collect_blocks_per_cu": 8, "collect_pop": 2, "collect_queue": 256, "skip_limit_micro": 250000,
     "collect_leaf_batch": 64, "collect_batch_tasks": 16, "collect_batch_fill": 256},
    {"collect_queue": 128, "collect_batch_tasks": 3, "collect_batch_fill": 17, "skip_rebase": 1,
     "skip_limit_micro": 2000},
])
def test_launch_shapes_do_not_change_results(dense, setting):
    from sand_atlas_dem_amd.hip import HipBackend

    ref = _run(dense, HipBackend(0)).state()
    m = HipBackend(0)
    for name, value in setting.items():
        m.tune(name, value)
    assert {k: m.tune(k) for k in setting} == setting
    _equal(_run(dense, m).state(), ref)


def test_tune_rejects_bad_values(pkg):
    from sand_atlas_dem_amd.hip import HipBackend

    m = HipBackend(0)
    for name, value in (("collect_queue", 64), ("collect_pop", 0), ("collect_blocks_per_cu", 99),
                        ("collect_leaf_batch", 0), ("collect_leaf_batch", 65), ("collect_batch_tasks", 0),
                        ("collect_batch_tasks", 17), ("collect_batch_fill", 0), ("collect_batch_fill", 257), ("skip_rebase", 2),
                        ("nope", 1)):
        with pytest.raises(RuntimeError):
            m.tune(name, value)


def test_autotune_restores_state_and_applies_best(dense, tmp_path):
    from sand_atlas_dem_amd import autotune
    from sand_atlas_dem_amd.hip import HipBackend

    m = HipBackend(0, graph=True)
    e = _run(dense, m)
    before, step = e.state(), e.step
    space = {"collect_blocks_per_cu": (2, 6), "collect_queue": (128, 256)}
    result = autotune.autotune(e, rate=dense.rate, steps=1, space=space, cache=tmp_path / "tune.json")
    assert not result["cached"] and len(result["trials"]) == 3
    assert {k: m.tune(k) for k in space} == result["best"]
    # the engine is left as restore(before) leaves it (forces recomputed at dt = 0,
    # gpu.py:1016-1092), whatever the trials did
    ref = dense.engine(HipBackend(0))
    ref.restore(before, step)
    _equal(e.state(), ref.state())
    assert e.step == step
    again = autotune.autotune(e, rate=dense.rate, steps=1, space=space, cache=tmp_path / "tune.json")
    assert again["cached"] and again["best"] == result["best"]


class _Scan:
    """The reference Engine's hierarchical scan driver (gpu.py:65-91), restated."""

    def __init__(self, m, n):
        self.m, self.n = m, n
        blocks = (n + 255) // 256
        self.totals = m.buffer(count=blocks)
        self.offsets = m.buffer(count=blocks)
        self.child = _Scan(m, blocks) if blocks > 1 else None

    def encode(self, e, source, dest):
        self.m.kernel("blockScan", "scan.slang").dispatch(
            thread_count=[((self.n + 255) // 256) * 256, 1, 1], command_encoder=e, N=self.n, input=source,
            output=dest, totals=self.totals)
        if self.child:
            self.child.encode(e, self.totals, self.offsets)
            self.m.kernel("addBlocks", "scan.slang").dispatch(
                thread_count=[self.n, 1, 1], command_encoder=e, N=self.n, output=dest, offsets=self.offsets)


@pytest.mark.parametrize("n", [1, 255, 256, 257, 65536, 70001, 1 << 20])
def test_reference_scan_kernels_through_the_adapter(n):
    from sand_atlas_dem_amd.hip import HipBackend

    m = HipBackend(0)
    rng = np.random.default_rng(n)
    values = rng.integers(0, 1000, n).astype(np.uint32)
    src, dst = m.buffer(values), m.buffer(count=n)
    e = m.create_command_encoder()
    _Scan(m, n).encode(e, src, dst)
    m.submit(e)
    m.wait()
    want = (np.cumsum(values.astype(np.uint64)) - values).astype(np.uint32)
    assert np.array_equal(m.read(dst, np.uint32)[:n], want)


def test_doctor_runs_the_compiled_kernel_check():
    """doctor() reports the smoke kernel (solver.slang:4-10) passed, as the
    reference's backends do (backend.py:103-119)."""
    from sand_atlas_dem_amd.hip import HipBackend

    report = HipBackend(0).doctor()
    assert report["backend"] == "hip" and report["smoke_passed"] and not report["fallback"]
    assert report["smoke_elements"] == 257 and report["smoke_max_error"] <= 1e-6


def test_serial_collector_name_matches_oracle(pkg, oracle_module):
    """collectMeshContacts (the reference's verification collector) dispatched on
    the HIP backend gives the oracle's contacts, forces and history exactly."""
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    class SerialEngine(pkg.Engine):
        def _encode_contacts(self, e, b):
            self.dispatch("resetMeshContacts", self.pair_capacity, e, b)
            self.dispatch("collectMeshContacts", 64 * self.pair_capacity, e, b)

    w = assembly.config_workload("dense", n=300, templates=4, nodes=256)
    engines = []
    for backend in (HipBackend(0), oracle_module.make_oracle(pkg, threads=8)):
        e = SerialEngine(w.shapes, w.spec, w.positions, w.cell, w.ids, w.scales, backend=backend)
        e.set_kinematics(rotation=w.rotations, velocity=w.velocities * e.time_unit / e.length_unit)
        e.initialize()
        engines.append(e.state())
    a, b = engines
    assert a["contacts"] == b["contacts"] > 1000
    for name in ("meta", "slip", "normalArea", "force", "torque"):
        assert np.array_equal(a[name], b[name]), name


def test_normal_probe_closed_form():
    """normalProbe through the adapter (test_cpu_solver.py:685-709 of the
    reference): the force of rows with positive depth is max(0, w k a d - w c a v),
    bit-equal to the same FP32 operations on the host, zero otherwise."""
    from sand_atlas_dem_amd.hip import HipBackend

    hip = HipBackend(0)
    values = np.array([[1e-7, 1e-8, -0.01, 0], [1e-7, 1e-8, 0.2, 0], [-1e-7, 1e-8, 0, 0], [2e-7, 3e-8, 0.0, 0]],
                      np.float32)
    source, result = hip.buffer(values), hip.buffer(count=len(values), stride=16)
    hip.kernel("normalProbe", "normal.slang").dispatch(thread_count=[len(values), 1, 1], N=len(values),
                                                       stiffness=1e12, viscosity=1e6, weight=0.5, input=source,
                                                       output=result)
    got = hip.read(result, shape=(-1, 4))
    f32 = np.float32
    d, a, v = values[:, 0], values[:, 1], values[:, 2]
    elastic = f32(0.5) * f32(1e12) * a * d
    force = np.where(d > 0, np.maximum(f32(0), elastic - f32(0.5) * f32(1e6) * a * v), f32(0))
    loss = np.where(d > 0, np.maximum(f32(0), (elastic - force) * v), f32(0))
    np.testing.assert_array_equal(got[:, 0], force)
    np.testing.assert_array_equal(got[:, 1], loss)
    reference = np.where(d > 0, np.maximum(0, 0.5 * 1e12 * a * d - 0.5 * 1e6 * a * v), 0)
    np.testing.assert_allclose(got[:, 0], reference, rtol=2e-5, atol=1e-9)
