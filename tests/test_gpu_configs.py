"""The benchmark configurations and the reference Engine's own stepping on the GPU,
against the oracle on identical seeded inputs (needs an MI355X).

* batched uneven meshes (320-5,120 triangles, 256-4,096 nodes per template) through
  the persistent work queues and hipGraph-captured steps;
* the learned cull on the dense, uneven and layer (planes across z) workloads: audited
  (no culled contact), then the culled run equal to the oracle's unculled run;
* the reference Engine's kernel-by-kernel dispatch (no native step encoding), mesh
  and SDF geometry;
* neighbour-list rebuilds while the temporal skip and warm-started queries are on;
* the small configuration at full size for 1,000 steps.

Every comparison is exact (value equality) unless a tolerance is written next to it.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("position", "positionRemainder", "velocity", "velocityRemainder", "rotation", "rotationRemainder",
         "momentum", "momentumRemainder", "force", "torque", "accumulated", "cell", "slip", "normalArea")


def assert_same(a, b, what=""):
    """Engines a (HIP) and b (oracle): committed history, every state row and the
    control words (error flags, pairs, contacts, rebuild flag, sliding contacts)."""
    sa, sb = a.state(), b.state()
    assert sa["contacts"] == sb["contacts"], what
    assert np.array_equal(sa["meta"], sb["meta"]), f"{what}: contact keys or order differ"
    for name in STATE:
        assert np.array_equal(sa[name], sb[name]), f"{what}: {name}"
    ca, cb = a.read("control", np.uint32, 1), b.read("control", np.uint32, 1)
    assert np.array_equal(ca[[0, 1, 2, 4, 6]], cb[[0, 1, 2, 4, 6]]), (what, ca[:8], cb[:8])


@pytest.fixture(scope="module")
def oracle_cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg, threads=16)


@pytest.fixture(scope="module")
def uneven(pkg):
    from sand_atlas_dem_amd import assembly

    return assembly.config_workload("uneven", n=500, templates=16)


def test_uneven_library_spans_the_configured_ranges(uneven):
    tris = np.array([len(s.faces) for s in uneven.shapes])
    nodes = np.array([len(s.samples) for s in uneven.shapes])
    assert tris.min() < 640 and tris.max() > 2560, tris
    assert nodes.min() == 256 and nodes.max() == 4096, nodes


def test_uneven_graph_steps_match_oracle(uneven, oracle_cpu):
    """Ten hipGraph-captured steps on the uneven library, compared every two steps."""
    from sand_atlas_dem_amd.hip import HipBackend

    a, b = uneven.engine(HipBackend(0, graph=True)), uneven.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    assert_same(a, b, "initial")
    for k in range(5):
        a.advance(2, rate=uneven.rate)
        b.advance(2, rate=uneven.rate)
        assert_same(a, b, f"after {2 * k + 2} steps")
    assert a.state()["contacts"] > 1000


@pytest.mark.parametrize("config,skip", [("dense", False), ("uneven", False), ("dense", True), ("uneven", True),
                                         ("layer", True)])
def test_learned_cull_matches_oracle(pkg, oracle_cpu, config, skip):
    """Train the per-template networks on the GPU, audit the cull (every culled node
    is still queried exactly: none may be a contact), then run culled and compare
    with the oracle's unculled run.  Without the temporal skip every queued node
    meets the network; with it (the benchmark's path) nodes known to lie within the
    network's reach of the surface are queried without it, and culled nodes'
    stored bounds stay conservative for the skip."""
    from sand_atlas_dem_amd import assembly, learned
    from sand_atlas_dem_amd.hip import HipBackend

    w = assembly.config_workload(config, n=400, templates=8)
    m = HipBackend(0, temporal_skip=skip)
    e = w.engine(m)
    e.initialize()
    lc = learned.LearnedCull(e, points=1 << 18, batch=8192, validation=8192, seed=2)
    lc.train(iterations=1500, lr=3e-3, log_every=0)
    lc.calibrate()
    assert lc.validation_report["validation_misses"] == 0
    state, step = e.state(), e.step
    e.restore(state, step)
    # Under the skip, anchored tasks pass over the far nodes the network would
    # drop; dropping the anchors makes the first collection re-anchor every task,
    # so the network meets every queued node there.
    m.invalidate()
    assert lc.attach(audit=True), "no template admits a cull"
    before = m.stats()
    e.advance(4, rate=w.rate)
    after = m.stats()
    audited = after["nn_audited"] - before["nn_audited"]
    assert after["nn_misses"] - before["nn_misses"] == 0
    assert audited > 0
    e.restore(state, step)
    m.invalidate()
    lc.attach(audit=False)
    before = m.stats()
    e.advance(4, rate=w.rate)
    after = m.stats()
    lc.detach()
    assert after["nn_culled"] - before["nn_culled"] > 0
    if skip:
        assert after["skipped_nodes"] > before["skipped_nodes"]
    b = w.engine(oracle_cpu)
    b.restore(state, step)
    b.advance(4, rate=w.rate)
    assert_same(e, b, f"{config} culled")


class KernelByKernel:
    """The HIP adapter without its native step encoder: the Engine then dispatches
    every kernel of every step by name through sad_hip_dispatch, as the reference
    Engine drives its backends (gpu.py:714-726)."""

    def __init__(self, inner):
        self.inner = inner
        self.kernels = []

    def __getattr__(self, name):
        if name == "encode_steps":
            raise AttributeError(name)
        return getattr(self.inner, name)

    @property
    def device(self):
        return self

    def dispatch(self, name, n, encoder=None, **kwargs):
        self.kernels.append(name)
        return self.inner.dispatch(name, n, encoder, **kwargs)


@pytest.mark.parametrize("config", ["dense", "sdf"])
def test_reference_engine_dispatch_path_matches_oracle(pkg, oracle_cpu, config):
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    kw = dict(resolution=32) if config == "sdf" else {}
    w = assembly.config_workload(config, n=300, templates=6, **kw)
    m = KernelByKernel(HipBackend(0))
    a, b = w.engine(m), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    m.kernels.clear()
    for k in range(3):
        a.advance(2, rate=w.rate)
        b.advance(2, rate=w.rate)
        assert_same(a, b, f"{config} step {2 * k + 2}")
    per_step = len(m.kernels) / 6
    assert m.kernels.count("beginStep") == m.kernels.count("finishStep") == 6
    assert per_step >= 15, m.kernels[:30]


def test_neighbour_rebuilds_with_skip_and_warm_start_match_oracle(pkg, oracle_cpu):
    """Fast isotropic compression moves the cell by more than half the skin every few
    dozen steps, so the neighbour list is rebuilt while the collection stage skips
    nodes by their anchored bounds and warm-starts its queries."""
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    w = assembly.config_workload("dense", n=300, templates=6)
    hip = HipBackend(0)
    assert hip.tune("temporal_skip") == 1
    rate = (-3e5, -3e5, -3e5)  # 1/s: about 2e-4 strain per step
    a, b = w.engine(hip), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    rebuilds = 0
    anchored_at_rebuilds = []  # anchors taken by the collection right after each rebuild
    derived_at_rebuilds = 0
    for step in range(120):  # one step per call: the anchors persist across calls
        before = hip.stats()
        a.advance(1, rate=rate)
        b.advance(1, rate=rate)
        after = hip.stats()
        rebuilt = int(a.read("control", np.uint32, 1)[4])
        # The per-grain adjacency and task buckets are rebuilt only when the pair
        # list changed (never between neighbour-list rebuilds, although every call
        # uploads the state again).
        derived = after["derived_rebuilds"] - before["derived_rebuilds"]
        if step > 0:
            assert derived <= rebuilt, (step, rebuilt, derived)
            derived_at_rebuilds += derived
        else:
            early = a.state(), b.state(), a.step
        if rebuilt or step % 20 == 19:
            assert_same(a, b, f"step {step + 1}")
            assert np.array_equal(a.read("referencePosition"), b.read("referencePosition"))
        if rebuilt and step > 0:
            tasks = after["directed_pairs"] - before["directed_pairs"]
            anchored_at_rebuilds.append((after["anchored_tasks"] - before["anchored_tasks"]) / tasks)
        rebuilds += rebuilt
    assert rebuilds >= 2, rebuilds
    assert derived_at_rebuilds >= 1, derived_at_rebuilds
    # Restoring the first step's state (an older pair list) into the same workspace:
    # the collection finds the list changed and rebuilds what derives from it.
    before = hip.stats()
    a.restore(early[0], early[2])
    b.restore(early[1], early[2])
    for k in range(3):
        a.advance(1, rate=rate)
        b.advance(1, rate=rate)
        assert_same(a, b, f"step {k + 1} after restoring step 1")
    assert hip.stats()["derived_rebuilds"] > before["derived_rebuilds"]
    assert hip.stats()["skipped_nodes"] > 0
    # A rebuilt list keeps each surviving pair's rows: only new pairs (and pairs at
    # their re-anchor distance) take anchors, not the whole list.
    assert max(anchored_at_rebuilds) < 0.25, anchored_at_rebuilds


def test_small_configuration_full_size_1000_steps(pkg, oracle_cpu):
    """1,000 icospheres, 512 nodes each, phi 0.30, v ~ N(0, 0.01 m/s): 1,000 steps on
    HIP and on the FP32 oracle, equal after every 250."""
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    w = assembly.config_workload("small")
    assert len(w.ids) == 1000
    a, b = w.engine(HipBackend(0, graph=True)), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    for k in range(4):
        a.advance(250)
        b.advance(250)
        assert_same(a, b, f"after {250 * (k + 1)} steps")
    assert a.state()["contacts"] > 100
