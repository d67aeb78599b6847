"""HIP backend vs the oracle on identical seeded inputs (needs an MI355X).

Gates: neighbour lists, contact key sets and order, and history mapping
bit-exact; FP32 state, forces and torques equal to the oracle.
The kernels reproduce the oracle's FP32 operation order, so the comparisons below
are exact (value equality, +0 == -0); where a comparison is made at a tolerance it
is written next to it.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from sand_atlas_dem_amd.hip import HipBackend

    return HipBackend(0)


@pytest.fixture(scope="module")
def oracle_cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg)


@pytest.fixture(scope="module")
def dense(pkg):
    from sand_atlas_dem_amd import assembly

    return assembly.config_workload("dense", n=300, templates=8)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def rel(a, b):
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(np.asarray(a, np.float64) - b).max()) / scale


def assert_states_equal(a, b, names=("position", "positionRemainder", "velocity", "velocityRemainder", "rotation",
                                     "rotationRemainder", "momentum", "force", "torque", "accumulated", "cell")):
    assert a["contacts"] == b["contacts"]
    assert np.array_equal(a["meta"], b["meta"]), "contact keys/order differ"
    for name in ("slip", "normalArea", *names):
        assert same(a[name], b[name]), f"{name}: max relative difference {rel(a[name], b[name]):.3g}"


def assert_engines_equal(a, b):
    """States equal, and the control words: error flags, pair count, contact count,
    rebuild flag and sliding-contact count (control[0, 1, 2, 4, 6])."""
    assert_states_equal(a.state(), b.state())
    ca, cb = a.read("control", np.uint32, 1), b.read("control", np.uint32, 1)
    assert np.array_equal(ca[[0, 1, 2, 4, 6]], cb[[0, 1, 2, 4, 6]]), (ca[:8], cb[:8])


GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"
GOLDEN_SETS = ("coherent", "brick", "adaptive", "breadth")


def test_mesh_probe_matches_oracle_on_reference_vectors(pkg, hip, oracle_cpu):
    """meshProbe (normal, signed distance) and meshCullProbe bit-equal to the oracle on
    every golden query point the reference holds for the box and prism library
    (4 sets, 9,484 points per shape), plus random points around both."""
    g = pkg.geometry
    box = g.box_shape([1e-3, 2e-3, 3e-3])
    concave = g.make_shape("concave", *g.concave_prism_mesh(), nodes=512, seed=0)
    spec = pkg.Experiment(particles=2, timestep=1e-9)
    engines = [pkg.Engine([box, concave], spec, [[0.005] * 3, [0.015] * 3], [0.03] * 3, [0, 1], [1, 1], backend=m)
               for m in (hip, oracle_cpu)]
    rng = np.random.default_rng(5)
    for sid, shape in ((0, "box"), (1, "concave")):
        golden = [np.load(GOLDEN / f"{tag}_probe_{shape}.npz")["points"] for tag in GOLDEN_SETS]
        pts = np.concatenate(golden + [rng.uniform(-1.5, 1.5, (4096, 3))]).astype(np.float32)
        pts = np.pad(pts, ((0, 0), (0, 1)))
        outs = [m.probe("meshProbe", [len(pts), sid], e.bindings(), pts, np.zeros((len(pts), 4), np.float32))
                for m, e in zip((hip, oracle_cpu), engines)]
        assert same(outs[0], outs[1])
        culls = [m.probe("meshCullProbe", [len(pts), sid], e.bindings(), pts, np.zeros(len(pts), np.uint32))
                 for m, e in zip((hip, oracle_cpu), engines)]
        assert same(culls[0], culls[1])


def test_initial_contacts_match_oracle_exactly(dense, hip, oracle_cpu):
    a, b = dense.engine(hip), dense.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    ctrl_a, ctrl_b = a.read("control", np.uint32, 1), b.read("control", np.uint32, 1)
    assert ctrl_a[1] == ctrl_b[1] and ctrl_a[2] == ctrl_b[2] and ctrl_a[0] == ctrl_b[0] == 0
    n = int(ctrl_b[1])
    assert np.array_equal(a.read("pairs", np.uint32, 2)[:n], b.read("pairs", np.uint32, 2)[:n])
    assert np.array_equal(a.read("contactCounts", np.uint32, 1)[:n], b.read("contactCounts", np.uint32, 1)[:n])
    ma, pa, na = a.contacts()
    mb, pb, nb = b.contacts()
    assert len(mb) > 1000
    assert np.array_equal(ma, mb)
    assert same(pa, pb) and same(na, nb)
    for name in ("force", "torque", "particleStress", "particleDiagnostic"):
        assert same(a.read(name), b.read(name)), name


def test_steps_match_oracle_exactly(dense, hip, oracle_cpu):
    a, b = dense.engine(hip), dense.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    for _ in range(4):
        a.advance(2, rate=dense.rate)
        b.advance(2, rate=dense.rate)
        assert_engines_equal(a, b)
    assert a.read("control", np.uint32, 1)[6] > 0  # some contacts slide: the count is compared


def test_restart_matches_oracle_and_continuous_run(dense, hip, oracle_cpu):
    """restore() rebuilds the hash from the committed prefix and recomputes forces at
    dt=0 (gpu.py:1016-1092).  HIP and oracle restores agree exactly; against the
    continuous run the reference's own gate applies (test_cpu_solver.py:212-238:
    rtol 2e-6, atol 1e-9), because the dt=0 recompute re-projects the slip."""
    runs = {}
    for name, m in (("hip", hip), ("cpu", oracle_cpu)):
        e = dense.engine(m)
        e.initialize()
        e.advance(3)
        state = e.state()
        e.advance(3)
        r = dense.engine(m)
        r.restore(state, 3)
        r.advance(3)
        runs[name] = (e.state(), r.state())
    assert_states_equal(runs["hip"][1], runs["cpu"][1])
    continuous, restored = runs["hip"]
    assert np.array_equal(continuous["meta"], restored["meta"])
    for name in ("position", "velocity", "rotation", "momentum", "force", "torque", "slip", "accumulated"):
        np.testing.assert_allclose(restored[name], continuous[name], rtol=2e-6, atol=1e-9)


def test_contact_capacity_growth_matches_large_capacity(pkg, dense, hip):
    from dataclasses import replace

    small = replace(dense, spec=replace(dense.spec, contacts_per_particle=8))
    a, b = small.engine(hip), dense.engine(hip)
    with pytest.raises(pkg.RejectedStep) as info:
        a.initialize()
    assert info.value.flags & 32
    a = replace(small, spec=replace(small.spec, contacts_per_particle=512)).engine(hip)
    a.initialize()
    b.initialize()
    assert same(a.read("force"), b.read("force"))


def test_collection_matches_reference_generated_collector_fixture(pkg, hip):
    """Contact rows after initial forces and two compressed steps equal, bit for bit,
    those the reference's own generated CPU collector wrote in the build container
    (tests/golden/ref_collector_contacts.py)."""
    from sand_atlas_dem_amd import assembly

    fixture = np.load(GOLDEN / "ref_collector_dense150.npz")
    w = assembly.config_workload("dense", n=150, templates=4, nodes=512, seed=9)
    e = w.engine(hip)
    e.initialize()
    e.advance(int(fixture["steps"]), rate=w.rate)
    meta, point_depth, normal_area = e.contacts()
    assert len(meta) == len(fixture["meta"]) > 1000
    assert np.array_equal(meta, fixture["meta"])
    assert np.array_equal(point_depth.view(np.uint32), fixture["pointDepth"].view(np.uint32))
    assert np.array_equal(normal_area.view(np.uint32), fixture["normalArea"].view(np.uint32))


def test_batched_hull_tests_equal_reference(pkg, hip):
    """The collection stage's batched hull-group tests return exactly the reference
    meshMayContain / meshClusterMayContact answers (smooth grains: 100-500 planes
    per group, several load batches per test)."""
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("uneven", n=300, templates=6)
    e = w.engine(hip)
    e.initialize()
    b = e.bindings()
    info = hip.read(e.buffers["meshHullInfo"], np.uint32, (-1, 4))
    assert info[:, 1].max() > 32
    rng = np.random.default_rng(9)
    for t in range(len(w.shapes)):
        v = w.shapes[t].vertices / e.length_unit
        near = v[rng.integers(0, len(v), 20000)] * rng.uniform(0.9, 1.1, (20000, 1))
        pts = np.concatenate([rng.uniform(v.min(0) - 0.1, v.max(0) + 0.1, (20000, 3)), near])
        pts = np.pad(pts, ((0, 0), (0, 1))).astype(np.float32)
        ref = hip.probe("meshCullProbe", [len(pts), t], b, pts, np.zeros(len(pts), np.uint32))
        fast = hip.probe("meshCullFastProbe", [len(pts), t], b, pts, np.zeros(len(pts), np.uint32))
        assert np.array_equal(ref, fast)
        assert 0.05 < ref.mean() < 0.95
    pairs = e.read("pairs", np.uint32, 2)[: int(e.read("control", np.uint32, 1)[1])]
    clusters = hip.read(e.buffers["meshClusterInfo"], np.uint32, (-1, 4))[:, 1]
    checked = 0
    for i, j in pairs[rng.choice(len(pairs), 200, replace=False)]:
        for src, dst in ((i, j), (j, i)):
            n = int(clusters[w.ids[src]])
            ref = hip.probe("meshClusterProbe", [n, src, dst], b, None, np.zeros(n, np.uint32))
            fast = hip.probe("meshClusterFastProbe", [n, src, dst], b, None, np.zeros(n, np.uint32))
            assert np.array_equal(ref, fast), (src, dst)
            checked += int(ref.sum())
    assert checked > 0


@pytest.mark.parametrize("seed,scales", [(531, [1.0, 0.7]), (532, [4.0, 0.2]), (533, [0.2, 4.0])])
def test_pair_bounds_match_oracle(pkg, hip, oracle_cpu, seed, scales):
    """meshPairBoundsProbe and the pair mask of prepareMeshPairs bit-equal to the oracle
    (which is pinned by FP64 corner enclosure, tests/test_oracle.py)."""
    import test_oracle as t

    rng = np.random.default_rng(seed)
    q = rng.normal(size=(2, 4))
    q = (q / np.linalg.norm(q, axis=1)[:, None]).astype(np.float32)
    out = [t.prepared_pair_bounds(t.pair_bounds_engine(pkg, m, scales), q) for m in (hip, oracle_cpu)]
    assert out[0][0] == out[1][0] == 3 and out[0][1] == out[1][1]
    for a, b in zip(out[0][2], out[1][2]):
        assert same(a, b)
