"""Pins of the oracle (the C++ CPU restatement) -- CPU only.

Each check is independent of the oracle's own formulas: the reference's golden
vectors, closed forms, brute force, or the reference's own generated collector.
"""

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg)


def probe_library(pkg, backend):
    """Box (1 x 2 x 3 mm) + concave prism in one library, as the reference probe did
    (evidence/coherent-mesh-query-probe-v1/driver.py:56-72)."""
    g = pkg.geometry
    box = g.box_shape([1e-3, 2e-3, 3e-3])
    concave = g.make_shape("concave", *g.concave_prism_mesh(), nodes=512, seed=0)
    spec = pkg.Experiment(particles=2, timestep=1e-9)
    return pkg.Engine([box, concave], spec, [[0.005] * 3, [0.015] * 3], [0.03] * 3, [0, 1], [1, 1], backend=backend)


@pytest.mark.parametrize("sid,name", [(0, "coherent_probe_box"), (1, "coherent_probe_concave")])
def test_mesh_distance_against_reference_golden_vectors(pkg, cpu, sid, name):
    d = np.load(GOLDEN / f"{name}.npz")
    e = probe_library(pkg, cpu)
    pts = np.pad(d["points"], ((0, 0), (0, 1))).astype(np.float32)
    out = cpu.probe("meshProbe", [len(pts), sid], e.bindings(), pts, np.zeros((len(pts), 4), np.float32))
    l = e.length_unit
    # Independent FP64 distance (analytic box / trimesh prism): the reference's gate 2e-9 m.
    assert np.abs(out[:, 3].astype(float) * l - d["expected"]).max() < 2e-9
    native = d["native"]
    # Same side of the surface as the reference's production query everywhere.
    assert np.array_equal(np.sign(out[:, 3]), np.sign(native[:, 3]))
    assert np.abs(out[:, 3] - native[:, 3]).max() * l < 2e-9
    # Normals = (p - nearest)/|d|; roundoff in p - nearest scales like eps/|d|.
    bound = 2e-5 + 64 * np.finfo(np.float32).eps * (1 + np.abs(pts[:, :3]).max(1)) / np.maximum(np.abs(native[:, 3]), 1e-12)
    assert np.all(np.abs(out[:, :3] - native[:, :3]).max(1) <= bound)


@pytest.mark.parametrize("tag", ["brick", "adaptive", "breadth"])
@pytest.mark.parametrize("sid,shape", [(0, "box"), (1, "concave")])
def test_mesh_distance_against_reference_edge_case_vectors(pkg, cpu, tag, sid, shape):
    """evidence/{brick,adaptive,breadth}-*-probe-v1: 2,820 points per shape at cell
    boundaries, ties, external points and offsets; FP64 truth and the reference's
    production query, both to the reference's own 2e-9 m gate."""
    d = np.load(GOLDEN / f"{tag}_probe_{shape}.npz")
    e = probe_library(pkg, cpu)
    pts = np.pad(d["points"], ((0, 0), (0, 1))).astype(np.float32)
    out = cpu.probe("meshProbe", [len(pts), sid], e.bindings(), pts, np.zeros((len(pts), 4), np.float32))
    l = e.length_unit
    assert np.abs(out[:, 3].astype(float) * l - d["expected"]).max() < 2e-9
    assert np.abs(out[:, 3].astype(float) - d["baseline"][:, 3]).max() * l < 2e-9
    resolved = np.abs(d["baseline"][:, 3]) * l > 2e-9  # off the surface by more than the gate
    assert np.array_equal(np.sign(out[resolved, 3]), np.sign(d["baseline"][resolved, 3]))


def test_box_face_normals_match_analytic_where_resolvable(pkg, cpu):
    """evidence/mesh-face-normal-probe-v2: 100 um cube, analytic distance and normal."""
    d = np.load(GOLDEN / "face_normal_probe_box.npz")
    cube = pkg.geometry.box_shape([1e-4] * 3)
    e = pkg.Engine([cube], pkg.Experiment(particles=1, timestep=1e-9), [[0.005] * 3], [0.01] * 3, [0], [1], backend=cpu)
    pts = np.pad(d["points"], ((0, 0), (0, 1))).astype(np.float32)
    out = cpu.probe("meshProbe", [len(pts), 0], e.bindings(), pts, np.zeros((len(pts), 4), np.float32))
    assert np.abs(out[:, 3] - d["expected_distance"]).max() < 1e-6  # protocol distance_atol
    depth = np.abs(d["expected_distance"])
    resolvable = depth >= 1e-4  # at 1e-6/1e-7 the production operator's normal is roundoff (report.json)
    assert np.abs(out[resolvable, :3] - d["expected_normal"][resolvable]).max() < 2e-5
    assert np.array_equal(np.sign(out[:, 3]), np.sign(d["actual"][:, 3]))


def test_enclosure_cull_never_rejects_inside_points(pkg, cpu):
    """meshMayContain keeps every point with exact signed distance < 0."""
    e = probe_library(pkg, cpu)
    rng = np.random.default_rng(3)
    for sid in (0, 1):
        pts = rng.uniform(-1.2, 1.2, (20000, 4)).astype(np.float32)
        d = cpu.probe("meshProbe", [len(pts), sid], e.bindings(), pts, np.zeros((len(pts), 4), np.float32))
        keep = cpu.probe("meshCullProbe", [len(pts), sid], e.bindings(), pts, np.zeros(len(pts), np.uint32))
        assert np.all(keep[d[:, 3] < 0] == 1)
        assert keep.mean() < 0.9  # it does cull


def test_box_signed_distance_closed_form(pkg, cpu):
    e = probe_library(pkg, cpu)
    rng = np.random.default_rng(11)
    pts = rng.uniform(-1.0, 1.0, (5000, 4)).astype(np.float32)
    d = cpu.probe("meshProbe", [len(pts), 0], e.bindings(), pts, np.zeros((len(pts), 4), np.float32))
    half = np.float32(np.array([0.5e-3, 1e-3, 1.5e-3]) / e.length_unit).astype(float)
    q = np.abs(pts[:, :3].astype(float)) - half
    exact = np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(1), 0)
    assert np.abs(d[:, 3] - exact).max() < 1e-6


def cube_pair(pkg, backend, separation=9.99e-5, friction=0.5):
    """Six-node exact cube pair (reference tests/test_cpu_solver.py:105-173)."""
    h = 1e-4
    shape = pkg.geometry.box_shape([h] * 3, nodes=64)
    shape.samples = np.concatenate([np.eye(3), -np.eye(3)]) * h / 2
    shape.areas = np.full(6, h * h)
    spec = pkg.Experiment(particles=2, timestep=1e-9, timestep_policy="active-contact", friction=friction,
                          max_overlap_fraction=0.1, contacts_per_particle=128)
    return pkg.Engine([shape], spec, [[0.005] * 3, [0.005 + separation, 0.005, 0.005]], [0.01] * 3, [0, 0], [1, 1],
                      backend=backend)


def test_six_node_cube_pair_closed_form_force(pkg, cpu):
    """F = kn * h^2 * overlap = 1e-12 * 1e-8 * 1e-7 = 1e-3 N, zero torque."""
    e = cube_pair(pkg, cpu)
    e.initialize()
    force = cpu.read(e.buffers["force"], shape=(2, 4))[:, :3] * e.force_unit
    torque = cpu.read(e.buffers["torque"], shape=(2, 4))[:, :3] * e.energy_unit
    np.testing.assert_allclose(force[0], [-1e-3, 0, 0], rtol=2e-4, atol=1e-9)
    np.testing.assert_allclose(force[1], [1e-3, 0, 0], rtol=2e-4, atol=1e-9)
    np.testing.assert_allclose(torque, 0, atol=1e-13)
    meta, _, _ = e.contacts()
    assert len(meta) == 2


def test_two_body_oscillator_closed_form(pkg, cpu):
    """Independent two-body oscillator through release (test_cpu_solver.py:23-97)."""
    e = cube_pair(pkg, cpu, separation=9.9e-5, friction=0)
    e.initialize()
    state = e.state()
    pos = (state["position"][:, :3].astype(float) + state["positionRemainder"][:, :3]) * e.length_unit
    h = 1e-4
    overlap = h - (pos[1, 0] - pos[0, 0])
    stiffness = e.spec.normal_stiffness * h * h
    mass = e.spec.density * e.shapes[0].volume
    omega = np.sqrt(2 * stiffness / mass)
    e.advance(500)
    t = 500 * 1e-9
    s = e.state()
    x = (s["position"][:, :3].astype(float) + s["positionRemainder"][:, :3]) * e.length_unit
    displacement = overlap * (1 - np.cos(omega * t)) / 2
    np.testing.assert_allclose(x[:, 0] - pos[:, 0], [-displacement, displacement], atol=1e-9)


def test_collection_equals_exhaustive_per_node_evaluation(pkg, cpu):
    """Cluster + enclosure culls are conservative (test_cpu_solver.py:794-876)."""
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("dense", n=120, templates=4, nodes=256, seed=7)
    e = w.engine(cpu)
    e.initialize()
    meta, pd, na = e.contacts()
    ctrl = e.read("control", np.uint32, 1)
    pairs = e.read("pairs", np.uint32, 2)[: int(ctrl[1])]
    rows = []
    counts = [len(s.samples) for s in w.shapes]
    for i, j in pairs:
        for src, dst in ((i, j), (j, i)):
            n = counts[int(w.ids[src])]
            rows.append(np.column_stack([np.full(n, src), np.full(n, dst), np.arange(n), np.zeros(n)]))
    rows = np.concatenate(rows).astype(np.uint32)
    out = np.zeros((3 * len(rows), 4), np.float32)
    cpu.probe("nodeContactProbe", [len(rows)], e.bindings(), rows, out)
    out = out.reshape(-1, 3, 4)
    hit = out[:, 2, 0] == 1
    exhaustive = {tuple(r[:3]): k for k, r in enumerate(rows) if hit[k]}
    collected = {tuple(m[:3]): k for k, m in enumerate(meta)}
    assert exhaustive.keys() == collected.keys() and len(collected) > 100
    order = [exhaustive[key] for key in collected]
    assert np.array_equal(out[order, 0], pd) and np.array_equal(out[order, 1], na)


def test_neighbour_list_equals_brute_force_minimum_image(pkg, cpu):
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("small", n=400, templates=4, nodes=64, seed=3)
    e = w.engine(cpu)
    e.initialize()
    ctrl = e.read("control", np.uint32, 1)
    pairs = e.read("pairs", np.uint32, 2)[: int(ctrl[1])]
    x = e.read("position")[:, :3].astype(float) + e.read("positionRemainder")[:, :3]
    box = e.read("cell")[0, :3].astype(float)
    r = e.read("body")[:, 2].astype(float)
    d = x[None] - x[:, None]
    d -= np.floor(d / box + 0.5) * box
    dist = np.linalg.norm(d, axis=2)
    i, j = np.nonzero((dist <= r[:, None] + r[None] + e.skin) & (np.arange(e.n)[None] > np.arange(e.n)[:, None]))
    expected = set(zip(i.tolist(), j.tolist()))
    got = [tuple(p) for p in pairs.tolist()]
    assert len(got) == len(set(got))  # no duplicates
    # Exact-threshold ties aside (FP32 vs FP64 range test), the sets agree.
    assert set(got) == expected


def test_slip_transport_closed_cases(pkg, cpu):
    """reference.transport_slip cases (tests/test_geometry.py:44-54)."""
    rows = np.array(
        [[1, 0, 0, 0], [0, 0, 1, 0], [1, 0, 0, 0],  # rotate z-normal to x-normal
         [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 1, np.pi / 2]],  # spin pi/2 about the normal
        np.float32,
    )
    e = probe_library(pkg, cpu)
    out = cpu.probe("transportProbe", [2], e.bindings(), rows, np.zeros((2, 4), np.float32))
    np.testing.assert_allclose(out[0, :3], [0, 0, -1], atol=1e-6)
    np.testing.assert_allclose(out[1, :3], [0, 1, 0], atol=1e-6)


def test_collector_matches_reference_generated_collector(pkg, oracle_module):
    """The reference's own Slang-generated CPU collector, compiled from its tree."""
    if not oracle_module.REF_COLLECTOR.exists():
        pytest.skip("reference collector not built (needs the reference tree)")
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("dense", n=150, templates=4, nodes=512, seed=9)
    out = []
    for collector in ("oracle", "reference"):
        e = w.engine(oracle_module.make_oracle(pkg, collector=collector))
        e.initialize()
        e.advance(2, rate=w.rate)
        out.append(e.contacts())
    for a, b in zip(*out):
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_parallel_collector_equals_serial(pkg, oracle_module):
    """Order-preserving pair-parallel collection (cpu_parallel.hpp:31-117)."""
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("dense", n=150, templates=4, nodes=256, seed=5)
    states = []
    for threads in (1, 4):
        e = w.engine(oracle_module.make_oracle(pkg, threads=threads))
        e.initialize()
        e.advance(2)
        states.append(e.state())
    for name in ("meta", "slip", "normalArea", "force", "torque", "position", "rotation"):
        assert np.array_equal(states[0][name], states[1][name]), name


def pair_bounds_engine(pkg, backend, scales):
    """Two 100 x 70 x 40 um boxes (1,024 nodes) at one point, the reference's fixture of
    test_parallel_mesh.py:135-200."""
    shape = pkg.geometry.box_shape([1e-4, 7e-5, 4e-5], nodes=1024)
    spec = pkg.Experiment(periodic=False, particles=2, contact_geometry="mesh", timestep=1e-10, pairs_per_particle=4)
    return pkg.Engine([shape], spec, [[0.003] * 3, [0.003] * 3], [0.01] * 3, [0, 0], scales, backend=backend)


def prepared_pair_bounds(e, q):
    """Rotations q, one candidate pair (0, 1), prepareMeshPairs; then every cluster's
    hoisted bounds in both directions (meshPairBoundsProbe)."""
    e.buffers["rotation"].copy_from_numpy(q.astype(np.float32))
    control = np.zeros(16, np.uint32)
    control[1] = 1
    e.buffers["control"].copy_from_numpy(control)
    e.m.write_prefix(e.buffers["pairs"], np.array([[0, 1]], np.uint32))
    e.m.dispatch("prepareMeshPairs", e.pair_capacity, r=e.bindings(dt=0))
    e.m.device.wait()
    mask = int(e.m.read(e.buffers["meshPairMask"], np.uint32)[0])
    count = e.buffers["meshClusterRanges"].size // 16
    bounds = [e.m.probe("meshPairBoundsProbe", [count, direction, 0], e.bindings(dt=0), None,
                        np.zeros((count, 2, 4), np.float32)) for direction in (0, 1)]
    return mask, count, bounds


@pytest.mark.parametrize("seed,scales", [(531, [1.0, 0.7]), (532, [4.0, 0.2]), (533, [0.2, 4.0])])
def test_pair_bounds_enclose_exact_fp64_corners(pkg, cpu, seed, scales):
    """meshPairBounds (solver.slang:404-417): each cluster box, carried into the other
    grain's frame by the hoisted affine map and padded, contains the FP64 images of
    all eight corners (the reference's gate, test_parallel_mesh.py:135-200)."""
    e = pair_bounds_engine(pkg, cpu, scales)
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(2, 4))
    q = (q / np.linalg.norm(q, axis=1)[:, None]).astype(np.float32)
    mask, count, bounds = prepared_pair_bounds(e, q)
    assert mask == 3
    low = e.read("meshClusterMin")[:count, :3].astype(float)
    high = e.read("meshClusterMax")[:count, :3].astype(float)
    corners = np.stack([np.where(np.array([c & 1, c & 2, c & 4], bool), high, low) for c in range(8)], axis=1)

    def rotate(r, v):
        return v + 2 * np.cross(r[:3], np.cross(r[:3], v) + r[3] * v)

    for direction, box in enumerate(bounds):
        src, dst = (0, 1) if direction == 0 else (1, 0)
        conj = q[dst].astype(float) * [-1, -1, -1, 1]
        exact = rotate(conj, rotate(q[src].astype(float), corners * float(e.body[src, 0]))) / float(e.body[dst, 0])
        lo, hi = box[:, 0, None, :3].astype(float), box[:, 1, None, :3].astype(float)
        assert np.all(exact >= lo) and np.all(exact <= hi)
        # and tight: no wider than the exact image's extent plus 1e-3 of the coordinates' size
        slack = 1e-3 * (1 + np.abs(exact).max())
        assert np.all((hi - lo)[:, 0] <= (exact.max(1) - exact.min(1)) + slack)
