"""SDF-grid geometry on the HIP backend vs the oracle (needs an MI355X).

The same gates as the mesh step: contact keys and their order, the history
mapping and every FP32 state bit-exact, on a packed SDF assembly under
compression, with graph replay, and with translating planes (every node of a
grain near a plane is tested, as the SDF geometry has no sample tree)."""

from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("slip", "normalArea", "position", "positionRemainder", "velocity", "velocityRemainder", "rotation",
         "rotationRemainder", "momentum", "force", "torque", "accumulated")


@pytest.fixture(scope="module")
def oracle_cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg, threads=8)


@pytest.fixture(scope="module")
def sdf_workload():
    from sand_atlas_dem_amd import assembly

    return assembly.config_workload("sdf", n=600, templates=6, nodes=256, resolution=32)


def same(a, b):
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def assert_states_equal(a, b):
    assert a["contacts"] == b["contacts"]
    assert np.array_equal(a["meta"], b["meta"]), "contact keys/order differ"
    for name in NAMES:
        assert same(a[name], b[name]), name
    for name in ("wallPlane", "wallRemainder", "wallAccumulated", "wallRows"):
        if name in a:
            assert same(a[name], b[name]), name


@pytest.mark.parametrize("graph", [False, True])
def test_sdf_steps_match_oracle_exactly(oracle_cpu, sdf_workload, graph):
    from sand_atlas_dem_amd.hip import HipBackend

    print(f"sdf steps graph={graph}", flush=True)
    w = sdf_workload
    a, b = w.engine(HipBackend(0, graph=graph)), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    assert b.state()["contacts"] > 1000
    assert_states_equal(a.state(), b.state())
    for _ in range(3):
        a.advance(4, rate=w.rate)
        b.advance(4, rate=w.rate)
        assert_states_equal(a.state(), b.state())
    x, y = a.diagnostics(), b.diagnostics()
    for key in ("stress", "kinetic_energy", "fabric", "active_grain_pairs", "active_surface_contacts"):
        assert np.array_equal(np.asarray(x[key]), np.asarray(y[key])), key


def test_sdf_with_planes_matches_oracle_exactly(oracle_cpu):
    from sand_atlas_dem_amd import assembly, geometry

    print("sdf walls", flush=True)
    w = assembly.config_workload("walls", n=300, templates=4, nodes=256)
    planes = [dict(p, offset=p["offset"] + 0.1 * 3e-4) for p in w.spec.planes]
    w = replace(w, shapes=[geometry.with_sdf(s, 24) for s in w.shapes],
                spec=replace(w.spec, planes=planes, contact_geometry="sdf"))
    from sand_atlas_dem_amd.hip import HipBackend

    a, b = w.engine(HipBackend(0)), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    meta, _, _ = b.contacts()
    assert (meta[:, 1] >= w.spec.particles).sum() > 50  # the case exercises plane contacts
    assert_states_equal(a.state(), b.state())
    for _ in range(3):
        a.advance(3)
        b.advance(3)
        assert_states_equal(a.state(), b.state())


def test_sdf_contact_overflow_is_flagged_then_fits(pkg, oracle_cpu, sdf_workload):
    """Too few contact slots raise flag 32 (contactTotal, solver.slang:568-574) and
    write nothing out of bounds; with room, the same step equals the oracle."""
    from sand_atlas_dem_amd.hip import HipBackend

    print("sdf overflow", flush=True)
    hip = HipBackend(0)
    tight = replace(sdf_workload, spec=replace(sdf_workload.spec, contacts_per_particle=4))
    with pytest.raises(pkg.RejectedStep) as info:
        tight.engine(hip).initialize()
    assert info.value.flags & 32
    roomy = replace(sdf_workload, spec=replace(sdf_workload.spec, contacts_per_particle=96))
    a, b = roomy.engine(hip), roomy.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    assert_states_equal(a.state(), b.state())


def test_sdf_small_templates_match_oracle_exactly(pkg, oracle_cpu):
    """Templates of 6, 65 and 200 nodes (one partial leaf, a leaf and one node, several
    partial leaves) and the six-node cube pair's closed form, on the HIP path."""
    from sand_atlas_dem_amd import geometry
    from sand_atlas_dem_amd.hip import HipBackend

    print("sdf small templates", flush=True)
    hip = HipBackend(0)
    h = 1e-4
    cube = geometry.with_sdf(geometry.box_shape([h] * 3, nodes=64), 16)
    cube.samples = np.concatenate([np.eye(3), -np.eye(3)]) * h / 2
    cube.areas = np.full(6, h * h)
    spec = pkg.Experiment(particles=2, timestep=1e-9, timestep_policy="active-contact", friction=0.5,
                          contact_geometry="sdf", max_overlap_fraction=0.1, contacts_per_particle=128)
    engines = [pkg.Engine([cube], spec, [[0.005] * 3, [0.005 + 9.99e-5, 0.005, 0.005]], [0.01] * 3, [0, 0], [1, 1],
                          backend=m) for m in (hip, oracle_cpu)]
    for e in engines:
        e.initialize()
    a, b = (e.state() for e in engines)
    assert_states_equal(a, b)
    np.testing.assert_allclose(a["force"][0, :3] * engines[0].force_unit, [-1e-3, 0, 0], rtol=2e-4, atol=1e-9)
    rng = np.random.default_rng(4)
    v, f = geometry.icosphere(2)
    shapes = [geometry.with_sdf(geometry.make_shape(f"s{k}", v * 3e-4 * (1 + 0.1 * k), f, nodes=n, seed=k), 20)
              for k, n in enumerate((6, 65, 200))]
    n = 200
    spec = pkg.Experiment(particles=n, timestep=1e-9, contact_geometry="sdf", max_overlap_fraction=0.5,
                          contacts_per_particle=128)
    grid = np.stack(np.meshgrid(*[np.arange(6)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    positions = (grid + 0.5) * 5.6e-4 + rng.uniform(-2e-5, 2e-5, (n, 3))
    ids = rng.integers(0, 3, n)
    engines = [pkg.Engine(shapes, spec, positions, [3.36e-3] * 3, ids, [1] * n, backend=m) for m in (hip, oracle_cpu)]
    q = rng.normal(size=(n, 4))
    for e in engines:
        e.set_kinematics(rotation=q / np.linalg.norm(q, axis=1)[:, None])
        e.initialize()
    a, b = (e.state() for e in engines)
    assert b["contacts"] > 50
    assert_states_equal(a, b)
    for e in engines:
        e.advance(3)
    assert_states_equal(engines[0].state(), engines[1].state())


def test_sdf_restart_matches_oracle_and_continuous_run(oracle_cpu, sdf_workload):
    """restore() on the SDF geometry (gpu.py:1016-1092): HIP and oracle restores agree
    exactly; against the continuous run the reference's own gate applies
    (test_cpu_solver.py:212-238: rtol 2e-6, atol 1e-9)."""
    from sand_atlas_dem_amd.hip import HipBackend

    print("sdf restart", flush=True)
    w = sdf_workload
    runs = {}
    for name, m in (("hip", HipBackend(0)), ("cpu", oracle_cpu)):
        e = w.engine(m)
        e.initialize()
        e.advance(3)
        state = e.state()
        e.advance(3)
        r = w.engine(m)
        r.restore(state, 3)
        r.advance(3)
        runs[name] = (e.state(), r.state())
    assert_states_equal(runs["hip"][1], runs["cpu"][1])
    continuous, restored = runs["hip"]
    assert np.array_equal(continuous["meta"], restored["meta"])
    for name in ("position", "velocity", "rotation", "momentum", "force", "torque", "slip", "accumulated"):
        np.testing.assert_allclose(restored[name], continuous[name], rtol=2e-6, atol=1e-9)


def test_sdf_layer_matches_oracle_exactly(oracle_cpu):
    """SDF grains between z planes in an x-y periodic cell (per-axis periodicity)."""
    from sand_atlas_dem_amd import assembly, geometry
    from sand_atlas_dem_amd.hip import HipBackend

    print("sdf layer", flush=True)
    w = assembly.config_workload("layer", n=300, templates=4, nodes=256)
    w = replace(w, shapes=[geometry.with_sdf(s, 24) for s in w.shapes], spec=replace(w.spec, contact_geometry="sdf"))
    a, b = w.engine(HipBackend(0)), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    assert b.state()["contacts"] > 100
    assert_states_equal(a.state(), b.state())
    for _ in range(3):
        a.advance(3, rate=w.rate)
        b.advance(3, rate=w.rate)
        assert_states_equal(a.state(), b.state())
