"""Translating planes on the HIP backend vs the oracle (needs an MI355X).

Same gates as the mesh step: plane-contact keys and order, pair contacts and the
history mapping bit-exact; forces, torques, poses, wall rows, wall work and plane
offsets equal to the oracle's FP32 values (exact comparisons, +0 == -0)."""

from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WALL_NAMES = ("wallPlane", "wallRemainder", "wallVelocity", "wallAccumulated", "wallRows")


@pytest.fixture(scope="module")
def hip():
    from sand_atlas_dem_amd.hip import HipBackend

    return HipBackend(0)


@pytest.fixture(scope="module")
def oracle_cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg, threads=8)


def boxed(viscosity=0.0, inset=0.04, n=300, grain_viscosity=None):
    """The walls workload at test size with every plane pushed `inset` grain
    diameters inward, so each wall starts with a few dozen node contacts."""
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("walls", n=n, templates=6, nodes=256)
    planes = []
    for p in w.spec.planes:
        p = dict(p, offset=p["offset"] + inset * 3e-4)
        if viscosity:
            p["normal_viscosity"] = viscosity
        planes.append(p)
    grain = viscosity if grain_viscosity is None else grain_viscosity
    return replace(w, spec=replace(w.spec, planes=planes, normal_viscosity=grain))


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def assert_states_equal(a, b):
    assert a["contacts"] == b["contacts"]
    assert np.array_equal(a["meta"], b["meta"]), "contact keys/order differ"
    for name in ("slip", "normalArea", "position", "positionRemainder", "velocity", "velocityRemainder", "rotation",
                 "rotationRemainder", "momentum", "force", "torque", "accumulated", *WALL_NAMES):
        assert same(a[name], b[name]), name


@pytest.mark.parametrize("viscosity,grain_viscosity", [(0.0, 0.0), (2e5, 2e5), (2e5, 0.0)])
def test_wall_steps_match_oracle_exactly(hip, oracle_cpu, viscosity, grain_viscosity):
    """Elastic planes; viscous planes and grains; viscous planes only (contact
    kinematics captured because a plane is viscous)."""
    w = boxed(viscosity, grain_viscosity=grain_viscosity)
    a, b = w.engine(hip), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    meta, _, _ = b.contacts()
    planes = meta[:, 1] >= w.spec.particles
    assert planes.sum() > 50, planes.sum()  # the case exercises plane contacts
    assert_states_equal(a.state(), b.state())
    # A lateral wall starts moving part way (refreshWallPower on the change).
    moving = np.zeros((6, 3))
    moving[5] = w.spec.planes[5]["velocity"]
    moving[0] = (0.02, 0.0, 0.0)
    for k in range(4):
        velocities = moving if k >= 2 else None
        a.advance(2, wall_velocities=velocities)
        b.advance(2, wall_velocities=velocities)
        assert_states_equal(a.state(), b.state())
    for x, y in zip(a.walls(), b.walls()):
        assert x["active_surface_contacts"] == y["active_surface_contacts"]
        for key in ("reaction", "reaction_moment_origin", "power", "work", "offset"):
            assert np.array_equal(x[key], y[key]), key


def test_wall_restart_matches_oracle(hip, oracle_cpu):
    w = boxed()
    runs = {}
    for name, m in (("hip", hip), ("cpu", oracle_cpu)):
        e = w.engine(m)
        e.initialize()
        e.advance(3)
        state = e.state()
        r = w.engine(m)
        r.restore(state, 3)
        r.advance(3)
        runs[name] = r.state()
    assert_states_equal(runs["hip"], runs["cpu"])


def test_wall_contacts_on_rotated_spheres_match_oracle(pkg, hip, oracle_cpu):
    """Thirty rotated icospheres part way through a floor: the plane-contact node
    set, order, arms, depths and areas equal the oracle's, which equals brute force
    (CPU wall tests)."""
    rng = np.random.default_rng(3)
    v, f = pkg.geometry.icosphere(2)
    shape = pkg.geometry.make_shape("s", v * 1e-4, f, nodes=512, seed=2)
    n = 30
    centers = np.column_stack([rng.uniform(0.002, 0.008, (n, 2)), rng.uniform(0.00505, 0.0052, n)])
    plane = dict(normal=[0, 0, 1], offset=0.00512, velocity=[0, 0, 0], normal_stiffness=1e12,
                 tangential_stiffness=5e11, friction=0.5)
    spec = pkg.Experiment(periodic=False, planes=[plane], particles=n, timestep=1e-9, max_overlap_fraction=0.9,
                          contacts_per_particle=512)
    q = rng.normal(size=(n, 4))
    out = []
    for m in (hip, oracle_cpu):
        e = pkg.Engine([shape], spec, centers, [0.01] * 3, [0] * n, [1] * n, backend=m)
        e.set_kinematics(rotation=q / np.linalg.norm(q, axis=1)[:, None])
        e.initialize()
        out.append((e.contacts(), e.read("force"), e.read("torque"), e.walls()[0]["reaction"]))
    (ca, fa, ta, ra), (cb, fb, tb, rb) = out
    assert (ca[0][:, 1] >= n).sum() > 100
    for x, y in zip(ca, cb):
        assert same(x, y)
    assert same(fa, fb) and same(ta, tb) and np.array_equal(ra, rb)


def test_box_check_rejects_crossed_walls(hip, oracle_cpu, pkg):
    """advanceWalls flags 4096 once the moving top wall passes the bottom wall."""
    w = boxed()
    planes = [dict(p) for p in w.spec.planes]
    planes[5]["velocity"] = [0.0, 0.0, -1e8]  # crosses the box height in one step
    w = replace(w, spec=replace(w.spec, planes=planes))
    flags = []
    for m in (hip, oracle_cpu):
        e = w.engine(m)
        e.initialize()
        with pytest.raises(pkg.RejectedStep) as info:
            e.advance(1)
        flags.append(info.value.flags)
    assert flags[0] & 4096 and flags[1] & 4096


def test_layer_with_per_axis_periodicity_matches_oracle(hip, oracle_cpu):
    """Planes across z in a cell periodic in x and y (per-axis periodicity), the top
    plane moving and an x-y affine rate: HIP equals the oracle step by step."""
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("layer", n=300, templates=6, nodes=256)
    planes = [dict(p, offset=p["offset"] + 0.04 * 3e-4) for p in w.spec.planes]
    w = replace(w, spec=replace(w.spec, planes=planes))
    a, b = w.engine(hip), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    meta, _, _ = b.contacts()
    assert (meta[:, 1] >= w.spec.particles).sum() > 20
    assert_states_equal(a.state(), b.state())
    for k in range(3):
        rate = (-0.5, -0.5, 0.0) if k else (0.0, 0.0, 0.0)
        a.advance(2, rate=rate)
        b.advance(2, rate=rate)
        assert_states_equal(a.state(), b.state())


def test_layer_servo_matches_oracle(hip, oracle_cpu):
    """Config 5's loading on the layer: a triaxial stage holding the x-y stress at the
    confining pressure (the reference's servo, simulation.py:138-161) while the top plane
    closes.  The HIP run and the oracle's take the same rates from the same diagnostics
    rows and reach the same state."""
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.servo import Stages

    w = assembly.config_workload("layer", n=300, templates=6, nodes=256)
    planes = [dict(p, offset=p["offset"] + 0.04 * 3e-4) for p in w.spec.planes]
    w = replace(w, spec=replace(w.spec, planes=planes))
    stage = dict(kind="triaxial", steps=1000, pressure=2e4, gain=1000.0, filter_time=1e-5, max_rate=1000.0,
                 rate=0.5, stop_strain=0.1)
    runs = []
    for backend in (hip, oracle_cpu):
        e = w.engine(backend)
        e.initialize()
        run = Stages(e, [stage], interval=3)
        assert run.run(max_chunks=4) == "paused"
        runs.append((e, run.rows))
    (a, rows_a), (b, rows_b) = runs
    for ra, rb in zip(rows_a, rows_b):
        assert ra.get("rate") == rb.get("rate")
        assert ra["stress"] == rb["stress"] and ra["layer_stress"] == rb["layer_stress"]
        assert ra["walls"] == rb["walls"]
    assert all(r["rate"][2] == 0.0 and r["rate"][0] != 0.0 for r in rows_a[1:])
    assert_states_equal(a.state(), b.state())
