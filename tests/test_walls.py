"""Translating penalty planes on the oracle (CPU only): the reference's closed forms
(tests/test_cpu_solver.py:240-276, 635-670 of the reference) and brute force."""

import numpy as np
import pytest



@pytest.fixture(scope="module")
def cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg)


def cube_shape(pkg):
    h = 1e-4
    shape = pkg.geometry.box_shape([h] * 3, nodes=64)
    shape.samples = np.concatenate([np.eye(3), -np.eye(3)]) * h / 2
    shape.areas = np.full(6, h * h)
    return shape


def plane(normal, offset, friction=0.5, viscosity=0.0, velocity=(0, 0, 0)):
    p = dict(normal=list(normal), offset=offset, velocity=list(velocity), normal_stiffness=1e12,
             tangential_stiffness=5e11, friction=friction)
    if viscosity:
        p["normal_viscosity"] = viscosity
    return p


def wall_cube(pkg, backend, viscosity=0.0):
    """One cube resting 1e-7 m into a floor (reference make(..., wall=True))."""
    spec = pkg.Experiment(particles=1, periodic=False, planes=[plane([0, 0, 1], 0.005 - 5e-5 + 1e-7, viscosity=viscosity)],
                          timestep=1e-9, timestep_policy="active-contact", normal_viscosity=viscosity,
                          max_overlap_fraction=0.1, contacts_per_particle=128)
    return pkg.Engine([cube_shape(pkg)], spec, [[0.005] * 3], [0.01] * 3, [0], [1], backend=backend)


@pytest.mark.parametrize("viscosity", [0.0, 1e6])
def test_wall_force_motion_loss_and_restart(pkg, cpu, viscosity):
    e = wall_cube(pkg, cpu, viscosity)
    velocity = np.array([[0.01, 0, -0.01]]) * e.time_unit / e.length_unit
    e.set_kinematics(velocity=velocity)
    e.initialize()
    expected = 0.001 + viscosity * 1e-8 * 0.01
    force = cpu.read(e.buffers["force"], shape=(1, 4))[0, :3] * e.force_unit
    assert force[2] == pytest.approx(expected, rel=2e-4, abs=1e-9)
    first = e.walls()[0]
    np.testing.assert_allclose(first["reaction"], -force, rtol=2e-4, atol=1e-9)
    assert first["active_surface_contacts"] == 1
    e.advance(3, wall_velocities=[[0, 0, 0.001]])
    state = e.state()
    e.advance(2)
    continuous = e.state()
    restored = wall_cube(pkg, cpu, viscosity)
    restored.restore(state, 3)
    restored.advance(2)
    again = restored.state()
    for name in ("position", "velocity", "force", "wallPlane", "wallRemainder", "wallAccumulated"):
        np.testing.assert_allclose(again[name], continuous[name], rtol=2e-6, atol=1e-9)
    assert e.walls()[0]["offset"] > first["offset"]


def test_wall_reaction_counts_force_and_moment(pkg, cpu):
    n = 257
    centers = np.array([[0.002 + (i % 17) * 0.0002, 0.002 + (i // 17) * 0.0002, 0.005] for i in range(n)])
    planes = [plane([0, 0, 1], 0.005 - 5e-5 + 1e-7), plane([0, 0, -1], -(0.005 + 5e-5 - 1e-7))]
    spec = pkg.Experiment(periodic=False, planes=planes, particles=n, timestep=1e-9, max_overlap_fraction=0.1)
    e = pkg.Engine([cube_shape(pkg)], spec, centers, [0.01] * 3, [0] * n, [1] * n, backend=cpu)
    e.initialize()
    for normal, wall in zip([np.array([0, 0, 1]), np.array([0, 0, -1])], e.walls()):
        reaction = -0.001 * n * normal
        moment = -np.cross(centers, np.tile(0.001 * normal, (n, 1))).sum(0)
        assert wall["active_surface_contacts"] == n
        np.testing.assert_allclose(wall["reaction"], reaction, rtol=2e-4, atol=1e-9)
        np.testing.assert_allclose(wall["reaction_moment_origin"], moment, rtol=2e-4, atol=1e-12)


def test_wall_contacts_equal_brute_force_node_set(pkg, cpu):
    """The cluster-culled serial collector finds exactly the nodes below the plane."""
    rng = np.random.default_rng(3)
    v, f = pkg.geometry.icosphere(2)
    shape = pkg.geometry.make_shape("s", v * 1e-4, f, nodes=512, seed=2)
    n = 30
    centers = np.column_stack([rng.uniform(0.002, 0.008, (n, 2)), rng.uniform(0.00505, 0.0052, n)])
    spec = pkg.Experiment(periodic=False, planes=[plane([0, 0, 1], 0.00512)], particles=n, timestep=1e-9,
                          max_overlap_fraction=0.9, contacts_per_particle=512)
    e = pkg.Engine([shape], spec, centers, [0.01] * 3, [0] * n, [1] * n, backend=cpu)
    q = rng.normal(size=(n, 4))
    e.set_kinematics(rotation=q / np.linalg.norm(q, axis=1)[:, None])
    e.initialize()
    meta, pd, na = e.contacts()
    wall = meta[:, 1] >= n
    got = {(int(a), int(c)) for a, c in zip(meta[wall, 0], meta[wall, 2])}
    rot = e.read("rotation")[:, :4].astype(np.float64)
    pos = (e.read("position")[:, :3].astype(np.float64) + e.read("positionRemainder")[:, :3])
    samples = shape.samples / e.length_unit
    want, near = set(), 0
    for b in range(n):
        x, y, z, w = rot[b]
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        phi = (pos[b] + samples @ R.T)[:, 2] - 0.00512 / e.length_unit
        near += int(np.sum(np.abs(phi) < 1e-5))
        want |= {(b, k) for k in np.nonzero(phi < -1e-5)[0]}
        got -= {(b, k) for k in np.nonzero(np.abs(phi) < 1e-5)[0]}  # FP32 ties decided either way
    assert len(want) > 100 and near < 20
    assert got == want
