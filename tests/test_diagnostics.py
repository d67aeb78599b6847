"""Diagnostics side path on the oracle (CPU only): the reduction tree's order and
integer fields, and the reference's closed forms for the evidence the step leaves
(tests/test_cpu_solver.py:199-208, 240-276, 635-682, 880-902 and
tests/test_wall_reductions.py:33-77 of the reference), plus Love-Weber stress,
fabric and kinetic energy of configurations whose values the mechanics fixes."""

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


def cube_pair(pkg, backend, separation=9.99e-5, viscosity=0.0):
    """Six-node cube pair overlapping 1e-7 m along x (reference make(...) at
    tests/test_cpu_solver.py:105-173)."""
    spec = pkg.Experiment(particles=2, timestep=1e-9, timestep_policy="active-contact", normal_viscosity=viscosity,
                          max_overlap_fraction=0.1, contacts_per_particle=128)
    return pkg.Engine([cube_shape(pkg)], spec, [[0.005] * 3, [0.005 + separation, 0.005, 0.005]], [0.01] * 3,
                      [0, 0], [1, 1], backend=backend)


def plane(normal, offset):
    return dict(normal=list(normal), offset=offset, velocity=[0, 0, 0], normal_stiffness=1e12,
                tangential_stiffness=5e11, friction=0.5)


# ---------------------------------------------------------------- the tree itself
def reduce(oracle_module, name, records, walls, n):
    return oracle_module.reduce_groups(name, np.asarray(records, np.float32).reshape(-1), walls, n, np.float32)


def test_tree_order_is_stride_halving(oracle_module):
    """2^24 at lane 0 and ones at lanes 1 and 129: the stride-halving tree adds the
    ones together first (lane 1 += lane 129) and returns 2^24 + 2 exactly; a
    sequential sum or an adjacent-pair tree rounds each +1 away and returns 2^24."""
    rows = np.zeros((256, 2, 4), np.float32)
    rows[0, 0, 0] = 2.0**24
    rows[1, 0, 0] = rows[129, 0, 0] = 1.0
    total = reduce(oracle_module, "reduceFabric", rows, 1, 256)
    assert total[0, 0, 0, 0] == 2.0**24 + 2
    sequential = np.float32(0)
    for value in rows[:, 0, 0]:
        sequential = np.float32(sequential + value)
    assert sequential == 2.0**24  # the alternative the check rules out


def test_tree_integer_totals_maximum_and_counts(oracle_module):
    """Exactly representable sums, the row-4 maximum and the wall count field over
    partial groups and two levels (n = 257, as the reference's two-stage case)."""
    rng = np.random.default_rng(5)
    n = 257
    rows = rng.integers(-50, 50, size=(n, 8, 4)).astype(np.float32)
    rows[:, 4, 3] = rng.uniform(0, 1, n).astype(np.float32)
    level = reduce(oracle_module, "reduceRows", rows, 1, n)[0]
    assert level.shape == (2, 8, 4)
    second = reduce(oracle_module, "reduceRows", level, 1, 2)[0, 0]
    want = rows.astype(np.float64).sum(0)
    want[4, 3] = rows[:, 4, 3].max()
    np.testing.assert_array_equal(second, want)
    walls = 3
    wall_rows = rng.integers(-9, 9, size=(walls, n, 3, 4)).astype(np.float32)
    counts = rng.integers(0, 1 << 20, size=(walls, n)).astype(np.uint32)
    wall_rows[:, :, 0, 3] = counts.view(np.float32)
    level = reduce(oracle_module, "reduceWallRows", wall_rows, walls, n)
    top = reduce(oracle_module, "reduceWallRows", level, walls, 2)[:, 0]
    got_counts = top[:, 0, 3].copy().view(np.uint32)
    np.testing.assert_array_equal(got_counts, counts.astype(np.uint64).sum(1))
    expect = wall_rows.astype(np.float64).sum(1)
    expect[:, 0, 3] = top[:, 0, 3]
    np.testing.assert_array_equal(top, expect)


# ---------------------------------------------------------------- engine evidence
def test_pair_at_rest_stress_fabric_and_overlap(pkg, cpu):
    """One pair along x pressed 1e-7 m: total normal force 1e-3 N over the branch
    9.99e-5 m gives sigma_xx = F d / V (Love-Weber), every other component zero;
    fabric n(x)n = e_x e_x; one active pair; overlap fraction 1e-7 / D_eq."""
    e = cube_pair(pkg, cpu)
    e.initialize()
    row = e.diagnostics()
    volume = 0.01**3
    stress = np.array(row["stress"])
    expected = np.zeros((3, 3))
    expected[0, 0] = 1e-3 * 9.99e-5 / volume
    np.testing.assert_allclose(stress, expected, rtol=2e-4, atol=1e-9)
    np.testing.assert_allclose(row["kinetic_stress"], 0, atol=1e-15)
    assert row["p"] == pytest.approx(expected[0, 0] / 3, rel=2e-4)
    assert row["q"] == pytest.approx(-expected[0, 0] / 2, rel=2e-4)
    assert row["active_grain_pairs"] == 1 and row["coordination"] == 1
    np.testing.assert_allclose(row["fabric"], np.diag([1.0, 0, 0]), atol=1e-7)
    diameter = e.shapes[0].equivalent_diameter
    assert row["max_overlap_fraction"] == pytest.approx(1e-7 / diameter, rel=2e-3)
    assert row["unbalanced_force"] == pytest.approx(1.0, rel=1e-5)
    assert row["candidate_pairs"] == 1 and row["active_surface_contacts"] == 2
    solid = 2 * 1e-12
    assert row["void_ratio"] == pytest.approx(volume / solid - 1, rel=1e-6)
    assert row["cell"] == pytest.approx([0.01] * 3, rel=1e-6)  # the FP32 cell, in metres
    assert abs(row["axial_strain"]) < 1e-6 and abs(row["volumetric_strain"]) < 1e-6


def test_pair_kinetic_energy_and_stress(pkg, cpu):
    """Receding pair at +-0.01 m/s: kinetic energy m v^2 and kinetic stress
    sum m v(x)v / V; then advance 10 steps (test_cpu_solver.py:199-208)."""
    e = cube_pair(pkg, cpu, separation=1.2e-4)
    v = np.array([[-0.01, 0, 0], [0.01, 0, 0]])
    e.set_kinematics(velocity=v * e.time_unit / e.length_unit)
    e.initialize()
    row = e.diagnostics()
    mass = e.spec.density * e.shapes[0].volume
    assert row["kinetic_energy"] == pytest.approx(mass * 1e-4, rel=1e-6)
    kinetic = np.zeros((3, 3))
    kinetic[0, 0] = 2 * mass * 1e-4 / 0.01**3
    np.testing.assert_allclose(row["kinetic_stress"], kinetic, rtol=1e-6, atol=1e-15)
    assert row["active_grain_pairs"] == 0 and row["coordination"] == 0
    pressed = cube_pair(pkg, cpu)
    pressed.initialize()
    pressed.advance(10)
    later = pressed.diagnostics()
    assert later["step"] == 10 and later["elapsed_time"] == pytest.approx(1e-8)
    assert later["active_grain_pairs"] == 1
    assert later["kinetic_energy"] > 0 and later["elastic_energy_proxy"] > 0
    assert np.isfinite(later["stress"]).all()


@pytest.mark.parametrize("common", [0.0, 0.7])
def test_viscous_pair_power_and_dissipation(pkg, cpu, common):
    """Approach speed 0.02 m/s, viscosity 1e6: normal power c h^2 v^2 = 1e6 * 1e-8 *
    0.02^2 W, independent of a common velocity (test_cpu_solver.py:880-902)."""
    e = cube_pair(pkg, cpu, viscosity=1e6)
    v = np.array([[0.01 + common, 0, 0], [-0.01 + common, 0, 0]])
    e.set_kinematics(velocity=v * e.time_unit / e.length_unit)
    e.initialize()
    row = e.diagnostics()
    assert row["normal_contact_power"] == pytest.approx(1e6 * 1e-8 * 0.02**2, rel=2e-4)
    e.advance(3)
    assert e.diagnostics()["normal_contact_dissipation"] > 0


def test_wall_reaction_through_diagnostics(pkg, cpu):
    """Cube on a floor (test_cpu_solver.py:240-276): the plane's reaction is minus
    the grain's force; a rising floor raises the plane offset."""
    spec = pkg.Experiment(particles=1, periodic=False, planes=[plane([0, 0, 1], 0.005 - 5e-5 + 1e-7)],
                          timestep=1e-9, timestep_policy="active-contact", max_overlap_fraction=0.1,
                          contacts_per_particle=128)
    e = pkg.Engine([cube_shape(pkg)], spec, [[0.005] * 3], [0.01] * 3, [0], [1], backend=cpu)
    e.set_kinematics(velocity=np.array([[0.01, 0, -0.01]]) * e.time_unit / e.length_unit)
    e.initialize()
    row = e.diagnostics()
    force = cpu.read(e.buffers["force"], shape=(1, 4))[0, :3] * e.force_unit
    assert force[2] == pytest.approx(0.001, rel=2e-4)
    wall = row["walls"][0]
    np.testing.assert_allclose(wall["reaction"], -force, rtol=2e-4, atol=1e-9)
    assert wall["active_surface_contacts"] == 1 and wall["index"] == 0
    e.advance(3, wall_velocities=[[0, 0, 0.001]])
    final = e.diagnostics()
    assert final["normal_contact_dissipation"] >= 0
    assert final["walls"][0]["offset"] > wall["offset"]
    assert final["walls"][0]["velocity"] == pytest.approx([0, 0, 0.001])


def test_two_level_wall_reduction_counts_and_moments(pkg, cpu):
    """257 cubes between two planes (test_cpu_solver.py:635-682): two reduction
    levels, exact integer counts, reaction -1e-3 n N and its moment about the origin;
    the per-plane counts add up to the surface contact total."""
    n = 257
    centers = np.array([[0.002 + (i % 17) * 0.0002, 0.002 + (i // 17) * 0.0002, 0.005] for i in range(n)])
    planes = [plane([0, 0, 1], 0.005 - 5e-5 + 1e-7), plane([0, 0, -1], -(0.005 + 5e-5 - 1e-7))]
    spec = pkg.Experiment(periodic=False, planes=planes, particles=n, timestep=1e-9, max_overlap_fraction=0.1)
    e = pkg.Engine([cube_shape(pkg)], spec, centers, [0.01] * 3, [0] * n, [1] * n, backend=cpu)
    e.initialize()
    row = e.diagnostics()
    assert len(e.wall_diag_buffers) == 2
    for normal, wall in zip([np.array([0, 0, 1]), np.array([0, 0, -1])], row["walls"]):
        reaction = -0.001 * n * normal
        moment = -np.cross(centers, np.tile(0.001 * normal, (n, 1))).sum(0)
        assert wall["active_surface_contacts"] == n
        np.testing.assert_allclose(wall["reaction"], reaction, rtol=2e-4, atol=1e-9)
        np.testing.assert_allclose(wall["reaction_moment_origin"], moment, rtol=2e-4, atol=1e-12)
    assert sum(w["active_surface_contacts"] for w in row["walls"]) == row["active_surface_contacts"]
    for device, host in zip(row["walls"], e.walls()):
        np.testing.assert_allclose(device["reaction"], host["reaction"], rtol=1e-5)
        assert device["active_surface_contacts"] == host["active_surface_contacts"]


def test_boundary_box_cell_follows_planes(pkg, cpu):
    """With a boundary box the reported cell spans the six plane offsets, and the
    strains are measured against that initial box."""
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("walls", n=64, templates=2, nodes=64)
    assert w.spec.boundary_box == [0, 1, 2, 3, 4, 5]
    e = w.engine(cpu)
    e.initialize()
    row = e.diagnostics()
    offsets = np.array([p["offset"] for p in w.spec.planes])
    want = pkg.spec.box_extents(offsets, w.spec.boundary_box, np.array(row["domain_cell"]))
    np.testing.assert_allclose(row["cell"], want, rtol=1e-6)
    assert abs(row["volumetric_strain"]) < 1e-6
    assert len(row["walls"]) == 6
