"""Experiment stages and the stress servo (`simulation.py:138-161, 233-315`; stage checks
`spec.py:113-166`): the filter's closed-form step response, the rate rule's fixed points
and limits, the stage sequence on a linear material whose stress follows the cell strain,
the layer's stress and axial strain between its planes, and the servo driving the oracle
on the x-y periodic layer."""

from types import SimpleNamespace

import numpy as np
import pytest

from sand_atlas_dem_amd import assembly
from sand_atlas_dem_amd.servo import Stages, StressServo, validate_stages

SERVO = dict(pressure=1e4, gain=500.0, filter_time=1e-7, max_rate=1e4)


def row(stress, contacts=1):
    return {"stress": np.diag(stress).tolist(), "active_surface_contacts": contacts}


def test_filter_step_response_is_the_first_order_closed_form():
    """Constant stress s for k intervals: filtered = s (1 - exp(-k dt / tau))."""
    servo = StressServo(interval=10, timestep=1e-6)
    stage = dict(SERVO, kind="isotropic", steps=10, filter_time=3e-5)
    s = np.array([3e4, 1e4, 2e4])
    for k in range(1, 21):
        servo.rate(stage, row(s))
        assert np.allclose(servo.filter_stress, s * (1 - np.exp(-k * 1e-5 / 3e-5)), rtol=1e-12, atol=0)


def test_rate_fixed_points_limits_and_contact_phase():
    stage = dict(SERVO, kind="isotropic", steps=10, max_rate=0.8, contact_max_rate=0.3)
    servo = StressServo(interval=10, timestep=1e-6)  # tau << interval: the filter follows the stress
    # Stress 0.1% above / below the pressure, at the pressure: expand / compress / hold,
    # at gain x relative excess.
    r = servo.rate(stage, row([1.001e4, 0.999e4, 1e4], contacts=0))
    assert np.allclose(r, [0.5, -0.5, 0.0], rtol=1e-9, atol=1e-9)
    r = servo.rate(stage, row([2e4, 0.0, 1.0002e4], contacts=0))
    assert np.allclose(r, [0.8, -0.8, 0.1], rtol=1e-9)  # clipped to max_rate
    r = servo.rate(stage, row([2e4, 0.0, 1.0002e4], contacts=3))
    assert np.allclose(r, [0.3, -0.3, 0.1], rtol=1e-9)  # contact phase: contact_max_rate
    r = servo.rate(stage, row([2e4, 0.0, 1.0002e4], contacts=0))
    assert np.allclose(r, [0.3, -0.3, 0.1], rtol=1e-9)  # the phase stays on
    tri = dict(stage, kind="triaxial", rate=0.25)
    assert servo.rate(tri, row([1e4, 1e4, 5e4]))[2] == -0.25
    assert np.array_equal(servo.rate(dict(kind="hold", steps=1), row([5e4] * 3)), np.zeros(3))
    assert np.array_equal(servo.rate(dict(kind="strain", steps=1, rate=[1, 2, 3]), row([5e4] * 3)), [1, 2, 3])


@pytest.mark.parametrize("periodic,walls,stage,message", [
    (False, 6, dict(SERVO, kind="isotropic", steps=5), "Nonperiodic"),
    (True, 0, dict(SERVO, kind="hold", steps=5, wall_velocities=[]), "wall stage"),
    (True, 0, dict(kind="wall", steps=5, wall_velocities=[]), "planes"),
    ((True, True, False), 2, dict(kind="wall", steps=5, wall_velocities=[[0, 0, 0]]), "one row per plane"),
    (True, 0, dict(SERVO, kind="triaxial", steps=5, rate=0.0), "compression rate"),
    (True, 0, dict(kind="isotropic", steps=5, pressure=1e4, gain=1, filter_time=1e-5), "max_rate"),
    (True, 0, dict(kind="hold", steps=0), "steps"),
    (True, 0, dict(kind="hold", steps=5, drag=-1.0), "Drag"),
    (True, 0, dict(kind="shear", steps=5), "Unsupported"),
    ((True, True, False), 2, dict(kind="strain", steps=5, rate=[0, 0, -1]), "periodic axes"),
    (True, 0, dict(SERVO, kind="isotropic", steps=5, converged_samples=0), "converged_samples"),
    (True, 0, dict(SERVO, kind="isotropic", steps=5, contact_max_rate=-1.0), "contact_max_rate"),
])
def test_stage_checks(periodic, walls, stage, message):
    with pytest.raises(ValueError, match=message):
        validate_stages([stage], periodic, walls)


def test_layer_stages_pass_the_checks():
    validate_stages([dict(SERVO, kind="isotropic", steps=5), dict(SERVO, kind="triaxial", steps=5, rate=0.5),
                     dict(kind="strain", steps=5, rate=[0.1, -0.1, 0]),
                     dict(kind="wall", steps=5, wall_velocities=[[0, 0, 0], [0, 0, -1e-3]])], (True, True, False), 2)


class LinearCell:
    """Stand-in engine: stress_ii = s0_ii - K ln(L_i / L0_i) (expansion relieves the
    compression), `advance` integrates the cell's log strain at the given rates.  With
    `layer`, z is bounded by two planes closing at `closing` m/s instead."""

    def __init__(self, *, K=1e6, s0=(5e4, 3e4, 2e4), dt=1e-6, layer=False, closing=0.0):
        periodic = (True, True, False) if layer else True
        self.spec = SimpleNamespace(periodic=periodic, timestep=dt, diagnostic_interval=10)
        self.wall_count = 2 if layer else 0
        self.K, self.s0, self.dt, self.closing = K, np.asarray(s0, float), dt, closing
        self.log = np.zeros(3)
        self.gap, self.cell_z = 1.0, 1.2
        self.initial_cell = np.array([1.0, 1.0, self.cell_z if layer else 1.0])
        self.calls = []

    def advance(self, steps, *, rate, drag=0, wall_velocities=None):
        self.calls.append((steps, tuple(rate), drag, wall_velocities))
        self.log += np.asarray(rate) * steps * self.dt
        if self.wall_count:
            self.gap -= self.closing * steps * self.dt

    def diagnostics(self):
        cell = np.exp(self.log) * [1.0, 1.0, self.cell_z if self.wall_count else 1.0]
        stress = self.s0 - self.K * self.log
        walls = []
        if self.wall_count:
            bottom = 0.1
            walls = [{"normal": [0.0, 0.0, 1.0], "offset": bottom},
                     {"normal": [0.0, 0.0, -1.0], "offset": -(bottom + self.gap)}]
            stress = stress * self.gap / self.cell_z  # the cell's sums over the whole cell volume
        zero = dict.fromkeys(("cell_work", "wall_work", "normal_contact_dissipation", "friction_dissipation",
                              "drag_dissipation", "kinetic_energy", "unbalanced_force"), 0.0)
        return dict(zero, stress=np.diag(stress).tolist(), cell=cell.tolist(), walls=walls,
                    active_surface_contacts=1, elastic_energy_proxy=1.0,
                    axial_strain=float(1 - cell[2] / self.initial_cell[2]))


def test_isotropic_then_triaxial_on_a_linear_cell():
    """gain K dt_interval / p = 0.5 halves the pressure error per interval: the isotropic
    stage converges and hands over; the triaxial stage holds the lateral pressure while
    compressing z at its rate and stops at its strain."""
    e = LinearCell()
    iso = dict(SERVO, kind="isotropic", steps=10_000, converged_samples=5, drag=3.0)
    tri = dict(SERVO, kind="triaxial", steps=10_000, rate=100.0, stop_strain=0.01)
    run = Stages(e, [iso, tri])
    assert run.run() == "complete"
    rows = run.rows
    stages = np.array([r["stage"] for r in rows[1:]])
    first_tri = int(np.argmax(stages == 1)) + 1
    iso_rows, tri_rows = rows[1:first_tri], rows[first_tri:]
    assert all(c[2] == 3.0 for c in e.calls[:len(iso_rows)]) and all(c[2] == 0 for c in e.calls[len(iso_rows):])
    for r in iso_rows[-5:]:
        assert np.allclose(np.diag(r["stress"]), 1e4, rtol=0.01)
    assert np.allclose([np.diag(r["stress"])[:2] for r in tri_rows], 1e4, rtol=0.01)
    assert all(r["rate"][2] == -100.0 for r in tri_rows)
    # Axial strain counts from the cell at the stage change.
    assert np.isclose(tri_rows[0]["axial_strain"], -np.expm1(-100.0 * 1e-5), rtol=1e-6)
    assert tri_rows[-1]["axial_strain"] >= 0.01 > tri_rows[-2]["axial_strain"]
    assert all("work_balance_proxy_residual" in r for r in rows[1:])


def test_isotropic_budget_exhausted_raises_and_pause_resumes():
    e = LinearCell()
    run = Stages(e, [dict(SERVO, kind="isotropic", steps=25, converged_samples=100)])
    assert run.run(max_chunks=2) == "paused"
    assert [c[0] for c in e.calls] == [10, 10]
    with pytest.raises(RuntimeError, match="exhausted"):
        run.run()
    assert [c[0] for c in e.calls] == [10, 10, 5]


def test_layer_servo_acts_on_the_layer_between_its_planes():
    """z bounded by planes: the servo sees the stress over the layer (cell stress x cell
    height / gap), drives x and y only, and the axial strain comes from the gap."""
    e = LinearCell(layer=True, closing=110.0)
    tri = dict(SERVO, kind="triaxial", steps=10_000, rate=0.5, stop_strain=0.01)
    run = Stages(e, [tri])
    assert run.run(max_chunks=40) == "complete"
    rows = run.rows
    assert rows[0]["layer_gap"] == 1.0
    for r in rows[1:]:
        assert r["rate"][2] == 0.0
        assert np.allclose(r["layer_stress"], np.asarray(r["stress"]) * 1.2 / r["layer_gap"], rtol=1e-12)
        assert np.isclose(r["layer_axial_strain"], 1 - r["layer_gap"])
    assert np.allclose(np.diag(rows[-1]["layer_stress"])[:2], 1e4, rtol=0.01)
    assert rows[-1]["layer_axial_strain"] >= 0.01 > rows[-2]["layer_axial_strain"]
    assert all(c[1][2] == 0.0 for c in e.calls)


def test_servo_drives_the_oracle_layer(pkg, oracle_module):
    """The oracle's x-y periodic layer under a triaxial stage: rates on x and y only, the
    step runs at them (its cell log rates), the planes close."""
    cpu = oracle_module.make_oracle(pkg, threads=8)
    w = assembly.config_workload("layer", n=150, templates=3, nodes=128)
    e = w.engine(cpu)
    e.initialize()
    stage = dict(kind="triaxial", steps=100, pressure=2e4, gain=1000.0, filter_time=1e-5, max_rate=1000.0, rate=0.5)
    run = Stages(e, [stage], interval=2)
    assert run.run(max_chunks=3) == "paused"
    rows = run.rows
    for before, r in zip(rows[:-1], rows[1:]):
        rate = np.asarray(r["rate"])
        assert rate[2] == 0.0 and np.all(rate[:2] != 0)
        assert np.allclose(r["cell_log_rate"], rate, rtol=1e-6)  # the rates the step ran at (FP32)
        assert r["layer_gap"] < before["layer_gap"]
        assert len(r["walls"]) == 2


def test_wall_and_strain_stages_pass_their_settings_to_the_step():
    """A wall stage hands its plane velocities to every interval of its steps and runs no
    cell strain; a strain stage runs its fixed rates; the last interval is cut to the
    stage's remaining steps."""
    e = LinearCell(layer=True)
    wall = dict(kind="wall", steps=25, wall_velocities=[[0, 0, 0], [0, 0, -1e-3]], drag=2.0)
    strain = dict(kind="strain", steps=10, rate=[0.5, -0.5, 0.0])
    run = Stages(e, [wall, strain])
    assert run.run() == "complete"
    assert [c[0] for c in e.calls] == [10, 10, 5, 10]
    for c in e.calls[:3]:
        assert c[1] == (0.0, 0.0, 0.0) and c[2] == 2.0 and c[3] == wall["wall_velocities"]
    assert e.calls[3][1] == (0.5, -0.5, 0.0) and e.calls[3][3] is None
    assert [r["stage"] for r in run.rows[1:]] == [0, 0, 0, 1]
