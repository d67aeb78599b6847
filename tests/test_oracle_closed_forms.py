"""Closed-form pins of the oracle's integrator and friction law -- CPU only.

Each expectation below comes from the mechanics, not from the oracle's formulas:

* a torque-free rigid body whose angular momentum lies along a principal axis spins
  at the constant rate |L|/I about that axis, so q(t) = exp(w t / 2) (x) q0 and L is
  conserved (solver.slang:124-129, 150-158, 859-888);
* a constant torque on a body with isotropic inertia gives L(t) = T t and a
  rotation angle T t^2 / (2 I) about the torque axis;
* in a cell strained at the log rate r a grain at rest follows the affine flow:
  x(t) = x0 e^{r t}, cell(t) = L0 e^{r t} (solver.slang:119-123, 859-897);
* linear drag gamma gives v(t) = v0 e^{-gamma t}, L(t) = L0 e^{-gamma t}, a
  displacement v0 (1 - e^{-gamma t}) / gamma and a rotation angle
  |w0| (1 - e^{-gamma t}) / gamma;
* two cubes pressed face to face and sheared at a constant relative speed v with
  their poses held: each contact's tangential spring force grows as
  kt A v t / 2 while it sticks, the pair's moment is that force times the lever
  h / 2, and once |ft| reaches mu fn it slides at |ft| = mu fn, dissipating
  mu fn v dt per step per contact (solver.slang:703-749).

A factor 1/2 dropped or doubled, a sign flip, a transposed inertia rotation or a
missing tangential projection changes each of these by far more than its tolerance.
"""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg)


def qmul(a, b):
    """Hamilton product, xyzw."""
    av, aw, bv, bw = a[:3], a[3], b[:3], b[3]
    return np.concatenate([aw * bv + bw * av + np.cross(av, bv), [aw * bw - av @ bv]])


def turn(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([np.sin(angle / 2) * axis, [np.cos(angle / 2)]])


def rotate(q, v):
    return qmul(qmul(q, np.concatenate([v, [0.0]])), q * [-1, -1, -1, 1])[:3]


def same_rotation(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return min(np.abs(a - b).max(), np.abs(a + b).max())


def lone_grain(pkg, backend, extents, *, cell=2e-3, position=(7e-4, 9e-4, 1.1e-3)):
    """One box-shaped grain alone in a periodic cell: no contacts ever form."""
    shape = pkg.geometry.box_shape(extents, nodes=64)
    spec = pkg.Experiment(particles=1, timestep=1e-9)
    return pkg.Engine([shape], spec, [list(position)], [cell] * 3, [0], [1], backend=backend)


def body_state(e):
    s = e.state()
    pos = s["position"][0, :3].astype(float) + s["positionRemainder"][0, :3]
    vel = s["velocity"][0, :3].astype(float) + s["velocityRemainder"][0, :3]
    mom = s["momentum"][0, :3].astype(float) + s["momentumRemainder"][0, :3]
    rot = s["rotation"][0].astype(float) + s["rotationRemainder"][0]
    return pos, vel, mom, rot / np.linalg.norm(rot), s


def inertia(e):
    """Normalized inertia tensor of grain 0 (inverse of the uploaded invI rows)."""
    return np.linalg.inv(e.read("invI")[:3, :3].astype(float))


def dt_norm(e):
    return e.spec.timestep / e.time_unit


def test_torque_free_spin_about_a_principal_axis(pkg, cpu):
    """Box 1 x 2 x 3 (principal axes = body axes), rotated by q0, L along the body
    x axis: w = R I^-1 R^T L stays |L|/I_xx along R e_x."""
    rng = np.random.default_rng(41)
    e = lone_grain(pkg, cpu, [1e-4, 2e-4, 3e-4])
    q0 = rng.normal(size=4)
    q0 /= np.linalg.norm(q0)
    axis = rotate(q0, np.array([1.0, 0.0, 0.0]))
    ixx = inertia(e)[0, 0]
    steps, per_step = 500, 2e-3  # angle per step, radians
    rate = per_step / dt_norm(e)
    L0 = axis * rate * ixx
    e.set_kinematics(rotation=q0[None], momentum=L0[None])
    e.initialize()
    e.advance(steps)
    pos, vel, mom, rot, _ = body_state(e)
    expected = qmul(turn(axis, steps * per_step), q0)
    assert same_rotation(rot, expected) < 2e-5
    # The inverse rotation instead (I^-1 applied in the wrong frame) or half the angle
    # would be far off.
    assert same_rotation(rot, qmul(turn(axis, 0.5 * steps * per_step), q0)) > 0.1
    np.testing.assert_allclose(mom, L0, rtol=1e-6)  # no torque: L conserved
    np.testing.assert_allclose(vel, 0, atol=0)
    np.testing.assert_allclose(pos * e.length_unit, [7e-4, 9e-4, 1.1e-3], rtol=1e-7)


def test_torque_free_spin_of_an_isotropic_cube(pkg, cpu):
    rng = np.random.default_rng(42)
    e = lone_grain(pkg, cpu, [1e-4] * 3)
    q0 = rng.normal(size=4)
    q0 /= np.linalg.norm(q0)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    i0 = inertia(e)[0, 0]
    steps, per_step = 1000, 1e-3
    L0 = axis * per_step / dt_norm(e) * i0
    e.set_kinematics(rotation=q0[None], momentum=L0[None])
    e.initialize()
    e.advance(steps)
    _, _, mom, rot, _ = body_state(e)
    assert same_rotation(rot, qmul(turn(axis, steps * per_step), q0)) < 2e-5
    assert same_rotation(rot, qmul(turn(-axis, steps * per_step), q0)) > 0.1  # sense of rotation
    assert abs(np.linalg.norm(mom) - np.linalg.norm(L0)) <= 1e-6 * np.linalg.norm(L0)


def integrate_only(e, steps):
    """beginStep + finishStep only (no force stage): the torque buffer stays as set."""
    encoder = e.m.device.create_command_encoder()
    for _ in range(steps):
        b = e.bindings()
        e.m.dispatch("beginStep", e.n, encoder, r=b)
        e.m.dispatch("finishStep", e.n, encoder, r=b)
    e.m.submit(encoder)
    e.m.device.wait()
    e.check_errors()


def test_constant_torque_on_an_isotropic_body(pkg, cpu):
    """L(t) = T t and angle T t^2 / (2 I): the half kicks around the rotation update
    make the per-step angles (k + 1/2) T dt^2 / I, which sum to the closed form."""
    rng = np.random.default_rng(43)
    e = lone_grain(pkg, cpu, [1e-4] * 3)
    q0 = rng.normal(size=4)
    q0 /= np.linalg.norm(q0)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    i0 = inertia(e)[0, 0]
    steps, dt = 1000, dt_norm(e)
    t = steps * dt
    torque = axis * 2.0 * i0 / t**2  # final angle 1 rad
    e.set_kinematics(rotation=q0[None])
    e.buffers["torque"].copy_from_numpy(np.pad(torque[None], ((0, 0), (0, 1))).astype(np.float32))
    integrate_only(e, steps)
    _, _, mom, rot, _ = body_state(e)
    torque32 = torque.astype(np.float32).astype(float)
    np.testing.assert_allclose(mom, torque32 * t, rtol=2e-6)
    angle = np.linalg.norm(torque32) * t**2 / (2 * i0)
    assert same_rotation(rot, qmul(turn(axis, angle), q0)) < 2e-5


def test_grain_at_rest_follows_the_affine_cell_flow(pkg, cpu):
    """x(t) = x0 e^{r t} per axis and cell(t) = L0 e^{r t}."""
    e = lone_grain(pkg, cpu, [1e-4] * 3)
    steps = 1000
    log_strain = np.array([-0.08, 0.05, 0.02])  # over the whole run
    rate = log_strain / (steps * e.spec.timestep)  # 1/s
    x0 = np.array([7e-4, 9e-4, 1.1e-3]) / e.length_unit
    L0 = np.full(3, 2e-3) / e.length_unit
    e.initialize()
    e.advance(steps, rate=rate)
    pos, vel, _, _, s = body_state(e)
    growth = np.exp(log_strain)
    np.testing.assert_allclose(s["cell"][0, :3], (L0 * growth).astype(np.float32), rtol=2e-7)
    np.testing.assert_allclose(pos, x0 * growth, rtol=1e-6)
    np.testing.assert_allclose(vel, 0, atol=0)  # the peculiar velocity stays zero
    # the opposite sense of the flow would be 2 |r t| away
    assert np.abs(pos / (x0 / growth) - 1).min() > 0.03


def test_peculiar_velocity_decays_in_a_straining_cell(pkg, cpu):
    """v is the velocity relative to the affine flow: dv/dt = -r v per axis."""
    e = lone_grain(pkg, cpu, [1e-4] * 3)
    steps = 1000
    log_strain = np.array([-0.1, 0.04, 0.07])
    rate = log_strain / (steps * e.spec.timestep)
    v0 = np.array([0.3, -0.2, 0.1]) * e.time_unit / e.length_unit
    e.set_kinematics(velocity=v0[None])
    e.initialize()
    e.advance(steps, rate=rate)
    _, vel, _, _, _ = body_state(e)
    np.testing.assert_allclose(vel, v0 * np.exp(-log_strain), rtol=1e-5)


def test_linear_drag_closed_forms(pkg, cpu):
    rng = np.random.default_rng(44)
    e = lone_grain(pkg, cpu, [1e-4] * 3)
    q0 = rng.normal(size=4)
    q0 /= np.linalg.norm(q0)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    steps, dt = 1000, dt_norm(e)
    gamma_si = 1.0 / (steps * e.spec.timestep)  # e-folding over the run
    gamma = gamma_si * e.time_unit
    t = steps * dt
    i0 = inertia(e)[0, 0]
    w0 = 0.8 / t  # 0.8 rad would be turned without drag
    v0 = np.array([0.2, -0.1, 0.05]) * e.time_unit / e.length_unit
    L0 = axis * w0 * i0
    x0 = np.array([7e-4, 9e-4, 1.1e-3]) / e.length_unit
    e.set_kinematics(rotation=q0[None], velocity=v0[None], momentum=L0[None])
    e.initialize()
    e.advance(steps, drag=gamma_si)
    pos, vel, mom, rot, _ = body_state(e)
    decay = np.exp(-gamma * t)
    np.testing.assert_allclose(vel, v0 * decay, rtol=1e-5)
    np.testing.assert_allclose(mom, L0 * decay, rtol=1e-5)
    np.testing.assert_allclose(pos - x0, v0 * (1 - decay) / gamma, rtol=1e-5)
    angle = w0 * (1 - decay) / gamma
    assert same_rotation(rot, qmul(turn(axis, angle), q0)) < 2e-5
    # damping each half kick by the full step would decay as e^{-2 gamma t}
    assert np.abs(vel / (v0 * decay**2) - 1).min() > 0.5


def cube_pair(pkg, backend, friction=0.5):
    """Two 100 um cubes with one quadrature node at each face centre, faces 1e-7 m
    into each other (the reference's six-node pair, tests/test_cpu_solver.py:105-173)."""
    h = 1e-4
    shape = pkg.geometry.box_shape([h] * 3, nodes=64)
    shape.samples = np.concatenate([np.eye(3), -np.eye(3)]) * h / 2
    shape.areas = np.full(6, h * h)
    spec = pkg.Experiment(particles=2, timestep=1e-9, timestep_policy="active-contact", friction=friction,
                          max_overlap_fraction=0.1, contacts_per_particle=128)
    return pkg.Engine([shape], spec, [[0.005] * 3, [0.005 + h - 1e-7, 0.005, 0.005]], [0.01] * 3, [0, 0], [1, 1],
                      backend=backend)


def force_pass(e):
    """One force evaluation with friction history and the poses held fixed."""
    encoder = e.m.device.create_command_encoder()
    e.forces(encoder, e.bindings())
    e.m.submit(encoder)
    e.m.device.wait()
    e.active = 1 - e.active
    e.history = True
    return e.check_errors()


def test_sheared_cube_pair_stick_then_coulomb_slip(pkg, cpu):
    h, area, depth = 1e-4, 1e-8, 1e-7
    e = cube_pair(pkg, cpu, friction=0.5)
    kn, kt, mu = e.spec.normal_stiffness, e.spec.tangential_stiffness, e.spec.friction
    # Grain 1 slides along +y and also recedes along the contact normal (x): only the
    # tangential part of the relative velocity may load the spring.
    v, u = 1.0, 0.5  # m/s
    vel = np.array([[0.0, 0.0, 0.0], [u, v, 0.0]]) * e.time_unit / e.length_unit
    e.set_kinematics(velocity=vel)
    e.initialize()  # dt = 0: history starts empty
    fn = 0.5 * kn * area * depth  # per contact (weight 1/2 per directed contact)
    dt = e.spec.timestep
    stick = mu * fn / (0.5 * kt * area * v)  # time at which |ft| = mu fn
    k_slip = int(np.ceil(stick / dt))
    assert 50 < k_slip < 200
    for k in range(1, k_slip + 60):
        ctrl = force_pass(e)
        force = e.read("force")[:2, :3].astype(float) * e.force_unit
        torque = e.read("torque")[:2, :3].astype(float) * e.energy_unit
        # normal force unchanged (the FP32 gap of the normalized centres sets its 2e-4)
        np.testing.assert_allclose(force[:, 0], [-2 * fn, 2 * fn], rtol=2e-4)
        fn_now = 0.5 * abs(force[0, 0])
        if k < k_slip - 1:
            ft = 0.5 * kt * area * v * k * dt  # per contact
            np.testing.assert_allclose(force[0, 1:], [2 * ft, 0.0], rtol=2e-5, atol=1e-12)
            np.testing.assert_allclose(force[1, 1:], [-2 * ft, 0.0], rtol=2e-5, atol=1e-12)
            # both contact points sit on the shared face, h/2 from each centre
            np.testing.assert_allclose(torque[0], [0, 0, 0.5 * h * force[0, 1]], rtol=2e-3, atol=1e-15)
            np.testing.assert_allclose(torque[1], [0, 0, 0.5 * h * force[0, 1]], rtol=2e-3, atol=1e-15)
            assert ctrl[6] == 0
        elif k > k_slip + 1:
            np.testing.assert_allclose(force[0, 1:], [2 * mu * fn_now, 0.0], rtol=2e-5, atol=1e-12)
            assert ctrl[6] == 2  # both directed contacts slide
            s = e.state()
            # friction work per contact and step: mu fn times the slip v dt
            np.testing.assert_allclose(s["slip"][:, 3].astype(float) * e.energy_unit, mu * fn_now * v * dt, rtol=1e-4)
