"""FP32 drift of the HIP path against the FP64 oracle (needs an MI355X).

The small plain workload at full size (1,000 icosphere grains, 512 nodes, phi 0.30,
velocities N(0, 0.01 m/s)) runs 1,000 steps on the HIP path and on the FP64 build of
the oracle from the same FP32 inputs.  The HIP path equals the FP32 oracle bit for
bit (parity tests), so the difference measured here is the FP32 arithmetic's own
drift.  Bounds are written next to each quantity; the measured values are printed.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.mark.timeout(900)
def test_drift_over_1000_steps_against_fp64_oracle(pkg, oracle_module):
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    w = assembly.config_workload("small")
    steps = 1000
    hip = HipBackend(0)
    ref = oracle_module.make_oracle(pkg, threads=_threads(), precision="fp64")
    a, b = w.engine(hip), w.engine(ref)
    a.initialize()
    b.initialize()

    def fp32(name):
        return a.read(name).astype(np.float64)

    def fp64(name):
        return b.buffers[name].read64().reshape(-1, 4)

    # One step's forces from the identical initial state.
    fa, fb = fp32("force")[:, :3], fp64("force")[:, :3]
    ta, tb = fp32("torque")[:, :3], fp64("torque")[:, :3]
    force_err = np.abs(fa - fb).max() / np.abs(fb).max()
    torque_err = np.abs(ta - tb).max() / np.abs(tb).max()
    loaded = np.linalg.norm(fb, axis=1) > 1e-3 * np.linalg.norm(fb, axis=1).max()
    grain_err = np.linalg.norm(fa - fb, axis=1)[loaded] / np.linalg.norm(fb, axis=1)[loaded]

    x0 = (fp64("position") + fp64("positionRemainder"))[:, :3]
    a.advance(steps)
    b.advance(steps)
    xa = (fp32("position") + fp32("positionRemainder"))[:, :3]
    xb = (fp64("position") + fp64("positionRemainder"))[:, :3]
    va = (fp32("velocity") + fp32("velocityRemainder"))[:, :3]
    vb = (fp64("velocity") + fp64("velocityRemainder"))[:, :3]
    la = (fp32("momentum") + fp32("momentumRemainder"))[:, :3]
    lb = (fp64("momentum") + fp64("momentumRemainder"))[:, :3]
    mass = fp64("body")[:, 1]
    box = fp64("cell")[0, :3]
    dx = xa - xb
    dx -= np.round(dx / box) * box
    moved = xb - x0
    moved -= np.round(moved / box) * box
    v_scale = np.sqrt((vb**2).sum(1).mean())
    l_scale = np.sqrt((lb**2).sum(1).mean())
    ke = [0.5 * (mass * (v**2).sum(1)).sum() for v in (va, vb)]
    drift = {
        "force_rel_max": force_err,
        "torque_rel_max": torque_err,
        "force_rel_per_grain_median": float(np.median(grain_err)),
        "position_max_diameters": float(np.abs(dx).max()),
        "largest_displacement_diameters": float(np.abs(moved).max()),
        "velocity_max_rel": float(np.abs(va - vb).max() / v_scale),
        "spin_momentum_max_rel": float(np.abs(la - lb).max() / l_scale),
        "kinetic_energy_rel": float(abs(ke[0] - ke[1]) / ke[1]),
        "contacts": (a.state()["contacts"], b.state()["contacts"]),
    }
    print("\nFP32 HIP vs FP64 oracle after", steps, "steps:", drift)
    # One step from the same state: the largest force/torque error relative to the
    # largest force/torque, and the typical grain's relative error.
    assert force_err < 5e-5 and torque_err < 2e-4
    assert drift["force_rel_per_grain_median"] < 1e-5
    # After 1,000 steps (grains move up to ~0.1 diameters): positions within 1e-4
    # diameters, velocities within 2e-3 of the RMS speed, kinetic energy within 1e-4.
    assert drift["position_max_diameters"] < 1e-4
    assert drift["velocity_max_rel"] < 2e-3
    assert drift["spin_momentum_max_rel"] < 1e-2
    assert drift["kinetic_energy_rel"] < 1e-4
