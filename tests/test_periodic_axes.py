"""Per-axis periodicity (no reference counterpart; the reference's periodic flag
covers every axis and forbids planes): validation, encoding, and the oracle checked
against a brute-force minimum-image neighbour set, translation invariance along a
periodic axis, and escapes through the open axis."""

import numpy as np
import pytest

from sand_atlas_dem_amd import assembly
from sand_atlas_dem_amd.spec import periodic_code


@pytest.fixture(scope="module")
def cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg, threads=8)


def test_encoding_and_validation(pkg):
    assert periodic_code(True) == 1 and periodic_code(False) == 0
    assert periodic_code((True, True, True)) == 1 and periodic_code((False, False, False)) == 0
    assert periodic_code((True, True, False)) == 8 | 3 and periodic_code((False, True, False)) == 8 | 2
    plane = dict(normal=[0, 0, 1], offset=0.0, velocity=[0, 0, 0], normal_stiffness=1.0, tangential_stiffness=1.0,
                 friction=0.5)
    pkg.Experiment(periodic=(True, True, False), planes=[plane])
    with pytest.raises(ValueError):
        pkg.Experiment(periodic=(True, False, True), planes=[plane])  # normal along a periodic axis
    with pytest.raises(ValueError):
        pkg.Experiment(periodic=True, planes=[plane])  # the reference's rule
    with pytest.raises(ValueError):
        pkg.Experiment(periodic=(True, 1, False))


def test_affine_rate_only_along_periodic_axes(cpu):
    w = assembly.config_workload("layer", n=150, templates=3, nodes=128)
    e = w.engine(cpu)
    e.initialize()
    with pytest.raises(ValueError, match="periodic"):
        e.advance(1, rate=(0, 0, -0.5))
    e.advance(2, rate=(-0.5, -0.5, 0))


def _pairs(e):
    n = int(e.read("control", np.uint32, 1)[1])
    return {tuple(sorted(p)) for p in e.read("pairs", np.uint32, 2)[:n].tolist()}


def test_neighbours_equal_brute_force_minimum_image(cpu):
    w = assembly.config_workload("layer", n=150, templates=3, nodes=128)
    e = w.engine(cpu)
    e.initialize()
    x = e.read("position")[:, :3].astype(np.float64) + e.read("positionRemainder")[:, :3]
    box = e.read("cell")[0, :3].astype(np.float64)
    reach = e.read("body")[:, 2].astype(np.float64)
    d = x[None, :, :] - x[:, None, :]
    d[..., :2] -= np.round(d[..., :2] / box[:2]) * box[:2]  # x and y periodic, z open
    r = np.linalg.norm(d, axis=2)
    limit = reach[:, None] + reach[None, :] + e.skin
    near = {(i, j) for i, j in zip(*np.nonzero(r < limit - 1e-6)) if i < j}
    far = {(i, j) for i, j in zip(*np.nonzero(r > limit + 1e-6)) if i < j}
    got = _pairs(e)
    assert near <= got and not (got & far)
    assert any(abs(d[i, j, 0]) < abs(x[j, 0] - x[i, 0]) for i, j in got)  # some pairs across the x seam


def test_translation_along_periodic_axis_changes_nothing(cpu):
    w = assembly.config_workload("layer", n=150, templates=3, nodes=128)
    a = w.engine(cpu)
    a.initialize()
    shifted = w.positions.copy()
    shifted[:, 0] = np.mod(shifted[:, 0] + 0.37 * w.cell[0], w.cell[0])
    from dataclasses import replace

    b = replace(w, positions=shifted).engine(cpu)
    b.initialize()
    fa, fb = a.read("force")[:, :3], b.read("force")[:, :3]
    np.testing.assert_allclose(fb, fa, rtol=1e-5, atol=1e-6 * np.abs(fa).max())
    assert a.state()["contacts"] == b.state()["contacts"]


def test_open_axis_escape_is_rejected(pkg, cpu):
    w = assembly.config_workload("layer", n=150, templates=3, nodes=128)
    e = w.engine(cpu)
    e.initialize()
    v = np.zeros((150, 3))
    v[0, 2] = 1e9  # leaves through the open z axis in one step
    e.set_kinematics(velocity=v)
    with pytest.raises(pkg.RejectedStep) as info:
        e.advance(1)
    assert info.value.flags & 2048
