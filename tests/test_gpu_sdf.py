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
