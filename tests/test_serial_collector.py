"""The reference's verification collector name (`collectMeshContacts`, G12) on the
oracle -- CPU only.  Like the reference CPU backend (cpu.py:80-96), both collector
names run the serial collector, so an engine that encodes the serial collection
(the reference's SerialEngine, tests/test_parallel_mesh.py:36-46) reaches the same
contacts, forces, history and motion as the frontier engine."""

import numpy as np


def serial_engine_class(pkg):
    class SerialEngine(pkg.Engine):
        def _encode_contacts(self, e, b):
            self.dispatch("resetMeshContacts", self.pair_capacity, e, b)
            self.dispatch("collectMeshContacts", 64 * self.pair_capacity, e, b)
            if self.wall_count:
                self.dispatch("collectWallContacts", 64 * self.n * self.wall_count, e, b)

    return SerialEngine


def test_serial_and_frontier_collectors_agree(pkg, oracle_module):
    from sand_atlas_dem_amd import assembly

    cpu = oracle_module.make_oracle(pkg)
    w = assembly.config_workload("concave", n=120, templates=3, nodes=256)
    states = []
    for cls in (serial_engine_class(pkg), pkg.Engine):
        e = cls(w.shapes, w.spec, w.positions, w.cell, w.ids, w.scales, backend=cpu)
        e.set_kinematics(rotation=w.rotations, velocity=w.velocities * e.time_unit / e.length_unit)
        e.initialize()
        e.advance(4)
        states.append(e.state())
    a, b = states
    assert a["contacts"] == b["contacts"] > 100
    for name in ("meta", "slip", "normalArea", "position", "velocity", "rotation", "force", "torque"):
        assert np.array_equal(a[name], b[name]), name
