"""Allocation-free stepping (sad_hip_advance_static) against the allocating entry
point (sad_hip_advance) through the C ABI (needs an MI355X).

Two calls in a row on one GPU, the second with a smaller assembly, must give states
identical to sad_hip_advance's; an assembly whose workspace exceeds the static array
is refused with a clear error before any launch.
"""

import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ("position", "positionRemainder", "velocity", "velocityRemainder", "rotation", "rotationRemainder",
         "momentum", "momentumRemainder", "force", "torque", "accumulated", "cell", "slip", "normalArea")


def static_backend():
    from sand_atlas_dem_amd.engine import config_values
    from sand_atlas_dem_amd.hip import SIDE_ORDER, HipBackend, _table

    class StaticStepping(HipBackend):
        """HipBackend whose whole-step encoder calls sad_hip_advance_static."""

        def encode_steps(self, encoder, steps, active, bindings, history_buffers):
            def execute():
                ptrs, sizes = _table(bindings)
                values = config_values(bindings["c"])
                sides = [history_buffers[s][name] for s in (0, 1) for name in SIDE_ORDER]
                sp = (ct.c_void_p * 8)(*(b.ptr for b in sides))
                ss = (ct.c_uint64 * 8)(*(b.size for b in sides))
                self._check(self.lib.sad_hip_advance_static(
                    steps, active, ptrs, sizes, values.ctypes.data_as(ct.POINTER(ct.c_double)), sp, ss,
                    self.stream), "advance_static")

            encoder.operations.append(execute)

    return StaticStepping(0)


def test_allocation_free_steps_equal_the_allocating_entry_point(pkg):
    from sand_atlas_dem_amd import assembly
    from sand_atlas_dem_amd.hip import HipBackend

    plain, static = HipBackend(0), static_backend()
    for n in (200, 120):  # the second call is smaller: nothing from the first may leak
        w = assembly.config_workload("dense", n=n, templates=4, nodes=256)
        a, b = w.engine(plain), w.engine(static)
        a.initialize()
        b.initialize()
        for _ in range(2):
            a.advance(3, rate=w.rate)
            b.advance(3, rate=w.rate)
            sa, sb = a.state(), b.state()
            assert sa["contacts"] == sb["contacts"] > 1000
            assert np.array_equal(sa["meta"], sb["meta"])
            for name in STATE:
                assert np.array_equal(sa[name], sb[name]), (n, name)
            ca, cb = a.read("control", np.uint32, 1), b.read("control", np.uint32, 1)
            assert np.array_equal(ca[[0, 1, 2, 4, 6]], cb[[0, 1, 2, 4, 6]])


def test_allocation_free_steps_refuse_an_oversized_workspace(pkg):
    from sand_atlas_dem_amd import assembly

    w = assembly.config_workload("dense", n=400, templates=4, nodes=256)  # 204,800 contact slots
    e = w.engine(static_backend())
    e.initialize()
    before = e.state()
    with pytest.raises(RuntimeError, match="allocation-free stepping needs"):
        e.advance(1, rate=w.rate)
    after = e.state()
    assert np.array_equal(before["position"], after["position"])  # rolled back, nothing stepped
