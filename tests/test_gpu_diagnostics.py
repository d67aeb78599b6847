"""Diagnostics side path on the HIP backend vs the oracle (needs an MI355X).

The evidence rows (diagnosticRows, fabricRows) and every level of the group-tree
reductions must equal the oracle's FP32 values exactly, so the host-side summary
(stress, energies, fabric, plane reactions) is the same float for float.  The
reduction entry point is also driven directly through the C ABI on random rows
(partial groups, three levels, several walls)."""

import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from sand_atlas_dem_amd.hip import HipBackend

    return HipBackend(0)


@pytest.fixture(scope="module")
def oracle_cpu(pkg, oracle_module):
    return oracle_module.make_oracle(pkg, threads=8)


def same(a, b):
    if isinstance(a, str) or isinstance(b, str):
        return a == b
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def assert_rows_equal(a, b):
    """Evidence rows and every reduction level bit for bit, then the summary."""
    for name in ("diagnosticRows", "fabricRows", "wallRows"):
        assert same(a.read(name), b.read(name)), name
    for levels in ("diag_buffers", "fabric_buffers", "wall_diag_buffers"):
        la, lb = getattr(a, levels), getattr(b, levels)
        assert len(la) == len(lb)
        for (size_a, buf_a), (size_b, buf_b) in zip(la, lb):
            assert size_a == size_b
            assert same(a.m.read(buf_a), b.m.read(buf_b)), levels


def assert_summary_equal(x, y):
    assert x.keys() == y.keys()
    for key in x:
        if key == "walls":
            assert len(x[key]) == len(y[key])
            for wx, wy in zip(x[key], y[key]):
                assert wx.keys() == wy.keys()
                for k in wx:
                    assert same(wx[k], wy[k]), ("walls", k)
        else:
            assert same(x[key], y[key]), key


@pytest.mark.parametrize("config,rate_steps", [("dense", 6), ("walls", 6), ("layer", 6)])
def test_diagnostics_match_oracle_exactly(hip, oracle_cpu, config, rate_steps):
    from sand_atlas_dem_amd import assembly

    print(f"diagnostics {config}", flush=True)
    w = assembly.config_workload(config, n=600, templates=6, nodes=256)
    a, b = w.engine(hip), w.engine(oracle_cpu)
    a.initialize()
    b.initialize()
    for _ in range(2):
        x, y = a.diagnostics(), b.diagnostics()
        assert_rows_equal(a, b)
        assert_summary_equal(x, y)
        assert y["active_grain_pairs"] > 0 and y["active_surface_contacts"] > 0
        a.advance(rate_steps, rate=w.rate)
        b.advance(rate_steps, rate=w.rate)
    assert a.diagnostics()["step"] == b.diagnostics()["step"] == 2 * rate_steps


def test_reduce_entry_point_matches_tree(hip, oracle_module):
    """sad_hip_reduce_kernel on random rows: 70,000 records (three levels), a partial
    last group, walls carrying integer counts, and the row-4 maximum."""
    torch = hip.torch
    rng = np.random.default_rng(11)
    lib = hip.lib

    def run(name, records, walls, n, output_count):
        rows = {"reduceRows": 8, "reduceFabric": 2, "reduceWallRows": 3}[name]
        src = torch.from_numpy(np.ascontiguousarray(records, np.float32).reshape(-1).copy()).to(hip.torch_device)
        dst = torch.zeros(walls * output_count * rows * 4, dtype=torch.float32, device=hip.torch_device)
        rc = lib.sad_hip_reduce_kernel(name.encode(), n, walls, output_count, ct.c_void_p(src.data_ptr()),
                                       src.numel() * 4, ct.c_void_p(dst.data_ptr()), dst.numel() * 4, hip.stream)
        assert rc == 0, lib.sad_hip_last_error()
        torch.cuda.synchronize()
        return dst.cpu().numpy().reshape(walls, output_count, rows, 4)

    for name, walls, n in (("reduceRows", 1, 70_000), ("reduceFabric", 1, 4097), ("reduceWallRows", 3, 300)):
        print(f"reduce {name} walls={walls} n={n}", flush=True)
        rows = {"reduceRows": 8, "reduceFabric": 2, "reduceWallRows": 3}[name]
        records = rng.normal(size=(walls, n, rows, 4)).astype(np.float32) * np.float32(1e3)
        if name == "reduceWallRows":
            records[:, :, 0, 3] = rng.integers(0, 1 << 16, size=(walls, n)).astype(np.uint32).view(np.float32)
        level, size = records, n
        while size > 1:
            groups = (size + 255) // 256
            got = run(name, level, walls, size, groups)
            want = oracle_module.reduce_groups(name, level.reshape(-1), walls, size, np.float32)
            assert same(got, want), (name, size)
            level, size = got, groups
    # Sizes the buffers cannot hold are refused before any launch.
    small = torch.zeros(16, dtype=torch.float32, device=hip.torch_device)
    rc = lib.sad_hip_reduce_kernel(b"reduceRows", 1000, 1, 4, ct.c_void_p(small.data_ptr()), 64,
                                   ct.c_void_p(small.data_ptr()), 64, hip.stream)
    assert rc != 0 and b"do not fit" in lib.sad_hip_last_error()
