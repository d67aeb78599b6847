"""Host-side contracts: packed layout, mass properties, C ABI exports, engine rules.

CPU only; the HIP library is loaded but never called for compute here.
"""

import ctypes as ct
import re
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def box(pkg):
    return pkg.geometry.box_shape([1e-3, 1e-3, 1e-3], nodes=32)


def test_box_mass_properties_closed_form(pkg):
    b = pkg.geometry.box_shape([1e-3, 2e-3, 3e-3])
    assert np.isclose(b.volume, 6e-9)
    expected = b.volume / 12 * np.array([2e-3**2 + 3e-3**2, 1e-3**2 + 3e-3**2, 1e-3**2 + 2e-3**2])
    np.testing.assert_allclose(np.diag(b.inertia), expected, rtol=1e-12)
    np.testing.assert_allclose(b.inertia - np.diag(np.diag(b.inertia)), 0, atol=1e-25)


def test_icosphere_converges_to_sphere(pkg):
    for s, tol in ((3, 1e-2), (4, 3e-3)):
        v, f = pkg.geometry.icosphere(s)
        volume, centroid, inertia = pkg.geometry.mass_properties(v, f)
        assert abs(volume / (4 / 3 * np.pi) - 1) < tol
        assert np.abs(centroid).max() < 1e-12
        assert np.allclose(np.diag(inertia) / volume, 0.4, rtol=tol)


def test_hull_sphere_triangle_counts(pkg):
    for t in (320, 1000, 5120):
        v, f = pkg.geometry.hull_sphere(t)
        assert len(f) == 2 * len(v) - 4 and abs(len(f) - t) <= 2
        volume, _, _ = pkg.geometry.mass_properties(v, f)
        assert volume > 0


def test_sobol_nodes_lie_on_surface_with_equal_areas(pkg, box):
    assert np.isclose(box.areas.sum(), 6e-6)
    assert np.allclose(box.areas, box.areas[0])
    assert np.all(np.isclose(np.abs(box.samples).max(1), 5e-4))


def test_packed_tree_covers_triangles_and_bounds(pkg, box):
    """Structure checks of the reference's tests/test_mesh.py:18-50."""
    data = pkg.geometry.pack_meshes([box, replace(box, id="box-2")], 0.001)
    triangles = data["meshTriangles"][:, :3].reshape(-1, 3, 3)
    links = data["meshLinks"]
    for root, end, _, _ in data["meshRoots"]:
        visits = []
        node = int(root)
        while node < end:
            escape, start, count, _ = links[node]
            assert node < escape <= end
            if count:
                leaf = triangles[start : start + count]
                assert np.all(leaf >= data["meshMin"][node, :3]) and np.all(leaf <= data["meshMax"][node, :3])
                visits.extend(range(int(start), int(start + count)))
            node += 1
        assert len(visits) == len(box.faces)
    for root in range(2):
        for start, count, _, _ in data["meshHullInfo"][6 * root : 6 * root + 6]:
            plane = data["meshPlanes"][start : start + count]
            assert np.all((box.vertices / 0.001) @ plane[:, :3].T + plane[:, 3] < 0)
    assert np.allclose(np.linalg.norm(data["meshNormals"][:, :3], axis=1), 1)


def test_open_inconsistent_or_fanned_meshes_are_rejected(pkg, box):
    with pytest.raises(ValueError, match="two faces"):
        pkg.geometry.pack_meshes([replace(box, faces=box.faces[:-1])], 0.001)
    faces = box.faces.copy()
    faces[0] = faces[0, ::-1]
    with pytest.raises(ValueError, match="winding"):
        pkg.geometry.pack_meshes([replace(box, faces=faces)], 0.001)


@pytest.mark.parametrize("count", [1, 32, 64, 65, 129, 1024])
def test_clusters_partition_node_keys_and_enclose_points(pkg, box, count):
    """tests/test_mesh.py:53-105 of the reference."""
    rng = np.random.default_rng(321)
    points = rng.normal(size=(count, 3)) * 0.001
    shape = replace(box, samples=points, areas=np.ones(count))
    data = pkg.geometry.pack_meshes([shape, replace(shape, id="second")], 0.001)
    for first, clusters, _, _ in data["meshClusterInfo"]:
        keys = []
        for cluster in range(first, first + clusters):
            start, size, escape, _ = data["meshClusterRanges"][cluster]
            assert cluster < escape <= first + clusters and 0 <= size <= 64
            if size == 0:
                left = cluster + 1
                right = data["meshClusterRanges"][left, 2]
                assert left < right < escape and data["meshClusterRanges"][right, 2] == escape
                continue
            node_ids = data["meshSampleIndices"][start : start + size]
            keys.extend(node_ids.tolist())
            rounded = (points[node_ids] / 0.001).astype(np.float32)
            assert np.all(rounded >= data["meshClusterMin"][cluster, :3])
            assert np.all(rounded <= data["meshClusterMax"][cluster, :3])
        assert sorted(keys) == list(range(count))


def test_frontier_partitions_the_cluster_tree(pkg):
    v, f = pkg.geometry.icosphere(3)
    s = pkg.geometry.make_shape("s", v * 1e-3, f, nodes=2048, seed=1)
    data = pkg.geometry.pack_meshes([s], 1e-3)
    first, count, _, _ = data["meshClusterInfo"][0]
    ranges = data["meshClusterRanges"]
    covered = []
    for start, end, _, _ in data["meshFrontier"][:32]:
        if end > start:
            covered.extend(c for c in range(int(start), int(end)) if ranges[c, 1] > 0)
    leaves = [c for c in range(int(first), int(first + count)) if ranges[c, 1] > 0]
    assert sorted(covered) == leaves  # disjoint complete subtrees cover every leaf once


def test_c_abi_library_exports_every_declared_symbol(pkg):
    """The HIP library loads without a GPU and exports the header's entry points."""
    header = (ROOT / "include" / "sand_dem_hip.h").read_text()
    declared = set(re.findall(r"\b(sad_hip_\w+)\s*\(", header))
    native = pkg.native
    assert declared == set(native.EXPORTED)
    if not native.LIBRARY.exists():
        pytest.skip("library not built")
    lib = ct.CDLL(str(native.LIBRARY))
    for name in declared:
        assert hasattr(lib, name), name
    lib.sad_hip_abi_version.restype = ct.c_int
    assert lib.sad_hip_abi_version() == 1


def test_packer_library_exports_every_declared_symbol(pkg):
    """The host mesh packer exports the entry points its header declares."""
    header = (ROOT / "include" / "sand_dem_pack.h").read_text()
    declared = set(re.findall(r"\b(sad_(?:pack|sdf)_\w+)\s*\(", header))
    assert declared == set(pkg.native.PACK_EXPORTED)
    lib = ct.CDLL(str(pkg.native.PACK_LIBRARY))
    for name in declared:
        assert hasattr(lib, name), name


def test_abi_tables_match_header(pkg):
    header = (ROOT / "include" / "sand_dem_hip.h").read_text()
    body = header[header.index("enum sad_buffer"):header.index("SAD_BUFFER_COUNT")]
    slots = re.findall(r"SAD_(\w+)", body)
    camel = ["".join(w.capitalize() if k else w.lower() for k, w in enumerate(s.split("_"))) for s in slots]
    fix = {"invI": "invI", "metaa": "meta"}
    names = [fix.get(c, c) for c in camel]
    names = [n.replace("InvI", "invI") for n in names]
    assert names == list(pkg.engine.BUFFER_ORDER)
    assert len(pkg.engine.VALUE_ORDER) + 1 == 26


def test_backend_selection_is_explicit(pkg):
    with pytest.raises(ValueError, match="reference"):
        pkg.backend("cpu")
    with pytest.raises(ValueError, match="Unknown"):
        pkg.backend("cuda")


def test_engine_rejects_unsupported_paths(pkg, oracle_module, box):
    cpu = oracle_module.make_oracle(pkg)
    with pytest.raises(NotImplementedError):
        pkg.Engine([box], pkg.Experiment(contact_normal="source-surface-experimental", friction=0, timestep=1e-9),
                   [[0.005] * 3], [0.01] * 3, [0], [1], backend=cpu)
    with pytest.raises(ValueError, match="no SDF grid"):  # the SDF geometry needs each template's grid
        pkg.Engine([box], pkg.Experiment(contact_geometry="sdf", timestep=1e-9), [[0.005] * 3], [0.01] * 3, [0], [1],
                   backend=cpu)
    with pytest.raises(ValueError, match="Timestep"):
        pkg.Engine([box], pkg.Experiment(timestep=1.0), [[0.005] * 3], [0.01] * 3, [0], [1], backend=cpu)


def test_rejected_step_restores_every_array(pkg, oracle_module):
    """Transactional rollback (gpu.py:729-743; test_cpu_solver.py:279-304)."""
    from sand_atlas_dem_amd import assembly

    cpu = oracle_module.make_oracle(pkg)
    w = assembly.config_workload("dense", n=120, templates=4, nodes=256, seed=4)
    e = w.engine(cpu)
    e.initialize()
    e.advance(2)
    before = e.state()
    with pytest.raises(pkg.RejectedStep):
        e.advance(3, rate=(-1e10, -1e10, -1e10))  # collapsing cell: periodic image ambiguity (256)
    after = e.state()
    for name in before:
        if name != "contacts":
            assert np.array_equal(before[name], after[name]), name
    assert before["contacts"] == after["contacts"]
    e.advance(1)


def test_capacity_growth_replays_to_large_capacity_result(pkg, oracle_module):
    from sand_atlas_dem_amd import assembly

    cpu = oracle_module.make_oracle(pkg)
    w = assembly.config_workload("dense", n=120, templates=4, nodes=256, seed=4)
    big = w.engine(cpu)
    big.initialize()
    big.advance(2)
    small = replace(w, spec=replace(w.spec, contacts_per_particle=512)).engine(cpu)
    small.initialize()
    tight = replace(w, spec=replace(w.spec, contacts_per_particle=4)).engine(cpu)
    with pytest.raises(pkg.RejectedStep):
        tight.initialize()


def test_autotune_coordinate_descent_and_cache(pkg, tmp_path):
    """Tuner logic on a stand-in engine whose step time depends on the knobs."""
    import time
    from types import SimpleNamespace

    from sand_atlas_dem_amd import autotune

    class Backend:
        def __init__(self):
            self.knobs = {"collect_blocks_per_cu": 6, "collect_pop": 1, "collect_queue": 128}
            self.torch = SimpleNamespace(cuda=SimpleNamespace(get_device_name=lambda d: "stand-in"))
            self.torch_device = None

        def tune(self, name, value=-1):
            old = self.knobs[name]
            if value >= 0:
                self.knobs[name] = value
            return old

        def wait(self):
            pass

    class Engine:
        def __init__(self):
            self.m = Backend()
            self.n, self.step = 4, 7
            self.shapes = [SimpleNamespace(faces=np.zeros((8, 3)), samples=np.zeros((16, 3)))]
            self.ids = np.zeros(4, np.uint32)
            self.restores = 0

        def state(self):
            return {"x": 1}

        def restore(self, state, step):
            self.restores += 1
            self.step = step

        def advance(self, k, rate=None):
            best = {"collect_blocks_per_cu": 4, "collect_pop": 2, "collect_queue": 256}
            miss = sum(self.m.knobs[n] != v for n, v in best.items())
            time.sleep(0.002 * (1 + 3 * miss) * k)
            self.step += k

    e = Engine()
    space = {"collect_blocks_per_cu": (2, 4, 6), "collect_pop": (1, 2), "collect_queue": (128, 256)}
    r = autotune.autotune(e, steps=1, space=space, cache=tmp_path / "t.json")
    assert r["best"] == {"collect_blocks_per_cu": 4, "collect_pop": 2, "collect_queue": 256}
    assert e.m.knobs == r["best"] and e.step == 7 and not r["cached"]
    settings = [tuple(sorted(t["setting"].items())) for t in r["trials"]]
    assert len(settings) == len(set(settings)) == 5  # base + 2 + 1 + 1, no repeats
    e2 = Engine()
    r2 = autotune.autotune(e2, steps=1, space=space, cache=tmp_path / "t.json")
    assert r2["cached"] and e2.m.knobs == r["best"] and e2.restores == 0
