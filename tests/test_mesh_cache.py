"""Mesh cache correctness (restating the reference's tests/test_mesh_cache.py:21-121)
-- CPU only: exact hits that skip packing, keys that change with every packer
input, damaged entries rebuilt exactly, display meshes refused, explicit disable,
JSON-equivalent provenance, and an unwritable directory."""

import tempfile
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture
def shapes(pkg):
    return [pkg.geometry.box_shape([1e-3, 1e-3, 1e-3], nodes=65, seed=1)]


@pytest.fixture
def directory():
    # A private scratch directory inside the checkout, removed afterwards.
    with tempfile.TemporaryDirectory(dir=ROOT, prefix=".mesh-cache-test-") as path:
        yield Path(path)


@pytest.fixture
def cache(pkg):
    from sand_atlas_dem_amd import mesh_cache

    return mesh_cache


def test_hit_is_exact_and_skips_packing(pkg, cache, shapes, directory, monkeypatch):
    expected = pkg.geometry.pack_meshes(shapes, 0.001)
    cold, first = cache.packed_meshes(shapes, 0.001, directory=directory)
    monkeypatch.setattr(cache.geometry, "pack_meshes", lambda *a, **k: pytest.fail("Repacked on hit"))
    warm, second = cache.packed_meshes(shapes, 0.001, directory=directory)
    assert first["status"] == "miss" and second["status"] == "hit"
    assert set(cold) == set(warm) == set(expected)
    for name in expected:
        assert expected[name].dtype == warm[name].dtype and expected[name].shape == warm[name].shape
        assert expected[name].tobytes() == warm[name].tobytes() == cold[name].tobytes()


@pytest.mark.parametrize("change", ["vertices", "faces", "samples", "units", "provenance", "order", "packer"])
def test_key_invalidates_inputs(cache, shapes, change, monkeypatch):
    before = cache.identity(shapes, 0.001)[0]
    shape, unit = shapes[0], 0.001
    if change in {"vertices", "samples", "faces"}:
        array = getattr(shape, change).copy()
        if change == "faces":
            array = array[::-1].copy()
        else:
            array[0, 0] += 1e-9
        shape = replace(shape, **{change: array})
    elif change == "units":
        unit *= 2
    elif change == "provenance":
        shape = replace(shape, provenance={"source_sha256": "different-source"})
    elif change == "packer":
        monkeypatch.setattr(cache.np, "__version__", "different-packing-runtime")
    changed = [shape] if change != "order" else [shape, replace(shape, id="second")]
    assert cache.identity(changed, unit)[0] != before


@pytest.mark.parametrize("damage", ["truncated", "array", "manifest", "manifest-type", "missing"])
def test_damaged_entry_rebuilt_exactly(cache, shapes, directory, damage):
    expected, receipt = cache.packed_meshes(shapes, 0.001, directory=directory)
    path = directory / (receipt["key"] + ".npz")
    if damage == "truncated":
        path.write_bytes(path.read_bytes()[:100])
    else:
        with np.load(path, allow_pickle=False) as data:
            arrays = {n: data[n] for n in data.files}
        if damage == "array":
            arrays["meshClusterMin"][0, 0] += 1
        elif damage == "manifest":
            arrays["manifest"] = np.array("{}")
        elif damage == "manifest-type":
            arrays["manifest"] = np.array("null")
        else:
            del arrays["meshFrontier"]
        np.savez(path, **arrays)
    with pytest.warns(RuntimeWarning, match="Invalid mesh cache"):
        actual, receipt = cache.packed_meshes(shapes, 0.001, directory=directory)
    assert receipt["status"] == "rebuilt-invalid"
    for name in expected:
        assert actual[name].tobytes() == expected[name].tobytes()


def test_display_mesh_never_accepted(cache, shapes, directory):
    cache.packed_meshes(shapes, 0.001, directory=directory)
    shapes[0].provenance["display_mesh_only"] = True
    with pytest.raises(ValueError, match="physical source"):
        cache.packed_meshes(shapes, 0.001, directory=directory)
    with pytest.raises(ValueError, match="physical source"):
        cache.packed_meshes(shapes, 0.001)


def test_explicit_disable(cache, shapes, monkeypatch):
    monkeypatch.setenv("SAND_DEM_MESH_CACHE", "off")
    _, receipt = cache.packed_meshes(shapes, 0.001)
    assert receipt == {"status": "disabled"}
    monkeypatch.delenv("SAND_DEM_MESH_CACHE")
    _, receipt = cache.packed_meshes(shapes, 0.001)
    assert receipt == {"status": "disabled"}


def test_environment_directory_and_engine_receipt(pkg, cache, shapes, directory, monkeypatch, oracle_module):
    monkeypatch.setenv("SAND_DEM_MESH_CACHE", str(directory))
    spec = pkg.Experiment(particles=1, timestep=1e-9, max_overlap_fraction=0.1)
    cpu = oracle_module.make_oracle(pkg)
    first = pkg.Engine(shapes, spec, [[0.005] * 3], [0.01] * 3, [0], [1], backend=cpu)
    second = pkg.Engine(shapes, spec, [[0.005] * 3], [0.01] * 3, [0], [1], backend=cpu)
    assert first.mesh_cache_receipt["status"] == "miss"
    assert second.mesh_cache_receipt["status"] == "hit"
    for name in ("meshLinks", "meshTriangles", "meshFrontier"):
        assert np.array_equal(first.read(name, np.uint8, 1), second.read(name, np.uint8, 1))


def test_json_equivalent_provenance_can_hit(cache, shapes, directory):
    shapes[0].provenance["source_units"] = ("mm", 1)
    cache.packed_meshes(shapes, 0.001, directory=directory)
    _, receipt = cache.packed_meshes(shapes, 0.001, directory=directory)
    assert receipt["status"] == "hit"


def test_failed_write_uses_exact_fresh_pack(pkg, cache, shapes, directory):
    path = directory / "file-not-directory"
    path.write_text("occupied")
    expected = pkg.geometry.pack_meshes(shapes, 0.001)
    with pytest.warns(RuntimeWarning, match="could not be saved"):
        actual, receipt = cache.packed_meshes(shapes, 0.001, directory=path)
    assert receipt["status"] == "uncached"
    assert all(actual[n].tobytes() == expected[n].tobytes() for n in expected)
