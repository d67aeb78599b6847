"""The verified store of packed template libraries -- CPU only.

Contract (the reference's cache behaviour, `tests/test_mesh_cache.py`, checked here
with tests of this package's own design): an entry read back equals a fresh pack
byte for byte and skips packing; the key moves with every packer input and only
with those; every kind of damage to an entry is detected and repacked; display-only
meshes are refused; the store is off unless asked for; and a store that cannot be
written still returns the fresh pack.
"""

import tempfile
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
UNIT = 1e-3


@pytest.fixture
def mc(pkg):
    from sand_atlas_dem_amd import mesh_cache

    return mesh_cache


@pytest.fixture
def library(pkg):
    g = pkg.geometry
    return [g.box_shape([1e-3, 1e-3, 1e-3], nodes=65, seed=1),
            g.make_shape("prism", *g.concave_prism_mesh(), nodes=40, seed=2)]


@pytest.fixture
def root():
    with tempfile.TemporaryDirectory(dir=ROOT, prefix=".mesh-store-test-") as path:
        yield Path(path)


def same_arrays(a, b):
    return set(a) == set(b) and all(a[n].dtype == b[n].dtype and a[n].tobytes() == b[n].tobytes() for n in a)


def test_stored_entry_reads_back_without_packing(pkg, mc, library, root, monkeypatch):
    fresh = pkg.geometry.pack_meshes(library, UNIT)
    calls = []
    real = mc.geometry.pack_meshes
    monkeypatch.setattr(mc.geometry, "pack_meshes", lambda *a, **k: calls.append(1) or real(*a, **k))
    first, r1 = mc.packed_meshes(library, UNIT, directory=root)
    second, r2 = mc.packed_meshes(library, UNIT, directory=root)
    assert (r1["status"], r2["status"], len(calls)) == ("miss", "hit", 1)
    assert r1["key"] == r2["key"] and Path(r2["path"]).exists()
    assert same_arrays(first, fresh) and same_arrays(second, fresh)
    assert set(fresh) == mc.ARRAYS


def variants(library):
    """Inputs that must each get a key of their own."""
    box, prism = library
    nudged = box.vertices.copy()
    nudged[3, 1] += 1e-12
    resampled = box.samples.copy()
    resampled[0, 2] -= 1e-12
    return {
        "base": (library, UNIT),
        "vertex moved": ([replace(box, vertices=nudged), prism], UNIT),
        "faces reordered": ([replace(box, faces=box.faces[::-1].copy()), prism], UNIT),
        "node moved": ([replace(box, samples=resampled), prism], UNIT),
        "length unit": (library, 2 * UNIT),
        "provenance": ([replace(box, provenance={"source": "elsewhere"}), prism], UNIT),
        "template order": ([prism, box], UNIT),
        "template id": ([replace(box, id="renamed"), prism], UNIT),
        "one template fewer": ([box], UNIT),
    }


def test_key_moves_with_every_packer_input(mc, library, monkeypatch):
    keys = {name: mc.Key.of(shapes, unit).hex for name, (shapes, unit) in variants(library).items()}
    assert len(set(keys.values())) == len(keys), keys
    base = keys["base"]
    monkeypatch.setattr(mc.np, "__version__", "0.0-other")  # the numerical runtime is part of the key
    assert mc.Key.of(library, UNIT).hex != base
    monkeypatch.undo()
    assert mc.Key.of(library, UNIT).hex == base  # and nothing else (deterministic)


def test_json_equivalent_provenance_hits(mc, library, root):
    library[0].provenance["units"] = ("mm", 1)  # reads back from JSON as a list
    mc.packed_meshes(library, UNIT, directory=root)
    assert mc.packed_meshes(library, UNIT, directory=root)[1]["status"] == "hit"


def _rewrite(path, edit):
    with np.load(path, allow_pickle=False) as npz:
        content = {name: npz[name] for name in npz.files}
    edit(content)
    np.savez(path, **content)


def _flip_array_bit(c):
    a = c["meshTriangles"]
    a.view(np.uint32)[1, 2] ^= 1


DAMAGE = {
    "truncated file": lambda p: p.write_bytes(p.read_bytes()[:512]),
    "not a zip": lambda p: p.write_text("garbage"),
    "one bit of an array": lambda p: _rewrite(p, _flip_array_bit),
    "array dropped": lambda p: _rewrite(p, lambda c: c.pop("meshLinks")),
    "array added": lambda p: _rewrite(p, lambda c: c.update(extra=np.zeros(4, np.uint32))),
    "array retyped": lambda p: _rewrite(p, lambda c: c.update(meshMin=c["meshMin"].astype(np.float64))),
    "manifest dropped": lambda p: _rewrite(p, lambda c: c.pop("manifest")),
    "manifest not an object": lambda p: _rewrite(p, lambda c: c.update(manifest=np.array("[1, 2]"))),
    "manifest of other inputs": lambda p: _rewrite(p, lambda c: c.update(
        manifest=np.array(str(c["manifest"]).replace('"length_unit": "', '"length_unit": "-')))),
}


@pytest.mark.parametrize("damage", sorted(DAMAGE))
def test_damaged_entry_is_detected_and_repacked(mc, library, root, damage):
    fresh, receipt = mc.packed_meshes(library, UNIT, directory=root)
    DAMAGE[damage](Path(receipt["path"]))
    with pytest.warns(RuntimeWarning, match="Invalid mesh cache entry"):
        arrays, again = mc.packed_meshes(library, UNIT, directory=root)
    assert again["status"] == "rebuilt-invalid" and same_arrays(arrays, fresh)
    assert mc.packed_meshes(library, UNIT, directory=root)[1]["status"] == "hit"  # the repack was stored


def test_display_only_meshes_are_refused_with_and_without_a_store(mc, library, root):
    library[1].provenance["display_mesh_only"] = True
    for directory in (root, None):
        with pytest.raises(ValueError, match="physical source triangles"):
            mc.packed_meshes(library, UNIT, directory=directory)


def test_store_is_off_unless_asked_for(mc, library, monkeypatch):
    for value in (None, "off", ""):
        if value is None:
            monkeypatch.delenv(mc.ENVIRONMENT, raising=False)
        else:
            monkeypatch.setenv(mc.ENVIRONMENT, value)
        assert mc.packed_meshes(library, UNIT)[1] == {"status": "disabled"}


def test_unwritable_store_returns_the_fresh_pack(pkg, mc, library, root):
    blocked = root / "a-file"
    blocked.write_text("not a directory")
    with pytest.warns(RuntimeWarning, match="could not be saved"):
        arrays, receipt = mc.packed_meshes(library, UNIT, directory=blocked)
    assert receipt["status"] == "uncached"
    assert same_arrays(arrays, pkg.geometry.pack_meshes(library, UNIT))
    assert not list(root.glob("*.part"))  # no partial files left behind


def test_engine_reads_the_environment_store(pkg, mc, library, root, monkeypatch, oracle_module):
    monkeypatch.setenv(mc.ENVIRONMENT, str(root))
    cpu = oracle_module.make_oracle(pkg)
    spec = pkg.Experiment(particles=2, timestep=1e-9, max_overlap_fraction=0.1)
    engines = [pkg.Engine(library, spec, [[0.004] * 3, [0.011] * 3], [0.016] * 3, [0, 1], [1, 1], backend=cpu)
               for _ in range(2)]
    assert [e.mesh_cache_receipt["status"] for e in engines] == ["miss", "hit"]
    for name in ("meshLinks", "meshTriangles", "meshPlanes", "meshFrontier"):
        assert np.array_equal(engines[0].read(name, np.uint8, 1), engines[1].read(name, np.uint8, 1))
