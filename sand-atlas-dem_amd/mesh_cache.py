"""Verified on-disk store of packed template libraries.

Packing a library (hull planes, pseudonormals, the triangle BVH and the sample tree
of every template) costs seconds to minutes for hundreds of templates, and its
result depends only on the templates, the length unit and the packer itself.  The
store keeps each packed library as one NPZ file named by a SHA-256 key of exactly
those inputs, and re-verifies every array against its recorded digest whenever it
reads an entry back.  A damaged or foreign entry is repacked (with a
RuntimeWarning); a directory that cannot be written still yields freshly packed
arrays.  Files are written under a temporary name and renamed into place, so a
reader never sees a partial entry.

Entry format (kept identical to the reference's cache, `mesh_cache.py:17-75`, so an
entry means the same thing to both): the format tag `sand-dem-mesh-cache/1`; the 14
packed arrays under their buffer names (uint32 or float32, rows of 4 except the
1-D `meshSampleIndices`); and a JSON `manifest` {"identity": <key record>,
"arrays": {name: digest}}, where a digest is SHA-256 over [dtype string, shape] as
JSON followed by the C-order bytes.  The key record names this package's packer
sources rather than the reference's `mesh.py`.

The directory is `directory=` or the environment variable SAND_DEM_MESH_CACHE
("off" or unset: no store; nothing is written to a default location).
"""

import hashlib
import json
import os
import platform
import tempfile
import warnings
import zipfile
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import scipy

from . import geometry

FORMAT = "sand-dem-mesh-cache/1"
UINT = frozenset(geometry._UINT_ARRAYS)
FLOAT = frozenset(("meshClusterMin", "meshClusterMax", "meshPlanes", "meshMin", "meshMax", "meshTriangles",
                   "meshNormals"))
ARRAYS = UINT | FLOAT
_HERE = Path(__file__).resolve().parent
PACKER_SOURCES = (_HERE / "geometry.py", _HERE / "csrc" / "sad_pack.cpp")
ENVIRONMENT = "SAND_DEM_MESH_CACHE"


def digest(array):
    """SHA-256 of an array's dtype, shape and C-order bytes."""
    a = np.ascontiguousarray(array)
    h = hashlib.sha256(json.dumps([a.dtype.str, list(a.shape)]).encode())
    h.update(memoryview(a).cast("B"))
    return h.hexdigest()


class EntryError(ValueError):
    """A stored entry that cannot be trusted (unreadable, foreign or damaged)."""


@dataclass(frozen=True)
class Key:
    """What the packer reads, as a JSON record, and its SHA-256 (the file name)."""

    hex: str
    record: dict

    @classmethod
    def of(cls, shapes, length_unit):
        if not (np.isfinite(length_unit) and length_unit > 0):
            raise ValueError("Mesh length unit must be finite and positive")
        templates = []
        for shape in shapes:
            if shape.provenance.get("display_mesh_only"):
                raise ValueError("Mesh query requires physical source triangles")
            templates.append({
                "id": shape.id,
                "provenance": shape.provenance,
                "arrays": {field: digest(getattr(shape, field)) for field in ("vertices", "faces", "samples")},
            })
        sources = hashlib.sha256(b"".join(path.read_bytes() for path in PACKER_SOURCES)).hexdigest()
        record = {
            "format": FORMAT,
            "length_unit": float(length_unit).hex(),
            "shapes": templates,
            "packer_sha256": sources,
            "numpy": np.__version__,
            "scipy": scipy.__version__,
            "machine": platform.machine(),
        }
        # The record is compared after a JSON round trip: tuples in a provenance
        # read back as lists, and must still match.
        text = json.dumps(record, sort_keys=True, allow_nan=False)
        return cls(hashlib.sha256(text.encode()).hexdigest(), json.loads(text))


def _checked(name, array, expected_digest):
    want = np.uint32 if name in UINT else np.float32
    rows_ok = array.ndim == 1 if name == "meshSampleIndices" else (array.ndim == 2 and array.shape[1] == 4)
    if array.dtype != want or not rows_ok or digest(array) != expected_digest:
        raise EntryError(f"stored array {name} does not match its digest, dtype or layout")
    return array


@dataclass(frozen=True)
class Store:
    root: Path

    def path(self, key):
        return self.root / f"{key.hex}.npz"

    def read(self, key):
        """The verified arrays of `key`, None if absent; EntryError if untrustworthy."""
        path = self.path(key)
        if not path.exists():
            return None
        try:
            with np.load(path, allow_pickle=False) as npz:
                names = set(npz.files)
                manifest = json.loads(str(npz["manifest"])) if "manifest" in names else None
                if not isinstance(manifest, dict) or manifest.get("identity") != key.record:
                    raise EntryError("manifest missing or written for other inputs")
                if names != ARRAYS | {"manifest"}:
                    raise EntryError("stored array set differs from the packed library's")
                digests = manifest.get("arrays")
                if not isinstance(digests, dict):
                    raise EntryError("manifest holds no array digests")
                return {name: _checked(name, npz[name], digests.get(name)) for name in sorted(ARRAYS)}
        except EntryError:
            raise
        except (OSError, ValueError, KeyError, TypeError, EOFError, zipfile.BadZipFile) as error:
            raise EntryError(f"unreadable entry: {error}") from error

    def write(self, key, arrays):
        manifest = json.dumps({"identity": key.record, "arrays": {n: digest(a) for n, a in arrays.items()}},
                              sort_keys=True)
        self.root.mkdir(parents=True, exist_ok=True, mode=0o700)
        fd, name = tempfile.mkstemp(dir=self.root, suffix=".npz.part")
        try:
            with os.fdopen(fd, "wb") as stream:
                np.savez(stream, manifest=np.array(manifest), **arrays)
            os.replace(name, self.path(key))
        except BaseException:
            Path(name).unlink(missing_ok=True)
            raise


def environment_store():
    value = os.environ.get(ENVIRONMENT, "")
    return None if value in ("", "off") else Store(Path(value).expanduser())


def packed_meshes(shapes, length_unit, *, directory=None, threads=None):
    """Packed arrays of `shapes` and a receipt {"status", "key", "path"}; status is
    "disabled" (no store), "hit", "miss", "rebuilt-invalid" (a damaged entry was
    replaced) or "uncached" (packed, but the store could not be written)."""
    shapes = list(shapes)
    store = environment_store() if directory is None else Store(Path(directory))
    if store is None:
        Key.of(shapes, length_unit)  # the same input checks as with a store
        return geometry.pack_meshes(shapes, length_unit, threads=threads), {"status": "disabled"}
    key = Key.of(shapes, length_unit)
    receipt = {"status": "miss", "key": key.hex, "path": str(store.path(key))}
    try:
        arrays = store.read(key)
        if arrays is not None:
            return arrays, dict(receipt, status="hit")
    except EntryError as error:
        warnings.warn(f"Invalid mesh cache entry {receipt['path']} repacked: {error}", RuntimeWarning)
        receipt["status"] = "rebuilt-invalid"
    arrays = geometry.pack_meshes(shapes, length_unit, threads=threads)
    try:
        store.write(key, arrays)
    except OSError as error:
        warnings.warn(f"Mesh cache entry could not be saved under {store.root}: {error}", RuntimeWarning)
        receipt["status"] = "uncached"
    return arrays, receipt
