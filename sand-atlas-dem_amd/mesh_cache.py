"""Content-addressed, verified cache of packed template libraries (NPZ).

Restates the reference's mesh cache (`src/sand_atlas_dem/mesh_cache.py:1-157`):
the key hashes what the packer actually reads (every template's id, provenance,
vertices, faces and nodes, the length unit) together with the packer's own source
and the numerical runtime; an entry holds the 14 packed arrays plus a JSON manifest
of that identity and of each array's hash, and is verified on every load.  Entries
are written to a temporary file and renamed, so readers only see complete ones; a
damaged or mismatching entry is rebuilt with a warning, and an unwritable cache
still returns the freshly packed arrays.

The entry format is the reference's (`sand-dem-mesh-cache/1`: same array names,
dtypes, shapes and manifest layout).  The directory comes from `directory=` or the
environment variable SAND_DEM_MESH_CACHE ("off" disables); with neither, caching is
off -- unlike the reference, nothing is written to a default per-user location.
"""

import hashlib
import json
import os
import platform
import tempfile
import warnings
import zipfile
from pathlib import Path

import numpy as np
import scipy

from . import geometry

FORMAT = "sand-dem-mesh-cache/1"
UINT = frozenset(geometry._UINT_ARRAYS)
FLOAT = frozenset(("meshClusterMin", "meshClusterMax", "meshPlanes", "meshMin", "meshMax", "meshTriangles",
                   "meshNormals"))
_HERE = Path(__file__).resolve().parent
PACKER_SOURCES = (_HERE / "geometry.py", _HERE / "csrc" / "sad_pack.cpp")


def array_hash(array):
    """SHA-256 over the dtype, the shape and the C-order bytes."""
    a = np.ascontiguousarray(array)
    h = hashlib.sha256()
    h.update(json.dumps([a.dtype.str, list(a.shape)]).encode())
    h.update(memoryview(a).cast("B"))
    return h.hexdigest()


def identity(shapes, length_unit):
    """(key, record) of what the packer reads; filenames and claimed hashes are never
    trusted.  The record is returned as it reads back from JSON, so tuple and list
    provenance compare equal on a hit."""
    if not np.isfinite(length_unit) or length_unit <= 0:
        raise ValueError("Mesh length unit must be finite and positive")
    entries = []
    for shape in shapes:
        if shape.provenance.get("display_mesh_only"):
            raise ValueError("Mesh query requires physical source triangles")
        entries.append(dict(
            id=shape.id,
            provenance=shape.provenance,
            arrays={name: array_hash(getattr(shape, name)) for name in ("vertices", "faces", "samples")},
        ))
    packer = hashlib.sha256()
    for source in PACKER_SOURCES:
        packer.update(source.read_bytes())
    record = dict(
        format=FORMAT,
        length_unit=float(length_unit).hex(),
        shapes=entries,
        packer_sha256=packer.hexdigest(),
        numpy=np.__version__,
        scipy=scipy.__version__,
        machine=platform.machine(),
    )
    encoded = json.dumps(record, sort_keys=True, allow_nan=False)
    return hashlib.sha256(encoded.encode()).hexdigest(), json.loads(encoded)


def cache_directory():
    value = os.environ.get("SAND_DEM_MESH_CACHE")
    if not value or value == "off":
        return None
    return Path(value).expanduser()


def _load(path, record):
    with np.load(path, allow_pickle=False) as archive:
        manifest = json.loads(str(archive["manifest"]))
        if not isinstance(manifest, dict) or manifest.get("identity") != record:
            raise ValueError("Mesh cache identity mismatch")
        if set(archive.files) != UINT | FLOAT | {"manifest"}:
            raise ValueError("Mesh cache array set mismatch")
        arrays = {name: archive[name] for name in UINT | FLOAT}
    for name, a in arrays.items():
        dtype = np.uint32 if name in UINT else np.float32
        shape_ok = a.ndim == 1 if name == "meshSampleIndices" else (a.ndim == 2 and a.shape[1] == 4)
        if a.dtype != dtype or not shape_ok or array_hash(a) != manifest["arrays"][name]:
            raise ValueError(f"Mesh cache array verification failed: {name}")
    return arrays


def packed_meshes(shapes, length_unit, *, directory=None, threads=None):
    """Packed arrays and a receipt {"status": disabled | hit | miss | rebuilt-invalid |
    uncached, "key", "path"} (`mesh_cache.py:115-157`)."""
    shapes = list(shapes)
    directory = cache_directory() if directory is None else Path(directory)
    if directory is None:
        return geometry.pack_meshes(shapes, length_unit, threads=threads), dict(status="disabled")
    key, record = identity(shapes, length_unit)
    path = directory / (key + ".npz")
    status = "miss"
    try:
        if path.exists():
            return _load(path, record), dict(status="hit", key=key, path=str(path))
    except (OSError, ValueError, KeyError, TypeError, EOFError, zipfile.BadZipFile) as error:
        warnings.warn(f"Invalid mesh cache; rebuilding {path}: {error}", RuntimeWarning)
        status = "rebuilt-invalid"
    arrays = geometry.pack_meshes(shapes, length_unit, threads=threads)
    manifest = json.dumps(dict(identity=record, arrays={n: array_hash(a) for n, a in arrays.items()}),
                          sort_keys=True)
    temporary = None
    try:
        directory.mkdir(parents=True, exist_ok=True, mode=0o700)
        with tempfile.NamedTemporaryFile(dir=directory, suffix=".npz", delete=False) as stream:
            temporary = Path(stream.name)
            np.savez(stream, manifest=np.array(manifest), **arrays)
        os.replace(temporary, path)
        temporary = None
    except OSError as error:
        warnings.warn(f"Mesh cache could not be saved: {error}", RuntimeWarning)
        status = "uncached"
    finally:
        if temporary is not None:
            temporary.unlink(missing_ok=True)
    return arrays, dict(status=status, key=key, path=str(path))
