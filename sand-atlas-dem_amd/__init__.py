"""MI355X-native `hip` backend for the DEM timestep of sand-atlas-dem.

Import name: `sand_atlas_dem_amd` (the directory is `sand-atlas-dem_amd/`; the repo
root's `load_package()` helpers register it under that name).
"""

from . import assembly, autotune, engine, geometry, learned, native, spec  # noqa: F401
from .engine import Engine, RejectedStep  # noqa: F401
from .spec import Experiment  # noqa: F401

__version__ = "0.1.0"


def backend(name):
    """Explicit backend selection (reference `simulation.py:60-68`).

    Only "hip" lives in this package.  "metal" and "cpu" are the reference's own
    backends; there is no implicit choice and no fallback.
    """
    if name == "hip":
        from .hip import HipBackend

        return HipBackend()
    if name in {"metal", "cpu"}:
        raise ValueError(f"backend {name!r} is provided by the reference package, not by sand_atlas_dem_amd")
    raise ValueError(f"Unknown backend {name!r}; this package provides 'hip'")
