"""Experiment contract for the DEM timestep (SI units).

Restates the fields and validation of the reference `Experiment`
(`src/sand_atlas_dem/spec.py:12-166`) that the timestep reads through
`Engine` (`gpu.py:95-368, 490-527`).  Field names, defaults and error behaviour are
kept, so a reference case JSON loads unchanged; stage/servo fields are accepted and
validated but driving stages is the run layer's job, outside this backend.
"""

from dataclasses import asdict, dataclass, field

import numpy as np


def finite_array(value, shape=None, positive=False):
    """Reference `util.finite_array` (`util.py:49-59`)."""
    a = np.asarray(value, dtype=np.float64)
    if (
        (shape is not None and a.shape != shape)
        or not np.isfinite(a).all()
        or (positive and np.any(a <= 0))
    ):
        raise ValueError(f"Expected finite {'positive ' if positive else ''}array with shape {shape}")
    return a


PLANE_FIELDS = {"normal", "offset", "velocity", "normal_stiffness", "tangential_stiffness", "friction"}
BOX_NORMALS = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])


def box_extents(offsets, indices, domain=None):
    """Gap of an axis-aligned plane box (reference `boundaries.py:62-69`)."""
    values = finite_array(offsets)[indices]
    low, high = values[::2], -values[1::2]
    if np.any(high <= low):
        raise ValueError("Boundary box has a nonpositive gap")
    if domain is not None and (np.any(low < 0) or np.any(high > domain)):
        raise ValueError("Boundary box must lie inside the computational domain")
    return high - low


def periodic_axes(periodic):
    """Per-axis periodicity: a bool (the reference's, every axis) or three bools."""
    if type(periodic) is bool:
        return np.array([periodic] * 3)
    if isinstance(periodic, (list, tuple)) and len(periodic) == 3 and all(type(v) is bool for v in periodic):
        return np.array(periodic)
    raise ValueError("periodic must be an explicit boolean or three booleans (x, y, z)")


def periodic_code(periodic):
    """Config `periodic`: 0 none, 1 every axis (the reference's values), 8 | mask for
    a subset (x = 1, y = 2, z = 4)."""
    axes = periodic_axes(periodic)
    if axes.all():
        return 1
    if not axes.any():
        return 0
    return 8 | int(axes[0]) | int(axes[1]) << 1 | int(axes[2]) << 2


def validate_planes(planes, periodic, boundary_box):
    """Fixed-normal translating penalty planes (reference `boundaries.py:20-59`).

    The reference forbids planes in a periodic domain.  With per-axis periodicity
    (no reference counterpart) a plane may bound the nonperiodic axes: its normal
    must have no component along a periodic axis.
    """
    axes = periodic_axes(periodic)
    if not isinstance(planes, list) or (planes and axes.all()):
        raise ValueError("Planes require a nonperiodic domain and a list")
    for plane in planes:
        if not isinstance(plane, dict) or not PLANE_FIELDS <= set(plane) or set(plane) - PLANE_FIELDS - {"normal_viscosity"}:
            raise ValueError("Every plane requires explicit normal, offset, velocity and contact properties")
        normal = finite_array(plane["normal"], (3,))
        if abs(np.linalg.norm(normal) - 1) > 1e-10:
            raise ValueError("Plane normal must have unit length; automatic normalization is forbidden")
        if np.any(np.asarray(normal)[axes] != 0):
            raise ValueError("A plane normal must not point along a periodic axis")
        finite_array(plane["velocity"], (3,))
        finite_array(plane["offset"], ())
        if finite_array(plane.get("normal_viscosity", 0), ()) < 0:
            raise ValueError("Plane normal viscosity must be nonnegative")
        for name in ("normal_stiffness", "tangential_stiffness", "friction"):
            value = finite_array(plane[name], ())
            if value < 0 or (name != "friction" and value == 0):
                raise ValueError("Plane stiffness must be positive and friction nonnegative")
    if boundary_box is not None:
        if axes.any():
            raise ValueError("boundary_box requires a nonperiodic domain")
        if (
            not isinstance(boundary_box, list)
            or len(boundary_box) != 6
            or any(type(i) is not int or not 0 <= i < len(planes) for i in boundary_box)
            or len(set(boundary_box)) != 6
        ):
            raise ValueError("boundary_box requires six distinct plane indices")
        if not np.array_equal([planes[i]["normal"] for i in boundary_box], BOX_NORMALS):
            raise ValueError(
                "boundary_box planes must be ordered xmin,xmax,ymin,ymax,zmin,zmax with exact inward axis normals")
        box_extents(np.array([p["offset"] for p in planes]), boundary_box)


@dataclass
class Experiment:
    format: str = "sand-dem-experiment/1"
    shape_library: str = ""
    contact_geometry: str = "mesh"
    contact_normal: str = "target-gradient"
    periodic: bool = True
    planes: list = field(default_factory=list)
    boundary_box: list | None = None
    particles: int = 1000
    shape_sampling: str = "random"
    seed: int = 42
    density: float = 2650.0
    normal_stiffness: float = 1e12  # traction / penetration, Pa/m
    tangential_stiffness: float = 5e11
    friction: float = 0.5
    normal_viscosity: float = 0.0  # N s / m^3
    timestep_policy: str = "conservative"
    timestep: float = 1e-7
    initial_solid_fraction: float = 0.15
    scale: float = 1.0
    cell: list | None = None
    skin: float = 0.05  # fraction of the maximum particle diameter
    working_memory_gib: float = 12.0
    contacts_per_particle: int = 256
    pairs_per_particle: int = 32
    bin_capacity: int = 128
    stages: list = field(default_factory=lambda: [{"kind": "hold", "steps": 1000, "drag": 0.0}])
    diagnostic_interval: int = 50
    snapshot_interval: int = 1000
    max_overlap_fraction: float = 0.01

    def __post_init__(self):
        if self.format != "sand-dem-experiment/1":
            raise ValueError("Unsupported experiment format")
        if self.contact_geometry not in {"sdf", "mesh"}:
            raise ValueError("Unsupported contact geometry")
        if self.contact_normal not in {"target-gradient", "source-surface-experimental"}:
            raise ValueError("Unsupported contact normal")
        if self.timestep_policy not in {"conservative", "active-contact"}:
            raise ValueError("Unsupported timestep policy")
        if self.shape_sampling not in {"random", "balanced"}:
            raise ValueError("shape_sampling must be random or balanced")
        periodic_axes(self.periodic)
        if isinstance(self.periodic, list):
            self.periodic = tuple(self.periodic)
        validate_planes(self.planes, self.periodic, self.boundary_box)
        for name in (
            "particles", "seed", "contacts_per_particle", "pairs_per_particle",
            "bin_capacity", "diagnostic_interval", "snapshot_interval",
        ):
            value = getattr(self, name)
            if isinstance(value, bool) or not isinstance(value, int) or value < (0 if name == "seed" else 1):
                raise ValueError(f"Invalid integer {name}")
        for name in (
            "density", "normal_stiffness", "tangential_stiffness", "timestep", "scale",
            "skin", "working_memory_gib", "max_overlap_fraction",
        ):
            value = getattr(self, name)
            if not np.isfinite(value) or value <= 0:
                raise ValueError(f"Invalid positive {name}")
        if (
            not np.isfinite(self.friction)
            or self.friction < 0
            or not np.isfinite(self.normal_viscosity)
            or self.normal_viscosity < 0
            or not 0 < self.initial_solid_fraction < 0.5
        ):
            raise ValueError("Invalid friction or initial solid fraction")
        if self.cell is not None:
            finite_array(self.cell, (3,), positive=True)

    def to_dict(self):
        return asdict(self)

    @property
    def viscous(self):
        return self.normal_viscosity > 0 or any(p.get("normal_viscosity", 0) > 0 for p in self.planes)

    def require_supported_path(self):
        """The `hip` backend implements the surface-mesh and SDF-grid steps (with or
        without planes) with target-gradient contact normals."""
        if self.contact_normal != "target-gradient":
            raise NotImplementedError("The hip backend implements target-gradient contact normals only")
