"""Host step driver for the DEM timestep (surface-mesh and SDF-grid geometry).

Restates the reference `Engine` (`src/sand_atlas_dem/gpu.py:94-1092`): SI ->
normalized units, capacities, the buffer inventory in the kernels'
`Resources` layout, the per-step encoder sequence, the device error word with
transactional rollback, capacity growth, the double-buffered friction history,
state/restore and the diagnostics summary.  It is adapter-agnostic exactly as the reference is: it talks to a
backend object through the duck-typed adapter protocol of `gpu.py` (buffer,
dispatch, create_command_encoder/copy_buffer, submit, wait, read, read_prefix,
write_prefix).  The product adapter is `HipBackend` (`hip.py`); when an adapter
offers `encode_steps`, whole timesteps are encoded natively instead of kernel by
kernel.
"""

import math
from dataclasses import replace

import numpy as np

from .geometry import _UINT_ARRAYS, sample_arrays
from .mesh_cache import packed_meshes
from .spec import box_extents, periodic_axes, periodic_code, finite_array

ERRORS = {  # gpu.py:11-26
    2: "degenerate contact distance gradient",
    4: "excessive penetration",
    8: "bin capacity",
    16: "pair capacity",
    32: "contact capacity",
    64: "history hash full",
    128: "contact-list cycle",
    256: "periodic image ambiguity",
    512: "nonfinite state",
    1024: "mesh traversal stack capacity",
    2048: "nonperiodic domain escape",
    4096: "invalid boundary box",
    8192: "active-contact timestep bound",
    16384: "coincident centres unsupported by experimental source-normal contacts",
}
CAPACITY_FLAGS = 8 | 16 | 32 | 64

DAMPING_STATE = ("normalDamping", "forceVelocity", "forceOmega")
RESIDUAL_STATE = (
    "positionRemainder",
    "velocityRemainder",
    "momentumRemainder",
    "rotationRemainder",
    "accumulatedRemainder",
    "cellLog",
)
STATE = ("position", "velocity", "rotation", "momentum", "cell", "force", "torque", "accumulated", *RESIDUAL_STATE)
WALL_STATE = ("wallPlane", "wallRemainder", "wallVelocity", "wallAccumulated", "wallRows")  # gpu.py:32-38

# Field order of the kernels' Resources block (solver.slang:20-100) = C ABI slots.
BUFFER_ORDER = (
    "control", "cell", "position", "positionRemainder", "velocityRemainder",
    "momentumRemainder", "rotationRemainder", "referencePositionRemainder", "cellLog",
    "accumulatedRemainder", "velocity", "rotation", "momentum", "force", "torque",
    "referencePosition", "referenceCell", "shapeId", "body", "invI", "grids", "sampleInfo",
    "gridOrigin", "distance", "sample", "meshRoots", "meshHullInfo", "meshPlanes", "meshMin",
    "meshMax", "meshLinks", "meshTriangles", "meshNormals", "meshFrontier", "meshPairMask",
    "meshPairTransform", "meshClusterInfo", "meshClusterRanges", "meshClusterMin",
    "meshClusterMax", "meshSampleIndices", "binCounts", "binParticles", "pairCounts",
    "pairOffsets", "pairs", "contactCounts", "contactOffsets", "meta", "pointDepth",
    "normalArea", "slip", "contactForce", "normalLossPower", "normalDamping", "forceVelocity",
    "forceOmega", "oldMeta", "oldSlip", "oldNormal", "oldHash", "hash", "heads", "next",
    "particleStress", "timestepRows", "particleDiagnostic", "diagnosticRows", "accumulated",
    "fabricRows", "wallPlane", "wallRemainder", "wallVelocity", "wallMaterial", "wallBoxIds",
    "wallRows", "wallAccumulated",
)
# Config scalars (solver.slang:12-19) with vector fields expanded per axis.
VALUE_ORDER = (
    ("activeTimestep", None), ("N", None), ("shapes", None), ("pairCapacity", None),
    ("contactCapacity", None), ("binCapacity", None), ("binCount", None), ("bins", 0),
    ("bins", 1), ("bins", 2), ("hashSize", None), ("dt", None), ("kn", None), ("kt", None),
    ("mu", None), ("normalViscosity", None), ("skin", None), ("drag", None),
    ("maxOverlap", None), ("historyStep", None), ("meshGeometry", None), ("periodic", None),
    ("wallCount", None), ("wallBoxEnabled", None), ("sourceSurfaceNormal", None),
)


def config_values(c, threads=1):
    """Flat `double values[]` of the C ABI (cpu.py:411-420)."""
    out = [float(c[name] if axis is None else c[name][axis]) for name, axis in VALUE_ORDER]
    out.append(float(threads))
    return np.asarray(out, np.float64)


class RejectedStep(RuntimeError):
    """Device error word was nonzero after a chunk (gpu.py:51-57)."""

    def __init__(self, flags):
        self.flags = int(flags)
        super().__init__("GPU step rejected: " + ", ".join(v for k, v in ERRORS.items() if self.flags & k))


def f4(x):
    a = np.asarray(x, np.float32)
    return np.pad(a, ((0, 0), (0, 4 - a.shape[1])))


class Engine:
    """Mesh-path engine over an explicit backend adapter (gpu.py:94-368)."""

    def __init__(self, shapes, spec, positions, cell, shape_ids, scales, *, backend):
        spec.require_supported_path()
        positions = finite_array(positions)
        cell = finite_array(cell, (3,), positive=True)
        if positions.ndim != 2 or positions.shape[1] != 3 or len(positions) < 1:
            raise ValueError("At least one three-dimensional position required")
        open_axes = ~periodic_axes(spec.periodic)
        if np.any((positions < 0)[:, open_axes]) or np.any((positions >= cell)[:, open_axes]):
            raise ValueError("Nonperiodic particle centres must lie in [0, cell)")
        self.m = backend
        self.spec = spec
        self.shapes = list(shapes)
        self.mesh_geometry = spec.contact_geometry == "mesh"
        self.n = len(positions)
        self.wall_count = len(spec.planes)
        self.viscous = spec.viscous
        self.step = 0
        self.history = False
        self.active = 0
        self.capacity_retries = []
        self.length_unit = float(np.median([s.equivalent_diameter for s in self.shapes]))
        self.time_unit = math.sqrt(spec.density * self.length_unit / spec.normal_stiffness)
        self.mass_unit = spec.density * self.length_unit**3
        self.force_unit = spec.normal_stiffness * self.length_unit**3
        self.stress_unit = spec.normal_stiffness * self.length_unit
        self.energy_unit = self.force_unit * self.length_unit
        self.ids = np.asarray(shape_ids, np.uint32)
        self.scales = np.asarray(scales, float)
        if len(self.ids) != self.n or len(self.scales) != self.n:
            raise ValueError("One shape id and scale per particle required")
        if np.any(self.ids >= len(self.shapes)):
            raise ValueError("Shape id outside the library")
        l = self.length_unit
        self.body = np.array(
            [
                [
                    s,
                    spec.density * self.shapes[int(i)].volume * s**3 / self.mass_unit,
                    self.shapes[int(i)].radius * s / l,
                    self.shapes[int(i)].equivalent_diameter * s / l,
                ]
                for i, s in zip(self.ids, self.scales)
            ],
            np.float32,
        )
        inertia = np.array(
            [
                np.linalg.inv(self.shapes[int(i)].inertia * spec.density * s**5 / (self.mass_unit * l * l))
                for i, s in zip(self.ids, self.scales)
            ]
        )
        self.skin = float(spec.skin * 2 * self.body[:, 2].max())
        self.cutoff = float(2 * self.body[:, 2].max() + self.skin)
        box = np.asarray(cell, float) / l
        axes = periodic_axes(spec.periodic)
        if axes.any() and self.cutoff >= 0.5 * box[axes].min():
            raise ValueError("Cell too small for unique periodic images")
        self.bins = np.maximum(1, np.minimum(64, np.floor(box / self.cutoff))).astype(np.uint32)
        self.bin_count = int(self.bins.prod())
        self.pair_capacity = max(1, self.n * spec.pairs_per_particle)
        self.contact_capacity = max(16, self.n * spec.contacts_per_particle)
        self.hash_size = 1 << math.ceil(math.log2(max(32, 2 * self.contact_capacity)))
        self.bin_capacity = spec.bin_capacity
        self._preflight_dt()
        estimated = (
            244 * self.contact_capacity
            + 8 * self.hash_size
            + 196 * self.pair_capacity
            + 8 * self.bin_count * self.bin_capacity
            + 1184 * self.n
        )
        if self.mesh_geometry:
            estimated += sum(224 * len(s.faces) + 736 + 8 * len(s.samples) for s in self.shapes)
        else:
            for shape in self.shapes:
                shape.validate_sdf()
            estimated += sum(np.asarray(s.sdf).nbytes + 32 * len(s.samples) for s in self.shapes)
        estimated += 160 * self.wall_count * self.n + 256 * self.wall_count
        if estimated > spec.working_memory_gib * 1024**3:
            raise MemoryError(f"Estimated working set {estimated / 1024**3:.2f} GiB exceeds budget")
        self.estimated_bytes = estimated
        self.buffers = {}
        self._static_geometry()
        self.allocate("meshPairMask", self.pair_capacity)
        self.allocate("meshPairTransform", 8 * self.pair_capacity, 16)
        self.add("shapeId", self.ids)
        self.add("body", self.body)
        self.add("invI", f4(inertia.reshape(-1, 3)))
        normalized = np.asarray(positions, float) / l
        self.add("position", f4(normalized))
        self.add("velocity", np.zeros((self.n, 4), np.float32))
        self.add("rotation", np.tile(np.array([0, 0, 0, 1], np.float32), (self.n, 1)))
        for name in ("momentum", "force", "torque", "referencePosition", "particleDiagnostic", "timestepRows", "accumulated"):
            self.add(name, np.zeros((self.n, 4), np.float32))
        self.add("particleStress", np.zeros((3 * self.n, 4), np.float32))
        self.add("diagnosticRows", np.zeros((8 * self.n, 4), np.float32))
        for name in DAMPING_STATE:
            self.add(name, np.zeros((self.n, 4), np.float32))
        for name in RESIDUAL_STATE:
            residual = np.zeros((2 if name == "cellLog" else self.n, 4), np.float32)
            if name == "positionRemainder":
                residual[:, :3] = normalized - normalized.astype(np.float32).astype(float)
            self.add(name, residual)
        self.add("referencePositionRemainder", np.zeros((self.n, 4), np.float32))
        self.allocate("fabricRows", 2 * self.pair_capacity, 16)
        self.add("cell", np.array([list(box) + [0], [0, 0, 0, 0], list(box) + [0]], np.float32))
        self.initial_cell = np.asarray(cell, float)
        if spec.boundary_box is not None:
            self.initial_cell = box_extents(np.array([p["offset"] for p in spec.planes]), spec.boundary_box, cell)
        self._wall_buffers()
        self.add("referenceCell", np.array([list(box) + [0], [0, 0, 0, self.cutoff]], np.float32))
        ctrl = np.zeros(16, np.uint32)
        ctrl[4] = 1
        self.add("control", ctrl)
        for name, count in (
            ("binCounts", self.bin_count),
            ("binParticles", self.bin_count * self.bin_capacity),
            ("pairCounts", self.n),
            ("pairOffsets", self.n),
            ("contactCounts", self.pair_capacity),
            ("contactOffsets", self.pair_capacity),
            ("heads", self.n),
            ("next", 2 * self.contact_capacity),
        ):
            self.allocate(name, count)
        self.allocate("pairs", self.pair_capacity, 8)
        self.history_buffers = []
        for _ in (0, 1):
            self.history_buffers.append(
                {
                    "meta": self.m.buffer(count=self.contact_capacity, stride=16),
                    "slip": self.m.buffer(count=self.contact_capacity, stride=16),
                    "normalArea": self.m.buffer(count=self.contact_capacity, stride=16),
                    "hash": self.m.buffer(np.full(self.hash_size, 0xFFFFFFFF, np.uint32)),
                }
            )
        self.allocate("pointDepth", self.contact_capacity, 16)
        self.allocate("contactForce", self.contact_capacity, 16)
        self.allocate("normalLossPower", self.contact_capacity)
        # Group-total levels of the diagnostics reductions (gpu.py:322-340): each
        # level holds one total per 256 records of the level below.
        self.diag_buffers = self._reduction_levels(self.n, 8)
        self.fabric_buffers = self._reduction_levels(self.pair_capacity, 2)
        self.wall_diag_buffers = self._reduction_levels(self.n, 3 * self.wall_count) if self.wall_count else []
        names = (
            "position", "velocity", "rotation", "momentum", "cell", "force", "torque",
            "referencePosition", "referenceCell", "control", "pairs", "pairCounts", "pairOffsets",
            "particleStress", "particleDiagnostic", "timestepRows", "accumulated",
            "referencePositionRemainder", *RESIDUAL_STATE, *DAMPING_STATE,
            *(WALL_STATE if self.wall_count else ()),
        )
        self.backups = {name: self.m.buffer(count=self.buffers[name].size // 4) for name in names}
        self.history_backup = {
            name: self.m.buffer(count=self.hash_size if name == "hash" else 4)
            for name in ("meta", "slip", "normalArea", "hash")
        }

    def _reduction_levels(self, size, rows):
        levels = []
        while size > 1:
            size = (size + 255) // 256
            levels.append((size, self.m.buffer(count=size * rows, stride=16)))
        return levels

    def _preflight_dt(self):  # gpu.py:370-413
        if self.spec.timestep_policy == "active-contact":
            return
        limits = []
        for i, s in zip(self.ids, self.scales):
            sh = self.shapes[int(i)]
            mass = self.spec.density * sh.volume * s**3
            stiffness = (self.spec.normal_stiffness + self.spec.tangential_stiffness) * sh.areas.sum() * s * s
            rotation = np.linalg.eigvalsh(sh.inertia * self.spec.density * s**5).min() / (
                stiffness * (sh.radius * s) ** 2
            )
            elastic_limit = math.sqrt(min(mass / stiffness, rotation))
            viscous_limit = math.inf
            viscosity = self.spec.normal_viscosity * sh.areas.sum() * s**2
            if viscosity > 0:
                viscous_limit = min(
                    mass / viscosity,
                    np.linalg.eigvalsh(sh.inertia * self.spec.density * s**5).min() / (viscosity * (sh.radius * s) ** 2),
                )
            limits.append(0.1 * min(elastic_limit, viscous_limit))
        if self.spec.timestep > min(limits):
            raise ValueError(f"Timestep exceeds conservative estimate {min(limits):.6g} s")

    def add(self, name, array):
        self.buffers[name] = self.m.buffer(array)

    def allocate(self, name, count, stride=4):
        self.buffers[name] = self.m.buffer(count=count, stride=stride)

    def _wall_buffers(self):  # gpu.py:232-285
        spec, l = self.spec, self.length_unit
        wall_plane = np.zeros((max(1, self.wall_count), 4), np.float64)
        wall_velocity = np.zeros_like(wall_plane)
        wall_material = np.zeros_like(wall_plane)
        for index, plane in enumerate(spec.planes):
            wall_plane[index] = [*plane["normal"], plane["offset"] / l]
            wall_velocity[index, :3] = np.array(plane["velocity"]) * self.time_unit / l
            wall_material[index, :3] = [
                plane["normal_stiffness"] / spec.normal_stiffness,
                plane["tangential_stiffness"] / spec.normal_stiffness,
                plane["friction"],
            ]
            wall_material[index, 3] = plane.get("normal_viscosity", 0) / (spec.normal_stiffness * self.time_unit)
        rounded_plane = wall_plane.astype(np.float32)
        residual_plane = (wall_plane - rounded_plane).astype(np.float32)
        if (
            not np.isfinite(rounded_plane).all()
            or not np.isfinite(residual_plane).all()
            or not np.isfinite(wall_velocity.astype(np.float32)).all()
            or not np.isfinite(wall_material.astype(np.float32)).all()
            or (self.wall_count and np.any(wall_material[: self.wall_count, :2].astype(np.float32) <= 0))
        ):
            raise ValueError("Plane data cannot be represented in normalized FP32 units")
        for index, plane in enumerate(spec.planes):
            if plane.get("normal_viscosity", 0) > 0 and np.float32(wall_material[index, 3]) <= 0:
                raise ValueError("Plane viscosity cannot be represented in normalized FP32 units")
        self.add("wallPlane", rounded_plane)
        self.add("wallRemainder", residual_plane)
        self.add("wallVelocity", wall_velocity.astype(np.float32))
        self.add("wallMaterial", wall_material.astype(np.float32))
        self.add("wallBoxIds", np.array(spec.boundary_box or [0] * 6, np.uint32))
        self.add("wallAccumulated", np.zeros((max(1, self.wall_count * self.n), 4), np.float32))
        self.add("wallRows", np.zeros((max(1, 3 * self.wall_count * self.n), 4), np.float32))

    def _static_geometry(self):  # gpu.py:421-488
        if not self.mesh_geometry:
            self._sdf_geometry()
            return
        mesh, self.mesh_cache_receipt = packed_meshes(self.shapes, self.length_unit)
        roots = mesh["meshRoots"]
        if np.any(roots[:, 1] - roots[:, 0] > 65535):
            raise ValueError("A template BVH exceeds 65,535 nodes (device traversal stack holds 16-bit offsets)")
        if np.any(roots[:, 2] > 14):
            raise ValueError("A template BVH is deeper than 14 levels (device traversal stack holds 16 entries)")
        if len(mesh["meshTriangles"]) // 3 >= 1 << 28:
            raise ValueError("More than 2^28 triangles in the template library")
        if np.any(mesh["meshClusterInfo"][:, 1] > 1024):
            raise ValueError("A template has more than 1,024 sample clusters (about 32k nodes)")
        for name, data in mesh.items():
            self.add(name, data)
        info, samples = sample_arrays(self.shapes, self.length_unit)
        self.add("sampleInfo", info)
        self.add("sample", samples)
        # SDF-path inputs are unused by the mesh step; bound as one-row placeholders.
        self.add("grids", np.zeros((1, 4), np.uint32))
        self.add("gridOrigin", np.zeros((1, 4), np.float32))
        self.add("distance", np.zeros(4, np.float32))

    def _sdf_geometry(self):
        """SDF branch (gpu.py:427-488): one-row mesh placeholders, then per template
        the grid row (first value, nx, ny, nz), origin and spacing in normalized
        units, and the grid values (normalized) concatenated in C order."""
        for name in ("meshHullInfo", "meshPlanes", "meshRoots", "meshLinks", "meshMin", "meshMax", "meshTriangles",
                     "meshNormals", "meshFrontier", "meshClusterInfo", "meshClusterRanges", "meshClusterMin",
                     "meshClusterMax", "meshSampleIndices"):
            self.add(name, np.zeros((1, 4), np.uint32 if name in _UINT_ARRAYS else np.float32))
        self.mesh_cache_receipt = dict(status="sdf")
        l = self.length_unit
        grids, origins, fields, first = [], [], [], 0
        for shape in self.shapes:
            grids.append([first, *np.asarray(shape.sdf).shape])
            origins.append([*(np.asarray(shape.origin) / l), shape.spacing / l])
            fields.append((np.asarray(shape.sdf) / l).ravel())
            first += np.asarray(shape.sdf).size
        if first >= 1 << 32:
            raise ValueError("SDF grids exceed 32-bit indices")
        info, samples = sample_arrays(self.shapes, l)
        self.add("grids", np.asarray(grids, np.uint32))
        self.add("sampleInfo", info)
        self.add("gridOrigin", np.asarray(origins, np.float32))
        self.add("distance", np.concatenate(fields).astype(np.float32))
        self.add("sample", samples)

    def config(self, dt=None, drag=0):  # gpu.py:493-518
        spec = self.spec
        return {
            "activeTimestep": int(spec.timestep_policy == "active-contact"),
            "N": self.n,
            "shapes": len(self.shapes),
            "pairCapacity": self.pair_capacity,
            "contactCapacity": self.contact_capacity,
            "binCapacity": self.bin_capacity,
            "binCount": self.bin_count,
            "bins": self.bins.tolist(),
            "hashSize": self.hash_size,
            "dt": (spec.timestep if dt is None else dt) / self.time_unit,
            "kn": 1.0,
            "kt": spec.tangential_stiffness / spec.normal_stiffness,
            "mu": spec.friction,
            "normalViscosity": spec.normal_viscosity / (spec.normal_stiffness * self.time_unit),
            "skin": self.skin,
            "drag": drag * self.time_unit,
            "maxOverlap": spec.max_overlap_fraction,
            "historyStep": int(self.history),
            "meshGeometry": int(self.mesh_geometry),
            "sourceSurfaceNormal": 0,
            "periodic": periodic_code(spec.periodic),
            "wallCount": self.wall_count,
            "wallBoxEnabled": int(self.spec.boundary_box is not None),
        }

    def bindings(self, dt=None, drag=0):  # gpu.py:490-527
        new = self.history_buffers[self.active]
        old = self.history_buffers[1 - self.active]
        return dict(
            self.buffers,
            c=self.config(dt, drag),
            **new,
            oldMeta=old["meta"],
            oldSlip=old["slip"],
            oldNormal=old["normalArea"],
            oldHash=old["hash"],
        )

    def dispatch(self, name, n, e, b):
        self.m.dispatch(name, n, e, r=b)

    def _encode_contacts(self, e, b):  # gpu.py:532-545
        if self.mesh_geometry:
            self.dispatch("resetMeshContacts", self.pair_capacity, e, b)
            self.dispatch("prepareMeshPairs", self.pair_capacity, e, b)
            self.dispatch("collectMeshFrontiers", 64 * 4096, e, b)
        else:
            self.dispatch("countContacts", 64 * self.pair_capacity, e, b)
            self.dispatch("contactScan", self.pair_capacity, e, b)
            self.dispatch("contactTotal", 1, e, b)
            self.dispatch("emitContacts", 64 * self.pair_capacity, e, b)
        if self.wall_count:
            self.dispatch("collectWallContacts", 64 * self.n * self.wall_count, e, b)

    def forces(self, e, b, *, capture_kinematics=True):  # gpu.py:547-565
        if self.viscous and capture_kinematics:
            self.dispatch("captureContactKinematics", self.n, e, b)
        self.dispatch("clearBins", self.bin_count, e, b)
        self.dispatch("fillBins", self.n, e, b)
        self.dispatch("countPairs", self.n, e, b)
        self.dispatch("pairScan", self.n, e, b)
        self.dispatch("emitPairs", self.n, e, b)
        self._encode_contacts(e, b)
        self.dispatch("clearContactStorage", max(self.n, self.hash_size), e, b)
        self.dispatch("resolveContacts", self.contact_capacity, e, b)
        self.dispatch("buildHistoryHash", self.contact_capacity, e, b)
        self.dispatch("gather", self.n, e, b)
        if self.wall_count:
            self.dispatch("gatherWallRows", self.n * self.wall_count, e, b)
        if self.spec.timestep_policy == "active-contact":
            self.dispatch("checkContactTimestep", self.n, e, b)

    def initialize(self):  # gpu.py:580-587
        e = self.m.device.create_command_encoder()
        self.forces(e, self.bindings(dt=0))
        self.m.submit(e)
        self.m.device.wait()
        self.check_errors()
        self.history = True
        self.active = 1 - self.active

    def check_errors(self):  # gpu.py:589-595
        ctrl = self.m.read(self.buffers["control"], np.uint32)
        if ctrl[15] != 0:
            raise RuntimeError("Arithmetic materialization mask must remain zero")
        if ctrl[0]:
            raise RejectedStep(ctrl[0])
        return ctrl

    def advance(self, steps, *, rate=(0, 0, 0), drag=0, wall_velocities=None):  # gpu.py:597-647
        if type(steps) is not int or steps < 0:
            raise ValueError("Nonnegative integer step count required")
        rate = finite_array(rate, (3,))
        if np.any(rate[~periodic_axes(self.spec.periodic)] != 0):
            raise ValueError("Affine cell rates require a periodic domain")
        if not np.isfinite(drag) or drag < 0:
            raise ValueError("Finite nonnegative drag required")
        if wall_velocities is not None:
            if not self.wall_count:
                raise ValueError("Wall velocities require planes")
            wall_velocities = finite_array(wall_velocities, (self.wall_count, 3))
        for attempt in range(5):
            try:
                return self._advance_once(steps, rate=rate, drag=drag, wall_velocities=wall_velocities)
            except RejectedStep as exc:
                capacity = exc.flags & CAPACITY_FLAGS
                if not capacity or exc.flags & ~CAPACITY_FLAGS or attempt == 4:
                    raise
                state = self.state()
                old = self.spec
                settings = {}
                if capacity & 8:
                    settings["bin_capacity"] = old.bin_capacity * 2
                if capacity & 16:
                    settings["pairs_per_particle"] = old.pairs_per_particle * 2
                if capacity & (32 | 64):
                    settings["contacts_per_particle"] = max(
                        old.contacts_per_particle * 2, math.ceil(2 * self.contact_capacity / self.n)
                    )
                log = self.capacity_retries + [{"step": self.step, "flags": exc.flags, "new_capacities": settings}]
                replacement = Engine(
                    self.shapes,
                    replace(old, **settings),
                    state["position"][:, :3] * self.length_unit,
                    state["cell"][0, :3] * self.length_unit,
                    self.ids,
                    self.scales,
                    backend=self.m,
                )
                replacement.initial_cell = self.initial_cell.copy()
                replacement.restore(state, self.step)
                self.__dict__.update(replacement.__dict__)
                self.capacity_retries = log

    def _reserve_history_backup(self):  # gpu.py:649-661
        count = int(self.m.read(self.buffers["control"], np.uint32)[2])
        if count < 0 or count > self.contact_capacity:
            raise ValueError("Committed contact count is outside capacity")
        size = max(16, count * 16)
        self.history_backup.update(
            {
                name: self.m.buffer(count=size // 4)
                for name in ("meta", "slip", "normalArea")
                if self.history_backup[name].size < size
            }
        )
        return count

    def _copy_history_backup(self, encoder, committed, count, *, restore=False):  # gpu.py:663-670
        for name, backup in self.history_backup.items():
            size = committed[name].size if name == "hash" else count * 16
            if size:
                src, dst = (backup, committed[name]) if restore else (committed[name], backup)
                encoder.copy_buffer(dst, 0, src, 0, size)

    def _encode_steps(self, e, steps, drag):
        """Per-step sequence (gpu.py:714-726); natively encoded when offered."""
        if hasattr(self.m, "encode_steps"):
            self.m.encode_steps(e, steps, self.active, self.bindings(drag=drag), self.history_buffers)
            if steps:
                self.active = (self.active + steps) % 2
                self.history = True
            return
        for _ in range(steps):
            b = self.bindings(drag=drag)
            if self.spec.timestep_policy == "active-contact":
                self.dispatch("checkContactTimestep", self.n, e, b)
            self.dispatch("beginStep", self.n, e, b)
            self.dispatch("advanceCell", 1, e, b)
            if self.wall_count:
                self.dispatch("advanceWalls", 1, e, b)
            self.dispatch("checkRebuild", self.n, e, b)
            self.forces(e, b)
            self.dispatch("finishStep", self.n, e, b)
            self.active = 1 - self.active
            self.history = True

    def _advance_once(self, steps, *, rate=(0, 0, 0), drag=0, wall_velocities=None):  # gpu.py:672-744
        if getattr(self, "invalid", False):
            raise RuntimeError("Engine is invalid; restore a checkpoint")
        if steps < 1:
            return
        history_count = self._reserve_history_backup()
        cell = self.m.read(self.buffers["cell"], shape=(3, 4))
        original_cell = cell.copy()
        cell[1, :3] = np.asarray(rate) * self.time_unit
        original_wall_velocity = normalized_velocity = None
        if wall_velocities is not None:
            original_wall_velocity = self.m.read(self.buffers["wallVelocity"], shape=(self.wall_count, 4))
            normalized_velocity = f4(wall_velocities * self.time_unit / self.length_unit)
        self.buffers["cell"].copy_from_numpy(cell)
        if normalized_velocity is not None:
            self.buffers["wallVelocity"].copy_from_numpy(normalized_velocity)
        e = self.m.device.create_command_encoder()
        for name, backup in self.backups.items():
            e.copy_buffer(backup, 0, self.buffers[name], 0, self.buffers[name].size)
        original_active, original_history = self.active, self.history
        committed = self.history_buffers[1 - self.active]
        self._copy_history_backup(e, committed, history_count)
        if original_wall_velocity is not None:
            if any(p.get("normal_viscosity", 0) > 0 for p in self.spec.planes) and not np.array_equal(
                original_wall_velocity, normalized_velocity
            ):
                self.forces(e, self.bindings(dt=0))
                self.active = 1 - self.active
                self.history = True
            else:
                self.dispatch("refreshWallPower", self.n * self.wall_count, e, self.bindings())
        self._encode_steps(e, steps, drag)
        self.m.submit(e)
        self.m.device.wait()
        try:
            self.check_errors()
        except RuntimeError:
            e = self.m.device.create_command_encoder()
            for name, backup in self.backups.items():
                e.copy_buffer(self.buffers[name], 0, backup, 0, backup.size)
            self._copy_history_backup(e, committed, history_count, restore=True)
            self.m.submit(e)
            self.m.device.wait()
            self.buffers["cell"].copy_from_numpy(original_cell)
            if original_wall_velocity is not None:
                self.buffers["wallVelocity"].copy_from_numpy(original_wall_velocity)
            self.active, self.history = original_active, original_history
            raise
        self.step += steps

    def state(self):  # gpu.py:893-936
        out = {
            name: self.m.read(self.buffers[name], shape=(3 if name == "cell" else 2 if name == "cellLog" else self.n, 4))
            for name in STATE
        }
        if self.viscous:
            for name in DAMPING_STATE:
                out[name] = self.m.read(self.buffers[name], shape=(self.n, 4))
        if self.wall_count:
            for name in WALL_STATE:
                size = self.wall_count * self.n * (3 if name == "wallRows" else 1) if name in {
                    "wallRows", "wallAccumulated"} else self.wall_count
                out[name] = self.m.read(self.buffers[name], shape=(size, 4))
        ctrl = self.check_errors()
        h = self.history_buffers[1 - self.active]
        count = int(ctrl[2])
        out["contacts"] = count
        for name in ("meta", "slip", "normalArea"):
            out[name] = self.m.read_prefix(
                h[name], 4 * count, np.uint32 if name == "meta" else np.float32, shape=(count, 4)
            )
        return out

    def restore(self, state, step):  # gpu.py:1016-1092
        if self.viscous:
            for name in DAMPING_STATE:
                finite_array(state[name], (self.n, 4))
        if self.wall_count:
            for name in WALL_STATE:
                value = finite_array(state[name])
                if value.shape != (self.buffers[name].size // 16, 4):
                    raise ValueError("Checkpoint wall state shape mismatch")
            normals = np.array([plane["normal"] for plane in self.spec.planes], np.float32)
            if not np.array_equal(state["wallPlane"][:, :3], normals):
                raise ValueError("Checkpoint plane normal identity mismatch")
            offsets = (np.asarray(state["wallPlane"][:, 3], float) + state["wallRemainder"][:, 3]) * self.length_unit
            if self.spec.boundary_box is not None:
                box_extents(offsets, self.spec.boundary_box, np.asarray(state["cell"][0, :3], float) * self.length_unit)
            for name in WALL_STATE:
                self.buffers[name].copy_from_numpy(np.asarray(state[name], np.float32))
        for name in DAMPING_STATE:
            self.buffers[name].copy_from_numpy(state[name] if self.viscous else np.zeros((self.n, 4), np.float32))
        for name in STATE:
            self.buffers[name].copy_from_numpy(np.asarray(state[name], np.float32))
        count = int(state["contacts"])
        if count < 0 or count > self.contact_capacity:
            raise ValueError("Checkpoint contact count is outside capacity")
        h = self.history_buffers[1]
        for name in ("meta", "slip", "normalArea"):
            values = np.asarray(state[name], dtype=np.uint32 if name == "meta" else np.float32)
            if values.shape != (count, 4):
                raise ValueError("Checkpoint contact history shape mismatch")
            self.m.write_prefix(h[name], values)
        ctrl = np.zeros(16, np.uint32)
        ctrl[2] = count
        ctrl[4] = 1
        self.buffers["control"].copy_from_numpy(ctrl)
        self.active = 1
        self.history = True
        e = self.m.device.create_command_encoder()
        b = self.bindings(dt=0)
        self.dispatch("clearContactStorage", max(self.n, self.hash_size), e, b)
        self.dispatch("buildHistoryHash", self.contact_capacity, e, b)
        self.m.submit(e)
        self.m.device.wait()
        self.check_errors()
        self.active = 0
        ctrl[4] = 1
        self.buffers["control"].copy_from_numpy(ctrl)
        e = self.m.device.create_command_encoder()
        self.forces(e, self.bindings(dt=0), capture_kinematics=False)
        self.m.submit(e)
        self.m.device.wait()
        self.check_errors()
        self.active = 1
        self.step = step

    # ---- convenience, not in the reference: direct state upload for generated workloads
    def set_kinematics(self, *, rotation=None, velocity=None, momentum=None):
        """Overwrite body rotations (xyzw), velocities or momenta in normalized units."""
        if rotation is not None:
            self.buffers["rotation"].copy_from_numpy(f4(rotation))
        if velocity is not None:
            self.buffers["velocity"].copy_from_numpy(f4(velocity))
        if momentum is not None:
            self.buffers["momentum"].copy_from_numpy(f4(momentum))

    def diagnostics(self):  # gpu.py:746-891
        """Whole-assembly evidence of the last completed step in SI units: stress,
        energies, fabric, coordination and per-plane reactions.  Per-grain and
        per-pair rows are totalled on the device by the diagnostic.slang tree."""
        e = self.m.device.create_command_encoder()
        self.dispatch("diagnosticRows", self.n, e, self.bindings())
        source = self._reduce(e, "reduceRows", self.buffers["diagnosticRows"], self.n, self.diag_buffers)
        self.dispatch("fabricRows", self.pair_capacity, e, self.bindings())
        fabric_source = self._reduce(e, "reduceFabric", self.buffers["fabricRows"], self.pair_capacity, self.fabric_buffers)
        wall_source = self._reduce(e, "reduceWallRows", self.buffers["wallRows"], self.n, self.wall_diag_buffers)
        self.m.submit(e)
        self.m.device.wait()
        data = self.m.read(source, shape=(-1, 4))[:8]
        fabric_data = self.m.read(fabric_source, shape=(-1, 4))[:2]
        active_pairs = round(float(fabric_data[1, 2]))
        fabric = np.array(
            [
                [fabric_data[0, 0], fabric_data[0, 3], fabric_data[1, 0]],
                [fabric_data[0, 3], fabric_data[0, 1], fabric_data[1, 1]],
                [fabric_data[1, 0], fabric_data[1, 1], fabric_data[0, 2]],
            ],
            float,
        ) / max(active_pairs, 1)
        cell = self.m.read(self.buffers["cell"], shape=(-1, 4))[:3]
        ctrl = self.check_errors()
        box = cell[0, :3].astype(float) * self.length_unit
        domain = box.copy()
        wall_results = []
        if self.wall_count:
            wall_data = self.m.read(wall_source, shape=(-1, 4))[: 3 * self.wall_count].reshape(self.wall_count, 3, 4)
            counts = wall_data[:, 0, 3].copy().view(np.uint32)
            plane = self.m.read(self.buffers["wallPlane"], shape=(-1, 4))[: self.wall_count].astype(float)
            remainder = self.m.read(self.buffers["wallRemainder"], shape=(-1, 4))[: self.wall_count].astype(float)
            wall_velocity = self.m.read(self.buffers["wallVelocity"], shape=(-1, 4))[: self.wall_count].astype(float)
            offsets = (plane[:, 3] + remainder[:, 3]) * self.length_unit
            for w in range(self.wall_count):
                wall_results.append(
                    dict(
                        index=w,
                        normal=plane[w, :3].tolist(),
                        offset=float(offsets[w]),
                        velocity=(wall_velocity[w, :3] * self.length_unit / self.time_unit).tolist(),
                        reaction=(wall_data[w, 0, :3].astype(float) * self.force_unit).tolist(),
                        reaction_moment_origin=(wall_data[w, 1, :3].astype(float) * self.energy_unit).tolist(),
                        active_surface_contacts=int(counts[w]),
                        power=float(wall_data[w, 1, 3] * self.energy_unit / self.time_unit),
                        work=float(wall_data[w, 2, 0] * self.energy_unit),
                    )
                )
            if self.spec.boundary_box is not None:
                box = box_extents(offsets, self.spec.boundary_box, domain)
        volume = float(box.prod())
        contact = data[:3, :3].astype(float) * self.energy_unit / volume
        kinetic = data[3:6, :3].astype(float) * self.energy_unit / volume
        total = contact + kinetic
        symmetric = 0.5 * (total + total.T)
        radial = 0.5 * (symmetric[0, 0] + symmetric[1, 1])
        solid = sum(self.shapes[int(i)].volume * s**3 for i, s in zip(self.ids, self.scales))
        return {
            "step": self.step,
            "elapsed_time": self.step * self.spec.timestep,
            "cell": box.tolist(),
            "domain_cell": domain.tolist(),
            "walls": wall_results,
            "stress": total.tolist(),
            "contact_stress": contact.tolist(),
            "kinetic_stress": kinetic.tolist(),
            "stress_antisymmetry": float(np.linalg.norm(total - total.T) / max(np.linalg.norm(total), 1e-30)),
            "p": float(np.trace(symmetric) / 3),
            "q": float(symmetric[2, 2] - radial),
            "axial_strain": float(1 - box[2] / self.initial_cell[2]),
            "volumetric_strain": float(np.sum(1 - box / self.initial_cell)),
            "volume_ratio": float(volume / np.prod(self.initial_cell)),
            "log_strain": (-np.log(box / self.initial_cell)).tolist(),
            "void_ratio": float(volume / solid - 1),
            "kinetic_energy": float(data[0, 3] * self.energy_unit),
            "elastic_energy_proxy": float(data[1, 3] * self.energy_unit),
            "contact_normal": self.spec.contact_normal,
            "elastic_energy_interpretation": "existing nodal spring energy proxy",
            "unbalanced_force": float(data[2, 3] / max(data[3, 3], 1e-30)),
            "max_overlap_fraction": float(data[4, 3]),
            "friction_dissipation_last_step": float(data[5, 3] * self.energy_unit),
            "candidate_pairs": int(ctrl[1]),
            "active_surface_contacts": int(ctrl[2]),
            "active_grain_pairs": active_pairs,
            "coordination": 2 * active_pairs / self.n,
            "fabric": fabric.tolist(),
            "sliding_fraction": float(ctrl[6] / max(ctrl[2], 1)),
            "cell_log_rate": (cell[1, :3].astype(float) / self.time_unit).tolist(),
            "friction_dissipation": float(data[6, 0] * self.energy_unit),
            "drag_dissipation": float(data[6, 1] * self.energy_unit),
            "cell_work": float(data[6, 2] * self.energy_unit),
            "wall_work": float(data[6, 3] * self.energy_unit),
            "normal_contact_dissipation": float(data[7, 0] * self.energy_unit),
            "normal_contact_power": float(data[7, 1] * self.energy_unit / self.time_unit),
        }

    def _reduce(self, e, name, source, n, levels):
        """Encode the reduction levels; returns the buffer holding the final totals."""
        kernel = self.m.kernel(name, "diagnostic.slang")
        for size, dest in levels:
            extra = {"outputCount": size} if name == "reduceWallRows" else {}
            walls = self.wall_count if name == "reduceWallRows" else 1
            kernel.dispatch(thread_count=[size * 256, walls, 1], command_encoder=e, N=n, input=source, output=dest, **extra)
            source, n = dest, size
        return source

    def walls(self):
        """Per-plane totals of the last force evaluation in SI units (host FP64 sums of
        the per-body wall rows; the reference reduces the same rows on the device,
        diagnostic.slang / gpu.py:774-835)."""
        if not self.wall_count:
            return []
        rows = self.m.read(self.buffers["wallRows"], shape=(self.wall_count, self.n, 3, 4)).astype(np.float64)
        counts = self.m.read(self.buffers["wallRows"], np.uint32, shape=(self.wall_count, self.n, 3, 4))[:, :, 0, 3]
        plane = self.m.read(self.buffers["wallPlane"], shape=(self.wall_count, 4)).astype(np.float64)
        rem = self.m.read(self.buffers["wallRemainder"], shape=(self.wall_count, 4)).astype(np.float64)
        l, f = self.length_unit, self.force_unit
        out = []
        for w in range(self.wall_count):
            out.append({
                "normal": plane[w, :3].tolist(),
                "offset": float((plane[w, 3] + rem[w, 3]) * l),
                "reaction": (rows[w, :, 0, :3].sum(0) * f).tolist(),
                "reaction_moment_origin": (rows[w, :, 1, :3].sum(0) * f * l).tolist(),
                "power": float(rows[w, :, 1, 3].sum() * f * l / self.time_unit),
                "work": float(rows[w, :, 2, 0].sum() * self.energy_unit),
                "active_surface_contacts": int(counts[w].sum()),
            })
        return out

    def read(self, name, dtype=np.float32, rows=4):
        """Whole buffer as an array of `rows`-wide records."""
        a = self.m.read(self.buffers[name], dtype)
        return a.reshape(-1, rows) if rows > 1 else a

    def contacts(self):
        """Committed contact rows of the last completed step (meta, pointDepth, normalArea)."""
        ctrl = self.m.read(self.buffers["control"], np.uint32)
        count = int(ctrl[2])
        h = self.history_buffers[1 - self.active]
        return (
            self.m.read_prefix(h["meta"], 4 * count, np.uint32, shape=(count, 4)),
            self.m.read_prefix(self.buffers["pointDepth"], 4 * count, np.float32, shape=(count, 4)),
            self.m.read_prefix(h["normalArea"], 4 * count, np.float32, shape=(count, 4)),
        )
