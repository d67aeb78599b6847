"""Host driver of the DEM timestep (surface-mesh and SDF-grid contact geometry).

Every per-step operation runs on the device.  The host decides what cannot be
decided there: the unit system, the capacities, the device buffers in the
kernels' `Resources` layout, the order of the kernels within a timestep, what
happens when a chunk of steps sets a bit in the device error word (roll back,
grow, retry), which of the friction history's two sides a step writes, and what a
checkpoint holds.

The driver reaches its backend only through the adapter protocol the reference
Engine uses (buffer / dispatch / command encoder with copy_buffer / submit / wait /
read / read_prefix / write_prefix: `gpu.py:75-90, 415-419, 529-530, 581`), so the
HIP adapter and the oracle run under the same host logic.  An adapter that offers
`encode_steps` receives whole chunks of timesteps and runs them natively.

Kept from the reference because the kernels, checkpoints and users depend on them,
each cited where it is implemented: the unit system, capacity formulas and growth
rule, the kernel order of a step, the error bits, what a rollback restores and what
a checkpoint carries.
"""

import math
from dataclasses import dataclass, replace

import numpy as np

from .geometry import _UINT_ARRAYS, sample_arrays
from .mesh_cache import packed_meshes
from .spec import box_extents, finite_array, periodic_axes, periodic_code

# ---------------------------------------------------------------- contract tables

# Bits of control[0] and their meaning (gpu.py:11-26).
ERRORS = {
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

# Field order of the kernels' Resources block (solver.slang:20-100) = the C ABI's slots.
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
# Config scalars (solver.slang:12-19), vector fields expanded per axis.
VALUE_ORDER = (
    ("activeTimestep", None), ("N", None), ("shapes", None), ("pairCapacity", None),
    ("contactCapacity", None), ("binCapacity", None), ("binCount", None), ("bins", 0),
    ("bins", 1), ("bins", 2), ("hashSize", None), ("dt", None), ("kn", None), ("kt", None),
    ("mu", None), ("normalViscosity", None), ("skin", None), ("drag", None),
    ("maxOverlap", None), ("historyStep", None), ("meshGeometry", None), ("periodic", None),
    ("wallCount", None), ("wallBoxEnabled", None), ("sourceSurfaceNormal", None),
)

# What a checkpoint carries (gpu.py:893-936) and what a rollback restores (gpu.py:490-527).
KINEMATIC = ("position", "velocity", "rotation", "momentum", "cell", "force", "torque", "accumulated")
RESIDUALS = ("positionRemainder", "velocityRemainder", "momentumRemainder", "rotationRemainder",
             "accumulatedRemainder", "cellLog")
STATE = KINEMATIC + RESIDUALS
DAMPING_STATE = ("normalDamping", "forceVelocity", "forceOmega")
WALL_STATE = ("wallPlane", "wallRemainder", "wallVelocity", "wallAccumulated", "wallRows")
HISTORY = ("meta", "slip", "normalArea")
ROLLBACK = ("position", "velocity", "rotation", "momentum", "cell", "force", "torque", "referencePosition",
            "referenceCell", "control", "pairs", "pairCounts", "pairOffsets", "particleStress",
            "particleDiagnostic", "timestepRows", "accumulated", "referencePositionRemainder", *RESIDUALS,
            *DAMPING_STATE)
EMPTY = 0xFFFFFFFF


def config_values(config, threads=1):
    """The C ABI's `double values[]`: Config scalars in VALUE_ORDER, then the
    worker-count slot the reference appends (cpu.py:411-420)."""
    row = [config[name] if axis is None else config[name][axis] for name, axis in VALUE_ORDER]
    return np.asarray([*map(float, row), float(threads)], np.float64)


class RejectedStep(RuntimeError):
    """A chunk ended with a nonzero device error word (gpu.py:51-57)."""

    def __init__(self, flags):
        self.flags = int(flags)
        names = [text for bit, text in ERRORS.items() if self.flags & bit]
        super().__init__("GPU step rejected: " + ", ".join(names))


def rows4(values):
    """(n, k <= 4) -> float32 (n, 4), zero padded."""
    a = np.asarray(values, np.float32)
    return np.pad(a, ((0, 0), (0, 4 - a.shape[1])))


# ---------------------------------------------------------------- units and sizes


@dataclass(frozen=True)
class Units:
    """Normalized units (gpu.py:118-127): length = median equivalent diameter,
    time = sqrt(rho l / kn), mass = rho l^3, force = kn l^3."""

    length: float
    time: float
    mass: float
    force: float
    stress: float

    @classmethod
    def of(cls, shapes, spec):
        length = float(np.median([s.equivalent_diameter for s in shapes]))
        return cls(length, math.sqrt(spec.density * length / spec.normal_stiffness), spec.density * length**3,
                   spec.normal_stiffness * length**3, spec.normal_stiffness * length)

    @property
    def energy(self):
        return self.force * self.length


@dataclass(frozen=True)
class Sizing:
    """Neighbour-search geometry and buffer capacities (gpu.py:157-169)."""

    skin: float
    cutoff: float
    bins: np.ndarray
    pairs: int
    contacts: int
    hash_slots: int
    bin_capacity: int

    @classmethod
    def plan(cls, spec, body, box, n):
        largest = body[:, 2].max()  # float32, as the reference computes these
        skin = float(spec.skin * 2 * largest)
        cutoff = float(2 * largest + skin)
        axes = periodic_axes(spec.periodic)
        if axes.any() and cutoff >= 0.5 * box[axes].min():
            raise ValueError("Cell too small for unique periodic images")
        bins = np.maximum(1, np.minimum(64, np.floor(box / cutoff))).astype(np.uint32)
        contacts = max(16, n * spec.contacts_per_particle)
        return cls(skin, cutoff, bins, max(1, n * spec.pairs_per_particle), contacts,
                   1 << math.ceil(math.log2(max(32, 2 * contacts))), spec.bin_capacity)


def grown_settings(spec, flags, contact_capacity, n):
    """Spec fields to double after a capacity overflow (gpu.py:618-633)."""
    settings = {}
    if flags & 8:
        settings["bin_capacity"] = spec.bin_capacity * 2
    if flags & 16:
        settings["pairs_per_particle"] = spec.pairs_per_particle * 2
    if flags & (32 | 64):
        settings["contacts_per_particle"] = max(spec.contacts_per_particle * 2, math.ceil(2 * contact_capacity / n))
    return settings


def grain_rows(shapes, spec, units, ids, scales):
    """Per grain: (scale, mass, bounding radius, equivalent diameter) and the inverse
    inertia tensor, both in normalized units (gpu.py:128-156)."""
    l = units.length
    body, inverse = [], []
    for i, s in zip(ids, scales):
        shape = shapes[int(i)]
        body.append([s, spec.density * shape.volume * s**3 / units.mass, shape.radius * s / l,
                     shape.equivalent_diameter * s / l])
        inverse.append(np.linalg.inv(shape.inertia * spec.density * s**5 / (units.mass * l * l)))
    return np.array(body, np.float32), np.array(inverse)


def check_timestep(shapes, spec, ids, scales):
    """The conservative preflight (gpu.py:370-413): dt at most 0.1 of the fastest
    elastic (and viscous) time of any grain."""
    if spec.timestep_policy == "active-contact":
        return
    limit = math.inf
    for i, s in zip(ids, scales):
        shape = shapes[int(i)]
        mass = spec.density * shape.volume * s**3
        soft = np.linalg.eigvalsh(shape.inertia * spec.density * s**5).min()
        lever = (shape.radius * s) ** 2
        stiffness = (spec.normal_stiffness + spec.tangential_stiffness) * shape.areas.sum() * s * s
        times = [math.sqrt(min(mass / stiffness, soft / (stiffness * lever)))]
        viscosity = spec.normal_viscosity * shape.areas.sum() * s**2
        if viscosity > 0:
            times.append(min(mass / viscosity, soft / (viscosity * lever)))
        limit = min(limit, 0.1 * min(times))
    if spec.timestep > limit:
        raise ValueError(f"Timestep exceeds conservative estimate {limit:.6g} s")


# ---------------------------------------------------------------- rollback


class _Rollback:
    """Device copies of everything a chunk of steps may change, taken by the chunk's
    own encoder before its first step and copied back if the chunk is rejected
    (gpu.py:649-744).  The committed friction history is saved up to its contact
    count; the whole hash of that side is saved."""

    def __init__(self, engine, names):
        m = engine.m
        self.engine = engine
        self.copies = {name: m.buffer(count=engine.buffers[name].size // 4) for name in names}
        self.history = {name: m.buffer(count=engine.hash_size if name == "hash" else 4)
                        for name in (*HISTORY, "hash")}

    def reserve(self):
        """Grow the history copies to the committed count; returns the count."""
        e = self.engine
        count = int(e.m.read(e.buffers["control"], np.uint32)[2])
        if not 0 <= count <= e.contact_capacity:
            raise ValueError("Committed contact count is outside capacity")
        need = max(16, 16 * count)
        for name in HISTORY:
            if self.history[name].size < need:
                self.history[name] = e.m.buffer(count=need // 4)
        return count

    def _pairs(self, committed, count):
        e = self.engine
        for name, copy in self.copies.items():
            yield e.buffers[name], copy, e.buffers[name].size
        for name, copy in self.history.items():
            yield committed[name], copy, committed[name].size if name == "hash" else 16 * count

    def save(self, encoder, committed, count):
        for live, copy, size in self._pairs(committed, count):
            if size:
                encoder.copy_buffer(copy, 0, live, 0, size)

    def restore(self, encoder, committed, count):
        for live, copy, size in self._pairs(committed, count):
            if size:
                encoder.copy_buffer(live, 0, copy, 0, size)


# ---------------------------------------------------------------- the engine


class Engine:
    """One assembly on one backend: upload, step, checkpoint, diagnose."""

    def __init__(self, shapes, spec, positions, cell, shape_ids, scales, *, backend):
        spec.require_supported_path()
        positions = finite_array(positions)
        cell = finite_array(cell, (3,), positive=True)
        if positions.ndim != 2 or positions.shape[1] != 3 or len(positions) < 1:
            raise ValueError("At least one three-dimensional position required")
        open_axes = ~periodic_axes(spec.periodic)
        outside = (positions < 0) | (positions >= cell)
        if np.any(outside[:, open_axes]):
            raise ValueError("Nonperiodic particle centres must lie in [0, cell)")
        self.m = backend
        self.spec = spec
        self.shapes = list(shapes)
        self.n = len(positions)
        self.ids = np.asarray(shape_ids, np.uint32)
        self.scales = np.asarray(scales, float)
        if len(self.ids) != self.n or len(self.scales) != self.n:
            raise ValueError("One shape id and scale per particle required")
        if np.any(self.ids >= len(self.shapes)):
            raise ValueError("Shape id outside the library")
        self.mesh_geometry = spec.contact_geometry == "mesh"
        self.wall_count = len(spec.planes)
        self.viscous = spec.viscous
        self.units = Units.of(self.shapes, spec)
        self.body, inverse_inertia = grain_rows(self.shapes, spec, self.units, self.ids, self.scales)
        box = np.asarray(cell, float) / self.length_unit
        self.sizing = Sizing.plan(spec, self.body, box, self.n)
        check_timestep(self.shapes, spec, self.ids, self.scales)
        self.estimated_bytes = self._working_set()
        if self.estimated_bytes > spec.working_memory_gib * 1024**3:
            raise MemoryError(f"Estimated working set {self.estimated_bytes / 1024**3:.2f} GiB exceeds budget")
        self.step = 0
        self.history = False  # does the next force evaluation carry friction history?
        self.active = 0  # history side the next force evaluation writes
        self.capacity_retries = []
        self.initial_cell = np.asarray(cell, float)
        if spec.boundary_box is not None:
            offsets = np.array([p["offset"] for p in spec.planes])
            self.initial_cell = box_extents(offsets, spec.boundary_box, cell)
        self.buffers = {}
        for name, data in self._geometry().items():
            self._upload(name, data)
        for name, data in self._grain_state(positions, box, inverse_inertia).items():
            self._upload(name, data)
        for name, data in self._wall_rows().items():
            self._upload(name, data)
        for name, (count, stride) in self._work_buffers().items():
            self.buffers[name] = self.m.buffer(count=count, stride=stride)
        self.history_buffers = [self._history_side() for _ in range(2)]
        # Totals per group of 256 records, level by level (gpu.py:322-340).
        self.diag_buffers = self._reduction_levels(self.n, 8)
        self.fabric_buffers = self._reduction_levels(self.pair_capacity, 2)
        self.wall_diag_buffers = self._reduction_levels(self.n, 3 * self.wall_count) if self.wall_count else []
        self.rollback = _Rollback(self, ROLLBACK + (WALL_STATE if self.wall_count else ()))

    # ---- sizes and units, as attributes
    length_unit = property(lambda self: self.units.length)
    time_unit = property(lambda self: self.units.time)
    mass_unit = property(lambda self: self.units.mass)
    force_unit = property(lambda self: self.units.force)
    stress_unit = property(lambda self: self.units.stress)
    energy_unit = property(lambda self: self.units.energy)
    skin = property(lambda self: self.sizing.skin)
    cutoff = property(lambda self: self.sizing.cutoff)
    bins = property(lambda self: self.sizing.bins)
    bin_count = property(lambda self: int(self.sizing.bins.prod()))
    bin_capacity = property(lambda self: self.sizing.bin_capacity)
    pair_capacity = property(lambda self: self.sizing.pairs)
    contact_capacity = property(lambda self: self.sizing.contacts)
    hash_size = property(lambda self: self.sizing.hash_slots)

    def _working_set(self):
        """Device bytes the buffers will take (gpu.py:171-188), plus geometry."""
        z = self.sizing
        total = (244 * z.contacts + 8 * z.hash_slots + 196 * z.pairs
                 + 8 * int(z.bins.prod()) * z.bin_capacity + 1184 * self.n)
        if self.mesh_geometry:
            total += sum(224 * len(s.faces) + 736 + 8 * len(s.samples) for s in self.shapes)
        else:
            for shape in self.shapes:
                shape.validate_sdf()
            total += sum(np.asarray(s.sdf).nbytes + 32 * len(s.samples) for s in self.shapes)
        return total + 160 * self.wall_count * self.n + 256 * self.wall_count

    def _upload(self, name, data):
        self.buffers[name] = self.m.buffer(data)

    def _geometry(self):
        """Static template arrays: the packed meshes (gpu.py:421-463), or the SDF grids
        (gpu.py:427-488), with one-row placeholders for the other geometry's slots."""
        out = {}
        info, samples = sample_arrays(self.shapes, self.length_unit)
        placeholder = lambda name: np.zeros((1, 4), np.uint32 if name in _UINT_ARRAYS else np.float32)  # noqa: E731
        if self.mesh_geometry:
            mesh, self.mesh_cache_receipt = packed_meshes(self.shapes, self.length_unit)
            roots = mesh["meshRoots"]
            limits = (
                (np.any(roots[:, 1] - roots[:, 0] > 65535),
                 "A template BVH exceeds 65,535 nodes (device traversal stack holds 16-bit offsets)"),
                (np.any(roots[:, 2] > 14), "A template BVH is deeper than 14 levels (device traversal stack holds 16 entries)"),
                (len(mesh["meshTriangles"]) // 3 >= 1 << 28, "More than 2^28 triangles in the template library"),
                (np.any(mesh["meshClusterInfo"][:, 1] > 1024),
                 "A template has more than 1,024 sample clusters (about 32k nodes)"),
            )
            for broken, message in limits:
                if broken:
                    raise ValueError(message)
            out.update(mesh)
            out.update(grids=np.zeros((1, 4), np.uint32), gridOrigin=np.zeros((1, 4), np.float32),
                       distance=np.zeros(4, np.float32))
        else:
            for name in ("meshHullInfo", "meshPlanes", "meshRoots", "meshLinks", "meshMin", "meshMax",
                         "meshTriangles", "meshNormals", "meshFrontier", "meshClusterInfo", "meshClusterRanges",
                         "meshClusterMin", "meshClusterMax", "meshSampleIndices"):
                out[name] = placeholder(name)
            self.mesh_cache_receipt = {"status": "sdf"}
            l = self.length_unit
            fields = [np.asarray(s.sdf) for s in self.shapes]
            starts = np.concatenate([[0], np.cumsum([f.size for f in fields])])
            if starts[-1] >= 1 << 32:
                raise ValueError("SDF grids exceed 32-bit indices")
            out["grids"] = np.array([[starts[k], *f.shape] for k, f in enumerate(fields)], np.uint32)
            out["gridOrigin"] = np.array([[*(np.asarray(s.origin) / l), s.spacing / l] for s in self.shapes],
                                         np.float32)
            out["distance"] = np.concatenate([(f / l).ravel() for f in fields]).astype(np.float32)
        out["sampleInfo"], out["sample"] = info, samples
        return out

    def _grain_state(self, positions, box, inverse_inertia):
        """Initial per-grain rows; positions split into FP32 value + FP32 remainder."""
        n = self.n
        zeros = lambda rows=n: np.zeros((rows, 4), np.float32)  # noqa: E731
        normalized = np.asarray(positions, float) / self.length_unit
        remainder = zeros()
        remainder[:, :3] = normalized - normalized.astype(np.float32).astype(float)
        control = np.zeros(16, np.uint32)
        control[4] = 1  # the first force evaluation builds the neighbour list
        rows = {
            "shapeId": self.ids,
            "body": self.body,
            "invI": rows4(inverse_inertia.reshape(-1, 3)),
            "position": rows4(normalized),
            "positionRemainder": remainder,
            "velocity": zeros(),
            "rotation": np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1)),
            "particleStress": zeros(3 * n),
            "diagnosticRows": zeros(8 * n),
            "cellLog": zeros(2),
            "cell": np.array([[*box, 0], [0, 0, 0, 0], [*box, 0]], np.float32),
            "referenceCell": np.array([[*box, 0], [0, 0, 0, self.cutoff]], np.float32),
            "control": control,
        }
        for name in ("momentum", "force", "torque", "referencePosition", "particleDiagnostic", "timestepRows",
                     "accumulated", "referencePositionRemainder", "velocityRemainder", "momentumRemainder",
                     "rotationRemainder", "accumulatedRemainder", *DAMPING_STATE):
            rows[name] = zeros()
        return rows

    def _wall_rows(self):
        """Planes in normalized units (gpu.py:232-285); one-row placeholders without walls."""
        spec, l = self.spec, self.length_unit
        rows = max(1, self.wall_count)
        plane, velocity, material = (np.zeros((rows, 4)) for _ in range(3))
        for k, p in enumerate(spec.planes):
            plane[k] = [*p["normal"], p["offset"] / l]
            velocity[k, :3] = np.array(p["velocity"]) * self.time_unit / l
            material[k] = [p["normal_stiffness"] / spec.normal_stiffness,
                           p["tangential_stiffness"] / spec.normal_stiffness, p["friction"],
                           p.get("normal_viscosity", 0) / (spec.normal_stiffness * self.time_unit)]
        plane32 = plane.astype(np.float32)
        residual = (plane - plane32).astype(np.float32)
        representable = all(np.isfinite(a).all() for a in (plane32, residual, velocity.astype(np.float32),
                                                             material.astype(np.float32)))
        if not representable or np.any(material[: self.wall_count, :2].astype(np.float32) <= 0):
            raise ValueError("Plane data cannot be represented in normalized FP32 units")
        for k, p in enumerate(spec.planes):
            if p.get("normal_viscosity", 0) > 0 and np.float32(material[k, 3]) <= 0:
                raise ValueError("Plane viscosity cannot be represented in normalized FP32 units")
        return {
            "wallPlane": plane32,
            "wallRemainder": residual,
            "wallVelocity": velocity.astype(np.float32),
            "wallMaterial": material.astype(np.float32),
            "wallBoxIds": np.array(spec.boundary_box or [0] * 6, np.uint32),
            "wallAccumulated": np.zeros((max(1, self.wall_count * self.n), 4), np.float32),
            "wallRows": np.zeros((max(1, 3 * self.wall_count * self.n), 4), np.float32),
        }

    def _work_buffers(self):
        """Device-written buffers: name -> (count, stride in bytes)."""
        z, n = self.sizing, self.n
        return {
            "meshPairMask": (z.pairs, 4), "meshPairTransform": (8 * z.pairs, 16), "fabricRows": (2 * z.pairs, 16),
            "binCounts": (self.bin_count, 4), "binParticles": (self.bin_count * z.bin_capacity, 4),
            "pairCounts": (n, 4), "pairOffsets": (n, 4), "pairs": (z.pairs, 8),
            "contactCounts": (z.pairs, 4), "contactOffsets": (z.pairs, 4),
            "heads": (n, 4), "next": (2 * z.contacts, 4),
            "pointDepth": (z.contacts, 16), "contactForce": (z.contacts, 16), "normalLossPower": (z.contacts, 4),
        }

    def _history_side(self):
        """One side of the double-buffered friction history (gpu.py:305-316)."""
        side = {name: self.m.buffer(count=self.contact_capacity, stride=16) for name in HISTORY}
        side["hash"] = self.m.buffer(np.full(self.hash_size, EMPTY, np.uint32))
        return side

    def _reduction_levels(self, records, rows):
        levels = []
        while records > 1:
            records = (records + 255) // 256
            levels.append((records, self.m.buffer(count=records * rows, stride=16)))
        return levels

    # ---- the Config block and the bindings of one dispatch

    def config(self, dt=None, drag=0):
        """Config values in normalized units (gpu.py:493-518)."""
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
            "wallBoxEnabled": int(spec.boundary_box is not None),
        }

    def bindings(self, dt=None, drag=0):
        """Every Resources buffer plus `c`; the active side is written, the other read
        as the previous step's history (gpu.py:519-527)."""
        write, read = self.history_buffers[self.active], self.history_buffers[1 - self.active]
        return {**self.buffers, "c": self.config(dt, drag), **write, "oldMeta": read["meta"],
                "oldSlip": read["slip"], "oldNormal": read["normalArea"], "oldHash": read["hash"]}

    # ---- kernel order

    def force_kernels(self, capture_kinematics=True):
        """(kernel, thread count) of one force evaluation, in order (gpu.py:532-565)."""
        n, p, walls = self.n, self.pair_capacity, self.wall_count
        plan = [("captureContactKinematics", n)] if self.viscous and capture_kinematics else []
        plan += [("clearBins", self.bin_count), ("fillBins", n), ("countPairs", n), ("pairScan", n),
                 ("emitPairs", n)]
        if self.mesh_geometry:
            plan += [("resetMeshContacts", p), ("prepareMeshPairs", p), ("collectMeshFrontiers", 64 * 4096)]
        else:
            plan += [("countContacts", 64 * p), ("contactScan", p), ("contactTotal", 1), ("emitContacts", 64 * p)]
        if walls:
            plan.append(("collectWallContacts", 64 * n * walls))
        plan += [("clearContactStorage", max(n, self.hash_size)), ("resolveContacts", self.contact_capacity),
                 ("buildHistoryHash", self.contact_capacity), ("gather", n)]
        if walls:
            plan.append(("gatherWallRows", n * walls))
        if self.spec.timestep_policy == "active-contact":
            plan.append(("checkContactTimestep", n))
        return plan

    def step_kernels(self):
        """(before, after): the kernels around a timestep's force evaluation (gpu.py:714-726)."""
        n = self.n
        before = [("checkContactTimestep", n)] if self.spec.timestep_policy == "active-contact" else []
        before += [("beginStep", n), ("advanceCell", 1)]
        before += [("advanceWalls", 1)] if self.wall_count else []
        return before + [("checkRebuild", n)], [("finishStep", n)]

    def _encode(self, encoder, plan, bindings):
        for kernel, count in plan:
            self.m.dispatch(kernel, count, encoder, r=bindings)

    def dispatch(self, name, count, encoder, bindings):
        self.m.dispatch(name, count, encoder, r=bindings)

    def forces(self, encoder, bindings, *, capture_kinematics=True):
        self._encode(encoder, self.force_kernels(capture_kinematics), bindings)

    def _flip(self):
        self.active = 1 - self.active
        self.history = True

    def _encode_steps(self, encoder, steps, drag):
        """`steps` timesteps; natively when the adapter offers it."""
        if steps and hasattr(self.m, "encode_steps"):
            self.m.encode_steps(encoder, steps, self.active, self.bindings(drag=drag), self.history_buffers)
            self.active = (self.active + steps) % 2
            self.history = True
            return
        before, after = self.step_kernels()
        for _ in range(steps):
            b = self.bindings(drag=drag)
            self._encode(encoder, before + self.force_kernels() + after, b)
            self._flip()

    def _run(self, plan, check=True):
        """Build one encoder with `plan(encoder)`, submit it, wait, check the word."""
        encoder = self.m.device.create_command_encoder()
        plan(encoder)
        self.m.submit(encoder)
        self.m.device.wait()
        return self.check_errors() if check else None

    def check_errors(self):
        """The control word after a submission (gpu.py:589-595)."""
        ctrl = self.m.read(self.buffers["control"], np.uint32)
        if ctrl[15] != 0:
            raise RuntimeError("Arithmetic materialization mask must remain zero")
        if ctrl[0]:
            raise RejectedStep(ctrl[0])
        return ctrl

    # ---- stepping

    def initialize(self):
        """First force evaluation at dt = 0, without history (gpu.py:580-587)."""
        self._run(lambda e: self.forces(e, self.bindings(dt=0)))
        self._flip()

    def advance(self, steps, *, rate=(0, 0, 0), drag=0, wall_velocities=None):
        """`steps` timesteps under a diagonal log strain rate (1/s), linear drag (1/s)
        and optional new wall velocities (m/s).  A rejected chunk is rolled back; a
        capacity overflow rebuilds the buffers at the grown capacities and retries,
        up to five attempts in all (gpu.py:597-647)."""
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
        attempt = 0
        while True:
            try:
                return self._chunk(steps, rate, drag, wall_velocities)
            except RejectedStep as rejected:
                flags = rejected.flags
                attempt += 1
                if not flags & CAPACITY_FLAGS or flags & ~CAPACITY_FLAGS or attempt == 5:
                    raise
                self._regrow(flags)

    def _regrow(self, flags):
        """Rebuild on the same backend with grown capacities, from the current
        (rolled back) state."""
        settings = grown_settings(self.spec, flags, self.contact_capacity, self.n)
        log = self.capacity_retries + [{"step": self.step, "flags": flags, "new_capacities": settings}]
        state = self.state()
        bigger = Engine(self.shapes, replace(self.spec, **settings), state["position"][:, :3] * self.length_unit,
                        state["cell"][0, :3] * self.length_unit, self.ids, self.scales, backend=self.m)
        bigger.initial_cell = self.initial_cell.copy()
        bigger.restore(state, self.step)
        self.__dict__.update(bigger.__dict__)
        self.rollback.engine = self
        self.capacity_retries = log

    def _chunk(self, steps, rate, drag, wall_velocities):
        """One transaction: save, step, check; restore everything on rejection."""
        if getattr(self, "invalid", False):
            raise RuntimeError("Engine is invalid; restore a checkpoint")
        if steps < 1:
            return
        committed_count = self.rollback.reserve()
        cell = self.m.read(self.buffers["cell"], shape=(3, 4))
        cell_before = cell.copy()
        cell[1, :3] = np.asarray(rate) * self.time_unit  # cell[1] holds the log rates (gpu.py:681)
        walls_before = None
        if wall_velocities is not None:
            walls_before = self.m.read(self.buffers["wallVelocity"], shape=(self.wall_count, 4))
            walls_now = rows4(wall_velocities * self.time_unit / self.length_unit)
        self.buffers["cell"].copy_from_numpy(cell)
        if walls_before is not None:
            self.buffers["wallVelocity"].copy_from_numpy(walls_now)
        side_before, history_before = self.active, self.history
        committed = self.history_buffers[1 - self.active]

        def encode(encoder):
            self.rollback.save(encoder, committed, committed_count)
            if walls_before is not None:
                viscous_planes = any(p.get("normal_viscosity", 0) > 0 for p in self.spec.planes)
                if viscous_planes and not np.array_equal(walls_before, walls_now):
                    # plane damping depends on the wall velocity: re-evaluate at dt = 0
                    self.forces(encoder, self.bindings(dt=0))
                    self._flip()
                else:
                    self.dispatch("refreshWallPower", self.n * self.wall_count, encoder, self.bindings())
            self._encode_steps(encoder, steps, drag)

        try:
            self._run(encode)
        except RuntimeError:
            self._run(lambda e: self.rollback.restore(e, committed, committed_count), check=False)
            self.buffers["cell"].copy_from_numpy(cell_before)
            if walls_before is not None:
                self.buffers["wallVelocity"].copy_from_numpy(walls_before)
            self.active, self.history = side_before, history_before
            raise
        self.step += steps

    # ---- checkpoints

    def state(self):
        """Checkpoint content (gpu.py:893-936): compensated grain state, damping rows,
        wall state and the committed friction history."""
        n, w = self.n, self.wall_count
        rows = {"cell": 3, "cellLog": 2}
        out = {name: self.m.read(self.buffers[name], shape=(rows.get(name, n), 4)) for name in STATE}
        if self.viscous:
            out.update({name: self.m.read(self.buffers[name], shape=(n, 4)) for name in DAMPING_STATE})
        if w:
            wall_rows = {"wallRows": 3 * w * n, "wallAccumulated": w * n}
            out.update({name: self.m.read(self.buffers[name], shape=(wall_rows.get(name, w), 4))
                        for name in WALL_STATE})
        count = int(self.check_errors()[2])
        side = self.history_buffers[1 - self.active]
        out["contacts"] = count
        for name in HISTORY:
            dtype = np.uint32 if name == "meta" else np.float32
            out[name] = self.m.read_prefix(side[name], 4 * count, dtype, shape=(count, 4))
        return out

    def restore(self, state, step):
        """Load a checkpoint (gpu.py:1016-1092): state rows, then the history hash
        rebuilt from the committed contacts, then forces at dt = 0 with that history,
        so the next step continues as the checkpointed run would."""
        if self.viscous:
            for name in DAMPING_STATE:
                finite_array(state[name], (self.n, 4))
        if self.wall_count:
            self._check_wall_state(state)
            for name in WALL_STATE:
                self.buffers[name].copy_from_numpy(np.asarray(state[name], np.float32))
        for name in DAMPING_STATE:
            rows = state[name] if self.viscous else np.zeros((self.n, 4), np.float32)
            self.buffers[name].copy_from_numpy(rows)
        for name in STATE:
            self.buffers[name].copy_from_numpy(np.asarray(state[name], np.float32))
        count = int(state["contacts"])
        if not 0 <= count <= self.contact_capacity:
            raise ValueError("Checkpoint contact count is outside capacity")
        side = self.history_buffers[1]
        for name in HISTORY:
            rows = np.asarray(state[name], dtype=np.uint32 if name == "meta" else np.float32)
            if rows.shape != (count, 4):
                raise ValueError("Checkpoint contact history shape mismatch")
            self.m.write_prefix(side[name], rows)
        control = np.zeros(16, np.uint32)
        control[2], control[4] = count, 1
        self.buffers["control"].copy_from_numpy(control)
        # Side 1 holds the committed history: index it ...
        self.active, self.history = 1, True
        b = self.bindings(dt=0)
        self._run(lambda e: self._encode(e, [("clearContactStorage", max(self.n, self.hash_size)),
                                              ("buildHistoryHash", self.contact_capacity)], b))
        # ... then evaluate forces at dt = 0 into side 0, reading side 1.
        self.active = 0
        self.buffers["control"].copy_from_numpy(control)
        self._run(lambda e: self.forces(e, self.bindings(dt=0), capture_kinematics=False))
        self.active = 1
        self.step = step

    def _check_wall_state(self, state):
        for name in WALL_STATE:
            if finite_array(state[name]).shape != (self.buffers[name].size // 16, 4):
                raise ValueError("Checkpoint wall state shape mismatch")
        normals = np.array([plane["normal"] for plane in self.spec.planes], np.float32)
        if not np.array_equal(state["wallPlane"][:, :3], normals):
            raise ValueError("Checkpoint plane normal identity mismatch")
        if self.spec.boundary_box is not None:
            offsets = (np.asarray(state["wallPlane"][:, 3], float) + state["wallRemainder"][:, 3]) * self.length_unit
            box_extents(offsets, self.spec.boundary_box, np.asarray(state["cell"][0, :3], float) * self.length_unit)

    def set_kinematics(self, *, rotation=None, velocity=None, momentum=None):
        """Overwrite rotations (xyzw), velocities or momenta (normalized units) -- for
        generated workloads; no reference counterpart."""
        for name, rows in (("rotation", rotation), ("velocity", velocity), ("momentum", momentum)):
            if rows is not None:
                self.buffers[name].copy_from_numpy(rows4(rows))

    # ---- reading results

    def read(self, name, dtype=np.float32, rows=4):
        """A whole buffer as records of `rows` values (rows = 1: flat)."""
        a = self.m.read(self.buffers[name], dtype)
        return a.reshape(-1, rows) if rows > 1 else a

    def contacts(self):
        """(meta, pointDepth, normalArea) rows of the last completed step's contacts."""
        count = int(self.m.read(self.buffers["control"], np.uint32)[2])
        side = self.history_buffers[1 - self.active]
        prefix = lambda buffer, dtype: self.m.read_prefix(buffer, 4 * count, dtype, shape=(count, 4))  # noqa: E731
        return (prefix(side["meta"], np.uint32), prefix(self.buffers["pointDepth"], np.float32),
                prefix(side["normalArea"], np.float32))

    def walls(self):
        """Per-plane totals of the last force evaluation in SI units, summed on the host
        in FP64 from the per-grain wall rows (the reference reduces the same rows on the
        device, gpu.py:774-835)."""
        if not self.wall_count:
            return []
        w, n = self.wall_count, self.n
        rows = self.m.read(self.buffers["wallRows"], shape=(w, n, 3, 4)).astype(np.float64)
        counts = self.m.read(self.buffers["wallRows"], np.uint32, shape=(w, n, 3, 4))[:, :, 0, 3].sum(1)
        plane = self.m.read(self.buffers["wallPlane"], shape=(w, 4)).astype(np.float64)
        plane[:, 3] += self.m.read(self.buffers["wallRemainder"], shape=(w, 4))[:, 3]
        total = rows.sum(1)
        u = self.units
        return [{
            "normal": plane[k, :3].tolist(),
            "offset": float(plane[k, 3] * u.length),
            "reaction": (total[k, 0, :3] * u.force).tolist(),
            "reaction_moment_origin": (total[k, 1, :3] * u.force * u.length).tolist(),
            "power": float(total[k, 1, 3] * u.force * u.length / u.time),
            "work": float(total[k, 2, 0] * u.energy),
            "active_surface_contacts": int(counts[k]),
        } for k in range(w)]

    def diagnostics(self):
        """Whole-assembly evidence of the last completed step in SI units (gpu.py:746-891):
        stress, energies, fabric, coordination and plane reactions, reduced on the
        device by the diagnostic.slang group tree."""
        totals, fabric_totals, wall_totals = self._reduce_rows()
        ctrl = self.check_errors()
        cell = self.m.read(self.buffers["cell"], shape=(-1, 4))[:3]
        domain = cell[0, :3].astype(float) * self.length_unit
        box, planes = domain.copy(), []
        if self.wall_count:
            planes, offsets = self._plane_summary(wall_totals)
            if self.spec.boundary_box is not None:
                box = box_extents(offsets, self.spec.boundary_box, domain)
        return summarize(self, totals, fabric_totals, ctrl, cell, box, domain, planes)

    def _reduce_rows(self):
        encoder = self.m.device.create_command_encoder()
        self.dispatch("diagnosticRows", self.n, encoder, self.bindings())
        grains = self._tree(encoder, "reduceRows", self.buffers["diagnosticRows"], self.n, self.diag_buffers)
        self.dispatch("fabricRows", self.pair_capacity, encoder, self.bindings())
        pairs = self._tree(encoder, "reduceFabric", self.buffers["fabricRows"], self.pair_capacity,
                           self.fabric_buffers)
        walls = self._tree(encoder, "reduceWallRows", self.buffers["wallRows"], self.n, self.wall_diag_buffers)
        self.m.submit(encoder)
        self.m.device.wait()
        read = lambda buffer, k: self.m.read(buffer, shape=(-1, 4))[:k]  # noqa: E731
        return read(grains, 8), read(pairs, 2), (read(walls, 3 * self.wall_count) if self.wall_count else None)

    def _tree(self, encoder, name, source, records, levels):
        """Encode the group-total levels (diagnostic.slang); returns the last level."""
        kernel = self.m.kernel(name, "diagnostic.slang")
        walls = self.wall_count if name == "reduceWallRows" else 1
        for groups, destination in levels:
            extra = {"outputCount": groups} if name == "reduceWallRows" else {}
            kernel.dispatch(thread_count=[groups * 256, walls, 1], command_encoder=encoder, N=records,
                            input=source, output=destination, **extra)
            source, records = destination, groups
        return source

    def _plane_summary(self, wall_totals):
        w, u = self.wall_count, self.units
        data = wall_totals.reshape(w, 3, 4)
        counts = data[:, 0, 3].copy().view(np.uint32)
        plane = self.m.read(self.buffers["wallPlane"], shape=(-1, 4))[:w].astype(float)
        remainder = self.m.read(self.buffers["wallRemainder"], shape=(-1, 4))[:w].astype(float)
        velocity = self.m.read(self.buffers["wallVelocity"], shape=(-1, 4))[:w].astype(float)
        offsets = (plane[:, 3] + remainder[:, 3]) * u.length
        planes = [{
            "index": k,
            "normal": plane[k, :3].tolist(),
            "offset": float(offsets[k]),
            "velocity": (velocity[k, :3] * u.length / u.time).tolist(),
            "reaction": (data[k, 0, :3].astype(float) * u.force).tolist(),
            "reaction_moment_origin": (data[k, 1, :3].astype(float) * u.energy).tolist(),
            "active_surface_contacts": int(counts[k]),
            "power": float(data[k, 1, 3] * u.energy / u.time),
            "work": float(data[k, 2, 0] * u.energy),
        } for k in range(w)]
        return planes, offsets


def summarize(engine, totals, fabric_totals, ctrl, cell, box, domain, planes):
    """SI summary of the reduced rows (gpu.py:836-891).

    totals rows: 0-2 contact stress (Love-Weber sum), 3-5 kinetic stress, .w of rows
    0-7 kinetic energy, elastic energy proxy, unbalanced force, force scale, max
    overlap fraction, friction dissipation of the step; row 6 cumulative friction,
    drag, cell and wall work; row 7 normal-contact dissipation and power.
    fabric rows: n (x) n components and the active grain-pair count.
    """
    u, e = engine.units, engine
    pairs = round(float(fabric_totals[1, 2]))
    f = fabric_totals
    fabric = np.array([[f[0, 0], f[0, 3], f[1, 0]], [f[0, 3], f[0, 1], f[1, 1]], [f[1, 0], f[1, 1], f[0, 2]]],
                      float) / max(pairs, 1)
    volume = float(box.prod())
    contact = totals[:3, :3].astype(float) * u.energy / volume
    kinetic = totals[3:6, :3].astype(float) * u.energy / volume
    stress = contact + kinetic
    symmetric = 0.5 * (stress + stress.T)
    lateral = 0.5 * (symmetric[0, 0] + symmetric[1, 1])
    solid = sum(e.shapes[int(i)].volume * s**3 for i, s in zip(e.ids, e.scales))
    energy = lambda value: float(value * u.energy)  # noqa: E731
    return {
        "step": e.step,
        "elapsed_time": e.step * e.spec.timestep,
        "cell": box.tolist(),
        "domain_cell": domain.tolist(),
        "walls": planes,
        "stress": stress.tolist(),
        "contact_stress": contact.tolist(),
        "kinetic_stress": kinetic.tolist(),
        "stress_antisymmetry": float(np.linalg.norm(stress - stress.T) / max(np.linalg.norm(stress), 1e-30)),
        "p": float(np.trace(symmetric) / 3),
        "q": float(symmetric[2, 2] - lateral),
        "axial_strain": float(1 - box[2] / e.initial_cell[2]),
        "volumetric_strain": float(np.sum(1 - box / e.initial_cell)),
        "volume_ratio": float(volume / np.prod(e.initial_cell)),
        "log_strain": (-np.log(box / e.initial_cell)).tolist(),
        "void_ratio": float(volume / solid - 1),
        "kinetic_energy": energy(totals[0, 3]),
        "elastic_energy_proxy": energy(totals[1, 3]),
        "contact_normal": e.spec.contact_normal,
        "elastic_energy_interpretation": "existing nodal spring energy proxy",
        "unbalanced_force": float(totals[2, 3] / max(totals[3, 3], 1e-30)),
        "max_overlap_fraction": float(totals[4, 3]),
        "friction_dissipation_last_step": energy(totals[5, 3]),
        "candidate_pairs": int(ctrl[1]),
        "active_surface_contacts": int(ctrl[2]),
        "active_grain_pairs": pairs,
        "coordination": 2 * pairs / e.n,
        "fabric": fabric.tolist(),
        "sliding_fraction": float(ctrl[6] / max(ctrl[2], 1)),
        "cell_log_rate": (cell[1, :3].astype(float) / u.time).tolist(),
        "friction_dissipation": energy(totals[6, 0]),
        "drag_dissipation": energy(totals[6, 1]),
        "cell_work": energy(totals[6, 2]),
        "wall_work": energy(totals[6, 3]),
        "normal_contact_dissipation": energy(totals[7, 0]),
        "normal_contact_power": float(totals[7, 1] * u.energy / u.time),
    }
