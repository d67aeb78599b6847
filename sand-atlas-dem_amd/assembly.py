"""Seeded synthetic workloads for the mesh step: template libraries and packings.

Everything derives from one integer seed through `numpy.random.default_rng`
streams split per purpose (templates, sizes, positions, orientations, velocities),
so a workload is reproducible on any host.  The reference has no generator for
these libraries or packings; its own assemblies come from `initial_assembly`
(`simulation.py:26-56`) and `benchmark.py:15-79`.  Orientations are normalized
4-D Gaussians as in `simulation.py:75-78`.
"""

from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import geometry
from .spec import Experiment


def streams(seed):
    root = np.random.SeedSequence(seed)
    names = ("templates", "sizes", "positions", "orientations", "velocities", "shapes", "spare")
    return {name: np.random.default_rng(s) for name, s in zip(names, root.spawn(len(names)))}


def icosphere_library(count, *, subdivisions, sigma, diameter, nodes, rng, seed):
    """`count` perturbed icospheres; `subdivisions` is an int or a sequence to draw from."""
    bases = {}
    shapes = []
    for k in range(count):
        s = int(subdivisions if np.isscalar(subdivisions) else rng.choice(subdivisions))
        if s not in bases:
            bases[s] = geometry.icosphere(s)
        v, f = bases[s]
        shapes.append(
            geometry.perturbed_template(
                f"ico{s}-{k}", v, f, sigma=sigma, diameter=diameter, nodes=nodes, rng=rng, seed=seed + k
            )
        )
    return shapes


def uneven_library(count, *, tri_range=(320, 5120), node_range=(256, 4096), surface="harmonic", sigma=0.15,
                   diameter, rng, seed):
    """Templates with log-uniform triangle counts and node counts proportional to area.

    surface "harmonic": sand-like spherical-harmonic grains (resolution-independent
    shape); "lognormal": per-vertex lognormal(0, sigma) radii.
    """
    tris = np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]), count)).astype(int)
    shapes = []
    descriptors = geometry.harmonic_descriptors()
    for k, t in enumerate(tris):
        v, f = geometry.hull_sphere(t)
        if surface == "harmonic":
            shapes.append(geometry.harmonic_template(f"sh{t}-{k}", v, f, descriptors=descriptors, diameter=diameter,
                                                     nodes=16, rng=rng, seed=seed + k))
        elif surface == "lognormal":
            shapes.append(geometry.perturbed_template(f"hull{t}-{k}", v, f, sigma=sigma, diameter=diameter, nodes=16,
                                                      rng=rng, seed=seed + k))
        else:
            raise ValueError(f"unknown surface {surface!r}")
    areas = np.array([s.areas.sum() for s in shapes])
    lo, hi = areas.min(), areas.max()
    counts = node_range[0] + (node_range[1] - node_range[0]) * (areas - lo) / max(hi - lo, 1e-30)
    out = []
    for k, (s, n) in enumerate(zip(shapes, counts)):
        out.append(geometry.make_shape(s.id, s.vertices, s.faces, nodes=int(round(n)), seed=seed + 7919 + k, provenance=s.provenance))
    return out


def sphere_packing(radii, box, rng, *, iterations=400, tolerance=1e-3, walled=False, clearance=None):
    """Overlap-relaxed sphere packing (growth from half size), periodic or walled.

    Radii grow from 0.5x to 1x over the first half of the iterations while pairwise
    overlaps are pushed apart symmetrically; stops once the largest overlap is below
    `tolerance` times the mean diameter.  `walled` (a bool, or one per axis) keeps
    every centre inside [c, box - c] on the walled axes instead of wrapping, with c
    the sphere radius or the given `clearance` (e.g. the grain's bounding radius),
    scaled like the radii.
    """
    from scipy.spatial import cKDTree

    n = len(radii)
    x = rng.uniform(0, box, (n, 3))
    mean = 2 * radii.mean()
    grow = max(1, iterations // 2)

    reach = radii if clearance is None else np.asarray(clearance, float)
    walls = np.broadcast_to(np.asarray(walled, bool), (3,))
    boxsize = None if walls.all() else (box if not walls.any() else np.where(walls, 1e9, box))

    def confine(x, scale):
        c = reach[:, None] * scale
        return np.where(walls, np.clip(x, c, box - c), np.mod(x, box))

    for it in range(iterations):
        scale = min(1.0, 0.5 + 0.5 * (it + 1) / grow)
        r = radii * scale
        x = confine(x, scale)
        tree = cKDTree(x, boxsize=boxsize)
        pairs = tree.query_pairs(2 * r.max(), output_type="ndarray")
        if len(pairs) == 0:
            if scale >= 1:
                break
            continue
        d = x[pairs[:, 1]] - x[pairs[:, 0]]
        d -= np.where(walls, 0.0, np.round(d / box) * box)
        dist = np.linalg.norm(d, axis=1)
        overlap = r[pairs[:, 0]] + r[pairs[:, 1]] - dist
        hit = overlap > 0
        if scale >= 1 and (not hit.any() or overlap[hit].max() < tolerance * mean):
            break
        pairs, d, dist, overlap = pairs[hit], d[hit], dist[hit], overlap[hit]
        u = d / np.maximum(dist, 1e-12)[:, None]
        push = 0.5 * overlap[:, None] * u
        move = np.zeros_like(x)
        np.add.at(move, pairs[:, 0], -push)
        np.add.at(move, pairs[:, 1], push)
        x = confine(x + move, scale)
    return confine(x, 1.0)


@dataclass
class Workload:
    name: str
    shapes: list
    spec: Experiment
    positions: np.ndarray  # metres
    cell: np.ndarray  # metres
    ids: np.ndarray
    scales: np.ndarray
    rotations: np.ndarray  # xyzw, body to world
    velocities: np.ndarray  # m/s
    rate: tuple  # imposed diagonal log strain rate, 1/s
    steps: int

    def engine(self, backend):
        from .engine import Engine

        e = Engine(self.shapes, self.spec, self.positions, self.cell, self.ids, self.scales, backend=backend)
        e.set_kinematics(rotation=self.rotations, velocity=self.velocities * e.time_unit / e.length_unit)
        return e


def random_rotations(n, rng):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1)[:, None]


def packed_workload(name, shapes, n, phi, *, seed, sphere_fraction=None, size_range=(0.8, 1.2),
                    velocity_sigma=0.0, rate=(0, 0, 0), steps=1000, spec_overrides=None, rng=None,
                    ids=None, scales=None):
    """Grains on an overlap-relaxed sphere packing at solids fraction `phi`.

    Contact spheres have the grains' equivalent diameters scaled so the sphere
    packing reaches `sphere_fraction` (default phi, i.e. equivalent spheres), which
    sets how closely neighbouring irregular grains approach.
    """
    rng = rng or streams(seed)
    if ids is None:
        ids = rng["shapes"].integers(0, len(shapes), n).astype(np.uint32)
    if scales is None:
        scales = rng["sizes"].uniform(*size_range, n)
    d = np.array([shapes[i].equivalent_diameter for i in ids]) * scales
    volume = float(np.sum(np.pi / 6 * d**3))
    box = (volume / phi) ** (1 / 3)
    alpha = ((sphere_fraction or phi) / phi) ** (1 / 3)
    x = sphere_packing(0.5 * alpha * d, box, rng["positions"])
    q = random_rotations(n, rng["orientations"])
    v = rng["velocities"].normal(0, velocity_sigma, (n, 3)) if velocity_sigma else np.zeros((n, 3))
    spec = Experiment(**dict(dict(contact_geometry="mesh"), particles=n, **(spec_overrides or {})))
    return Workload(name, shapes, spec, x, np.full(3, box), ids, scales, q, v, tuple(rate), steps)


def box_planes(low, high, *, stiffness, friction, viscosity=0.0, velocities=None):
    """Six inward axis planes xmin,xmax,ymin,ymax,zmin,zmax bounding [low, high]
    (the reference `boundary_box` order, `boundaries.py:14-16, 45-55`)."""
    planes = []
    for axis in range(3):
        for side, sign in ((low, 1.0), (high, -1.0)):
            normal = [0.0, 0.0, 0.0]
            normal[axis] = sign
            plane = dict(normal=normal, offset=sign * float(side[axis]), velocity=[0.0, 0.0, 0.0],
                         normal_stiffness=stiffness, tangential_stiffness=0.5 * stiffness, friction=friction)
            if viscosity:
                plane["normal_viscosity"] = viscosity
            planes.append(plane)
    for index, v in (velocities or {}).items():
        planes[index]["velocity"] = [float(c) for c in v]
    return planes


def walled_workload(name, shapes, n, phi, *, seed, sphere_fraction=None, size_range=(0.8, 1.2), axial_rate=0.5,
                    margin=0.05, steps=1000, spec_overrides=None, rng=None, ids=None, scales=None, viscosity=0.0):
    """Grains packed inside a box of six planes in a nonperiodic cell.

    The top plane (zmax) translates down at `axial_rate` times the box height per
    second; the other five are fixed.  The cell is the box plus `margin` of its
    side on every face, so grains stay inside the computational domain.
    """
    rng = rng or streams(seed)
    if ids is None:
        ids = rng["shapes"].integers(0, len(shapes), n).astype(np.uint32)
    if scales is None:
        scales = rng["sizes"].uniform(*size_range, n)
    d = np.array([shapes[i].equivalent_diameter for i in ids]) * scales
    side = (float(np.sum(np.pi / 6 * d**3)) / phi) ** (1 / 3)
    alpha = ((sphere_fraction or phi) / phi) ** (1 / 3)
    # Centres sit between the contact-sphere and bounding radii from the walls: some
    # grains start touching a wall, with overlaps well below the pair overlaps.
    bounding = 0.5 * (0.5 * alpha * d + np.array([shapes[i].radius for i in ids]) * scales)
    x = sphere_packing(0.5 * alpha * d, side, rng["positions"], walled=True, clearance=bounding)
    pad = margin * side
    low, high = np.full(3, pad), np.full(3, pad + side)
    overrides = dict(spec_overrides or {})
    planes = box_planes(low, high, stiffness=overrides["normal_stiffness"], friction=overrides.get("friction", 0.5),
                        viscosity=viscosity, velocities={5: (0.0, 0.0, -axial_rate * side)})
    spec = Experiment(contact_geometry="mesh", particles=n, periodic=False, planes=planes,
                      boundary_box=[0, 1, 2, 3, 4, 5], **overrides)
    q = random_rotations(n, rng["orientations"])
    return Workload(name, shapes, spec, x + pad, np.full(3, side + 2 * pad), ids, scales, q, np.zeros((n, 3)),
                    (0, 0, 0), steps)


def layer_workload(name, shapes, n, phi, *, seed, sphere_fraction=None, size_range=(0.8, 1.2), axial_rate=0.5,
                   margin=0.05, steps=1000, spec_overrides=None, rng=None, ids=None, scales=None, viscosity=0.0):
    """Grains between two z planes in a cell periodic in x and y (per-axis periodicity).

    The top plane translates down at `axial_rate` times the layer height per second;
    the bottom plane is fixed.  Along z the cell is the layer plus `margin` of its
    height on each side.
    """
    rng = rng or streams(seed)
    if ids is None:
        ids = rng["shapes"].integers(0, len(shapes), n).astype(np.uint32)
    if scales is None:
        scales = rng["sizes"].uniform(*size_range, n)
    d = np.array([shapes[i].equivalent_diameter for i in ids]) * scales
    side = (float(np.sum(np.pi / 6 * d**3)) / phi) ** (1 / 3)
    alpha = ((sphere_fraction or phi) / phi) ** (1 / 3)
    bounding = 0.5 * (0.5 * alpha * d + np.array([shapes[i].radius for i in ids]) * scales)
    x = sphere_packing(0.5 * alpha * d, side, rng["positions"], walled=(False, False, True), clearance=bounding)
    pad = margin * side
    overrides = dict(spec_overrides or {})
    kn, mu = overrides["normal_stiffness"], overrides.get("friction", 0.5)
    planes = []
    for normal, offset, velocity in (((0.0, 0.0, 1.0), pad, (0.0, 0.0, 0.0)),
                                     ((0.0, 0.0, -1.0), -(pad + side), (0.0, 0.0, -axial_rate * side))):
        plane = dict(normal=list(normal), offset=offset, velocity=list(velocity), normal_stiffness=kn,
                     tangential_stiffness=0.5 * kn, friction=mu)
        if viscosity:
            plane["normal_viscosity"] = viscosity
        planes.append(plane)
    spec = Experiment(contact_geometry="mesh", particles=n, periodic=(True, True, False), planes=planes, **overrides)
    q = random_rotations(n, rng["orientations"])
    return Workload(name, shapes, spec, x + np.array([0.0, 0.0, pad]), np.array([side, side, side + 2 * pad]), ids,
                    scales, q, np.zeros((n, 3)), (0, 0, 0), steps)


def conservative_timestep(shapes, ids, scales, spec):
    """Largest dt the conservative preflight accepts (`gpu.py:370-413`), rounded down."""
    limits = []
    for i in np.unique(ids):
        sh = shapes[int(i)]
        for s in (scales[ids == i].min(), scales[ids == i].max()):
            mass = spec.density * sh.volume * s**3
            stiffness = (spec.normal_stiffness + spec.tangential_stiffness) * sh.areas.sum() * s * s
            rotation = np.linalg.eigvalsh(sh.inertia * spec.density * s**5).min() / (stiffness * (sh.radius * s) ** 2)
            limits.append(0.1 * np.sqrt(min(mass / stiffness, rotation)))
    dt = min(limits)
    exponent = np.floor(np.log10(dt))
    return float(np.floor(dt / 10**exponent * 100) / 100 * 10**exponent)


RELAXED = Path(__file__).resolve().parent / "data"


def relaxed_path(config, n, seed):
    return RELAXED / f"relaxed_{config}_{n}_{seed}.npz"


def with_relaxed_state(w, path):
    """The workload with the poses of a relaxed fixture (written by `relax.py`) and the
    reference's default overlap gate (max_overlap_fraction 0.01, `spec.py:44`)."""
    from dataclasses import replace

    from .relax import identity

    data = np.load(path, allow_pickle=False)
    if str(data["identity"]) != identity(w):
        raise ValueError(f"{path.name} was relaxed from a different assembly")
    spec = replace(w.spec, max_overlap_fraction=0.01)
    return replace(w, spec=spec, positions=data["positions"].astype(np.float64), cell=data["cell"].astype(np.float64),
                   rotations=data["rotations"].astype(np.float64), velocities=np.zeros_like(w.velocities))


def config_workload(config, *, seed=2024, n=None, templates=None, nodes=None, steps=None, resolution=None,
                    relaxed=False):
    """The named benchmark configurations (sizes overridable for tests).

    small  : 1,000 icospheres (subdiv 2, sigma 0.1), 512 nodes, phi 0.30, v ~ N(0, 0.01 m/s)
    dense  : 10,000 grains from 64 perturbed icospheres (subdiv 2-4, sigma 0.15), 2,048
             nodes, phi 0.60, isotropic compression at 0.5/s
    uneven : 50,000 grains from 424 sand-like spherical-harmonic templates (320-5,120
             triangles log-uniform, 256-4,096 nodes by area), phi 0.60
    concave: 1,000 grains from 8 strongly non-convex low-degree harmonic templates
             (1,280 triangles, 1,024 nodes), phi 0.55 -- loose hull bounds
    walls  : 10,000 grains from the dense library at phi 0.58 in a nonperiodic box
             of six planes, the top plane moving down at 0.5 box heights per second
    layer  : the same grains between two z planes in a cell periodic in x and y, the
             top plane moving down at 0.5 layer heights per second
    sdf    : the dense assembly with the reference's default contact geometry, SDF
             grids (resolution 64: 64 points across the largest extent, 3 of padding)

    relaxed=True replaces the sphere-packing poses by the committed relaxed state of
    the same assembly (largest overlap below 1% of a diameter) and raises if there is
    none for these sizes and seed.
    """
    if relaxed:
        w = config_workload(config, seed=seed, n=n, templates=templates, nodes=nodes, steps=steps,
                            resolution=resolution)
        path = relaxed_path(config, len(w.ids), seed)
        if not path.exists():
            raise FileNotFoundError(f"no relaxed state {path.name}; write it with relax.py on a GPU")
        return with_relaxed_state(w, path)
    rng = streams(seed)
    d0 = 3e-4
    base = dict(friction=0.5, density=2650.0, tangential_stiffness=0.5)
    # kn* = E / ((1 - nu^2) d) with E = 70 GPa, nu = 0.2 (traction per unit penetration).
    kn = 70e9 / ((1 - 0.2**2) * d0)
    base.update(normal_stiffness=kn, tangential_stiffness=0.5 * kn)
    if config == "small":
        n = n or 1000
        shapes = icosphere_library(templates or 16, subdivisions=2, sigma=0.1, diameter=d0,
                                   nodes=nodes or 512, rng=rng["templates"], seed=seed)
        phi, sphere_fraction, vsig, rate = 0.30, 0.30, 0.01, (0, 0, 0)
        steps = steps or 1000
        cpp = 256
    elif config == "dense":
        n = n or 10_000
        shapes = icosphere_library(templates or 64, subdivisions=(2, 3, 4), sigma=0.15, diameter=d0,
                                   nodes=nodes or 2048, rng=rng["templates"], seed=seed)
        phi, sphere_fraction, vsig, rate = 0.60, 0.62, 0.0, (-0.5, -0.5, -0.5)
        steps = steps or 2000
        cpp = 512  # ~190 node contacts per grain at phi 0.60: headroom over the default 256
    elif config == "sdf":
        n = n or 10_000
        shapes = icosphere_library(templates or 64, subdivisions=(2, 3, 4), sigma=0.15, diameter=d0,
                                   nodes=nodes or 2048, rng=rng["templates"], seed=seed)
        shapes = [geometry.with_sdf(shape, resolution or 64) for shape in shapes]
        phi, sphere_fraction, vsig, rate = 0.60, 0.62, 0.0, (-0.5, -0.5, -0.5)
        steps = steps or 2000
        cpp = 512
    elif config == "uneven":
        n = n or 50_000
        shapes = uneven_library(templates or 424, diameter=d0, rng=rng["templates"], seed=seed)
        phi, sphere_fraction, vsig, rate = 0.60, 0.62, 0.0, (0, 0, 0)
        steps = steps or 1000
        cpp = 512
    elif config == "concave":
        n = n or 1000
        rngt = rng["templates"]
        descriptors = geometry.harmonic_descriptors(d2=0.3, top=6, decay=1.0)
        shapes = []
        for k in range(templates or 8):
            v, f = geometry.icosphere(3)
            shapes.append(geometry.harmonic_template(f"concave-{k}", v, f, descriptors=descriptors, diameter=d0,
                                                     nodes=nodes or 1024, rng=rngt, seed=seed + k))
        phi, sphere_fraction, vsig, rate = 0.55, 0.57, 0.0, (0, 0, 0)
        steps = steps or 1000
        cpp = 512
    elif config in ("walls", "layer"):
        n = n or 10_000
        shapes = icosphere_library(templates or 64, subdivisions=(2, 3, 4), sigma=0.15, diameter=d0,
                                   nodes=nodes or 2048, rng=rng["templates"], seed=seed)
        phi, sphere_fraction = 0.58, 0.60
        steps = steps or 5000
        cpp = 512
    else:
        raise ValueError(f"unknown workload {config!r}")
    ids = rng["shapes"].integers(0, len(shapes), n).astype(np.uint32)
    probe = Experiment(contact_geometry="mesh", particles=n, **base)
    scales = rng["sizes"].uniform(0.8, 1.2, n)
    dt = conservative_timestep(shapes, ids, scales, probe)
    overrides = dict(base, timestep=dt, max_overlap_fraction=0.5, contacts_per_particle=cpp)
    if config == "sdf":
        overrides["contact_geometry"] = "sdf"
    if config == "layer":
        return layer_workload(config, shapes, n, phi, seed=seed, sphere_fraction=sphere_fraction, steps=steps,
                              spec_overrides=overrides, rng=rng, ids=ids, scales=scales)
    if config == "walls":
        return walled_workload(config, shapes, n, phi, seed=seed, sphere_fraction=sphere_fraction, steps=steps,
                               spec_overrides=overrides, rng=rng, ids=ids, scales=scales)
    return packed_workload(config, shapes, n, phi, seed=seed, sphere_fraction=sphere_fraction,
                           velocity_sigma=vsig, rate=rate, steps=steps, spec_overrides=overrides, rng=rng,
                           ids=ids, scales=scales)
