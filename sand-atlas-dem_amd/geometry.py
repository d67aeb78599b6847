"""Grain templates, surface quadrature nodes and the packed static mesh layout.

The packed layout is the one the reference's mesh path binds
(`src/sand_atlas_dem/mesh.py:10-222`, bound by `gpu.py:421-488`): a median-split
triangle BVH with stackless escape links, feature pseudonormals (face, three edges,
three angle-weighted vertices; Baerentzen & Aanaes 2005), outward-rounded convex-hull
planes in six dominant-axis groups, a preorder sample-cluster tree (<=64 nodes per
leaf) and a 32-subtree frontier.  Every array keeps the reference's row format so a
packed library is interchangeable with the reference's own buffers.

Template generation (icospheres, Fibonacci-hull spheres, lognormal radial noise) and
mass properties are this package's own: the reference obtains them from trimesh
(`geometry.py:166-259`), which is not a dependency here.
"""

import ctypes as ct
import os
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


@dataclass
class Shape:
    """One grain template in metres, centred at its centre of mass.

    Mirrors the reference `Shape` (`geometry.py:17-31`).  The SDF-geometry field
    (`sdf`, `origin`, `spacing`: values in metres on the grid origin + spacing *
    index, negative inside) is optional here and made by `with_sdf`; the mesh path
    does not read it.
    """

    id: str
    vertices: np.ndarray
    faces: np.ndarray
    samples: np.ndarray
    areas: np.ndarray
    volume: float
    inertia: np.ndarray  # unit density; multiply by kg/m^3
    radius: float
    provenance: dict = field(default_factory=dict)
    sdf: np.ndarray = None
    origin: np.ndarray = None
    spacing: float = 0.0

    def validate(self):
        for a in (self.vertices, self.samples, self.areas, self.inertia):
            if not np.isfinite(a).all():
                raise ValueError("Nonfinite shape data")
        if self.volume <= 0 or self.radius <= 0 or np.any(self.areas <= 0):
            raise ValueError("Invalid shape measure")
        if len(self.samples) != len(self.areas) or self.samples.shape[1:] != (3,):
            raise ValueError("Invalid surface samples")
        if np.linalg.eigvalsh(self.inertia).min() <= 0:
            raise ValueError("Inertia is not positive definite")

    def validate_sdf(self):
        """The reference's field checks (`geometry.py:32-74`): a finite 3-D grid at
        least 4 wide, negative somewhere and positive somewhere, with an exterior
        boundary layer."""
        if self.sdf is None:
            raise ValueError(f"Shape {self.id} has no SDF grid (make one with geometry.with_sdf)")
        sdf = np.asarray(self.sdf)
        if not (np.isfinite(sdf).all() and np.isfinite(self.origin).all()):
            raise ValueError("Nonfinite shape data")
        if sdf.ndim != 3 or min(sdf.shape) < 4 or not (sdf.min() < 0 < sdf.max()):
            raise ValueError("SDF must be a 3D signed field, negative inside")
        if not self.spacing > 0:
            raise ValueError("Invalid shape measure")
        boundary = np.concatenate([sdf[0].ravel(), sdf[-1].ravel(), sdf[:, 0].ravel(), sdf[:, -1].ravel(),
                                   sdf[:, :, 0].ravel(), sdf[:, :, -1].ravel()])
        if np.any(boundary <= 0):
            raise ValueError("SDF exterior padding is not outside the grain")

    def query(self, points):
        """Trilinear value and exact piecewise gradient of the grid (FP64;
        `geometry.py:76-103`): +inf outside the grid."""
        p = np.atleast_2d(np.asarray(points, float))
        g = (p - self.origin) / self.spacing
        ijk = np.floor(g).astype(np.int64)
        valid = np.all((ijk >= 0) & (ijk < np.array(self.sdf.shape) - 1), axis=1)
        values = np.full(len(p), np.inf)
        gradients = np.zeros((len(p), 3))
        if np.any(valid):
            base, t = ijk[valid], g[valid] - ijk[valid]
            v, grad = np.zeros(len(base)), np.zeros((len(base), 3))
            for bits in np.ndindex(2, 2, 2):
                w = np.where(np.array(bits, bool), t, 1 - t)
                f = self.sdf[tuple((base + np.array(bits)).T)].astype(float)
                v += f * w.prod(axis=1)
                for axis in range(3):
                    other = [d for d in range(3) if d != axis]
                    grad[:, axis] += f * (2 * bits[axis] - 1) * w[:, other].prod(axis=1) / self.spacing
            values[valid], gradients[valid] = v, grad
        return values, gradients

    @property
    def equivalent_diameter(self):
        return (6 * self.volume / np.pi) ** (1 / 3)


def mass_properties(vertices, faces):
    """Volume, centroid and unit-density inertia of a closed outward-wound mesh.

    Each face spans a tetrahedron with the origin.  For a tetrahedron with edge
    matrix A = [a b c] the second moment is det(A)/120 * A (I + 11^T) A^T
    (canonical covariance of the unit simplex), summed and shifted to the centroid
    with the parallel-axis rule.  Inertia = tr(C) I - C.
    """
    v = np.asarray(vertices, float)
    f = np.asarray(faces, np.int64)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    det = np.einsum("ij,ij->i", a, np.cross(b, c))
    volume = det.sum() / 6
    if volume <= 0:
        raise ValueError("Mesh must enclose positive volume with outward winding")
    centroid = (det[:, None] * (a + b + c)).sum(0) / 24 / volume
    s = a + b + c
    cov = np.einsum("k,ki,kj->ij", det, a, a)
    cov += np.einsum("k,ki,kj->ij", det, b, b)
    cov += np.einsum("k,ki,kj->ij", det, c, c)
    cov += np.einsum("k,ki,kj->ij", det, s, s)
    cov /= 120
    cov -= volume * np.outer(centroid, centroid)
    inertia = np.trace(cov) * np.eye(3) - cov
    return float(volume), centroid, inertia


def face_areas(vertices, faces):
    v = np.asarray(vertices, float)
    tri = v[np.asarray(faces)]
    return 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)


def icosphere(subdivisions):
    """Unit icosphere by midpoint subdivision (10*4^s + 2 vertices, 20*4^s faces)."""
    t = (1 + 5**0.5) / 2
    v = [
        [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
        [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
        [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
    ]
    f = [
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ]
    vertices = [np.asarray(p, float) / np.linalg.norm(p) for p in v]
    faces = np.asarray(f, np.int64)
    for _ in range(subdivisions):
        midpoint = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in midpoint:
                p = vertices[i] + vertices[j]
                vertices.append(p / np.linalg.norm(p))
                midpoint[key] = len(vertices) - 1
            return midpoint[key]

        out = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            out += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        faces = np.asarray(out, np.int64)
    return np.asarray(vertices), faces


def hull_sphere(triangles):
    """Unit sphere with an arbitrary even triangle count (Fibonacci points + hull).

    A convex hull of V points on a sphere is a closed triangulation with 2V-4 faces,
    so any target count of at least 4 is reachable; faces are oriented outward.
    """
    from scipy.spatial import ConvexHull

    count = max(4, int(triangles) // 2 + 2)
    k = np.arange(count) + 0.5
    z = 1 - 2 * k / count
    angle = np.pi * (1 + 5**0.5) * k
    r = np.sqrt(1 - z * z)
    points = np.column_stack([r * np.cos(angle), r * np.sin(angle), z])
    hull = ConvexHull(points)
    faces = hull.simplices.astype(np.int64).copy()
    tri = points[faces]
    outward = np.einsum("ij,ij->i", np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), tri.mean(1))
    faces[outward < 0] = faces[outward < 0][:, ::-1]
    return points, faces


def sobol_area_samples(vertices, faces, count, seed):
    """Area-weighted scrambled-Sobol surface nodes with equal weights A/count.

    Same construction as the reference's `from_mesh` (`geometry.py:216-232`):
    first Sobol coordinate selects a face by cumulative area, the other two map to
    barycentric coordinates with the square-root warp.
    """
    from scipy.stats import qmc

    area = face_areas(vertices, faces)
    u = qmc.Sobol(3, scramble=True, seed=seed).random_base2(int(np.ceil(np.log2(count))))[:count]
    face = np.searchsorted(np.cumsum(area) / area.sum(), u[:, 0], side="right")
    face = np.minimum(face, len(faces) - 1)
    tri = np.asarray(vertices, float)[np.asarray(faces)[face]]
    root = np.sqrt(u[:, 1])
    v = u[:, 2]
    points = (
        (1 - root[:, None]) * tri[:, 0]
        + (root * (1 - v))[:, None] * tri[:, 1]
        + (root * v)[:, None] * tri[:, 2]
    )
    return points, np.full(count, area.sum() / count)


def make_shape(shape_id, vertices, faces, *, nodes, seed, scale=1.0, provenance=None):
    """Centre a closed mesh at its centre of mass and attach surface nodes."""
    v = np.asarray(vertices, float) * scale
    faces = np.asarray(faces, np.int64)
    volume, centroid, _ = mass_properties(v, faces)
    v = v - centroid
    volume, _, inertia = mass_properties(v, faces)
    samples, areas = sobol_area_samples(v, faces, nodes, seed)
    shape = Shape(
        shape_id,
        v,
        faces,
        samples,
        areas,
        volume,
        inertia,
        float(np.linalg.norm(v, axis=1).max()),
        dict(provenance or {}, sample_count=nodes, seed=seed, sampling="scrambled_sobol_area"),
    )
    shape.validate()
    return shape


def perturbed_template(shape_id, base_vertices, faces, *, sigma, diameter, nodes, rng, seed):
    """Radially perturbed closed mesh (per-vertex lognormal(0, sigma) radii).

    The result is rescaled so its equivalent diameter equals `diameter` metres.
    """
    radii = rng.lognormal(0.0, sigma, len(base_vertices))
    v = base_vertices * radii[:, None]
    volume, _, _ = mass_properties(v, faces)
    v = v * (diameter / ((6 * volume / np.pi) ** (1 / 3)))
    return make_shape(
        shape_id,
        v,
        faces,
        nodes=nodes,
        seed=seed,
        provenance={"generator": "radially_perturbed_sphere", "sigma": sigma},
    )


def harmonic_descriptors(d2=0.12, top=12, decay=2.0):
    """Sand-like spherical-harmonic shape descriptors D_n = d2 (n / 2)^-decay, n = 2..top.

    D_2 sets elongation/flatness, D_3..D_7 the grain-scale irregularity and the tail
    the roughness, after the harmonic descriptor decay measured on sand scans
    (Zhou & Wang, Geotechnique 2015; Wei et al., Powder Technology 2018).
    """
    return {n: d2 * (n / 2.0) ** -decay for n in range(2, top + 1)}


def harmonic_template(shape_id, base_vertices, faces, *, descriptors, diameter, nodes, rng, seed):
    """Closed mesh displaced radially by a random real spherical-harmonic field.

    r(u) = 1 + sum_n D_n g_n(u) with g_n = sum_m c_nm Y_nm(u), c_nm ~ N(0, 4 pi / (2n + 1)),
    so by the addition theorem each degree adds an RMS radial deviation of D_n, and the
    field, unlike per-vertex noise, does not roughen as the tessellation is refined.
    The result is rescaled so its equivalent diameter equals `diameter` metres.
    """
    from scipy.special import sph_harm_y

    u = base_vertices / np.linalg.norm(base_vertices, axis=1)[:, None]
    polar = np.arccos(np.clip(u[:, 2], -1.0, 1.0))
    azimuth = np.arctan2(u[:, 1], u[:, 0])
    r = np.ones(len(u))
    for n, d in sorted(descriptors.items()):
        g = np.zeros(len(u))
        for m in range(0, n + 1):
            y = sph_harm_y(n, m, polar, azimuth)
            scale = np.sqrt(4 * np.pi / (2 * n + 1))
            if m == 0:
                g += rng.normal(0, scale) * y.real
            else:
                g += rng.normal(0, scale) * np.sqrt(2) * y.real + rng.normal(0, scale) * np.sqrt(2) * y.imag
        r += d * g
    v = u * np.maximum(r, 0.25)[:, None]
    volume, _, _ = mass_properties(v, faces)
    v = v * (diameter / ((6 * volume / np.pi) ** (1 / 3)))
    return make_shape(
        shape_id,
        v,
        faces,
        nodes=nodes,
        seed=seed,
        provenance={"generator": "spherical_harmonic_sphere", "descriptors": {str(k): v for k, v in descriptors.items()}},
    )


def box_shape(extents, *, nodes=32, seed=0, shape_id="box"):
    """Axis-aligned box with the vertex/face layout of trimesh's `creation.box`."""
    v = np.array(
        [[0, 0, 0], [0, 0, 1], [0, 1, 0], [0, 1, 1], [1, 0, 0], [1, 0, 1], [1, 1, 0], [1, 1, 1]],
        float,
    )
    v = (v - 0.5) * np.asarray(extents, float)
    f = np.array(
        [[1, 3, 0], [4, 1, 0], [0, 3, 2], [2, 4, 0], [1, 7, 3], [5, 1, 4],
         [5, 7, 1], [3, 7, 2], [6, 4, 2], [2, 7, 6], [6, 5, 4], [7, 5, 6]],
        np.int64,
    )
    return make_shape(shape_id, v, f, nodes=nodes, seed=seed)


def concave_prism_mesh():
    """The reference's extruded-L fixture mesh in metres (`examples/create_fixtures.py:16-29`)."""
    polygon = np.array([[0, 0], [1, 0], [1, 0.5], [0.5, 0.5], [0.5, 1], [0, 1]], float)
    v = np.vstack(
        [np.column_stack([polygon, np.full(6, -0.3)]), np.column_stack([polygon, np.full(6, 0.3)])]
    ) * 0.001
    triangles = np.array([[0, 1, 3], [1, 2, 3], [0, 3, 5], [3, 4, 5]])
    faces = [*triangles[:, ::-1], *(triangles + 6)]
    for i in range(6):
        j = (i + 1) % 6
        faces.extend([[i, j, j + 6], [i, j + 6, i + 6]])
    return v, np.asarray(faces, np.int64)


# --------------------------------------------------------------------------------------
# Packed static geometry (reference `mesh.py:10-222`)
# --------------------------------------------------------------------------------------

_UINT_ARRAYS = (
    "meshRoots",
    "meshLinks",
    "meshHullInfo",
    "meshFrontier",
    "meshClusterInfo",
    "meshClusterRanges",
    "meshSampleIndices",
)


def _round_out(values, direction):
    return np.nextafter(values.astype(F32), F32(direction))


def _check_manifold(faces):
    """Two faces per edge, opposite orientation, one fan per vertex (`mesh.py:111-144`)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    edges = faces[:, [[0, 1], [1, 2], [2, 0]]].reshape(-1, 2)
    unique, inverse, count = np.unique(np.sort(edges, axis=1), axis=0, return_inverse=True, return_counts=True)
    inverse = inverse.reshape(-1)
    if np.any(count != 2):
        raise ValueError("Mesh query requires exactly two faces per edge")
    direction = np.where(edges[:, 0] < edges[:, 1], 1, -1)
    oriented = np.zeros(len(unique), int)
    np.add.at(oriented, inverse, direction)
    if np.any(oriented):
        raise ValueError("Inconsistent source winding")
    corner = np.arange(3 * len(faces))
    a = 3 * (corner // 3) + corner % 3
    b = 3 * (corner // 3) + (corner % 3 + 1) % 3
    corners = np.stack([a, b], axis=1)
    corners[direction < 0] = corners[direction < 0, ::-1]
    paired = corners[np.argsort(inverse, kind="stable")].reshape(-1, 2, 2)
    graph = coo_matrix(
        (np.ones(2 * len(paired)), (paired[:, 0].ravel(), paired[:, 1].ravel())),
        shape=(3 * len(faces),) * 2,
    )
    components, _ = connected_components(graph, directed=False)
    if components != len(np.unique(faces)):
        raise ValueError("Disconnected vertex fans in source mesh")
    return unique, inverse


def _pseudonormals(vertices, faces, inverse):
    tri = vertices[faces]
    cross = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    norm = np.linalg.norm(cross, axis=1)
    if np.any(norm <= 0):
        raise ValueError("Degenerate source triangle")
    face_normal = cross / norm[:, None]
    angle = np.empty((len(faces), 3))
    for k in range(3):
        u = tri[:, (k + 1) % 3] - tri[:, k]
        w = tri[:, (k + 2) % 3] - tri[:, k]
        angle[:, k] = np.arctan2(np.linalg.norm(np.cross(u, w), axis=1), (u * w).sum(1))
    vertex_normal = np.zeros_like(vertices)
    np.add.at(vertex_normal, faces.ravel(), (face_normal[:, None, :] * angle[:, :, None]).reshape(-1, 3))
    edge_normal = np.zeros((inverse.max() + 1, 3))
    np.add.at(edge_normal, inverse, np.repeat(face_normal, 3, axis=0))
    feature = np.concatenate(
        [face_normal[:, None, :], edge_normal[inverse.reshape(-1, 3)], vertex_normal[faces]], axis=1
    )
    lengths = np.linalg.norm(feature, axis=2)
    if np.any(lengths < 1e-12):
        raise ValueError("Undefined mesh feature pseudonormal")
    return feature / lengths[:, :, None]


def _hull_groups(vertices):
    """Outward-rounded hull planes grouped by dominant normal axis (`mesh.py:74-90`)."""
    from scipy.spatial import ConvexHull

    rounded = vertices.astype(F32).astype(float)
    hull = ConvexHull(rounded)
    margin = 16 * np.finfo(F32).eps * (1 + abs(rounded).max())
    planes = hull.equations.astype(F32)
    planes[:, 3] = _round_out(hull.equations[:, 3] - margin, -np.inf)
    planes = np.unique(planes, axis=0)
    dominant = np.abs(planes[:, :3]).argmax(axis=1)
    group = 2 * dominant + (planes[np.arange(len(planes)), dominant] >= 0)
    return [planes[group == g] for g in range(6)]


def _prepare(shape, length_unit):
    """Per-template inputs of the packer, in the reference's check order
    (`mesh.py:22-30, 74-160`): FP32 nodes, hull plane groups, manifold check,
    pseudonormals, FP32 triangles and their vertex means."""
    if shape.provenance.get("display_mesh_only"):
        raise ValueError("Mesh query requires physical source triangles, not a display-only mesh")
    vertices = np.asarray(shape.vertices, float) / length_unit
    faces = np.asarray(shape.faces, np.int64)
    points = np.ascontiguousarray((np.asarray(shape.samples) / length_unit).astype(F32))
    if len(points) == 0:
        raise ValueError("A template needs at least one surface node")
    groups = _hull_groups(vertices)
    _, inverse = _check_manifold(faces)
    feature = _pseudonormals(vertices, faces, inverse)
    tri = np.ascontiguousarray(vertices[faces].astype(F32))
    rounded = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    if np.any(np.linalg.norm(rounded, axis=1) == 0):
        raise ValueError("Triangle degenerates in normalized FP32")
    centres = np.ascontiguousarray(tri.mean(axis=1))
    return points, groups, tri, feature, centres


def sdf_grid(vertices, faces, resolution=64, *, threads=None):
    """Signed distance grid of a closed mesh on the reference's grid
    (`geometry.py:189-215`): spacing = largest extent / (resolution - 1), origin =
    lower bound - 3 spacings, ceil(extents / spacing) + 7 points per axis.  Values
    (FP32, metres, negative inside) are exact FP64 distances from the native
    producer (`sad_sdf_grid`).  Returns (sdf, origin, spacing)."""
    from . import native

    if resolution < 8:
        raise ValueError("SDF resolution >= 8 required")
    vertices = np.asarray(vertices, float)
    faces = np.asarray(faces, np.int64)
    _, inverse = _check_manifold(faces)
    feature = np.ascontiguousarray(_pseudonormals(vertices, faces, inverse))
    low, high = vertices.min(0), vertices.max(0)
    extents = high - low
    spacing = float(extents.max() / (resolution - 1))
    origin = low - 3 * spacing
    dims = (np.ceil(extents / spacing).astype(np.int64) + 7).astype(np.uint32)
    tri = np.ascontiguousarray(vertices[faces])
    out = np.empty(tuple(int(d) for d in dims), np.float32)
    threads = max(1, min(16, os.cpu_count() or 1) if threads is None else int(threads))
    o = np.ascontiguousarray(origin, float)
    rc = native.packer().sad_sdf_grid(_pointer(tri), _pointer(feature), len(tri), _pointer(o), spacing,
                                      _pointer(dims), threads, _pointer(out))
    if rc:
        raise ValueError(f"SDF grid failed ({rc})")
    return out, origin, spacing


def with_sdf(shape, resolution=64, *, threads=None):
    """`shape` with its SDF grid (the reference's `from_mesh` field)."""
    from dataclasses import replace

    sdf, origin, spacing = sdf_grid(shape.vertices, shape.faces, resolution, threads=threads)
    out = replace(shape, sdf=sdf, origin=origin, spacing=spacing,
                  provenance=dict(shape.provenance, resolution=resolution, sdf_sign="negative_inside"))
    out.validate_sdf()
    return out


def _pointer(array, row=0):
    return ct.c_void_p(int(array.ctypes.data) + int(row) * int(array.strides[0]))


def pack_meshes(shapes, length_unit, *, threads=None):
    """Pack a template library into the reference's 14 mesh arrays (normalized units;
    `mesh.py:10-222`).

    Templates are independent: their inputs are prepared and their two hierarchies
    built (native, `libsad_pack.so`) on a thread pool, each writing its own rows of
    the preallocated arrays at offsets fixed up front, so the result is the serial
    packer's, row for row, for any thread count.
    """
    from concurrent.futures import ThreadPoolExecutor

    from . import native

    lib = native.packer()
    shapes = list(shapes)
    threads = max(1, min(16, os.cpu_count() or 1) if threads is None else int(threads))
    with ThreadPoolExecutor(threads) as pool:
        prepared = list(pool.map(lambda shape: _prepare(shape, length_unit), shapes))
        t = len(shapes)
        tris = np.array([len(p[2]) for p in prepared], np.int64)
        points = np.array([len(p[0]) for p in prepared], np.int64)
        nodes = np.array([lib.sad_pack_node_count(int(n), 4) for n in tris], np.int64)
        clusters = np.array([lib.sad_pack_node_count(int(n), 64) for n in points], np.int64)
        planes = np.array([sum(len(g) for g in p[1]) for p in prepared], np.int64)
        if max(nodes.sum(), clusters.sum(), 7 * tris.sum(), points.sum()) >= 1 << 32:
            raise ValueError("Packed template library exceeds 32-bit indices")

        def starts(counts):
            return np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)

        node0, tri0, cluster0, key0, plane0 = (starts(c) for c in (nodes, tris, clusters, points, planes))
        u4 = lambda rows: np.zeros((int(rows), 4), np.uint32)  # noqa: E731
        f4 = lambda rows: np.zeros((int(rows), 4), F32)  # noqa: E731
        out = {
            "meshFrontier": u4(32 * t), "meshClusterInfo": u4(t), "meshClusterRanges": u4(clusters.sum()),
            "meshClusterMin": f4(clusters.sum()), "meshClusterMax": f4(clusters.sum()),
            "meshSampleIndices": np.zeros(int(points.sum()), np.uint32), "meshHullInfo": u4(6 * t),
            "meshPlanes": f4(planes.sum()), "meshRoots": u4(t), "meshMin": f4(nodes.sum()),
            "meshMax": f4(nodes.sum()), "meshLinks": u4(nodes.sum()), "meshTriangles": f4(3 * tris.sum()),
            "meshNormals": f4(7 * tris.sum()),
        }

        def pack(k):
            p_points, groups, tri, feature, centres = prepared[k]
            depth = ct.c_uint32(0)
            rc = lib.sad_pack_clusters(
                _pointer(p_points), len(p_points), int(cluster0[k]), int(key0[k]),
                _pointer(out["meshClusterRanges"], cluster0[k]), _pointer(out["meshClusterMin"], cluster0[k]),
                _pointer(out["meshClusterMax"], cluster0[k]), _pointer(out["meshSampleIndices"], key0[k]),
                _pointer(out["meshFrontier"], 32 * k))
            if rc:
                raise ValueError(f"cluster tree packing failed ({rc})")
            order = np.zeros(len(tri), np.uint32)
            rc = lib.sad_pack_bvh(
                _pointer(tri), _pointer(centres), len(tri), int(node0[k]), int(tri0[k]),
                _pointer(out["meshLinks"], node0[k]), _pointer(out["meshMin"], node0[k]),
                _pointer(out["meshMax"], node0[k]), _pointer(order), ct.byref(depth))
            if rc == 5:
                raise ValueError("Mesh hierarchy exceeds GPU traversal stack depth")
            if rc:
                raise ValueError(f"triangle hierarchy packing failed ({rc})")
            out["meshClusterInfo"][k] = [cluster0[k], clusters[k], 0, 0]
            out["meshRoots"][k] = [node0[k], node0[k] + nodes[k], depth.value, 0]
            out["meshTriangles"][3 * tri0[k] : 3 * (tri0[k] + tris[k]), :3] = tri[order].reshape(-1, 3)
            out["meshNormals"][7 * tri0[k] : 7 * (tri0[k] + tris[k]), :3] = feature[order].reshape(-1, 3)
            at = int(plane0[k])
            for g, group in enumerate(groups):
                out["meshHullInfo"][6 * k + g] = [at, len(group), 0, 0]
                out["meshPlanes"][at : at + len(group)] = group
                at += len(group)

        list(pool.map(pack, range(t)))
    return out


def sample_arrays(shapes, length_unit):
    """`sampleInfo` (offset, count) and `sample` (xyz, area) rows (`gpu.py:464-488`)."""
    info, samples, offset = [], [], 0
    for s in shapes:
        info.append([offset, len(s.samples), 0, 0])
        samples.append(np.column_stack([s.samples / length_unit, s.areas / length_unit**2]).astype(F32))
        offset += len(s.samples)
    return np.asarray(info, np.uint32), np.concatenate(samples)
