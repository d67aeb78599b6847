"""The warm-started exact query's pruning bound is safe in FP32 -- CPU only.

The collection kernel starts each node's BVH traversal with the bound
(sqrt(d_last) + WARM_SLACK)^2, where d_last is the FP32 squared distance to the
triangle that won the node's previous query.  A box is pruned when its FP32 box
distance exceeds the bound.  The winner of the current query is never lost if no
box holding a triangle T is farther, in FP32, than the bound built from T's own
FP32 distance: the current winner W has d(W) <= d(T_last), ancestors' boxes contain
the leaf box of W, and FP32 subtraction, squaring and summation are monotone.

This test evaluates the kernel's own FP32 arithmetic (closestTriangle and boxDistance
of csrc/dem_math.h, no contraction, correctly rounded division and square root) with
numpy float32 on adversarial points -- on and just off every vertex, edge and face,
and along the box axes -- of every packed template of four libraries, and checks
box(leaf(T), q) <= bound(T, q).  On these 8 M points even the unwidened bound d(T)
holds, thanks to the packer's one-ulp outward box padding; the slack keeps a margin
of about 30 ulps of distance beyond that.
"""

import numpy as np
import pytest

F = np.float32
WARM_SLACK = F(4e-6)  # csrc/sad_hip.hip


def dot32(a, b):
    return ((F(0) + a[:, 0] * b[:, 0]) + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def closest_triangle32(p, a, b, c):
    """closestTriangle (dem_math.h; solver.slang:203-219) in float32, vectorized."""
    with np.errstate(all="ignore"):
        ab, ac, ap = b - a, c - a, p - a
        d1, d2 = dot32(ab, ap), dot32(ac, ap)
        bp = p - b
        d3, d4 = dot32(ab, bp), dot32(ac, bp)
        vc = d1 * d4 - d3 * d2
        cp = p - c
        d5, d6 = dot32(ab, cp), dot32(ac, cp)
        vb = d5 * d2 - d1 * d6
        va = d3 * d6 - d5 * d4
        inv = F(1) / (va + vb + vc)
        face = (a + ab * (vb * inv)[:, None]) + ac * (vc * inv)[:, None]
        edge_bc = b + ((d4 - d3) / ((d4 - d3) + (d5 - d6)))[:, None] * (c - b)
        edge_ac = a + (d2 / (d2 - d6))[:, None] * ac
        edge_ab = a + (d1 / (d1 - d3))[:, None] * ab
        x = face
        x = np.where(((va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0))[:, None], edge_bc, x)
        x = np.where(((vb <= 0) & (d2 >= 0) & (d6 <= 0))[:, None], edge_ac, x)
        x = np.where(((d6 >= 0) & (d5 <= d6))[:, None], c, x)
        x = np.where(((vc <= 0) & (d1 >= 0) & (d3 <= 0))[:, None], edge_ab, x)
        x = np.where(((d3 >= 0) & (d4 <= d3))[:, None], b, x)
        x = np.where(((d1 <= 0) & (d2 <= 0))[:, None], a, x)
    return x.astype(F)


def box_distance32(lo, hi, p):
    d = np.maximum(np.maximum(lo - p, p - hi), F(0))
    return dot32(d, d)


def adversarial_points(a, b, c, rng):
    """Per triangle: its vertices, edge points and face points, each moved by offsets
    from 0 to 0.1 along random directions, the face normal and the coordinate axes."""
    t = len(a)
    n = np.cross(b - a, c - a).astype(np.float64)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    bases = [a, b, c, 0.5 * (a + b), 0.5 * (b + c), 0.5 * (c + a)]
    w = rng.dirichlet(np.ones(3), t)
    bases.append(w[:, :1] * a + w[:, 1:2] * b + w[:, 2:] * c)
    u = rng.uniform(0, 1, (t, 1))
    bases.append(a + u * (b - a))
    axes = np.eye(3)
    out = []
    for base in bases:
        base = base.astype(np.float64)
        for size in (0.0, 1e-9, 1e-8, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2, 1e-1):
            r = rng.normal(size=(t, 3))
            r /= np.linalg.norm(r, axis=1, keepdims=True)
            out.append(base + size * r)
            out.append(base + size * n * rng.choice([-1.0, 1.0], (t, 1)))
            out.append(base + size * axes[rng.integers(0, 3, t)] * rng.choice([-1.0, 1.0], (t, 1)))
    return np.stack(out, 1).astype(F)  # (t, points, 3)


def libraries(pkg):
    from sand_atlas_dem_amd import assembly

    g = pkg.geometry
    rng = assembly.streams(7)
    dense = assembly.icosphere_library(6, subdivisions=(2, 3, 4), sigma=0.15, diameter=3e-4, nodes=256,
                                       rng=rng["templates"], seed=7)
    uneven = assembly.uneven_library(6, diameter=3e-4, rng=rng["shapes"], seed=7)
    box = g.box_shape([1e-3, 2e-3, 3e-3])
    concave = g.make_shape("concave", *g.concave_prism_mesh(), nodes=256, seed=0)
    return {"dense": dense, "uneven": uneven, "box": [box], "concave": [concave]}


@pytest.mark.parametrize("library", ["dense", "uneven", "box", "concave"])
def test_widened_warm_bound_never_prunes_a_holding_box(pkg, library):
    shapes = libraries(pkg)[library]
    length = float(np.median([s.equivalent_diameter for s in shapes]))
    mesh = pkg.geometry.pack_meshes(shapes, length)
    links, lo, hi = mesh["meshLinks"], mesh["meshMin"][:, :3], mesh["meshMax"][:, :3]
    tri = mesh["meshTriangles"][:, :3].reshape(-1, 3, 3)
    leaf_of = np.full(len(tri), -1, np.int64)
    for node in np.nonzero(links[:, 2])[0]:
        leaf_of[links[node, 1] : links[node, 1] + links[node, 2]] = node
    assert (leaf_of >= 0).all()
    rng = np.random.default_rng(13)
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    pts = adversarial_points(a, b, c, rng)
    m = pts.shape[1]
    rep = lambda x: np.repeat(x, m, axis=0)  # noqa: E731
    p = pts.reshape(-1, 3)
    x = closest_triangle32(p, rep(a), rep(b), rep(c))
    d = dot32(p - x, p - x)
    box = box_distance32(rep(lo[leaf_of]), rep(hi[leaf_of]), p)
    reach = np.sqrt(d) + WARM_SLACK
    bound = reach * reach
    assert np.isfinite(d).all()
    bad = box > bound
    assert not bad.any(), f"{bad.sum()} of {len(p)} points: box {box[bad][:4]} > bound {bound[bad][:4]}"


def test_float32_emulation_equals_the_oracle_query(pkg, oracle_module):
    """The emulated closest-point distance is the oracle's meshDistance bit for bit:
    the minimum over all triangles of the emulated FP32 squared distance, square-
    rooted, equals |meshProbe| for points around two templates."""
    cpu = oracle_module.make_oracle(pkg)
    g = pkg.geometry
    shapes = [g.box_shape([1e-3, 2e-3, 3e-3]), g.make_shape("concave", *g.concave_prism_mesh(), nodes=256, seed=0)]
    e = pkg.Engine(shapes, pkg.Experiment(particles=2, timestep=1e-9), [[0.005] * 3, [0.015] * 3], [0.03] * 3,
                   [0, 1], [1, 1], backend=cpu)
    tri = e.read("meshTriangles")[:, :3].reshape(-1, 3, 3)
    roots = e.read("meshRoots", np.uint32)
    links = e.read("meshLinks", np.uint32)
    rng = np.random.default_rng(17)
    for sid in (0, 1):
        first = links[roots[sid, 0] : roots[sid, 1], 1]
        count = links[roots[sid, 0] : roots[sid, 1], 2]
        own = np.concatenate([np.arange(f, f + c) for f, c in zip(first, count) if c])
        pts = rng.uniform(-1.2, 1.2, (2000, 3)).astype(F)
        best = np.full(len(pts), np.inf, F)
        for t in own:
            a, b, c = (np.broadcast_to(tri[t, k], pts.shape).astype(F) for k in range(3))
            x = closest_triangle32(pts, a, b, c)
            best = np.minimum(best, dot32(pts - x, pts - x))
        out = cpu.probe("meshProbe", [len(pts), sid], e.bindings(), np.pad(pts, ((0, 0), (0, 1))),
                        np.zeros((len(pts), 4), F))
        assert np.array_equal(np.sqrt(best), np.abs(out[:, 3]))
