"""ORACLE serial mesh packer -- test infrastructure only.

A plain-Python restatement of the reference's per-template packing loop
(`src/sand_atlas_dem/mesh.py:10-222`: `partition` 28-70 for the sample-cluster tree
and frontier, `build` 160-202 for the triangle BVH), one template after another with
lists that grow as the reference's do.  Tests check the package's native, threaded
packer against it row for row; the product never imports it.
"""

import numpy as np


def pack_meshes_serial(geometry, shapes, length_unit):
    """Pack `shapes` serially; `geometry` is the package's geometry module (its
    hull-plane, manifold and pseudonormal helpers are shared inputs, not under test)."""
    F32 = geometry.F32
    _round_out = geometry._round_out
    _hull_groups = geometry._hull_groups
    _check_manifold = geometry._check_manifold
    _pseudonormals = geometry._pseudonormals
    _UINT_ARRAYS = geometry._UINT_ARRAYS

    class _Packer:
        def __init__(self):
            self.a = {name: [] for name in (
                "meshFrontier", "meshClusterInfo", "meshClusterRanges", "meshClusterMin",
                "meshClusterMax", "meshSampleIndices", "meshHullInfo", "meshPlanes",
                "meshRoots", "meshMin", "meshMax", "meshLinks", "meshTriangles", "meshNormals",
            )}

        def clusters(self, points):
            """Preorder median-split sample tree; leaves hold <=64 node keys."""
            ranges, lo, hi, keys = (self.a[k] for k in (
                "meshClusterRanges", "meshClusterMin", "meshClusterMax", "meshSampleIndices"))
            first = len(ranges)

            def split(ids):
                node = len(ranges)
                ranges.append([0, 0, 0, 0])
                lo.append([*_round_out(points[ids].min(0), -np.inf), 0])
                hi.append([*_round_out(points[ids].max(0), np.inf), 0])
                if len(ids) > 64:
                    axis = np.ptp(points[ids], axis=0).argmax()
                    order = ids[np.argsort(points[ids, axis], kind="stable")]
                    half = len(order) // 2
                    split(order[:half])
                    split(order[half:])
                else:
                    ranges[node][0:2] = [len(keys), len(ids)]
                    keys.extend(ids.tolist())
                ranges[node][2] = len(ranges)

            split(np.arange(len(points)))
            self.a["meshClusterInfo"].append([first, len(ranges) - first, 0, 0])
            # Up to 32 disjoint complete subtrees; split the largest internal one first.
            pending = [first]
            while len(pending) < 32:
                internal = [n for n in pending if ranges[n][1] == 0]
                if not internal:
                    break
                node = max(internal, key=lambda n: ranges[n][2] - n)
                at = pending.index(node)
                pending[at : at + 1] = [node + 1, ranges[node + 1][2]]
            self.a["meshFrontier"].extend([[n, ranges[n][2], 0, 0] for n in pending])
            self.a["meshFrontier"].extend([[0, 0, 0, 0]] * (32 - len(pending)))

        def bvh(self, tri, feature):
            """Median-split triangle BVH, <=4 triangles per leaf, escape links."""
            links, lo, hi = self.a["meshLinks"], self.a["meshMin"], self.a["meshMax"]
            root = len(links)
            first_triangle = len(self.a["meshTriangles"]) // 3
            order, depths = [], []

            def build(ids, depth):
                node = len(links)
                depths.append(depth)
                lo.append([*_round_out(tri[ids].min(axis=(0, 1)), -np.inf), 0])
                hi.append([*_round_out(tri[ids].max(axis=(0, 1)), np.inf), 0])
                links.append([0, 0, 0, 0])
                if len(ids) <= 4:
                    links[node][1:3] = [first_triangle + len(order), len(ids)]
                    order.extend(ids.tolist())
                else:
                    centres = tri[ids].mean(axis=1)
                    axis = np.ptp(centres, axis=0).argmax()
                    ids = ids[np.argsort(centres[:, axis], kind="stable")]
                    half = len(ids) // 2
                    build(ids[:half], depth + 1)
                    build(ids[half:], depth + 1)
                links[node][0] = len(links)

            build(np.arange(len(tri)), 0)
            if max(depths) > 30:
                raise ValueError("Mesh hierarchy exceeds GPU traversal stack depth")
            self.a["meshRoots"].append([root, len(links), max(depths), 0])
            self.a["meshTriangles"].extend(tri[order].reshape(-1, 3))
            self.a["meshNormals"].extend(feature[order].reshape(-1, 3))

        def add(self, shape, length_unit):
            vertices = np.asarray(shape.vertices, float) / length_unit
            faces = np.asarray(shape.faces, np.int64)
            self.clusters((np.asarray(shape.samples) / length_unit).astype(F32))
            for group in _hull_groups(vertices):
                self.a["meshHullInfo"].append([len(self.a["meshPlanes"]), len(group), 0, 0])
                self.a["meshPlanes"].extend(group)
            _, inverse = _check_manifold(faces)
            feature = _pseudonormals(vertices, faces, inverse)
            tri = vertices[faces].astype(F32)
            rounded = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
            if np.any(np.linalg.norm(rounded, axis=1) == 0):
                raise ValueError("Triangle degenerates in normalized FP32")
            self.bvh(tri, feature)

        def arrays(self):
            out = {}
            for name, rows in self.a.items():
                if name in _UINT_ARRAYS:
                    out[name] = np.asarray(rows, np.uint32)
                elif name in ("meshTriangles", "meshNormals"):
                    a = np.asarray(rows, F32).reshape(-1, 3)
                    out[name] = np.pad(a, ((0, 0), (0, 1)))
                else:
                    out[name] = np.asarray(rows, F32).reshape(-1, 4) if rows else np.zeros((0, 4), F32)
            return out


    def pack_meshes(shapes, length_unit):
        """Pack a template library into the reference's 14 mesh arrays (normalized units)."""
        packer = _Packer()
        for shape in shapes:
            packer.add(shape, length_unit)
        return packer.arrays()

    packer = _Packer()
    for shape in shapes:
        packer.add(shape, length_unit)
    return packer.arrays()
