/* Native static-geometry packer: the two hierarchies of the reference's packed mesh
 * layout, built per template (the per-shape loop of `pack_meshes`,
 * src/sand_atlas_dem/mesh.py:10-222, whose output mesh_cache.py:115-157 caches).
 *
 * Host code only (libsad_pack.so, plain C++): the Python side prepares each
 * template's FP32 triangles and nodes, sizes the global arrays with
 * sad_pack_node_count, and calls these per template, from several threads at once
 * if it likes (the calls share nothing).  Indices written are global: the caller
 * passes where the template's nodes, triangles, clusters and node keys start.
 * Results are identical to the serial reference packer row for row.
 *
 * Return codes: 0 ok, 2 argument error, 5 hierarchy deeper than 30 levels. */
#ifndef SAND_DEM_PACK_H
#define SAND_DEM_PACK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Nodes of a median-split tree over `items` with at most `leaf_size` items per
 * leaf (1 + both halves' counts above the leaf size). */
uint32_t sad_pack_node_count(uint32_t items, uint32_t leaf_size);

/* Triangle BVH (mesh.py:160-202 `build`): preorder median split on the widest
 * FP32 centroid extent, stable order, <= 4 triangles per leaf.  `triangles` holds
 * count x 3 vertices x xyz FP32, `centres` count x xyz FP32 (each triangle's vertex
 * mean, as the caller's FP32 arithmetic gives it: the split keys are compared, not
 * computed, here).  Writes sad_pack_node_count(count, 4) rows:
 * links (escape, first triangle, triangle count, 0), lower/upper bounds rounded
 * outward by one FP32 step (x, y, z, 0); `order` receives the template-local
 * triangle index of each packed triangle; `depth` the deepest level. */
int sad_pack_bvh(const float* triangles, const float* centres, uint32_t count, uint32_t node_base, uint32_t triangle_base,
                 uint32_t* links, float* lower, float* upper, uint32_t* order, uint32_t* depth);

/* Sample-cluster tree and 32-subtree frontier (mesh.py:28-70 `partition`): preorder
 * median split on the widest FP32 extent, <= 64 node keys per leaf.  `points`
 * holds count x xyz FP32.  Writes sad_pack_node_count(count, 64) rows of ranges
 * (first key, key count, escape, 0) and lower/upper bounds, `count` template-local
 * node keys, and 32 frontier rows (subtree root, escape, 0, 0; zero rows pad). */
int sad_pack_clusters(const float* points, uint32_t count, uint32_t cluster_base, uint32_t key_base,
                      uint32_t* ranges, float* lower, float* upper, uint32_t* keys, uint32_t* frontier);

/* Signed distance grid of one closed, consistently wound triangle mesh (the SDF
 * geometry's field; the reference samples trimesh's signed distance on the same
 * grid, geometry.py:189-215).  Grid point (x, y, z) is origin + spacing * (x, y, z)
 * and goes to out[(x * ny + y) * nz + z]; its value is the exact distance to the
 * nearest triangle (FP64, BVH), negative inside by the angle-weighted pseudonormal
 * of the nearest feature.  `triangles`: count x 3 vertices x xyz; `normals`: count
 * x 7 x xyz unit pseudonormals (face, edges AB/BC/CA, vertices A/B/C).  `threads`
 * host threads share the x slabs. */
int sad_sdf_grid(const double* triangles, const double* normals, uint32_t count, const double* origin,
                 double spacing, const uint32_t* dims, uint32_t threads, float* out);

#ifdef __cplusplus
}
#endif

#endif
