/*
 * sand_dem_hip.h -- C ABI of the MI355X `hip` backend for the DEM timestep
 * (surface-mesh and SDF-grid contact geometry, translating planes, diagnostics).
 *
 * The boundary mirrors the reference's native CPU backend ABI
 * (src/sand_atlas_dem/cpu.py:288-293, generated `cpu_dispatch`):
 *
 *     int cpu_dispatch(uint32_t count, void** buffers, uint64_t* bytes,
 *                      double* values[, void* scratch]);
 *
 * Buffers are bound by position in the reference's `Resources` field order
 * (src/sand_atlas_dem/shaders/solver.slang:20-100, minus the `c` config block);
 * `values` carries the `Config` scalars with vector fields expanded per axis
 * (solver.slang:12-19; cpu.py:217-228) followed by a thread-count slot.  All
 * pointers here are device pointers (HBM) owned by the caller; the library keeps
 * only its scratch workspace.  Every entry point returns 0 on success and a
 * nonzero SAD_E_* code on a contract or HIP runtime failure; physics failures are
 * reported, as in the reference, by bits OR-ed into control[0] (gpu.py:11-26).
 * No C++ exception crosses this boundary.
 */
#ifndef SAND_DEM_HIP_H
#define SAND_DEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Buffer slots: position = field index in solver.slang `Resources` (lines 22-98). */
enum sad_buffer {
    SAD_CONTROL = 0, SAD_CELL, SAD_POSITION, SAD_POSITION_REMAINDER,
    SAD_VELOCITY_REMAINDER, SAD_MOMENTUM_REMAINDER, SAD_ROTATION_REMAINDER,
    SAD_REFERENCE_POSITION_REMAINDER, SAD_CELL_LOG, SAD_ACCUMULATED_REMAINDER,
    SAD_VELOCITY, SAD_ROTATION, SAD_MOMENTUM, SAD_FORCE, SAD_TORQUE,
    SAD_REFERENCE_POSITION, SAD_REFERENCE_CELL, SAD_SHAPE_ID, SAD_BODY, SAD_INV_I,
    SAD_GRIDS, SAD_SAMPLE_INFO, SAD_GRID_ORIGIN, SAD_DISTANCE, SAD_SAMPLE,
    SAD_MESH_ROOTS, SAD_MESH_HULL_INFO, SAD_MESH_PLANES, SAD_MESH_MIN, SAD_MESH_MAX,
    SAD_MESH_LINKS, SAD_MESH_TRIANGLES, SAD_MESH_NORMALS, SAD_MESH_FRONTIER,
    SAD_MESH_PAIR_MASK, SAD_MESH_PAIR_TRANSFORM, SAD_MESH_CLUSTER_INFO,
    SAD_MESH_CLUSTER_RANGES, SAD_MESH_CLUSTER_MIN, SAD_MESH_CLUSTER_MAX,
    SAD_MESH_SAMPLE_INDICES, SAD_BIN_COUNTS, SAD_BIN_PARTICLES, SAD_PAIR_COUNTS,
    SAD_PAIR_OFFSETS, SAD_PAIRS, SAD_CONTACT_COUNTS, SAD_CONTACT_OFFSETS, SAD_META,
    SAD_POINT_DEPTH, SAD_NORMAL_AREA, SAD_SLIP, SAD_CONTACT_FORCE,
    SAD_NORMAL_LOSS_POWER, SAD_NORMAL_DAMPING, SAD_FORCE_VELOCITY, SAD_FORCE_OMEGA,
    SAD_OLD_META, SAD_OLD_SLIP, SAD_OLD_NORMAL, SAD_OLD_HASH, SAD_HASH, SAD_HEADS,
    SAD_NEXT, SAD_PARTICLE_STRESS, SAD_TIMESTEP_ROWS, SAD_PARTICLE_DIAGNOSTIC,
    SAD_DIAGNOSTIC_ROWS, SAD_ACCUMULATED, SAD_FABRIC_ROWS, SAD_WALL_PLANE,
    SAD_WALL_REMAINDER, SAD_WALL_VELOCITY, SAD_WALL_MATERIAL, SAD_WALL_BOX_IDS,
    SAD_WALL_ROWS, SAD_WALL_ACCUMULATED,
    SAD_BUFFER_COUNT /* = 77 */
};

/* values[] slots: Config fields (solver.slang:12-19) expanded per axis. */
enum sad_value {
    SAD_V_ACTIVE_TIMESTEP = 0, SAD_V_N, SAD_V_SHAPES, SAD_V_PAIR_CAPACITY,
    SAD_V_CONTACT_CAPACITY, SAD_V_BIN_CAPACITY, SAD_V_BIN_COUNT, SAD_V_BINS_X,
    SAD_V_BINS_Y, SAD_V_BINS_Z, SAD_V_HASH_SIZE, SAD_V_DT, SAD_V_KN, SAD_V_KT,
    SAD_V_MU, SAD_V_NORMAL_VISCOSITY, SAD_V_SKIN, SAD_V_DRAG, SAD_V_MAX_OVERLAP,
    SAD_V_HISTORY_STEP, SAD_V_MESH_GEOMETRY, SAD_V_PERIODIC, SAD_V_WALL_COUNT,
    SAD_V_WALL_BOX_ENABLED, SAD_V_SOURCE_SURFACE_NORMAL,
    SAD_V_THREADS, /* reference cpu.py:419 appends the worker count here */
    SAD_VALUE_COUNT /* = 26 */
};

/* History side slots for sad_hip_advance (gpu.py:305-316 double buffer). */
enum sad_side { SAD_SIDE_META = 0, SAD_SIDE_SLIP, SAD_SIDE_NORMAL_AREA, SAD_SIDE_HASH,
                SAD_SIDE_COUNT };

/* Return codes. */
#define SAD_OK 0
#define SAD_E_HIP 1          /* HIP runtime error; sad_hip_last_error() names it */
#define SAD_E_ARGUMENT 2     /* contract violation: size, null pointer, unknown kernel */
#define SAD_E_UNSUPPORTED 3  /* configuration not provided (source-surface normals, limits) */

/* Device workspace owned by the library (replaces the reference's collector
 * scratch object, cpu_parallel.hpp:3-25 / cpu.py:364-376).  Holds staging,
 * per-task counts, scan partials and the per-grain pair adjacency; grows on
 * demand, never shrinks.  Returns NULL on failure. */
void* sad_hip_scratch_create(void);
void sad_hip_scratch_destroy(void* scratch);
/* stats[0..11] (the caller passes 12 words): cumulative steps (collections),
 * directed candidate pairs, exactly queried nodes and contacts since creation;
 * workspace bytes; learned-cull culled, audited and missed nodes; nodes passed over
 * by the temporal skip; queried nodes whose nearest triangle was proven unchanged
 * without a traversal; directed-pair tasks that (re)took a temporal-skip anchor;
 * mesh collections that rebuilt the per-grain pair adjacency and task buckets
 * (the others reused them: the pair list and templates were unchanged). */
int sad_hip_scratch_stats(void* scratch, uint64_t* stats);

/* Per-kernel HIP-event timing on the launch stream (no reference counterpart;
 * the reference's Engine.profile_step, gpu.py:938-1014, times stage groups).
 * enable != 0 clears the accumulators and brackets every launch with events. */
int sad_hip_profile(void* scratch, int enable);
/* Copies accumulated milliseconds and launch counts per kernel id (up to
 * `capacity`); returns the number of kernel ids or a negative error. */
int sad_hip_profile_read(void* scratch, double* ms, uint64_t* launches, uint32_t capacity);
const char* sad_hip_kernel_name(uint32_t kernel);

/* One reference kernel by name, `count` threads as the reference dispatches it
 * (gpu.py:529-565).  Replaces cpu_dispatch (cpu.py:391-428) / Metal dispatch
 * (backend.py:62-83).  `stream` is a hipStream_t (NULL = default stream).
 * Supported names: beginStep advanceCell checkRebuild clearBins fillBins
 * countPairs pairScan emitPairs resetMeshContacts prepareMeshPairs
 * collectMeshFrontiers collectMeshContacts (the reference's verification
 * collector; like the CPU backend, the same deterministic collection)
 * clearContactStorage resolveContacts buildHistoryHash
 * gather finishStep captureContactKinematics checkContactTimestep advanceWalls
 * collectWallContacts gatherWallRows refreshWallPower diagnosticRows fabricRows,
 * and for the SDF geometry (values[SAD_V_MESH_GEOMETRY] == 0) countContacts
 * contactTotal emitContacts.  "pairScan" and "contactScan" are the reference
 * Scan's exclusive scans (gpu.py:553-555, 539-541) of pairCounts and
 * contactCounts as single names. */
int sad_hip_dispatch(const char* kernel, uint32_t count, void* const* buffers,
                     const uint64_t* bytes, const double* values, void* scratch,
                     void* stream);

/* `steps` explicit timesteps, the reference's per-step encoder sequence
 * (gpu.py:714-726) on device, with the history sides swapped every step:
 * step k writes side[(active + k) % 2] and reads the other.  `sides` holds
 * 2*SAD_SIDE_COUNT device pointers (side 0 then side 1); `buffers` entries for
 * meta/slip/normalArea/hash and old* are ignored.  values[SAD_V_HISTORY_STEP]
 * applies to the first step; later steps always carry history (gpu.py:726).
 * `graph` != 0 replays a captured hipGraph of the step pair when possible. */
int sad_hip_advance(uint32_t steps, uint32_t active, void* const* buffers,
                    const uint64_t* bytes, const double* values,
                    void* const* sides, const uint64_t* side_bytes,
                    void* scratch, void* stream, int graph);

/* The timesteps of sad_hip_advance without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * library workspace is carved from a file-scope static device array of 64 MiB per
 * GPU instead of the scratch object; the caller still passes every Resources buffer
 * and both history sides as for sad_hip_advance.  Mesh geometry only, plain launches
 * (no captured graph, temporal skip or learned cull); results equal sad_hip_advance's.
 * An assembly whose workspace exceeds the array (about 448 bytes per contact slot
 * plus per-pair and per-grain rows) returns SAD_E_UNSUPPORTED before any launch,
 * with sad_hip_last_error() giving the bytes needed.  All calls on one GPU share its
 * array, so calls that overlap on one GPU race. */
int sad_hip_advance_static(uint32_t steps, uint32_t active, void* const* buffers,
                           const uint64_t* bytes, const double* values,
                           void* const* sides, const uint64_t* side_bytes, void* stream);

/* Probes over the reference's query functions (solver.slang:253-257, 306-318,
 * 418-424): meshProbe (float4 points -> float4 normal,signed distance),
 * meshCullProbe (float4 points -> uint), meshClusterProbe (src,dst -> uint per
 * cluster), meshPairBoundsProbe (direction, first -> 2 float4 per cluster).
 * uniforms[] = {count, shape|src|direction, dst|first}. */
int sad_hip_probe(const char* kernel, const uint32_t* uniforms, void* const* buffers,
                  const uint64_t* bytes, const double* values, const void* input,
                  void* output, void* stream);

/* The reference Engine's hierarchical scan (scan.slang, driven by the Scan class
 * at gpu.py:65-91): "blockScan" writes the exclusive scan of each 256-element
 * block of input[0..n) to output and the block totals to aux; "addBlocks" adds
 * aux[id / 256] to output[id].  Unsigned wrap-around, as the GPU kernels. */
int sad_hip_scan_kernel(const char* kernel, uint32_t n, const uint32_t* input,
                        uint32_t* output, uint32_t* aux, void* stream);

/* normalProbe (normal.slang:12-20), the reference's probe of its normal force law
 * normalResponse (normal.slang:3-11): input rows (depth, area, normal velocity, -)
 * -> output rows (force, loss power, 0, 0) for i < n, FP32. */
int sad_hip_normal_probe(uint32_t n, float stiffness, float viscosity, float weight,
                         const void* input, void* output, void* stream);

/* The adapter's compiled-kernel check (solver.slang:4-10, dispatched by the
 * backends' doctor at backend.py:103-119 and cpu.py:668-689): b[i] = a[i] * a[i] + 1
 * for i < n, FP32. */
int sad_hip_smoke(uint32_t n, const void* a, void* b, void* stream);

/* The diagnostics reductions (diagnostic.slang, driven by Engine.diagnostics at
 * gpu.py:746-792): one total per group of 256 records, summed by the pairwise
 * halving tree 128 -> 1 of that shader.  "reduceRows": 8 float4 rows per record,
 * row 4 .w takes the maximum; "reduceFabric": 2 rows; "reduceWallRows": 3 rows
 * per (wall, record) at input[3 * (wall * n + id) + row], output[3 * (wall *
 * output_count + group) + row], row 0 .w an integer count carried as FP32 bits
 * (walls and output_count are ignored by the other two).  Byte sizes are checked
 * against what the launch touches. */
int sad_hip_reduce_kernel(const char* kernel, uint32_t n, uint32_t walls, uint32_t output_count,
                          const void* input, uint64_t input_bytes, void* output,
                          uint64_t output_bytes, void* stream);

/* Tunables of the collection stage (no reference counterpart; none changes any
 * result): "collect_blocks_per_cu" (1..32 workgroups per CU in the persistent
 * launch), "collect_pop" (1..16 tasks taken per queue atomic), "collect_queue"
 * (128 or 256 queued nodes per wavefront), "collect_leaf_batch" (1..64, default 32:
 * the exact queries stop a descent pass once this many lanes hold a leaf to test),
 * "collect_batch_tasks" (1..16, default 16; at most 8 with the 128-node queue: tasks
 * whose queued nodes a wavefront queries together), "collect_batch_fill" (1..256,
 * default 96: queued traversals that end such a batch early), "gap_shortcut" (0/1,
 * default 1, with "temporal_skip": a query whose node moved less than half its stored
 * nearest-triangle gap since the previous collection takes the previous nearest
 * triangle without a traversal), "gap_reach_micro" (0..100000, default 300: a traversal
 * also visits what lies within this many 1e-6 template units beyond its bound, so the
 * gap it stores is the true one up to that reach), "temporal_skip" (0/1, default 0: pass
 * over nodes whose signed distance at an anchor pose exceeds a rigorous bound on
 * their motion since), "skip_limit_micro" (2000..250000, default 20000: re-anchor a
 * directed pair once its motion bound reaches this many 1e-6 template units), "skip_rebase" (0/1,
 * default 0: re-anchor an expired pair by moving its stored bounds toward zero by its
 * motion bound instead of recomputing them).  value < 0 only
 * reads; *previous (if non-null) receives the value before the call. */
int sad_hip_tune(void* scratch, const char* name, int64_t value, int64_t* previous);

/* Forget every temporal-skip anchor of this workspace (no reference counterpart).
 * A caller that enables "temporal_skip" calls this whenever it rewrites template
 * geometry, body scales or shape ids in buffers it keeps passing (a new binding of
 * those buffers is detected by address). */
int sad_hip_invalidate(void* scratch);

/* ---- Learned signed-distance cull (no reference counterpart).  One MLP per
 * template: positional encoding (xyz + sin/cos of 2^k pi x, k < 6) -> 64 -> 64 ->
 * 64 -> 1, ReLU, BF16 MFMA with FP32 accumulation.  `weights` = BF16 blocks of
 * 11,264 values per template, `blocks` = FP32 blocks of 264 floats per template
 * (biases, output layer, input centre/scale, cull margin); masters, gradients and
 * Adam moments use 11,008 floats per template.  All pointers are device pointers. */
/* Enable the cull in the collection stage (audit != 0: culled nodes are still
 * queried exactly and counted; stats[5..7] = culled, audited, audit misses). */
int sad_hip_nn_configure(void* scratch, int enabled, int audit, const void* weights,
                         const void* blocks, uint32_t templates);
/* Mean-squared-error gradient over template-major points (float4 xyz; counts per
 * template multiples of 64) into [template][slice] partials; adds squared errors
 * to loss[template]. */
int sad_hip_nn_grad(const void* points, const float* labels, const uint32_t* tstart,
                    const uint32_t* tcount, uint32_t templates, uint32_t slices,
                    const void* weights, const float* blocks, float* partial,
                    float* loss, float loss_scale, void* stream);
/* Sum the partials into grad; when lr > 0 take one Adam step (bias-corrected at
 * `step`) on the FP32 masters. */
int sad_hip_nn_adam(float* master, float* m1, float* m2, const float* partial,
                    float* grad, uint32_t slices, uint32_t templates, float lr,
                    float beta1, float beta2, float eps, uint32_t step, void* stream);
/* FP32 masters -> BF16 weight blocks + FP32 layer parameters of the blocks. */
int sad_hip_nn_pack(const float* master, void* weights, float* blocks,
                    uint32_t templates, void* stream);
/* Predictions for 64-point tiles (tile_template[t] = template of tile t). */
int sad_hip_nn_infer(const void* points, const uint32_t* tile_template, uint32_t tiles,
                     const void* weights, const float* blocks, float* out, void* stream);
/* Training points for every template: uniforms = (templates, points per template,
 * seed high, seed low); `buffers`/`values` as for sad_hip_dispatch (the packed mesh
 * arrays).  triangles: 3 float4 rows per triangle (normalized units) with per-
 * template ranges tri_start/tri_count and inclusive area prefix `cdf`; box: per
 * template (lo, grain diameter), (hi, label scale).  Writes float4 points and
 * labels = exact signed distance / label scale clamped to [-1, 1]. */
int sad_hip_nn_sample(const uint32_t* uniforms, void* const* buffers, const uint64_t* bytes,
                      const double* values, const void* triangles, const float* cdf,
                      const uint32_t* tri_start, const uint32_t* tri_count, const void* box,
                      void* points, float* labels, void* stream);
/* One BF16 MFMA 16x16xk tile, k = 16 or 32 (row-major A 16 x k, B k x 16; FP32 D
 * 16 x 16) for operand-layout tests. */
int sad_hip_mfma_probe(const void* a, const void* b, float* d, int k, void* stream);

/* Name of the last HIP error seen by this library (static string). */
const char* sad_hip_last_error(void);
/* ABI version, currently 1. */
int sad_hip_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SAND_DEM_HIP_H */
