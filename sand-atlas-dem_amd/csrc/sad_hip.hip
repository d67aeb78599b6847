// sad_hip.hip -- MI355X (gfx950) kernels of the surface-mesh DEM step and the C ABI
// declared in include/sand_dem_hip.h.
//
// Semantics follow the reference CPU backend (solver.slang + cpu_scheduling.slang as
// spliced by cpu.py:46-103), so neighbour lists, contact sets and order, history
// lookups and per-grain sums are reproduced exactly; the schedule is the GPU's own:
//   * neighbour search: atomic bin fill + per-bin id sort (= serial fill order),
//     27-cell stencil count / scan / emit, per-grain pair adjacency for the gather;
//   * contact collection: persistent wave64 work queue over directed-pair tasks;
//     cluster culls lane-parallel, ballot-compacted query queue in LDS, per-lane
//     BVH stack in LDS; hits ranked by wave prefix count, staged, then placed by a
//     task-count scan (= the CPU's pair/pass/preorder/lane order);
//   * friction history: per-contact resolve, CAS linear-probe hash;
//   * force/torque gather: one lane per grain walking its pairs in descending
//     contact id with the compensated FP32 sums of the reference;
//   * integrator: per-grain compensated kicks, quaternion drift, affine streaming.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/sand_dem_hip.h"
#include "dem_math.h"
#include "nn.h"

namespace {

struct Cfg {
    uint32_t activeTimestep, N, shapes, pairCapacity, contactCapacity, binCapacity, binCount;
    uint32_t bins[3];
    uint32_t hashSize;
    float dt, kn, kt, mu, normalViscosity, skin, drag, maxOverlap;
    uint32_t historyStep, meshGeometry, periodic, wallCount, wallBoxEnabled, sourceSurfaceNormal;
};

// Device pointers of one binding set (reference Resources order) plus config.
struct P {
    Cfg c;
    uint32_t meshNodes;  // BVH nodes over all templates (rows of meshLinks)
    uint32_t templates;  // rows of meshRoots
    uint32_t planes;     // rows of meshPlanes
    uint32_t triangles;  // meshTriangles rows / 3
    uint32_t sampleKeys; // rows of meshSampleIndices
    uint32_t clusters;   // rows of meshClusterRanges
    uint32_t* control;
    float4 *cell, *position, *positionRemainder, *velocityRemainder, *momentumRemainder;
    float4 *rotationRemainder, *referencePositionRemainder, *cellLog, *accumulatedRemainder;
    float4 *velocity, *rotation, *momentum, *force, *torque, *referencePosition, *referenceCell;
    const uint32_t* shapeId;
    const float4 *body, *invI;
    const uint4* sampleInfo;
    const float4* sample;
    const uint4 *meshRoots, *meshHullInfo;
    const float4 *meshPlanes, *meshMin, *meshMax;
    const uint4* meshLinks;
    const float4 *meshTriangles, *meshNormals;
    uint32_t* meshPairMask;
    float4* meshPairTransform;
    const uint4 *meshClusterInfo, *meshClusterRanges;
    const float4 *meshClusterMin, *meshClusterMax;
    const uint32_t* meshSampleIndices;
    uint32_t *binCounts, *binParticles, *pairCounts, *pairOffsets;
    uint2* pairs;
    uint32_t *contactCounts, *contactOffsets;
    // SDF geometry: per template (first value, nx, ny, nz) and (origin, spacing),
    // and the grid values.
    const uint4* grids;
    const float4* gridOrigin;
    const float* distance;
    uint64_t distanceCount;
    uint4* meta;
    float4 *pointDepth, *normalArea, *slip, *contactForce;
    float* normalLossPower;
    float4 *normalDamping, *forceVelocity, *forceOmega;
    const uint4* oldMeta;
    const float4 *oldSlip, *oldNormal;
    const uint32_t* oldHash;
    uint32_t* hash;
    float4 *particleStress, *timestepRows, *particleDiagnostic, *accumulated;
    // Evidence rows for the diagnostics reductions (8 per grain, 2 per pair slot);
    // optional bindings, checked when their kernels run.
    float4 *diagnosticRows, *fabricRows;
    uint64_t diagnosticBytes, fabricBytes;
    // Translating planes (c.wallCount rows; per-(wall, grain) rows at wall * N + grain).
    float4 *wallPlane, *wallRemainder, *wallRows, *wallAccumulated;
    const float4 *wallVelocity, *wallMaterial;
    const uint32_t* wallBoxIds;
};

// Library workspace (device arrays), grown on demand.
struct W {
    uint32_t* counters;   // [0] task cursor, [1] staged count, [2] queried nodes, [3] rebuilds
    uint32_t* taskCount;  // 2P
    uint32_t* taskOffset; // 2P + 1
    uint32_t* adjCount;   // N
    float4* omegaRow;     // N: each grain's angular velocity for the resolve (omega(), computed once)
    uint32_t* adjOffset;  // N
    uint32_t* adjCursor;  // N
    uint32_t* adjList;    // 2P
    uint32_t* sums;       // scan tile sums
    uint2* stageKey;      // C
    uint4* stageMeta;     // C
    float4* stagePd;      // C
    float4* stageNa;      // C
    unsigned long long* totals;  // steps, directed pairs, queried nodes, contacts (cumulative)
    float4* wide;         // 4 rows per BVH node: both child boxes + child references
    // Per-step copies of template geometry in the order the collection reads it:
    // triangle rows with the face normal in their w components (vertex k's row
    // carries normal component k), and the surface nodes in sample-tree key order
    // (key k holds node meshSampleIndices[k] of its template), so a leaf's nodes
    // are contiguous rows.
    float4* triPacked;    // 3 rows per triangle
    float4* sampleKeyed;  // 1 row per sample key
    uint4* clusterPacked; // per sample-tree cluster: first key, node count, |box corner| bound (float bits),
                          // first key less the template's first key (the node's row slot base)
    uint4* templateRec;   // TREC rows per template (see k_pack_samples)
    uint32_t* taskFar;    // per task: the source's node count when k_task_far passes it over whole, else 0
    float4* taskFrame;    // TFRAME rows per task, written by k_task_far for the tasks it does not pass over
    uint32_t* xcd;        // per queue q: count at [q*LINE], fill cursor at [(8+q)*LINE], pop cursor at [(16+q)*LINE]
    uint32_t* stageCount; // staged hits per queue at [q*LINE]; queue q stages into [q*stageCap, (q+1)*stageCap)
    uint32_t* xcdTasks;   // 2P task ids grouped by XCD queue, ascending target template within a queue
    uint32_t* keyCount;   // tasks per (target rank, source) key
    uint32_t* keyStart;   // exclusive scan of keyCount; consumed as fill cursors
    // Temporal node skip (opt-in): per directed-pair task an anchor (both
    // rotations, the branch, and the pair/epoch key) and one byte per source node:
    // its signed distance to the target at the anchor poses, quantized down.
    float4* anchor;       // 4 rows per task
    int8_t* nodeDist;     // nodeStride bytes per task, indexed by the node's slot (key position in
                          // its template, so a leaf's nodes are contiguous)
    uint16_t* nodeTri;    // nodeStride entries per task (by slot): 1 + the node's last nearest triangle
                          // (relative to its template's first), 0 = none
    uint16_t* leafPlane;  // clusterStride words per task: a culled leaf's separating plane (tagged), or 0
    float4* clusterSphere;  // per sample-tree cluster: box centre (xyz) and half diagonal, source frame
    int8_t* leafDist;     // clusterStride bytes per task: per sample-tree leaf, a lower bound of
                          // its nodes' signed distances at the anchor poses
    uint32_t clusterStride;
    const uint32_t* triBase;  // per template: first triangle, triangle count, first sample key
    uint32_t warm;        // 1 = warm-start the exact queries from nodeTri
    uint32_t planeCache;  // 1 = nodeTri also caches a hull-test node's separating plane (tagged)
    uint32_t nodeStride;  // bytes per task (>= nodes of any template, multiple of 64)
    uint32_t skip;        // 1 = on
    float skipLimit;      // re-anchor once the motion bound reaches this (template units)
    uint32_t rebase;      // 1 = an expired anchor is moved by its motion bound, not recomputed
    float* sdfReach;      // SDF geometry, per template: squared radius beyond which the field is >= 0
    uint32_t* sdfHits;    // SDF geometry, per directed task: bitmask of its contact nodes (count pass)
    float4* sdfLeaves;    // SDF geometry, per template 64 rows: node leaves' bounding spheres (centre, radius)
    uint16_t* sdfOrder;   // SDF geometry, per template 4,096 node ids, leaf after leaf (64 per leaf)
    uint32_t sdfWords;    // 32-bit words per task in sdfHits (>= largest node count / 32, even)
    uint32_t* wallCount;  // plane contacts per (wall, grain) task, task = wall * N + grain
    uint32_t* wallOffset; // exclusive scan of wallCount (ids follow the mesh contacts)
    // Learned cull (nn.h): per-template BF16 weights and FP32 blocks; stats on separate lines.
    uint32_t nnEnabled, nnAudit, nnTemplates;
    const uint16_t* nnW;
    const float* nnF;
    unsigned long long* nnStats;  // [0] culled, [LINE/2] audited culled nodes, [LINE] audit misses
    uint32_t stageCap, taskCap, nCap, sumsCap, wallCap;
    uint32_t pop;         // tasks taken per queue atomic (tunable)
    uint32_t leafBatch;   // exact queries: lanes holding a leaf that end a descent pass (tunable)
    uint32_t batchTasks;  // exact queries: tasks queried together per wave, 1..BATCH_MAX (tunable)
    uint32_t batchFill;   // exact queries: traversals queued that end a batch early (tunable)
    float gapReach;       // exact queries: extra search distance that measures the winner's gap (tunable)
    uint32_t stageChunk;  // staging slots a wave reserves per atomic (0: the hits of each call)
    // Unchanged-winner shortcut (with the temporal skip and the warm start): per task
    // the frame of its previous collection (qi, qj, (branch, generation), (i, j,
    // epoch, 1)); per node slot (generation << 16) | gap, the gap in GAP_QUANTUM
    // units: a lower bound, at the last query, of how much nearer its nearest
    // triangle was than any other.
    float4* lastFrame;    // 4 rows per task
    uint32_t* nodeGap;    // nodeStride words per task
    // Persistent task rows (with the temporal skip).  The per-task rows above are
    // indexed by row, not by task: pair q's tasks use rows 2 pairRow[q] and
    // 2 pairRow[q] + 1.  A pair keeps its row pair across neighbour-list rebuilds,
    // through a hash of its key (i, j); rows of pairs that left the list are freed
    // and new pairs take free ones, so only new pairs start without anchors, warm
    // triangles or gaps.
    uint32_t* pairRow;    // per listed pair: its row pair
    unsigned long long* rowKeys;  // hash slots: (i << 32) | j of a listed pair, or EMPTY_KEY
    uint32_t* rowOfSlot;  // hash slots: that pair's row pair
    uint32_t* rowMark;    // per row pair: the assignment generation that kept it
    unsigned long long* rowOwner;  // per row pair: key of the pair it was assigned to
    uint32_t* rowNeed;    // per pair: 1 when it needs a free row pair (scan input)
    uint32_t* rowNeedRank;
    uint32_t* rowFree;    // per row pair: 1 when free (scan input)
    uint32_t* rowFreeRank;
    uint32_t* rowState;   // [0] rows assigned once, [3] generation, [4] assigning in this step,
                          // [5] a task found its row pair owned by another pair (stale map)
    uint32_t rowHashMask; // hash slots - 1 (a power of two)
    uint32_t rowPairs;    // row pairs (= pair capacity)
    // Persistent pair-derived structures: the per-grain pair adjacency and the task
    // buckets depend only on the pair list and the endpoints' templates, so they are
    // rebuilt only when those differ from the ones they were built from (a neighbour
    // list rebuild, a host upload).  builtKey[k] = (i, j, shapeId[i], shapeId[j]) of
    // pair k when they were built; the header counters[DERIVED_HEAD * LINE + 0..3]
    // holds (pair count, N, templates, builtGen) and counters[DERIVED_FLAG * LINE]
    // is 1 while this step rebuilds them.  Null builtKey: rebuilt every step.
    uint4* builtKey;      // pair capacity
    uint32_t builtGen;    // host generation of the derived arrays' allocations
};

constexpr int WAVE = 64;
constexpr int COLLECT_WAVES = 4;
constexpr int QCAP = 256;             // per-wave query queue (node keys)
constexpr int MAX_CLUSTERS = 1024;    // per source template
constexpr uint32_t TREC = 8;          // rows of a template record
constexpr uint32_t TFRAME = 7;        // rows of a task frame record (k_task_far)
constexpr int STACK = 32;             // BVH traversal stack (reference depth limit)
constexpr int SCAN_TILE = 1024;       // 256 threads x 4
// Device counters that many waves update atomically each sit on their own 256-byte
// line: same-line atomics serialize in one L2 channel (~10 ns each).
constexpr int LINE = 64;
#ifndef SAD_POP
#define SAD_POP 1
#endif
#ifndef SAD_COLLECT_NN_WAVES_PER_EU
#define SAD_COLLECT_NN_WAVES_PER_EU 4
#endif
#ifndef SAD_COLLECT_WAVES_PER_EU
#define SAD_COLLECT_WAVES_PER_EU 4
#endif
#ifndef SAD_COLLECT_BLOCKS_PER_CU
#define SAD_COLLECT_BLOCKS_PER_CU 6
#endif
constexpr int POP = SAD_POP;          // default tasks taken per queue atomic

__device__ __forceinline__ void fail(const P& p, uint32_t bit) { atomicOr(p.control, bit); }
__device__ __forceinline__ uint32_t ld_volatile(const uint32_t* a) { return __atomic_load_n(a, __ATOMIC_RELAXED); }
// Streaming stores and loads (the non-temporal hint): data written once and read once
// by a later kernel, kept from displacing the reused template geometry in L2.
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t nt_u4 __attribute__((ext_vector_type(4)));
typedef uint32_t nt_u2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_stream(float4* a, float4 v) {
    __builtin_nontemporal_store(nt_f4{v.x, v.y, v.z, v.w}, reinterpret_cast<nt_f4*>(a));
}
__device__ __forceinline__ void st_stream(uint4* a, uint4 v) {
    __builtin_nontemporal_store(nt_u4{v.x, v.y, v.z, v.w}, reinterpret_cast<nt_u4*>(a));
}
__device__ __forceinline__ void st_stream(uint2* a, uint2 v) {
    __builtin_nontemporal_store(nt_u2{v.x, v.y}, reinterpret_cast<nt_u2*>(a));
}
__device__ __forceinline__ float4 ld_stream(const float4* a) {
    const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(a));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ld_stream(const uint4* a) {
    const nt_u4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u4*>(a));
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 ld_stream(const uint2* a) {
    const nt_u2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u2*>(a));
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------- shared helpers (solver.slang:130-186)
// Periodicity: Config `periodic` is 0 (none) or 1 (every axis) as the reference
// sends it; 8 | mask (x = 1, y = 2, z = 4) selects axes (no reference counterpart:
// walls across one axis of an otherwise periodic cell).  0 and 1 take exactly the
// reference's paths.
__device__ __forceinline__ bool periodicAxis(const Cfg& c, int a) {
    return c.periodic == 1u || (c.periodic >= 8u && ((c.periodic >> a) & 1u));
}
__device__ __forceinline__ bool anyPeriodic(const Cfg& c) {
    return c.periodic == 1u || (c.periodic >= 8u && (c.periodic & 7u) != 0u);
}
__device__ __forceinline__ float axisOf(v3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ v3 centerBranch(const P& p, uint32_t i, uint32_t j) {  // FP64 (cpu_scheduling.slang:10-15)
    float4 c = p.cell[0], pj = p.position[j], pi = p.position[i];
    float4 rj = p.positionRemainder[j], ri = p.positionRemainder[i];
    double bx = c.x, by = c.y, bz = c.z;
    double dx = (double(pj.x) - double(pi.x)) + (double(rj.x) - double(ri.x));
    double dy = (double(pj.y) - double(pi.y)) + (double(rj.y) - double(ri.y));
    double dz = (double(pj.z) - double(pi.z)) + (double(rj.z) - double(ri.z));
    if (periodicAxis(p.c, 0)) dx = dx - floor(dx / bx + 0.5) * bx;
    if (periodicAxis(p.c, 1)) dy = dy - floor(dy / by + 0.5) * by;
    if (periodicAxis(p.c, 2)) dz = dz - floor(dz / bz + 0.5) * bz;
    return {float(dx), float(dy), float(dz)};
}
__device__ v3 omega(const P& p, uint32_t i) {
    v4 q = v4of(p.rotation[i]);
    v3 b = qinv(q, xyz(p.momentum[i]) + xyz(p.momentumRemainder[i]));
    v3 w = {dot(xyz(p.invI[3 * i]), b), dot(xyz(p.invI[3 * i + 1]), b), dot(xyz(p.invI[3 * i + 2]), b)};
    return qrot(q, w);
}
__device__ float rotationalEnergy(const P& p, uint32_t i, v3 l) {
    v3 b = qinv(v4of(p.rotation[i]), l);
    v3 w = {dot(xyz(p.invI[3 * i]), b), dot(xyz(p.invI[3 * i + 1]), b), dot(xyz(p.invI[3 * i + 2]), b)};
    return 0.5f * dot(b, w);
}
__device__ void accumulateRow(const P& p, uint32_t id, v4 inc) {
    v4 high = v4of(p.accumulated[id]), low = v4of(p.accumulatedRemainder[id]);
    compensated4(high, low, inc);
    p.accumulated[id] = f4(high);
    p.accumulatedRemainder[id] = f4(low);
}
__device__ void accumulateNormalLoss(const P& p, uint32_t id, float dt) {
    float4 value = p.normalDamping[id];
    v3 high = {value.x, 0, 0}, low = {value.y, 0, 0};
    compensated3(high, low, {value.z * dt, 0, 0});
    value.x = high.x;
    value.y = low.x;
    p.normalDamping[id] = value;
}
__device__ float cellPower(const P& p, uint32_t i) {
    v3 diagonal = {p.particleStress[3 * i].x, p.particleStress[3 * i + 1].y, p.particleStress[3 * i + 2].z};
    v3 v = xyz(p.velocity[i]);
    return -dot(diagonal + p.body[i].y * v * v, xyz(p.cell[1]));
}

// ---------------------------------------------------------------- translating planes (solver.slang:135-145, 578-687)
// Work done on a grain by the walls over dt (solver.slang:135-145): each wall's
// compensated running total advances by dt times the last gathered wall power.
__device__ float wallStepWork(const P& p, uint32_t id, float dt) {
    float total = 0;
    for (uint32_t wall = 0; wall < p.c.wallCount; wall++) {
        uint32_t index = wall * p.c.N + id;
        float increment = dt * p.wallRows[3 * index + 1].w;
        float4 state = p.wallAccumulated[index];
        v3 high = {state.x, 0, 0}, low = {state.y, 0, 0};
        compensated3(high, low, {increment, 0, 0});
        p.wallAccumulated[index] = make_float4(high.x, low.x, 0, 0);
        p.wallRows[3 * index + 2] = make_float4(high.x + low.x, 0, 0, 0);
        total += increment;
    }
    return total;
}
// Double-float plane distance (solver.slang:578-610): Knuth two-sums with exact
// FMA product errors, so a far-off plane offset does not swamp the node arm.
struct f2 { float hi, lo; };
__device__ __forceinline__ f2 wallTwoSum(float a, float b) {
    float high = a + b, v = high - a;
    return {high, (a - (high - v)) + (b - v)};
}
__device__ __forceinline__ f2 wallAdd(f2 a, float b) {
    f2 sum = wallTwoSum(a.hi, b);
    return wallTwoSum(sum.hi, sum.lo + a.lo);
}
__device__ __forceinline__ f2 wallAddDot(f2 sum, v3 a, v3 b) {
    const float pr[3] = {a.x * b.x, a.y * b.y, a.z * b.z};
    const float er[3] = {__builtin_fmaf(a.x, b.x, -pr[0]), __builtin_fmaf(a.y, b.y, -pr[1]), __builtin_fmaf(a.z, b.z, -pr[2])};
    for (int axis = 0; axis < 3; axis++) {
        sum = wallAdd(sum, pr[axis]);
        sum = wallAdd(sum, er[axis]);
    }
    return sum;
}
__device__ f2 wallCentreDistance(const P& p, uint32_t body, uint32_t wall) {
    float4 plane = p.wallPlane[wall], low = p.wallRemainder[wall];
    v3 x = xyz(p.position[body]), xl = xyz(p.positionRemainder[body]);
    f2 sum = {-plane.w, -low.w};
    sum = wallAddDot(sum, xyz(plane), x);
    sum = wallAddDot(sum, xyz(plane), xl);
    sum = wallAddDot(sum, xyz(low), x);
    return wallAddDot(sum, xyz(low), xl);
}
__device__ float localWallDistance(f2 centre, float4 plane, float4 low, v3 arm) {
    f2 sum = wallAddDot(centre, xyz(plane), arm);
    sum = wallAddDot(sum, xyz(low), arm);
    return sum.hi + sum.lo;
}
__device__ void binIndex(const P& p, uint32_t i, uint32_t b[3]) {
    v3 g = floor3(xyz(p.position[i]) / xyz(p.cell[0]) * mk3(float(p.c.bins[0]), float(p.c.bins[1]), float(p.c.bins[2])));
    b[0] = min(uint32_t(g.x), p.c.bins[0] - 1);
    b[1] = min(uint32_t(g.y), p.c.bins[1] - 1);
    b[2] = min(uint32_t(g.z), p.c.bins[2] - 1);
}
__device__ __forceinline__ uint32_t binFlat(const P& p, const uint32_t b[3]) { return (b[0] * p.c.bins[1] + b[1]) * p.c.bins[2] + b[2]; }
__device__ __forceinline__ uint32_t modIndex(int a, uint32_t n) { return uint32_t((a % int(n) + int(n)) % int(n)); }

// ---------------------------------------------------------------- mesh predicates (solver.slang:259-305)
__device__ bool meshMayContain(const P& p, uint32_t shape, v3 q) {
    uint32_t root = p.meshRoots[shape].x;
    v3 low = xyz(p.meshMin[root]), high = xyz(p.meshMax[root]);
    if (q.x < low.x || q.y < low.y || q.z < low.z || q.x > high.x || q.y > high.y || q.z > high.z) return false;
    v3 direction = (q - 0.5f * (low + high)) / (high - low);
    uint4 info = p.meshHullInfo[6 * shape + dominantGroup(direction)];
    for (uint32_t k = 0; k < info.y; k++) {
        float4 plane = p.meshPlanes[info.x + k];
        if (dot(xyz(plane), q) + plane.w > 0) return false;
    }
    return true;
}
__device__ bool meshBoxMayContact(const P& p, uint32_t shape, v3 low, v3 high) {
    uint32_t root = p.meshRoots[shape].x;
    v3 tl = xyz(p.meshMin[root]), th = xyz(p.meshMax[root]);
    if (high.x < tl.x || high.y < tl.y || high.z < tl.z || low.x > th.x || low.y > th.y || low.z > th.z) return false;
    v3 center = 0.5f * (low + high), half = 0.5f * (high - low);
    v3 direction = (center - 0.5f * (tl + th)) / (th - tl);
    uint4 info = p.meshHullInfo[6 * shape + dominantGroup(direction)];
    for (uint32_t k = 0; k < info.y; k++) {
        float4 plane = p.meshPlanes[info.x + k];
        if (dot(xyz(plane), center) + plane.w - dot(fabs3(xyz(plane)), half) > 0) return false;
    }
    return true;
}
__device__ bool meshClusterMayContact(const P& p, v4 qs, v4 qd, float ss, float sd, v3 branch, uint32_t dstShape,
                                      uint32_t cluster) {
    v3 sl = xyz(p.meshClusterMin[cluster]), sh = xyz(p.meshClusterMax[cluster]);
    v3 low = {1e30f, 1e30f, 1e30f}, high = {-1e30f, -1e30f, -1e30f};
    for (uint32_t corner = 0; corner < 8; corner++) {
        v3 q = {(corner & 1) ? sh.x : sl.x, (corner & 2) ? sh.y : sl.y, (corner & 4) ? sh.z : sl.z};
        q = qinv(qd, qrot(qs, q * ss) - branch) / sd;
        low = fmin3(low, q);
        high = fmax3(high, q);
    }
    float pad = 64.0f * DEM_EPS * (1.0f + maxabs(low) + maxabs(high) + length(branch) / sd);
    low = low - mk3(pad, pad, pad);
    high = high + mk3(pad, pad, pad);
    return meshBoxMayContact(p, dstShape, low, high);
}

// ---------------------------------------------------------------- hull-group tests, batched loads
// meshMayContain / meshBoxMayContact (solver.slang:259-288) are conjunctions over
// the planes of one group, so any evaluation order gives the reference's answer.
// The reference loop waits for each plane's load before testing it; smooth grains
// have 100-500 planes per group, so here four planes are loaded together (indices
// clamped into the group) and tested with the reference's exact FP32 expressions.
constexpr uint32_t PLANE_BATCH = 4;
__device__ __forceinline__ bool planesPass(const float4* __restrict__ planes, uint32_t first, uint32_t count, v3 x,
                                           v3 half, bool box) {
    for (uint32_t k = 0; k < count; k += PLANE_BATCH) {
        float4 pl[PLANE_BATCH];
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++) pl[j] = planes[first + min(k + j, count - 1)];
        bool bad = false;
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++) {
            float v = box ? dot(xyz(pl[j]), x) + pl[j].w - dot(fabs3(xyz(pl[j])), half) : dot(xyz(pl[j]), x) + pl[j].w;
            bad |= v > 0;
        }
        if (bad) return false;
    }
    return true;
}
// meshMayContainFast's test with its evidence: 0 when q passes, otherwise a
// positive lower bound on q's distance to the mesh (how far q lies outside the
// root box, or outside the first violated hull plane of its group; the planes
// have unit normals and contain the mesh).
__device__ __forceinline__ float planesExcess(const float4* __restrict__ planes, uint32_t first, uint32_t count, v3 x) {
    for (uint32_t k = 0; k < count; k += PLANE_BATCH) {
        float4 pl[PLANE_BATCH];
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++) pl[j] = planes[first + min(k + j, count - 1)];
        float worst = 0.0f;
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++) worst = fmaxf(worst, dot(xyz(pl[j]), x) + pl[j].w);
        if (worst > 0) return worst;
    }
    return 0.0f;
}
__device__ float meshOutsideFast(const P& p, uint32_t shape, v3 q) {
    uint32_t root = p.meshRoots[shape].x;
    v3 low = xyz(p.meshMin[root]), high = xyz(p.meshMax[root]);
    if (q.x < low.x || q.y < low.y || q.z < low.z || q.x > high.x || q.y > high.y || q.z > high.z)
        return fmaxf(fmaxf(fmaxf(low.x - q.x, q.x - high.x), fmaxf(low.y - q.y, q.y - high.y)),
                     fmaxf(low.z - q.z, q.z - high.z));
    v3 direction = (q - 0.5f * (low + high)) / (high - low);
    uint4 info = p.meshHullInfo[6 * shape + dominantGroup(direction)];
    return info.y == 0 ? 0.0f : planesExcess(p.meshPlanes, info.x, info.y, q);
}
__device__ bool meshMayContainFast(const P& p, const W& w, uint32_t shape, v3 q) {
    uint32_t root = p.meshRoots[shape].x;
    v3 low = xyz(p.meshMin[root]), high = xyz(p.meshMax[root]);
    if (q.x < low.x || q.y < low.y || q.z < low.z || q.x > high.x || q.y > high.y || q.z > high.z) return false;
    v3 direction = (q - 0.5f * (low + high)) / (high - low);
    uint4 info = p.meshHullInfo[6 * shape + dominantGroup(direction)];
    return info.y == 0 || planesPass(p.meshPlanes, info.x, info.y, q, mk3(0, 0, 0), false);
}
// meshClusterMayContactFast's test with its evidence: 0 when the cluster passes,
// otherwise a positive lower bound on the distance from any point of the cluster
// box to the mesh (the gap between the padded transformed box and the root box,
// or how far the box lies outside the first violated hull plane of its group).
__device__ float meshClusterOutsideFast(const P& p, v4 qs, v4 qd, float ss, float sd, v3 branch, uint32_t dstShape,
                                        uint32_t cluster) {
    v3 sl = xyz(p.meshClusterMin[cluster]), sh = xyz(p.meshClusterMax[cluster]);
    v3 low = {1e30f, 1e30f, 1e30f}, high = {-1e30f, -1e30f, -1e30f};
    for (uint32_t corner = 0; corner < 8; corner++) {
        v3 q = {(corner & 1) ? sh.x : sl.x, (corner & 2) ? sh.y : sl.y, (corner & 4) ? sh.z : sl.z};
        q = qinv(qd, qrot(qs, q * ss) - branch) / sd;
        low = fmin3(low, q);
        high = fmax3(high, q);
    }
    float pad = 64.0f * DEM_EPS * (1.0f + maxabs(low) + maxabs(high) + length(branch) / sd);
    low = low - mk3(pad, pad, pad);
    high = high + mk3(pad, pad, pad);
    uint32_t root = p.meshRoots[dstShape].x;
    v3 tl = xyz(p.meshMin[root]), th = xyz(p.meshMax[root]);
    if (high.x < tl.x || high.y < tl.y || high.z < tl.z || low.x > th.x || low.y > th.y || low.z > th.z)
        return fmaxf(fmaxf(fmaxf(tl.x - high.x, low.x - th.x), fmaxf(tl.y - high.y, low.y - th.y)),
                     fmaxf(tl.z - high.z, low.z - th.z));
    v3 center = 0.5f * (low + high), half = 0.5f * (high - low);
    v3 direction = (center - 0.5f * (tl + th)) / (th - tl);
    uint4 info = p.meshHullInfo[6 * dstShape + dominantGroup(direction)];
    if (info.y == 0) return 0.0f;
    for (uint32_t k = 0; k < info.y; k += PLANE_BATCH) {
        float4 pl[PLANE_BATCH];
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++) pl[j] = p.meshPlanes[info.x + min(k + j, info.y - 1)];
        float worst = 0.0f;
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++)
            worst = fmaxf(worst, dot(xyz(pl[j]), center) + pl[j].w - dot(fabs3(xyz(pl[j])), half));
        if (worst > 0) return worst;
    }
    return 0.0f;
}
__device__ bool meshClusterMayContactFast(const P& p, const W& w, v4 qs, v4 qd, float ss, float sd, v3 branch,
                                          uint32_t dstShape, uint32_t cluster) {
    v3 sl = xyz(p.meshClusterMin[cluster]), sh = xyz(p.meshClusterMax[cluster]);
    v3 low = {1e30f, 1e30f, 1e30f}, high = {-1e30f, -1e30f, -1e30f};
    for (uint32_t corner = 0; corner < 8; corner++) {
        v3 q = {(corner & 1) ? sh.x : sl.x, (corner & 2) ? sh.y : sl.y, (corner & 4) ? sh.z : sl.z};
        q = qinv(qd, qrot(qs, q * ss) - branch) / sd;
        low = fmin3(low, q);
        high = fmax3(high, q);
    }
    float pad = 64.0f * DEM_EPS * (1.0f + maxabs(low) + maxabs(high) + length(branch) / sd);
    low = low - mk3(pad, pad, pad);
    high = high + mk3(pad, pad, pad);
    uint32_t root = p.meshRoots[dstShape].x;
    v3 tl = xyz(p.meshMin[root]), th = xyz(p.meshMax[root]);
    if (high.x < tl.x || high.y < tl.y || high.z < tl.z || low.x > th.x || low.y > th.y || low.z > th.z) return false;
    v3 center = 0.5f * (low + high), half = 0.5f * (high - low);
    v3 direction = (center - 0.5f * (tl + th)) / (th - tl);
    uint4 info = p.meshHullInfo[6 * dstShape + dominantGroup(direction)];
    return info.y == 0 || planesPass(p.meshPlanes, info.x, info.y, center, half, true);
}

// Nearest-feature BVH query (cpu_scheduling.slang:18-48).  The pruning distance
// of a popped node is recomputed (identical to the CPU's cached push-time value);
// the stack holds node offsets from the template root, lane-strided in LDS.
__device__ v4 meshDistance(const P& p, uint32_t shape, v3 q, uint16_t* stack) {
    const uint32_t root = p.meshRoots[shape].x;
    uint32_t top = 1;
    stack[0] = 0;
    float best = 1e30f;
    v3 nearest = {0, 0, 0};
    uint32_t nearestNormal = 0;
    bool found = false;
    while (top > 0) {
        --top;
        uint32_t node = root + stack[top * WAVE];
        if (boxDistance(xyz(p.meshMin[node]), xyz(p.meshMax[node]), q) > best) continue;
        uint4 link = p.meshLinks[node];
        if (link.z == 0) {
            uint32_t a = node + 1, b = p.meshLinks[a].x;
            float da = boxDistance(xyz(p.meshMin[a]), xyz(p.meshMax[a]), q);
            float db = boxDistance(xyz(p.meshMin[b]), xyz(p.meshMax[b]), q);
            if (da < db) { uint32_t s = a; a = b; b = s; float t = da; da = db; db = t; }
            if (top + 2 > STACK) { fail(p, 1024); return {0, 0, 0, 1e20f}; }
            if (da <= best) stack[(top++) * WAVE] = uint16_t(a - root);
            if (db <= best) stack[(top++) * WAVE] = uint16_t(b - root);
            continue;
        }
        for (uint32_t k = 0; k < link.z; k++) {
            uint32_t triangle = link.y + k, feature;
            v3 c = closestTriangle(q, xyz(p.meshTriangles[3 * triangle]), xyz(p.meshTriangles[3 * triangle + 1]),
                                   xyz(p.meshTriangles[3 * triangle + 2]), feature);
            float distance = dot(q - c, q - c);
            if (distance < best) { best = distance; nearest = c; nearestNormal = 7 * triangle + feature; found = true; }
        }
    }
    v3 pseudo = found ? xyz(p.meshNormals[nearestNormal]) : mk3(0, 0, 0);
    float distance = sqrtf(best);
    float side = dot(q - nearest, pseudo) < 0 ? -1.0f : 1.0f;
    v3 normal = distance > 1e-12f ? side * (q - nearest) / distance : pseudo;
    return mk4(normal, side * distance);
}

// Wide node rows derived from the reference BVH arrays (meshMin/meshMax/meshLinks):
// row 0 = (left child min, left ref), row 1 = (left child max, right ref),
// row 2 = (right child min, -), row 3 = (right child max, -).  A ref is a node index,
// or LEAF | (count - 1) << 28 | first triangle for a leaf.  The traversal below
// visits nodes in the reference order and compares the same FP32 box distances,
// cached at push time as the CPU variant does (cpu_scheduling.slang:18-48).
constexpr uint32_t LEAF = 0x80000000u;
constexpr int WSTACK = 16;  // max depth 14 => at most 15 pending entries (host-checked)

__device__ __forceinline__ uint32_t nodeRef(const P& p, uint32_t node) {
    uint4 l = p.meshLinks[node];
    return l.z == 0 ? node : (LEAF | ((l.z - 1) << 28) | l.y);
}
__global__ void k_wide_nodes(P p, W w, uint32_t nodes) {
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < nodes; n += gridDim.x * blockDim.x) {
        uint4 link = p.meshLinks[n];
        if (link.z != 0) continue;
        uint32_t a = n + 1, b = p.meshLinks[a].x;
        float4 la = p.meshMin[a], ha = p.meshMax[a], lb = p.meshMin[b], hb = p.meshMax[b];
        w.wide[4 * size_t(n)] = make_float4(la.x, la.y, la.z, __uint_as_float(nodeRef(p, a)));
        w.wide[4 * size_t(n) + 1] = make_float4(ha.x, ha.y, ha.z, __uint_as_float(nodeRef(p, b)));
        w.wide[4 * size_t(n) + 2] = make_float4(lb.x, lb.y, lb.z, 0.0f);
        w.wide[4 * size_t(n) + 3] = make_float4(hb.x, hb.y, hb.z, 0.0f);
    }
}
__global__ void k_pack_triangles(P p, W w) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < p.triangles; t += gridDim.x * blockDim.x) {
        const float4 n = p.meshNormals[7 * size_t(t)];
        const float4 a = p.meshTriangles[3 * size_t(t)], b = p.meshTriangles[3 * size_t(t) + 1],
                     c = p.meshTriangles[3 * size_t(t) + 2];
        w.triPacked[3 * size_t(t)] = make_float4(a.x, a.y, a.z, n.x);
        w.triPacked[3 * size_t(t) + 1] = make_float4(b.x, b.y, b.z, n.y);
        w.triPacked[3 * size_t(t) + 2] = make_float4(c.x, c.y, c.z, n.z);
    }
}
// One block per template: its record, its clusters' packed rows, then its leaves'
// keys, 64 lanes per leaf.  A cluster's bound on |p| over its box is the skip
// test's expression.  The record holds what a collection task reads of its target
// template, one row per lane of a single load:
//   0  root box low (xyz), root node (w bits)   1  root box high (xyz), root node ref (w bits)
//   2-4  hull groups 0-5 as (first, count) pairs
//   5  first triangle and triangle count of the BVH (0, 0 without warm starts), first key
//   6  the largest corner-radius bound over its leaves (float bits)
__global__ __launch_bounds__(256) void k_pack_samples(P p, W w) {
    __shared__ uint32_t widest;  // float bits of a nonnegative value: ordered as unsigned
    if (threadIdx.x == 0) widest = 0u;
    __syncthreads();
    const uint32_t t = blockIdx.x;
    if (t >= p.c.shapes) return;
    const uint4 info = p.meshClusterInfo[t];
    const uint32_t base = p.sampleInfo[t].x;
    uint32_t first = info.x;  // the template's first leaf (leaves are laid out in preorder)
    while (first + 1 < info.x + info.y && p.meshClusterRanges[first].y == 0) first++;
    const uint32_t keysBase = info.y ? p.meshClusterRanges[first].x : 0u;
    if (threadIdx.x < 6 && t < p.templates) {
        const uint32_t row = threadIdx.x, root = p.meshRoots[t].x;
        uint4 r = make_uint4(0u, 0u, 0u, 0u);
        if (row == 0) {
            const float4 lo = p.meshMin[root];
            r = make_uint4(__float_as_uint(lo.x), __float_as_uint(lo.y), __float_as_uint(lo.z), root);
        } else if (row == 1) {
            const float4 hi = p.meshMax[root];
            r = make_uint4(__float_as_uint(hi.x), __float_as_uint(hi.y), __float_as_uint(hi.z), nodeRef(p, root));
        } else if (row < 5) {
            const uint4 a = p.meshHullInfo[6 * t + 2 * (row - 2)], b = p.meshHullInfo[6 * t + 2 * (row - 2) + 1];
            r = make_uint4(a.x, a.y, b.x, b.y);
        } else if (row == 5) {
            r = w.warm ? make_uint4(w.triBase[3 * t], w.triBase[3 * t + 1], keysBase, 0u)
                       : make_uint4(0u, 0u, keysBase, 0u);
        }
        w.templateRec[TREC * size_t(t) + row] = r;
    }
    for (uint32_t c = info.x + threadIdx.x; c < info.x + info.y && c < p.clusters; c += blockDim.x) {
        const uint4 range = p.meshClusterRanges[c];
        const v3 lo = xyz(p.meshClusterMin[c]), hi = xyz(p.meshClusterMax[c]);
        const float radius = length(fmax3(fabs3(lo), fabs3(hi)));
        w.clusterPacked[c] = make_uint4(range.x, range.y, __float_as_uint(radius), range.y ? range.x - keysBase : 0u);
        w.clusterSphere[c] = f4(0.5f * (lo + hi), 0.5f * length(hi - lo));
        if (range.y > 0) atomicMax(&widest, __float_as_uint(radius));
    }
    __syncthreads();
    if (threadIdx.x == 0 && t < p.templates) {
        w.templateRec[TREC * size_t(t) + 6] = make_uint4(widest, 0u, 0u, 0u);
        w.templateRec[TREC * size_t(t) + 7] = make_uint4(0u, 0u, 0u, 0u);
    }
    for (uint32_t k = threadIdx.x; k < info.y * 64u; k += blockDim.x) {
        const uint4 range = p.meshClusterRanges[info.x + k / 64];
        const uint32_t lane = k % 64;
        if (lane < range.y && range.x + lane < p.sampleKeys)
            w.sampleKeyed[range.x + lane] = p.sample[base + p.meshSampleIndices[range.x + lane]];
    }
}
// Cached push distances on the LDS stack: exact FP32, or BF16 truncated toward
// zero (every distance is >= 0, so the stored value never exceeds the true one).
// A rounded-down distance can only make the pop-time test `dist > best` fail more
// often, i.e. visit nodes the exact test would skip; such a node's triangles are
// all farther than `best`, so they cannot replace it under the strict `<`, and the
// traversal order (decided at push time from the exact values) is unchanged: the
// result is identical, with half the stack LDS.
__device__ __forceinline__ void dstore(float* p, float v) { *p = v; }
__device__ __forceinline__ float dload(const float* p) { return *p; }
__device__ __forceinline__ void dstore(uint16_t* p, float v) { *p = uint16_t(__float_as_uint(v) >> 16); }
__device__ __forceinline__ float dload(const uint16_t* p) { return __uint_as_float(uint32_t(*p) << 16); }

template <typename D>
__device__ v4 meshDistanceWide(const P& p, const float4* __restrict__ wide, uint32_t shape, v3 q, uint32_t* sref,
                               D* sdist) {
    const uint32_t root = p.meshRoots[shape].x;
    uint32_t top = 1;
    sref[0] = nodeRef(p, root);
    dstore(sdist, boxDistance(xyz(p.meshMin[root]), xyz(p.meshMax[root]), q));
    float best = 1e30f;
    v3 nearest = {0, 0, 0};
    uint32_t nearestNormal = 0;
    bool found = false;
    while (top > 0) {
        --top;
        uint32_t ref = sref[top * WAVE];
        if (dload(sdist + top * WAVE) > best) continue;
        if ((ref & LEAF) == 0) {
            const float4* node = wide + 4 * size_t(ref);
            float4 w0 = node[0], w1 = node[1], w2 = node[2], w3 = node[3];
            float da = boxDistance(xyz(w0), xyz(w1), q), db = boxDistance(xyz(w2), xyz(w3), q);
            uint32_t a = __float_as_uint(w0.w), b = __float_as_uint(w1.w);
            if (da < db) { uint32_t s = a; a = b; b = s; float t = da; da = db; db = t; }
            if (top + 2 > WSTACK) { fail(p, 1024); return {0, 0, 0, 1e20f}; }
            if (da <= best) { sref[top * WAVE] = a; dstore(sdist + top * WAVE, da); top++; }
            if (db <= best) { sref[top * WAVE] = b; dstore(sdist + top * WAVE, db); top++; }
            continue;
        }
        uint32_t first = ref & 0x0fffffffu, count = ((ref >> 28) & 3u) + 1;
        for (uint32_t k = 0; k < count; k++) {
            uint32_t triangle = first + k, feature;
            const float4* t = p.meshTriangles + 3 * size_t(triangle);
            v3 c = closestTriangle(q, xyz(t[0]), xyz(t[1]), xyz(t[2]), feature);
            float distance = dot(q - c, q - c);
            if (distance < best) { best = distance; nearest = c; nearestNormal = 7 * triangle + feature; found = true; }
        }
    }
    v3 pseudo = found ? xyz(p.meshNormals[nearestNormal]) : mk3(0, 0, 0);
    float distance = sqrtf(best);
    float side = dot(q - nearest, pseudo) < 0 ? -1.0f : 1.0f;
    v3 normal = distance > 1e-12f ? side * (q - nearest) / distance : pseudo;
    return mk4(normal, side * distance);
}

// ---------------------------------------------------------------- integrator (solver.slang:859-920)
__global__ void k_begin_step(P p) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N || ld_volatile(p.control) != 0) return;
    float dt = p.c.dt;
    v3 rate = xyz(p.cell[1]);
    accumulateNormalLoss(p, id, 0.5f * dt);
    float4 body = p.body[id];
    v3 v = xyz(p.velocity[id]), vl = xyz(p.velocityRemainder[id]);
    v3 l = xyz(p.momentum[id]), ll = xyz(p.momentumRemainder[id]);
    compensated3(v, vl, 0.5f * dt * xyz(p.force[id]) / body.y);
    compensated3(l, ll, 0.5f * dt * xyz(p.torque[id]));
    float loss = -expm1Stable(-p.c.drag * dt) * (0.5f * body.y * dot(v + vl, v + vl) + rotationalEnergy(p, id, l + ll));
    accumulateRow(p, id, {0, loss, 0.5f * dt * cellPower(p, id), wallStepWork(p, id, 0.5f * dt)});
    float damping = expm1Stable(-0.5f * p.c.drag * dt);
    compensated3(v, vl, damping * (v + vl));
    compensated3(l, ll, damping * (l + ll));
    p.momentum[id] = f4(l, 0);
    p.momentumRemainder[id] = f4(ll, 0);
    v4 q = v4of(p.rotation[id]), ql = v4of(p.rotationRemainder[id]);
    v3 w = omega(p, id);
    p.rotation[id] = f4(normalize4(q + rotationIncrement(w, 0.5f * dt, q)));
    w = omega(p, id);
    compensated4(q, ql, rotationIncrement(w, dt, q));
    if (fabsf(dot4(q, q) - 1.0f) > 2e-6f) { q = normalize4(q); ql = {0, 0, 0, 0}; }
    p.rotation[id] = f4(q);
    p.rotationRemainder[id] = f4(ql);
    v3 logarithm = xyz(p.cellLog[0]), residual = xyz(p.cellLog[1]);
    compensated3(logarithm, residual, rate * dt);
    v3 box = xyz(p.cell[2]) * mk3(dem_expf(logarithm.x), dem_expf(logarithm.y), dem_expf(logarithm.z));
    v3 x = xyz(p.position[id]), xl = xyz(p.positionRemainder[id]);
    compensated3(x, xl, expm1Stable3(rate * dt) * (x + xl) + dt * (v + vl));
    if (!finite3(x) || !finite3(xl)) { fail(p, 512); return; }
    if (p.c.periodic == 1) x = x - floor3(x / box) * box;
    else {
        if (anyPeriodic(p.c)) {
            v3 wrapped = x - floor3(x / box) * box;
            x = {periodicAxis(p.c, 0) ? wrapped.x : x.x, periodicAxis(p.c, 1) ? wrapped.y : x.y,
                 periodicAxis(p.c, 2) ? wrapped.z : x.z};
        }
        v3 s = x + xl, t = (x - box) + xl;
        for (int a = 0; a < 3; a++)
            if (!periodicAxis(p.c, a) && (axisOf(s, a) < 0 || axisOf(t, a) >= 0)) { fail(p, 2048); return; }
    }
    p.position[id] = f4(x, 0);
    p.positionRemainder[id] = f4(xl, 0);
    compensated3(v, vl, expm1Stable3(-rate * dt) * (v + vl));
    p.velocity[id] = f4(v, 0);
    p.velocityRemainder[id] = f4(vl, 0);
}
__global__ void k_advance_cell(P p) {
    if (ld_volatile(p.control) != 0) return;
    v3 logarithm = xyz(p.cellLog[0]), residual = xyz(p.cellLog[1]);
    compensated3(logarithm, residual, xyz(p.cell[1]) * p.c.dt);
    p.cellLog[0] = f4(logarithm, 0);
    p.cellLog[1] = f4(residual, 0);
    v3 box = xyz(p.cell[2]) * mk3(dem_expf(logarithm.x), dem_expf(logarithm.y), dem_expf(logarithm.z));
    p.cell[0] = f4(box, 0);
    p.control[4] = 0;
    if (p.c.periodic == 1 && p.referenceCell[1].w >= 0.5f * fminf(box.x, fminf(box.y, box.z))) fail(p, 256);
    if (p.c.periodic >= 8u)
        for (int a = 0; a < 3; a++)
            if (periodicAxis(p.c, a) && p.referenceCell[1].w >= 0.5f * axisOf(box, a)) fail(p, 256);
}
// advanceWalls (solver.slang:669-687): one thread; each plane offset moves by
// dt (n . v) in compensated form, then the optional box of six walls must still
// enclose a positive volume inside the cell.
__global__ void k_advance_walls(P p) {
    if (ld_volatile(p.control) != 0) return;
    for (uint32_t wall = 0; wall < p.c.wallCount; wall++) {
        float4 plane = p.wallPlane[wall], residual = p.wallRemainder[wall];
        v3 high = {plane.w, 0, 0}, low = {residual.w, 0, 0};
        compensated3(high, low, {dot(xyz(plane), xyz(p.wallVelocity[wall])) * p.c.dt, 0, 0});
        plane.w = high.x;
        residual.w = low.x;
        if (!isfinite(high.x) || !isfinite(low.x)) { fail(p, 512); return; }
        p.wallPlane[wall] = plane;
        p.wallRemainder[wall] = residual;
    }
    if (p.c.wallBoxEnabled != 0) {
        const float4 box = p.cell[0];
        const float extent[3] = {box.x, box.y, box.z};
        for (uint32_t axis = 0; axis < 3; axis++) {
            uint32_t lo = p.wallBoxIds[2 * axis], hi = p.wallBoxIds[2 * axis + 1];
            if (lo >= p.c.wallCount || hi >= p.c.wallCount) { fail(p, 4096); return; }  // ids validated by the host
            float low = p.wallPlane[lo].w + p.wallRemainder[lo].w;
            float high = -(p.wallPlane[hi].w + p.wallRemainder[hi].w);
            if (low < 0 || high > extent[axis] || high <= low) fail(p, 4096);
        }
    }
}
__global__ void k_check_rebuild(P p) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N) return;
    v3 d = (xyz(p.position[id]) - xyz(p.referencePosition[id])) +
           (xyz(p.positionRemainder[id]) - xyz(p.referencePositionRemainder[id]));
    v3 box = xyz(p.cell[0]);
    if (anyPeriodic(p.c)) {
        v3 w = d - floor3(d / box + mk3(0.5f, 0.5f, 0.5f)) * box;
        d = {periodicAxis(p.c, 0) ? w.x : d.x, periodicAxis(p.c, 1) ? w.y : d.y, periodicAxis(p.c, 2) ? w.z : d.z};
    }
    float deformation = length(box - xyz(p.referenceCell[0]));
    if (length(d) + deformation > 0.5f * p.c.skin) atomicOr(p.control + 4, 1u);
}
__global__ void k_finish_step(P p) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N || ld_volatile(p.control) != 0) return;
    accumulateNormalLoss(p, id, 0.5f * p.c.dt);
    float4 body = p.body[id];
    v3 v = xyz(p.velocity[id]), vl = xyz(p.velocityRemainder[id]);
    v3 l = xyz(p.momentum[id]), ll = xyz(p.momentumRemainder[id]);
    compensated3(v, vl, 0.5f * p.c.dt * xyz(p.force[id]) / body.y);
    compensated3(l, ll, 0.5f * p.c.dt * xyz(p.torque[id]));
    float loss = -expm1Stable(-p.c.drag * p.c.dt) * (0.5f * body.y * dot(v + vl, v + vl) + rotationalEnergy(p, id, l + ll));
    accumulateRow(p, id, {0, loss, 0.5f * p.c.dt * cellPower(p, id), wallStepWork(p, id, 0.5f * p.c.dt)});
    float damping = expm1Stable(-0.5f * p.c.drag * p.c.dt);
    compensated3(v, vl, damping * (v + vl));
    compensated3(l, ll, damping * (l + ll));
    if (!finite3(v) || !finite3(l)) { fail(p, 512); return; }
    p.velocity[id] = f4(v, 0);
    p.velocityRemainder[id] = f4(vl, 0);
    p.momentum[id] = f4(l, 0);
    p.momentumRemainder[id] = f4(ll, 0);
}
// gated: the whole-step path launches this whenever planes exist and skips the
// work unless some contact is viscous (spec `viscous`: the grain-grain or any
// plane's normal viscosity is positive), as the Engine does on the host.
__global__ void k_capture_kinematics(P p, int gated) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N || ld_volatile(p.control) != 0) return;
    if (gated && !(p.c.normalViscosity > 0)) {
        bool viscous = false;
        for (uint32_t wall = 0; wall < p.c.wallCount; wall++) viscous |= p.wallMaterial[wall].w > 0;
        if (!viscous) return;
    }
    p.forceVelocity[id] = f4(xyz(p.velocity[id]) + xyz(p.velocityRemainder[id]), 0);
    p.forceOmega[id] = f4(omega(p, id), 0);
}
__device__ float inverseInertiaBound(const P& p, uint32_t id) {
    v3 one = {1, 1, 1};
    return fmaxf(fmaxf(dot(fabs3(xyz(p.invI[3 * id])), one), dot(fabs3(xyz(p.invI[3 * id + 1])), one)),
                 dot(fabs3(xyz(p.invI[3 * id + 2])), one));
}
__global__ void k_check_timestep(P p) {  // solver.slang:816-832
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N || ld_volatile(p.control) != 0 || p.c.dt == 0) return;
    float dt = p.c.dt;
    float rx = p.timestepRows[id].x, ry = p.timestepRows[id].y;
    v3 v = xyz(p.velocity[id]) + xyz(p.velocityRemainder[id]);
    float speed = length(v) + dt * length(xyz(p.force[id])) / p.body[id].y;
    float spin = inverseInertiaBound(p, id) * (length(xyz(p.momentum[id]) + xyz(p.momentumRemainder[id])) + dt * length(xyz(p.torque[id])));
    float wallSpeed = 0;
    for (uint32_t wall = 0; wall < p.c.wallCount; wall++) wallSpeed = fmaxf(wallSpeed, length(xyz(p.wallVelocity[wall])));
    float motion = dt * (speed + p.body[id].z * spin + wallSpeed);
    if (!(isfinite(rx) && isfinite(ry) && isfinite(motion) && isfinite(spin))) { fail(p, 512); return; }
    v3 rate = fabs3(xyz(p.cell[1])) * dt;
    if (dt * sqrtf(rx) > 0.1f || dt * ry > 0.1f || dt * spin > 0.01f || motion > 0.25f * p.c.skin ||
        rate.x > 0.01f || rate.y > 0.01f || rate.z > 0.01f)
        fail(p, 8192);
}

// ---------------------------------------------------------------- neighbour search (solver.slang:346-392)
__global__ void k_clear_bins(P p) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (p.control[4] == 0) return;
    if (id < p.c.binCount) p.binCounts[id] = 0;
}
__global__ void k_fill_bins(P p) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N || p.control[4] == 0) return;
    v3 x = xyz(p.position[id]), xl = xyz(p.positionRemainder[id]);
    if (!finite3(x) || !finite3(xl)) { fail(p, 512); return; }
    if (p.c.periodic != 1) {
        v3 s = x + xl, t = (x - xyz(p.cell[0])) + xl;
        for (int a = 0; a < 3; a++)
            if (!periodicAxis(p.c, a) && (axisOf(s, a) < 0 || axisOf(t, a) >= 0)) { fail(p, 2048); return; }
    }
    uint32_t b[3];
    binIndex(p, id, b);
    uint32_t bin = binFlat(p, b);
    uint32_t slot = atomicAdd(p.binCounts + bin, 1u);
    if (slot < p.c.binCapacity) p.binParticles[bin * p.c.binCapacity + slot] = id;
    else fail(p, 8);
}
// Ascending ids within each bin: the order the serial fill produces.
__global__ void k_sort_bins(P p) {
    uint32_t bin = blockIdx.x * blockDim.x + threadIdx.x;
    if (bin >= p.c.binCount || p.control[4] == 0) return;
    uint32_t n = min(p.binCounts[bin], p.c.binCapacity);
    uint32_t* a = p.binParticles + size_t(bin) * p.c.binCapacity;
    for (uint32_t k = 1; k < n; k++) {
        uint32_t v = a[k];
        uint32_t m = k;
        while (m > 0 && a[m - 1] > v) { a[m] = a[m - 1]; m--; }
        a[m] = v;
    }
}
__device__ uint32_t findPairs(const P& p, uint32_t id, bool emit) {
    uint32_t b[3];
    binIndex(p, id, b);
    float cutoff = p.referenceCell[1].w;
    v3 g = mk3(cutoff, cutoff, cutoff) / xyz(p.cell[0]) * mk3(float(p.c.bins[0]), float(p.c.bins[1]), float(p.c.bins[2]));
    int reach[3] = {int(ceilf(g.x)), int(ceilf(g.y)), int(ceilf(g.z))};
    int low[3], width[3];
    for (int a = 0; a < 3; a++) {
        low[a] = max(int(b[a]) - reach[a], 0);
        int high = min(int(b[a]) + reach[a], int(p.c.bins[a]) - 1);
        width[a] = periodicAxis(p.c, a) ? int(min(uint32_t(2 * reach[a] + 1), p.c.bins[a])) : high - low[a] + 1;
    }
    float4 bi = p.body[id];
    uint32_t count = 0;
    uint32_t base = emit ? p.pairOffsets[id] : 0;
    for (int x = 0; x < width[0]; x++)
        for (int y = 0; y < width[1]; y++)
            for (int z = 0; z < width[2]; z++) {
                uint32_t c[3];
                const int at[3] = {x, y, z};
                for (int a = 0; a < 3; a++)
                    c[a] = periodicAxis(p.c, a) ? modIndex(int(b[a]) - reach[a] + at[a], p.c.bins[a])
                                                : uint32_t(low[a] + at[a]);
                uint32_t flat = binFlat(p, c);
                uint32_t occupancy = min(p.binCounts[flat], p.c.binCapacity);
                const uint32_t* slots = p.binParticles + size_t(flat) * p.c.binCapacity;
                for (uint32_t s = 0; s < occupancy; s++) {
                    uint32_t j = slots[s];
                    if (j <= id) continue;
                    v3 d = centerBranch(p, id, j);
                    float range = bi.z + p.body[j].z + p.c.skin;
                    if (dot(d, d) > range * range) continue;
                    if (emit) {
                        uint32_t index = base + count;
                        if (index < p.c.pairCapacity) p.pairs[index] = make_uint2(id, j);
                        else fail(p, 16);
                    }
                    count++;
                }
            }
    return count;
}
__global__ void k_count_pairs(P p) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id < p.c.N && p.control[4] != 0) p.pairCounts[id] = findPairs(p, id, false);
}
__global__ void k_emit_pairs(P p) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N || p.control[4] == 0) return;
    findPairs(p, id, true);
    p.referencePosition[id] = p.position[id];
    p.referencePositionRemainder[id] = p.positionRemainder[id];
    if (id == 0) {
        p.control[1] = p.pairOffsets[p.c.N - 1] + p.pairCounts[p.c.N - 1];
        p.referenceCell[0] = p.cell[0];
    }
}

// ---------------------------------------------------------------- generic exclusive scan
// Length = min(cap, lenPtr ? *lenPtr * mul : cap); gated on control[4] when `gate`.
struct ScanArgs {
    const uint32_t* in;
    uint32_t* out;
    uint32_t* sums;
    uint32_t cap;
    const uint32_t* lenPtr;
    uint32_t mul;
    const uint32_t* gate;  // skip unless *gate != 0 (null = always)
    uint32_t* total;       // optional total
};
__device__ __forceinline__ uint32_t scan_len(const ScanArgs& a) {
    if (!a.lenPtr) return a.cap;
    uint64_t n = uint64_t(*a.lenPtr) * a.mul;
    return uint32_t(min<uint64_t>(n, a.cap));
}
__device__ __forceinline__ uint32_t wave_inclusive(uint32_t v, int lane) {
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t o = __shfl_up(v, d, WAVE);
        if (lane >= d) v += o;
    }
    return v;
}
// Block (256) exclusive scan; returns exclusive prefix, writes block total.
__device__ uint32_t block_exclusive(uint32_t v, uint32_t& total) {
    __shared__ uint32_t waveSums[4];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive(v, lane);
    if (lane == 63) waveSums[wv] = inc;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < wv; k++) before += waveSums[k];
    total = waveSums[0] + waveSums[1] + waveSums[2] + waveSums[3];
    __syncthreads();
    return before + inc - v;
}
__global__ __launch_bounds__(256) void k_scan_reduce(ScanArgs a) {
    if (a.gate && *a.gate == 0) return;
    uint32_t n = scan_len(a);
    uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    uint32_t s = 0;
    for (int k = 0; k < 4; k++) {
        uint32_t idx = base + threadIdx.x * 4 + k;
        if (idx < n) s += a.in[idx];
    }
    uint32_t total;
    block_exclusive(s, total);
    if (threadIdx.x == 0) a.sums[blockIdx.x] = total;
}
__global__ __launch_bounds__(256) void k_scan_sums(ScanArgs a) {
    if (a.gate && *a.gate == 0) return;
    uint32_t n = scan_len(a);
    uint32_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t carry = 0;
    for (uint32_t start = 0; start < tiles; start += 256) {
        uint32_t idx = start + threadIdx.x;
        uint32_t v = idx < tiles ? a.sums[idx] : 0;
        uint32_t total;
        uint32_t ex = block_exclusive(v, total);
        if (idx < tiles) a.sums[idx] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && a.total) *a.total = carry;
}
__global__ __launch_bounds__(256) void k_scan_apply(ScanArgs a) {
    if (a.gate && *a.gate == 0) return;
    uint32_t n = scan_len(a);
    uint32_t base = blockIdx.x * SCAN_TILE;
    if (base >= n) return;
    uint32_t v[4], s = 0;
    for (int k = 0; k < 4; k++) {
        uint32_t idx = base + threadIdx.x * 4 + k;
        v[k] = idx < n ? a.in[idx] : 0;
        s += v[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive(s, total) + a.sums[blockIdx.x];
    for (int k = 0; k < 4; k++) {
        uint32_t idx = base + threadIdx.x * 4 + k;
        if (idx < n) a.out[idx] = ex;
        ex += v[k];
    }
}

// ---------------------------------------------------------------- persistent pair-derived structures
// The mesh collection's per-grain pair adjacency and task buckets persist between
// steps.  Each collection first compares the pair list and the endpoints' template
// ids with those the structures were built from (builtKey and its header); the
// builders run only when anything differs and record what they built from.  The
// comparison reads the list itself, so a neighbour-list rebuild, a restored
// checkpoint or any other host write is caught without host bookkeeping.  (The
// reference's experimental persistent endpoint slots validate cached keys through
// the current hash instead: examples/persistent_gather_shader.py:6-71.)
constexpr int DERIVED_FLAG = 14;  // counters line: 1 while this step rebuilds
constexpr int DERIVED_HEAD = 15;  // counters line: pair count, N, templates, builtGen
__global__ void k_derived_check(P p, W w) {
    uint32_t* flag = w.counters + DERIVED_FLAG * LINE;
    const uint32_t* head = w.counters + DERIVED_HEAD * LINE;
    const uint32_t pairs = min(p.control[1], p.c.pairCapacity);
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    if (!w.builtKey || head[0] != pairs || head[1] != p.c.N || head[2] != p.templates || head[3] != w.builtGen) {
        if (first) *flag = 1u;
        return;
    }
    bool differ = false;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < pairs; k += gridDim.x * blockDim.x) {
        const uint2 ij = p.pairs[k];
        const uint4 b = w.builtKey[k];
        differ |= b.x != ij.x || b.y != ij.y || b.z != p.shapeId[ij.x] || b.w != p.shapeId[ij.y];
    }
    if (differ) atomicOr(flag, 1u);
}
// Builders run when `gate` is null (always) or *gate != 0.
__device__ __forceinline__ bool gated(const uint32_t* gate) { return gate && ld_volatile(gate) == 0u; }

// ---------------------------------------------------------------- per-grain pair adjacency
// Pairs touching each grain in ascending pair index.
__global__ void k_adj_clear(P p, W w, const uint32_t* gate) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N || gated(gate)) return;
    w.adjCount[id] = 0;
    w.adjCursor[id] = 0;
}
__global__ void k_adj_count(P p, W w, const uint32_t* gate) {
    if (gated(gate)) return;
    uint32_t pairs = min(p.control[1], p.c.pairCapacity);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < pairs; k += gridDim.x * blockDim.x) {
        uint2 ij = p.pairs[k];
        atomicAdd(w.adjCount + ij.x, 1u);
        atomicAdd(w.adjCount + ij.y, 1u);
    }
}
// With a gate (the mesh collection) it also records what the structures are built
// from; without one (the SDF path rebuilds every step) it marks the record stale.
__global__ void k_adj_fill(P p, W w, const uint32_t* gate) {
    if (gated(gate)) return;
    uint32_t pairs = min(p.control[1], p.c.pairCapacity);
    const bool record = gate && w.builtKey;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < pairs; k += gridDim.x * blockDim.x) {
        uint2 ij = p.pairs[k];
        w.adjList[w.adjOffset[ij.x] + atomicAdd(w.adjCursor + ij.x, 1u)] = k;
        w.adjList[w.adjOffset[ij.y] + atomicAdd(w.adjCursor + ij.y, 1u)] = k;
        if (record) w.builtKey[k] = make_uint4(ij.x, ij.y, p.shapeId[ij.x], p.shapeId[ij.y]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t* head = w.counters + DERIVED_HEAD * LINE;
        head[0] = record ? pairs : 0xffffffffu;
        head[1] = p.c.N;
        head[2] = p.templates;
        head[3] = w.builtGen;
    }
}
// Each grain's pair ids in ascending order (the order the gather walks).  One wave
// per grain: the ids are staged in LDS and each is placed by its rank (ids are
// distinct), 64 at a time, so a list of n ids costs n^2/64 LDS reads per lane.
// Lists longer than ADJ_SORT_LDS (none in practice: a grain has tens of candidate
// pairs) fall back to an insertion sort by one lane.
constexpr uint32_t ADJ_SORT_LDS = 1024;
__global__ __launch_bounds__(256) void k_adj_sort(P p, W w, const uint32_t* gate) {
    __shared__ uint32_t idsMem[4][ADJ_SORT_LDS];
    const int lane = threadIdx.x & 63;
    const uint32_t id = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (id >= p.c.N || gated(gate)) return;
    uint32_t* a = w.adjList + w.adjOffset[id];
    const uint32_t n = w.adjCount[id];
    if (n <= ADJ_SORT_LDS) {
        uint32_t* ids = idsMem[threadIdx.x >> 6];
        for (uint32_t k = lane; k < n; k += WAVE) ids[k] = a[k];
        wave_sync();
        for (uint32_t k = lane; k < n; k += WAVE) {
            const uint32_t mine = ids[k];
            uint32_t rank = 0;
            for (uint32_t m = 0; m < n; m++) rank += ids[m] < mine ? 1u : 0u;
            a[rank] = mine;
        }
        return;
    }
    if (lane != 0) return;
    for (uint32_t k = 1; k < n; k++) {
        uint32_t v = a[k];
        uint32_t m = k;
        while (m > 0 && a[m - 1] > v) { a[m] = a[m - 1]; m--; }
        a[m] = v;
    }
}

// ---------------------------------------------------------------- contact collection
__global__ void k_reset_mesh(P p, W w) {  // resetMeshContacts (solver.slang:397-403) + cursors
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id == 0) {
        w.counters[2 * LINE] = 0;
        w.counters[11 * LINE] = 0;
        w.counters[12 * LINE] = 0;
        w.counters[13 * LINE] = 0;
        for (int q = 0; q < 8; q++) w.stageCount[q * LINE] = 0;
    }
    if (ld_volatile(p.control) != 0) return;
    if (id < p.c.pairCapacity) p.contactCounts[id] = 0;
    if (id == 0) { p.control[2] = 0; p.control[6] = 0; }
}
__global__ void k_prepare_pairs(P p) {  // solver.slang:469-489 (Metal frontier transforms)
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= min(p.control[1], p.c.pairCapacity) || ld_volatile(p.control) != 0) return;
    uint2 ij = p.pairs[id];
    uint32_t mask = 0;
    for (uint32_t pass = 0; pass < 2; pass++) {
        uint32_t i = pass == 0 ? ij.x : ij.y, j = pass == 0 ? ij.y : ij.x;
        v3 branch = centerBranch(p, i, j);
        if (meshClusterMayContact(p, v4of(p.rotation[i]), v4of(p.rotation[j]), p.body[i].x, p.body[j].x, branch,
                                  p.shapeId[j], p.meshClusterInfo[p.shapeId[i]].x))
            mask |= 1u << pass;
    }
    for (uint32_t pass = 0; pass < 2; pass++) {
        if ((mask & (1u << pass)) == 0) continue;
        uint32_t i = pass == 0 ? ij.x : ij.y, j = pass == 0 ? ij.y : ij.x;
        v4 qi = v4of(p.rotation[i]), qj = v4of(p.rotation[j]);
        float ratio = p.body[i].x / p.body[j].x;
        v3 a = qinv(qj, qrot(qi, mk3(ratio, 0, 0)));
        v3 b = qinv(qj, qrot(qi, mk3(0, ratio, 0)));
        v3 c = qinv(qj, qrot(qi, mk3(0, 0, ratio)));
        v3 branch = centerBranch(p, i, j), offset = qinv(qj, -branch) / p.body[j].x;
        uint32_t base = 4 * (2 * id + pass);
        p.meshPairTransform[base] = make_float4(a.x, b.x, c.x, offset.x);
        p.meshPairTransform[base + 1] = make_float4(a.y, b.y, c.y, offset.y);
        p.meshPairTransform[base + 2] = make_float4(a.z, b.z, c.z, offset.z);
        p.meshPairTransform[base + 3] = make_float4(length(branch) / p.body[j].x, 0, 0, 0);
    }
    p.meshPairMask[id] = mask;
}

// Task queues per XCD: a counting sort of the directed-pair tasks by (target
// template, source template).  Queue q = target & 7 holds its targets in ascending
// order, each target's tasks grouped by source template, so the waves of one XCD
// sweep one target BVH and, within it, one source's cluster tree and samples at a
// time, and the XCD's L2 holds both.  Libraries above SORT_PAIRS_MAX templates are
// sorted by target only.  Order inside a bucket is irrelevant to the results (a
// task's hits are ranked within the task, tasks by id).
constexpr uint32_t MAX_TEMPLATES = 4096;
constexpr uint32_t SORT_PAIRS_MAX = 1024;
struct BucketShape {
    uint32_t perQueue;  // target ranks per queue = ceil(T / 8)
    uint32_t sources;   // secondary key range (T, or 1)
    __device__ __host__ uint32_t keys() const { return 8 * perQueue * sources; }
};
__device__ __host__ inline BucketShape bucketShape(uint32_t templates) {
    return {(templates + 7) / 8, templates <= SORT_PAIRS_MAX ? templates : 1u};
}
__device__ __forceinline__ uint32_t taskKey(const P& p, uint32_t task, BucketShape b) {
    uint2 ij = p.pairs[task >> 1];
    uint32_t target = p.shapeId[(task & 1) ? ij.x : ij.y], source = p.shapeId[(task & 1) ? ij.y : ij.x];
    uint32_t rank = (target & 7) * b.perQueue + (target >> 3);
    return rank * b.sources + (b.sources > 1 ? source : 0u);
}
// The queues' pop cursors restart every collection; counts, fill cursors and
// key counts only when the buckets are rebuilt.
__global__ void k_bucket_reset(P p, W w, const uint32_t* gate) {
    if (blockIdx.x == 0 && threadIdx.x < 8) w.xcd[(16 + threadIdx.x) * LINE] = 0;
    if (gated(gate)) return;
    const BucketShape b = bucketShape(p.templates);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < b.keys(); k += gridDim.x * blockDim.x)
        w.keyCount[k] = 0;
    if (blockIdx.x == 0 && threadIdx.x < 16) w.xcd[threadIdx.x * LINE] = 0;
}
__global__ void k_bucket_count(P p, W w, const uint32_t* gate) {
    if (gated(gate)) return;
    const BucketShape b = bucketShape(p.templates);
    const uint32_t tasks = 2 * min(p.control[1], p.c.pairCapacity);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tasks; t += gridDim.x * blockDim.x)
        atomicAdd(w.keyCount + taskKey(p, t, b), 1u);
}
// Queue totals from the scanned key offsets (w.keyStart, exclusive).
__global__ void k_bucket_offsets(P p, W w, const uint32_t* gate) {
    if (gated(gate)) return;
    const BucketShape b = bucketShape(p.templates);
    if (threadIdx.x < 8) {
        const uint32_t q = threadIdx.x, span = b.perQueue * b.sources;
        const uint32_t lo = w.keyStart[q * span];
        const uint32_t hi = q == 7 ? w.counters[6 * LINE] : w.keyStart[(q + 1) * span];
        w.xcd[q * LINE] = hi - lo;
    }
}
__global__ void k_bucket_fill(P p, W w, const uint32_t* gate) {
    if (gated(gate)) return;
    const BucketShape b = bucketShape(p.templates);
    const uint32_t tasks = 2 * min(p.control[1], p.c.pairCapacity);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tasks; t += gridDim.x * blockDim.x)
        w.xcdTasks[atomicAdd(w.keyStart + taskKey(p, t, b), 1u)] = t;
}

// ---------------------------------------------------------------- persistent task rows
// Assigned when the neighbour list was rebuilt in this step (or never yet): every
// listed pair looks its key up in the hash of the previous list and keeps that
// pair's row pair; the row pairs no listed pair kept are free, and the k-th new pair
// (in list order) takes the k-th free row pair; then the hash is rebuilt from the
// new list.  Which rows a pair gets never changes a result: a task uses its row
// pair only while that row pair's owner key is its own pair (a pair list restored
// from a checkpoint may not match the map; such a task runs without the skip and
// flags the map for re-assignment at the next collection), and every row is checked
// against its anchor's key (pair and epoch) before it is trusted.
constexpr unsigned long long EMPTY_KEY = ~0ull;
__device__ __forceinline__ uint32_t rowHash(unsigned long long key) {
    key ^= key >> 33;
    key *= 0xff51afd7ed558ccdull;
    key ^= key >> 33;
    key *= 0xc4ceb9fe1a85ec53ull;
    key ^= key >> 33;
    return uint32_t(key);
}
__device__ __forceinline__ unsigned long long pairKey(uint2 ij) {
    return (static_cast<unsigned long long>(ij.x) << 32) | ij.y;
}
__global__ void k_rows_lookup(P p, W w) {
    const bool assign = p.control[4] != 0 || ld_volatile(w.rowState) == 0u || ld_volatile(w.rowState + 5) != 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) w.rowState[4] = assign ? 1u : 0u;
    if (!assign) return;
    const bool listed = ld_volatile(w.rowState) != 0u;  // the hash holds the previous list
    const uint32_t gen = w.rowState[3] + 1u;
    const uint32_t pairs = min(p.control[1], p.c.pairCapacity);
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < w.rowPairs; q += gridDim.x * blockDim.x) {
        uint32_t need = q < pairs ? 1u : 0u;
        if (need && listed) {
            const unsigned long long key = pairKey(p.pairs[q]);
            uint32_t h = rowHash(key) & w.rowHashMask;
            for (uint32_t probe = 0; probe <= w.rowHashMask; probe++) {
                const unsigned long long k = w.rowKeys[h];
                if (k == key) {
                    const uint32_t r = w.rowOfSlot[h];
                    w.pairRow[q] = r;
                    w.rowMark[r] = gen;
                    need = 0u;
                    break;
                }
                if (k == EMPTY_KEY) break;
                h = (h + 1u) & w.rowHashMask;
            }
        }
        w.rowNeed[q] = need;
    }
}
__global__ void k_rows_free(P p, W w) {
    if (w.rowState[4] == 0u) return;
    const uint32_t gen = w.rowState[3] + 1u, slots = w.rowHashMask + 1u;
    const uint32_t n = max(w.rowPairs, slots);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (k < w.rowPairs) w.rowFree[k] = w.rowMark[k] != gen ? 1u : 0u;
        if (k < slots) w.rowKeys[k] = EMPTY_KEY;  // every lookup is done: the hash is rebuilt below
    }
}
__global__ void k_rows_assign(P p, W w) {
    if (w.rowState[4] == 0u) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.rowState[3] = w.rowState[3] + 1u;
        w.rowState[0] = 1u;
        w.rowState[5] = 0u;
    }
    const uint32_t pairs = min(p.control[1], p.c.pairCapacity);
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < pairs; q += gridDim.x * blockDim.x) {
        uint32_t r = w.pairRow[q];
        if (w.rowNeed[q]) {
            // The rank-th free row pair: the last row pair whose exclusive free rank is
            // <= rank (free row pairs outnumber new pairs: there are pairCapacity of them).
            const uint32_t rank = w.rowNeedRank[q];
            uint32_t lo = 0, hi = w.rowPairs;  // first index whose rank exceeds `rank`, in (0, rowPairs]
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (w.rowFreeRank[mid] > rank) hi = mid; else lo = mid + 1;
            }
            r = lo - 1u;
            w.pairRow[q] = r;
        }
        const unsigned long long key = pairKey(p.pairs[q]);
        w.rowOwner[r] = key;
        uint32_t h = rowHash(key) & w.rowHashMask;
        for (uint32_t probe = 0; probe <= w.rowHashMask; probe++) {
            if (atomicCAS(w.rowKeys + h, EMPTY_KEY, key) == EMPTY_KEY) {
                w.rowOfSlot[h] = r;
                break;
            }
            h = (h + 1u) & w.rowHashMask;
        }
    }
}

struct TaskFrame {
    uint32_t i, j, si, sj, sampleBase;
    v4 qi, qj;
    float s_i, s_j, wi, wj;  // body scale and equivalent diameter
    v3 branch;
    float4 plane;
};
// Every lane computes the same task frame; reading it back from the first lane
// tells the compiler so, which moves it to scalar registers.
__device__ __forceinline__ uint32_t uniformU(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uniformF(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
__device__ __forceinline__ void uniformFrame(TaskFrame& f) {
    f.i = uniformU(f.i); f.j = uniformU(f.j); f.si = uniformU(f.si); f.sj = uniformU(f.sj);
    f.sampleBase = uniformU(f.sampleBase);
    f.qi = {uniformF(f.qi.x), uniformF(f.qi.y), uniformF(f.qi.z), uniformF(f.qi.w)};
    f.qj = {uniformF(f.qj.x), uniformF(f.qj.y), uniformF(f.qj.z), uniformF(f.qj.w)};
    f.s_i = uniformF(f.s_i); f.s_j = uniformF(f.s_j); f.wi = uniformF(f.wi); f.wj = uniformF(f.wj);
    f.branch = {uniformF(f.branch.x), uniformF(f.branch.y), uniformF(f.branch.z)};
    f.plane = make_float4(uniformF(f.plane.x), uniformF(f.plane.y), uniformF(f.plane.z), uniformF(f.plane.w));
}
template <typename T>
__device__ __forceinline__ T* uniformPtr(T* ptr) {
    const uint64_t v = reinterpret_cast<uint64_t>(ptr);
    return reinterpret_cast<T*>((uint64_t(uniformU(uint32_t(v >> 32))) << 32) | uniformU(uint32_t(v)));
}

// Exact part of nodeContact (solver.slang:319-343) for one queued node.
template <typename D>
__device__ bool queryNode(const P& p, const W& w, const TaskFrame& f, uint32_t node, uint32_t* sref, D* sdist,
                          float4& pd, float4& na) {
    float4 sp = p.sample[f.sampleBase + node];
    v3 arm = qrot(f.qi, xyz(sp) * f.s_i);
    v3 local = qinv(f.qj, arm - f.branch) / f.s_j;
    v4 r = meshDistanceWide(p, w.wide, f.sj, local, sref, sdist);
    float depth = -r.w * f.s_j;
    pd = f4(arm, depth);
    if (depth <= 0) return false;
    v3 grad = qrot(f.qj, xyz(r));
    float len = length(grad);
    if (len < 1e-7f || !isfinite(depth)) { fail(p, 2); return false; }
    na = f4(grad / len, sp.w * f.s_i * f.s_i);
    if (depth > p.c.maxOverlap * fminf(f.wi, f.wj)) fail(p, 4);
    return true;
}

// Rank-ordered staging of one batch of hits.  A wave reserves its staging slots
// w.stageChunk at a time (0: exactly the hits of each call), so its queue's counter
// takes one atomic per chunk instead of one per call: atomics on one word serialise.
// [resNext, resEnd) is the wave's unused reservation (wave-uniform); the slots it
// still holds when the wave exits are marked unused (task NO_TASK) for k_scatter.
constexpr uint32_t NO_TASK = 0xffffffffu;
constexpr uint32_t STAGE_CHUNK = 256;
__device__ void stageHits(const P& p, const W& w, const TaskFrame& f, uint32_t task, uint32_t queue, bool hit,
                          uint32_t node, float4 pd, float4 na, uint32_t& running, int lane, uint32_t& resNext,
                          uint32_t& resEnd) {
    uint64_t mask = __ballot(hit);
    uint32_t count = __popcll(mask);
    if (count == 0) return;
    const uint32_t avail = resEnd - resNext, first = resNext;  // ranks below avail use the current chunk
    uint32_t more = 0;
    if (count > avail) {
        const uint32_t need = count - avail, take = max(need, w.stageChunk);
        if (lane == 0) more = atomicAdd(w.stageCount + queue * LINE, take);
        more = uniformU(__shfl(more, 0));
        resNext = more + need;
        resEnd = more + take;
    } else {
        resNext += count;
    }
    if (hit) {
        uint32_t rank = __popcll(mask & lanes_below(lane));
        uint32_t local = rank < avail ? first + rank : more + (rank - avail);
        size_t pos = size_t(queue) * w.stageCap + local;
        if (local < w.stageCap) {
            st_stream(w.stageKey + pos, make_uint2(task, running + rank));
            st_stream(w.stageMeta + pos, make_uint4(f.i, f.j, node, 0));
            st_stream(w.stagePd + pos, pd);
            st_stream(w.stageNa + pos, na);
        } else {
            fail(p, 32);
        }
    }
    running += count;
}

// Widening of the warm-start bound, in template units of distance: about 30 ulps
// of a coordinate of magnitude 1, far above the closest point's rounding.
constexpr float WARM_SLACK = 4e-6f;
// Unchanged-winner shortcut.  A node's query found nearest triangle W at the point
// x0 and a lower bound L on every other triangle's squared distance (from the
// triangles it tested, the boxes it pruned and the triangles whose plane it
// skipped).  In exact arithmetic every other triangle was at least
// G = sqrt(L) - sqrt(d_W) - 2 GAP_EVAL farther, GAP_EVAL covering the FP32 error of
// one distance evaluation.  Distances are 1-Lipschitz, so after the node's point
// has moved by at most m the gap is at least G - 2m; while that stays above
// 2 GAP_EVAL, W is the unique FP32 minimum at the new point, which is the winner the
// traversal would return, so the query is just closestTriangle(W) and the stored
// gap becomes G - 2m.  The points are the FP32 query points themselves: the task's
// frame of its previous collection is kept, so the node's previous point is
// recomputed bit for bit from it, and m is their computed distance widened by
// 1e-4 relative and GAP_MOTION_ABS (1 + |x|), far above the rounding of the
// difference and of the length.
constexpr float GAP_EVAL = WARM_SLACK;
constexpr float GAP_QUANTUM = 1.0f / 1048576.0f;  // 2^-20 template units
constexpr float GAP_MOTION_ABS = 2e-6f;
struct GapCtx {
    uint32_t* row;       // this task's per-slot words, or null when the shortcut is off
    v4 qiPrev, qjPrev;   // the frame of the task's previous collection
    v3 branchPrev;
    uint32_t genPrev;    // generation written by that collection (0: none usable)
    uint32_t genNew;     // generation this collection writes
};
__device__ __forceinline__ uint32_t gapWord(uint32_t gen, float gap) {
    const float units = gap > 0.0f ? fminf(gap / GAP_QUANTUM, 65535.0f) : 0.0f;
    return (gen << 16) | uint32_t(units);  // truncation: rounds the gap down
}
__device__ __forceinline__ v3 queryPoint(const TaskFrame& f, float4 sp, v3& arm) {
    arm = qrot(f.qi, xyz(sp) * f.s_i);
    return qinv(f.qj, arm - f.branch) / f.s_j;
}
// Temporal skip constants: signed distances are stored in units of 2^-9 template
// units (clamped to +-127); a node is passed over (or queued untested) only when
// its stored distance clears the motion bound by SKIP_MARGIN, which covers the
// FP32 error of both distance evaluations and of the node transforms (~1e-5 at
// these magnitudes).
constexpr float SKIP_QUANTUM = 1.0f / 512.0f;
constexpr float SKIP_MARGIN = 1e-3f;
// The temporal skip's motion bound since a task's anchor (rows a0, a1, a2: q_i, q_j,
// branch): a node p of the source has moved at most c0 + c1 |p| in the target frame,
// and every node of the grain (bounding radius / s_i) at most `all`.
__device__ __forceinline__ void motionBound(v4 qi, v4 qj, v3 branch, float s_i, float s_j, float radius, float4 a0,
                                            float4 a1, float4 a2, float& c0, float& c1, float& all) {
    const float dqi = sqrtf(dot4(qi - v4of(a0), qi - v4of(a0)));
    const float dqj = sqrtf(dot4(qj - v4of(a1), qj - v4of(a1)));
    const float db = length(branch - xyz(a2));
    c0 = 1.001f * (2.0f * dqj * length(branch) + db) / s_j;
    c1 = 1.001f * (2.0f * dqj + 2.0f * dqi) * s_i / s_j;
    all = c0 + c1 * (radius / s_i);
}
// The anchor's least leaf bound (bits 8-15 of row 3's w) against the motion bound of
// the source's widest leaf: true when the leaf test would pass over every leaf.
__device__ __forceinline__ bool allLeavesFar(float4 a3, float c0, float c1, float widest) {
    const int leafMin = int(int8_t(uint8_t(__float_as_uint(a3.w) >> 8)));
    return float(leafMin) * SKIP_QUANTUM > c0 + c1 * widest + SKIP_MARGIN;
}
// Before the collection, one thread per task: the tasks it can pass over whole (their
// anchor valid, the skip on, every leaf far), with the collection's own inputs and
// FP32 expressions.  About half the tasks of a dense packing; the collection then
// spends one load on each of them instead of its wave-wide setup.  For every other
// task it writes the frame the collection would compute (one record instead of a
// chain of dependent loads):
//   0  i, j, source and target template        1  q_i            2  q_j
//   3  branch (xyz), s_i                        4  s_j, d_i, d_j, first sample of the source (bits)
//   5  row task (bits), flags (1 rows, 2 skip, 4 re-anchor by shrinking), c0, c1
//   6  motion bound of every node, anchored rejection plane (bits)
__global__ void k_task_far(P p, W w) {
    const uint32_t tasks = 2 * min(p.control[1], p.c.pairCapacity);
    const uint32_t epoch = ld_volatile(w.counters + 10 * LINE);
    for (uint32_t task = blockIdx.x * blockDim.x + threadIdx.x; task < tasks; task += gridDim.x * blockDim.x) {
        uint32_t far = 0u;
        const uint2 ij = p.pairs[task >> 1];
        const uint32_t r = w.pairRow[task >> 1];
        const uint32_t i = (task & 1) ? ij.y : ij.x, j = (task & 1) ? ij.x : ij.y;
        const uint32_t si = p.shapeId[i], sj = p.shapeId[j];
        const float4 bi = p.body[i], bj = p.body[j];
        const v4 qi = v4of(p.rotation[i]), qj = v4of(p.rotation[j]);
        const v3 branch = centerBranch(p, i, j);
        const bool rows = w.rowOwner[r] == pairKey(ij);
        const uint32_t rowTask = 2u * r + (task & 1u);
        float c0 = 0.0f, c1 = 0.0f, all = 0.0f;
        uint32_t flags = rows ? 1u : 0u, plane = 0xffffffffu;
        if (rows) {
            const float4* an = w.anchor + 4 * size_t(rowTask);
            const float4 a0 = an[0], a1 = an[1], a2 = an[2], a3 = an[3];
            const bool valid = (__float_as_uint(a3.w) & 0xffu) == 1u && __float_as_uint(a3.x) == i &&
                               __float_as_uint(a3.y) == j && __float_as_uint(a3.z) == epoch;
            motionBound(qi, qj, branch, bi.x, bj.x, bi.z, a0, a1, a2, c0, c1, all);
            const bool skipMode = valid && all + SKIP_MARGIN < w.skipLimit;
            const bool rebased = valid && !skipMode && w.rebase != 0u && all < 0.25f;
            flags |= (skipMode ? 2u : 0u) | (rebased ? 4u : 0u);
            plane = __float_as_uint(a2.w);
            if (skipMode) {
                const float widest = __uint_as_float(w.templateRec[TREC * size_t(si) + 6].x);
                if (allLeavesFar(a3, c0, c1, widest)) far = p.sampleInfo[si].y;
            }
        } else {
            w.rowState[5] = 1u;  // a listed pair without its row pair: re-assign at the next collection
        }
        w.taskFar[task] = far;
        if (far == 0u) {
            float4* fr = w.taskFrame + TFRAME * size_t(task);
            fr[0] = make_float4(__uint_as_float(i), __uint_as_float(j), __uint_as_float(si), __uint_as_float(sj));
            fr[1] = f4(qi);
            fr[2] = f4(qj);
            fr[3] = f4(branch, bi.x);
            fr[4] = make_float4(bj.x, bi.w, bj.w, __uint_as_float(p.sampleInfo[si].x));
            fr[5] = make_float4(__uint_as_float(rowTask), __uint_as_float(flags), c0, c1);
            fr[6] = make_float4(all, __uint_as_float(plane), 0.0f, 0.0f);
        }
    }
}
// Rounded toward zero: a lower bound of a positive distance, an upper bound of a
// negative one; NaN stores 0 (neither far nor inside).
// Four stored bounds moved toward zero by `shrink` quanta (clamped at zero).  Leaf
// bounds are read only as lower bounds of positive distances: their negative values
// are kept.
__device__ __forceinline__ uint32_t shrinkBounds(uint32_t word, uint32_t shrink, bool positiveOnly) {
    uint32_t out = 0;
    const int s = int(min(shrink, 255u));
    for (int b = 0; b < 4; b++) {
        int v = int(int8_t(uint8_t(word >> (8 * b))));
        if (v > 0) v = max(0, v - s);
        else if (v < 0 && !positiveOnly) v = min(0, v + s);
        out |= uint32_t(uint8_t(int8_t(v))) << (8 * b);
    }
    return out;
}
__device__ __forceinline__ int8_t quantizeDistance(float d) {
    if (!(d == d)) return 0;
    return int8_t(fminf(fmaxf(d * 512.0f, -127.0f), 127.0f));
}
// Wave-cooperative hull tests.  Under the temporal skip a leaf or node loop leaves
// only a few lanes with a hull test to run, and a passing test walks every plane of
// its group (50-200 on smooth grains) one lane at a time.  These forms run one
// test with the whole wave: the eight box corners are transformed one per lane and
// reduced across each group of eight lanes, and the planes are tested 64 at a time.
// Every value is formed by the one-lane forms' FP32 expressions (a min/max
// reduction is exact in any order once NaN corners are replaced by the one-lane
// form's initial +-1e30), so the verdict -- no plane value above zero, inside the
// root box -- is the same bit for bit.  A failing test returns the largest plane
// excess instead of the first, an equally valid bound on the distance.
constexpr uint32_t COOP_LEAVES = 8;  // at most this many leaves per 64-leaf block: cooperative
constexpr uint32_t COOP_NODES = 16;  // at most this many hull-tested nodes per leaf: cooperative
__device__ __forceinline__ float waveMax(float v) {
    for (int d = 1; d < WAVE; d <<= 1) v = fmaxf(v, __shfl_xor(v, d));
    return v;
}
__device__ __forceinline__ float broadcastF(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
// meshClusterOutsideFast of one wave-uniform cluster, all lanes active.
__device__ float meshClusterOutsideWave(const P& p, v4 qs, v4 qd, float ss, float sd, v3 branch, uint32_t dstShape,
                                        uint32_t cluster, int lane) {
    const v3 sl = xyz(p.meshClusterMin[cluster]), sh = xyz(p.meshClusterMax[cluster]);
    const uint32_t corner = uint32_t(lane) & 7u;
    v3 q = {(corner & 1) ? sh.x : sl.x, (corner & 2) ? sh.y : sl.y, (corner & 4) ? sh.z : sl.z};
    q = qinv(qd, qrot(qs, q * ss) - branch) / sd;
    v3 low = fmin3(mk3(1e30f, 1e30f, 1e30f), q), high = fmax3(mk3(-1e30f, -1e30f, -1e30f), q);
    for (int d = 1; d < 8; d <<= 1) {
        low = fmin3(low, mk3(__shfl_xor(low.x, d), __shfl_xor(low.y, d), __shfl_xor(low.z, d)));
        high = fmax3(high, mk3(__shfl_xor(high.x, d), __shfl_xor(high.y, d), __shfl_xor(high.z, d)));
    }
    float pad = 64.0f * DEM_EPS * (1.0f + maxabs(low) + maxabs(high) + length(branch) / sd);
    low = low - mk3(pad, pad, pad);
    high = high + mk3(pad, pad, pad);
    uint32_t root = p.meshRoots[dstShape].x;
    v3 tl = xyz(p.meshMin[root]), th = xyz(p.meshMax[root]);
    if (high.x < tl.x || high.y < tl.y || high.z < tl.z || low.x > th.x || low.y > th.y || low.z > th.z)
        return fmaxf(fmaxf(fmaxf(tl.x - high.x, low.x - th.x), fmaxf(tl.y - high.y, low.y - th.y)),
                     fmaxf(tl.z - high.z, low.z - th.z));
    v3 center = 0.5f * (low + high), half = 0.5f * (high - low);
    v3 direction = (center - 0.5f * (tl + th)) / (th - tl);
    const uint4 info = p.meshHullInfo[6 * dstShape + __builtin_amdgcn_readfirstlane(dominantGroup(direction))];
    float worst = 0.0f;
    for (uint32_t k = lane; k < info.y; k += WAVE) {
        const float4 pl = p.meshPlanes[info.x + k];
        worst = fmaxf(worst, dot(xyz(pl), center) + pl.w - dot(fabs3(xyz(pl)), half));
    }
    return waveMax(worst);
}
// meshOutsideFast of one wave-uniform point, all lanes active.
__device__ float meshOutsideWave(const P& p, uint32_t shape, v3 q, int lane) {
    uint32_t root = p.meshRoots[shape].x;
    v3 low = xyz(p.meshMin[root]), high = xyz(p.meshMax[root]);
    if (q.x < low.x || q.y < low.y || q.z < low.z || q.x > high.x || q.y > high.y || q.z > high.z)
        return fmaxf(fmaxf(fmaxf(low.x - q.x, q.x - high.x), fmaxf(low.y - q.y, q.y - high.y)),
                     fmaxf(low.z - q.z, q.z - high.z));
    v3 direction = (q - 0.5f * (low + high)) / (high - low);
    const uint4 info = p.meshHullInfo[6 * shape + __builtin_amdgcn_readfirstlane(dominantGroup(direction))];
    float worst = 0.0f;
    for (uint32_t k = lane; k < info.y; k += WAVE) {
        const float4 pl = p.meshPlanes[info.x + k];
        worst = fmaxf(worst, dot(xyz(pl), q) + pl.w);
    }
    return waveMax(worst);
}
// The target template's root box and hull-plane groups, written to the wave's LDS
// record once per task: a hull test then reads only its group's planes from global
// memory instead of a chain of root, box and group loads.  The group's (first,
// count) is indexed by the per-lane group.  mid() and ext() are the one-lane
// form's 0.5 (low + high) and high - low (same bits).
// A node's cached separating plane, kept in its warm-start slot (nodeTri) with the
// top bit set: 3 bits of hull group, 12 of index in the group.  NO_PLANE: none.
constexpr uint32_t NO_PLANE = 0u;
constexpr uint32_t PLANE_TAG = 0x8000u;
__device__ __forceinline__ uint32_t planeCode(uint32_t group, uint32_t k) {
    return k < 4096u ? PLANE_TAG | (group << 12) | k : NO_PLANE;
}
struct HullCtx {
    v3 low, high;
    uint2 range[6];
    __device__ __forceinline__ v3 mid() const { return 0.5f * (low + high); }
    __device__ __forceinline__ v3 ext() const { return high - low; }
};
// From the target's template record, row `lane` of which is `rec`.
__device__ __forceinline__ void hullCtx(HullCtx* h, uint4 rec, int lane) {
    if (lane < 2) {
        const v3 corner = mk3(__uint_as_float(rec.x), __uint_as_float(rec.y), __uint_as_float(rec.z));
        if (lane == 0) h->low = corner;
        else h->high = corner;
    } else if (lane < 5) {
        h->range[2 * (lane - 2)] = make_uint2(rec.x, rec.y);
        h->range[2 * (lane - 2) + 1] = make_uint2(rec.z, rec.w);
    }
    wave_sync();
}
__device__ __forceinline__ bool outsideRoot(const HullCtx& h, v3 q) {
    return q.x < h.low.x || q.y < h.low.y || q.z < h.low.z || q.x > h.high.x || q.y > h.high.y || q.z > h.high.z;
}
__device__ __forceinline__ float rootExcess(const HullCtx& h, v3 q) {
    return fmaxf(fmaxf(fmaxf(h.low.x - q.x, q.x - h.high.x), fmaxf(h.low.y - q.y, q.y - h.high.y)),
                 fmaxf(h.low.z - q.z, q.z - h.high.z));
}
// meshOutsideFast with the task's hull context; `code` names the plane of a rejection
// (the largest of the first batch that has one).
__device__ __forceinline__ float nodeOutside(const P& p, const HullCtx& h, v3 q, uint32_t& code) {
    code = NO_PLANE;
    if (outsideRoot(h, q)) return rootExcess(h, q);
    const uint32_t g = dominantGroup((q - h.mid()) / h.ext());
    const uint2 r = h.range[g];
    for (uint32_t k = 0; k < r.y; k += PLANE_BATCH) {
        float worst = 0.0f;
        uint32_t best = NO_PLANE;
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++) {
            const uint32_t kk = min(k + j, r.y - 1);
            const float4 pl = p.meshPlanes[r.x + kk];
            const float e = dot(xyz(pl), q) + pl.w;
            if (e > worst) { worst = e; best = planeCode(g, kk); }
            worst = fmaxf(worst, e);
        }
        if (worst > 0) { code = best; return worst; }
    }
    return 0.0f;
}
// Hull tests of the lanes in `tests` (at most 16) together: the tested points are
// spread over groups of 64 / 2^ceil(log2 n) lanes, each group evaluates its point's
// planes (or root-box excess) and reduces the largest, and every tested lane gets
// its own point's value back.  The verdict (no plane value above zero, inside the
// root box) is the one-lane form's bit for bit; a failing point reports its largest
// excess rather than the first batch's, an equally valid distance bound.
__device__ __forceinline__ float nodesOutsideWave(const P& p, const HullCtx& h, uint64_t tests, v3 q, int lane,
                                                  uint32_t& code) {
    const uint32_t n = uint32_t(__popcll(tests));
    const uint32_t shift = n <= 1 ? 6u : n <= 2 ? 5u : n <= 4 ? 4u : n <= 8 ? 3u : 2u;
    const uint32_t span = 1u << shift, group = uint32_t(lane) >> shift, sub = uint32_t(lane) & (span - 1u);
    int src = -1;
    uint64_t bits = tests;
    for (uint32_t k = 0; bits != 0; k++) {
        const int s = __builtin_ctzll(bits);
        bits &= bits - 1;
        if (group == k) src = s;
    }
    const int from = src < 0 ? lane : src;
    const v3 x = mk3(__shfl(q.x, from), __shfl(q.y, from), __shfl(q.z, from));
    float worst = 0.0f;
    uint32_t best = NO_PLANE;
    if (src >= 0) {
        if (outsideRoot(h, x)) {
            worst = rootExcess(h, x);
        } else {
            const uint32_t g = dominantGroup((x - h.mid()) / h.ext());
            const uint2 r = h.range[g];
            for (uint32_t k = sub; k < r.y; k += span) {
                const float4 pl = p.meshPlanes[r.x + k];
                const float e = dot(xyz(pl), x) + pl.w;
                if (e > worst) { worst = e; best = planeCode(g, k); }
                worst = fmaxf(worst, e);
            }
        }
    }
    for (uint32_t d = 1; d < span; d <<= 1) {
        const float ow = __shfl_xor(worst, int(d));
        const uint32_t ob = __shfl_xor(best, int(d));
        if (ow > worst) best = ob;
        worst = fmaxf(worst, ow);
    }
    const uint32_t rank = uint32_t(__popcll(tests & lanes_below(lane)));
    code = __shfl(best, int(rank << shift));
    return __shfl(worst, int(rank << shift));
}
// The transformed, padded box of a source cluster in the target frame
// (meshClusterOutsideFast's corner loop, same expressions).
__device__ __forceinline__ void clusterBox(v3 sl, v3 sh, v4 qs, v4 qd, float ss, float sd, v3 branch, v3& low,
                                           v3& high) {
    low = {1e30f, 1e30f, 1e30f};
    high = {-1e30f, -1e30f, -1e30f};
    for (uint32_t corner = 0; corner < 8; corner++) {
        v3 q = {(corner & 1) ? sh.x : sl.x, (corner & 2) ? sh.y : sl.y, (corner & 4) ? sh.z : sl.z};
        q = qinv(qd, qrot(qs, q * ss) - branch) / sd;
        low = fmin3(low, q);
        high = fmax3(high, q);
    }
}
__device__ __forceinline__ bool boxOutsideRoot(const HullCtx& h, v3 low, v3 high) {
    return high.x < h.low.x || high.y < h.low.y || high.z < h.low.z || low.x > h.high.x || low.y > h.high.y ||
           low.z > h.high.z;
}
__device__ __forceinline__ float boxRootExcess(const HullCtx& h, v3 low, v3 high) {
    return fmaxf(fmaxf(fmaxf(h.low.x - high.x, low.x - h.high.x), fmaxf(h.low.y - high.y, low.y - h.high.y)),
                 fmaxf(h.low.z - high.z, low.z - h.high.z));
}
__device__ __forceinline__ float clusterPad(v3 low, v3 high, v3 branch, float sd) {
    return 64.0f * DEM_EPS * (1.0f + maxabs(low) + maxabs(high) + length(branch) / sd);
}
// meshClusterOutsideFast with the task's hull context.
__device__ __forceinline__ float clusterOutside(const P& p, const HullCtx& h, v4 qs, v4 qd, float ss, float sd, v3 branch,
                                uint32_t cluster, uint32_t& code) {
    code = NO_PLANE;
    v3 low, high;
    clusterBox(xyz(p.meshClusterMin[cluster]), xyz(p.meshClusterMax[cluster]), qs, qd, ss, sd, branch, low, high);
    const float pad = clusterPad(low, high, branch, sd);
    low = low - mk3(pad, pad, pad);
    high = high + mk3(pad, pad, pad);
    if (boxOutsideRoot(h, low, high)) return boxRootExcess(h, low, high);
    const v3 center = 0.5f * (low + high), half = 0.5f * (high - low);
    const uint32_t g = dominantGroup((center - h.mid()) / h.ext());
    const uint2 r = h.range[g];
    for (uint32_t k = 0; k < r.y; k += PLANE_BATCH) {
        float4 pl[PLANE_BATCH];
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++) pl[j] = p.meshPlanes[r.x + min(k + j, r.y - 1)];
        float worst = 0.0f;
        uint32_t best = NO_PLANE;
#pragma unroll
        for (uint32_t j = 0; j < PLANE_BATCH; j++) {
            const float e = dot(xyz(pl[j]), center) + pl[j].w - dot(fabs3(xyz(pl[j])), half);
            if (e > worst) { worst = e; best = planeCode(g, min(k + j, r.y - 1)); }
            worst = fmaxf(worst, e);
        }
        if (worst > 0) { code = best; return worst; }
    }
    return 0.0f;
}
// Leaf tests of the lanes in `tests` (at most 8) together, as nodesOutsideWave: each
// leaf gets 64 / 2^ceil(log2 n) lanes, eight of which transform its box corners
// (reduced as meshClusterOutsideWave does) before the group tests its planes.
// Lane c of `tests` is cluster clusterBase + c.
__device__ __forceinline__ float leavesOutsideWave(const P& p, const HullCtx& h, uint64_t tests, uint32_t clusterBase, v4 qs,
                                   v4 qd, float ss, float sd, v3 branch, int lane, uint32_t& code) {
    const uint32_t n = uint32_t(__popcll(tests));
    const uint32_t shift = n <= 1 ? 6u : n <= 2 ? 5u : n <= 4 ? 4u : 3u;
    const uint32_t span = 1u << shift, group = uint32_t(lane) >> shift, sub = uint32_t(lane) & (span - 1u);
    int src = -1;
    uint64_t bits = tests;
    for (uint32_t k = 0; bits != 0; k++) {
        const int s = __builtin_ctzll(bits);
        bits &= bits - 1;
        if (group == k) src = s;
    }
    const uint32_t cluster = clusterBase + uint32_t(src < 0 ? 0 : src);
    const v3 sl = xyz(p.meshClusterMin[cluster]), sh = xyz(p.meshClusterMax[cluster]);
    const uint32_t corner = sub & 7u;
    v3 q = {(corner & 1) ? sh.x : sl.x, (corner & 2) ? sh.y : sl.y, (corner & 4) ? sh.z : sl.z};
    q = qinv(qd, qrot(qs, q * ss) - branch) / sd;
    v3 low = fmin3(mk3(1e30f, 1e30f, 1e30f), q), high = fmax3(mk3(-1e30f, -1e30f, -1e30f), q);
    for (int d = 1; d < 8; d <<= 1) {
        low = fmin3(low, mk3(__shfl_xor(low.x, d), __shfl_xor(low.y, d), __shfl_xor(low.z, d)));
        high = fmax3(high, mk3(__shfl_xor(high.x, d), __shfl_xor(high.y, d), __shfl_xor(high.z, d)));
    }
    const float pad = clusterPad(low, high, branch, sd);
    low = low - mk3(pad, pad, pad);
    high = high + mk3(pad, pad, pad);
    float worst = 0.0f;
    uint32_t best = NO_PLANE;
    if (src >= 0) {
        if (boxOutsideRoot(h, low, high)) {
            worst = boxRootExcess(h, low, high);
        } else {
            const v3 center = 0.5f * (low + high), half = 0.5f * (high - low);
            const uint32_t g = dominantGroup((center - h.mid()) / h.ext());
            const uint2 r = h.range[g];
            for (uint32_t k = sub; k < r.y; k += span) {
                const float4 pl = p.meshPlanes[r.x + k];
                const float e = dot(xyz(pl), center) + pl.w - dot(fabs3(xyz(pl)), half);
                if (e > worst) { worst = e; best = planeCode(g, k); }
                worst = fmaxf(worst, e);
            }
        }
    }
    for (uint32_t d = 1; d < span; d <<= 1) {
        const float ow = __shfl_xor(worst, int(d));
        const uint32_t ob = __shfl_xor(best, int(d));
        if (ow > worst) best = ob;
        worst = fmaxf(worst, ow);
    }
    const uint32_t rank = uint32_t(__popcll(tests & lanes_below(lane)));
    code = __shfl(best, int(rank << shift));
    return __shfl(worst, int(rank << shift));
}
// A wave's query batch.  Under the temporal skip a task queues about 10-16 nodes,
// so querying one task at a time would leave most of the wave's 64 lanes idle.
// The wave instead filters several tasks in turn, appending each one's queued
// nodes to its queue, and queries them together: each queue entry carries the
// batch slot of its task (held in nearestTri[] until the query overwrites it with
// the winner), and each lane reads its task's frame from the slot's record.  Hits
// are then staged task by task, each task's in its queue order, with the task's
// own running count, so every contact, value and rank is the same as querying
// each task alone.
constexpr uint32_t BATCH_MAX = 16;  // with the 256-node queue; 8 with the 128-node queue (LDS)
struct BatchRec {  // 112 bytes in LDS per task of the batch
    uint32_t task, root, rootRef, sampleBase;
    uint32_t triFirst, triCount, gens, end;  // gens = genPrev | genNew << 16; end = queue end | refresh << 16
    float4 qi, qj;
    float4 branchSi;  // branch, s_i
    float4 tail;      // s_j, min(wi, wj), unused, unused
    uint32_t i, j, row, sj;  // row: the task's row in the persistent per-task rows, or NO_ROW; sj: target template
};
constexpr uint32_t NO_ROW = 0xffffffffu;
static_assert(sizeof(BatchRec) == 112, "batch record layout");

// The target's root node, its reference, and its first triangle and triangle count
// (0, 0 without warm starts), from its template record.
struct TargetRoot {
    uint32_t root, rootRef, triFirst, triCount;
};
__device__ __forceinline__ void batchPut(BatchRec* r, const TaskFrame& f, const TargetRoot& tg, uint32_t task,
                                         uint32_t row, bool refresh, const GapCtx& gc, int lane) {
    const uint32_t root = tg.root, rootRef = tg.rootRef, triFirst = tg.triFirst, triCount = tg.triCount;
    if (lane == 0) {
        r->task = task; r->root = root; r->rootRef = rootRef; r->sampleBase = f.sampleBase;
        r->triFirst = triFirst; r->triCount = triCount;
        r->gens = gc.genPrev | (gc.genNew << 16);
        r->end = refresh ? 1u << 16 : 0u;
        r->qi = f4(f.qi); r->qj = f4(f.qj);
        r->branchSi = f4(f.branch, f.s_i);
        r->tail = make_float4(f.s_j, fminf(f.wi, f.wj), 0.0f, 0.0f);
        r->i = f.i; r->j = f.j; r->row = row; r->sj = f.sj;
    }
}

// Exact queries of queue[0..qn), the entries of batch slots 0..nb-1 in slot order,
// in two phases.
//
// Phase 1 finds every queued node's nearest triangle with lane refill: each lane
// runs one BVH traversal (meshDistanceWide's exact visit order) for the node of its
// entry, in its task's frame, and when it finishes takes the next unqueried entry,
// so lanes do not idle while the wave's slowest traversal runs (per-query visit
// counts differ by ~3x).  The winner's triangle index goes to nearestTri[k].
//
// Phase 2 takes each slot's entries 64 at a time in order, rebuilds each query's
// result from its nearest triangle (closestTriangle is deterministic, so the point,
// squared distance and feature equal the traversal's) and stages hits by wave rank:
// the same contacts, values and ranks as querying each task to completion.  Every
// slot's hit count goes to taskCount, except the last slot's when its task is still
// being filtered (`lastOpen`): its count is returned and carried into the next batch.
constexpr uint32_t NO_TRIANGLE = 0xffffffffu;

// Queue entries carry their state in nearestTri[]: PENDING | batch slot until the
// entry's traversal ends, then the winner (a triangle index < 2^28, or NO_TRIANGLE).
// Entries whose winner the unchanged-winner shortcut proved hold it from the start.
constexpr uint32_t PENDING = 0x40000000u;
__device__ __forceinline__ bool isPending(uint32_t v) { return (v & 0xc0000000u) == PENDING; }
// A pending entry the learned cull may drop (its network is evaluated before the
// traversals): set at admission, ignored by the traversal (it reads the batch slot
// from the low bits).
constexpr uint32_t NN_CAND = 0x20000000u;
// Position of the r-th set bit (from 0) of m; m holds more than r set bits.
__device__ __forceinline__ uint32_t selectBit(uint64_t m, uint32_t r) {
    uint32_t pos = 0, c = __popc(uint32_t(m));
    if (r >= c) { r -= c; m >>= 32; pos += 32; }
    c = __popc(uint32_t(m) & 0xffffu);
    if (r >= c) { r -= c; m >>= 16; pos += 16; }
    c = __popc(uint32_t(m) & 0xffu);
    if (r >= c) { r -= c; m >>= 8; pos += 8; }
    c = __popc(uint32_t(m) & 0xfu);
    if (r >= c) { r -= c; m >>= 4; pos += 4; }
    c = __popc(uint32_t(m) & 0x3u);
    if (r >= c) { r -= c; m >>= 2; pos += 2; }
    if (r >= (uint32_t(m) & 1u)) pos += 1;
    return pos;
}
template <bool NN, typename D>
__device__ uint32_t queryBatch(const P& p, const W& w, const BatchRec* batch, uint32_t nb, uint32_t stageQueue,
                               const uint32_t* queue, uint32_t* nearestTri, uint32_t qn, uint32_t* sref, D* sdist,
                               uint32_t carry, bool lastOpen, int lane, const uint32_t* auditBits,
                               uint32_t& resNext, uint32_t& resEnd) {
    const float4* __restrict__ wide = w.wide;
    const bool gapOn = w.warm && w.nodeGap;
    uint32_t k = 0, top = 0, bestTri = NO_TRIANGLE, slotNow = 0, rowNow = 0, genNow = 0;
    float best = 1e30f, prune = 1e30f, second = 1e30f;
    // Gap reach G: the traversal also visits what lies within G of its bound, so a
    // recorded gap is at least min(G, the true gap) instead of the distance to the
    // nearest pruned box; `ext` = (sqrt(best) + G)^2 replaces best in the bound.
    // Visiting more never changes the winner (the extra triangles lie farther than
    // the minimum, and the record test stays strict).
    const float reachG = gapOn ? w.gapReach : 0.0f;
    float ext = 1e30f;
    v3 q = {0, 0, 0};
    bool active = false;
    // Warm start: the squared distance to the node's nearest triangle of an earlier
    // step is an upper bound on this query's minimum, so subtrees whose box lies
    // strictly farther can be pruned from the start.  Triangles at the minimum are
    // never pruned (their boxes are no farther than the minimum), and the record
    // test still starts from 1e30, so the winner -- the first triangle in visit
    // order at the minimum distance -- is the one the plain traversal finds.
    // In FP32 a box's distance can exceed the computed distance of a triangle it
    // holds (the closest point carries a few ulps of rounding), so the bound is
    // widened by WARM_SLACK in distance: (sqrt(d) + WARM_SLACK)^2.  A looser bound
    // only visits more boxes; it never changes the winner.
    auto start = [&]() {
        const uint32_t entry = queue[k];
        const BatchRec& r = batch[nearestTri[k] & 0xffffu];
        const uint32_t node = entry & 0xffffu, slot = entry >> 16;
        const uint4 h0 = make_uint4(r.row, r.root, r.rootRef, r.sampleBase);
        const uint4 h1 = make_uint4(r.triFirst, r.triCount, r.gens, r.end);
        const float4 qi = r.qi, qj = r.qj, bs = r.branchSi, tail = r.tail;
        const float4 sp = p.sample[h0.w + node];
        // queryPoint in the task's frame
        const v3 arm = qrot(v4of(qi), xyz(sp) * bs.w);
        q = qinv(v4of(qj), arm - xyz(bs)) / tail.x;
        sref[0] = h0.z;
        dstore(sdist, boxDistance(xyz(p.meshMin[h0.y]), xyz(p.meshMax[h0.y]), q));
        top = 1;
        best = 1e30f;
        prune = 1e30f;
        second = 1e30f;
        ext = 1e30f;
        slotNow = slot;
        rowNow = h0.x;
        genNow = h1.z >> 16;
        bestTri = NO_TRIANGLE;
        if (w.warm && h0.x != NO_ROW) {
            const uint32_t last = w.nodeTri[size_t(h0.x) * w.nodeStride + slot];
            if (last != 0u && last <= h1.y) {
                uint32_t feature;
                const float4* t = w.triPacked + 3 * size_t(h1.x + last - 1);
                const v3 x = closestTriangle(q, xyz(t[0]), xyz(t[1]), xyz(t[2]), feature);
                const float reach = sqrtf(dot(q - x, q - x)) + WARM_SLACK + reachG;
                prune = reach * reach;
            }
        }
    };
    // Idle lanes take the next pending entries in queue order.
    uint32_t next = 0;  // first queue entry not yet scanned (wave-uniform)
    auto refill = [&]() {
        for (;;) {
            const uint64_t idle = __ballot(!active);
            if (idle == 0 || next >= qn) break;
            const uint32_t kk = next + lane;
            const uint64_t pend = __ballot(kk < qn && isPending(nearestTri[kk]));
            const uint32_t nIdle = __popcll(idle), nPend = __popcll(pend);
            if (!active) {
                const uint32_t rank = __popcll(idle & lanes_below(lane));
                if (rank < nPend) {
                    k = next + selectBit(pend, rank);
                    active = true;
                    start();
                }
            }
            next += nPend <= nIdle ? uint32_t(WAVE) : selectBit(pend, nIdle - 1) + 1;
        }
    };
    refill();
    // Each lane walks its own stack in its own order; the wave alternates between
    // descending (lanes pop and test internal nodes until they hold a leaf or their
    // stack is empty) and testing the held leaves' triangles together, so the cheap
    // box tests and the costly triangle tests no longer diverge within one pass.  A
    // lane stops descending while it holds a leaf and resumes only after that leaf's
    // triangles updated its bound: its pops, pruning decisions and record updates
    // happen in exactly the order of the one-entry-per-pass walk.
    const uint32_t leafBatch = w.leafBatch ? w.leafBatch : uint32_t(WAVE);
    bool hasLeaf = false;
    uint32_t leafRef = 0;
    while (__ballot(active) != 0) {
        for (;;) {
            const bool descend = active && !hasLeaf && top > 0;
            if (__ballot(descend) == 0 || uint32_t(__popcll(__ballot(hasLeaf))) >= leafBatch) break;
            if (descend) {
                --top;
                const uint32_t ref = sref[top * WAVE];
                const float bound = fminf(ext, prune);
                const float popped = dload(sdist + top * WAVE);
                if (!(popped > bound)) {
                    if ((ref & LEAF) == 0) {
                        // The node's rows, and the next node's: in the depth-first layout
                        // the left child is node ref + 1, so its rows (unused when it is a
                        // leaf) are read in the same round trip.
                        const float4* node = wide + 4 * size_t(ref);
                        float4 w0 = node[0], w1 = node[1], w2 = node[2], w3 = node[3];
                        const float4 v0 = node[4], v1 = node[5], v2 = node[6], v3 = node[7];
                        float da = boxDistance(xyz(w0), xyz(w1), q), db = boxDistance(xyz(w2), xyz(w3), q);
                        uint32_t a = __float_as_uint(w0.w), b = __float_as_uint(w1.w);
                        const bool leftNear = da < db;
                        if (da < db) { uint32_t t = a; a = b; b = t; float u = da; da = db; db = u; }
                        if (top + 2 > WSTACK) {
                            fail(p, 1024);
                            top = 0;
                            bestTri = NO_TRIANGLE;
                        } else {
                            if (da <= bound) { sref[top * WAVE] = a; dstore(sdist + top * WAVE, da); top++; }
                            else second = fminf(second, da);
                            if (!(db <= bound)) {
                                second = fminf(second, db);
                            } else if ((b & LEAF) != 0) {
                                // The near child is popped next and passes the same bound
                                // (no triangle is tested in between): hold its leaf now.
                                hasLeaf = true;
                                leafRef = b;
                            } else if (leftNear) {
                                // The near child is the left one, whose rows are in hand:
                                // pop it in this pass, exactly as the next pass would.
                                float ea = boxDistance(xyz(v0), xyz(v1), q), eb = boxDistance(xyz(v2), xyz(v3), q);
                                uint32_t c = __float_as_uint(v0.w), d = __float_as_uint(v1.w);
                                if (ea < eb) { uint32_t t = c; c = d; d = t; float u = ea; ea = eb; eb = u; }
                                if (top + 2 > WSTACK) {
                                    fail(p, 1024);
                                    top = 0;
                                    bestTri = NO_TRIANGLE;
                                } else {
                                    if (ea <= bound) { sref[top * WAVE] = c; dstore(sdist + top * WAVE, ea); top++; }
                                    else second = fminf(second, ea);
                                    if (eb <= bound) { sref[top * WAVE] = d; dstore(sdist + top * WAVE, eb); top++; }
                                    else second = fminf(second, eb);
                                }
                            } else {
                                sref[top * WAVE] = b;
                                dstore(sdist + top * WAVE, db);
                                top++;
                            }
                        }
                    } else {
                        hasLeaf = true;
                        leafRef = ref;
                    }
                } else {
                    second = fminf(second, popped);  // a pruned subtree: a lower bound for all its triangles
                }
            }
        }
        if (hasLeaf) {
            const uint32_t first = leafRef & 0x0fffffffu, count = ((leafRef >> 28) & 3u) + 1;
            for (uint32_t c = 0; c < count; c++) {
                uint32_t triangle = first + c, feature;
                const float4* t = w.triPacked + 3 * size_t(triangle);
                const float4 t0 = t[0], t1 = t[1], t2 = t[2];
                const v3 a = xyz(t0);
                // A triangle whose plane is farther than the current bound cannot
                // reach the minimum (the distance to a triangle is at least the
                // distance to its plane; 1e-6 absolute and 1e-4 relative cover the
                // FP32 normal and both evaluations), so skipping it cannot change
                // which triangle wins.  The face normal rides in the rows' w.
                const float h = fabsf(dot(q - a, mk3(t0.w, t1.w, t2.w))) - 1e-6f;
                if (h > 0.0f && h * h > 1.0001f * fminf(ext, prune)) {
                    second = fminf(second, h * h);
                    continue;
                }
                v3 x = closestTriangle(q, a, xyz(t1), xyz(t2), feature);
                float distance = dot(q - x, q - x);
                if (distance < best) {
                    second = fminf(second, best);
                    best = distance;
                    bestTri = triangle;
                    if (reachG > 0.0f) {
                        const float e = sqrtf(best) + reachG;
                        ext = e * e;
                    } else {
                        ext = best;
                    }
                } else {
                    second = fminf(second, distance);
                }
            }
            hasLeaf = false;
        }
        if (active && top == 0) {
            nearestTri[k] = bestTri;
            if (gapOn && rowNow != NO_ROW) {
                // The gap this query leaves for the next collection (0 when there is none).
                const float gap = bestTri != NO_TRIANGLE && second < 1e30f
                                      ? sqrtf(second) - sqrtf(best) - 2.0f * GAP_EVAL - GAP_MOTION_ABS
                                      : 0.0f;
                w.nodeGap[size_t(rowNow) * w.nodeStride + slotNow] = gapWord(genNow, gap);
            }
            active = false;
        }
        refill();
    }
    wave_sync();
    uint32_t from = 0, open = 0;
    for (uint32_t s = 0; s < nb; s++) {
        // The slot's frame, read uniformly (scalar registers).
        const BatchRec& r = batch[s];
        TaskFrame f;
        const uint32_t task = uniformU(r.task), end = uniformU(r.end);
        f.i = r.i; f.j = r.j; f.sampleBase = r.sampleBase;
        f.qi = v4of(r.qi); f.qj = v4of(r.qj);
        const float4 bs = r.branchSi, tail = r.tail;
        f.branch = xyz(bs); f.s_i = bs.w; f.s_j = tail.x; f.wi = tail.y; f.wj = tail.y;
        f.si = 0; f.sj = 0; f.plane = make_float4(0, 0, 0, 0);
        uniformFrame(f);
        const uint32_t triFirst = uniformU(r.triFirst);
        const uint32_t to = end & 0xffffu;
        const uint32_t rowIndex = uniformU(r.row);
        const size_t row = size_t(rowIndex) * w.nodeStride;
        int8_t* refreshDist = (end >> 16) && rowIndex != NO_ROW ? w.nodeDist + row : nullptr;
        uint16_t* warmTri = w.warm && rowIndex != NO_ROW ? w.nodeTri + row : nullptr;
        uint32_t running = s == 0 ? carry : 0u;
        for (uint32_t b = from; b < to; b += WAVE) {
            const uint32_t kk = b + lane;
            bool hit = false, boundViolation = false;
            uint32_t node = 0;
            float4 pd = make_float4(0, 0, 0, 0), na = make_float4(0, 0, 0, 0);
            if (kk < to) {
                node = queue[kk] & 0xffffu;
                const uint32_t slot = queue[kk] >> 16;
                const float4 sp = p.sample[f.sampleBase + node];
                v3 arm;
                const v3 x = queryPoint(f, sp, arm);
                // meshDistanceWide's result from the nearest triangle.
                const uint32_t triangle = nearestTri[kk];
                float bestD = 1e30f;
                v3 nearest = {0, 0, 0}, pseudo = {0, 0, 0};
                if (triangle != NO_TRIANGLE) {
                    uint32_t feature;
                    const float4* t = w.triPacked + 3 * size_t(triangle);
                    nearest = closestTriangle(x, xyz(t[0]), xyz(t[1]), xyz(t[2]), feature);
                    bestD = dot(x - nearest, x - nearest);
                    pseudo = xyz(p.meshNormals[7 * triangle + feature]);
                }
                const float distance = sqrtf(bestD);
                const float side = dot(x - nearest, pseudo) < 0 ? -1.0f : 1.0f;
                const v3 normal = distance > 1e-12f ? side * (x - nearest) / distance : pseudo;
                bool keepBound = false;  // an audit-marked entry keeps the network's bound (as when culled)
                if constexpr (NN) {
                    if (auditBits && refreshDist && ((auditBits[kk / 32] >> (kk % 32)) & 1u) != 0u) {
                        keepBound = true;
                        if (side * distance < float(refreshDist[slot]) * SKIP_QUANTUM) boundViolation = true;
                    }
                }
                if (refreshDist && !keepBound) refreshDist[slot] = quantizeDistance(side * distance);
                if (warmTri && triangle != NO_TRIANGLE) warmTri[slot] = uint16_t(triangle - triFirst + 1);
                // The rest of queryNode.
                const float depth = -(side * distance) * f.s_j;
                pd = f4(arm, depth);
                if (depth > 0) {
                    const v3 grad = qrot(f.qj, normal);
                    const float len = length(grad);
                    if (len < 1e-7f || !isfinite(depth)) {
                        fail(p, 2);
                    } else {
                        na = f4(grad / len, sp.w * f.s_i * f.s_i);
                        if (depth > p.c.maxOverlap * fminf(f.wi, f.wj)) fail(p, 4);
                        hit = true;
                    }
                }
            }
            if constexpr (NN) {  // audit: contacts, or distances below the network's bound, among dropped entries
                if (auditBits) {
                    const bool marked = kk < to && ((auditBits[kk / 32] >> (kk % 32)) & 1u) != 0u;
                    const uint32_t missed = __popcll(__ballot(marked && (hit || boundViolation)));
                    if (missed && lane == 0) atomicAdd(w.nnStats + LINE, (unsigned long long)missed);
                }
            }
            stageHits(p, w, f, task, stageQueue, hit, node, pd, na, running, lane, resNext, resEnd);
        }
        from = to;
        if (s + 1 < nb || !lastOpen) {
            if (lane == 0) w.taskCount[task] = running;
        } else {
            open = running;
        }
    }
    return open;
}

// Learned cull of a query batch, before its traversals.  The entries admitted as
// candidates (NN_CAND) are evaluated 16 at a time per task, in that task's frame with
// its target template's network; an entry whose predicted signed distance exceeds
// the template's margin is dropped, and the survivors are compacted in place in
// order (the hit ranks are those of the unculled walk, since a dropped entry is not
// a contact when the network is right).  In audit mode nothing is dropped: the
// entries are marked in `auditBits`, every entry is queried and marked hits are
// counted as misses.  Returns the surviving queue length.
__device__ uint32_t cullEntries(const P& p, const W& w, BatchRec* batch, uint32_t nb, uint32_t* queue,
                                uint32_t* nearestTri, uint32_t qn, uint32_t* auditBits, uint16_t* nnA,
                                uint16_t* nnB, int lane) {
    const bool audit = w.nnAudit != 0u;
    if (audit)
        for (uint32_t k = lane; k < (qn + 31) / 32; k += WAVE) auditBits[k] = 0u;
    wave_sync();
    uint32_t from = 0, out = 0, culledTotal = 0;
    for (uint32_t s = 0; s < nb; s++) {
        BatchRec& r = batch[s];
        const uint32_t end = uniformU(r.end), to = end & 0xffffu, sj = uniformU(r.sj);
        const bool on = sj < w.nnTemplates && w.nnF[size_t(sj) * nn::FSTRIDE + nn::FMARGIN] < 1e30f;
        const uint16_t* wb = w.nnW + size_t(sj) * nn::BSTRIDE;
        const float* fb = w.nnF + size_t(sj) * nn::FSTRIDE;
        const float margin = on ? fb[nn::FMARGIN] : 0.0f;
        // A re-anchoring task records a distance bound for each node it drops, so the
        // temporal skip passes the node over in later steps as it would after an
        // exact query: (min(y, 1) - margin) x label scale for a prediction y.  It
        // fails only where the network errs by more than the margin, the same
        // error that would drop a contact (the audit checks both).
        int8_t* refreshDist = (end >> 16) != 0u && uniformU(r.row) != NO_ROW
                                  ? w.nodeDist + size_t(uniformU(r.row)) * w.nodeStride
                                  : nullptr;
        const float labelScale = on ? fb[nn::FLABEL] : 0.0f;
        for (uint32_t b = from; b < to; b += WAVE) {
            const uint32_t k = b + lane;
            const uint32_t word = k < to ? queue[k] : 0u, state = k < to ? nearestTri[k] : 0u;
            uint64_t cand = __ballot(on && k < to && isPending(state) && (state & NN_CAND) != 0u);
            uint64_t culled = 0;
            while (cand != 0) {
                const uint32_t rows = min(uint32_t(__popcll(cand)), 16u);
                const uint32_t pos = lane < 16 && uint32_t(lane) < rows ? selectBit(cand, lane) : 0u;
                if (lane < 16) {
                    uint16_t* row = nnA + lane * nn::K1;
                    if (uint32_t(lane) < rows) {  // the node's point in the task's frame (read here: few live registers)
                        const float4 sp = p.sample[r.sampleBase + (queue[b + pos] & 0xffffu)];
                        const float4 bs = r.branchSi;
                        const v3 arm = qrot(v4of(r.qi), xyz(sp) * bs.w);
                        nn_features(qinv(v4of(r.qj), arm - xyz(bs)) / r.tail.x, fb, row);
                    } else {
                        for (int c = 0; c < nn::K1; c++) row[c] = 0;
                    }
                }
                wave_sync();
                const float y = nn_forward<1, false>(nnA, nnB, wb, fb, lane);  // 16x16x16 MFMA here (nn_layer)
                const bool dropMine = lane < 16 && uint32_t(lane) < rows && y > margin;
                if (dropMine && refreshDist)
                    refreshDist[queue[b + pos] >> 16] = quantizeDistance((fminf(y, 1.0f) - margin) * labelScale);
                const uint64_t drop = __ballot(dropMine);
                for (uint64_t m = drop; m != 0; m &= m - 1) culled |= 1ull << __shfl(pos, __builtin_ctzll(m));
                for (uint32_t c = 0; c < rows; c++) cand &= cand - 1;  // the rows just evaluated
                wave_sync();
            }
            culledTotal += __popcll(culled);
            const bool mine = ((culled >> lane) & 1ull) != 0ull;
            if (audit) {
                if (mine) atomicOr(auditBits + k / 32, 1u << (k % 32));
                out = to;
            } else {
                const bool keep = k < to && !mine;
                const uint64_t m = __ballot(keep);
                if (keep) {
                    const uint32_t at = out + __popcll(m & lanes_below(lane));
                    queue[at] = word;
                    nearestTri[at] = state;
                }
                out += __popcll(m);
            }
            wave_sync();
        }
        if (lane == 0) r.end = (end & 0xffff0000u) | out;
        from = to;
    }
    wave_sync();
    if (lane == 0 && culledTotal) {
        atomicAdd(w.nnStats, (unsigned long long)culledTotal);
        if (audit) atomicAdd(w.nnStats + LINE / 2, (unsigned long long)culledTotal);
    }
    return out;
}

template <bool NN, typename D>
__device__ uint32_t flushBatch(const P& p, const W& w, uint32_t stageQueue, uint32_t* queue, uint32_t* nearestTri,
                               BatchRec* batch, uint32_t nb, uint32_t qn, uint32_t* sref, D* sdist, uint32_t carry,
                               bool lastOpen, int lane, uint32_t& nnPending, uint32_t* auditBits, uint16_t* nnA,
                               uint16_t* nnB, uint32_t& resNext, uint32_t& resEnd) {
    const uint32_t* audited = nullptr;
    if constexpr (NN) {
        if (nnPending != 0u) {
            qn = cullEntries(p, w, batch, nb, queue, nearestTri, qn, auditBits, nnA, nnB, lane);
            if (w.nnAudit) audited = auditBits;
        }
        nnPending = 0u;
    }
    return queryBatch<NN, D>(p, w, batch, nb, stageQueue, queue, nearestTri, qn, sref, sdist, carry, lastOpen, lane,
                             audited, resNext, resEnd);
}

// One directed pair per wave (cpuCollectMesh, cpu_scheduling.slang:53-91).
// NN = learned-cull variant (a separate instantiation keeps the plain kernel's registers);
// QN = query-queue capacity per wave (tunable: results do not depend on it, since hits
// are ranked by their running count within the task, not by flush boundaries).
template <bool NN, int QN>
__device__ __forceinline__ void collectBody(P p, W w) {
    using D = uint16_t;
    // The learned-cull variant evaluates its network before a batch's traversals:
    // the traversal stack's 4 KB then holds two 16 x 64 BF16 activation buffers.
    __shared__ __align__(16) uint32_t stackRef[COLLECT_WAVES][WSTACK * WAVE];
    __shared__ D stackDist[COLLECT_WAVES][WSTACK * WAVE];
    __shared__ uint32_t queues[COLLECT_WAVES][QN];
    __shared__ uint32_t nearestTris[COLLECT_WAVES][QN];
    __shared__ uint64_t candidates[COLLECT_WAVES][MAX_CLUSTERS / 64];
    // The 128-node queue leaves LDS for 8 batch records.
    constexpr uint32_t batchCap = QN >= 256 ? BATCH_MAX : 8u;
    __shared__ BatchRec batches[COLLECT_WAVES][batchCap];
    __shared__ uint32_t auditWords[COLLECT_WAVES][NN ? QN / 32 : 1];  // learned-cull audit marks per entry
    __shared__ HullCtx hulls[COLLECT_WAVES];                          // the task's target root box and hull groups
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t* sref = &stackRef[wv][lane];
    D* sdist = &stackDist[wv][lane];
    uint16_t* nnA = reinterpret_cast<uint16_t*>(stackRef[wv]);
    uint16_t* nnB = nnA + 16 * nn::H;
    uint32_t* auditBits = auditWords[wv];
    uint32_t nnPending = 0;  // learned-cull candidates in the open batch
    uint32_t resNext = 0, resEnd = 0;  // this wave's unused staging reservation
    uint32_t* queue = queues[wv];
    uint32_t* nearestTri = nearestTris[wv];
    uint64_t* cand = candidates[wv];
    BatchRec* batch = batches[wv];
    uint32_t queried = 0, skippedNodes = 0, shortcutNodes = 0;  // statistics, added once per wave at its exit
    // The query batch: nb tasks with qn queued nodes; carry = hits already staged
    // for the batch's first task when an earlier batch ended inside it.
    uint32_t nb = 0, qn = 0, carry = 0, pending = 0;  // pending: queued nodes that need a traversal
    const uint32_t batchLimit = max(1u, min(w.batchTasks, batchCap));
    TaskFrame f{};
    // XCD-affine queues: tasks are grouped by target template; a wave drains the
    // queue of its own XCD (workgroups are dispatched round-robin over 8 XCDs)
    // and then steals from the others.
    uint32_t qcount[8], qoff[8];
    {
        uint32_t acc = 0;
        for (int k = 0; k < 8; k++) { qcount[k] = w.xcd[k * LINE]; qoff[k] = acc; acc += qcount[k]; }
    }
    const uint32_t home = blockIdx.x & 7;
    uint32_t tried = 0, next = 0, left = 0, from = 0;
    for (;;) {
        uint32_t task = 0xffffffffu;
        if (left == 0) {
            while (tried < 8) {
                uint32_t qi = (home + tried) & 7;
                uint32_t k = 0;
                if (lane == 0) k = atomicAdd(w.xcd + (16 + qi) * LINE, w.pop);
                k = __shfl(k, 0);
                if (k < qcount[qi]) { from = qi; next = k; left = min(w.pop, qcount[qi] - k); break; }
                tried++;
            }
        }
        if (left == 0) {
            if (nb > 0) {
                wave_sync();
                flushBatch<NN, D>(p, w, home, queue, nearestTri, batch, nb, qn, sref, sdist, carry, false, lane,
                                  nnPending, auditBits, nnA, nnB, resNext, resEnd);
                queried += qn;
            }
            break;
        }
        task = w.xcdTasks[qoff[from] + next];
        next++;
        left--;
        if (w.skip) {  // passed over whole by k_task_far: no node can be a contact
            const uint32_t far = uniformU(w.taskFar[task]);
            if (far != 0u) {
                if (lane == 0) w.taskCount[task] = 0u;
                skippedNodes += far;
                continue;
            }
        }
        if (ld_volatile(p.control) == 0) {
            // The task's frame and row decision: k_task_far's record under the skip
            // (one row per lane), else from the grains' rows.
            float4 fr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (w.skip) {
                fr = w.taskFrame[TFRAME * size_t(task) + min(uint32_t(lane), TFRAME - 1)];
                auto U = [&](float v, int l) { return uint32_t(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
                auto F = [&](float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
                f.i = U(fr.x, 0); f.j = U(fr.y, 0); f.si = U(fr.z, 0); f.sj = U(fr.w, 0);
                f.qi = {F(fr.x, 1), F(fr.y, 1), F(fr.z, 1), F(fr.w, 1)};
                f.qj = {F(fr.x, 2), F(fr.y, 2), F(fr.z, 2), F(fr.w, 2)};
                f.branch = {F(fr.x, 3), F(fr.y, 3), F(fr.z, 3)};
                f.s_i = F(fr.w, 3);
                f.s_j = F(fr.x, 4); f.wi = F(fr.y, 4); f.wj = F(fr.z, 4); f.sampleBase = U(fr.w, 4);
            } else {
                const uint2 ij = p.pairs[task >> 1];
                const uint32_t pass = task & 1;
                f.i = pass == 0 ? ij.x : ij.y;
                f.j = pass == 0 ? ij.y : ij.x;
                f.si = p.shapeId[f.i];
                f.sj = p.shapeId[f.j];
                f.sampleBase = p.sampleInfo[f.si].x;
                f.qi = v4of(p.rotation[f.i]);
                f.qj = v4of(p.rotation[f.j]);
                const float4 bi = p.body[f.i], bj = p.body[f.j];
                f.s_i = bi.x; f.s_j = bj.x; f.wi = bi.w; f.wj = bj.w;
                f.branch = centerBranch(p, f.i, f.j);
            }
            const uint4 trec = w.templateRec[TREC * size_t(f.sj) + (uint32_t(lane) & (TREC - 1))];
            // Temporal skip.  The node's target-frame point is
            //   x = R_j^T (R_i a - b) / s_j,   a = s_i p,
            // and for unit quaternions |R(q) v - R(q') v| <= 2 |q - q'| |v|, so since
            // the anchor it has moved at most
            //   (2 |dq_j| (|a| + |b|) + 2 |dq_i| |a| + |db|) / s_j = c0 + c1 |p|.
            // The exact signed distance is 1-Lipschitz in x.  A node whose anchor
            // distance exceeds that bound (plus SKIP_MARGIN) is still outside the
            // target: not a contact, passed over.  One whose anchor distance is
            // negative was inside the target: it is queued without the rejection plane
            // and hull tests (below minus the bound it is still inside, so they would
            // pass it; otherwise they are only filters, and a node they would have
            // rejected costs one exact query that finds it outside).  Either way the
            // queue holds every contact in walk order and the hits and their ranks are
            // unchanged.
            // A task without a valid anchor, or whose bound has grown past
            // skipLimit, tests and queries every node and re-anchors; the rejection
            // plane of the anchor (any hull plane is a valid filter) is reused.
            bool skipMode = false, rebased = false;
            float c0 = 0.0f, c1 = 0.0f;
            int8_t* dist = nullptr;
            int8_t* refreshDist = nullptr;
            uint32_t skipped = 0, anchoredPlane = 0xffffffffu;
            // The task's persistent rows, used only while their owner is this pair.
            uint32_t rowTask = task;
            bool rows = false;
            uint32_t flags = 0u;
            if (w.skip) {  // rows exist only under the skip
                rowTask = uniformU(__builtin_amdgcn_readlane(__float_as_int(fr.x), 5));
                flags = uniformU(__builtin_amdgcn_readlane(__float_as_int(fr.y), 5));
                rows = (flags & 1u) != 0u;
            }
            uint16_t* warmTri = w.warm && rows ? uniformPtr(w.nodeTri + size_t(rowTask) * w.nodeStride) : nullptr;
            if (rows) {
                dist = uniformPtr(w.nodeDist + size_t(rowTask) * w.nodeStride);
                c0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fr.z), 5));
                c1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fr.w), 5));
                const float reachAll = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fr.x), 6));
                skipMode = (flags & 2u) != 0u;
                rebased = (flags & 4u) != 0u;
                if (skipMode) {
                    anchoredPlane = uniformU(__builtin_amdgcn_readlane(__float_as_int(fr.y), 6));
                } else if (rebased) {
                    // The anchor has expired: re-anchor at the current poses by moving
                    // every stored bound toward zero by the motion since the old anchor
                    // (rounded up to whole quanta).  A positive value stays a lower bound
                    // of its node's distance and a negative one an upper bound, so they
                    // hold at the new anchor with zero motion; this collection then
                    // overwrites the value of every node and leaf its tests reach.
                    const uint32_t shrink = uint32_t(ceilf(uniformF(reachAll) / SKIP_QUANTUM)) + 1u;
                    const uint32_t words = (p.sampleInfo[f.si].y + 3) / 4;
                    uint32_t* d32 = reinterpret_cast<uint32_t*>(dist);
                    for (uint32_t k = lane; k < words; k += WAVE) d32[k] = shrinkBounds(d32[k], shrink, false);
                    uint32_t* l32 = reinterpret_cast<uint32_t*>(w.leafDist + size_t(rowTask) * w.clusterStride);
                    const uint32_t leafWords = (p.meshClusterInfo[f.si].y + 3) / 4;
                    for (uint32_t k = lane; k < leafWords; k += WAVE) l32[k] = shrinkBounds(l32[k], shrink, true);
                    __threadfence_block();
                    refreshDist = dist;
                    skipMode = true;
                    c0 = 0.0f;
                    c1 = 0.0f;
                } else {
                    refreshDist = dist;
                    const uint32_t words = (p.sampleInfo[f.si].y + 3) / 4;
                    uint32_t* d32 = reinterpret_cast<uint32_t*>(dist);
                    for (uint32_t k = lane; k < words; k += WAVE) d32[k] = 0u;
                    __threadfence_block();
                }
            }
            // Unchanged-winner shortcut: the motion bound since this task's previous
            // collection and the generations that tag each node's stored gap.
            // Generations run 1..65535 per task; on wrap the task's words are
            // cleared, so a stale word can never carry the expected generation.
            GapCtx gc{nullptr, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0}, 0u, 0u};
            if (warmTri && w.nodeGap) {
                const float4* lf = w.lastFrame + 4 * size_t(rowTask);
                const float4 l0 = lf[0], l1 = lf[1], l2 = lf[2], l3 = lf[3];
                uint32_t* row = w.nodeGap + size_t(rowTask) * w.nodeStride;
                const uint32_t genPrev = __float_as_uint(l2.w) & 0xffffu;
                uint32_t genNew = genPrev + 1u;
                if (genNew > 0xffffu) {
                    const uint32_t nodes = p.sampleInfo[f.si].y;
                    for (uint32_t k = lane; k < nodes; k += WAVE) row[k] = 0u;
                    __threadfence_block();
                    genNew = 1u;
                }
                const bool valid = __float_as_uint(l3.w) == 1u && __float_as_uint(l3.x) == f.i &&
                                   __float_as_uint(l3.y) == f.j &&
                                   __float_as_uint(l3.z) == ld_volatile(w.counters + 10 * LINE) && genNew != 1u;
                gc.row = uniformPtr(row);
                gc.qiPrev = {uniformF(l0.x), uniformF(l0.y), uniformF(l0.z), uniformF(l0.w)};
                gc.qjPrev = {uniformF(l1.x), uniformF(l1.y), uniformF(l1.z), uniformF(l1.w)};
                gc.branchPrev = {uniformF(l2.x), uniformF(l2.y), uniformF(l2.z)};
                gc.genPrev = uniformU(valid ? genPrev : 0u);
                gc.genNew = uniformU(genNew);
            }
            // Rejection plane: largest violation at the source centre, first in
            // plane order on ties (strict >, initial -1e30), as a wave reduction
            // over (value, sequence) pairs.
            if (skipMode && !rebased) {
                f.plane = anchoredPlane == 0xffffffffu ? make_float4(0, 0, 0, 0) : p.meshPlanes[anchoredPlane];
            } else {
                v3 centre = -qinv(f.qj, f.branch) / f.s_j;
                float v0 = -1e30f;
                uint32_t s0 = 0xffffffffu, i0 = 0xffffffffu;
                uint32_t sq = 0;
                for (uint32_t g = 0; g < 6; g++) {
                    uint4 hull = p.meshHullInfo[6 * f.sj + g];
                    for (uint32_t k = lane; k < hull.y; k += WAVE) {
                        float4 e = p.meshPlanes[hull.x + k];
                        float v = dot(xyz(e), centre) + e.w;
                        if (v > v0) { v0 = v; s0 = sq + k; i0 = hull.x + k; }
                    }
                    sq += hull.y;
                }
                for (int d = 1; d < WAVE; d <<= 1) {
                    float ov = __shfl_xor(v0, d);
                    uint32_t os = __shfl_xor(s0, d), oi = __shfl_xor(i0, d);
                    if (ov > v0 || (ov == v0 && os < s0)) { v0 = ov; s0 = os; i0 = oi; }
                }
                anchoredPlane = i0;
                f.plane = i0 == 0xffffffffu ? make_float4(0, 0, 0, 0) : p.meshPlanes[i0];
            }
            uniformFrame(f);
            hullCtx(&hulls[wv], trec, lane);
            const HullCtx& hc = hulls[wv];
            const TargetRoot tg{uniformU(__shfl(trec.w, 0)), uniformU(__shfl(trec.w, 1)),
                                w.warm ? uniformU(__shfl(trec.x, 5)) : 0u, w.warm ? uniformU(__shfl(trec.y, 5)) : 0u};
            const bool refresh = refreshDist != nullptr;
            // Learned cull: on for targets whose network admits a cull.  A task with
            // persistent rows evaluates it only when it re-anchors: the nodes it drops
            // then keep the network's distance bound, so the temporal skip passes
            // them over until the next anchor, and the nodes it keeps were within its
            // reach of the surface, which the motion allowed before the next anchor
            // barely changes.  A node whose signed distance is known to be at most
            // nnNear is never dropped (the network's validated error), so it is
            // queried without evaluating it.
            bool nnOn = false;
            float nnNear = 0.0f;
            if constexpr (NN) {
                if (f.sj < w.nnTemplates && (!rows || refresh)) {
                    const float* fb = w.nnF + size_t(f.sj) * nn::FSTRIDE;
                    nnOn = fb[nn::FMARGIN] < 1e30f;
                    nnNear = fb[nn::FNEAR];
                }
                nnOn = uniformU(nnOn ? 1u : 0u) != 0u;
                nnNear = uniformF(nnNear);
            }
            const bool gapTask = gc.row != nullptr && gc.genPrev != 0u;
            const uint32_t triFirst = gapTask ? tg.triFirst : 0u;
            const uint32_t triCount = gapTask ? tg.triCount : 0u;
            batchPut(batch + nb, f, tg, task, rows ? rowTask : NO_ROW, refresh, gc, lane);
            nb++;
            // Leaf culls, lane-parallel.  Internal clusters are not tested: a leaf's
            // own test is a valid filter, and a leaf the preorder walk would prune
            // with an ancestor holds only nodes outside the target, so the hits and
            // their walk order (leaves in preorder, nodes in leaf order) are the
            // same.  Under the temporal skip a leaf whose stored bound clears the
            // motion bound of its box is passed over untested; a re-anchoring task
            // records the cull's evidence for each rejected leaf.
            const uint4 info = p.meshClusterInfo[f.si];
            int8_t* leafDist = rows ? uniformPtr(w.leafDist + size_t(rowTask) * w.clusterStride) : nullptr;
            uint16_t* leafPlane = rows && w.leafPlane ? uniformPtr(w.leafPlane + size_t(rowTask) * w.clusterStride) : nullptr;
            uint32_t leafSkipped = 0;  // per lane, summed below
            uint4 cp0 = make_uint4(0u, 0u, 0u, 0u);  // lane c: cluster c's packed row (the first 64 clusters)
            for (uint32_t base = 0; base < info.y; base += WAVE) {
                const uint32_t c = base + lane;
                bool test = false, nearLeaf = false;
                if (c < info.y) {
                    const uint4 cp = w.clusterPacked[info.x + c];
                    if (base == 0) cp0 = cp;
                    if (cp.y > 0) {
                        bool farLeaf = false;
                        if (skipMode) {
                            const float radius = __uint_as_float(cp.z);
                            farLeaf = float(leafDist[c]) * SKIP_QUANTUM > c0 + c1 * radius + SKIP_MARGIN;
                            if (farLeaf) leafSkipped += cp.y;
                            // A leaf with a node inside the target at the anchor is a
                            // candidate without its cull test (the test only filters).
                            nearLeaf = !farLeaf && leafDist[c] < 0;
                        }
                        test = !farLeaf && !nearLeaf;
                        // The leaf's cached separating plane: its box lies more than the
                        // skip margin outside that hull plane (bounded through the box's
                        // circumscribed sphere), so none of its nodes can be a contact.
                        if (test && leafPlane) {
                            const uint32_t cached = leafPlane[c];
                            const uint32_t g = (cached >> 12) & 7u, k = cached & 0xfffu;
                            if ((cached & PLANE_TAG) != 0u && g < 6u && k < hc.range[g].y) {
                                const float4 pl = p.meshPlanes[hc.range[g].x + k];
                                const float4 sph = w.clusterSphere[info.x + c];
                                const v3 cc = qinv(f.qj, qrot(f.qi, xyz(sph) * f.s_i) - f.branch) / f.s_j;
                                const float e = dot(xyz(pl), cc) + pl.w - (sph.w * f.s_i / f.s_j * 1.0001f + 1e-6f);
                                if (e > SKIP_MARGIN) {
                                    test = false;
                                    if (refreshDist) leafDist[c] = quantizeDistance(e);
                                }
                            }
                        }
                    }
                }
                const uint64_t tests = __ballot(test);
                uint64_t m = __ballot(nearLeaf);
                if (tests != 0) {
                    float out = 0.0f;
                    uint32_t code = NO_PLANE;
                    if (uint32_t(__popcll(tests)) <= COOP_LEAVES)
                        out = leavesOutsideWave(p, hc, tests, info.x + base, f.qi, f.qj, f.s_i, f.s_j, f.branch, lane, code);
                    else if (test)
                        out = clusterOutside(p, hc, f.qi, f.qj, f.s_i, f.s_j, f.branch, info.x + c, code);
                    const bool ok = test && !(out > 0.0f);
                    if (test && !ok && refreshDist) leafDist[c] = quantizeDistance(out);
                    if (test && !ok && leafPlane && code != NO_PLANE) leafPlane[c] = uint16_t(code);
                    m |= __ballot(ok);
                }
                if (lane == 0) cand[base / WAVE] = m;
            }
            if (skipMode) {
                for (int d = 1; d < WAVE; d <<= 1) leafSkipped += __shfl_xor(leafSkipped, d);
                skipped += leafSkipped;
            }
            wave_sync();
            // Candidate leaves in preorder feed the query queue.
            const uint32_t words = (info.y + WAVE - 1) / WAVE;
            for (uint32_t word = 0; word < words; word++) {
                uint64_t bits = cand[word];
                while (bits != 0) {
                    const uint32_t rel = word * WAVE + uint32_t(__builtin_ctzll(bits));
                    bits &= bits - 1;
                    const uint4 range =
                        rel < uint32_t(WAVE)
                            ? make_uint4(uint32_t(__builtin_amdgcn_readlane(int(cp0.x), int(rel))),
                                         uint32_t(__builtin_amdgcn_readlane(int(cp0.y), int(rel))), 0u,
                                         uint32_t(__builtin_amdgcn_readlane(int(cp0.w), int(rel))))
                            : w.clusterPacked[info.x + rel];
                    bool pass_ = false, test = false;
                    uint32_t node = 0;
                    bool far = false;
                    float nearBound = 1e30f;  // an upper bound of the node's signed distance, when one is known
                    uint32_t slot = 0, cached = 0;  // cached: the slot's warm-start word
                    uint32_t gapWordNow = 0;        // the slot's gap word, read ahead of the tests
                    v3 local = mk3(0.0f, 0.0f, 0.0f);
                    float4 sp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (uint32_t(lane) < range.y) {
                        node = p.meshSampleIndices[range.x + lane];
                        if (rows) slot = range.w + lane;
                        sp = w.sampleKeyed[range.x + lane];  // = p.sample[f.sampleBase + node]
                        if (warmTri) cached = warmTri[slot];
                        bool inside = false;
                        if (skipMode) {
                            const float stored = float(dist[slot]) * SKIP_QUANTUM;
                            const float reach = c0 + c1 * length(xyz(sp)) + SKIP_MARGIN;
                            far = stored > reach;
                            // Inside the target at the anchor: queued without the plane
                            // and hull tests, which only filter (a node that has since
                            // left the hull costs one exact query, never a contact).
                            inside = stored < 0.0f;
                            if (stored < 0.0f) nearBound = stored + reach;  // stored rounds toward zero
                        }
                        if (gapTask && !far) gapWordNow = gc.row[slot];
                        if (inside) {
                            pass_ = true;
                        } else if (!far) {
                            v3 arm = qrot(f.qi, xyz(sp) * f.s_i);
                            local = qinv(f.qj, arm - f.branch) / f.s_j;
                            float planePad = 64.0f * DEM_EPS * (1.0f + maxabs(local));
                            // The rejection plane, then the hull test; a node they
                            // reject keeps their evidence as its distance lower bound.
                            const float rejection = dot(xyz(f.plane), local) + f.plane.w;
                            if (rejection > planePad) {
                                if (refreshDist) refreshDist[slot] = quantizeDistance(rejection);
                            } else {
                                test = true;
                            }
                        }
                    }
                    // The node's cached separating plane: a hull plane of the target that it
                    // lies more than SKIP_MARGIN outside of keeps it out of every contact (its
                    // distance to the mesh exceeds that, far above the FP32 error of a
                    // query), so the group test, itself only a filter, is not needed.
                    if (test && (cached & PLANE_TAG) != 0u && w.planeCache) {
                        const uint32_t g = (cached >> 12) & 7u, k = cached & 0xfffu;
                        if (g < 6u) {
                            const uint2 r = hc.range[g];
                            if (k < r.y) {
                                const float4 pl = p.meshPlanes[r.x + k];
                                const float e = dot(xyz(pl), local) + pl.w;
                                if (e > SKIP_MARGIN) {
                                    test = false;
                                    if (refreshDist) refreshDist[slot] = quantizeDistance(e);
                                }
                            }
                        }
                    }
                    const uint64_t tests = __ballot(test);
                    if (tests != 0) {
                        float outside = 0.0f;
                        uint32_t code = NO_PLANE;
                        if (uint32_t(__popcll(tests)) <= COOP_NODES)
                            outside = nodesOutsideWave(p, hc, tests, local, lane, code);
                        else if (test)
                            outside = nodeOutside(p, hc, local, code);
                        if (test) {
                            pass_ = !(outside > 0.0f);
                            if (refreshDist && !pass_) refreshDist[slot] = quantizeDistance(outside);
                            if (!pass_ && warmTri && w.planeCache && code != NO_PLANE && code != cached)
                                warmTri[slot] = uint16_t(code);
                        }
                    }
                    if (skipMode) skipped += __popcll(__ballot(far));
                    // Unchanged-winner shortcut: while the node's stored gap less
                    // twice its motion since its last query stays above 2 GAP_EVAL,
                    // its nearest triangle is the same, so it needs no traversal and
                    // its gap shrinks by twice that motion.
                    uint32_t done = 0u;  // 1 + the proven nearest triangle, 0: traverse
                    if (pass_ && gapTask) {
                        const uint32_t last = cached;
                        if (last != 0u && last <= triCount) {
                            const uint32_t word = gapWordNow;
                            if ((word >> 16) == gc.genPrev) {
                                const v3 a = xyz(sp) * f.s_i;
                                const v3 now = qinv(f.qj, qrot(f.qi, a) - f.branch) / f.s_j;  // queryPoint
                                const v3 then = qinv(gc.qjPrev, qrot(gc.qiPrev, a) - gc.branchPrev) / f.s_j;
                                const float moved =
                                    1.0001f * length(now - then) + GAP_MOTION_ABS * (1.0f + maxabs(now));
                                const float left = float(word & 0xffffu) * GAP_QUANTUM - 2.0f * moved;
                                if (left > 2.0f * GAP_EVAL) {
                                    gc.row[slot] = gapWord(gc.genNew, left);
                                    done = triFirst + last;
                                }
                            }
                        }
                    }
                    if (gapTask) shortcutNodes += __popcll(__ballot(done != 0u));
                    uint64_t m = __ballot(pass_);
                    uint32_t n = __popcll(m);
                    if (qn + n > QN) {  // the batch ends inside this task
                        wave_sync();
                        if (lane == 0) batch[nb - 1].end = qn | (refresh ? 1u << 16 : 0u);
                        wave_sync();
                        carry = flushBatch<NN, D>(p, w, home, queue, nearestTri, batch, nb, qn, sref, sdist, carry,
                                                  true, lane, nnPending, auditBits, nnA, nnB, resNext,
                                                  resEnd);
                        queried += qn;
                        qn = 0;
                        pending = 0;
                        wave_sync();
                        batchPut(batch, f, tg, task, rows ? rowTask : NO_ROW, refresh, gc, lane);
                        nb = 1;
                        wave_sync();
                    }
                    bool nnCand = false;
                    if constexpr (NN) {
                        if (nnOn) {
                            nnCand = pass_ && done == 0u && nearBound > nnNear;
                            nnPending += __popcll(__ballot(nnCand));
                        }
                    }
                    if (pass_) {
                        const uint32_t at = qn + __popcll(m & lanes_below(lane));
                        queue[at] = node | (slot << 16);
                        nearestTri[at] = done != 0u ? done - 1 : PENDING | (nnCand ? NN_CAND : 0u) | (nb - 1);
                    }
                    qn += n;
                    pending += __popcll(__ballot(pass_ && done == 0u));
                }
            }
            wave_sync();
            if (lane == 0) batch[nb - 1].end = qn | (refresh ? 1u << 16 : 0u);
            // A re-anchoring task records its leaves' bounds from its queries' results
            // below, so its batch ends with it.
            if (refresh || nb >= batchLimit || pending >= w.batchFill) {
                wave_sync();
                flushBatch<NN, D>(p, w, home, queue, nearestTri, batch, nb, qn, sref, sdist, carry, false, lane,
                                  nnPending, auditBits, nnA, nnB, resNext, resEnd);
                queried += qn;
                qn = 0;
                nb = 0;
                carry = 0;
                pending = 0;
            }
            wave_sync();
            if (refreshDist) {
                // Each candidate leaf's bound: the least of its nodes' stored values.
                __threadfence_block();
                for (uint32_t word = 0; word < words; word++) {
                    uint64_t bits = cand[word];
                    while (bits != 0) {
                        const uint32_t rel = word * WAVE + uint32_t(__builtin_ctzll(bits));
                        bits &= bits - 1;
                        const uint4 range = w.clusterPacked[info.x + rel];
                        int v = 127;
                        if (uint32_t(lane) < range.y) v = dist[range.w + lane];
                        for (int d = 1; d < WAVE; d <<= 1) v = min(v, __shfl_xor(v, d));
                        if (lane == 0) leafDist[rel] = int8_t(v);
                    }
                }
            }
            if (gc.row && lane == 0) {  // this collection's frame, for the next one's motion bound
                float4* lf = w.lastFrame + 4 * size_t(rowTask);
                lf[0] = f4(f.qi);
                lf[1] = f4(f.qj);
                lf[2] = f4(f.branch, __uint_as_float(gc.genNew));
                lf[3] = make_float4(__uint_as_float(f.i), __uint_as_float(f.j),
                                    __uint_as_float(ld_volatile(w.counters + 10 * LINE)), __uint_as_float(1u));
            }
            if (refreshDist) {
                // The task's least leaf bound, for the all-far test of later collections.
                __threadfence_block();
                int leafMin = 127;
                for (uint32_t c = lane; c < info.y; c += WAVE)
                    if (w.clusterPacked[info.x + c].y > 0) leafMin = min(leafMin, int(leafDist[c]));
                for (int d = 1; d < WAVE; d <<= 1) leafMin = min(leafMin, __shfl_xor(leafMin, d));
                if (lane == 0) {
                    atomicAdd(w.counters + 13 * LINE, 1u);
                    float4* an = w.anchor + 4 * size_t(rowTask);
                    an[0] = f4(f.qi);
                    an[1] = f4(f.qj);
                    an[2] = f4(f.branch, __uint_as_float(anchoredPlane));
                    an[3] = make_float4(__uint_as_float(f.i), __uint_as_float(f.j),
                                        __uint_as_float(ld_volatile(w.counters + 10 * LINE)),
                                        __uint_as_float(1u | (uint32_t(uint8_t(int8_t(leafMin))) << 8)));
                }
            }
            skippedNodes += skipped;
        } else if (lane == 0) {
            w.taskCount[task] = 0u;
        }
    }
    for (uint32_t k = resNext + lane; k < min(resEnd, w.stageCap); k += WAVE)  // reserved, never used
        w.stageKey[size_t(home) * w.stageCap + k] = make_uint2(NO_TASK, 0u);
    if (lane == 0 && queried) atomicAdd(w.counters + 2 * LINE, queried);
    if (lane == 0 && skippedNodes) atomicAdd(w.counters + 11 * LINE, skippedNodes);
    if (lane == 0 && shortcutNodes) atomicAdd(w.counters + 12 * LINE, shortcutNodes);
}

// The plain kernel keeps <= 96 VGPRs and 29 KB of LDS per workgroup so five
// workgroups (20 waves) share a CU: the exact queries are L2-latency bound.
template <int QN>
__global__ __launch_bounds__(COLLECT_WAVES * WAVE) __attribute__((amdgpu_waves_per_eu(SAD_COLLECT_WAVES_PER_EU)))
void k_collect(P p, W w) { collectBody<false, QN>(p, w); }
template <int QN>
__global__ __launch_bounds__(COLLECT_WAVES * WAVE) __attribute__((amdgpu_waves_per_eu(SAD_COLLECT_NN_WAVES_PER_EU)))
void k_collect_nn(P p, W w) { collectBody<true, QN>(p, w); }

// control[2] = total contacts, per-pair counts (solver.slang:397-403 semantics).
__global__ void k_contact_total(P p, W w) {
    uint32_t pairs = min(p.control[1], p.c.pairCapacity);
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id == 0) {  // this step's builders have all run
        if (w.counters[DERIVED_FLAG * LINE]) atomicAdd(w.totals + 7, 1ull);
        w.counters[DERIVED_FLAG * LINE] = 0u;
    }
    if (ld_volatile(p.control) != 0) return;
    if (id < pairs) p.contactCounts[id] = w.taskCount[2 * id] + w.taskCount[2 * id + 1];
    if (id == 0) {
        uint32_t total = w.counters[4 * LINE];
        if (total > p.c.contactCapacity) { fail(p, 32); total = p.c.contactCapacity; }
        p.control[2] = total;
        atomicAdd(w.totals + 0, 1ull);
        atomicAdd(w.totals + 1, 2ull * pairs);
        atomicAdd(w.totals + 2, (unsigned long long)w.counters[2 * LINE]);
        atomicAdd(w.totals + 3, (unsigned long long)total);
        atomicAdd(w.totals + 4, (unsigned long long)w.counters[11 * LINE]);
        atomicAdd(w.totals + 5, (unsigned long long)w.counters[12 * LINE]);
        atomicAdd(w.totals + 6, (unsigned long long)w.counters[13 * LINE]);
    }
}
__global__ void k_scatter(P p, W w) {
    if (ld_volatile(p.control) != 0) return;
    const uint32_t queue = blockIdx.y;
    uint32_t staged = min(w.stageCount[queue * LINE], w.stageCap);
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < staged; r += gridDim.x * blockDim.x) {
        size_t k = size_t(queue) * w.stageCap + r;
        uint2 key = ld_stream(w.stageKey + k);
        if (key.x == NO_TASK) continue;  // a slot its wave reserved but did not use
        uint32_t id = w.taskOffset[key.x] + key.y;
        if (id < p.c.contactCapacity) {
            p.meta[id] = ld_stream(w.stageMeta + k);
            p.pointDepth[id] = ld_stream(w.stagePd + k);
            p.normalArea[id] = ld_stream(w.stageNa + k);
        }
    }
}

// ---------------------------------------------------------------- SDF contacts (solver.slang:188-200, 319-343, 536-574)
// sdf: the template grid's trilinear value and piecewise gradient, 1e20 outside the
// grid; lerp(x, y, s) is x + s (y - x).  A corner index beyond the distance buffer
// (inconsistent grid rows) raises flag 2 and reads as outside.  It comes in two parts,
// so a caller can issue the corner loads of several points before it uses any:
// sdfLoad fetches the 8 corners (in-bounds reads only; the cell is taken only for
// points inside the grid), sdfFinish forms the value and gradient.
struct SdfCorners {
    float f[8];
    v3 g;
    float spacing;
    bool inside, ok;
};
__device__ SdfCorners sdfLoad(const P& p, uint4 meta, float4 o, v3 q) {  // meta, o: the template's grid rows
    SdfCorners c;
    c.g = (q - xyz(o)) / o.w;
    c.spacing = o.w;
    c.inside = c.g.x >= 0.0f && c.g.y >= 0.0f && c.g.z >= 0.0f && !(c.g.x >= float(int64_t(meta.y) - 1)) &&
               !(c.g.y >= float(int64_t(meta.z) - 1)) && !(c.g.z >= float(int64_t(meta.w) - 1));
    const uint32_t ax = uint32_t(int(floorf(c.inside ? c.g.x : 0.0f)));
    const uint32_t ay = uint32_t(int(floorf(c.inside ? c.g.y : 0.0f)));
    const uint32_t az = uint32_t(int(floorf(c.inside ? c.g.z : 0.0f)));
    c.ok = true;
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
        const uint32_t index = meta.x + ((ax + (k >> 2)) * meta.z + (ay + ((k >> 1) & 1u))) * meta.w + (az + (k & 1u));
        const bool fits = index < p.distanceCount;
        c.ok = c.ok && fits;
        c.f[k] = p.distance[fits ? index : 0u];
    }
    return c;
}
__device__ v4 sdfFinish(const P& p, const SdfCorners& c) {
    const v4 outside = {0, 0, 0, 1e20f};
    if (!c.inside) return outside;
    if (!c.ok) { fail(p, 2); return outside; }
    const v3 g = c.g;
    const int ax = int(floorf(g.x)), ay = int(floorf(g.y)), az = int(floorf(g.z));
    const v3 t = {g.x - float(ax), g.y - float(ay), g.z - float(az)};
    const v3 x = mk3(1.0f, 1.0f, 1.0f) - t;
    float value = 0;
    v3 grad = {0, 0, 0};
    for (uint32_t i = 0; i < 2; i++)
        for (uint32_t j = 0; j < 2; j++)
            for (uint32_t k = 0; k < 2; k++) {
                const v3 b = mk3(float(i), float(j), float(k));
                const v3 w = x + b * (t - x);
                const float f = c.f[4 * i + 2 * j + k];
                value += f * w.x * w.y * w.z;
                const v3 d = {(2.0f * float(i) - 1.0f) * w.y * w.z, (2.0f * float(j) - 1.0f) * w.x * w.z,
                              (2.0f * float(k) - 1.0f) * w.x * w.y};
                grad = grad + f * d / c.spacing;
            }
    return {grad.x, grad.y, grad.z, value};
}
// Exact evaluations per lane per batch; the queue holds one batch plus a row of 64;
// per wave the contact-node bitmask covers 4,096 nodes.
constexpr uint32_t SDF_PER_LANE = 2, SDF_QUEUE = 64 * SDF_PER_LANE + 64, SDF_MASK_WORDS = 128;
// Per template, the squared radius (template units) outside which every grid cell
// has only nonnegative corners: the largest |corner| over negative grid values, plus
// two spacings (more than a cell diagonal, covering the rounding of the cell index).
// Trilinear weights are nonnegative, so beyond it the field is >= 0 and no node can
// be in contact; no value below zero gives -1 (nothing reaches).
__global__ __launch_bounds__(256) void k_sdf_reach(P p, W w, uint32_t* largest) {
    __shared__ float best[256];
    const uint32_t shape = blockIdx.x;
    if (threadIdx.x == 0) atomicMax(largest, p.sampleInfo[shape].y);
    const uint4 meta = p.grids[shape];
    const float4 o = p.gridOrigin[shape];
    const uint64_t count = uint64_t(meta.y) * meta.z * meta.w;
    float r = -1.0f;
    for (uint64_t k = threadIdx.x; k < count; k += 256) {
        const uint64_t index = meta.x + k;
        if (index >= p.distanceCount || !(p.distance[index] < 0.0f)) continue;
        const uint32_t ix = uint32_t(k / (uint64_t(meta.z) * meta.w)), rest = uint32_t(k % (uint64_t(meta.z) * meta.w));
        const v3 g = xyz(o) + o.w * mk3(float(ix), float(rest / meta.w), float(rest % meta.w));
        r = fmaxf(r, length(g));
    }
    best[threadIdx.x] = r;
    __syncthreads();
    for (uint32_t step = 128; step > 0; step >>= 1) {
        if (threadIdx.x < step) best[threadIdx.x] = fmaxf(best[threadIdx.x], best[threadIdx.x + step]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float reach = best[0] < 0.0f ? -1.0f : (best[0] + 2.0f * o.w) * 1.0001f;
        w.sdfReach[shape] = reach < 0.0f ? -1.0f : reach * reach;
    }
}
// Per template, a spatial index of its nodes for the count pass's screen: nodes
// sorted by the Morton code of their position (10 bits per axis over the node
// bounding box; ties by node id), cut into leaves of 64, each with a bounding sphere
// (radius padded by 1e-5 relative).  A leaf whose sphere lies beyond the reach
// holds no node that can be in contact.  Templates with more than 4,096 nodes are
// refused before this runs.
__global__ __launch_bounds__(256) void k_sdf_index(P p, W w) {
    __shared__ uint64_t keys[32 * SDF_MASK_WORDS];
    __shared__ float part[6][256];
    __shared__ float box[6];
    const uint32_t shape = blockIdx.x, t = threadIdx.x;
    const uint4 info = p.sampleInfo[shape];
    const uint32_t n = min(info.y, 32u * SDF_MASK_WORDS);
    float lo3[3] = {INFINITY, INFINITY, INFINITY}, hi3[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = t; k < n; k += 256) {
        const float4 q = p.sample[info.x + k];
        const float c[3] = {q.x, q.y, q.z};
        for (int a = 0; a < 3; a++) lo3[a] = fminf(lo3[a], c[a]), hi3[a] = fmaxf(hi3[a], c[a]);
    }
    for (int a = 0; a < 3; a++) part[a][t] = lo3[a], part[3 + a][t] = hi3[a];
    __syncthreads();
    if (t < 6) {
        float v = part[t][0];
        for (uint32_t k = 1; k < 256; k++) v = t < 3 ? fminf(v, part[t][k]) : fmaxf(v, part[t][k]);
        box[t] = v;
    }
    __syncthreads();
    uint32_t size = 1;
    while (size < n) size <<= 1;
    for (uint32_t k = t; k < size; k += 256) {
        uint64_t key = ~0ull;
        if (k < n) {
            const float4 q = p.sample[info.x + k];
            const float c[3] = {q.x, q.y, q.z};
            uint32_t code = 0;
            for (int a = 0; a < 3; a++) {
                const float extent = box[3 + a] - box[a];
                const float u = extent > 0.0f ? (c[a] - box[a]) / extent : 0.0f;
                const uint32_t cell = min(1023u, uint32_t(fmaxf(u, 0.0f) * 1024.0f));
                for (int bit = 0; bit < 10; bit++) code |= ((cell >> bit) & 1u) << (3 * bit + a);
            }
            key = (uint64_t(code) << 16) | k;
        }
        keys[k] = key;
    }
    __syncthreads();
    for (uint32_t span = 2; span <= size; span <<= 1)  // bitonic sort, ascending
        for (uint32_t stride = span >> 1; stride > 0; stride >>= 1) {
            for (uint32_t k = t; k < size; k += 256) {
                const uint32_t partner = k ^ stride;
                if (partner > k) {
                    const bool up = (k & span) == 0;
                    const uint64_t a = keys[k], b = keys[partner];
                    if ((a > b) == up) { keys[k] = b; keys[partner] = a; }
                }
            }
            __syncthreads();
        }
    uint16_t* order = w.sdfOrder + size_t(shape) * (32 * SDF_MASK_WORDS);
    for (uint32_t k = t; k < n; k += 256) order[k] = uint16_t(keys[k] & 0xffffu);
    const uint32_t leaves = (n + 63) / 64;
    if (t < leaves) {
        v3 lo = {INFINITY, INFINITY, INFINITY}, hi = {-INFINITY, -INFINITY, -INFINITY};
        const uint32_t end = min(n, 64 * t + 64);
        for (uint32_t k = 64 * t; k < end; k++) {
            const v3 q = xyz(p.sample[info.x + uint32_t(keys[k] & 0xffffu)]);
            lo = {fminf(lo.x, q.x), fminf(lo.y, q.y), fminf(lo.z, q.z)};
            hi = {fmaxf(hi.x, q.x), fmaxf(hi.y, q.y), fmaxf(hi.z, q.z)};
        }
        const v3 centre = 0.5f * (lo + hi);
        float radius = 0.0f;
        for (uint32_t k = 64 * t; k < end; k++)
            radius = fmaxf(radius, length(xyz(p.sample[info.x + uint32_t(keys[k] & 0xffffu)]) - centre));
        w.sdfLeaves[size_t(shape) * 64 + t] = make_float4(centre.x, centre.y, centre.z, radius * 1.00001f + 1e-30f);
    }
}
// nodeContact (solver.slang:319-343), SDF branch, after the node's local point.
__device__ bool sdfNodeContact(const P& p, const SdfCorners& corners, v3 arm, v4 qj, float area, float si,
                               float wi, float sj, float wj, float4& pd, float4& na) {
    const v4 f = sdfFinish(p, corners);
    const float depth = -f.w * sj;
    pd = f4(arm, depth);
    na = make_float4(0, 0, 0, 0);
    if (depth <= 0) return false;
    const v3 grad = qrot(qj, xyz(f));
    const float len = length(grad);
    if (len < 1e-7f || !isfinite(depth)) { fail(p, 2); return false; }
    na = f4(grad / len, area * si * si);
    if (depth > p.c.maxOverlap * fminf(wi, wj)) fail(p, 4);
    return true;
}
// countContacts / emitContacts in the CPU's order (cpu_scheduling.slang:102-131):
// one wave per pair takes the first grain's nodes (source, the partner the target),
// then the second's; contacts are written in node order.
//
// Count pass: a node whose local point lies beyond the target's reach is not in
// contact: |local| = s_i |p - R_i^T b / s_i| / s_j (rotations keep lengths), so the
// screen costs one subtraction and a dot product per node, with a 1e-4 relative
// margin over the FP32 error of both evaluations.  Survivors are queued in node order
// in LDS and evaluated exactly 64 at a time (the node's local point as the reference
// computes it, then the field); each direction's contact nodes are recorded as a
// bitmask.  Emit pass: only the recorded nodes are evaluated again (the evaluation is
// deterministic, so the same nodes hit with the same values), 64 node ids per step,
// ranked by ballot.  The counts and offsets of the two directions also go to the
// per-direction task rows the gather walks, as after a mesh collection.
template <bool EMIT>
__global__ __launch_bounds__(256) void k_sdf_contacts(P p, W w) {
    // The emit pass evaluates only recorded contact nodes (a few per direction): one
    // per lane keeps its registers low.
    constexpr uint32_t PER_LANE = EMIT ? 1u : SDF_PER_LANE;
    __shared__ uint32_t queueMem[4][SDF_QUEUE];
    __shared__ uint32_t maskMem[4][SDF_MASK_WORDS];
    const int lane = threadIdx.x & 63;
    uint32_t* queue = queueMem[threadIdx.x >> 6];
    uint32_t* bits = maskMem[threadIdx.x >> 6];
    // The pair index is the same for the whole wave: held in a scalar register, the
    // pair's rows (ids, bodies, poses, grid rows) come through scalar loads.
    const uint32_t pair = uniformU(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (pair >= p.c.pairCapacity) return;
    const bool live = pair < p.control[1] && ld_volatile(p.control) == 0;
    if (EMIT && !live) return;
    uint32_t count = 0, forward = 0, evaluated = 0;  // hits, hits of the first direction, nodes
    if (live) {
        const uint2 ij = p.pairs[pair];
        const uint32_t base = EMIT ? p.contactOffsets[pair] : 0;
        const uint32_t words = w.sdfWords;
        for (uint32_t pass = 0; pass < 2; pass++) {
            const uint32_t i = pass ? ij.y : ij.x, j = pass ? ij.x : ij.y;
            const uint32_t targetShape = p.shapeId[j];
            const uint4 info = p.sampleInfo[p.shapeId[i]];
            const float4 bi = p.body[i], bj = p.body[j];
            const v4 qi = v4of(p.rotation[i]), qj = v4of(p.rotation[j]);
            const v3 branch = centerBranch(p, i, j);
            const uint4 gridRow = p.grids[targetShape];
            const float4 gridOrigin = p.gridOrigin[targetShape];
            uint32_t* record = w.sdfHits + size_t(2 * pair + pass) * words;
            evaluated += info.y;
            // Exact evaluation of one node per active lane; hits ranked by ballot.
            // Exact evaluation of queue entries [0, n), n <= 64 * PER_LANE, several per
            // lane (entries lane, 64 + lane, ...): every point's corner loads are issued
            // before any is used.  A lane without an entry in a row repeats entry 0
            // (same flags) and discards it.  Hits are ranked by ballot in entry order.
            auto evaluate = [&](uint32_t n) {
                uint32_t node[PER_LANE];
                v3 arm[PER_LANE];
                float area[PER_LANE];
                SdfCorners corners[PER_LANE];
#pragma unroll
                for (uint32_t u = 0; u < PER_LANE; u++) {
                    if (64 * u >= n) break;
                    const uint32_t e = 64 * u + lane;
                    node[u] = queue[e < n ? e : 0u];
                    const float4 sp = p.sample[info.x + node[u]];
                    arm[u] = qrot(qi, xyz(sp) * bi.x);
                    area[u] = sp.w;
                    corners[u] = sdfLoad(p, gridRow, gridOrigin, qinv(qj, arm[u] - branch) / bj.x);
                }
                for (uint32_t u = 0; u < PER_LANE && 64 * u < n; u++) {
                    float4 pd, na;
                    const bool hit = sdfNodeContact(p, corners[u], arm[u], qj, area[u], bi.x, bi.w, bj.x, bj.w, pd, na) &&
                                     64 * u + lane < n;
                    const uint64_t mask = __ballot(hit);
                    if (EMIT && hit) {
                        const uint32_t id = base + count + __popcll(mask & lanes_below(lane));
                        if (id < p.c.contactCapacity) {
                            p.meta[id] = make_uint4(i, j, node[u], 0);
                            p.pointDepth[id] = pd;
                            p.normalArea[id] = na;
                        } else {
                            fail(p, 32);
                        }
                    }
                    if (!EMIT && hit) atomicOr(bits + node[u] / 32, 1u << (node[u] % 32));
                    count += __popcll(mask);
                }
            };
            // Node-ordered queue of candidates, evaluated 64 * PER_LANE at a time.
            uint32_t queued = 0;
            auto push = [&](bool keep, uint32_t node) {
                const uint64_t mask = __ballot(keep);
                if (keep) queue[queued + __popcll(mask & lanes_below(lane))] = node;
                queued += __popcll(mask);
                wave_sync();
                if (queued >= 64 * PER_LANE) {
                    evaluate(64 * PER_LANE);
                    wave_sync();
                    const uint32_t rest = queued - 64 * PER_LANE;
                    const uint32_t moved = uint32_t(lane) < rest ? queue[64 * PER_LANE + lane] : 0;
                    wave_sync();
                    if (uint32_t(lane) < rest) queue[lane] = moved;
                    queued = rest;
                    wave_sync();
                }
            };
            if (EMIT) {
                // The direction's bitmask (<= 128 words) is held two words per lane.
                const uint32_t low = uint32_t(lane) < words ? record[lane] : 0u;
                const uint32_t high = uint32_t(lane) + 64 < words ? record[lane + 64] : 0u;
                for (uint32_t first = 0; first < info.y && first / 32 < words; first += 64) {
                    const uint32_t k = first / 32;  // even
                    const uint32_t w0 = __shfl(k < 64 ? low : high, int(k % 64));
                    const uint32_t w1 = __shfl(k + 1 < 64 ? low : high, int((k + 1) % 64));
                    const uint64_t set = uint64_t(w0) | (uint64_t(w1) << 32);
                    if (set) push((set >> lane) & 1u, first + lane);
                }
                if (queued) evaluate(queued);
                wave_sync();
            } else {
                const v3 centre = qinv(qi, branch) / bi.x;  // the target's centre in source template units
                const float bound = w.sdfReach[targetShape] * (bj.x * bj.x) * 1.0001f, scale = bi.x * bi.x;
                const bool recorded = info.y <= 32 * min(words, SDF_MASK_WORDS);
                for (uint32_t k = lane; k < words && k < SDF_MASK_WORDS; k += 64) bits[k] = 0;
                wave_sync();
                // Only the count and the set of contact nodes come out of this pass, so
                // the nodes are visited leaf by leaf: a leaf whose bounding sphere
                // stays beyond the reach (1e-3 relative margin over the node screen)
                // is passed over, the nodes of the others are screened one per lane.
                const uint32_t source = p.shapeId[i];
                const float limit = bound > 0.0f ? sqrtf(bound / scale) * 1.001f : -1.0f;
                const uint32_t leaves = (min(info.y, 32u * SDF_MASK_WORDS) + 63) / 64;
                bool near = false;
                if (uint32_t(lane) < leaves) {
                    const float4 leaf = w.sdfLeaves[size_t(source) * 64 + lane];
                    near = !(length(xyz(leaf) - centre) - leaf.w > limit);
                }
                uint64_t nearLeaves = __ballot(near);
                const uint16_t* order = w.sdfOrder + size_t(source) * (32 * SDF_MASK_WORDS);
                while (nearLeaves) {
                    const uint32_t leaf = uint32_t(__builtin_ctzll(nearLeaves));
                    nearLeaves &= nearLeaves - 1;
                    const uint32_t k = 64 * leaf + lane;
                    uint32_t node = 0;
                    bool keep = false;
                    if (k < info.y) {
                        node = order[k];
                        const v3 d = xyz(p.sample[info.x + node]) - centre;
                        keep = !(scale * dot(d, d) > bound);
                    }
                    push(keep, node);
                }
                if (queued) evaluate(queued);
                wave_sync();
                if (!recorded) fail(p, 2);  // more nodes than the bitmask rows hold
                for (uint32_t k = lane; k < words; k += 64) record[k] = k < SDF_MASK_WORDS ? bits[k] : 0;
                wave_sync();
            }
            if (pass == 0) forward = count;
        }
        if (EMIT && lane == 0) {
            w.taskOffset[2 * pair] = base;
            w.taskOffset[2 * pair + 1] = base + forward;
        }
    }
    if (!EMIT && lane == 0) {
        if (live) atomicAdd(w.totals + 2, (unsigned long long)evaluated);  // the "queried nodes" statistic
        p.contactCounts[pair] = count;
        w.taskCount[2 * pair] = forward;
        w.taskCount[2 * pair + 1] = count - forward;
    }
}
// contactTotal (solver.slang:568-574).
__global__ void k_sdf_total(P p, W w) {
    const uint32_t n = p.c.pairCapacity;
    const uint32_t total = p.contactOffsets[n - 1] + p.contactCounts[n - 1];
    p.control[2] = total;
    p.control[6] = 0;
    if (total > p.c.contactCapacity) fail(p, 32);
    atomicAdd(w.totals + 0, 1ull);
    atomicAdd(w.totals + 1, 2ull * min(p.control[1], n));
    atomicAdd(w.totals + 3, (unsigned long long)total);
}

// ---------------------------------------------------------------- plane contacts (solver.slang:596-667)
// collectWallContacts in the CPU's serial order (cpu_scheduling.slang:133-154):
// task t = wall * N + grain, the grain's sample clusters in preorder, a leaf's
// nodes in leaf order, ids after every mesh contact.  One wave per task walks the
// cluster tree under wave-uniform control; a leaf's nodes (<= 64) go one per lane
// and hits are ranked by ballot.  The count pass sizes each task, a scan gives the
// offsets, and the emit pass re-walks and writes the same hits.
template <bool EMIT>
__global__ __launch_bounds__(256) void k_wall_collect(P p, W w) {
    const int lane = threadIdx.x & 63;
    const uint32_t task = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (task >= p.c.wallCount * p.c.N) return;
    if (ld_volatile(p.control) != 0) {
        if (!EMIT && lane == 0) w.wallCount[task] = 0;
        return;
    }
    const uint32_t body = task % p.c.N, wall = task / p.c.N;
    const float4 plane = p.wallPlane[wall], planeLow = p.wallRemainder[wall];
    const f2 centre = wallCentreDistance(p, body, wall);
    const float phi = centre.hi + centre.lo;
    const float4 bd = p.body[body];
    const float lengthX = length(xyz(p.position[body]));
    const float eps64 = 64.0f * 1.1920928955078125e-7f;
    uint32_t count = 0;
    if (!(phi > bd.z + eps64 * (1.0f + lengthX + fabsf(plane.w)))) {
        const v4 q = v4of(p.rotation[body]);
        const v3 normal = qinv(q, xyz(plane)) * bd.x;  // wallClusterMayContact (solver.slang:613-623)
        const float normalLength = length(normal);
        const uint32_t shape = p.shapeId[body];
        const uint32_t samples = p.sampleInfo[shape].x;
        const uint32_t out = EMIT ? p.control[2] + w.wallOffset[task] : 0;
        // appendWallContact (solver.slang:625-638) for up to 64 nodes, one per lane,
        // ranked by ballot.
        auto appendNodes = [&](bool active, uint32_t node) {
            bool hit = false;
            v3 arm = {0, 0, 0};
            float depth = 0, area = 0;
            if (active) {
                const float4 point = p.sample[samples + node];
                arm = qrot(q, xyz(point) * bd.x);
                const float d = localWallDistance(centre, plane, planeLow, arm);
                if (!isfinite(d)) fail(p, 512);
                else if (d < 0) {
                    hit = true;
                    depth = -d;
                    area = point.w * bd.x * bd.x;
                    if (depth > p.c.maxOverlap * bd.w) fail(p, 4);
                }
            }
            const uint64_t mask = __ballot(hit);
            if (EMIT && hit) {
                const uint32_t id = out + count + __popcll(mask & lanes_below(lane));
                if (id < p.c.contactCapacity) {
                    p.meta[id] = make_uint4(body, p.c.N + wall, node, 1);
                    p.pointDepth[id] = f4(arm, depth);
                    p.normalArea[id] = make_float4(plane.x, plane.y, plane.z, area);
                }
            }
            count += __popcll(mask);
        };
        if (p.c.meshGeometry == 0) {  // every node in order (SDF geometry has no sample tree)
            const uint32_t nodes = p.sampleInfo[shape].y;
            for (uint32_t first = 0; first < nodes; first += 64) appendNodes(first + lane < nodes, first + lane);
        } else {
            const uint4 info = p.meshClusterInfo[shape];
            for (uint32_t cluster = info.x; cluster < info.x + info.y;) {
                const v3 low = xyz(p.meshClusterMin[cluster]), high = xyz(p.meshClusterMax[cluster]);
                const v3 center = 0.5f * (low + high), half = 0.5f * (high - low);
                const float support = dot(normal, center) - dot(fabs3(normal), half);
                const float pad =
                    eps64 * (1.0f + fabsf(plane.w) + lengthX + normalLength * (length(center) + length(half)));
                const bool candidate = phi + support <= pad;
                const uint4 range = p.meshClusterRanges[cluster];
                if (candidate && range.y != 0) {
                    const bool active = uint32_t(lane) < range.y;
                    appendNodes(active, active ? p.meshSampleIndices[range.x + lane] : 0u);
                }
                cluster = candidate ? cluster + 1 : range.z;
            }
        }
    }
    if (!EMIT && lane == 0) w.wallCount[task] = count;
}
// Plane contacts follow the mesh contacts: control[2] grows by their total (flag 32
// past capacity, as the serial appends); the mesh total is kept for the gather.
__global__ void k_wall_total(P p, W w) {
    if (ld_volatile(p.control) != 0) return;
    const uint32_t base = p.control[2], added = w.counters[7 * LINE];
    w.counters[8 * LINE] = base;
    uint64_t total = uint64_t(base) + added;
    if (total > p.c.contactCapacity) { fail(p, 32); total = p.c.contactCapacity; }
    p.control[2] = uint32_t(total);
    atomicAdd(w.totals + 3, (unsigned long long)added);
}

// ---------------------------------------------------------------- friction history (solver.slang:689-759)
__device__ __forceinline__ uint32_t hashKey(uint32_t size, uint4 key) {
    uint32_t h = key.x * 73856093u ^ key.y * 19349663u ^ key.z * 83492791u;
    h ^= h >> 16;
    return h & (size - 1);
}
__device__ v3 transported(v3 x, v3 a, v3 b, float spin) {
    float c = fminf(fmaxf(dot(a, b), -1.0f), 1.0f);
    if (c < -0.999999f) return {0, 0, 0};
    v3 v = cross(a, b);
    v3 out = x + cross(v, x) + cross(v, cross(v, x)) / (1.0f + c);
    float cs = dem_cosf(spin), sn = dem_sinf(spin);
    out = out * cs + cross(b, out) * sn + b * dot(b, out) * (1.0f - cs);
    return out - b * dot(out, b);
}
__global__ void k_clear_storage(P p) {
    uint32_t n = max(p.c.hashSize, p.c.N);
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < n; id += gridDim.x * blockDim.x)
        if (id < p.c.hashSize) p.hash[id] = DEM_EMPTY;
}
// Sliding contacts are counted into control[6] once per workgroup: each lane counts
// its own contacts, the workgroup sums them in LDS, and one atomic per workgroup
// adds the total.  Atomics on that one word serialise in L2, so one per wave (tens
// of thousands per launch once a fifth of the contacts slide) tripled the kernel.
__device__ __forceinline__ void addSliding(const P& p, uint32_t mine, uint32_t* total) {
    if (threadIdx.x == 0) *total = 0u;
    __syncthreads();
    for (int d = 32; d > 0; d >>= 1) mine += __shfl_xor(mine, d);
    if ((threadIdx.x & 63) == 0 && mine) atomicAdd(total, mine);
    __syncthreads();
    if (threadIdx.x == 0 && *total) atomicAdd(p.control + 6, *total);
}
// Each grain's omega() for the resolve, which would otherwise form it for both grains
// of every contact (same function, same inputs, same bits).
__global__ void k_omega_rows(P p, W w) {
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < p.c.N; id += gridDim.x * blockDim.x)
        w.omegaRow[id] = f4(omega(p, id), 0.0f);
}
__global__ __launch_bounds__(256) void k_resolve(P p, W w) {
    __shared__ uint32_t slidingTotal;
    // Every lane reaches the workgroup sum below, also after an error flag.
    const uint32_t count = ld_volatile(p.control) != 0 ? 0u : min(p.control[2], p.c.contactCapacity);
    uint32_t slidingMine = 0;
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < count; id += gridDim.x * blockDim.x) {
        bool slid = false;
        uint4 key = p.meta[id];
        uint32_t i = key.x, j = key.y;
        float4 na = p.normalArea[id];
        v3 n = xyz(na);
        const bool wall = j >= p.c.N;
        float viscosity = wall ? p.wallMaterial[j - p.c.N].w : p.c.normalViscosity;
        v3 wi = xyz(viscosity > 0 ? p.forceOmega[i] : w.omegaRow[i]);
        v3 wj = {0, 0, 0};
        if (!wall) wj = xyz(viscosity > 0 ? p.forceOmega[j] : w.omegaRow[j]);
        v3 elastic = {0, 0, 0};
        if (p.c.historyStep != 0) {
            uint32_t slot = hashKey(p.c.hashSize, key);
            for (uint32_t probe = 0; probe < p.c.hashSize; probe++) {
                uint32_t old = p.oldHash[slot];
                if (old == DEM_EMPTY) break;
                uint4 om = p.oldMeta[old];
                if (om.x == key.x && om.y == key.y && om.z == key.z) {
                    elastic = transported(xyz(p.oldSlip[old]), xyz(p.oldNormal[old]), n, 0.5f * dot(wi + wj, n) * p.c.dt);
                    break;
                }
                slot = (slot + 1) & (p.c.hashSize - 1);
            }
        }
        float4 pdv = p.pointDepth[id];
        v3 arm = xyz(pdv);
        if (wall) {  // planeViscousResponse (plane.slang:13-44), normalResponse weight 1
            const uint32_t k = j - p.c.N;
            const float4 material = p.wallMaterial[k];
            v3 v = viscosity > 0 ? xyz(p.forceVelocity[i]) : xyz(p.velocity[i]) + xyz(p.velocityRemainder[i]);
            v3 rv = v + cross(wi, arm) - xyz(p.wallVelocity[k]);
            float fn = 0, lossPower = 0;
            float4 forceEnergy = {0, 0, 0, 0}, slipLoss = {0, 0, 0, 0};
            bool sliding = false;
            if (!(-pdv.w >= 0)) {
                const float depth = -(-pdv.w), area = na.w, kn = material.x, kt = material.y, mu = material.z;
                v3 tangent = rv - n * dot(rv, n);
                v3 trial = elastic - n * dot(elastic, n) + p.c.dt * tangent;
                v3 slip = trial;
                float normalVelocity = dot(rv, n);
                float el = 1.0f * kn * area * depth;
                fn = viscosity == 0 ? el : fmaxf(0.0f, el - 1.0f * viscosity * area * normalVelocity);
                lossPower = fmaxf(0.0f, (el - fn) * normalVelocity);
                v3 ft = -kt * area * slip;
                float magnitude = length(ft);
                if (magnitude > mu * fn && magnitude > 0) {
                    sliding = true;
                    ft = ft * (mu * fn / magnitude);
                    slip = -ft / (kt * area);
                }
                forceEnergy = f4(fn * n + ft, 0.5f * area * (kn * depth * depth + kt * dot(slip, slip)));
                slipLoss = f4(slip, fmaxf(0.0f, -dot(ft, trial - slip)));
            }
            p.slip[id] = slipLoss;
            p.contactForce[id] = forceEnergy;
            p.normalLossPower[id] = lossPower;
            if (!isfinite(forceEnergy.x) || !isfinite(forceEnergy.y) || !isfinite(forceEnergy.z) ||
                !isfinite(forceEnergy.w) || !isfinite(lossPower))
                fail(p, 512);
            else
                slid = sliding;
        } else {
            v3 branch = centerBranch(p, i, j);
            v3 translation = xyz(p.velocity[i]) - xyz(p.velocity[j]);
            if (viscosity > 0) translation = xyz(p.forceVelocity[i]) - xyz(p.forceVelocity[j]);
            v3 rv = translation - xyz(p.cell[1]) * branch + cross(wi, arm) - cross(wj, arm - branch);
            elastic = elastic + (rv - n * dot(rv, n)) * p.c.dt;
            v3 trial = elastic;
            float area = na.w, depth = pdv.w;
            float fn = 0, lossPower = 0;
            float normalVelocity = dot(rv, n);
            if (depth > 0) {  // normalResponse (normal.slang:3-11), weight 0.5
                float el = 0.5f * p.c.kn * area * depth;
                fn = viscosity == 0 ? el : fmaxf(0.0f, el - 0.5f * viscosity * area * normalVelocity);
                lossPower = fmaxf(0.0f, (el - fn) * normalVelocity);
            }
            p.normalLossPower[id] = lossPower;
            v3 ft = -0.5f * p.c.kt * area * elastic;
            float magnitude = length(ft);
            if (magnitude > p.c.mu * fn && magnitude > 0) {
                ft = ft * (p.c.mu * fn / magnitude);
                elastic = -ft / (0.5f * p.c.kt * area);
                slid = true;
            }
            p.slip[id] = f4(elastic, fmaxf(0.0f, -dot(ft, trial - elastic)));
            float4 cf = f4(fn * n + ft, 0.25f * area * (p.c.kn * depth * depth + p.c.kt * dot(elastic, elastic)));
            p.contactForce[id] = cf;
            if (!isfinite(cf.x) || !isfinite(cf.y) || !isfinite(cf.z) || !isfinite(cf.w) || !isfinite(lossPower)) fail(p, 512);
        }
        slidingMine += slid ? 1u : 0u;
    }
    addSliding(p, slidingMine, &slidingTotal);
}
__global__ void k_build_hash(P p) {
    if (ld_volatile(p.control) != 0) return;
    uint32_t count = min(p.control[2], p.c.contactCapacity);
    for (uint32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < count; id += gridDim.x * blockDim.x) {
        uint32_t slot = hashKey(p.c.hashSize, p.meta[id]);
        bool placed = false;
        for (uint32_t probe = 0; probe < p.c.hashSize; probe++) {
            if (atomicCAS(p.hash + slot, DEM_EMPTY, id) == DEM_EMPTY) { placed = true; break; }
            slot = (slot + 1) & (p.c.hashSize - 1);
        }
        if (!placed) fail(p, 64);
    }
}

// ---------------------------------------------------------------- force / torque gather (solver.slang:761-811)
// One wave per grain.  The grain's contact endpoints are visited in the order of
// the CPU's linked list (descending contact id: pairs descending, pass 1 then 0,
// contacts descending), staged in LDS 64 at a time, and each of the reference's
// sequential sums runs on its own lane with the reference's exact FP32 operations:
// lanes 0-2 force (compensated), 3-5 torque (compensated), 6-14 stress rows,
// 15 energy, 16 |F| sum, 17 max depth, 18 dissipation, 19 normal power,
// 20-21 frozen-contact rates.  Sums over different quantities are independent, so
// splitting them across lanes leaves every sum's operation sequence unchanged.
constexpr int GCHUNK = 64;
constexpr int GFIELDS = 16;
constexpr int GSTRIDE = GCHUNK + 1;  // padded rows: lanes reading different fields hit different banks
constexpr int GSUMS = 22;            // reference sums per grain, one per lane
enum GField { G_FX, G_FY, G_FZ, G_E, G_AX, G_AY, G_AZ, G_DEPTH, G_SLIP, G_NLP, G_AREA, G_SRC, G_BX, G_BY, G_BZ, G_MINW };

__device__ __forceinline__ float gget(const float* st, int field, int e) { return st[field * GSTRIDE + e]; }

__global__ __launch_bounds__(256) void k_gather(P p, W w) {
    __shared__ float stage[4][GFIELDS * GSTRIDE];
    __shared__ float contrib[4][GSUMS * GSTRIDE];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t id = blockIdx.x * 4 + wv;
    if (id >= p.c.N || ld_volatile(p.control) != 0) return;
    float* st = stage[wv];
    const float mass = p.body[id].y;
    const float inverseInertia = p.c.activeTimestep ? inverseInertiaBound(p, id) : 0.0f;
    float hi = 0.0f, lo = 0.0f;  // compensated chains (lanes 0-5) or plain sums
    int fill = 0;
    // Two passes per chunk of endpoints: lane e evaluates every contribution of
    // endpoint e (the reference's expressions, term by term), then lane k folds
    // quantity k over the chunk in order -- compensated (force, torque), plain
    // (stress, energy, |F|, dissipation, power, timestep rates) or max (depth).
    float* ct = contrib[wv];
    auto process = [&](int count) {
        wave_sync();
        if (lane < count) {
            const int e = lane;
            const float kind = gget(st, G_SRC, e);  // 0 target, 1 source, 2 + k plane k
            const bool source = kind != 0.0f, wall = kind >= 2.0f;
            const v3 force = {gget(st, G_FX, e), gget(st, G_FY, e), gget(st, G_FZ, e)};
            const v3 arm = {gget(st, G_AX, e), gget(st, G_AY, e), gget(st, G_AZ, e)};
            const v3 branch = {gget(st, G_BX, e), gget(st, G_BY, e), gget(st, G_BZ, e)};
            ct[0 * GSTRIDE + e] = source ? force.x : -force.x;
            ct[1 * GSTRIDE + e] = source ? force.y : -force.y;
            ct[2 * GSTRIDE + e] = source ? force.z : -force.z;
            const v3 c = source ? cross(arm, force) : cross(arm - branch, -force);
            ct[3 * GSTRIDE + e] = c.x;
            ct[4 * GSTRIDE + e] = c.y;
            ct[5 * GSTRIDE + e] = c.z;
            // A plane contact carries its whole force and energy (no half share)
            // and its stress moment arm is the contact arm itself.
            const float fr[3] = {force.x, force.y, force.z}, bc[3] = {branch.x, branch.y, branch.z};
            const float ac[3] = {arm.x, arm.y, arm.z};
            for (int r = 0; r < 3; r++)
                for (int k = 0; k < 3; k++)
                    ct[(6 + 3 * r + k) * GSTRIDE + e] = wall ? fr[r] * ac[k] : 0.5f * fr[r] * bc[k];
            ct[15 * GSTRIDE + e] = wall ? gget(st, G_E, e) : 0.5f * gget(st, G_E, e);
            ct[16 * GSTRIDE + e] = length(force);
            ct[17 * GSTRIDE + e] = gget(st, G_DEPTH, e) / gget(st, G_MINW, e);
            ct[18 * GSTRIDE + e] = wall ? gget(st, G_SLIP, e) : 0.5f * gget(st, G_SLIP, e);
            ct[19 * GSTRIDE + e] = wall ? gget(st, G_NLP, e) : 0.5f * gget(st, G_NLP, e);
            if (p.c.activeTimestep != 0) {
                const v3 localArm = source ? arm : arm - branch;
                const float mobility = 1.0f / mass + dot(localArm, localArm) * inverseInertia;
                float4 material = make_float4(p.c.kn, p.c.kt, p.c.mu, p.c.normalViscosity);
                if (wall) material = p.wallMaterial[uint32_t(kind) - 2];
                const float area = gget(st, G_AREA, e) * (wall ? 1.0f : 0.5f);
                const float coupling = wall ? 1.0f : 2.0f;
                ct[20 * GSTRIDE + e] = coupling * area * (material.x * (1.0f + material.z) + material.y) * mobility;
                ct[21 * GSTRIDE + e] = coupling * area * material.w * (1.0f + material.z) * mobility;
            }
        }
        wave_sync();
        const float* row = ct + lane * GSTRIDE;
        if (lane < 6) {
            for (int e = 0; e < count; e++) {
                v3 h = {hi, 0, 0}, l = {lo, 0, 0};
                compensated3(h, l, {row[e], 0, 0});
                hi = h.x; lo = l.x;
            }
        } else if (lane < 15) {
            for (int e = 0; e < count; e++) hi = hi - row[e];
        } else if (lane == 17) {
            for (int e = 0; e < count; e++) hi = fmaxf(hi, row[e]);
        } else if (lane < 20 || (lane < 22 && p.c.activeTimestep != 0)) {
            for (int e = 0; e < count; e++) hi += row[e];
        }
        wave_sync();
    };
    // Plane contacts carry the highest ids (after every mesh contact): the list
    // starts with them, planes descending, each plane's contacts descending.
    const uint32_t wallBase = p.c.wallCount ? w.counters[8 * LINE] : 0;
    for (uint32_t k = p.c.wallCount; k-- > 0;) {
        const uint32_t task = k * p.c.N + id;
        const uint32_t cnt = w.wallCount[task];
        const uint32_t off = wallBase + w.wallOffset[task];
        const float kind = 2.0f + float(k), minW = p.body[id].w;
        for (uint32_t done = 0; done < cnt;) {
            uint32_t m = min(cnt - done, uint32_t(GCHUNK - fill));
            if (uint32_t(lane) < m) {
                uint32_t contact = off + cnt - 1 - (done + lane);
                float4 cf = p.contactForce[contact];
                float4 pdv = p.pointDepth[contact];
                int e = fill + lane;
                st[G_FX * GSTRIDE + e] = cf.x; st[G_FY * GSTRIDE + e] = cf.y; st[G_FZ * GSTRIDE + e] = cf.z;
                st[G_E * GSTRIDE + e] = cf.w;
                st[G_AX * GSTRIDE + e] = pdv.x; st[G_AY * GSTRIDE + e] = pdv.y; st[G_AZ * GSTRIDE + e] = pdv.z;
                st[G_DEPTH * GSTRIDE + e] = pdv.w;
                st[G_SLIP * GSTRIDE + e] = p.slip[contact].w;
                st[G_NLP * GSTRIDE + e] = p.normalLossPower[contact];
                st[G_AREA * GSTRIDE + e] = p.c.activeTimestep ? p.normalArea[contact].w : 0.0f;
                st[G_SRC * GSTRIDE + e] = kind;
                st[G_BX * GSTRIDE + e] = 0.0f; st[G_BY * GSTRIDE + e] = 0.0f; st[G_BZ * GSTRIDE + e] = 0.0f;
                st[G_MINW * GSTRIDE + e] = minW;
            }
            fill += int(m);
            done += m;
            if (fill == GCHUNK) { process(fill); fill = 0; }
        }
    }
    // The grain's pair endpoints: the tasks of its adjacency list in list order (pairs
    // descending, pass 1 then pass 0), 64 per round, one lane each, so a round reads
    // every task's count, offset, branch and radii together; then each lane of a chunk
    // finds its endpoint's task by a search over the running counts.  The endpoint
    // sequence, and so every sum, is the one of a task-by-task walk.
    const uint32_t first = w.adjOffset[id], tasksHere = 2 * w.adjCount[id];
    for (uint32_t s0 = 0; s0 < tasksHere; s0 += WAVE) {
        const uint32_t s = s0 + uint32_t(lane);
        uint32_t cnt = 0, off = 0;
        float source = 0.0f, minW = 0.0f;
        v3 branch = {0.0f, 0.0f, 0.0f};
        if (s < tasksHere) {
            const uint32_t k = tasksHere / 2 - 1 - s / 2, pass = 1 - (s & 1);
            const uint32_t pair = w.adjList[first + k];
            const uint2 ij = p.pairs[pair];
            const uint32_t task = 2 * pair + pass;
            cnt = w.taskCount[task];
            if (cnt != 0) {
                const uint32_t src = pass == 0 ? ij.x : ij.y, dst = pass == 0 ? ij.y : ij.x;
                source = id == src ? 1.0f : 0.0f;
                branch = centerBranch(p, src, dst);
                minW = fminf(p.body[src].w, p.body[dst].w);
                off = w.taskOffset[task];
            }
        }
        uint32_t incl = cnt;  // running count through this lane's task
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d);
            if (lane >= d) incl += o;
        }
        const uint32_t total = __shfl(incl, WAVE - 1);
        for (uint32_t e0 = 0; e0 < total;) {
            const uint32_t m = min(total - e0, uint32_t(GCHUNK - fill));
            const uint32_t e = e0 + uint32_t(lane);
            uint32_t t = 0;  // the first task lane whose running count exceeds e
            for (uint32_t step = WAVE / 2; step > 0; step >>= 1)
                if (__shfl(incl, int(t + step - 1)) <= e) t += step;
            const uint32_t tIncl = __shfl(incl, int(t)), tCnt = __shfl(cnt, int(t)), tOff = __shfl(off, int(t));
            const float tSource = __shfl(source, int(t)), tMinW = __shfl(minW, int(t));
            const v3 tBranch = mk3(__shfl(branch.x, int(t)), __shfl(branch.y, int(t)), __shfl(branch.z, int(t)));
            if (uint32_t(lane) < m) {
                const uint32_t i = e - (tIncl - tCnt);  // the endpoint's place in its task, from its first
                const uint32_t contact = tOff + tCnt - 1 - i;
                float4 cf = p.contactForce[contact];
                float4 pdv = p.pointDepth[contact];
                int at = fill + lane;
                st[G_FX * GSTRIDE + at] = cf.x; st[G_FY * GSTRIDE + at] = cf.y; st[G_FZ * GSTRIDE + at] = cf.z;
                st[G_E * GSTRIDE + at] = cf.w;
                st[G_AX * GSTRIDE + at] = pdv.x; st[G_AY * GSTRIDE + at] = pdv.y; st[G_AZ * GSTRIDE + at] = pdv.z;
                st[G_DEPTH * GSTRIDE + at] = pdv.w;
                st[G_SLIP * GSTRIDE + at] = p.slip[contact].w;
                st[G_NLP * GSTRIDE + at] = p.normalLossPower[contact];
                st[G_AREA * GSTRIDE + at] = p.c.activeTimestep ? p.normalArea[contact].w : 0.0f;
                st[G_SRC * GSTRIDE + at] = tSource;
                st[G_BX * GSTRIDE + at] = tBranch.x; st[G_BY * GSTRIDE + at] = tBranch.y; st[G_BZ * GSTRIDE + at] = tBranch.z;
                st[G_MINW * GSTRIDE + at] = tMinW;
            }
            fill += int(m);
            e0 += m;
            if (fill == GCHUNK) { process(fill); fill = 0; }
        }
    }
    if (fill) process(fill);
    // Collect the chain results (compensated chains round hi + lo once, as the reference).
    float total = lane < 6 ? hi + lo : hi;
    float fx = __shfl(total, 0), fy = __shfl(total, 1), fz = __shfl(total, 2);
    float tx = __shfl(total, 3), ty = __shfl(total, 4), tz = __shfl(total, 5);
    float sv[9];
    for (int k = 0; k < 9; k++) sv[k] = __shfl(total, 6 + k);
    float energy = __shfl(total, 15), denom = __shfl(total, 16), maxDepth = __shfl(total, 17);
    float dissipation = __shfl(total, 18), normalPower = __shfl(total, 19);
    float stiffnessRate = __shfl(total, 20), viscousRate = __shfl(total, 21);
    if (lane != 0) return;
    v3 f = {fx, fy, fz}, t = {tx, ty, tz};
    p.timestepRows[id] = make_float4(p.c.activeTimestep ? stiffnessRate : 0.0f, p.c.activeTimestep ? viscousRate : 0.0f, 0, 0);
    float4 normalState = p.normalDamping[id];
    normalState.z = normalPower;
    p.normalDamping[id] = normalState;
    p.force[id] = f4(f, maxDepth);
    p.torque[id] = f4(t, dissipation);
    p.particleStress[3 * id] = make_float4(sv[0], sv[1], sv[2], 0);
    p.particleStress[3 * id + 1] = make_float4(sv[3], sv[4], sv[5], 0);
    p.particleStress[3 * id + 2] = make_float4(sv[6], sv[7], sv[8], 0);
    v3 vel = xyz(p.velocity[id]);
    float kinetic = 0.5f * (mass * dot(vel, vel) + dot(xyz(p.momentum[id]), omega(p, id)));
    p.particleDiagnostic[id] = make_float4(kinetic, energy, length(f), denom);
    accumulateRow(p, id, {dissipation, 0, 0, 0});
}

// gatherWallRows (solver.slang:837-856): one thread per (wall, grain) task sums
// that plane's contacts on the grain in list order (descending id): reaction
// -F, moment -(x + arm) x F about the origin, power F . v_wall, the contact count
// (stored as the bits of row 0's w) and the accumulated work.
__global__ void k_gather_wall_rows(P p, W w) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= p.c.N * p.c.wallCount || ld_volatile(p.control) != 0) return;
    const uint32_t body = index % p.c.N, wall = index / p.c.N;
    const uint32_t cnt = w.wallCount[index], off = w.counters[8 * LINE] + w.wallOffset[index];
    const v3 x = xyz(p.position[body]) + xyz(p.positionRemainder[body]);
    const v3 velocity = xyz(p.wallVelocity[wall]);
    v3 reaction = {0, 0, 0}, moment = {0, 0, 0};
    float power = 0;
    for (uint32_t k = cnt; k-- > 0;) {
        const uint32_t contact = off + k;
        const v3 force = xyz(p.contactForce[contact]);
        const v3 point = x + xyz(p.pointDepth[contact]);
        reaction = reaction - force;
        moment = moment - cross(point, force);
        power += dot(force, velocity);
    }
    const float4 work = p.wallAccumulated[index];
    p.wallRows[3 * index] = f4(reaction, __uint_as_float(cnt));
    p.wallRows[3 * index + 1] = f4(moment, power);
    p.wallRows[3 * index + 2] = make_float4(work.x + work.y, 0, 0, 0);
}
// refreshWallPower (solver.slang:950-955): wall power after a velocity change,
// from the stored reaction.
__global__ void k_refresh_wall_power(P p) {
    const uint32_t index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= p.c.N * p.c.wallCount) return;
    float4 value = p.wallRows[3 * index + 1];
    value.w = -dot(xyz(p.wallRows[3 * index]), xyz(p.wallVelocity[index / p.c.N]));
    p.wallRows[3 * index + 1] = value;
}

// ---------------------------------------------------------------- evidence rows (solver.slang:922-947)
__global__ void k_diagnostic_rows(P p) {  // :922-936
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.N) return;
    v3 v = xyz(p.velocity[id]);
    float m = p.body[id].y;
    float kinetic = 0.5f * (m * dot(v, v) + dot(xyz(p.momentum[id]), omega(p, id)));
    float4* row = p.diagnosticRows + 8 * size_t(id);
    row[0] = f4(xyz(p.particleStress[3 * id]), kinetic);
    row[1] = f4(xyz(p.particleStress[3 * id + 1]), p.particleDiagnostic[id].y);
    row[2] = f4(xyz(p.particleStress[3 * id + 2]), length(xyz(p.force[id])));
    row[3] = f4(m * v.x * v, p.particleDiagnostic[id].w);
    row[4] = f4(m * v.y * v, p.force[id].w);
    row[5] = f4(m * v.z * v, p.torque[id].w);
    row[6] = p.accumulated[id];
    float4 normalState = p.normalDamping[id];
    row[7] = make_float4(normalState.x + normalState.y, normalState.z, 0, 0);
}
__global__ void k_fabric_rows(P p) {  // :938-947
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= p.c.pairCapacity) return;
    float4 a = make_float4(0, 0, 0, 0), b = a;
    if (id < p.control[1] && p.contactCounts[id] > 0) {
        uint2 ij = p.pairs[id];
        v3 n = normalize3(centerBranch(p, ij.x, ij.y));
        a = f4(n * n, n.x * n.y);
        b = make_float4(n.x * n.z, n.y * n.z, 1, 0);
    }
    p.fabricRows[2 * size_t(id)] = a;
    p.fabricRows[2 * size_t(id) + 1] = b;
}

// Group totals of evidence rows (diagnostic.slang).  One 256-lane group per 256
// records; the LDS tree halves 128 -> 1 pairwise in exactly that order, so the FP32
// totals are reproducible.  MODE 0: 8 rows per grain, row 4 .w keeps the maximum;
// MODE 1: 2 rows per pair slot; MODE 2: 3 rows per (wall, grain) with blockIdx.y the
// wall, row 0 .w an integer count carried as FP32 bits.
template <int MODE>
__global__ __launch_bounds__(256) void k_reduce_rows(uint32_t n, uint32_t outputCount, const float4* in, float4* out) {
    constexpr int ROWS = MODE == 0 ? 8 : (MODE == 1 ? 2 : 3);
    __shared__ float4 values[256];
    __shared__ uint32_t counts[256];
    const uint32_t lane = threadIdx.x, id = blockIdx.x * 256 + lane;
    const size_t base = size_t(blockIdx.y) * n;
    for (int row = 0; row < ROWS; row++) {
        float4 value = make_float4(0, 0, 0, 0);
        if (id < n) value = in[ROWS * (base + id) + row];
        if (MODE == 2 && row == 0) {
            counts[lane] = __float_as_uint(value.w);
            value.w = 0;
        }
        values[lane] = value;
        __syncthreads();
        for (uint32_t step = 128; step > 0; step >>= 1) {
            if (lane < step) {
                float4 a = values[lane], b = values[lane + step];
                float4 sum = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
                if (MODE == 0 && row == 4) sum.w = (a.w > b.w || a.w != a.w) ? a.w : b.w;
                if (MODE == 2 && row == 0) counts[lane] += counts[lane + step];
                values[lane] = sum;
            }
            __syncthreads();
        }
        if (lane == 0) {
            float4 total = values[0];
            if (MODE == 2 && row == 0) total.w = __uint_as_float(counts[0]);
            out[ROWS * (size_t(blockIdx.y) * outputCount + blockIdx.x) + row] = total;
        }
        __syncthreads();
    }
}

// Per template: the largest sample-node count (sizes the temporal rows), the first
// triangle and triangle count of its BVH (leftmost and rightmost leaves of the
// preorder tree; an internal node's children are n + 1 and that child's escape
// link), its first sample key, and the largest triangle and cluster counts.
__global__ void k_template_extent(P p, uint32_t* out, uint32_t* triBase) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.c.shapes || t >= p.templates) return;
    atomicMax(out, p.sampleInfo[t].y);
    const uint32_t root = p.meshRoots[t].x;
    uint32_t n = root;
    for (int d = 0; d < 64 && p.meshLinks[n].z == 0; d++) n = n + 1;
    const uint32_t first = p.meshLinks[n].y;
    n = root;
    for (int d = 0; d < 64 && p.meshLinks[n].z == 0; d++) n = p.meshLinks[n + 1].x;
    const uint32_t count = p.meshLinks[n].y + p.meshLinks[n].z - first;
    // The template's sample keys start at its first leaf (leaves are laid out in
    // preorder): a node's slot is its key position minus this base.
    const uint4 info = p.meshClusterInfo[t];
    uint32_t c = info.x;
    while (c + 1 < info.x + info.y && p.meshClusterRanges[c].y == 0) c++;
    triBase[3 * t] = first;
    triBase[3 * t + 1] = count;
    triBase[3 * t + 2] = p.meshClusterRanges[c].x;
    atomicMax(out + 1, count);
    atomicMax(out + 2, info.y);
}
__global__ void k_set_word(uint32_t* word, uint32_t value) { *word = value; }

// ---------------------------------------------------------------- scan.slang (the reference Engine's Scan)
// blockScan: exclusive uint scan of each 256-element block, block totals out
// (scan.slang:1-21, unsigned wrap-around as on the GPU); addBlocks: add the scanned
// block totals (scan.slang:22-28).  Wave64 prefix via DPP-free shuffles, then the
// four wave totals through LDS.
__global__ __launch_bounds__(256) void k_block_scan(uint32_t n, const uint32_t* in, uint32_t* out, uint32_t* totals) {
    __shared__ uint32_t waves[4];
    const uint32_t id = blockIdx.x * 256 + threadIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t v = id < n ? in[id] : 0u;
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d);
        if (lane >= uint32_t(d)) x += y;
    }
    if (lane == 63) waves[wv] = x;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < wv; k++) base += waves[k];
    x += base;
    if (id < n) out[id] = x - v;
    if (threadIdx.x == 255) totals[blockIdx.x] = x;
}
__global__ void k_add_blocks(uint32_t n, uint32_t* out, const uint32_t* offsets) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) out[id] += offsets[id / 256];
}

// ---------------------------------------------------------------- learned-cull training data
// Counter-based uniforms (splitmix64 finaliser): reproducible per (seed, template, point, draw).
__device__ __forceinline__ uint64_t nn_mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float nn_unit(uint64_t bits) { return float(uint32_t(bits >> 40)) * (1.0f / 16777216.0f); }

struct NNSampleArgs {
    const float4* triangles;   // per template: 3 rows per triangle, normalized units
    const float* cdf;          // inclusive area prefix per triangle, last = 1 per template
    const uint32_t* triStart;  // first triangle of each template
    const uint32_t* triCount;
    const float4* box;         // per template: (lo, diameter), (hi, label scale)
    float4* points;            // [template][count]
    float* labels;
    uint32_t count;
    uint64_t seed;
};
// One candidate per point, redrawn until it passes the hull-group test the
// collection applies before any exact query (at most 64 draws): even points lie
// within one grain diameter of the surface (offset d s^3 along the face normal,
// s ~ U(-1, 1)), odd points are uniform in the bounding box padded by 10%.  The
// label is the exact BVH signed distance, clamped to +-1 label units.
__global__ void k_nn_sample(P p, NNSampleArgs a) {
    __shared__ uint16_t stacks[4][STACK * WAVE];
    const uint32_t t = blockIdx.y, id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= a.count) return;
    uint16_t* stack = &stacks[threadIdx.x >> 6][threadIdx.x & 63];
    const float4 b0 = a.box[2 * t], b1 = a.box[2 * t + 1];
    const v3 lo = xyz(b0), hi = xyz(b1), pad = 0.1f * (hi - lo);
    const uint32_t first = a.triStart[t], n = a.triCount[t];
    v3 q = {0, 0, 0};
    for (uint32_t draw = 0; draw < 64; draw++) {
        uint64_t base = nn_mix(a.seed ^ nn_mix((uint64_t(t) << 40) ^ (uint64_t(id) << 8) ^ draw));
        float u0 = nn_unit(base), u1 = nn_unit(nn_mix(base + 1)), u2 = nn_unit(nn_mix(base + 2)),
              u3 = nn_unit(nn_mix(base + 3));
        if ((id & 1) == 0) {
            uint32_t l = 0, r = n - 1;  // first triangle whose inclusive prefix exceeds u0
            while (l < r) {
                uint32_t m = (l + r) >> 1;
                if (a.cdf[first + m] > u0) r = m; else l = m + 1;
            }
            const float4* tri = a.triangles + 3 * size_t(first + l);
            v3 A = xyz(tri[0]), B = xyz(tri[1]), C = xyz(tri[2]);
            float sq = sqrtf(u1);
            v3 x = (1.0f - sq) * A + (sq * (1.0f - u2)) * B + (sq * u2) * C;
            v3 nrm = cross(B - A, C - A);
            nrm = nrm / length(nrm);
            float s = 2.0f * u3 - 1.0f;
            q = x + (b0.w * s * s * s) * nrm;
        } else {
            q = mk3(lo.x - pad.x + u0 * (hi.x - lo.x + 2 * pad.x), lo.y - pad.y + u1 * (hi.y - lo.y + 2 * pad.y),
                    lo.z - pad.z + u2 * (hi.z - lo.z + 2 * pad.z));
        }
        if (meshMayContain(p, t, q)) break;
    }
    const size_t out = size_t(t) * a.count + id;
    a.points[out] = f4(q, 0.0f);
    float d = meshDistance(p, t, q, stack).w / b1.w;
    a.labels[out] = fminf(fmaxf(d, -1.0f), 1.0f);
}

// ---------------------------------------------------------------- probes
__global__ void k_mesh_probe(P p, uint32_t count, uint32_t shape, const float4* points, float4* out) {
    __shared__ uint16_t stacks[4][STACK * WAVE];
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    uint16_t* stack = &stacks[threadIdx.x >> 6][threadIdx.x & 63];
    if (id < count) out[id] = f4(meshDistance(p, shape, xyz(points[id]), stack));
}
__global__ void k_cull_probe(P p, uint32_t count, uint32_t shape, const float4* points, uint32_t* out) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id < count) out[id] = meshMayContain(p, shape, xyz(points[id])) ? 1 : 0;
}
__global__ void k_cull_fast_probe(P p, W w, uint32_t count, uint32_t shape, const float4* points, uint32_t* out) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id < count) out[id] = meshMayContainFast(p, w, shape, xyz(points[id])) ? 1 : 0;
}
__global__ void k_cluster_fast_probe(P p, W w, uint32_t count, uint32_t src, uint32_t dst, uint32_t* out) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) return;
    v3 branch = centerBranch(p, src, dst);
    out[id] = meshClusterMayContactFast(p, w, v4of(p.rotation[src]), v4of(p.rotation[dst]), p.body[src].x,
                                        p.body[dst].x, branch, p.shapeId[dst],
                                        p.meshClusterInfo[p.shapeId[src]].x + id) ? 1 : 0;
}
__global__ void k_cluster_probe(P p, uint32_t count, uint32_t src, uint32_t dst, uint32_t* out) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) return;
    v3 branch = centerBranch(p, src, dst);
    out[id] = meshClusterMayContact(p, v4of(p.rotation[src]), v4of(p.rotation[dst]), p.body[src].x, p.body[dst].x, branch,
                                    p.shapeId[dst], p.meshClusterInfo[p.shapeId[src]].x + id) ? 1 : 0;
}
__global__ void k_pair_bounds_probe(P p, uint32_t count, uint32_t direction, uint32_t first, float4* out) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) return;
    uint32_t cluster = first + id;
    v3 center = 0.5f * (xyz(p.meshClusterMin[cluster]) + xyz(p.meshClusterMax[cluster]));
    v3 half = 0.5f * (xyz(p.meshClusterMax[cluster]) - xyz(p.meshClusterMin[cluster]));
    float c[3], e[3];
    for (int axis = 0; axis < 3; axis++) {
        float4 row = p.meshPairTransform[4 * direction + axis];
        c[axis] = dot(xyz(row), center) + row.w;
        e[axis] = dot(fabs3(xyz(row)), half);
    }
    v3 low = mk3(c[0], c[1], c[2]) - mk3(e[0], e[1], e[2]), high = mk3(c[0], c[1], c[2]) + mk3(e[0], e[1], e[2]);
    float pad = 256.0f * DEM_EPS * (1.0f + maxabs(low) + maxabs(high) + p.meshPairTransform[4 * direction + 3].x);
    out[2 * id] = f4(low - mk3(pad, pad, pad), 0);
    out[2 * id + 1] = f4(high + mk3(pad, pad, pad), 0);
}

// ================================================================= host side
thread_local char g_error[256] = "no error";

// Kernel identities for the per-kernel HIP-event profile.
#define SAD_KERNELS(X) X(begin_step) X(advance_cell) X(check_rebuild) X(clear_bins) X(fill_bins) X(sort_bins) \
    X(count_pairs) X(scan_reduce) X(scan_sums) X(scan_apply) X(emit_pairs) X(adj_clear) X(adj_count) X(adj_fill) \
    X(adj_sort) X(reset_mesh) X(prepare_pairs) X(collect) X(contact_total) X(scatter) X(clear_storage) X(resolve) \
    X(build_hash) X(gather) X(finish_step) X(capture_kinematics) X(check_timestep) \
    X(wide_nodes) X(bucket_reset) X(bucket_count) X(bucket_offsets) X(bucket_fill) \
    X(advance_walls) X(wall_count) X(wall_emit) X(wall_total) X(gather_wall_rows) X(refresh_wall_power) \
    X(diagnostic_rows) X(fabric_rows) X(sdf_count) X(sdf_total) X(sdf_emit) X(rows) X(derived_check) \
    X(pack_geometry) X(task_far) X(omega_rows)
enum KernelId {
#define SAD_ID(name) KID_##name,
    SAD_KERNELS(SAD_ID)
#undef SAD_ID
    KID_COUNT
};
const char* const kKernelNames[] = {
#define SAD_NAME(name) "k_" #name,
    SAD_KERNELS(SAD_NAME)
#undef SAD_NAME
};

int hip_fail(hipError_t e, const char* where) {
    std::snprintf(g_error, sizeof(g_error), "%s: %s", where, hipGetErrorString(e));
    return SAD_E_HIP;
}
#define HIP_OK(call)                                          \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return hip_fail(_e, #call);     \
    } while (0)

struct Scratch {
    W w{};
    uint64_t stats[8] = {};
    size_t bytes = 0;
    int device = -1;
    int cus = 0;
    unsigned long long* totals = nullptr;  // device: steps, directed pairs, queried nodes, contacts
    // HIP-event profile: pending (kernel, start, stop) triples resolved on read.
    bool profiling = false;
    struct Mark { int kernel; hipEvent_t start, stop; };
    std::vector<Mark> pending;
    std::vector<hipEvent_t> pool;
    double ms[KID_COUNT] = {};
    uint64_t launches[KID_COUNT] = {};
};

template <class T>
int grow(T*& ptr, uint32_t& cap, size_t need, size_t& bytes) {
    if (ptr && cap >= need) return SAD_OK;
    if (ptr) { (void)hipFree(ptr); bytes -= size_t(cap) * sizeof(T); ptr = nullptr; cap = 0; }
    size_t n = std::max<size_t>(need, 1);
    HIP_OK(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)));
    cap = uint32_t(n);
    bytes += n * sizeof(T);
    return SAD_OK;
}

struct Caps {
    uint32_t counters = 0, taskCount = 0, taskOffset = 0, adjCount = 0, adjOffset = 0, adjCursor = 0, adjList = 0;
    uint32_t sums = 0, stageKey = 0, stageMeta = 0, stagePd = 0, stageNa = 0, xcd = 0, xcdTasks = 0, wide = 0;
    uint32_t stageCount = 0, keyCount = 0, keyStart = 0, wallCount = 0, wallOffset = 0, sdfReach = 0, sdfHits = 0;
    uint32_t sdfLeaves = 0, sdfOrder = 0, builtKey = 0, triPacked = 0, sampleKeyed = 0, clusterPacked = 0;
    uint32_t templateRec = 0, omegaRow = 0, clusterSphere = 0;
};

struct ScratchImpl : Scratch {
    Caps caps;
    // Tunables of the collection launch (sad_hip_tune); none changes results.
    int blocksPerCU = SAD_COLLECT_BLOCKS_PER_CU;
    uint32_t pop = POP;
    int queue = QCAP;
    uint32_t leafBatch = 32;
    uint32_t batchTasks = 16;
    uint32_t batchFill = 96;
    int gapReachMicro = 300;  // best of 0-10000 over 1,000-step runs at dense; 1000 is best at uneven
    // Temporal node skip (sad_hip_tune "temporal_skip", "skip_limit_micro"): off by
    // default; the caller that enables it calls sad_hip_invalidate whenever it
    // rewrites template geometry, body scales or shape ids in place.
    int skip = 0;
    int skipLimitMicro = 20000;  // re-anchor distance: best of 2000-40000 over 2,000-step runs at uneven
    int rebase = 0;
    uint32_t epoch = 1, epochWritten = 0;  // anchors carry the epoch they were taken in
    std::vector<uint64_t> lastTable;        // binding table of the previous call
    const void* maxNodesFor = nullptr;
    uint32_t maxNodesShapes = 0, maxNodes = 0;
    size_t nodeDistCap = 0;
    uint32_t anchorCap = 0, taskFarCap = 0;
    size_t taskFrameCap = 0;
    uint32_t* probeWord = nullptr;
    uint32_t* triBase = nullptr;
    uint32_t triBaseCap = 0, maxTriangles = 0, maxClusters = 0;
    size_t leafDistCap = 0, leafPlaneCap = 0;
    // Unchanged-winner shortcut rows (sad_hip_tune "gap_shortcut", default on).
    int gapShortcut = 1;
    uint32_t* gapRows = nullptr;
    float4* lastFrames = nullptr;
    size_t gapCap = 0;
    uint32_t lastFrameCap = 0;
    uint32_t rowMapPairs = 0;  // pair capacity of the persistent-row map
    const void* triBaseFor = nullptr;
    // Captured timesteps keyed by everything a step's launches bake in.
    struct Graph { std::vector<uint64_t> key; hipGraphExec_t exec; };
    std::vector<Graph> graphs;
    hipStream_t captureStream = nullptr;
    void drop_graphs() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    int ensure_nodes(uint32_t nodes) { return grow(w.wide, caps.wide, 4 * size_t(std::max(nodes, 1u)), bytes); }
    int ensure_packed(const P& p) {
        int rc = grow(w.triPacked, caps.triPacked, 3 * size_t(std::max(p.triangles, 1u)), bytes);
        if (rc || (rc = grow(w.sampleKeyed, caps.sampleKeyed, std::max(p.sampleKeys, 1u), bytes))) return rc;
        if ((rc = grow(w.clusterPacked, caps.clusterPacked, std::max(p.clusters, 1u), bytes))) return rc;
        if ((rc = grow(w.clusterSphere, caps.clusterSphere, std::max(p.clusters, 1u), bytes))) return rc;
        return grow(w.templateRec, caps.templateRec, TREC * size_t(std::max(p.c.shapes, 1u)), bytes);
    }
    // SDF geometry: per-template reach and the hit-bitmask rows, refreshed whenever
    // the geometry epoch changes (an upload, or new bindings); outside any capture.
    uint32_t reachEpoch = 0, reachShapes = 0;
    int ensure_sdf(P& p, hipStream_t s) {
        if (p.c.meshGeometry) return SAD_OK;
        int rc = grow(w.sdfReach, caps.sdfReach, std::max(p.c.shapes, 1u), bytes);
        if (rc) return rc;
        const uint32_t shapes = std::max(p.c.shapes, 1u);
        if ((rc = grow(w.sdfLeaves, caps.sdfLeaves, size_t(64) * shapes, bytes))) return rc;
        if ((rc = grow(w.sdfOrder, caps.sdfOrder, size_t(32 * SDF_MASK_WORDS) * shapes, bytes))) return rc;
        if (reachEpoch != epoch || reachShapes != p.c.shapes) {
            if (!probeWord) HIP_OK(hipMalloc(reinterpret_cast<void**>(&probeWord), 3 * sizeof(uint32_t)));
            HIP_OK(hipMemsetAsync(probeWord, 0, sizeof(uint32_t), s));
            hipLaunchKernelGGL(k_sdf_reach, dim3(std::max(p.c.shapes, 1u)), dim3(256), 0, s, p, w, probeWord);
            HIP_OK(hipGetLastError());
            uint32_t largest = 0;
            HIP_OK(hipMemcpyAsync(&largest, probeWord, sizeof(largest), hipMemcpyDeviceToHost, s));
            HIP_OK(hipStreamSynchronize(s));
            if (largest > 32 * SDF_MASK_WORDS) {
                std::snprintf(g_error, sizeof(g_error), "SDF geometry: a template has %u nodes (at most %u)", largest,
                              32 * SDF_MASK_WORDS);
                return SAD_E_UNSUPPORTED;
            }
            w.sdfWords = std::max(2u, (largest + 63) / 64 * 2);
            hipLaunchKernelGGL(k_sdf_index, dim3(shapes), dim3(256), 0, s, p, w);
            HIP_OK(hipGetLastError());
            reachEpoch = epoch;
            reachShapes = p.c.shapes;
        }
        return grow(w.sdfHits, caps.sdfHits, size_t(2) * std::max(p.c.pairCapacity, 1u) * w.sdfWords, bytes);
    }
    // A new allocation of any persistent pair-derived array (adjacency, buckets, the
    // record of what they were built from) starts a new generation: the device
    // header no longer matches, so the next collection rebuilds them.
    const void* derivedPtrs[9] = {};
    void note_derived() {
        const void* now[9] = {w.adjCount, w.adjOffset, w.adjCursor, w.adjList, w.xcd,
                              w.xcdTasks, w.keyCount, w.keyStart, w.builtKey};
        if (std::memcmp(now, derivedPtrs, sizeof(now)) != 0) {
            std::memcpy(derivedPtrs, now, sizeof(now));
            w.builtGen++;
        }
    }
    int ensure_keys(uint32_t templates) {
        const uint32_t keys = bucketShape(templates).keys();
        int rc = grow(w.keyCount, caps.keyCount, keys, bytes);
        if (rc || (rc = grow(w.keyStart, caps.keyStart, keys, bytes))) return rc;
        note_derived();
        const size_t sums = (size_t(keys) + SCAN_TILE - 1) / SCAN_TILE + 1;
        if (sums > caps.sums && (rc = grow(w.sums, caps.sums, sums, bytes))) return rc;
        w.sumsCap = caps.sums;
        return SAD_OK;
    }
    int ensure(const Cfg& c) {
        size_t tasks = size_t(2) * c.pairCapacity;
        int rc;
        if (!w.counters) {
            if ((rc = grow(w.counters, caps.counters, 16 * LINE, bytes))) return rc;
            HIP_OK(hipMemset(w.counters, 0, 16 * LINE * sizeof(uint32_t)));
        }
        if ((rc = grow(w.stageCount, caps.stageCount, 8 * LINE, bytes))) return rc;
        if (!w.totals) {
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.totals), 8 * sizeof(unsigned long long)));
            HIP_OK(hipMemset(w.totals, 0, 8 * sizeof(unsigned long long)));
            totals = w.totals;
        }
        if ((rc = grow(w.taskCount, caps.taskCount, tasks, bytes))) return rc;
        if ((rc = grow(w.taskOffset, caps.taskOffset, tasks + 1, bytes))) return rc;
        if ((rc = grow(w.adjCount, caps.adjCount, c.N, bytes))) return rc;
        if ((rc = grow(w.omegaRow, caps.omegaRow, c.N, bytes))) return rc;
        if ((rc = grow(w.adjOffset, caps.adjOffset, c.N, bytes))) return rc;
        if ((rc = grow(w.adjCursor, caps.adjCursor, c.N, bytes))) return rc;
        if ((rc = grow(w.adjList, caps.adjList, tasks, bytes))) return rc;
        const size_t wallTasks = size_t(c.wallCount) * c.N;
        if (wallTasks > 0xffffffffull) {
            std::snprintf(g_error, sizeof(g_error), "too many (wall, grain) pairs");
            return SAD_E_ARGUMENT;
        }
        if (caps.wallCount < std::max<size_t>(wallTasks, 1)) {  // zeroed: no plane contacts before a collection
            if ((rc = grow(w.wallCount, caps.wallCount, wallTasks, bytes))) return rc;
            HIP_OK(hipMemset(w.wallCount, 0, size_t(caps.wallCount) * sizeof(uint32_t)));
        }
        if ((rc = grow(w.wallOffset, caps.wallOffset, wallTasks, bytes))) return rc;
        w.wallCap = uint32_t(std::max<size_t>(wallTasks, 1));
        size_t sums = (std::max<size_t>(std::max<size_t>(tasks + 1, c.N), wallTasks) + SCAN_TILE - 1) / SCAN_TILE + 1;
        if ((rc = grow(w.sums, caps.sums, sums, bytes))) return rc;
        if ((rc = grow(w.stageKey, caps.stageKey, size_t(8) * c.contactCapacity, bytes))) return rc;
        if ((rc = grow(w.stageMeta, caps.stageMeta, size_t(8) * c.contactCapacity, bytes))) return rc;
        if ((rc = grow(w.stagePd, caps.stagePd, size_t(8) * c.contactCapacity, bytes))) return rc;
        if ((rc = grow(w.stageNa, caps.stageNa, size_t(8) * c.contactCapacity, bytes))) return rc;
        if ((rc = grow(w.xcd, caps.xcd, 24 * LINE, bytes))) return rc;
        if ((rc = grow(w.xcdTasks, caps.xcdTasks, tasks, bytes))) return rc;
        if ((rc = grow(w.builtKey, caps.builtKey, c.pairCapacity, bytes))) return rc;
        note_derived();

        w.stageCap = c.contactCapacity;
        w.taskCap = uint32_t(tasks);
        w.nCap = c.N;
        w.sumsCap = caps.sums;
        if (device < 0) {
            HIP_OK(hipGetDevice(&device));
            HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        }
        return SAD_OK;
    }
    // Binding-table identity and the temporal-skip workspace.  Called outside any
    // graph capture (it may synchronize once per new template set).
    int ensure_skip(const P& p, void* const* table, const uint64_t* sizes) {
        // Identity of the geometry the stored distances refer to: the template,
        // body and shape buffers (the history sides alternate every step and are
        // not part of it) and the grain state's address.
        static const int slots[] = {SAD_SHAPE_ID, SAD_BODY, SAD_SAMPLE_INFO, SAD_SAMPLE, SAD_MESH_ROOTS,
                                    SAD_MESH_HULL_INFO, SAD_MESH_PLANES, SAD_MESH_MIN, SAD_MESH_MAX, SAD_MESH_LINKS,
                                    SAD_MESH_TRIANGLES, SAD_MESH_NORMALS, SAD_MESH_CLUSTER_INFO,
                                    SAD_MESH_CLUSTER_RANGES, SAD_MESH_CLUSTER_MIN, SAD_MESH_CLUSTER_MAX,
                                    SAD_MESH_SAMPLE_INDICES, SAD_GRIDS, SAD_GRID_ORIGIN, SAD_DISTANCE,
                                    SAD_POSITION};
        std::vector<uint64_t> key;
        key.reserve(2 * (sizeof(slots) / sizeof(slots[0])) + 1);
        for (int k : slots) {
            key.push_back(reinterpret_cast<uint64_t>(table[k]));
            key.push_back(sizes[k]);
        }
        key.push_back(p.c.N);
        if (key != lastTable) { lastTable.swap(key); epoch++; }
        w.skip = 0;
        w.warm = 0;
        w.planeCache = 0;
        w.nodeGap = nullptr;
        w.lastFrame = nullptr;
        if (!skip || p.c.meshGeometry == 0) return SAD_OK;  // the SDF step has no node queries to skip
        if (maxNodesFor != p.sampleInfo || maxNodesShapes != p.c.shapes || triBaseFor != p.meshLinks) {
            if (!probeWord) HIP_OK(hipMalloc(reinterpret_cast<void**>(&probeWord), 3 * sizeof(uint32_t)));
            if (triBaseCap < 3 * std::max(p.c.shapes, 1u)) {
                if (triBase) (void)hipFree(triBase);
                triBase = nullptr;
                triBaseCap = 3 * std::max(p.c.shapes, 1u);
                HIP_OK(hipMalloc(reinterpret_cast<void**>(&triBase), triBaseCap * sizeof(uint32_t)));
            }
            HIP_OK(hipMemset(probeWord, 0, 3 * sizeof(uint32_t)));
            HIP_OK(hipMemset(triBase, 0, triBaseCap * sizeof(uint32_t)));
            hipLaunchKernelGGL(k_template_extent, dim3((std::max(p.c.shapes, 1u) + 255) / 256), dim3(256), 0, nullptr,
                               p, probeWord, triBase);
            HIP_OK(hipGetLastError());
            uint32_t extent[3] = {0, 0, 0};
            HIP_OK(hipMemcpy(extent, probeWord, sizeof(extent), hipMemcpyDeviceToHost));
            maxNodes = extent[0];
            maxTriangles = extent[1];
            maxClusters = extent[2];
            maxNodesFor = p.sampleInfo;
            maxNodesShapes = p.c.shapes;
            triBaseFor = p.meshLinks;
        }
        const uint32_t stride = (std::max(maxNodes, 1u) + 63) / 64 * 64;
        const size_t tasks = std::max<size_t>(w.taskCap, 1);
        const size_t bytesNeeded = tasks * stride;
        if (3 * bytesNeeded > (size_t(64) << 30)) return SAD_OK;  // more than 64 GiB of rows: leave it off
        // The unchanged-winner shortcut's rows (4 bytes per node slot) come on top,
        // up to 96 GiB more; beyond that the shortcut alone stays off.
        const bool gaps = maxTriangles <= 65534u && 4 * bytesNeeded <= (size_t(96) << 30);
        if (anchorCap < 4 * tasks) {
            if (w.anchor) { (void)hipFree(w.anchor); bytes -= size_t(anchorCap) * sizeof(float4); }
            w.anchor = nullptr;
            anchorCap = 0;
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.anchor), 4 * tasks * sizeof(float4)));
            HIP_OK(hipMemset(w.anchor, 0, 4 * tasks * sizeof(float4)));
            anchorCap = uint32_t(4 * tasks);
            bytes += size_t(anchorCap) * sizeof(float4);
            if (w.taskFar) { (void)hipFree(w.taskFar); bytes -= size_t(taskFarCap) * sizeof(uint32_t); }
            w.taskFar = nullptr;
            taskFarCap = 0;
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.taskFar), tasks * sizeof(uint32_t)));
            taskFarCap = uint32_t(tasks);
            bytes += size_t(taskFarCap) * sizeof(uint32_t);
            if (w.taskFrame) { (void)hipFree(w.taskFrame); bytes -= size_t(taskFrameCap) * sizeof(float4); }
            w.taskFrame = nullptr;
            taskFrameCap = 0;
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.taskFrame), TFRAME * tasks * sizeof(float4)));
            taskFrameCap = TFRAME * tasks;
            bytes += taskFrameCap * sizeof(float4);
        }
        if (nodeDistCap < bytesNeeded || w.nodeStride != stride) {
            if (w.nodeDist) { (void)hipFree(w.nodeDist); bytes -= nodeDistCap; }
            if (w.nodeTri) { (void)hipFree(w.nodeTri); bytes -= 2 * nodeDistCap; }
            w.nodeDist = nullptr;
            w.nodeTri = nullptr;
            nodeDistCap = 0;
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.nodeDist), bytesNeeded));
            // Any content is safe for the warm start (a triangle index in range is an
            // upper bound of the minimum; out of range is ignored): no clearing.
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.nodeTri), 2 * bytesNeeded));
            nodeDistCap = bytesNeeded;
            bytes += 3 * bytesNeeded;
            epoch++;  // stored distances are gone: no anchor may be trusted
        }
        w.nodeStride = stride;
        // The persistent-row map (pair capacity = tasks / 2): a new map starts with
        // no pair listed, so the next collection assigns every row.
        const uint32_t pairCap = uint32_t(tasks / 2);
        if (rowMapPairs < pairCap) {
            free_row_map();
            uint32_t slots = 1;
            while (slots < 2 * pairCap) slots <<= 1;
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.pairRow), size_t(pairCap) * sizeof(uint32_t)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowMark), size_t(pairCap) * sizeof(uint32_t)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowNeed), size_t(pairCap) * sizeof(uint32_t)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowNeedRank), size_t(pairCap) * sizeof(uint32_t)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowFree), size_t(pairCap) * sizeof(uint32_t)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowFreeRank), size_t(pairCap) * sizeof(uint32_t)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowKeys), size_t(slots) * sizeof(unsigned long long)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowOfSlot), size_t(slots) * sizeof(uint32_t)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowOwner), size_t(pairCap) * sizeof(unsigned long long)));
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.rowState), 8 * sizeof(uint32_t)));
            HIP_OK(hipMemset(w.pairRow, 0, size_t(pairCap) * sizeof(uint32_t)));
            HIP_OK(hipMemset(w.rowMark, 0, size_t(pairCap) * sizeof(uint32_t)));
            HIP_OK(hipMemset(w.rowOwner, 0xff, size_t(pairCap) * sizeof(unsigned long long)));  // no owner
            HIP_OK(hipMemset(w.rowState, 0, 8 * sizeof(uint32_t)));
            rowMapPairs = pairCap;
            w.rowPairs = pairCap;
            w.rowHashMask = slots - 1;
            bytes += size_t(pairCap) * 32 + size_t(slots) * 12 + 32;
        }
        const uint32_t clusterStride = (std::max(maxClusters, 1u) + 63) / 64 * 64;
        if (leafDistCap < tasks * clusterStride || w.clusterStride != clusterStride) {
            if (w.leafDist) { (void)hipFree(w.leafDist); bytes -= leafDistCap; }
            w.leafDist = nullptr;
            leafDistCap = 0;
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.leafDist), tasks * clusterStride));
            leafDistCap = tasks * clusterStride;
            bytes += leafDistCap;
            if (w.leafPlane) { (void)hipFree(w.leafPlane); bytes -= 2 * leafPlaneCap; }
            w.leafPlane = nullptr;
            leafPlaneCap = 0;
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&w.leafPlane), 2 * tasks * clusterStride));
            HIP_OK(hipMemset(w.leafPlane, 0, 2 * tasks * clusterStride));
            leafPlaneCap = tasks * clusterStride;
            bytes += 2 * leafPlaneCap;
            epoch++;
        }
        w.clusterStride = clusterStride;
        if (gaps && gapShortcut && (gapCap < 4 * bytesNeeded || lastFrameCap < 4 * tasks)) {
            if (gapRows) { (void)hipFree(gapRows); bytes -= gapCap; }
            if (lastFrames) { (void)hipFree(lastFrames); bytes -= size_t(lastFrameCap) * sizeof(float4); }
            gapRows = nullptr;
            lastFrames = nullptr;
            gapCap = 0;
            lastFrameCap = 0;
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&gapRows), 4 * bytesNeeded));
            HIP_OK(hipMemset(gapRows, 0, 4 * bytesNeeded));  // generation 0: never a usable gap
            HIP_OK(hipMalloc(reinterpret_cast<void**>(&lastFrames), 4 * tasks * sizeof(float4)));
            HIP_OK(hipMemset(lastFrames, 0, 4 * tasks * sizeof(float4)));
            gapCap = 4 * bytesNeeded;
            lastFrameCap = uint32_t(4 * tasks);
            bytes += gapCap + size_t(lastFrameCap) * sizeof(float4);
        }
        // A null row pointer turns the shortcut off in the kernel.
        const bool gapsOn = gaps && gapShortcut && gapRows && gapCap >= 4 * bytesNeeded;
        w.nodeGap = gapsOn ? gapRows : nullptr;
        w.lastFrame = gapsOn ? lastFrames : nullptr;
        w.triBase = triBase;
        w.warm = maxTriangles <= 65534u ? 1u : 0u;
        w.planeCache = maxTriangles < PLANE_TAG ? 1u : 0u;  // a tag never reads as a triangle
        w.skipLimit = float(skipLimitMicro) * 1e-6f;
        w.rebase = uint32_t(rebase);
        w.skip = 1;
        return SAD_OK;
    }
    // The device copy of the epoch, written on the launch stream before any
    // collection that follows a change (never inside a capture).
    int sync_epoch(hipStream_t s) {
        if (epochWritten == epoch || !w.counters) return SAD_OK;
        hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, s, w.counters + 10 * LINE, epoch);
        HIP_OK(hipGetLastError());
        epochWritten = epoch;
        return SAD_OK;
    }
    void free_row_map() {
        if (!rowMapPairs) return;
        (void)hipFree(w.pairRow); (void)hipFree(w.rowMark); (void)hipFree(w.rowNeed); (void)hipFree(w.rowNeedRank);
        (void)hipFree(w.rowFree); (void)hipFree(w.rowFreeRank); (void)hipFree(w.rowKeys); (void)hipFree(w.rowOfSlot);
        (void)hipFree(w.rowState); (void)hipFree(w.rowOwner);
        bytes -= size_t(rowMapPairs) * 32 + size_t(w.rowHashMask + 1) * 12 + 32;
        w.pairRow = w.rowMark = w.rowNeed = w.rowNeedRank = w.rowFree = w.rowFreeRank = w.rowOfSlot = w.rowState = nullptr;
        w.rowKeys = w.rowOwner = nullptr;
        rowMapPairs = 0;
        w.rowPairs = 0;
        w.rowHashMask = 0;
    }
    ~ScratchImpl() {
        free_row_map();
        drop_graphs();
        if (captureStream) (void)hipStreamDestroy(captureStream);
        for (auto& m : pending) { pool.push_back(m.start); pool.push_back(m.stop); }
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        (void)hipFree(w.totals);
        (void)hipFree(w.counters); (void)hipFree(w.taskCount); (void)hipFree(w.taskOffset); (void)hipFree(w.adjCount);
        (void)hipFree(w.adjOffset); (void)hipFree(w.adjCursor); (void)hipFree(w.adjList); (void)hipFree(w.sums); (void)hipFree(w.sdfReach); (void)hipFree(w.sdfHits);
        (void)hipFree(w.sdfLeaves); (void)hipFree(w.sdfOrder);
        (void)hipFree(w.xcd); (void)hipFree(w.xcdTasks); (void)hipFree(w.wide); (void)hipFree(w.stageCount); (void)hipFree(w.nnStats);
        (void)hipFree(w.keyCount); (void)hipFree(w.keyStart); (void)hipFree(w.builtKey); (void)hipFree(w.triPacked);
        (void)hipFree(w.sampleKeyed); (void)hipFree(w.clusterPacked); (void)hipFree(w.clusterSphere); (void)hipFree(w.templateRec); (void)hipFree(w.omegaRow); (void)hipFree(w.wallCount); (void)hipFree(w.wallOffset);
        (void)hipFree(w.anchor); (void)hipFree(w.taskFar); (void)hipFree(w.taskFrame); (void)hipFree(w.nodeDist); (void)hipFree(w.nodeTri); (void)hipFree(probeWord);
        (void)hipFree(triBase); (void)hipFree(w.leafDist); (void)hipFree(w.leafPlane); (void)hipFree(gapRows); (void)hipFree(lastFrames);
        (void)hipFree(w.stageKey); (void)hipFree(w.stageMeta); (void)hipFree(w.stagePd); (void)hipFree(w.stageNa);
    }
};

template <class T>
T* slot(void* const* b, int i) { return reinterpret_cast<T*>(b[i]); }

int bind(P& p, void* const* b, const uint64_t* bytes, const double* v) {
    if (!b || !bytes || !v) { std::snprintf(g_error, sizeof(g_error), "null binding table"); return SAD_E_ARGUMENT; }
    Cfg& c = p.c;
    c.activeTimestep = uint32_t(v[SAD_V_ACTIVE_TIMESTEP]);
    c.N = uint32_t(v[SAD_V_N]);
    c.shapes = uint32_t(v[SAD_V_SHAPES]);
    c.pairCapacity = uint32_t(v[SAD_V_PAIR_CAPACITY]);
    c.contactCapacity = uint32_t(v[SAD_V_CONTACT_CAPACITY]);
    c.binCapacity = uint32_t(v[SAD_V_BIN_CAPACITY]);
    c.binCount = uint32_t(v[SAD_V_BIN_COUNT]);
    c.bins[0] = uint32_t(v[SAD_V_BINS_X]);
    c.bins[1] = uint32_t(v[SAD_V_BINS_Y]);
    c.bins[2] = uint32_t(v[SAD_V_BINS_Z]);
    c.hashSize = uint32_t(v[SAD_V_HASH_SIZE]);
    c.dt = float(v[SAD_V_DT]);
    c.kn = float(v[SAD_V_KN]);
    c.kt = float(v[SAD_V_KT]);
    c.mu = float(v[SAD_V_MU]);
    c.normalViscosity = float(v[SAD_V_NORMAL_VISCOSITY]);
    c.skin = float(v[SAD_V_SKIN]);
    c.drag = float(v[SAD_V_DRAG]);
    c.maxOverlap = float(v[SAD_V_MAX_OVERLAP]);
    c.historyStep = uint32_t(v[SAD_V_HISTORY_STEP]);
    c.meshGeometry = uint32_t(v[SAD_V_MESH_GEOMETRY]);
    c.periodic = uint32_t(v[SAD_V_PERIODIC]);
    c.wallCount = uint32_t(v[SAD_V_WALL_COUNT]);
    c.wallBoxEnabled = uint32_t(v[SAD_V_WALL_BOX_ENABLED]);
    c.sourceSurfaceNormal = uint32_t(v[SAD_V_SOURCE_SURFACE_NORMAL]);
    if (c.sourceSurfaceNormal != 0) {
        std::snprintf(g_error, sizeof(g_error), "source-surface contact normals are not implemented");
        return SAD_E_UNSUPPORTED;
    }
    if (c.N == 0 || c.hashSize == 0 || (c.hashSize & (c.hashSize - 1)) != 0) {
        std::snprintf(g_error, sizeof(g_error), "invalid N or hash size");
        return SAD_E_ARGUMENT;
    }
    struct Need { int slot; uint64_t bytes; const char* name; };
    const Need needs[] = {
        {SAD_CONTROL, 64, "control"}, {SAD_CELL, 48, "cell"}, {SAD_POSITION, 16ull * c.N, "position"},
        {SAD_POSITION_REMAINDER, 16ull * c.N, "positionRemainder"}, {SAD_VELOCITY, 16ull * c.N, "velocity"},
        {SAD_ROTATION, 16ull * c.N, "rotation"}, {SAD_MOMENTUM, 16ull * c.N, "momentum"},
        {SAD_INV_I, 48ull * c.N, "invI"}, {SAD_BODY, 16ull * c.N, "body"}, {SAD_SHAPE_ID, 4ull * c.N, "shapeId"},
        {SAD_PAIRS, 8ull * c.pairCapacity, "pairs"}, {SAD_CONTACT_COUNTS, 4ull * c.pairCapacity, "contactCounts"},
        {SAD_META, 16ull * c.contactCapacity, "meta"}, {SAD_POINT_DEPTH, 16ull * c.contactCapacity, "pointDepth"},
        {SAD_NORMAL_AREA, 16ull * c.contactCapacity, "normalArea"}, {SAD_SLIP, 16ull * c.contactCapacity, "slip"},
        {SAD_CONTACT_FORCE, 16ull * c.contactCapacity, "contactForce"},
        {SAD_NORMAL_LOSS_POWER, 4ull * c.contactCapacity, "normalLossPower"},
        {SAD_HASH, 4ull * c.hashSize, "hash"}, {SAD_OLD_HASH, 4ull * c.hashSize, "oldHash"},
        {SAD_BIN_COUNTS, 4ull * c.binCount, "binCounts"},
        {SAD_BIN_PARTICLES, 4ull * c.binCount * c.binCapacity, "binParticles"},
        {SAD_PARTICLE_STRESS, 48ull * c.N, "particleStress"}, {SAD_MESH_PAIR_MASK, 4ull * c.pairCapacity, "meshPairMask"},
        {SAD_MESH_PAIR_TRANSFORM, 128ull * c.pairCapacity, "meshPairTransform"},
        {SAD_FORCE, 16ull * c.N, "force"}, {SAD_TORQUE, 16ull * c.N, "torque"},
        {SAD_ACCUMULATED, 16ull * c.N, "accumulated"}, {SAD_ACCUMULATED_REMAINDER, 16ull * c.N, "accumulatedRemainder"},
        {SAD_PARTICLE_DIAGNOSTIC, 16ull * c.N, "particleDiagnostic"}, {SAD_TIMESTEP_ROWS, 16ull * c.N, "timestepRows"},
        {SAD_VELOCITY_REMAINDER, 16ull * c.N, "velocityRemainder"},
        {SAD_MOMENTUM_REMAINDER, 16ull * c.N, "momentumRemainder"},
        {SAD_ROTATION_REMAINDER, 16ull * c.N, "rotationRemainder"},
        {SAD_REFERENCE_POSITION, 16ull * c.N, "referencePosition"},
        {SAD_REFERENCE_POSITION_REMAINDER, 16ull * c.N, "referencePositionRemainder"},
        {SAD_NORMAL_DAMPING, 16ull * c.N, "normalDamping"}, {SAD_FORCE_VELOCITY, 16ull * c.N, "forceVelocity"},
        {SAD_FORCE_OMEGA, 16ull * c.N, "forceOmega"}, {SAD_CELL_LOG, 32, "cellLog"},
        {SAD_REFERENCE_CELL, 32, "referenceCell"},
        {SAD_CONTACT_OFFSETS, 4ull * c.pairCapacity, "contactOffsets"},
        {SAD_SAMPLE_INFO, 16ull * c.shapes, "sampleInfo"},
        {SAD_GRIDS, c.meshGeometry ? 0ull : 16ull * c.shapes, "grids"},
        {SAD_GRID_ORIGIN, c.meshGeometry ? 0ull : 16ull * c.shapes, "gridOrigin"},
        {SAD_DISTANCE, c.meshGeometry ? 0ull : 4ull, "distance"},
    };
    for (const Need& n : needs) {
        if (!b[n.slot] || bytes[n.slot] < n.bytes) {
            std::snprintf(g_error, sizeof(g_error), "buffer %s too small (%llu < %llu bytes)", n.name,
                          (unsigned long long)bytes[n.slot], (unsigned long long)n.bytes);
            return SAD_E_ARGUMENT;
        }
    }
    p.meshNodes = uint32_t(bytes[SAD_MESH_LINKS] / 16);
    p.templates = uint32_t(bytes[SAD_MESH_ROOTS] / 16);
    p.planes = uint32_t(bytes[SAD_MESH_PLANES] / 16);
    p.triangles = uint32_t(bytes[SAD_MESH_TRIANGLES] / 48);
    p.sampleKeys = uint32_t(bytes[SAD_MESH_SAMPLE_INDICES] / 4);
    p.clusters = uint32_t(bytes[SAD_MESH_CLUSTER_RANGES] / 16);
    p.control = slot<uint32_t>(b, SAD_CONTROL);
    p.cell = slot<float4>(b, SAD_CELL);
    p.position = slot<float4>(b, SAD_POSITION);
    p.positionRemainder = slot<float4>(b, SAD_POSITION_REMAINDER);
    p.velocityRemainder = slot<float4>(b, SAD_VELOCITY_REMAINDER);
    p.momentumRemainder = slot<float4>(b, SAD_MOMENTUM_REMAINDER);
    p.rotationRemainder = slot<float4>(b, SAD_ROTATION_REMAINDER);
    p.referencePositionRemainder = slot<float4>(b, SAD_REFERENCE_POSITION_REMAINDER);
    p.cellLog = slot<float4>(b, SAD_CELL_LOG);
    p.accumulatedRemainder = slot<float4>(b, SAD_ACCUMULATED_REMAINDER);
    p.velocity = slot<float4>(b, SAD_VELOCITY);
    p.rotation = slot<float4>(b, SAD_ROTATION);
    p.momentum = slot<float4>(b, SAD_MOMENTUM);
    p.force = slot<float4>(b, SAD_FORCE);
    p.torque = slot<float4>(b, SAD_TORQUE);
    p.referencePosition = slot<float4>(b, SAD_REFERENCE_POSITION);
    p.referenceCell = slot<float4>(b, SAD_REFERENCE_CELL);
    p.shapeId = slot<const uint32_t>(b, SAD_SHAPE_ID);
    p.body = slot<const float4>(b, SAD_BODY);
    p.invI = slot<const float4>(b, SAD_INV_I);
    p.sampleInfo = slot<const uint4>(b, SAD_SAMPLE_INFO);
    p.sample = slot<const float4>(b, SAD_SAMPLE);
    p.meshRoots = slot<const uint4>(b, SAD_MESH_ROOTS);
    p.meshHullInfo = slot<const uint4>(b, SAD_MESH_HULL_INFO);
    p.meshPlanes = slot<const float4>(b, SAD_MESH_PLANES);
    p.meshMin = slot<const float4>(b, SAD_MESH_MIN);
    p.meshMax = slot<const float4>(b, SAD_MESH_MAX);
    p.meshLinks = slot<const uint4>(b, SAD_MESH_LINKS);
    p.meshTriangles = slot<const float4>(b, SAD_MESH_TRIANGLES);
    p.meshNormals = slot<const float4>(b, SAD_MESH_NORMALS);
    p.meshPairMask = slot<uint32_t>(b, SAD_MESH_PAIR_MASK);
    p.meshPairTransform = slot<float4>(b, SAD_MESH_PAIR_TRANSFORM);
    p.meshClusterInfo = slot<const uint4>(b, SAD_MESH_CLUSTER_INFO);
    p.meshClusterRanges = slot<const uint4>(b, SAD_MESH_CLUSTER_RANGES);
    p.meshClusterMin = slot<const float4>(b, SAD_MESH_CLUSTER_MIN);
    p.meshClusterMax = slot<const float4>(b, SAD_MESH_CLUSTER_MAX);
    p.meshSampleIndices = slot<const uint32_t>(b, SAD_MESH_SAMPLE_INDICES);
    p.binCounts = slot<uint32_t>(b, SAD_BIN_COUNTS);
    p.binParticles = slot<uint32_t>(b, SAD_BIN_PARTICLES);
    p.pairCounts = slot<uint32_t>(b, SAD_PAIR_COUNTS);
    p.pairOffsets = slot<uint32_t>(b, SAD_PAIR_OFFSETS);
    p.pairs = slot<uint2>(b, SAD_PAIRS);
    p.contactCounts = slot<uint32_t>(b, SAD_CONTACT_COUNTS);
    p.meta = slot<uint4>(b, SAD_META);
    p.pointDepth = slot<float4>(b, SAD_POINT_DEPTH);
    p.normalArea = slot<float4>(b, SAD_NORMAL_AREA);
    p.slip = slot<float4>(b, SAD_SLIP);
    p.contactForce = slot<float4>(b, SAD_CONTACT_FORCE);
    p.normalLossPower = slot<float>(b, SAD_NORMAL_LOSS_POWER);
    p.normalDamping = slot<float4>(b, SAD_NORMAL_DAMPING);
    p.forceVelocity = slot<float4>(b, SAD_FORCE_VELOCITY);
    p.forceOmega = slot<float4>(b, SAD_FORCE_OMEGA);
    p.oldMeta = slot<const uint4>(b, SAD_OLD_META);
    p.oldSlip = slot<const float4>(b, SAD_OLD_SLIP);
    p.oldNormal = slot<const float4>(b, SAD_OLD_NORMAL);
    p.oldHash = slot<const uint32_t>(b, SAD_OLD_HASH);
    p.hash = slot<uint32_t>(b, SAD_HASH);
    p.particleStress = slot<float4>(b, SAD_PARTICLE_STRESS);
    p.timestepRows = slot<float4>(b, SAD_TIMESTEP_ROWS);
    p.particleDiagnostic = slot<float4>(b, SAD_PARTICLE_DIAGNOSTIC);
    p.accumulated = slot<float4>(b, SAD_ACCUMULATED);
    p.contactOffsets = slot<uint32_t>(b, SAD_CONTACT_OFFSETS);
    p.grids = slot<const uint4>(b, SAD_GRIDS);
    p.gridOrigin = slot<const float4>(b, SAD_GRID_ORIGIN);
    p.distance = slot<const float>(b, SAD_DISTANCE);
    p.distanceCount = b[SAD_DISTANCE] ? bytes[SAD_DISTANCE] / 4 : 0;
    p.diagnosticRows = slot<float4>(b, SAD_DIAGNOSTIC_ROWS);
    p.diagnosticBytes = b[SAD_DIAGNOSTIC_ROWS] ? bytes[SAD_DIAGNOSTIC_ROWS] : 0;
    p.fabricRows = slot<float4>(b, SAD_FABRIC_ROWS);
    p.fabricBytes = b[SAD_FABRIC_ROWS] ? bytes[SAD_FABRIC_ROWS] : 0;
    if (c.wallCount != 0) {
        const uint64_t planes = 16ull * c.wallCount, tasks = uint64_t(c.wallCount) * c.N;
        const Need walls[] = {
            {SAD_WALL_PLANE, planes, "wallPlane"}, {SAD_WALL_REMAINDER, planes, "wallRemainder"},
            {SAD_WALL_VELOCITY, planes, "wallVelocity"}, {SAD_WALL_MATERIAL, planes, "wallMaterial"},
            {SAD_WALL_BOX_IDS, c.wallBoxEnabled ? 24ull : 0ull, "wallBoxIds"},
            {SAD_WALL_ROWS, 48ull * tasks, "wallRows"}, {SAD_WALL_ACCUMULATED, 16ull * tasks, "wallAccumulated"},
            {SAD_NORMAL_DAMPING, 16ull * c.N, "normalDamping"},
        };
        for (const Need& n : walls) {
            if (!b[n.slot] || bytes[n.slot] < n.bytes) {
                std::snprintf(g_error, sizeof(g_error), "buffer %s too small (%llu < %llu bytes)", n.name,
                              (unsigned long long)bytes[n.slot], (unsigned long long)n.bytes);
                return SAD_E_ARGUMENT;
            }
        }
        p.wallPlane = slot<float4>(b, SAD_WALL_PLANE);
        p.wallRemainder = slot<float4>(b, SAD_WALL_REMAINDER);
        p.wallVelocity = slot<const float4>(b, SAD_WALL_VELOCITY);
        p.wallMaterial = slot<const float4>(b, SAD_WALL_MATERIAL);
        p.wallBoxIds = slot<const uint32_t>(b, SAD_WALL_BOX_IDS);
        p.wallRows = slot<float4>(b, SAD_WALL_ROWS);
        p.wallAccumulated = slot<float4>(b, SAD_WALL_ACCUMULATED);
    }
    return SAD_OK;
}

inline unsigned blocks(uint64_t n, unsigned t = 256) { return unsigned(std::max<uint64_t>(1, (n + t - 1) / t)); }

hipEvent_t take_event(Scratch& sc) {
    if (!sc.pool.empty()) { hipEvent_t e = sc.pool.back(); sc.pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
// Launch with optional start/stop events on the same stream.
#define LAUNCH(SC, ID, KERNEL, GRID, BLOCK, STREAM, ...)                                        \
    do {                                                                                        \
        hipEvent_t _a = nullptr, _b = nullptr;                                                  \
        if ((SC).profiling) { _a = take_event(SC); _b = take_event(SC); }                       \
        if (_a) (void)hipEventRecord(_a, STREAM);                                               \
        hipLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, STREAM, __VA_ARGS__);                        \
        if (_b) { (void)hipEventRecord(_b, STREAM); (SC).pending.push_back({ID, _a, _b}); }      \
    } while (0)

int launch_scan(Scratch& sc, const ScanArgs& a, uint32_t cap, hipStream_t s) {
    unsigned tiles = blocks(cap, SCAN_TILE);
    LAUNCH(sc, KID_scan_reduce, k_scan_reduce, dim3(tiles), dim3(256), s, a);
    LAUNCH(sc, KID_scan_sums, k_scan_sums, dim3(1), dim3(256), s, a);
    LAUNCH(sc, KID_scan_apply, k_scan_apply, dim3(tiles), dim3(256), s, a);
    return SAD_OK;
}

// Staging slots a collection wave reserves per atomic.  Reserved-but-unused slots
// are at most STAGE_CHUNK - 1 per wave: chunks only where that leaves the queues'
// capacity ample.
uint32_t stage_chunk(const Cfg& c, const ScratchImpl& sc) {
    const uint64_t waves = uint64_t(std::max(1, sc.cus)) * sc.blocksPerCU * COLLECT_WAVES;
    return uint64_t(c.contactCapacity) >= 2 * waves * STAGE_CHUNK ? STAGE_CHUNK : 0u;
}

// Reference kernels by name (gpu.py:547-565 / solver.slang entry points).
int run_kernel(const std::string& name, P& p, ScratchImpl& sc, hipStream_t s) {
    const Cfg& c = p.c;
    W& w = sc.w;
    const dim3 T(256);
    const dim3 perGrain(blocks(c.N)), gridStride(std::max(1, sc.cus) * 8);
    if (name == "beginStep") LAUNCH(sc, KID_begin_step, k_begin_step, perGrain, T, s, p);
    else if (name == "advanceCell") LAUNCH(sc, KID_advance_cell, k_advance_cell, dim3(1), dim3(1), s, p);
    else if (name == "checkRebuild") LAUNCH(sc, KID_check_rebuild, k_check_rebuild, perGrain, T, s, p);
    else if (name == "clearBins") LAUNCH(sc, KID_clear_bins, k_clear_bins, dim3(blocks(c.binCount)), T, s, p);
    else if (name == "fillBins") {
        LAUNCH(sc, KID_fill_bins, k_fill_bins, perGrain, T, s, p);
        LAUNCH(sc, KID_sort_bins, k_sort_bins, dim3(blocks(c.binCount)), T, s, p);
    } else if (name == "countPairs") LAUNCH(sc, KID_count_pairs, k_count_pairs, perGrain, T, s, p);
    else if (name == "pairScan") {
        ScanArgs a{p.pairCounts, p.pairOffsets, w.sums, c.N, nullptr, 1, p.control + 4, nullptr};
        launch_scan(sc, a, c.N, s);
    } else if (name == "emitPairs") {
        LAUNCH(sc, KID_emit_pairs, k_emit_pairs, perGrain, T, s, p);
    } else if (name == "resetMeshContacts") {
        LAUNCH(sc, KID_reset_mesh, k_reset_mesh, dim3(blocks(c.pairCapacity)), T, s, p, w);
    } else if (name == "prepareMeshPairs") {
        LAUNCH(sc, KID_prepare_pairs, k_prepare_pairs, dim3(blocks(c.pairCapacity)), T, s, p);
    } else if (name == "collectMeshFrontiers" || name == "collectMeshContacts") {
        // collectMeshContacts is the reference's verification collector (same contact
        // set, atomic order); the CPU backend maps both names to its serial collector
        // (cpu.py:80-96), and so does this one, to the deterministic collection.
        const uint32_t* gate = w.counters + DERIVED_FLAG * LINE;
        LAUNCH(sc, KID_derived_check, k_derived_check, gridStride, T, s, p, w);
        LAUNCH(sc, KID_adj_clear, k_adj_clear, perGrain, T, s, p, w, gate);
        LAUNCH(sc, KID_adj_count, k_adj_count, gridStride, T, s, p, w, gate);
        ScanArgs adj{w.adjCount, w.adjOffset, w.sums, c.N, nullptr, 1, gate, nullptr};
        launch_scan(sc, adj, c.N, s);
        LAUNCH(sc, KID_adj_fill, k_adj_fill, gridStride, T, s, p, w, gate);
        LAUNCH(sc, KID_adj_sort, k_adj_sort, dim3(blocks(c.N, 4)), T, s, p, w, gate);
        int rc = sc.ensure_nodes(p.meshNodes);
        if (rc || (rc = sc.ensure_packed(p))) return rc;
        LAUNCH(sc, KID_wide_nodes, k_wide_nodes, dim3(blocks(p.meshNodes)), T, s, p, w, p.meshNodes);
        LAUNCH(sc, KID_pack_geometry, k_pack_triangles, dim3(std::min(blocks(p.triangles), 4096u)), T, s, p, w);
        LAUNCH(sc, KID_pack_geometry, k_pack_samples, dim3(std::max(c.shapes, 1u)), T, s, p, w);
        if (p.templates > MAX_TEMPLATES) {
            std::snprintf(g_error, sizeof(g_error), "more than %u templates", MAX_TEMPLATES);
            return SAD_E_UNSUPPORTED;
        }
        if ((rc = sc.ensure_keys(p.templates))) return rc;
        const uint32_t keys = bucketShape(p.templates).keys();
        LAUNCH(sc, KID_bucket_reset, k_bucket_reset, dim3(std::min(blocks(keys), 1024u)), T, s, p, w, gate);
        LAUNCH(sc, KID_bucket_count, k_bucket_count, gridStride, T, s, p, w, gate);
        ScanArgs keyScan{w.keyCount, w.keyStart, w.sums, keys, nullptr, 1, gate, w.counters + 6 * LINE};
        launch_scan(sc, keyScan, keys, s);
        LAUNCH(sc, KID_bucket_offsets, k_bucket_offsets, dim3(1), dim3(64), s, p, w, gate);
        LAUNCH(sc, KID_bucket_fill, k_bucket_fill, gridStride, T, s, p, w, gate);
        if (w.skip) {  // persistent task rows (assigned when the list was rebuilt)
            LAUNCH(sc, KID_rows, k_rows_lookup, gridStride, T, s, p, w);
            ScanArgs need{w.rowNeed, w.rowNeedRank, w.sums, w.rowPairs, nullptr, 1, w.rowState + 4, nullptr};
            launch_scan(sc, need, w.rowPairs, s);
            LAUNCH(sc, KID_rows, k_rows_free, gridStride, T, s, p, w);
            ScanArgs fr{w.rowFree, w.rowFreeRank, w.sums, w.rowPairs, nullptr, 1, w.rowState + 4, nullptr};
            launch_scan(sc, fr, w.rowPairs, s);
            LAUNCH(sc, KID_rows, k_rows_assign, gridStride, T, s, p, w);
            LAUNCH(sc, KID_task_far, k_task_far, dim3(blocks(2 * c.pairCapacity)), T, s, p, w);
        }
        const dim3 persistent(std::max(1, sc.cus) * sc.blocksPerCU);
        w.stageChunk = stage_chunk(c, sc);
        w.pop = sc.pop;
        w.leafBatch = sc.leafBatch;
        w.batchTasks = sc.batchTasks;
        w.batchFill = sc.batchFill;
        w.gapReach = float(sc.gapReachMicro) * 1e-6f;
        if (w.nnEnabled) {
            if (sc.queue == 128)
                LAUNCH(sc, KID_collect, (k_collect_nn<128>), persistent, dim3(COLLECT_WAVES * WAVE), s, p, w);
            else
                LAUNCH(sc, KID_collect, (k_collect_nn<256>), persistent, dim3(COLLECT_WAVES * WAVE), s, p, w);
        } else {
            if (sc.queue == 128)
                LAUNCH(sc, KID_collect, (k_collect<128>), persistent, dim3(COLLECT_WAVES * WAVE), s, p, w);
            else
                LAUNCH(sc, KID_collect, (k_collect<256>), persistent, dim3(COLLECT_WAVES * WAVE), s, p, w);
        }
        ScanArgs a{w.taskCount, w.taskOffset, w.sums, 2 * c.pairCapacity, p.control + 1, 2, nullptr, w.counters + 4 * LINE};
        launch_scan(sc, a, 2 * c.pairCapacity, s);
        LAUNCH(sc, KID_contact_total, k_contact_total, dim3(blocks(c.pairCapacity)), T, s, p, w);
        LAUNCH(sc, KID_scatter, k_scatter, dim3(gridStride.x / 8 + 1, 8), T, s, p, w);
    } else if (name == "clearContactStorage") LAUNCH(sc, KID_clear_storage, k_clear_storage, gridStride, T, s, p);
    else if (name == "resolveContacts") {
        LAUNCH(sc, KID_omega_rows, k_omega_rows, perGrain, T, s, p, w);
        LAUNCH(sc, KID_resolve, k_resolve, gridStride, T, s, p, w);
    }
    else if (name == "buildHistoryHash") LAUNCH(sc, KID_build_hash, k_build_hash, gridStride, T, s, p);
    else if (name == "gather") LAUNCH(sc, KID_gather, k_gather, dim3(blocks(c.N, 4)), T, s, p, w);
    else if (name == "finishStep") LAUNCH(sc, KID_finish_step, k_finish_step, perGrain, T, s, p);
    else if (name == "captureContactKinematics") LAUNCH(sc, KID_capture_kinematics, k_capture_kinematics, perGrain, T, s, p, 0);
    else if (name == "checkContactTimestep") LAUNCH(sc, KID_check_timestep, k_check_timestep, perGrain, T, s, p);
    else if (name == "advanceWalls") {
        if (c.wallCount) LAUNCH(sc, KID_advance_walls, k_advance_walls, dim3(1), dim3(1), s, p);
    } else if (name == "collectWallContacts") {
        if (c.wallCount) {
            const uint32_t tasks = c.wallCount * c.N;
            LAUNCH(sc, KID_wall_count, k_wall_collect<false>, dim3(blocks(tasks, 4)), T, s, p, w);
            ScanArgs a{w.wallCount, w.wallOffset, w.sums, tasks, nullptr, 1, nullptr, w.counters + 7 * LINE};
            launch_scan(sc, a, tasks, s);
            LAUNCH(sc, KID_wall_emit, k_wall_collect<true>, dim3(blocks(tasks, 4)), T, s, p, w);
            LAUNCH(sc, KID_wall_total, k_wall_total, dim3(1), dim3(1), s, p, w);
        }
    } else if (name == "gatherWallRows") {
        if (c.wallCount) LAUNCH(sc, KID_gather_wall_rows, k_gather_wall_rows, dim3(blocks(c.wallCount * c.N)), T, s, p, w);
    } else if (name == "refreshWallPower") {
        if (c.wallCount) LAUNCH(sc, KID_refresh_wall_power, k_refresh_wall_power, dim3(blocks(c.wallCount * c.N)), T, s, p);
    } else if (name == "countContacts") {
        // The gather walks each grain's pairs in pair order (as after a mesh collection).
        LAUNCH(sc, KID_adj_clear, k_adj_clear, perGrain, T, s, p, w, nullptr);
        LAUNCH(sc, KID_adj_count, k_adj_count, gridStride, T, s, p, w, nullptr);
        ScanArgs adj{w.adjCount, w.adjOffset, w.sums, c.N, nullptr, 1, nullptr, nullptr};
        launch_scan(sc, adj, c.N, s);
        LAUNCH(sc, KID_adj_fill, k_adj_fill, gridStride, T, s, p, w, nullptr);
        LAUNCH(sc, KID_adj_sort, k_adj_sort, dim3(blocks(c.N, 4)), T, s, p, w, nullptr);
        LAUNCH(sc, KID_sdf_count, k_sdf_contacts<false>, dim3(blocks(c.pairCapacity, 4)), T, s, p, w);
    } else if (name == "contactScan") {
        ScanArgs a{p.contactCounts, p.contactOffsets, w.sums, c.pairCapacity, nullptr, 1, nullptr, nullptr};
        launch_scan(sc, a, c.pairCapacity, s);
    } else if (name == "contactTotal") {
        LAUNCH(sc, KID_sdf_total, k_sdf_total, dim3(1), dim3(1), s, p, w);
    } else if (name == "emitContacts") {
        LAUNCH(sc, KID_sdf_emit, k_sdf_contacts<true>, dim3(blocks(c.pairCapacity, 4)), T, s, p, w);
    } else if (name == "diagnosticRows") {
        if (!p.diagnosticRows || p.diagnosticBytes < 128ull * c.N) {
            std::snprintf(g_error, sizeof(g_error), "buffer diagnosticRows too small for %u grains", c.N);
            return SAD_E_ARGUMENT;
        }
        LAUNCH(sc, KID_diagnostic_rows, k_diagnostic_rows, perGrain, T, s, p);
    } else if (name == "fabricRows") {
        if (!p.fabricRows || p.fabricBytes < 32ull * c.pairCapacity) {
            std::snprintf(g_error, sizeof(g_error), "buffer fabricRows too small for %u pair slots", c.pairCapacity);
            return SAD_E_ARGUMENT;
        }
        LAUNCH(sc, KID_fabric_rows, k_fabric_rows, dim3(blocks(c.pairCapacity)), T, s, p);
    } else {
        if (name == "smoke") {
            std::snprintf(g_error, sizeof(g_error), "smoke binds its own buffers: call sad_hip_smoke");
            return SAD_E_UNSUPPORTED;
        }
        std::snprintf(g_error, sizeof(g_error), "unknown kernel %s", name.c_str());
        return SAD_E_ARGUMENT;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, name.c_str());
    return SAD_OK;
}

// Resolve pending event pairs into per-kernel totals (synchronizes the events).
int resolve_profile(Scratch& sc) {
    for (auto& m : sc.pending) {
        HIP_OK(hipEventSynchronize(m.stop));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, m.start, m.stop));
        sc.ms[m.kernel] += ms;
        sc.launches[m.kernel] += 1;
        sc.pool.push_back(m.start);
        sc.pool.push_back(m.stop);
    }
    sc.pending.clear();
    return SAD_OK;
}

}  // namespace

extern "C" {

void* sad_hip_scratch_create(void) {
    try { return new ScratchImpl(); } catch (...) { return nullptr; }
}
void sad_hip_scratch_destroy(void* scratch) { delete static_cast<ScratchImpl*>(scratch); }
int sad_hip_scratch_stats(void* scratch, uint64_t* stats) {
    if (!scratch || !stats) return SAD_E_ARGUMENT;
    auto* sc = static_cast<ScratchImpl*>(scratch);
    unsigned long long totals[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (sc->w.totals) {
        hipError_t e = hipMemcpy(totals, sc->w.totals, sizeof(totals), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "stats");
    }
    for (int k = 0; k < 4; k++) stats[k] = totals[k];
    stats[8] = totals[4];
    stats[9] = totals[5];
    stats[10] = totals[6];
    stats[11] = totals[7];
    stats[4] = sc->bytes;
    unsigned long long nnv[3] = {0, 0, 0};
    if (sc->w.nnStats) {
        for (int k = 0; k < 3; k++) {
            hipError_t e = hipMemcpy(nnv + k, sc->w.nnStats + k * (LINE / 2), 8, hipMemcpyDeviceToHost);
            if (e != hipSuccess) return hip_fail(e, "stats");
        }
    }
    for (int k = 0; k < 3; k++) stats[5 + k] = nnv[k];
    return SAD_OK;
}

int sad_hip_profile(void* scratch, int enable) {
    if (!scratch) return SAD_E_ARGUMENT;
    auto* sc = static_cast<ScratchImpl*>(scratch);
    int rc = resolve_profile(*sc);
    if (rc) return rc;
    if (enable) {
        std::memset(sc->ms, 0, sizeof(sc->ms));
        std::memset(sc->launches, 0, sizeof(sc->launches));
    }
    sc->profiling = enable != 0;
    return SAD_OK;
}

int sad_hip_profile_read(void* scratch, double* ms, uint64_t* launches, uint32_t capacity) {
    if (!scratch) return -SAD_E_ARGUMENT;
    auto* sc = static_cast<ScratchImpl*>(scratch);
    if (resolve_profile(*sc)) return -SAD_E_HIP;
    uint32_t n = std::min<uint32_t>(capacity, KID_COUNT);
    for (uint32_t k = 0; k < n; k++) {
        if (ms) ms[k] = sc->ms[k];
        if (launches) launches[k] = sc->launches[k];
    }
    return int(KID_COUNT);
}

const char* sad_hip_kernel_name(uint32_t kernel) { return kernel < KID_COUNT ? kKernelNames[kernel] : ""; }

int sad_hip_dispatch(const char* kernel, uint32_t count, void* const* buffers, const uint64_t* bytes,
                     const double* values, void* scratch, void* stream) {
    (void)count;
    if (!kernel || !scratch) return SAD_E_ARGUMENT;
    try {
        P p{};
        int rc = bind(p, buffers, bytes, values);
        if (rc) return rc;
        auto* sc = static_cast<ScratchImpl*>(scratch);
        if ((rc = sc->ensure(p.c))) return rc;
        if ((rc = sc->ensure_skip(p, buffers, bytes))) return rc;
        if ((rc = sc->ensure_sdf(p, static_cast<hipStream_t>(stream)))) return rc;
        if ((rc = sc->sync_epoch(static_cast<hipStream_t>(stream)))) return rc;
        return run_kernel(kernel, p, *sc, static_cast<hipStream_t>(stream));
    } catch (...) {
        std::snprintf(g_error, sizeof(g_error), "exception in dispatch");
        return SAD_E_HIP;
    }
}

// One timestep's launches (per-step order of gpu.py:714-726).
int launch_step(P& p, ScratchImpl& sc, hipStream_t s) {
    static const char* mesh[] = {"beginStep", "advanceCell", "advanceWalls", "checkRebuild", "clearBins",
                                 "fillBins", "countPairs", "pairScan", "emitPairs", "resetMeshContacts",
                                 "collectMeshFrontiers", "collectWallContacts", "clearContactStorage",
                                 "resolveContacts", "buildHistoryHash", "gather", "gatherWallRows"};
    static const char* sdf[] = {"beginStep", "advanceCell", "advanceWalls", "checkRebuild", "clearBins",
                                "fillBins", "countPairs", "pairScan", "emitPairs", "countContacts", "contactScan",
                                "contactTotal", "emitContacts", "collectWallContacts", "clearContactStorage",
                                "resolveContacts", "buildHistoryHash", "gather", "gatherWallRows"};
    int rc;
    if (p.c.activeTimestep && (rc = run_kernel("checkContactTimestep", p, sc, s))) return rc;
    const char* const* sequence = p.c.meshGeometry ? mesh : sdf;
    const size_t steps = p.c.meshGeometry ? sizeof(mesh) / sizeof(mesh[0]) : sizeof(sdf) / sizeof(sdf[0]);
    for (size_t k = 0; k < steps; k++) {
        const char* name = sequence[k];
        if (std::strcmp(name, "clearBins") == 0) {
            // The Engine captures contact kinematics when any contact is viscous;
            // with planes that depends on device-resident wall materials.
            if (p.c.normalViscosity > 0) {
                if ((rc = run_kernel("captureContactKinematics", p, sc, s))) return rc;
            } else if (p.c.wallCount) {
                LAUNCH(sc, KID_capture_kinematics, k_capture_kinematics, dim3(blocks(p.c.N)), dim3(256), s, p, 1);
            }
        }
        if ((rc = run_kernel(name, p, sc, s))) return rc;
    }
    if (p.c.activeTimestep && (rc = run_kernel("checkContactTimestep", p, sc, s))) return rc;
    return run_kernel("finishStep", p, sc, s);
}

// Graph key: every value a captured step bakes into its launches.
std::vector<uint64_t> graph_key(void* const* table, const uint64_t* sizes, const double* vals, const ScratchImpl& sc) {
    std::vector<uint64_t> key;
    key.reserve(2 * SAD_BUFFER_COUNT + SAD_VALUE_COUNT + sizeof(W) / 8 + 4);
    for (int k = 0; k < SAD_BUFFER_COUNT; k++) key.push_back(reinterpret_cast<uint64_t>(table[k]));
    for (int k = 0; k < SAD_BUFFER_COUNT; k++) key.push_back(sizes[k]);
    for (int k = 0; k < SAD_VALUE_COUNT; k++) {
        uint64_t bits;
        std::memcpy(&bits, vals + k, 8);
        key.push_back(bits);
    }
    uint64_t words[(sizeof(W) + 7) / 8] = {};
    std::memcpy(words, &sc.w, sizeof(W));
    for (uint64_t v : words) key.push_back(v);
    key.push_back(uint64_t(sc.blocksPerCU));
    key.push_back(uint64_t(sc.pop));
    key.push_back(uint64_t(sc.queue));
    key.push_back(uint64_t(sc.cus));
    return key;
}

// Replay (capturing on first use) the step for this key on stream s.
int graph_step(P& p, ScratchImpl& sc, const std::vector<uint64_t>& key, hipStream_t s) {
    hipGraphExec_t exec = nullptr;
    for (size_t k = 0; k < sc.graphs.size(); k++)
        if (sc.graphs[k].key == key) {
            exec = sc.graphs[k].exec;
            std::rotate(sc.graphs.begin(), sc.graphs.begin() + k, sc.graphs.begin() + k + 1);  // most recent first
            break;
        }
    if (!exec) {
        if (!sc.captureStream) HIP_OK(hipStreamCreateWithFlags(&sc.captureStream, hipStreamNonBlocking));
        HIP_OK(hipStreamBeginCapture(sc.captureStream, hipStreamCaptureModeRelaxed));
        int rc = launch_step(p, sc, sc.captureStream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(sc.captureStream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
        if (sc.graphs.size() >= 4) {
            (void)hipGraphExecDestroy(sc.graphs.back().exec);
            sc.graphs.pop_back();
        }
        sc.graphs.insert(sc.graphs.begin(), ScratchImpl::Graph{key, exec});
    }
    HIP_OK(hipGraphLaunch(exec, s));
    return SAD_OK;
}

int sad_hip_advance(uint32_t steps, uint32_t active, void* const* buffers, const uint64_t* bytes,
                    const double* values, void* const* sides, const uint64_t* side_bytes, void* scratch,
                    void* stream, int graph) {
    if (!scratch || !sides || !side_bytes) return SAD_E_ARGUMENT;
    try {
        void* table[SAD_BUFFER_COUNT];
        uint64_t sizes[SAD_BUFFER_COUNT];
        std::memcpy(table, buffers, sizeof(table));
        std::memcpy(sizes, bytes, sizeof(sizes));
        double vals[SAD_VALUE_COUNT];
        std::memcpy(vals, values, sizeof(vals));
        auto* sc = static_cast<ScratchImpl*>(scratch);
        hipStream_t s = static_cast<hipStream_t>(stream);
        for (uint32_t k = 0; k < steps; k++) {
            uint32_t side = (active + k) & 1, old = 1 - side;
            const int slotsNew[4] = {SAD_META, SAD_SLIP, SAD_NORMAL_AREA, SAD_HASH};
            const int slotsOld[4] = {SAD_OLD_META, SAD_OLD_SLIP, SAD_OLD_NORMAL, SAD_OLD_HASH};
            for (int q = 0; q < SAD_SIDE_COUNT; q++) {
                table[slotsNew[q]] = sides[side * SAD_SIDE_COUNT + q];
                sizes[slotsNew[q]] = side_bytes[side * SAD_SIDE_COUNT + q];
                table[slotsOld[q]] = sides[old * SAD_SIDE_COUNT + q];
                sizes[slotsOld[q]] = side_bytes[old * SAD_SIDE_COUNT + q];
            }
            if (k > 0) vals[SAD_V_HISTORY_STEP] = 1;
            P p{};
            int rc = bind(p, table, sizes, vals);
            if (rc) return rc;
            if ((rc = sc->ensure(p.c))) return rc;
            if ((rc = sc->ensure_nodes(p.meshNodes))) return rc;
            if ((rc = sc->ensure_packed(p))) return rc;
            if (p.templates <= MAX_TEMPLATES && (rc = sc->ensure_keys(p.templates))) return rc;
            if ((rc = sc->ensure_skip(p, table, sizes))) return rc;
            if ((rc = sc->ensure_sdf(p, s))) return rc;  // no allocation inside a captured step
            if ((rc = sc->sync_epoch(s))) return rc;
            if (graph && !sc->profiling) {
                sc->w.stageChunk = stage_chunk(p.c, *sc);
                sc->w.pop = sc->pop;
                sc->w.leafBatch = sc->leafBatch;
                sc->w.batchTasks = sc->batchTasks;
                sc->w.batchFill = sc->batchFill;
                sc->w.gapReach = float(sc->gapReachMicro) * 1e-6f;
                rc = graph_step(p, *sc, graph_key(table, sizes, vals, *sc), s);
            } else {
                rc = launch_step(p, *sc, s);
            }
            if (rc) return rc;
        }
        return SAD_OK;
    } catch (...) {
        std::snprintf(g_error, sizeof(g_error), "exception in advance");
        return SAD_E_HIP;
    }
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free stepping
// sad_hip_advance_static: the timesteps of sad_hip_advance with the library
// workspace carved from one file-scope device array instead of hipMalloc'd
// buffers, for callers that must not allocate (graph capture, serving loops).
// Plain launches only: no captured graph, no temporal skip, no learned cull.
// Each GPU holds its own copy of the array; all calls on one GPU share it, so
// calls that overlap on one GPU race.  Every call clears the counters it needs.
namespace {

constexpr size_t STATIC_WORKSPACE_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char g_static_workspace[STATIC_WORKSPACE_BYTES];
constexpr int STATIC_DEVICES = 64;

__global__ void k_fill_words(uint32_t* words, uint32_t count, uint32_t value) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x)
        words[k] = value;
}

struct StaticDevice {
    unsigned char* base = nullptr;  // this device's address of g_static_workspace
    ScratchImpl* context = nullptr; // launch settings; its workspace points into base
};
StaticDevice g_static[STATIC_DEVICES];
std::mutex g_static_lock;

// Per device, with that device current: the array's device address (the host-side
// name is only a handle) and a launch context that owns no device memory.
int static_device(StaticDevice*& out) {
    int device = -1;
    HIP_OK(hipGetDevice(&device));
    if (device < 0 || device >= STATIC_DEVICES) {
        std::snprintf(g_error, sizeof(g_error), "device %d beyond the %d the static workspace serves", device,
                      STATIC_DEVICES);
        return SAD_E_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> guard(g_static_lock);
    StaticDevice& d = g_static[device];
    if (!d.base) {
        void* address = nullptr;
        HIP_OK(hipGetSymbolAddress(&address, HIP_SYMBOL(g_static_workspace)));
        if (!address) {
            std::snprintf(g_error, sizeof(g_error), "no device address for the static workspace");
            return SAD_E_HIP;
        }
        d.base = static_cast<unsigned char*>(address);
    }
    if (!d.context) {
        auto* context = new ScratchImpl();  // host object; never destroyed (its workspace is static)
        context->device = device;
        HIP_OK(hipDeviceGetAttribute(&context->cus, hipDeviceAttributeMultiprocessorCount, device));
        d.context = context;
    }
    out = &d;
    return SAD_OK;
}

// Lay the workspace out in the static array: the zero-initialised head (counters,
// staged counts, totals, queue counters, plane-contact counts) first, then the rest.
// Returns the bytes used, or 0 with g_error set when the array is too small.
size_t carve_static(ScratchImpl& sc, unsigned char* base, const P& p, size_t& head) {
    const Cfg& c = p.c;
    W& w = sc.w;
    Caps& caps = sc.caps;
    size_t used = 0;
    auto take = [&](auto*& ptr, size_t count, uint32_t& cap) {
        using T = std::remove_reference_t<decltype(*ptr)>;
        count = std::max<size_t>(count, 1);
        ptr = reinterpret_cast<T*>(base + used);
        used += (count * sizeof(T) + 255) & ~size_t(255);
        cap = uint32_t(std::min<size_t>(count, 0xffffffffu));
    };
    uint32_t unused = 0;
    const size_t tasks = size_t(2) * c.pairCapacity, wallTasks = size_t(c.wallCount) * c.N;
    const uint32_t keys = bucketShape(p.templates).keys();
    take(w.counters, 16 * LINE, caps.counters);
    take(w.stageCount, 8 * LINE, caps.stageCount);
    take(w.totals, 8, unused);
    take(w.xcd, 24 * LINE, caps.xcd);
    take(w.wallCount, wallTasks, caps.wallCount);
    head = used;
    take(w.wallOffset, wallTasks, caps.wallOffset);
    take(w.taskCount, tasks, caps.taskCount);
    take(w.taskOffset, tasks + 1, caps.taskOffset);
    take(w.adjCount, c.N, caps.adjCount);
    take(w.omegaRow, c.N, caps.omegaRow);
    take(w.adjOffset, c.N, caps.adjOffset);
    take(w.adjCursor, c.N, caps.adjCursor);
    take(w.adjList, tasks, caps.adjList);
    take(w.xcdTasks, tasks, caps.xcdTasks);
    take(w.keyCount, keys, caps.keyCount);
    take(w.keyStart, keys, caps.keyStart);
    const size_t widest = std::max<size_t>(std::max<size_t>(tasks + 1, c.N), std::max<size_t>(wallTasks, keys));
    take(w.sums, (widest + SCAN_TILE - 1) / SCAN_TILE + 1, caps.sums);
    take(w.wide, 4 * size_t(std::max(p.meshNodes, 1u)), caps.wide);
    take(w.triPacked, 3 * size_t(p.triangles), caps.triPacked);
    take(w.sampleKeyed, p.sampleKeys, caps.sampleKeyed);
    take(w.clusterPacked, p.clusters, caps.clusterPacked);
    take(w.clusterSphere, p.clusters, caps.clusterSphere);
    take(w.templateRec, TREC * size_t(p.c.shapes), caps.templateRec);
    take(w.stageKey, size_t(8) * c.contactCapacity, caps.stageKey);
    take(w.stageMeta, size_t(8) * c.contactCapacity, caps.stageMeta);
    take(w.stagePd, size_t(8) * c.contactCapacity, caps.stagePd);
    take(w.stageNa, size_t(8) * c.contactCapacity, caps.stageNa);
    if (used > STATIC_WORKSPACE_BYTES) {
        std::snprintf(g_error, sizeof(g_error),
                      "allocation-free stepping needs %zu bytes of workspace for this assembly; the static "
                      "workspace holds %zu (N %u, pair capacity %u, contact capacity %u)",
                      used, STATIC_WORKSPACE_BYTES, c.N, c.pairCapacity, c.contactCapacity);
        return 0;
    }
    w.stageCap = c.contactCapacity;
    w.taskCap = uint32_t(tasks);
    w.nCap = c.N;
    w.sumsCap = caps.sums;
    w.wallCap = uint32_t(std::max<size_t>(wallTasks, 1));
    w.skip = 0;
    w.warm = 0;
    w.planeCache = 0;
    w.nnEnabled = 0;
    w.nnAudit = 0;
    w.nnTemplates = 0;
    w.anchor = nullptr;
    w.taskFar = nullptr;
    w.taskFrame = nullptr;
    w.nodeDist = nullptr;
    w.nodeTri = nullptr;
    w.leafDist = nullptr;
    w.leafPlane = nullptr;
    w.triBase = nullptr;
    w.nodeGap = nullptr;
    w.lastFrame = nullptr;
    w.builtKey = nullptr;  // the adjacency and buckets are rebuilt in every step
    return used;
}

}  // namespace

extern "C" {

int sad_hip_advance_static(uint32_t steps, uint32_t active, void* const* buffers, const uint64_t* bytes,
                           const double* values, void* const* sides, const uint64_t* side_bytes, void* stream) {
    if (!buffers || !bytes || !values || !sides || !side_bytes) return SAD_E_ARGUMENT;
    try {
        StaticDevice* d = nullptr;
        int rc = static_device(d);
        if (rc) return rc;
        ScratchImpl& sc = *d->context;
        hipStream_t s = static_cast<hipStream_t>(stream);
        void* table[SAD_BUFFER_COUNT];
        uint64_t sizes[SAD_BUFFER_COUNT];
        std::memcpy(table, buffers, sizeof(table));
        std::memcpy(sizes, bytes, sizeof(sizes));
        double vals[SAD_VALUE_COUNT];
        std::memcpy(vals, values, sizeof(vals));
        for (uint32_t k = 0; k < steps; k++) {
            const uint32_t side = (active + k) & 1, old = 1 - side;
            const int slotsNew[4] = {SAD_META, SAD_SLIP, SAD_NORMAL_AREA, SAD_HASH};
            const int slotsOld[4] = {SAD_OLD_META, SAD_OLD_SLIP, SAD_OLD_NORMAL, SAD_OLD_HASH};
            for (int q = 0; q < SAD_SIDE_COUNT; q++) {
                table[slotsNew[q]] = sides[side * SAD_SIDE_COUNT + q];
                sizes[slotsNew[q]] = side_bytes[side * SAD_SIDE_COUNT + q];
                table[slotsOld[q]] = sides[old * SAD_SIDE_COUNT + q];
                sizes[slotsOld[q]] = side_bytes[old * SAD_SIDE_COUNT + q];
            }
            if (k > 0) vals[SAD_V_HISTORY_STEP] = 1;
            P p{};
            if ((rc = bind(p, table, sizes, vals))) return rc;
            if (!p.c.meshGeometry) {
                std::snprintf(g_error, sizeof(g_error), "allocation-free stepping covers the mesh geometry only");
                return SAD_E_UNSUPPORTED;
            }
            if (p.templates > MAX_TEMPLATES) {
                std::snprintf(g_error, sizeof(g_error), "more than %u templates", MAX_TEMPLATES);
                return SAD_E_UNSUPPORTED;
            }
            size_t head = 0;
            if (!carve_static(sc, d->base, p, head)) return SAD_E_UNSUPPORTED;
            if (k == 0) {  // what a previous call left must not leak into this one
                const uint32_t words = uint32_t(head / 4);
                hipLaunchKernelGGL(k_fill_words, dim3(std::min<uint32_t>(blocks(words), 1024u)), dim3(256), 0, s,
                                   reinterpret_cast<uint32_t*>(d->base), words, 0u);
                HIP_OK(hipGetLastError());
            }
            if ((rc = launch_step(p, sc, s))) return rc;
        }
        return SAD_OK;
    } catch (...) {
        std::snprintf(g_error, sizeof(g_error), "exception in advance_static");
        return SAD_E_HIP;
    }
}

int sad_hip_tune(void* scratch, const char* name, int64_t value, int64_t* previous) {
    if (!scratch || !name) return SAD_E_ARGUMENT;
    auto* sc = static_cast<ScratchImpl*>(scratch);
    std::string key(name);
    int64_t old;
    if (key == "collect_blocks_per_cu") {
        old = sc->blocksPerCU;
        if (value >= 0) {
            if (value < 1 || value > 32) return SAD_E_ARGUMENT;
            sc->blocksPerCU = int(value);
        }
    } else if (key == "collect_pop") {
        old = sc->pop;
        if (value >= 0) {
            if (value < 1 || value > 16) return SAD_E_ARGUMENT;
            sc->pop = uint32_t(value);
        }
    } else if (key == "collect_queue") {
        old = sc->queue;
        if (value >= 0) {
            if (value != 128 && value != 256) return SAD_E_ARGUMENT;
            sc->queue = int(value);
        }
    } else if (key == "collect_leaf_batch") {
        old = sc->leafBatch;
        if (value >= 0) {
            if (value < 1 || value > 64) return SAD_E_ARGUMENT;
            sc->leafBatch = uint32_t(value);
        }
    } else if (key == "collect_batch_tasks") {
        old = sc->batchTasks;
        if (value >= 0) {
            if (value < 1 || value > int64_t(BATCH_MAX)) return SAD_E_ARGUMENT;
            sc->batchTasks = uint32_t(value);
        }
    } else if (key == "collect_batch_fill") {
        old = sc->batchFill;
        if (value >= 0) {
            if (value < 1 || value > QCAP) return SAD_E_ARGUMENT;
            sc->batchFill = uint32_t(value);
        }
    } else if (key == "gap_reach_micro") {
        old = sc->gapReachMicro;
        if (value >= 0) {
            if (value > 100000) return SAD_E_ARGUMENT;
            sc->gapReachMicro = int(value);
        }
    } else if (key == "skip_rebase") {
        old = sc->rebase;
        if (value >= 0) {
            if (value > 1) return SAD_E_ARGUMENT;
            sc->rebase = int(value);
        }
    } else if (key == "gap_shortcut") {
        old = sc->gapShortcut;
        if (value >= 0) {
            if (value > 1) return SAD_E_ARGUMENT;
            sc->gapShortcut = int(value);
        }
    } else if (key == "temporal_skip") {
        old = sc->skip;
        if (value >= 0) {
            if (value > 1) return SAD_E_ARGUMENT;
            sc->skip = int(value);
        }
    } else if (key == "skip_limit_micro") {
        old = sc->skipLimitMicro;
        if (value >= 0) {
            if (value < 2000 || value > 250000) return SAD_E_ARGUMENT;  // above the margin, inside the byte range
            sc->skipLimitMicro = int(value);
        }
    } else {
        std::snprintf(g_error, sizeof(g_error), "unknown tunable %s", name);
        return SAD_E_ARGUMENT;
    }
    if (previous) *previous = old;
    return SAD_OK;
}

int sad_hip_invalidate(void* scratch) {
    if (!scratch) return SAD_E_ARGUMENT;
    static_cast<ScratchImpl*>(scratch)->epoch++;
    return SAD_OK;
}

int sad_hip_probe(const char* kernel, const uint32_t* uniforms, void* const* buffers, const uint64_t* bytes,
                  const double* values, const void* input, void* output, void* stream) {
    if (!kernel || !uniforms || !output) return SAD_E_ARGUMENT;
    P p{};
    int rc = bind(p, buffers, bytes, values);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::string name(kernel);
    uint32_t count = uniforms[0];
    if (count == 0) return SAD_OK;
    if (name == "meshProbe")
        hipLaunchKernelGGL(k_mesh_probe, dim3(blocks(count)), dim3(256), 0, s, p, count, uniforms[1],
                           static_cast<const float4*>(input), static_cast<float4*>(output));
    else if (name == "meshCullProbe")
        hipLaunchKernelGGL(k_cull_probe, dim3(blocks(count)), dim3(256), 0, s, p, count, uniforms[1],
                           static_cast<const float4*>(input), static_cast<uint32_t*>(output));
    else if (name == "meshClusterProbe")
        hipLaunchKernelGGL(k_cluster_probe, dim3(blocks(count)), dim3(256), 0, s, p, count, uniforms[1], uniforms[2],
                           static_cast<uint32_t*>(output));
    else if (name == "meshPairBoundsProbe")
        hipLaunchKernelGGL(k_pair_bounds_probe, dim3(blocks(count)), dim3(256), 0, s, p, count, uniforms[1],
                           uniforms[2], static_cast<float4*>(output));
    else if (name == "meshCullFastProbe")  // the collection stage's batched hull tests
        hipLaunchKernelGGL(k_cull_fast_probe, dim3(blocks(count)), dim3(256), 0, s, p, W{}, count, uniforms[1],
                           static_cast<const float4*>(input), static_cast<uint32_t*>(output));
    else if (name == "meshClusterFastProbe")
        hipLaunchKernelGGL(k_cluster_fast_probe, dim3(blocks(count)), dim3(256), 0, s, p, W{}, count, uniforms[1],
                           uniforms[2], static_cast<uint32_t*>(output));
    else {
        std::snprintf(g_error, sizeof(g_error), "unknown probe %s", kernel);
        return SAD_E_ARGUMENT;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, kernel);
    return SAD_OK;
}

int sad_hip_nn_configure(void* scratch, int enabled, int audit, const void* weights, const void* blocks,
                         uint32_t templates) {
    if (!scratch) return SAD_E_ARGUMENT;
    if (enabled && (!weights || !blocks || templates == 0)) return SAD_E_ARGUMENT;
    auto* sc = static_cast<ScratchImpl*>(scratch);
    if (!sc->w.nnStats) {
        HIP_OK(hipMalloc(reinterpret_cast<void**>(&sc->w.nnStats), 3 * LINE * sizeof(uint32_t)));
        sc->bytes += 3 * LINE * sizeof(uint32_t);
    }
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemset(sc->w.nnStats, 0, 3 * LINE * sizeof(uint32_t)));
    sc->w.nnEnabled = enabled ? 1u : 0u;
    sc->w.nnAudit = audit ? 1u : 0u;
    sc->w.nnW = static_cast<const uint16_t*>(weights);
    sc->w.nnF = static_cast<const float*>(blocks);
    sc->w.nnTemplates = enabled ? templates : 0u;
    return SAD_OK;
}

int sad_hip_nn_grad(const void* points, const float* labels, const uint32_t* tstart, const uint32_t* tcount,
                    uint32_t templates, uint32_t slices, const void* weights, const float* blocks, float* partial,
                    float* loss, float loss_scale, void* stream) {
    if (!points || !labels || !tstart || !tcount || !weights || !blocks || !partial || !loss || templates == 0 ||
        slices == 0)
        return SAD_E_ARGUMENT;
    NNTrainArgs a{static_cast<const float4*>(points), labels, tstart, tcount, static_cast<const uint16_t*>(weights),
                  blocks, partial, loss, slices, loss_scale};
    hipLaunchKernelGGL(k_nn_grad, dim3(slices, templates), dim3(128), 0, static_cast<hipStream_t>(stream), a);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, "nn_grad");
}

int sad_hip_nn_adam(float* master, float* m1, float* m2, const float* partial, float* grad, uint32_t slices,
                    uint32_t templates, float lr, float beta1, float beta2, float eps, uint32_t step, void* stream) {
    if (!master || !m1 || !m2 || !partial || !grad || step == 0) return SAD_E_ARGUMENT;
    float bc1 = 1.0f - powf(beta1, float(step)), bc2 = 1.0f - powf(beta2, float(step));
    size_t total = size_t(templates) * nn::PARAMS;
    unsigned grid = unsigned(std::min<size_t>((total + 255) / 256, 4096));
    hipLaunchKernelGGL(k_nn_adam, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), master, m1, m2, partial,
                       grad, slices, templates, lr, beta1, beta2, eps, bc1, bc2);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, "nn_adam");
}

int sad_hip_nn_pack(const float* master, void* weights, float* blocks, uint32_t templates, void* stream) {
    if (!master || !weights || !blocks || templates == 0) return SAD_E_ARGUMENT;
    hipLaunchKernelGGL(k_nn_pack, dim3(16, templates), dim3(256), 0, static_cast<hipStream_t>(stream), master,
                       static_cast<uint16_t*>(weights), blocks, templates);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, "nn_pack");
}

int sad_hip_nn_infer(const void* points, const uint32_t* tile_template, uint32_t tiles, const void* weights,
                     const float* blocks, float* out, void* stream) {
    if (tiles == 0) return SAD_OK;
    if (!points || !tile_template || !weights || !blocks || !out) return SAD_E_ARGUMENT;
    hipLaunchKernelGGL(k_nn_infer, dim3(tiles), dim3(64), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float4*>(points), tile_template, static_cast<const uint16_t*>(weights), blocks,
                       out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, "nn_infer");
}

int sad_hip_scan_kernel(const char* kernel, uint32_t n, const uint32_t* input, uint32_t* output, uint32_t* aux,
                        void* stream) {
    if (!kernel || !output || !aux) return SAD_E_ARGUMENT;
    if (n == 0) return SAD_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (std::strcmp(kernel, "blockScan") == 0) {
        if (!input) return SAD_E_ARGUMENT;
        hipLaunchKernelGGL(k_block_scan, dim3(blocks(n)), dim3(256), 0, s, n, input, output, aux);
    } else if (std::strcmp(kernel, "addBlocks") == 0) {
        hipLaunchKernelGGL(k_add_blocks, dim3(blocks(n)), dim3(256), 0, s, n, output, aux);
    } else {
        std::snprintf(g_error, sizeof(g_error), "unknown scan kernel %s", kernel);
        return SAD_E_ARGUMENT;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, kernel);
}

// normalProbe (normal.slang:12-20): rows (depth, area, normal velocity, -) ->
// (force, loss power, 0, 0) by normalResponse (normal.slang:3-11).
__global__ void k_normal_probe(uint32_t n, float stiffness, float viscosity, float weight, const float4* in,
                               float4* out) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const float4 data = in[id];
    float force = 0.0f, lossPower = 0.0f;
    if (data.x > 0.0f) {
        const float elastic = weight * stiffness * data.y * data.x;
        force = viscosity == 0.0f ? elastic : fmaxf(0.0f, elastic - weight * viscosity * data.y * data.z);
        lossPower = fmaxf(0.0f, (elastic - force) * data.z);
    }
    out[id] = make_float4(force, lossPower, 0.0f, 0.0f);
}
int sad_hip_normal_probe(uint32_t n, float stiffness, float viscosity, float weight, const void* input,
                         void* output, void* stream) {
    if (n == 0) return SAD_OK;
    if (!input || !output) return SAD_E_ARGUMENT;
    hipLaunchKernelGGL(k_normal_probe, dim3(blocks(n)), dim3(256), 0, static_cast<hipStream_t>(stream), n, stiffness,
                       viscosity, weight, static_cast<const float4*>(input), static_cast<float4*>(output));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, "normalProbe");
}

// The adapter's compiled-kernel check (solver.slang:4-10): b = a * a + 1.
__global__ void k_smoke(uint32_t n, const float* a, float* b) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) b[id] = a[id] * a[id] + 1.0f;
}
int sad_hip_smoke(uint32_t n, const void* a, void* b, void* stream) {
    if (n == 0) return SAD_OK;
    if (!a || !b) return SAD_E_ARGUMENT;
    hipLaunchKernelGGL(k_smoke, dim3(blocks(n)), dim3(256), 0, static_cast<hipStream_t>(stream), n,
                       static_cast<const float*>(a), static_cast<float*>(b));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, "smoke");
}

int sad_hip_reduce_kernel(const char* kernel, uint32_t n, uint32_t walls, uint32_t output_count, const void* input,
                          uint64_t input_bytes, void* output, uint64_t output_bytes, void* stream) {
    if (!kernel || !input || !output) return SAD_E_ARGUMENT;
    int mode = -1;
    uint64_t rows = 0;
    if (std::strcmp(kernel, "reduceRows") == 0) mode = 0, rows = 8;
    else if (std::strcmp(kernel, "reduceFabric") == 0) mode = 1, rows = 2;
    else if (std::strcmp(kernel, "reduceWallRows") == 0) mode = 2, rows = 3;
    if (mode < 0) {
        std::snprintf(g_error, sizeof(g_error), "unknown reduction kernel %s", kernel);
        return SAD_E_ARGUMENT;
    }
    if (n == 0) return SAD_OK;
    const uint32_t groups = blocks(n);
    if (mode != 2) walls = 1, output_count = groups;
    if (walls == 0 || walls > 65535 || output_count < groups ||
        input_bytes < 16 * rows * uint64_t(walls) * n || output_bytes < 16 * rows * uint64_t(walls) * output_count) {
        std::snprintf(g_error, sizeof(g_error), "%s: %u walls, %u records, %u outputs do not fit the buffers", kernel,
                      walls, n, output_count);
        return SAD_E_ARGUMENT;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const float4* in = static_cast<const float4*>(input);
    float4* out = static_cast<float4*>(output);
    if (mode == 0) hipLaunchKernelGGL(k_reduce_rows<0>, dim3(groups), dim3(256), 0, s, n, output_count, in, out);
    else if (mode == 1) hipLaunchKernelGGL(k_reduce_rows<1>, dim3(groups), dim3(256), 0, s, n, output_count, in, out);
    else hipLaunchKernelGGL(k_reduce_rows<2>, dim3(groups, walls), dim3(256), 0, s, n, output_count, in, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, kernel);
}

int sad_hip_nn_sample(const uint32_t* uniforms, void* const* buffers, const uint64_t* bytes, const double* values,
                      const void* triangles, const float* cdf, const uint32_t* tri_start, const uint32_t* tri_count,
                      const void* box, void* points, float* labels, void* stream) {
    if (!uniforms || !triangles || !cdf || !tri_start || !tri_count || !box || !points || !labels)
        return SAD_E_ARGUMENT;
    P p{};
    int rc = bind(p, buffers, bytes, values);
    if (rc) return rc;
    const uint32_t templates = uniforms[0], count = uniforms[1];
    if (templates == 0 || count == 0) return SAD_OK;
    NNSampleArgs a{static_cast<const float4*>(triangles), cdf, tri_start, tri_count, static_cast<const float4*>(box),
                   static_cast<float4*>(points), labels, count,
                   (uint64_t(uniforms[2]) << 32) | uniforms[3]};
    hipLaunchKernelGGL(k_nn_sample, dim3(blocks(count), templates), dim3(256), 0, static_cast<hipStream_t>(stream), p, a);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, "nn_sample");
}

int sad_hip_mfma_probe(const void* a, const void* b, float* d, int k, void* stream) {
    if (!a || !b || !d || (k != 16 && k != 32)) return SAD_E_ARGUMENT;
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint16_t*>(a), static_cast<const uint16_t*>(b), d, k == 32 ? 1 : 0);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAD_OK : hip_fail(e, "mfma_probe");
}

const char* sad_hip_last_error(void) { return g_error; }
int sad_hip_abi_version(void) { return 1; }

}  // extern "C"
