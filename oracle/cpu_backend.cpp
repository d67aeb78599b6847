// ORACLE -- test infrastructure only.  Nothing in the product imports, links or
// runs this file: only tests/, the smoke check and bench.py's cpu_baseline leg do,
// as the checker.
//
// A serial C++17 restatement of the reference's CPU backend for the surface-mesh
// step: the Slang kernels of src/sand_atlas_dem/shaders/solver.slang with the CPU
// overrides spliced in by src/sand_atlas_dem/cpu.py:46-103 (FP64 centerBranch,
// cached-box-distance meshDistance, per-directed-pair rejection plane, serial
// pair-order collector cpuCollectMesh, non-atomic atomics from
// shaders/cpu_scheduling.slang:3-99).  FP32 arithmetic follows the generated C++:
// dot() accumulates from 0.0f left to right, length = sqrtf(dot), normalize =
// x * (1/sqrtf(dot)), min/max = fminf/fmaxf, sinf/cosf/expf as the correctly rounded
// FP32 value (below; evidence/.../instrumented-collector.cpp:869-1027).  Build with
// -O3 -fno-fast-math -ffp-contract=off (cpu.py:323-333).  Built with
// -DSAD_ORACLE_FP64 (liboracle64.so) every FP32 quantity is FP64 instead: the same
// algorithm at double precision, used only as the yardstick for FP32 drift.
//
// Exported C ABI (binding order of include/sand_dem_hip.h, same as the
// reference's generated cpu_dispatch):
//   int oracle_dispatch(const char* kernel, uint32_t count, void** buffers,
//                       const uint64_t* bytes, const double* values);
//   int oracle_probe(const char* kernel, const uint32_t* uniforms, void** buffers,
//                    const uint64_t* bytes, const double* values,
//                    const void* input, void* output);
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../include/sand_dem_hip.h"

namespace {

// Arithmetic type of the restatement: FP32 (the reference's, and the HIP path's)
// unless built with -DSAD_ORACLE_FP64, which runs the same algorithm in FP64 on
// the same FP32 inputs (liboracle64.so: the drift yardstick).  Buffers the FP64
// build reads hold doubles in place of floats, element for element.
#ifdef SAD_ORACLE_FP64
using real = double;
#else
using real = float;
#endif

struct float3 { real x, y, z; };
struct float4 { real x, y, z, w; };
struct uint4 { uint32_t x, y, z, w; };
struct uint2 { uint32_t x, y; };
struct double3 { double x, y, z; };

inline float3 f3(real x, real y, real z) { return {x, y, z}; }
inline float3 xyz(const float4& a) { return {a.x, a.y, a.z}; }
inline float4 f4(const float3& a, real w) { return {a.x, a.y, a.z, w}; }
inline float3 operator+(float3 a, float3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float3 operator-(float3 a, float3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float3 operator*(float3 a, float3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float3 operator/(float3 a, float3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline float3 operator*(real s, float3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float3 operator*(float3 a, real s) { return {a.x * s, a.y * s, a.z * s}; }
inline float3 operator/(float3 a, real s) { return {a.x / s, a.y / s, a.z / s}; }
inline float3 operator-(float3 a) { return {-a.x, -a.y, -a.z}; }
inline float4 operator+(float4 a, float4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline float4 operator-(float4 a, float4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
inline float4 operator*(float4 a, real s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline real get(const float3& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline void set(float3& a, int i, real v) { (i == 0 ? a.x : (i == 1 ? a.y : a.z)) = v; }

inline real dot(float3 a, float3 b) {
    real r = real(0.0);
    r = r + a.x * b.x;
    r = r + a.y * b.y;
    r = r + a.z * b.z;
    return r;
}
inline real dot4(float4 a, float4 b) {
    real r = real(0.0);
    r = r + a.x * b.x;
    r = r + a.y * b.y;
    r = r + a.z * b.z;
    r = r + a.w * b.w;
    return r;
}
inline float3 cross(float3 l, float3 r) {
    return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x};
}
inline real length(float3 a) { return std::sqrt(dot(a, a)); }
inline float3 fabs3(float3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline float3 fmax3(float3 a, float3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline float3 fmin3(float3 a, float3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline float3 floor3(float3 a) { return {std::floor(a.x), std::floor(a.y), std::floor(a.z)}; }
inline bool finite3(float3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
// FP32 transcendentals: the reference calls the platform's ::sinf/::cosf/::expf
// (instrumented-collector.cpp:877-943), whose last bit differs between libms
// (glibc is off by one ulp on ~1% of arguments).  The oracle fixes them to the
// correctly rounded value, obtained from the FP64 libm functions.
inline real sin32(real x) { return real(std::sin(double(x))); }
inline real cos32(real x) { return real(std::cos(double(x))); }
inline real exp32(real x) { return real(std::exp(double(x))); }

inline float3 normalize3(float3 a) {
    real s = real(1.0) / std::sqrt(dot(a, a));
    return a * s;
}
inline float4 normalize4(float4 a) {
    real s = real(1.0) / std::sqrt(dot4(a, a));
    return a * s;
}

constexpr uint32_t EMPTY = 0xffffffffu;
constexpr real EPS = real(1.1920928955078125e-7);

struct Config {
    uint32_t activeTimestep, N, shapes, pairCapacity, contactCapacity, binCapacity, binCount;
    uint32_t bins[3];
    uint32_t hashSize;
    real dt, kn, kt, mu, normalViscosity, skin, drag, maxOverlap;
    uint32_t historyStep, meshGeometry, periodic, wallCount, wallBoxEnabled, sourceSurfaceNormal;
    uint32_t threads;
};

struct R {
    Config c;
    uint32_t* control;
    float4 *cell, *position, *positionRemainder, *velocityRemainder, *momentumRemainder;
    float4 *rotationRemainder, *referencePositionRemainder, *cellLog, *accumulatedRemainder;
    float4 *velocity, *rotation, *momentum, *force, *torque, *referencePosition, *referenceCell;
    uint32_t* shapeId;
    float4 *body, *invI;
    uint4* sampleInfo;
    float4* sample;
    uint4 *meshRoots, *meshHullInfo;
    float4 *meshPlanes, *meshMin, *meshMax;
    uint4* meshLinks;
    float4 *meshTriangles, *meshNormals;
    uint4* meshFrontier;
    uint32_t* meshPairMask;
    float4* meshPairTransform;
    uint4 *meshClusterInfo, *meshClusterRanges;
    float4 *meshClusterMin, *meshClusterMax;
    uint32_t* meshSampleIndices;
    uint32_t *binCounts, *binParticles, *pairCounts, *pairOffsets;
    uint2* pairs;
    uint32_t *contactCounts, *contactOffsets;
    uint4* meta;
    float4 *pointDepth, *normalArea, *slip, *contactForce;
    real* normalLossPower;
    float4 *normalDamping, *forceVelocity, *forceOmega;
    uint4* oldMeta;
    float4 *oldSlip, *oldNormal;
    uint32_t *oldHash, *hash, *heads, *next;
    float4 *particleStress, *timestepRows, *particleDiagnostic, *accumulated, *diagnosticRows, *fabricRows;
    float4 *wallPlane, *wallRemainder, *wallVelocity, *wallMaterial, *wallRows, *wallAccumulated;
    uint32_t* wallBoxIds;
    uint64_t metaCount;  // capacity of the meta/pointDepth/normalArea views
    uint4* grids;        // SDF path: per template (first value, nx, ny, nz)
    float4* gridOrigin;  // per template (origin xyz, spacing)
    real* distance;
    uint64_t distanceCount;
};

template <class T>
T* at(void** b, int i) { return reinterpret_cast<T*>(b[i]); }

R bind(void** b, const uint64_t* bytes, const double* v) {
    R r{};
    Config& c = r.c;
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
    c.dt = real(v[SAD_V_DT]);
    c.kn = real(v[SAD_V_KN]);
    c.kt = real(v[SAD_V_KT]);
    c.mu = real(v[SAD_V_MU]);
    c.normalViscosity = real(v[SAD_V_NORMAL_VISCOSITY]);
    c.skin = real(v[SAD_V_SKIN]);
    c.drag = real(v[SAD_V_DRAG]);
    c.maxOverlap = real(v[SAD_V_MAX_OVERLAP]);
    c.historyStep = uint32_t(v[SAD_V_HISTORY_STEP]);
    c.meshGeometry = uint32_t(v[SAD_V_MESH_GEOMETRY]);
    c.periodic = uint32_t(v[SAD_V_PERIODIC]);
    c.wallCount = uint32_t(v[SAD_V_WALL_COUNT]);
    c.wallBoxEnabled = uint32_t(v[SAD_V_WALL_BOX_ENABLED]);
    c.sourceSurfaceNormal = uint32_t(v[SAD_V_SOURCE_SURFACE_NORMAL]);
    c.threads = uint32_t(v[SAD_V_THREADS]);
    r.control = at<uint32_t>(b, SAD_CONTROL);
    r.cell = at<float4>(b, SAD_CELL);
    r.position = at<float4>(b, SAD_POSITION);
    r.positionRemainder = at<float4>(b, SAD_POSITION_REMAINDER);
    r.velocityRemainder = at<float4>(b, SAD_VELOCITY_REMAINDER);
    r.momentumRemainder = at<float4>(b, SAD_MOMENTUM_REMAINDER);
    r.rotationRemainder = at<float4>(b, SAD_ROTATION_REMAINDER);
    r.referencePositionRemainder = at<float4>(b, SAD_REFERENCE_POSITION_REMAINDER);
    r.cellLog = at<float4>(b, SAD_CELL_LOG);
    r.accumulatedRemainder = at<float4>(b, SAD_ACCUMULATED_REMAINDER);
    r.velocity = at<float4>(b, SAD_VELOCITY);
    r.rotation = at<float4>(b, SAD_ROTATION);
    r.momentum = at<float4>(b, SAD_MOMENTUM);
    r.force = at<float4>(b, SAD_FORCE);
    r.torque = at<float4>(b, SAD_TORQUE);
    r.referencePosition = at<float4>(b, SAD_REFERENCE_POSITION);
    r.referenceCell = at<float4>(b, SAD_REFERENCE_CELL);
    r.shapeId = at<uint32_t>(b, SAD_SHAPE_ID);
    r.body = at<float4>(b, SAD_BODY);
    r.invI = at<float4>(b, SAD_INV_I);
    r.sampleInfo = at<uint4>(b, SAD_SAMPLE_INFO);
    r.sample = at<float4>(b, SAD_SAMPLE);
    r.meshRoots = at<uint4>(b, SAD_MESH_ROOTS);
    r.meshHullInfo = at<uint4>(b, SAD_MESH_HULL_INFO);
    r.meshPlanes = at<float4>(b, SAD_MESH_PLANES);
    r.meshMin = at<float4>(b, SAD_MESH_MIN);
    r.meshMax = at<float4>(b, SAD_MESH_MAX);
    r.meshLinks = at<uint4>(b, SAD_MESH_LINKS);
    r.meshTriangles = at<float4>(b, SAD_MESH_TRIANGLES);
    r.meshNormals = at<float4>(b, SAD_MESH_NORMALS);
    r.meshFrontier = at<uint4>(b, SAD_MESH_FRONTIER);
    r.meshPairMask = at<uint32_t>(b, SAD_MESH_PAIR_MASK);
    r.meshPairTransform = at<float4>(b, SAD_MESH_PAIR_TRANSFORM);
    r.meshClusterInfo = at<uint4>(b, SAD_MESH_CLUSTER_INFO);
    r.meshClusterRanges = at<uint4>(b, SAD_MESH_CLUSTER_RANGES);
    r.meshClusterMin = at<float4>(b, SAD_MESH_CLUSTER_MIN);
    r.meshClusterMax = at<float4>(b, SAD_MESH_CLUSTER_MAX);
    r.meshSampleIndices = at<uint32_t>(b, SAD_MESH_SAMPLE_INDICES);
    r.binCounts = at<uint32_t>(b, SAD_BIN_COUNTS);
    r.binParticles = at<uint32_t>(b, SAD_BIN_PARTICLES);
    r.pairCounts = at<uint32_t>(b, SAD_PAIR_COUNTS);
    r.pairOffsets = at<uint32_t>(b, SAD_PAIR_OFFSETS);
    r.pairs = at<uint2>(b, SAD_PAIRS);
    r.contactCounts = at<uint32_t>(b, SAD_CONTACT_COUNTS);
    r.contactOffsets = at<uint32_t>(b, SAD_CONTACT_OFFSETS);
    r.grids = at<uint4>(b, SAD_GRIDS);
    r.gridOrigin = at<float4>(b, SAD_GRID_ORIGIN);
    r.distance = at<real>(b, SAD_DISTANCE);
    r.distanceCount = bytes[SAD_DISTANCE] / 4;
    r.meta = at<uint4>(b, SAD_META);
    r.pointDepth = at<float4>(b, SAD_POINT_DEPTH);
    r.normalArea = at<float4>(b, SAD_NORMAL_AREA);
    r.slip = at<float4>(b, SAD_SLIP);
    r.contactForce = at<float4>(b, SAD_CONTACT_FORCE);
    r.normalLossPower = at<real>(b, SAD_NORMAL_LOSS_POWER);
    r.normalDamping = at<float4>(b, SAD_NORMAL_DAMPING);
    r.forceVelocity = at<float4>(b, SAD_FORCE_VELOCITY);
    r.forceOmega = at<float4>(b, SAD_FORCE_OMEGA);
    r.oldMeta = at<uint4>(b, SAD_OLD_META);
    r.oldSlip = at<float4>(b, SAD_OLD_SLIP);
    r.oldNormal = at<float4>(b, SAD_OLD_NORMAL);
    r.oldHash = at<uint32_t>(b, SAD_OLD_HASH);
    r.hash = at<uint32_t>(b, SAD_HASH);
    r.heads = at<uint32_t>(b, SAD_HEADS);
    r.next = at<uint32_t>(b, SAD_NEXT);
    r.particleStress = at<float4>(b, SAD_PARTICLE_STRESS);
    r.timestepRows = at<float4>(b, SAD_TIMESTEP_ROWS);
    r.particleDiagnostic = at<float4>(b, SAD_PARTICLE_DIAGNOSTIC);
    r.diagnosticRows = at<float4>(b, SAD_DIAGNOSTIC_ROWS);
    r.fabricRows = at<float4>(b, SAD_FABRIC_ROWS);
    r.accumulated = at<float4>(b, SAD_ACCUMULATED);
    r.wallPlane = at<float4>(b, SAD_WALL_PLANE);
    r.wallRemainder = at<float4>(b, SAD_WALL_REMAINDER);
    r.wallVelocity = at<float4>(b, SAD_WALL_VELOCITY);
    r.wallMaterial = at<float4>(b, SAD_WALL_MATERIAL);
    r.wallBoxIds = at<uint32_t>(b, SAD_WALL_BOX_IDS);
    r.wallRows = at<float4>(b, SAD_WALL_ROWS);
    r.wallAccumulated = at<float4>(b, SAD_WALL_ACCUMULATED);
    r.metaCount = bytes[SAD_META] / 16;
    return r;
}

// ---------------------------------------------------------------- helpers (solver.slang:104-186)
inline float3 qrot(float4 q, float3 v) {  // solver.slang:104
    float3 u = xyz(q);
    return v + real(2.0) * cross(u, cross(u, v) + q.w * v);
}
inline float3 qinv(float4 q, float3 v) { return qrot({-q.x, -q.y, -q.z, q.w}, v); }  // :105
inline float4 qmul(float4 a, float4 b) {  // :106
    float3 v = a.w * xyz(b) + b.w * xyz(a) + cross(xyz(a), xyz(b));
    return f4(v, a.w * b.w - dot(xyz(a), xyz(b)));
}
// rounded3/4 XOR with control[15] (always zero): identity on the CPU (:109-110).
inline void compensated3(float3& high, float3& low, float3 inc) {  // :111-114
    float3 y = inc + low;
    float3 next = high + y;
    float3 moved = next - high;
    low = y - moved;
    high = next;
}
inline void compensated4(float4& high, float4& low, float4 inc) {  // :115-118
    float4 y = inc + low;
    float4 next = high + y;
    float4 moved = next - high;
    low = y - moved;
    high = next;
}
inline real expm1Stable(real x) {  // :119-122
    if (std::fabs(x) < real(0.01))
        return x * (real(1.0) + x * (real(0.5) + x * (real(1.0) / real(6.0) + x * (real(1.0) / real(24.0) + x / real(120.0)))));
    return exp32(x) - real(1.0);
}
inline float3 expm1Stable3(float3 x) { return {expm1Stable(x.x), expm1Stable(x.y), expm1Stable(x.z)}; }
inline float4 rotationIncrement(float3 w, real dt, float4 q) {  // :124-129
    real half = real(0.5) * length(w) * dt;
    real coefficient = std::fabs(half) < real(0.001) ? real(0.5) * dt * (real(1.0) - half * half / real(6.0))
                                                 : sin32(half) / length(w);
    real cm1 = -real(2.0) * sin32(real(0.5) * half) * sin32(real(0.5) * half);
    return qmul(f4(coefficient * w, cm1), q);
}
inline void accumulate(const R& r, uint32_t id, float4 inc) {  // :130-134
    float4 high = r.accumulated[id], low = r.accumulatedRemainder[id];
    compensated4(high, low, inc);
    r.accumulated[id] = high;
    r.accumulatedRemainder[id] = low;
}
// Periodicity: Config `periodic` is 0 (none) or 1 (every axis) as the reference
// sends it; 8 | mask (x = 1, y = 2, z = 4) selects axes (no reference counterpart:
// walls across one axis of an otherwise periodic cell).  0 and 1 take exactly the
// reference's paths.
inline bool periodicAxis(const Config& c, int a) { return c.periodic == 1u || (c.periodic >= 8u && ((c.periodic >> a) & 1u)); }
inline bool anyPeriodic(const Config& c) { return c.periodic == 1u || (c.periodic >= 8u && (c.periodic & 7u) != 0u); }
// CPU override (cpu_scheduling.slang:10-15): the branch subtraction in FP64.
inline float3 centerBranch(const R& r, uint32_t i, uint32_t j) {
    float4 c = r.cell[0], pj = r.position[j], pi = r.position[i];
    float4 rj = r.positionRemainder[j], ri = r.positionRemainder[i];
    double bx = c.x, by = c.y, bz = c.z;
    double dx = (double(pj.x) - double(pi.x)) + (double(rj.x) - double(ri.x));
    double dy = (double(pj.y) - double(pi.y)) + (double(rj.y) - double(ri.y));
    double dz = (double(pj.z) - double(pi.z)) + (double(rj.z) - double(ri.z));
    if (periodicAxis(r.c, 0)) dx = dx - std::floor(dx / bx + 0.5) * bx;
    if (periodicAxis(r.c, 1)) dy = dy - std::floor(dy / by + 0.5) * by;
    if (periodicAxis(r.c, 2)) dz = dz - std::floor(dz / bz + 0.5) * bz;
    return {real(dx), real(dy), real(dz)};
}
inline float3 omega(const R& r, uint32_t i) {  // :154-158
    float3 b = qinv(r.rotation[i], xyz(r.momentum[i]) + xyz(r.momentumRemainder[i]));
    float3 w = {dot(xyz(r.invI[3 * i]), b), dot(xyz(r.invI[3 * i + 1]), b), dot(xyz(r.invI[3 * i + 2]), b)};
    return qrot(r.rotation[i], w);
}
inline real rotationalEnergy(const R& r, uint32_t i, float3 l) {  // :159-163
    float3 b = qinv(r.rotation[i], l);
    float3 w = {dot(xyz(r.invI[3 * i]), b), dot(xyz(r.invI[3 * i + 1]), b), dot(xyz(r.invI[3 * i + 2]), b)};
    return real(0.5) * dot(b, w);
}
inline void accumulateNormalLoss(const R& r, uint32_t id, real dt) {  // :164-169
    float4 value = r.normalDamping[id];
    float3 high = {value.x, 0, 0}, low = {value.y, 0, 0};
    compensated3(high, low, {value.z * dt, 0, 0});
    value.x = high.x;
    value.y = low.x;
    r.normalDamping[id] = value;
}
inline real cellPower(const R& r, uint32_t i) {  // :176-180
    float3 diagonal = {r.particleStress[3 * i].x, r.particleStress[3 * i + 1].y, r.particleStress[3 * i + 2].z};
    float3 v = xyz(r.velocity[i]);
    return -dot(diagonal + r.body[i].y * v * v, xyz(r.cell[1]));
}
inline float3 image(const R& r, float3 d) {  // :181
    float3 box = xyz(r.cell[0]);
    if (!anyPeriodic(r.c)) return d;
    float3 w = d - floor3(d / box + f3(real(0.5), real(0.5), real(0.5))) * box;
    return {periodicAxis(r.c, 0) ? w.x : d.x, periodicAxis(r.c, 1) ? w.y : d.y, periodicAxis(r.c, 2) ? w.z : d.z};
}
inline void fail(const R& r, uint32_t bit) { r.control[0] |= bit; }  // :183 (serial)
inline void binIndex(const R& r, uint32_t i, uint32_t b[3]) {  // :184
    float3 g = floor3(xyz(r.position[i]) / xyz(r.cell[0]) * f3(real(r.c.bins[0]), real(r.c.bins[1]), real(r.c.bins[2])));
    b[0] = std::min(uint32_t(g.x), r.c.bins[0] - 1);
    b[1] = std::min(uint32_t(g.y), r.c.bins[1] - 1);
    b[2] = std::min(uint32_t(g.z), r.c.bins[2] - 1);
}
inline uint32_t binFlat(const R& r, const uint32_t b[3]) { return (b[0] * r.c.bins[1] + b[1]) * r.c.bins[2] + b[2]; }
inline uint32_t modIndex(int a, uint32_t n) { return uint32_t((a % int(n) + int(n)) % int(n)); }

// ---------------------------------------------------------------- mesh queries (solver.slang:203-343)
inline float3 closestTriangle(float3 p, float3 a, float3 b, float3 c, uint32_t& feature) {  // :203-219
    float3 ab = b - a, ac = c - a, ap = p - a;
    real d1 = dot(ab, ap), d2 = dot(ac, ap);
    feature = 4;
    if (d1 <= 0 && d2 <= 0) return a;
    float3 bp = p - b;
    real d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { feature = 5; return b; }
    real vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { feature = 1; return a + (d1 / (d1 - d3)) * ab; }
    float3 cp = p - c;
    real d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { feature = 6; return c; }
    real vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { feature = 3; return a + (d2 / (d2 - d6)) * ac; }
    real va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        feature = 2;
        return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    }
    real inv = real(1.0) / (va + vb + vc);
    feature = 0;
    return a + ab * (vb * inv) + ac * (vc * inv);
}
inline real meshBoxDistance(const R& r, uint32_t node, float3 p) {  // :220-223
    float3 d = fmax3(fmax3(xyz(r.meshMin[node]) - p, p - xyz(r.meshMax[node])), f3(0, 0, 0));
    return dot(d, d);
}
// CPU override (cpu_scheduling.slang:18-48).
float4 meshDistance(const R& r, uint32_t shape, float3 p, bool surfaceNormal = false) {
    uint32_t root = r.meshRoots[shape].x;
    uint32_t pending[32];
    real pendingDistance[32];
    uint32_t top = 1;
    pending[0] = root;
    pendingDistance[0] = meshBoxDistance(r, root, p);
    real best = real(1e30);
    float3 nearest = {0, 0, 0};
    uint32_t nearestNormal = 0;
    bool found = false;
    while (top > 0) {
        uint32_t slot = --top, node = pending[slot];
        if (pendingDistance[slot] > best) continue;
        uint4 link = r.meshLinks[node];
        if (link.z == 0) {
            uint32_t a = node + 1, b = r.meshLinks[a].x;
            real da = meshBoxDistance(r, a, p), db = meshBoxDistance(r, b, p);
            if (da < db) { std::swap(a, b); std::swap(da, db); }
            if (top + 2 > 32) { fail(r, 1024); return {0, 0, 0, real(1e20)}; }
            if (da <= best) { pending[top] = a; pendingDistance[top++] = da; }
            if (db <= best) { pending[top] = b; pendingDistance[top++] = db; }
            continue;
        }
        for (uint32_t j = 0; j < link.z; j++) {
            uint32_t triangle = link.y + j, feature;
            float3 q = closestTriangle(p, xyz(r.meshTriangles[3 * triangle]), xyz(r.meshTriangles[3 * triangle + 1]),
                                       xyz(r.meshTriangles[3 * triangle + 2]), feature);
            real distance = dot(p - q, p - q);
            if (distance < best) { best = distance; nearest = q; nearestNormal = 7 * triangle + feature; found = true; }
        }
    }
    float3 pseudo = found ? xyz(r.meshNormals[nearestNormal]) : f3(0, 0, 0);
    real distance = std::sqrt(best);
    real side = dot(p - nearest, pseudo) < 0 ? -real(1.0) : real(1.0);
    float3 normal = surfaceNormal ? pseudo : (distance > real(1e-12) ? side * (p - nearest) / distance : pseudo);
    return f4(normal, side * distance);
}
inline uint32_t dominantGroup(float3 direction) {
    float3 a = fabs3(direction);
    uint32_t axis = a.x >= a.y && a.x >= a.z ? 0 : (a.y >= a.z ? 1 : 2);
    return 2 * axis + (get(direction, axis) >= 0 ? 1 : 0);
}
bool meshMayContain(const R& r, uint32_t shape, float3 p) {  // :259-272
    uint32_t root = r.meshRoots[shape].x;
    float3 low = xyz(r.meshMin[root]), high = xyz(r.meshMax[root]);
    if (p.x < low.x || p.y < low.y || p.z < low.z || p.x > high.x || p.y > high.y || p.z > high.z) return false;
    float3 direction = (p - real(0.5) * (low + high)) / (high - low);
    uint4 info = r.meshHullInfo[6 * shape + dominantGroup(direction)];
    for (uint32_t i = 0; i < info.y; i++) {
        float4 plane = r.meshPlanes[info.x + i];
        if (dot(xyz(plane), p) + plane.w > 0) return false;
    }
    return true;
}
bool meshBoxMayContact(const R& r, uint32_t shape, float3 low, float3 high) {  // :274-288
    uint32_t root = r.meshRoots[shape].x;
    float3 tl = xyz(r.meshMin[root]), th = xyz(r.meshMax[root]);
    if (high.x < tl.x || high.y < tl.y || high.z < tl.z || low.x > th.x || low.y > th.y || low.z > th.z) return false;
    float3 center = real(0.5) * (low + high), half = real(0.5) * (high - low);
    float3 direction = (center - real(0.5) * (tl + th)) / (th - tl);
    uint4 info = r.meshHullInfo[6 * shape + dominantGroup(direction)];
    for (uint32_t i = 0; i < info.y; i++) {
        float4 plane = r.meshPlanes[info.x + i];
        if (dot(xyz(plane), center) + plane.w - dot(fabs3(xyz(plane)), half) > 0) return false;
    }
    return true;
}
inline real maxabs(float3 a) { return std::fmax(std::fmax(std::fabs(a.x), std::fabs(a.y)), std::fabs(a.z)); }
bool meshClusterMayContact(const R& r, uint32_t src, uint32_t dst, uint32_t cluster) {  // :290-305
    float3 sl = xyz(r.meshClusterMin[cluster]), sh = xyz(r.meshClusterMax[cluster]);
    float3 branch = centerBranch(r, src, dst);
    float3 low = {real(1e30), real(1e30), real(1e30)}, high = {-real(1e30), -real(1e30), -real(1e30)};
    for (uint32_t corner = 0; corner < 8; corner++) {
        float3 p = {(corner & 1) ? sh.x : sl.x, (corner & 2) ? sh.y : sl.y, (corner & 4) ? sh.z : sl.z};
        p = qinv(r.rotation[dst], qrot(r.rotation[src], p * r.body[src].x) - branch) / r.body[dst].x;
        low = fmin3(low, p);
        high = fmax3(high, p);
    }
    real pad = real(64.0) * EPS * (real(1.0) + maxabs(low) + maxabs(high) + length(branch) / r.body[dst].x);
    low = low - f3(pad, pad, pad);
    high = high + f3(pad, pad, pad);
    return meshBoxMayContact(r, r.shapeId[dst], low, high);
}
// sdf (solver.slang:188-200): trilinear value and its piecewise gradient; 1e20 outside
// the grid.  lerp(x, y, s) is x + s (y - x).  A corner index beyond the distance
// buffer (inconsistent grid rows) raises flag 2 and reads as outside.
float4 sdf(const R& r, uint32_t s, float3 p) {
    uint4 meta = r.grids[s];
    float4 o = r.gridOrigin[s];
    float3 g = (p - xyz(o)) / o.w;
    const float4 outside = {0, 0, 0, real(1e20)};
    real dims[3] = {real(int64_t(meta.y) - 1), real(int64_t(meta.z) - 1), real(int64_t(meta.w) - 1)};
    real gs[3] = {g.x, g.y, g.z};
    for (int a = 0; a < 3; a++)
        if (!(gs[a] >= 0) || gs[a] >= dims[a]) return outside;
    int ax = int(std::floor(g.x)), ay = int(std::floor(g.y)), az = int(std::floor(g.z));
    float3 t = {g.x - real(ax), g.y - real(ay), g.z - real(az)};
    float3 x = f3(real(1.0), real(1.0), real(1.0)) - t;
    real value = 0;
    float3 grad = {0, 0, 0};
    for (uint32_t i = 0; i < 2; i++)
        for (uint32_t j = 0; j < 2; j++)
            for (uint32_t k = 0; k < 2; k++) {
                float3 b = {real(i), real(j), real(k)};
                float3 w = x + b * (t - x);
                uint32_t index = meta.x + ((uint32_t(ax) + i) * meta.z + (uint32_t(ay) + j)) * meta.w + (uint32_t(az) + k);
                if (index >= r.distanceCount) { fail(r, 2); return outside; }
                real f = r.distance[index];
                value += f * w.x * w.y * w.z;
                float3 d = {(real(2.0) * real(i) - 1) * w.y * w.z, (real(2.0) * real(j) - 1) * w.x * w.z,
                            (real(2.0) * real(k) - 1) * w.x * w.y};
                grad = grad + f * d / o.w;
            }
    return f4(grad, value);
}
// nodeContact (solver.slang:319-343) with the CPU rejection plane (cpu.py:59-76).
bool nodeContact(const R& r, uint32_t src, uint32_t dst, uint32_t node, float4& pd, float4& na,
                 float4 rejectionPlane = {0, 0, 0, 0}) {
    uint32_t sid = r.shapeId[src];
    float4 sp = r.sample[r.sampleInfo[sid].x + node];
    float3 arm = qrot(r.rotation[src], xyz(sp) * r.body[src].x);
    float3 branch = centerBranch(r, src, dst);
    float3 local = qinv(r.rotation[dst], arm - branch) / r.body[dst].x;
    uint32_t targetShape = r.shapeId[dst];
    float4 f;
    if (r.c.meshGeometry != 0) {
        real planePad = real(64.0) * EPS * (real(1.0) + maxabs(local));
        if (dot(xyz(rejectionPlane), local) + rejectionPlane.w > planePad) { pd = {0, 0, 0, 0}; na = {0, 0, 0, 0}; return false; }
        if (!meshMayContain(r, targetShape, local)) { pd = {0, 0, 0, 0}; na = {0, 0, 0, 0}; return false; }
        f = meshDistance(r, targetShape, local);
    } else {
        f = sdf(r, targetShape, local);
    }
    pd = f4(arm, -f.w * r.body[dst].x);
    na = {0, 0, 0, 0};
    if (pd.w <= 0) return false;
    float3 grad = qrot(r.rotation[dst], xyz(f));
    real len = length(grad);
    if (len < real(1e-7) || !std::isfinite(pd.w)) { fail(r, 2); return false; }
    na = f4(grad / len, sp.w * r.body[src].x * r.body[src].x);
    if (pd.w > r.c.maxOverlap * std::fmin(r.body[src].w, r.body[dst].w)) fail(r, 4);
    return true;
}
// cpuCollectMesh (cpu_scheduling.slang:53-91): serial append in pair order.
// Output is written through the views in `r` (meta/pointDepth/normalArea,
// control[2] as the running count, contactCounts[pair]).
void cpuCollectMesh(const R& r, uint32_t pair) {
    if (pair >= std::min(r.control[1], r.c.pairCapacity) || r.control[0] != 0) return;
    uint2 ij = r.pairs[pair];
    for (uint32_t pass = 0; pass < 2; pass++) {
        uint32_t i = pass == 0 ? ij.x : ij.y, j = pass == 0 ? ij.y : ij.x;
        float3 sourceCentre = -qinv(r.rotation[j], centerBranch(r, i, j)) / r.body[j].x;
        float4 rejectionPlane = {0, 0, 0, 0};
        real largest = -real(1e30);
        for (uint32_t group = 0; group < 6; group++) {
            uint4 hull = r.meshHullInfo[6 * r.shapeId[j] + group];
            for (uint32_t plane = 0; plane < hull.y; plane++) {
                float4 enclosure = r.meshPlanes[hull.x + plane];
                real violation = dot(xyz(enclosure), sourceCentre) + enclosure.w;
                if (violation > largest) { largest = violation; rejectionPlane = enclosure; }
            }
        }
        uint4 info = r.meshClusterInfo[r.shapeId[i]];
        uint32_t cluster = info.x;
        while (cluster < info.x + info.y) {
            bool candidate = meshClusterMayContact(r, i, j, cluster);
            uint4 range = r.meshClusterRanges[cluster];
            if (candidate) {
                for (uint32_t lane = 0; lane < range.y; lane++) {
                    uint32_t node = r.meshSampleIndices[range.x + lane];
                    float4 pd, na;
                    if (nodeContact(r, i, j, node, pd, na, rejectionPlane)) {
                        uint32_t offset = r.control[2]++;
                        if (offset >= r.c.contactCapacity) { fail(r, 32); return; }
                        r.meta[offset] = {i, j, node, 0};
                        r.pointDepth[offset] = pd;
                        r.normalArea[offset] = na;
                        r.contactCounts[pair]++;
                    }
                }
            }
            cluster = candidate ? cluster + 1 : range.z;
        }
    }
}

// ---------------------------------------------------------------- SDF contacts (cpu_scheduling.slang:102-132)
// Both grains' nodes in order, s < a the source grain's: pair-major, node order.
uint32_t pairNodeContacts(const R& r, uint32_t pair, bool emit) {
    uint2 ij = r.pairs[pair];
    uint32_t a = r.sampleInfo[r.shapeId[ij.x]].y;
    uint32_t total = a + r.sampleInfo[r.shapeId[ij.y]].y, count = 0;
    uint32_t offset = emit ? r.contactOffsets[pair] : 0;
    for (uint32_t s = 0; s < total; s++) {
        uint32_t i = s < a ? ij.x : ij.y, j = s < a ? ij.y : ij.x, node = s < a ? s : s - a;
        float4 pd, na;
        if (!nodeContact(r, i, j, node, pd, na)) continue;
        count++;
        if (!emit) continue;
        if (offset >= r.c.contactCapacity) { fail(r, 32); return count; }
        r.meta[offset] = {i, j, node, 0};
        r.pointDepth[offset] = pd;
        r.normalArea[offset] = na;
        offset++;
    }
    return count;
}
void countContacts(const R& r, uint32_t pair) {  // :102-116
    if (pair >= r.c.pairCapacity) return;
    uint32_t count = 0;
    if (pair < r.control[1] && r.control[0] == 0) count = pairNodeContacts(r, pair, false);
    r.contactCounts[pair] = count;
}
void emitContacts(const R& r, uint32_t pair) {  // :118-131
    if (pair >= std::min(r.control[1], r.c.pairCapacity) || r.control[0] != 0) return;
    pairNodeContacts(r, pair, true);
}
void contactScan(const R& r) {  // Scan over contactCounts (gpu.py:539-541): exclusive, wrapping
    uint32_t sum = 0;
    for (uint32_t pair = 0; pair < r.c.pairCapacity; pair++) {
        r.contactOffsets[pair] = sum;
        sum += r.contactCounts[pair];
    }
}
void contactTotal(const R& r) {  // solver.slang:568-574
    uint32_t n = r.c.pairCapacity;
    r.control[2] = r.contactOffsets[n - 1] + r.contactCounts[n - 1];
    r.control[6] = 0;
    if (r.control[2] > r.c.contactCapacity) fail(r, 32);
}

// ---------------------------------------------------------------- translating planes (solver.slang:135-145, 578-667)
real wallStepWork(const R& r, uint32_t id, real dt) {  // :135-145
    real total = 0;
    for (uint32_t w = 0; w < r.c.wallCount; w++) {
        uint32_t index = w * r.c.N + id;
        real increment = dt * r.wallRows[3 * index + 1].w;
        float4 state = r.wallAccumulated[index];
        float3 high = {state.x, 0, 0}, low = {state.y, 0, 0};
        compensated3(high, low, {increment, 0, 0});
        r.wallAccumulated[index] = {high.x, low.x, 0, 0};
        r.wallRows[3 * index + 2] = {high.x + low.x, 0, 0, 0};
        total += increment;
    }
    return total;
}
struct float2p { real x, y; };
inline float2p wallTwoSum(real a, real b) {  // :578-582 (materialization mask 0)
    real high = a + b, v = high - a;
    real low = (a - (high - v)) + (b - v);
    return {high, low};
}
inline float2p wallAdd(float2p a, real b) {  // :583-586
    float2p sum = wallTwoSum(a.x, b);
    return wallTwoSum(sum.x, sum.y + a.y);
}
inline float2p wallAddDot(float2p sum, float3 a, float3 b) {  // :587-594
    real product[3] = {a.x * b.x, a.y * b.y, a.z * b.z};
    real error[3] = {std::fma(a.x, b.x, -product[0]), std::fma(a.y, b.y, -product[1]), std::fma(a.z, b.z, -product[2])};
    for (int axis = 0; axis < 3; axis++) {
        sum = wallAdd(sum, product[axis]);
        sum = wallAdd(sum, error[axis]);
    }
    return sum;
}
float2p wallCentreDistance(const R& r, uint32_t body, uint32_t wall) {  // :596-604
    float4 plane = r.wallPlane[wall], low = r.wallRemainder[wall];
    float3 p = xyz(r.position[body]), pl = xyz(r.positionRemainder[body]);
    float2p sum = {-plane.w, -low.w};
    sum = wallAddDot(sum, xyz(plane), p);
    sum = wallAddDot(sum, xyz(plane), pl);
    sum = wallAddDot(sum, xyz(low), p);
    return wallAddDot(sum, xyz(low), pl);
}
real localWallDistance(const R& r, uint32_t body, uint32_t wall, float3 arm) {  // :605-610
    float2p sum = wallCentreDistance(r, body, wall);
    sum = wallAddDot(sum, xyz(r.wallPlane[wall]), arm);
    sum = wallAddDot(sum, xyz(r.wallRemainder[wall]), arm);
    return sum.x + sum.y;
}
bool wallClusterMayContact(const R& r, uint32_t body, uint32_t wall, uint32_t cluster) {  // :613-623
    float3 low = xyz(r.meshClusterMin[cluster]), high = xyz(r.meshClusterMax[cluster]);
    float4 plane = r.wallPlane[wall];
    float3 normal = qinv(r.rotation[body], xyz(plane)) * r.body[body].x;
    float3 center = real(0.5) * (low + high), half = real(0.5) * (high - low);
    float2p centreDistance = wallCentreDistance(r, body, wall);
    real base = centreDistance.x + centreDistance.y;
    real support = dot(normal, center) - dot(fabs3(normal), half);
    real pad = real(64.0) * real(1.1920928955078125e-7) *
                (real(1.0) + std::fabs(plane.w) + length(xyz(r.position[body])) + length(normal) * (length(center) + length(half)));
    return base + support <= pad;
}
void appendWallContact(const R& r, uint32_t body, uint32_t wall, uint32_t node) {  // :625-638
    float4 point = r.sample[r.sampleInfo[r.shapeId[body]].x + node];
    float3 arm = qrot(r.rotation[body], xyz(point) * r.body[body].x);
    float4 plane = r.wallPlane[wall];
    real phi = localWallDistance(r, body, wall, arm);
    if (!std::isfinite(phi)) { fail(r, 512); return; }
    if (phi >= 0) return;
    if (-phi > r.c.maxOverlap * r.body[body].w) fail(r, 4);
    uint32_t offset = r.control[2]++;
    if (offset >= r.c.contactCapacity) { fail(r, 32); return; }
    r.meta[offset] = {body, r.c.N + wall, node, 1};
    r.pointDepth[offset] = f4(arm, -phi);
    r.normalArea[offset] = f4(xyz(plane), point.w * r.body[body].x * r.body[body].x);
}
void collectWallContacts(const R& r, uint32_t id) {  // CPU serial order, cpu_scheduling.slang:133-154
    uint32_t body = id % r.c.N, wall = id / r.c.N;
    if (wall >= r.c.wallCount || r.control[0] != 0) return;
    float4 plane = r.wallPlane[wall];
    float2p centreDistance = wallCentreDistance(r, body, wall);
    real phi = centreDistance.x + centreDistance.y;
    real pad = real(64.0) * real(1.1920928955078125e-7) * (real(1.0) + length(xyz(r.position[body])) + std::fabs(plane.w));
    if (phi > r.body[body].z + pad) return;
    if (r.c.meshGeometry == 0) {
        uint32_t count = r.sampleInfo[r.shapeId[body]].y;
        for (uint32_t node = 0; node < count; node++) appendWallContact(r, body, wall, node);
        return;
    }
    uint4 info = r.meshClusterInfo[r.shapeId[body]];
    uint32_t cluster = info.x;
    while (cluster < info.x + info.y) {
        bool candidate = wallClusterMayContact(r, body, wall, cluster);
        uint4 range = r.meshClusterRanges[cluster];
        if (candidate)
            for (uint32_t lane = 0; lane < range.y; lane++) appendWallContact(r, body, wall, r.meshSampleIndices[range.x + lane]);
        cluster = candidate ? cluster + 1 : range.z;
    }
}
void advanceWalls(const R& r) {  // :669-687
    if (r.control[0] != 0) return;
    for (uint32_t w = 0; w < r.c.wallCount; w++) {
        float4 plane = r.wallPlane[w], residual = r.wallRemainder[w];
        float3 high = {plane.w, 0, 0}, low = {residual.w, 0, 0};
        compensated3(high, low, {dot(xyz(plane), xyz(r.wallVelocity[w])) * r.c.dt, 0, 0});
        plane.w = high.x;
        residual.w = low.x;
        if (!std::isfinite(high.x) || !std::isfinite(low.x)) { fail(r, 512); return; }
        r.wallPlane[w] = plane;
        r.wallRemainder[w] = residual;
    }
    if (r.c.wallBoxEnabled != 0) {
        for (uint32_t axis = 0; axis < 3; axis++) {
            uint32_t lo = r.wallBoxIds[2 * axis], hi = r.wallBoxIds[2 * axis + 1];
            real low = r.wallPlane[lo].w + r.wallRemainder[lo].w;
            real high = -(r.wallPlane[hi].w + r.wallRemainder[hi].w);
            if (low < 0 || high > get(xyz(r.cell[0]), int(axis)) || high <= low) fail(r, 4096);
        }
    }
}
// planeViscousResponse (plane.slang:13-44) with normalResponse weight 1 (normal.slang:3-11)
struct PlaneResponse { float4 forceEnergy, slipLoss; bool sliding; real normalLossPower; };
PlaneResponse planeViscousResponse(float3 normal, real signedDistance, real area, float3 rv, float3 oldSlip, real kn,
                                   real kt, real mu, real viscosity, real dt) {
    PlaneResponse value{{0, 0, 0, 0}, {0, 0, 0, 0}, false, 0};
    if (signedDistance >= 0) return value;
    real depth = -signedDistance;
    float3 tangent = rv - normal * dot(rv, normal);
    float3 trial = oldSlip - normal * dot(oldSlip, normal) + dt * tangent;
    float3 slip = trial;
    real normalVelocity = dot(rv, normal);
    real fn = 0;
    {
        real elastic = real(1.0) * kn * area * depth;
        fn = viscosity == 0 ? elastic : std::fmax(real(0.0), elastic - real(1.0) * viscosity * area * normalVelocity);
        value.normalLossPower = std::fmax(real(0.0), (elastic - fn) * normalVelocity);
    }
    float3 ft = -kt * area * slip;
    real magnitude = length(ft);
    if (magnitude > mu * fn && magnitude > 0) {
        value.sliding = true;
        ft = ft * (mu * fn / magnitude);
        slip = -ft / (kt * area);
    }
    value.forceEnergy = f4(fn * normal + ft, real(0.5) * area * (kn * depth * depth + kt * dot(slip, slip)));
    value.slipLoss = f4(slip, std::fmax(real(0.0), -dot(ft, trial - slip)));
    return value;
}
void gatherWallRows(const R& r, uint32_t index) {  // :837-856
    if (index >= r.c.N * r.c.wallCount || r.control[0] != 0) return;
    uint32_t body = index % r.c.N, wall = index / r.c.N, count = 0, visited = 0;
    float3 reaction = {0, 0, 0}, moment = {0, 0, 0};
    real power = 0;
    uint32_t entry = r.heads[body];
    while (entry != EMPTY) {
        if (visited++ > 2 * r.c.contactCapacity) { fail(r, 128); return; }
        uint32_t contact = entry / 2;
        if (r.meta[contact].y == r.c.N + wall) {
            float3 force = xyz(r.contactForce[contact]);
            float3 point = xyz(r.position[body]) + xyz(r.positionRemainder[body]) + xyz(r.pointDepth[contact]);
            reaction = reaction - force;
            moment = moment - cross(point, force);
            power += dot(force, xyz(r.wallVelocity[wall]));
            count++;
        }
        entry = r.next[entry];
    }
    float4 work = r.wallAccumulated[index];
    float countBits;  // the count's 32-bit pattern, stored as an FP32 value
    std::memcpy(&countBits, &count, 4);
    r.wallRows[3 * index] = f4(reaction, real(countBits));
    r.wallRows[3 * index + 1] = f4(moment, power);
    r.wallRows[3 * index + 2] = {work.x + work.y, 0, 0, 0};
}
void refreshWallPower(const R& r, uint32_t index) {  // :950-955
    if (index >= r.c.N * r.c.wallCount) return;
    float4 value = r.wallRows[3 * index + 1];
    value.w = -dot(xyz(r.wallRows[3 * index]), xyz(r.wallVelocity[index / r.c.N]));
    r.wallRows[3 * index + 1] = value;
}

// ---------------------------------------------------------------- kernels
void beginStep(const R& r, uint32_t id) {  // solver.slang:859-888
    if (id >= r.c.N || r.control[0] != 0) return;
    real dt = r.c.dt;
    float3 rate = xyz(r.cell[1]);
    accumulateNormalLoss(r, id, real(0.5) * dt);
    float3 v = xyz(r.velocity[id]), vl = xyz(r.velocityRemainder[id]);
    float3 l = xyz(r.momentum[id]), ll = xyz(r.momentumRemainder[id]);
    compensated3(v, vl, real(0.5) * dt * xyz(r.force[id]) / r.body[id].y);
    compensated3(l, ll, real(0.5) * dt * xyz(r.torque[id]));
    real loss = -expm1Stable(-r.c.drag * dt) * (real(0.5) * r.body[id].y * dot(v + vl, v + vl) + rotationalEnergy(r, id, l + ll));
    accumulate(r, id, {0, loss, real(0.5) * dt * cellPower(r, id), wallStepWork(r, id, real(0.5) * dt)});
    real damping = expm1Stable(-real(0.5) * r.c.drag * dt);
    compensated3(v, vl, damping * (v + vl));
    compensated3(l, ll, damping * (l + ll));
    r.momentum[id] = f4(l, 0);
    r.momentumRemainder[id] = f4(ll, 0);
    float4 q = r.rotation[id], ql = r.rotationRemainder[id];
    float3 w = omega(r, id);
    r.rotation[id] = normalize4(q + rotationIncrement(w, real(0.5) * dt, q));
    w = omega(r, id);
    compensated4(q, ql, rotationIncrement(w, dt, q));
    if (std::fabs(dot4(q, q) - real(1.0)) > real(2e-6)) { q = normalize4(q); ql = {0, 0, 0, 0}; }
    r.rotation[id] = q;
    r.rotationRemainder[id] = ql;
    float3 logarithm = xyz(r.cellLog[0]), residual = xyz(r.cellLog[1]);
    compensated3(logarithm, residual, rate * dt);
    float3 box = xyz(r.cell[2]) * f3(exp32(logarithm.x), exp32(logarithm.y), exp32(logarithm.z));
    float3 p = xyz(r.position[id]), pl = xyz(r.positionRemainder[id]);
    compensated3(p, pl, expm1Stable3(rate * dt) * (p + pl) + dt * (v + vl));
    if (!finite3(p) || !finite3(pl)) { fail(r, 512); return; }
    if (r.c.periodic == 1) p = p - floor3(p / box) * box;
    else {
        if (anyPeriodic(r.c)) {
            float3 w = p - floor3(p / box) * box;
            p = {periodicAxis(r.c, 0) ? w.x : p.x, periodicAxis(r.c, 1) ? w.y : p.y, periodicAxis(r.c, 2) ? w.z : p.z};
        }
        float3 s = p + pl, t = (p - box) + pl;
        for (int a = 0; a < 3; a++)
            if (!periodicAxis(r.c, a) && (get(s, a) < 0 || get(t, a) >= 0)) { fail(r, 2048); return; }
    }
    r.position[id] = f4(p, 0);
    r.positionRemainder[id] = f4(pl, 0);
    compensated3(v, vl, expm1Stable3(-rate * dt) * (v + vl));
    r.velocity[id] = f4(v, 0);
    r.velocityRemainder[id] = f4(vl, 0);
}
void advanceCell(const R& r) {  // :890-897
    if (r.control[0] != 0) return;
    float3 logarithm = xyz(r.cellLog[0]), residual = xyz(r.cellLog[1]);
    compensated3(logarithm, residual, xyz(r.cell[1]) * r.c.dt);
    r.cellLog[0] = f4(logarithm, 0);
    r.cellLog[1] = f4(residual, 0);
    r.cell[0] = f4(xyz(r.cell[2]) * f3(exp32(logarithm.x), exp32(logarithm.y), exp32(logarithm.z)), 0);
    r.control[4] = 0;
    if (r.c.periodic == 1 && r.referenceCell[1].w >= real(0.5) * std::fmin(r.cell[0].x, std::fmin(r.cell[0].y, r.cell[0].z)))
        fail(r, 256);
    if (r.c.periodic >= 8u)
        for (int a = 0; a < 3; a++)
            if (periodicAxis(r.c, a) && r.referenceCell[1].w >= real(0.5) * get(xyz(r.cell[0]), a)) fail(r, 256);
}
void checkRebuild(const R& r, uint32_t id) {  // :899-904
    if (id >= r.c.N) return;
    float3 d = image(r, (xyz(r.position[id]) - xyz(r.referencePosition[id])) +
                            (xyz(r.positionRemainder[id]) - xyz(r.referencePositionRemainder[id])));
    real deformation = length(xyz(r.cell[0]) - xyz(r.referenceCell[0]));
    if (length(d) + deformation > real(0.5) * r.c.skin) r.control[4] |= 1;
}
void finishStep(const R& r, uint32_t id) {  // :906-920
    if (id >= r.c.N || r.control[0] != 0) return;
    accumulateNormalLoss(r, id, real(0.5) * r.c.dt);
    float3 v = xyz(r.velocity[id]), vl = xyz(r.velocityRemainder[id]);
    float3 l = xyz(r.momentum[id]), ll = xyz(r.momentumRemainder[id]);
    compensated3(v, vl, real(0.5) * r.c.dt * xyz(r.force[id]) / r.body[id].y);
    compensated3(l, ll, real(0.5) * r.c.dt * xyz(r.torque[id]));
    real loss = -expm1Stable(-r.c.drag * r.c.dt) * (real(0.5) * r.body[id].y * dot(v + vl, v + vl) + rotationalEnergy(r, id, l + ll));
    accumulate(r, id, {0, loss, real(0.5) * r.c.dt * cellPower(r, id), wallStepWork(r, id, real(0.5) * r.c.dt)});
    real damping = expm1Stable(-real(0.5) * r.c.drag * r.c.dt);
    compensated3(v, vl, damping * (v + vl));
    compensated3(l, ll, damping * (l + ll));
    if (!finite3(v) || !finite3(l)) { fail(r, 512); return; }
    r.velocity[id] = f4(v, 0);
    r.velocityRemainder[id] = f4(vl, 0);
    r.momentum[id] = f4(l, 0);
    r.momentumRemainder[id] = f4(ll, 0);
}
void captureContactKinematics(const R& r, uint32_t id) {  // :171-175
    if (id >= r.c.N || r.control[0] != 0) return;
    r.forceVelocity[id] = f4(xyz(r.velocity[id]) + xyz(r.velocityRemainder[id]), 0);
    r.forceOmega[id] = f4(omega(r, id), 0);
}
void clearBins(const R& r, uint32_t id) {  // :346-349
    if (r.control[4] == 0) return;
    if (id < r.c.binCount) r.binCounts[id] = 0;
}
void fillBins(const R& r, uint32_t id) {  // :351-359
    if (id >= r.c.N || r.control[4] == 0) return;
    float3 p = xyz(r.position[id]), pl = xyz(r.positionRemainder[id]);
    if (!finite3(p) || !finite3(pl)) { fail(r, 512); return; }
    if (r.c.periodic != 1) {
        float3 s = p + pl, t = (p - xyz(r.cell[0])) + pl;
        for (int a = 0; a < 3; a++)
            if (!periodicAxis(r.c, a) && (get(s, a) < 0 || get(t, a) >= 0)) { fail(r, 2048); return; }
    }
    uint32_t b[3];
    binIndex(r, id, b);
    uint32_t bin = binFlat(r, b);
    uint32_t slot = r.binCounts[bin]++;
    if (slot < r.c.binCapacity) r.binParticles[bin * r.c.binCapacity + slot] = id;
    else fail(r, 8);
}
uint32_t findPairs(const R& r, uint32_t id, bool emit) {  // :360-384
    uint32_t b[3];
    binIndex(r, id, b);
    real cutoff = r.referenceCell[1].w;
    float3 cellBox = xyz(r.cell[0]);
    float3 g = f3(cutoff, cutoff, cutoff) / cellBox * f3(real(r.c.bins[0]), real(r.c.bins[1]), real(r.c.bins[2]));
    int reach[3] = {int(std::ceil(g.x)), int(std::ceil(g.y)), int(std::ceil(g.z))};
    int low[3], width[3];
    for (int a = 0; a < 3; a++) {
        low[a] = std::max(int(b[a]) - reach[a], 0);
        int high = std::min(int(b[a]) + reach[a], int(r.c.bins[a]) - 1);
        width[a] = periodicAxis(r.c, a) ? int(std::min(uint32_t(2 * reach[a] + 1), r.c.bins[a])) : high - low[a] + 1;
    }
    uint32_t count = 0;
    for (int x = 0; x < width[0]; x++)
        for (int y = 0; y < width[1]; y++)
            for (int z = 0; z < width[2]; z++) {
                uint32_t c[3];
                const int at[3] = {x, y, z};
                for (int a = 0; a < 3; a++)
                    c[a] = periodicAxis(r.c, a) ? modIndex(int(b[a]) - reach[a] + at[a], r.c.bins[a]) : uint32_t(low[a] + at[a]);
                uint32_t bi = binFlat(r, c);
                uint32_t occupancy = std::min(r.binCounts[bi], r.c.binCapacity);
                for (uint32_t slot = 0; slot < occupancy; slot++) {
                    uint32_t j = r.binParticles[bi * r.c.binCapacity + slot];
                    if (j <= id) continue;
                    float3 d = centerBranch(r, id, j);
                    real range = r.body[id].z + r.body[j].z + r.c.skin;
                    if (dot(d, d) > range * range) continue;
                    if (emit) {
                        uint32_t index = r.pairOffsets[id] + count;
                        if (index < r.c.pairCapacity) r.pairs[index] = {id, j};
                        else fail(r, 16);
                    }
                    count++;
                }
            }
    return count;
}
void countPairs(const R& r, uint32_t id) {  // :386
    if (id < r.c.N && r.control[4] != 0) r.pairCounts[id] = findPairs(r, id, false);
}
// Exclusive scan of pairCounts (scan.slang:1-28; CPU replica cpu.py:476-490).
int pairScan(const R& r) {
    uint64_t total = 0;
    for (uint32_t i = 0; i < r.c.N; i++) {
        uint32_t v = r.pairCounts[i];
        r.pairOffsets[i] = uint32_t(total);
        total += v;
        if (total > 0xffffffffull) return 1;
    }
    return 0;
}
void emitPairs(const R& r, uint32_t id) {  // :388-392
    if (id >= r.c.N || r.control[4] == 0) return;
    findPairs(r, id, true);
    r.referencePosition[id] = r.position[id];
    r.referencePositionRemainder[id] = r.positionRemainder[id];
    if (id == 0) {
        r.control[1] = r.pairOffsets[r.c.N - 1] + r.pairCounts[r.c.N - 1];
        r.referenceCell[0] = r.cell[0];
    }
}
void resetMeshContacts(const R& r, uint32_t id) {  // :397-403
    if (r.control[0] != 0) return;
    if (id < r.c.pairCapacity) r.contactCounts[id] = 0;
    if (id == 0) { r.control[2] = 0; r.control[6] = 0; }
}
void meshPairBounds(const R& r, uint32_t direction, uint32_t cluster, float3& low, float3& high) {  // :404-417
    float3 center = real(0.5) * (xyz(r.meshClusterMin[cluster]) + xyz(r.meshClusterMax[cluster]));
    float3 half = real(0.5) * (xyz(r.meshClusterMax[cluster]) - xyz(r.meshClusterMin[cluster]));
    float3 c, e;
    for (int axis = 0; axis < 3; axis++) {
        float4 row = r.meshPairTransform[4 * direction + axis];
        set(c, axis, dot(xyz(row), center) + row.w);
        set(e, axis, dot(fabs3(xyz(row)), half));
    }
    low = c - e;
    high = c + e;
    real pad = real(256.0) * EPS * (real(1.0) + maxabs(low) + maxabs(high) + r.meshPairTransform[4 * direction + 3].x);
    low = low - f3(pad, pad, pad);
    high = high + f3(pad, pad, pad);
}
void prepareMeshPairs(const R& r, uint32_t id) {  // :469-489
    if (id >= std::min(r.control[1], r.c.pairCapacity) || r.control[0] != 0) return;
    uint2 ij = r.pairs[id];
    uint32_t mask = 0;
    if (meshClusterMayContact(r, ij.x, ij.y, r.meshClusterInfo[r.shapeId[ij.x]].x)) mask |= 1;
    if (meshClusterMayContact(r, ij.y, ij.x, r.meshClusterInfo[r.shapeId[ij.y]].x)) mask |= 2;
    for (uint32_t pass = 0; pass < 2; pass++) {
        if ((mask & (1u << pass)) == 0) continue;
        uint32_t i = pass == 0 ? ij.x : ij.y, j = pass == 0 ? ij.y : ij.x;
        real ratio = r.body[i].x / r.body[j].x;
        float3 a = qinv(r.rotation[j], qrot(r.rotation[i], f3(ratio, 0, 0)));
        float3 b = qinv(r.rotation[j], qrot(r.rotation[i], f3(0, ratio, 0)));
        float3 c = qinv(r.rotation[j], qrot(r.rotation[i], f3(0, 0, ratio)));
        float3 branch = centerBranch(r, i, j), offset = qinv(r.rotation[j], -branch) / r.body[j].x;
        uint32_t base = 4 * (2 * id + pass);
        r.meshPairTransform[base] = {a.x, b.x, c.x, offset.x};
        r.meshPairTransform[base + 1] = {a.y, b.y, c.y, offset.y};
        r.meshPairTransform[base + 2] = {a.z, b.z, c.z, offset.z};
        r.meshPairTransform[base + 3] = {length(branch) / r.body[j].x, 0, 0, 0};
    }
    r.meshPairMask[id] = mask;
}
void clearContactStorage(const R& r, uint32_t id) {  // :698-701
    if (id < r.c.hashSize) r.hash[id] = EMPTY;
    if (id < r.c.N) r.heads[id] = EMPTY;
}
inline uint32_t hashKey(const R& r, uint4 key) {  // :689
    uint32_t h = key.x * 73856093u ^ key.y * 19349663u ^ key.z * 83492791u;
    h ^= h >> 16;
    return h & (r.c.hashSize - 1);
}
inline float3 transported(float3 x, float3 a, float3 b, real spin) {  // :691-696
    real c = std::fmin(std::fmax(dot(a, b), -real(1.0)), real(1.0));
    if (c < -real(0.999999)) return {0, 0, 0};
    float3 v = cross(a, b);
    float3 out = x + cross(v, x) + cross(v, cross(v, x)) / (real(1.0) + c);
    out = out * cos32(spin) + cross(b, out) * sin32(spin) + b * dot(b, out) * (real(1.0) - cos32(spin));
    return out - b * dot(out, b);
}
void resolveContacts(const R& r, uint32_t id) {  // :703-749
    if (id >= std::min(r.control[2], r.c.contactCapacity) || r.control[0] != 0) return;
    uint4 key = r.meta[id];
    uint32_t i = key.x, j = key.y;
    float3 n = xyz(r.normalArea[id]);
    bool wall = j >= r.c.N;
    real viscosity = r.c.normalViscosity;
    if (wall) viscosity = r.wallMaterial[j - r.c.N].w;
    float3 wi = viscosity > 0 ? xyz(r.forceOmega[i]) : omega(r, i), wj = {0, 0, 0};
    if (!wall) wj = viscosity > 0 ? xyz(r.forceOmega[j]) : omega(r, j);
    float3 elastic = {0, 0, 0};
    uint32_t slot = hashKey(r, key);
    if (r.c.historyStep != 0) {
        for (uint32_t probe = 0; probe < r.c.hashSize; probe++) {
            uint32_t old = r.oldHash[slot];
            if (old == EMPTY) break;
            uint4 om = r.oldMeta[old];
            if (om.x == key.x && om.y == key.y && om.z == key.z) {
                elastic = transported(xyz(r.oldSlip[old]), xyz(r.oldNormal[old]), n, real(0.5) * dot(wi + wj, n) * r.c.dt);
                break;
            }
            slot = (slot + 1) & (r.c.hashSize - 1);
        }
    }
    float3 arm = xyz(r.pointDepth[id]);
    if (wall) {
        uint32_t w = j - r.c.N;
        float4 material = r.wallMaterial[w];
        float3 v = viscosity > 0 ? xyz(r.forceVelocity[i]) : xyz(r.velocity[i]) + xyz(r.velocityRemainder[i]);
        float3 rv = v + cross(wi, arm) - xyz(r.wallVelocity[w]);
        PlaneResponse response = planeViscousResponse(n, -r.pointDepth[id].w, r.normalArea[id].w, rv, elastic, material.x,
                                                      material.y, material.z, material.w, r.c.dt);
        r.slip[id] = response.slipLoss;
        r.contactForce[id] = response.forceEnergy;
        r.normalLossPower[id] = response.normalLossPower;
        float4 fe = response.forceEnergy;
        if (!std::isfinite(fe.x) || !std::isfinite(fe.y) || !std::isfinite(fe.z) || !std::isfinite(fe.w) ||
            !std::isfinite(response.normalLossPower)) { fail(r, 512); return; }
        if (response.sliding) r.control[6] += 1;
        uint32_t oldHead = r.heads[i];
        r.heads[i] = 2 * id;
        r.next[2 * id] = oldHead;
        r.next[2 * id + 1] = EMPTY;
        return;
    }
    float3 branch = centerBranch(r, i, j);
    float3 translation = xyz(r.velocity[i]) - xyz(r.velocity[j]);
    if (viscosity > 0) translation = xyz(r.forceVelocity[i]) - xyz(r.forceVelocity[j]);
    float3 rv = translation - xyz(r.cell[1]) * branch + cross(wi, arm) - cross(wj, arm - branch);
    elastic = elastic + (rv - n * dot(rv, n)) * r.c.dt;
    float3 trial = elastic;
    real area = r.normalArea[id].w, depth = r.pointDepth[id].w;
    // normalResponse (normal.slang:3-11) with weight 0.5
    real fn = 0, lossPower = 0;
    real normalVelocity = dot(rv, n);
    if (depth > 0) {
        real el = real(0.5) * r.c.kn * area * depth;
        fn = viscosity == 0 ? el : std::fmax(real(0.0), el - real(0.5) * viscosity * area * normalVelocity);
        lossPower = std::fmax(real(0.0), (el - fn) * normalVelocity);
    }
    r.normalLossPower[id] = lossPower;
    float3 ft = -real(0.5) * r.c.kt * area * elastic;
    real magnitude = length(ft);
    if (magnitude > r.c.mu * fn && magnitude > 0) {
        ft = ft * (r.c.mu * fn / magnitude);
        elastic = -ft / (real(0.5) * r.c.kt * area);
        r.control[6] += 1;
    }
    r.slip[id] = f4(elastic, std::fmax(real(0.0), -dot(ft, trial - elastic)));
    r.contactForce[id] = f4(fn * n + ft, real(0.25) * area * (r.c.kn * depth * depth + r.c.kt * dot(elastic, elastic)));
    float4 cf = r.contactForce[id];
    if (!std::isfinite(cf.x) || !std::isfinite(cf.y) || !std::isfinite(cf.z) || !std::isfinite(cf.w) ||
        !std::isfinite(lossPower)) { fail(r, 512); return; }
    uint32_t oldHead = r.heads[i];
    r.heads[i] = 2 * id;
    r.next[2 * id] = oldHead;
    oldHead = r.heads[j];
    r.heads[j] = 2 * id + 1;
    r.next[2 * id + 1] = oldHead;
}
void buildHistoryHash(const R& r, uint32_t id) {  // :751-759
    if (id >= std::min(r.control[2], r.c.contactCapacity) || r.control[0] != 0) return;
    uint32_t slot = hashKey(r, r.meta[id]);
    for (uint32_t probe = 0; probe < r.c.hashSize; probe++) {
        if (r.hash[slot] == EMPTY) { r.hash[slot] = id; return; }
        slot = (slot + 1) & (r.c.hashSize - 1);
    }
    fail(r, 64);
}
void gather(const R& r, uint32_t id) {  // :761-811
    if (id >= r.c.N || r.control[0] != 0) return;
    float3 f = {0, 0, 0}, t = {0, 0, 0}, fLow = {0, 0, 0}, tLow = {0, 0, 0};
    float3 stress0 = {0, 0, 0}, stress1 = {0, 0, 0}, stress2 = {0, 0, 0};
    real energy = 0, denom = 0, maxDepth = 0, dissipation = 0, normalPower = 0;
    real stiffnessRate = 0, viscousRate = 0;
    float3 one = {1, 1, 1};
    real inverseInertia = std::fmax(std::fmax(dot(fabs3(xyz(r.invI[3 * id])), one), dot(fabs3(xyz(r.invI[3 * id + 1])), one)),
                                     dot(fabs3(xyz(r.invI[3 * id + 2])), one));
    uint32_t entry = r.heads[id];
    uint32_t visited = 0;
    while (entry != EMPTY) {
        if (visited++ > 2 * r.c.contactCapacity) { fail(r, 128); return; }
        uint32_t contact = entry / 2;
        bool source = (entry & 1) == 0;
        uint4 key = r.meta[contact];
        float3 force = xyz(r.contactForce[contact]);
        float3 arm = xyz(r.pointDepth[contact]);
        bool wall = key.y >= r.c.N;
        if (r.c.activeTimestep != 0) {
            float3 localArm = source ? arm : arm - centerBranch(r, key.x, key.y);
            real mobility = real(1.0) / r.body[id].y + dot(localArm, localArm) * inverseInertia;
            float4 material = {r.c.kn, r.c.kt, r.c.mu, r.c.normalViscosity};
            if (wall) material = r.wallMaterial[key.y - r.c.N];
            real area = r.normalArea[contact].w * (wall ? real(1.0) : real(0.5));
            real coupling = wall ? real(1.0) : real(2.0);
            stiffnessRate += coupling * area * (material.x * (real(1.0) + material.z) + material.y) * mobility;
            viscousRate += coupling * area * material.w * (real(1.0) + material.z) * mobility;
        }
        if (wall) {
            compensated3(f, fLow, force);
            compensated3(t, tLow, cross(arm, force));
            stress0 = stress0 - force.x * arm;
            stress1 = stress1 - force.y * arm;
            stress2 = stress2 - force.z * arm;
            energy += r.contactForce[contact].w;
            denom += length(force);
            maxDepth = std::fmax(maxDepth, r.pointDepth[contact].w / r.body[key.x].w);
            dissipation += r.slip[contact].w;
            normalPower += r.normalLossPower[contact];
            entry = r.next[entry];
            continue;
        }
        float3 branch = centerBranch(r, key.x, key.y);
        compensated3(f, fLow, source ? force : -force);
        compensated3(t, tLow, source ? cross(arm, force) : cross(arm - branch, -force));
        stress0 = stress0 - real(0.5) * force.x * branch;
        stress1 = stress1 - real(0.5) * force.y * branch;
        stress2 = stress2 - real(0.5) * force.z * branch;
        energy += real(0.5) * r.contactForce[contact].w;
        denom += length(force);
        maxDepth = std::fmax(maxDepth, r.pointDepth[contact].w / std::fmin(r.body[key.x].w, r.body[key.y].w));
        dissipation += real(0.5) * r.slip[contact].w;
        normalPower += real(0.5) * r.normalLossPower[contact];
        entry = r.next[entry];
    }
    f = f + fLow;
    t = t + tLow;
    r.timestepRows[id] = {stiffnessRate, viscousRate, 0, 0};
    float4 normalState = r.normalDamping[id];
    normalState.z = normalPower;
    r.normalDamping[id] = normalState;
    r.force[id] = f4(f, maxDepth);
    r.torque[id] = f4(t, dissipation);
    r.particleStress[3 * id] = f4(stress0, 0);
    r.particleStress[3 * id + 1] = f4(stress1, 0);
    r.particleStress[3 * id + 2] = f4(stress2, 0);
    real kinetic = real(0.5) * (r.body[id].y * dot(xyz(r.velocity[id]), xyz(r.velocity[id])) + dot(xyz(r.momentum[id]), omega(r, id)));
    r.particleDiagnostic[id] = {kinetic, energy, length(f), denom};
    accumulate(r, id, {dissipation, 0, 0, 0});
}
void checkContactTimestep(const R& r, uint32_t id) {  // :816-832
    if (id >= r.c.N || r.control[0] != 0 || r.c.dt == 0) return;
    real dt = r.c.dt;
    real rx = r.timestepRows[id].x, ry = r.timestepRows[id].y;
    float3 v = xyz(r.velocity[id]) + xyz(r.velocityRemainder[id]);
    real speed = length(v) + dt * length(xyz(r.force[id])) / r.body[id].y;
    float3 one = {1, 1, 1};
    real inverseInertia = std::fmax(std::fmax(dot(fabs3(xyz(r.invI[3 * id])), one), dot(fabs3(xyz(r.invI[3 * id + 1])), one)),
                                     dot(fabs3(xyz(r.invI[3 * id + 2])), one));
    real spin = inverseInertia * (length(xyz(r.momentum[id]) + xyz(r.momentumRemainder[id])) + dt * length(xyz(r.torque[id])));
    real wallSpeed = 0;
    for (uint32_t w = 0; w < r.c.wallCount; w++) wallSpeed = std::fmax(wallSpeed, length(xyz(r.wallVelocity[w])));
    real motion = dt * (speed + r.body[id].z * spin + wallSpeed);
    bool finite = std::isfinite(rx) && std::isfinite(ry) && std::isfinite(motion) && std::isfinite(spin);
    if (!finite) { fail(r, 512); return; }
    float3 rate = fabs3(xyz(r.cell[1])) * dt;
    if (dt * std::sqrt(rx) > real(0.1) || dt * ry > real(0.1) || dt * spin > real(0.01) || motion > real(0.25) * r.c.skin ||
        rate.x > real(0.01) || rate.y > real(0.01) || rate.z > real(0.01))
        fail(r, 8192);
}

// Pair-parallel collector with order-preserving concatenation
// (cpu_parallel.hpp:31-117): per-pair private output, serial copy in pair order.
int collectParallel(R& r, uint32_t count, uint32_t threads) {
    struct Hits { std::vector<uint4> meta; std::vector<float4> point, normal; uint32_t flags = 0, count = 0; };
    std::vector<Hits> hits(count);
    std::atomic<uint32_t> next{0};
    const uint32_t limit = r.c.contactCapacity;
    auto worker = [&]() {
        for (;;) {
            uint32_t pair = next.fetch_add(1, std::memory_order_relaxed);
            if (pair >= count) return;
            Hits& hit = hits[pair];
            uint32_t capacity = std::min<uint32_t>(limit, 1024);
            for (;;) {
                hit.meta.resize(std::max<size_t>(capacity, hit.meta.size()));
                hit.point.resize(std::max<size_t>(capacity, hit.point.size()));
                hit.normal.resize(std::max<size_t>(capacity, hit.normal.size()));
                R local = r;
                uint32_t control[16];
                std::memcpy(control, r.control, sizeof(control));
                control[2] = 0;
                local.control = control;
                local.c.contactCapacity = capacity;
                local.meta = hit.meta.data();
                local.pointDepth = hit.point.data();
                local.normalArea = hit.normal.data();
                r.contactCounts[pair] = 0;
                cpuCollectMesh(local, pair);
                if (control[0] == 32 && capacity < limit) {
                    capacity = uint32_t(std::min<uint64_t>(limit, uint64_t(capacity) * 2));
                    continue;
                }
                hit.flags = control[0];
                hit.count = control[2];
                break;
            }
        }
    };
    std::vector<std::thread> pool;
    uint32_t workers = std::min(threads, count);
    for (uint32_t i = 0; i < workers; ++i) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    uint64_t total = 0;
    uint32_t flags = 0;
    for (auto& hit : hits) { total += hit.count; flags |= hit.flags; }
    r.control[2] = uint32_t(std::min<uint64_t>(total, 0xffffffffu));
    if (total > limit) flags |= 32;
    r.control[0] |= flags;
    if (flags) return 0;
    size_t offset = 0;
    for (auto& hit : hits) {
        if (hit.count) {
            std::memcpy(r.meta + offset, hit.meta.data(), hit.count * sizeof(uint4));
            std::memcpy(r.pointDepth + offset, hit.point.data(), hit.count * sizeof(float4));
            std::memcpy(r.normalArea + offset, hit.normal.data(), hit.count * sizeof(float4));
        }
        offset += hit.count;
    }
    return 0;
}

// ---------------------------------------------------------------- evidence rows (solver.slang:922-947)
void diagnosticRows(const R& r, uint32_t id) {  // :922-936
    if (id >= r.c.N) return;
    float3 v = xyz(r.velocity[id]);
    real m = r.body[id].y;
    real kinetic = real(0.5) * (m * dot(v, v) + dot(xyz(r.momentum[id]), omega(r, id)));
    r.diagnosticRows[id * 8] = f4(xyz(r.particleStress[id * 3]), kinetic);
    r.diagnosticRows[id * 8 + 1] = f4(xyz(r.particleStress[id * 3 + 1]), r.particleDiagnostic[id].y);
    r.diagnosticRows[id * 8 + 2] = f4(xyz(r.particleStress[id * 3 + 2]), length(xyz(r.force[id])));
    r.diagnosticRows[id * 8 + 3] = f4(m * v.x * v, r.particleDiagnostic[id].w);
    r.diagnosticRows[id * 8 + 4] = f4(m * v.y * v, r.force[id].w);
    r.diagnosticRows[id * 8 + 5] = f4(m * v.z * v, r.torque[id].w);
    r.diagnosticRows[id * 8 + 6] = r.accumulated[id];
    float4 normalState = r.normalDamping[id];
    r.diagnosticRows[id * 8 + 7] = {normalState.x + normalState.y, normalState.z, 0, 0};
}
void fabricRows(const R& r, uint32_t id) {  // :938-947
    if (id >= r.c.pairCapacity) return;
    float4 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
    if (id < r.control[1] && r.contactCounts[id] > 0) {
        uint2 ij = r.pairs[id];
        float3 n = normalize3(centerBranch(r, ij.x, ij.y));
        a = f4(n * n, n.x * n.y);
        b = {n.x * n.z, n.y * n.z, 1, 0};
    }
    r.fabricRows[2 * id] = a;
    r.fabricRows[2 * id + 1] = b;
}

}  // namespace

extern "C" int oracle_dispatch(const char* kernel, uint32_t count, void** buffers, const uint64_t* bytes,
                               const double* values) {
    try {
        R r = bind(buffers, bytes, values);
        if (r.c.sourceSurfaceNormal != 0) return 3;
        if (r.c.meshGeometry == 0 && (!r.grids || !r.gridOrigin || !r.distance)) return 2;
        std::string name(kernel);
        const uint32_t wallTasks = r.c.N * r.c.wallCount;
        auto each = [&](auto fn) { for (uint32_t id = 0; id < count; id++) fn(r, id); };
        if (name == "beginStep") each(beginStep);
        else if (name == "advanceCell") advanceCell(r);
        else if (name == "checkRebuild") each(checkRebuild);
        else if (name == "clearBins") each(clearBins);
        else if (name == "fillBins") each(fillBins);
        else if (name == "countPairs") each(countPairs);
        else if (name == "pairScan") { if (r.control[4] != 0 && pairScan(r)) return 4; }
        else if (name == "emitPairs") each(emitPairs);
        else if (name == "resetMeshContacts") each(resetMeshContacts);
        else if (name == "prepareMeshPairs") each(prepareMeshPairs);
        else if (name == "collectMeshFrontiers" || name == "collectMeshContacts") {  // cpu.py:80-96: both the serial collector
            uint32_t pairs = std::min(r.control[1], r.c.pairCapacity);
            if (r.c.threads > 1 && pairs > 1 && r.control[0] == 0) return collectParallel(r, pairs, r.c.threads);
            for (uint32_t pair = 0; pair < pairs; pair++) cpuCollectMesh(r, pair);
        }
        else if (name == "clearContactStorage") each(clearContactStorage);
        else if (name == "resolveContacts") each(resolveContacts);
        else if (name == "buildHistoryHash") each(buildHistoryHash);
        else if (name == "gather") each(gather);
        else if (name == "finishStep") each(finishStep);
        else if (name == "captureContactKinematics") each(captureContactKinematics);
        else if (name == "checkContactTimestep") each(checkContactTimestep);
        else if (name == "countContacts") each(countContacts);
        else if (name == "contactScan") contactScan(r);
        else if (name == "contactTotal") contactTotal(r);
        else if (name == "emitContacts") each(emitContacts);
        else if (name == "diagnosticRows") each(diagnosticRows);
        else if (name == "fabricRows") each(fabricRows);
        else if (name == "advanceWalls") advanceWalls(r);
        else if (name == "collectWallContacts") { for (uint32_t id = 0; id < wallTasks; id++) collectWallContacts(r, id); }
        else if (name == "gatherWallRows") { for (uint32_t id = 0; id < wallTasks; id++) gatherWallRows(r, id); }
        else if (name == "refreshWallPower") { for (uint32_t id = 0; id < wallTasks; id++) refreshWallPower(r, id); }
        else return 2;
        return 0;
    } catch (...) {
        return 1;
    }
}

extern "C" int oracle_probe(const char* kernel, const uint32_t* uniforms, void** buffers, const uint64_t* bytes,
                            const double* values, const void* input, void* output) {
#ifdef SAD_ORACLE_FP64
    (void)kernel, (void)uniforms, (void)buffers, (void)bytes, (void)values, (void)input, (void)output;
    return 3;  // probe inputs and outputs are FP32 rows; probes are for the FP32 build
#endif
    try {
        R r = bind(buffers, bytes, values);
        std::string name(kernel);
        uint32_t count = uniforms[0];
        if (name == "meshProbe") {
            const float4* points = static_cast<const float4*>(input);
            float4* out = static_cast<float4*>(output);
            for (uint32_t id = 0; id < count; id++) out[id] = meshDistance(r, uniforms[1], xyz(points[id]));
        } else if (name == "meshCullProbe") {
            const float4* points = static_cast<const float4*>(input);
            uint32_t* out = static_cast<uint32_t*>(output);
            for (uint32_t id = 0; id < count; id++) out[id] = meshMayContain(r, uniforms[1], xyz(points[id])) ? 1 : 0;
        } else if (name == "meshClusterProbe") {
            uint32_t* out = static_cast<uint32_t*>(output);
            uint32_t src = uniforms[1], dst = uniforms[2];
            for (uint32_t id = 0; id < count; id++)
                out[id] = meshClusterMayContact(r, src, dst, r.meshClusterInfo[r.shapeId[src]].x + id) ? 1 : 0;
        } else if (name == "meshPairBoundsProbe") {
            float4* out = static_cast<float4*>(output);
            for (uint32_t id = 0; id < count; id++) {
                float3 low, high;
                meshPairBounds(r, uniforms[1], uniforms[2] + id, low, high);
                out[2 * id] = f4(low, 0);
                out[2 * id + 1] = f4(high, 0);
            }
        } else if (name == "nodeContactProbe") {
            // input: uint4 (src, dst, node, usePlane) rows; output: pd, na, hit flag per row.
            const uint4* rows = static_cast<const uint4*>(input);
            float4* out = static_cast<float4*>(output);
            for (uint32_t id = 0; id < count; id++) {
                float4 pd, na;
                bool hit = nodeContact(r, rows[id].x, rows[id].y, rows[id].z, pd, na);
                out[3 * id] = pd;
                out[3 * id + 1] = na;
                out[3 * id + 2] = {hit ? real(1.0) : real(0.0), 0, 0, 0};
            }
        } else if (name == "transportProbe") {
            // input float4 rows (x, a, b, spin in b.w) -> transported vector.
            const float4* rows = static_cast<const float4*>(input);
            float4* out = static_cast<float4*>(output);
            for (uint32_t id = 0; id < count; id++)
                out[id] = f4(transported(xyz(rows[3 * id]), xyz(rows[3 * id + 1]), xyz(rows[3 * id + 2]), rows[3 * id + 2].w), 0);
        } else {
            return 2;
        }
        return 0;
    } catch (...) {
        return 1;
    }
}
