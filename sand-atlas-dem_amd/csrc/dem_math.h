// dem_math.h -- FP32 vector algebra and geometric predicates of the mesh step,
// device side.  Operation order is fixed so results are bit-identical to the
// reference CPU backend's generated C++ (compiled -ffp-contract=off): dot()
// accumulates from +0.0f left to right, length = sqrtf(dot), normalize multiplies
// by 1/sqrtf(dot), min/max are IEEE fminf/fmaxf, and the centre branch is formed
// in FP64 (reference src/sand_atlas_dem/shaders/cpu_scheduling.slang:10-15).
// Build with -ffp-contract=off and correctly rounded FP32 divide/sqrt.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEM_D __device__ __forceinline__

struct v3 { float x, y, z; };
struct v4 { float x, y, z, w; };

DEM_D v3 mk3(float x, float y, float z) { return {x, y, z}; }
DEM_D v3 xyz(const float4& a) { return {a.x, a.y, a.z}; }
DEM_D v3 xyz(const v4& a) { return {a.x, a.y, a.z}; }
DEM_D v4 mk4(v3 a, float w) { return {a.x, a.y, a.z, w}; }
DEM_D float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
DEM_D float4 f4(v4 a) { return make_float4(a.x, a.y, a.z, a.w); }
DEM_D v4 v4of(float4 a) { return {a.x, a.y, a.z, a.w}; }
DEM_D v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
DEM_D v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
DEM_D v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
DEM_D v3 operator/(v3 a, v3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
DEM_D v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
DEM_D v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
DEM_D v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
DEM_D v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
DEM_D v4 operator+(v4 a, v4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
DEM_D v4 operator-(v4 a, v4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
DEM_D v4 operator*(v4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
DEM_D float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

DEM_D float dot(v3 a, v3 b) {
    float r = 0.0f;
    r = r + a.x * b.x;
    r = r + a.y * b.y;
    r = r + a.z * b.z;
    return r;
}
DEM_D float dot4(v4 a, v4 b) {
    float r = 0.0f;
    r = r + a.x * b.x;
    r = r + a.y * b.y;
    r = r + a.z * b.z;
    r = r + a.w * b.w;
    return r;
}
DEM_D v3 cross(v3 l, v3 r) { return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x}; }
DEM_D float length(v3 a) { return sqrtf(dot(a, a)); }
DEM_D v3 fabs3(v3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
DEM_D v3 fmax3(v3 a, v3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
DEM_D v3 fmin3(v3 a, v3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
DEM_D v3 floor3(v3 a) { return {floorf(a.x), floorf(a.y), floorf(a.z)}; }
DEM_D bool finite3(v3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
DEM_D float maxabs(v3 a) { return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fabsf(a.z)); }
DEM_D v4 normalize4(v4 a) { float s = 1.0f / sqrtf(dot4(a, a)); return a * s; }
DEM_D v3 normalize3(v3 a) { float s = 1.0f / sqrtf(dot(a, a)); return a * s; }

constexpr float DEM_EPS = 1.1920928955078125e-7f;
constexpr uint32_t DEM_EMPTY = 0xffffffffu;

#include "dem_transcendental.h"

// ---------------------------------------------------------------- rigid body (solver.slang:104-169)
DEM_D v3 qrot(v4 q, v3 v) { v3 u = xyz(q); return v + 2.0f * cross(u, cross(u, v) + q.w * v); }
DEM_D v3 qinv(v4 q, v3 v) { return qrot({-q.x, -q.y, -q.z, q.w}, v); }
DEM_D v4 qmul(v4 a, v4 b) {
    v3 v = a.w * xyz(b) + b.w * xyz(a) + cross(xyz(a), xyz(b));
    return mk4(v, a.w * b.w - dot(xyz(a), xyz(b)));
}
DEM_D void compensated3(v3& high, v3& low, v3 inc) {
    v3 y = inc + low; v3 next = high + y; v3 moved = next - high;
    low = y - moved; high = next;
}
DEM_D void compensated4(v4& high, v4& low, v4 inc) {
    v4 y = inc + low; v4 next = high + y; v4 moved = next - high;
    low = y - moved; high = next;
}
DEM_D float expm1Stable(float x) {
    if (fabsf(x) < 0.01f) return x * (1.0f + x * (0.5f + x * (1.0f / 6.0f + x * (1.0f / 24.0f + x / 120.0f))));
    return dem_expf(x) - 1.0f;
}
DEM_D v3 expm1Stable3(v3 x) { return {expm1Stable(x.x), expm1Stable(x.y), expm1Stable(x.z)}; }
DEM_D v4 rotationIncrement(v3 w, float dt, v4 q) {
    float half = 0.5f * length(w) * dt;
    float coefficient = fabsf(half) < 0.001f ? 0.5f * dt * (1.0f - half * half / 6.0f) : dem_sinf(half) / length(w);
    float s = dem_sinf(0.5f * half);
    float cm1 = -2.0f * s * s;
    return qmul(mk4(coefficient * w, cm1), q);
}

// ---------------------------------------------------------------- triangle / box queries (solver.slang:203-305)
DEM_D v3 closestTriangle(v3 p, v3 a, v3 b, v3 c, uint32_t& feature) {
    v3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    feature = 4;
    if (d1 <= 0 && d2 <= 0) return a;
    v3 bp = p - b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { feature = 5; return b; }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { feature = 1; return a + (d1 / (d1 - d3)) * ab; }
    v3 cp = p - c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { feature = 6; return c; }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { feature = 3; return a + (d2 / (d2 - d6)) * ac; }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        feature = 2;
        return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    }
    float inv = 1.0f / (va + vb + vc);
    feature = 0;
    return a + ab * (vb * inv) + ac * (vc * inv);
}
DEM_D float boxDistance(v3 lo, v3 hi, v3 p) {
    v3 d = fmax3(fmax3(lo - p, p - hi), mk3(0.0f, 0.0f, 0.0f));
    return dot(d, d);
}
DEM_D uint32_t dominantGroup(v3 direction) {
    v3 a = fabs3(direction);
    uint32_t axis = a.x >= a.y && a.x >= a.z ? 0 : (a.y >= a.z ? 1 : 2);
    return 2 * axis + (comp(direction, axis) >= 0 ? 1 : 0);
}
