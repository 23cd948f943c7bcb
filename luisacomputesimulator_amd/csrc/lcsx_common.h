// Shared device/host helpers for the lcsx HIP kernels (gfx950, wave64).
//
// All kernels are compiled with -ffp-contract=off: every floating-point expression below is
// evaluated exactly in the written order (left to right), which is what makes the per-element
// results reproducible bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LCSX_HD __host__ __device__ __forceinline__
#define LCSX_D __device__ __forceinline__

namespace lcsx
{

constexpr float kEpsilon        = 1.19209290E-7f;  // FLT_EPSILON (reference Core/constant_value.h:12)
constexpr float kLineSearchMaxT = 1.25f;           // ACCD step cap (reference accd.hpp:47)
constexpr float kCcdReduction   = 0.01f;           // ACCD gap reduction (reference accd.hpp:46)
constexpr uint32_t kInvalid     = 0xffffffffu;

struct f3
{
    float x, y, z;
};
LCSX_HD f3 mk(float a, float b, float c) { return f3{a, b, c}; }
LCSX_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LCSX_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LCSX_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
LCSX_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
LCSX_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
LCSX_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
LCSX_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LCSX_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LCSX_HD float length(f3 a) { return sqrtf(dot(a, a)); }
LCSX_HD f3 normalize(f3 a) { return a * (1.0f / length(a)); }
// min/max with std::min / std::max semantics ((b < a) ? b : a), so signed zeros and ties
// resolve the same way on the device and on the host.
LCSX_HD float smin(float a, float b) { return (b < a) ? b : a; }
LCSX_HD float smax(float a, float b) { return (a < b) ? b : a; }
LCSX_HD f3 vmin(f3 a, f3 b) { return {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
LCSX_HD f3 vmax(f3 a, f3 b) { return {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
LCSX_HD float clampf(float v, float lo, float hi) { return smax(smin(v, hi), lo); }
LCSX_HD float get(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

LCSX_HD f3 ld3(const float* p, uint32_t i) { return {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]}; }
LCSX_HD void st3(float* p, uint32_t i, f3 v)
{
    p[3 * (size_t)i]     = v.x;
    p[3 * (size_t)i + 1] = v.y;
    p[3 * (size_t)i + 2] = v.z;
}

// Row-major 3x3.
struct m3
{
    float m[9];
};
LCSX_HD m3 m3_zero()
{
    m3 a;
    for (int i = 0; i < 9; i++) a.m[i] = 0.0f;
    return a;
}
LCSX_HD f3 mul(const m3& a, f3 x)
{
    return {a.m[0] * x.x + a.m[1] * x.y + a.m[2] * x.z, a.m[3] * x.x + a.m[4] * x.y + a.m[5] * x.z,
            a.m[6] * x.x + a.m[7] * x.y + a.m[8] * x.z};
}

// ---------------------------------------------------------------- ordered float atomics
// Map a float to an unsigned key with the same order (used for global min/max of AABBs
// and of step bounds; min/max is order-independent, so the result is deterministic).
LCSX_HD uint32_t f2ord(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
LCSX_HD float ord2f(uint32_t u) { return __builtin_bit_cast(float, (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

// ---------------------------------------------------------------- wave64 primitives
LCSX_D uint32_t lane_id() { return __lane_id(); }
LCSX_D uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Appends `pred` lanes to a global list with one atomic per wave; returns this lane's slot.
LCSX_D uint32_t wave_append(uint32_t* counter, bool pred)
{
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    int      leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base   = 0;
    if ((int)lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & lanemask_lt());
}

template <class T>
LCSX_D T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o);
    return v;
}
LCSX_D float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
LCSX_D float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

}  // namespace lcsx
