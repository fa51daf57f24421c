// oracle_math.hpp -- TEST INFRASTRUCTURE ONLY (the CPU oracle).
//
// Scalar FP32 vector helpers and the reference's random number generator,
// restated from the reference's HLSL/Slang utility modules. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// oracle; the product path never links or calls it.
#pragma once
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };
struct i3 { int x, y, z; };

static inline f3 mk3(float a) { return {a, a, a}; }
static inline f3 mk3(float a, float b, float c) { return {a, b, c}; }
static inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
static inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
static inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline f3 operator+(f3 a, float s) { return {a.x + s, a.y + s, a.z + s}; }
static inline f3 operator-(f3 a, float s) { return {a.x - s, a.y - s, a.z - s}; }
static inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline float length(f3 a) { return std::sqrt(dot(a, a)); }
static inline f3 normalize(f3 a) { return a / length(a); }
static inline f3 f3floor(f3 a) { return {std::floor(a.x), std::floor(a.y), std::floor(a.z)}; }
static inline f3 f3abs(f3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
static inline f3 f3of(i3 a) { return {(float)a.x, (float)a.y, (float)a.z}; }
static inline i3 i3floor(f3 a) { return {(int)std::floor(a.x), (int)std::floor(a.y), (int)std::floor(a.z)}; }
static inline bool anyPositive(f3 a) { return a.x > 0.f || a.y > 0.f || a.z > 0.f; }
static inline bool allZero(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
static inline float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// UT/Color/ColorHelpers.slang:39-42
static inline float luminance(f3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

static const float kRayTMax = FLT_MAX;  // VR/VolumeBase.slang:11
static const float kPi = 3.14159265358979323846f;
static const float kInv4Pi = 0.07957747154594766788f;

// Row-vector affine transform as in HLSL mul(float4(v, w), M) (VR/VolumeBase.slang:103-130).
static inline f4 mulRow(f4 v, const float* M) {
    f4 r;
    r.x = v.x * M[0] + v.y * M[4] + v.z * M[8] + v.w * M[12];
    r.y = v.x * M[1] + v.y * M[5] + v.z * M[9] + v.w * M[13];
    r.z = v.x * M[2] + v.y * M[6] + v.z * M[10] + v.w * M[14];
    r.w = v.x * M[3] + v.y * M[7] + v.z * M[11] + v.w * M[15];
    return r;
}

static inline float asfloat(int v) { float f; std::memcpy(&f, &v, 4); return f; }
static inline int asint(float v) { int i; std::memcpy(&i, &v, 4); return i; }

// ---------------------------------------------------------------------------
// Random numbers.
// Morton interleave: UT/Math/BitTricks.slang:45-61.
static inline uint32_t interleave32(uint32_t x, uint32_t y) {
    x &= 0xffffu; y &= 0xffffu;
    x = (x | (x << 8)) & 0x00FF00FFu; x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u; x = (x | (x << 1)) & 0x55555555u;
    y = (y | (y << 8)) & 0x00FF00FFu; y = (y | (y << 4)) & 0x0F0F0F0Fu;
    y = (y | (y << 2)) & 0x33333333u; y = (y | (y << 1)) & 0x55555555u;
    return x | (y << 1);
}
// SplitMix64: UT/Sampling/Pseudorandom/SplitMix64.slang:54-78.
static inline uint64_t splitmix64Next(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// xoshiro128**: UT/Sampling/Pseudorandom/Xoshiro.slang:45-66.
struct SampleGen { uint32_t s[4]; };
static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
static inline uint32_t xoshiroNext(SampleGen& g) {
    const uint32_t result = rotl32(g.s[0] * 5u, 7) * 9u;
    const uint32_t t = g.s[1] << 9;
    g.s[2] ^= g.s[0]; g.s[3] ^= g.s[1]; g.s[1] ^= g.s[2]; g.s[0] ^= g.s[3];
    g.s[2] ^= t; g.s[3] = rotl32(g.s[3], 11);
    return result;
}
// UniformSampleGenerator::create: UT/Sampling/UniformSampleGenerator.slang:49-68.
static inline SampleGen makeSampleGen(uint32_t px, uint32_t py, uint32_t sampleNumber) {
    uint64_t st = (uint64_t)interleave32(px, py) | ((uint64_t)sampleNumber << 32);
    uint64_t s0 = splitmix64Next(st), s1 = splitmix64Next(st);
    SampleGen g;
    g.s[0] = (uint32_t)s0; g.s[1] = (uint32_t)(s0 >> 32);
    g.s[2] = (uint32_t)s1; g.s[3] = (uint32_t)(s1 >> 32);
    return g;
}
// sampleNext1D: UT/Sampling/SampleGenerator.slang:58-64.
static inline float next1D(SampleGen& g) { return (float)(xoshiroNext(g) >> 8) * 0x1p-24f; }

}  // namespace orc
