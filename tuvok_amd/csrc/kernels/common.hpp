// Device helpers shared by the march kernels: 3-vectors, unorm texel values,
// the blocked brick layout's element offsets.  Float operations here are part
// of what the rendered parity depends on; the files that include this are
// compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include "../tvk_types.hpp"

namespace tvk {

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 mk(float a, float b, float c) { V3 r; r.x = a; r.y = b; r.z = c; return r; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ V3 operator/(V3 a, V3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }

// sqrtf(x) for x >= 0: the hardware square root (within 1 ulp) moved to the
// neighbour whose residual brackets x, which is the correctly rounded root (the
// same steps as the compiler's sqrtf).  sqrtf also rescales every input below
// 2^-96, where the residuals would underflow; here such inputs (rare: the
// squared gradient of a nearly flat neighbourhood) take sqrtf itself.
__device__ __forceinline__ float sqrt_rn(float x) {
  if (__builtin_expect(x < 0x1p-96f, 0)) return sqrtf(x);
  const float s = __builtin_amdgcn_sqrtf(x);
  const float lo = __builtin_bit_cast(float, __builtin_bit_cast(int, s) - 1);
  const float hi = __builtin_bit_cast(float, __builtin_bit_cast(int, s) + 1);
  const float rlo = fmaf(-lo, s, x), rhi = fmaf(-hi, s, x);
  return rhi > 0.0f ? hi : (rlo <= 0.0f ? lo : s);
}
// GLSL pow(y, s) of the opacity correction 1 - (1 - a)^s: exponents that are
// powers of two from 1/4 to 16 (a sample-rate modifier of 2^k, a level's 2^k
// domain ratio) by correctly rounded square roots and repeated squaring, any
// other by powf -- the oracle's steps, so the two agree bit for bit.  s is the
// same for every lane of a brick, so the branches are uniform in the march.
__device__ __forceinline__ float opacity_pow(float y, float s) {
  if (s == 1.0f) return y;
  if (s == 0.5f) return sqrt_rn(y);
  if (s == 0.25f) return sqrt_rn(sqrt_rn(y));
  if (s == 2.0f || s == 4.0f || s == 8.0f || s == 16.0f) {
    float r = y * y;
    for (float e = 4.0f; e <= s; e *= 2.0f) r = r * r;
    return r;
  }
  return powf(y, s);
}
// GLSL normalize(): one correctly rounded reciprocal of the length, then three products
__device__ __forceinline__ V3 normalize(V3 a) { const float r = 1.0f / length(a); return a * r; }
__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
__device__ __forceinline__ float pow8(float x) { float x2 = x * x; float x4 = x2 * x2; return x4 * x4; }
// The entry point goes through an RGBA16F target: the fp32 value is rounded to
// fp16.  The empty asm keeps the product that feeds it an fp32 value (otherwise
// the mixed-precision FMA rounds the exact product to fp16 in one step).
__device__ __forceinline__ float fp16_round(float x) {
  asm volatile("" : "+v"(x));
  return (float)(_Float16)x;
}

// value of a stored voxel as the unorm texture returns it
template <typename T> struct Unorm;
template <> struct Unorm<uint8_t> {
  static __device__ __forceinline__ float inv() { return 1.0f / 255.0f; }
  static __device__ __forceinline__ uint32_t max_value() { return 255u; }
};
template <> struct Unorm<uint16_t> {
  static __device__ __forceinline__ float inv() { return 1.0f / 65535.0f; }
  static __device__ __forceinline__ uint32_t max_value() { return 65535u; }
};

// Element offset parts of the blocked brick layout (see tvk_types.hpp): the
// offset of voxel (x,y,z) is ox(x) + oy(y) + oz(z); rows hold 2 << LX elements.
template <typename T> struct Blk { static constexpr int LX = sizeof(T) == 1 ? 3 : 2; };
template <typename T>
__device__ __forceinline__ uint32_t ox(int x) {
  constexpr int L = Blk<T>::LX;
  return ((uint32_t)(x >> L) << (L + 5)) + (uint32_t)(x & ((1 << L) - 1));
}
template <typename T>
__device__ __forceinline__ uint32_t oy(int y, uint32_t nbx) {
  constexpr int L = Blk<T>::LX;
  return (((uint32_t)(y >> 2) * nbx) << (L + 5)) + ((uint32_t)(y & 3) << (L + 1));
}
template <typename T>
__device__ __forceinline__ uint32_t oz(int z, uint32_t nbxy) {
  constexpr int L = Blk<T>::LX;
  return (((uint32_t)(z >> 2) * nbxy) << (L + 5)) + ((uint32_t)(z & 3) << (L + 3));
}

__device__ __forceinline__ V3 reflect(V3 I, V3 N) { return I - N * (2.0f * dot(N, I)); }

}  // namespace tvk
