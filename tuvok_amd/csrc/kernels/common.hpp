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

// Runs of voxels along x of the blocked layout.  A row holds its block's W
// voxels followed by the next block's first W (16 bytes), so voxels s .. s+m-1
// of a row (s < W, s + m <= 2W) are contiguous.
//   run4: voxels s .. s+3 of the row at byte offset r, from the dword-aligned
//         12 (uint16) or 8 (uint8) bytes that start at the dword holding voxel s;
//   run2: voxels x .. x+1 from byte offset r of voxel x itself, one 4-byte
//         (uint16) or 2-byte (uint8) under-aligned load: the amdhsa code
//         object runs with unaligned global access enabled, and the two voxels
//         never leave their 16-byte row.
// Measured on the lit 2D-TF 1024^3 uint16 workload (16 views, 1080p): all
// twelve loads as 16-byte rows 74.9 ms; these forms 71.0 ms; under-aligned
// 8-byte loads for the four central runs as well 76.9 ms.  Values are stored
// integers.
typedef uint32_t tvk_u32x3_a4 __attribute__((ext_vector_type(3), aligned(4)));
typedef uint32_t tvk_u32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef uint32_t tvk_u32_a1 __attribute__((aligned(1)));
typedef uint16_t tvk_u16_a1 __attribute__((aligned(1)));
template <typename T>
__device__ __forceinline__ void run4(const uint8_t* __restrict__ b, uint32_t r, int s, uint32_t o[4]) {
  if (sizeof(T) == 2) {
    const tvk_u32x3_a4 w = *reinterpret_cast<const tvk_u32x3_a4*>(b + (r + 4u * (uint32_t)(s >> 1)));
    const uint32_t sh = (uint32_t)(s & 1) * 2u;
    const uint32_t lo = __builtin_amdgcn_alignbyte(w.y, w.x, sh), hi = __builtin_amdgcn_alignbyte(w.z, w.y, sh);
    o[0] = lo & 0xffffu; o[1] = lo >> 16; o[2] = hi & 0xffffu; o[3] = hi >> 16;
  } else {
    const tvk_u32x2_a4 w = *reinterpret_cast<const tvk_u32x2_a4*>(b + (r + 4u * (uint32_t)(s >> 2)));
    const uint32_t v = __builtin_amdgcn_alignbyte(w.y, w.x, (uint32_t)(s & 3));
    o[0] = v & 0xffu; o[1] = (v >> 8) & 0xffu; o[2] = (v >> 16) & 0xffu; o[3] = v >> 24;
  }
}
template <typename T>
__device__ __forceinline__ void run2(const uint8_t* __restrict__ b, uint32_t r, uint32_t o[2]) {
  if (sizeof(T) == 2) {
    const uint32_t w = *reinterpret_cast<const tvk_u32_a1*>(b + r);
    o[0] = w & 0xffffu; o[1] = w >> 16;
  } else {
    const uint32_t w = *reinterpret_cast<const tvk_u16_a1*>(b + r);
    o[0] = w & 0xffu; o[1] = w >> 8;
  }
}

__device__ __forceinline__ V3 reflect(V3 I, V3 N) { return I - N * (2.0f * dot(N, I)); }

}  // namespace tvk
