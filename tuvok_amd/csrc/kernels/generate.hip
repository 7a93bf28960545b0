// Seeded synthetic volumes, generated directly in HBM (x fastest).
// Voxel (i,j,k) sits at p = ((i,j,k)+0.5)/N in the unit cube.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "generate.hpp"

namespace tvk {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float u01(uint64_t h) { return (float)(h >> 40) * (1.0f / 16777216.0f); }

__device__ __forceinline__ void voxel_pos(uint64_t i, uint64_t X, uint64_t Y, uint64_t Z, float p[3]) {
  const uint64_t x = i % X, r = i / X, y = r % Y, z = r / Y;
  p[0] = ((float)x + 0.5f) / (float)X;
  p[1] = ((float)y + 0.5f) / (float)Y;
  p[2] = ((float)z + 0.5f) / (float)Z;
}

// ---- Gaussian blobs ---------------------------------------------------------
__device__ __forceinline__ float blob_field(const GenBlobs& g, const float p[3]) {
  float f = 0.f;
  for (int b = 0; b < g.n; ++b) {
    const float dx = p[0] - g.c[b][0], dy = p[1] - g.c[b][1], dz = p[2] - g.c[b][2];
    f += g.amp[b] * __expf(-(dx * dx + dy * dy + dz * dz) * g.inv2s2[b]);
  }
  return f;
}

__global__ __launch_bounds__(256) void blobs_max(GenBlobs g, uint64_t X, uint64_t Y, uint64_t Z, unsigned* mx) {
  float best = 0.f;
  const uint64_t n = X * Y * Z;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    float p[3];
    voxel_pos(i, X, Y, Z, p);
    best = fmaxf(best, blob_field(g, p));
  }
  for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(mx, __float_as_uint(best));  // positive floats order as uints
}

template <typename T>
__global__ __launch_bounds__(256) void blobs_fill(GenBlobs g, uint64_t X, uint64_t Y, uint64_t Z, const unsigned* mx,
                                                  T* out, float full, uint64_t seed) {
  const float inv = 1.0f / __uint_as_float(*mx);
  const uint64_t n = X * Y * Z;
  const float unit = full / 255.0f;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    float p[3];
    voxel_pos(i, X, Y, Z, p);
    float v = rintf(full * blob_field(g, p) * inv);
    const int noise = (int)(mix64(seed ^ (i * 0x9e3779b97f4a7c15ull)) % 9ull) - 4;  // uniform in [-4, 4]
    v += (float)noise * unit;
    v = fminf(fmaxf(v, 0.f), full);
    out[i] = (T)v;
  }
}

// ---- fBm value noise + hollow shells -----------------------------------------
__device__ __forceinline__ float lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
  uint64_t h = mix64(seed ^ ((uint64_t)x * 0x8da6b343ull) ^ ((uint64_t)y * 0xd8163841ull) ^ ((uint64_t)z * 0xcb1ab31full));
  return u01(h);
}
__device__ __forceinline__ float value_noise(float x, float y, float z, uint64_t seed) {
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
  float tx = x - fx, ty = y - fy, tz = z - fz;
  tx = tx * tx * (3.f - 2.f * tx);
  ty = ty * ty * (3.f - 2.f * ty);
  tz = tz * tz * (3.f - 2.f * tz);
  float c[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) c[k] = lattice(ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2), seed);
  const float c00 = c[0] + (c[1] - c[0]) * tx, c10 = c[2] + (c[3] - c[2]) * tx;
  const float c01 = c[4] + (c[5] - c[4]) * tx, c11 = c[6] + (c[7] - c[6]) * tx;
  const float c0 = c00 + (c10 - c00) * ty, c1 = c01 + (c11 - c01) * ty;
  return c0 + (c1 - c0) * tz;
}

template <typename T>
__global__ __launch_bounds__(256) void fbm_shells(GenShells g, uint64_t X, uint64_t Y, uint64_t Z, T* out, float full) {
  const uint64_t n = X * Y * Z;
  const float nmax = (float)(X > Y ? (X > Z ? X : Z) : (Y > Z ? Y : Z));
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    float p[3];
    voxel_pos(i, X, Y, Z, p);
    float f = 0.f, amp = 1.f, freq = g.base_freq, norm = 0.f;
    for (int o = 0; o < g.octaves; ++o) {
      f += amp * value_noise(p[0] * freq, p[1] * freq, p[2] * freq, g.seed + (uint64_t)o * 7919ull);
      norm += amp;
      amp *= 0.5f;
      freq *= 2.f;
    }
    f /= norm;
    // stretch the fBm histogram (concentrated around 0.5) over the full range
    f = fminf(fmaxf((f - 0.2f) / 0.6f, 0.f), 1.f);
    float v = f * full;
    for (int s = 0; s < g.n; ++s) {
      const float dx = p[0] - g.c[s][0], dy = p[1] - g.c[s][1], dz = p[2] - g.c[s][2];
      const float r = sqrtf(dx * dx + dy * dy + dz * dz);
      if (fabsf(r - g.r[s]) * nmax < g.half_thickness_vox) v = g.shell_value;
    }
    out[i] = (T)fminf(v, full);
  }
}

// ---- signed distance of random ellipsoids: clamp(128 + 20 * d_voxels) ---------
__global__ __launch_bounds__(256) void ellipsoid_sdf(const GenEllipsoid* __restrict__ e, int ne, uint64_t X, uint64_t Y,
                                                     uint64_t Z, uint8_t* out, float scale) {
  const uint64_t n = X * Y * Z;
  const float nmax = (float)(X > Y ? (X > Z ? X : Z) : (Y > Z ? Y : Z));
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    float p[3];
    voxel_pos(i, X, Y, Z, p);
    float best = -1e30f;  // positive inside
    for (int k = 0; k < ne; ++k) {
      const float qx = (p[0] - e[k].c[0]) / e[k].a[0], qy = (p[1] - e[k].c[1]) / e[k].a[1],
                  qz = (p[2] - e[k].c[2]) / e[k].a[2];
      const float q = sqrtf(qx * qx + qy * qy + qz * qz);
      const float d = (1.f - q) * e[k].amin;  // first-order distance to the surface
      best = fmaxf(best, d);
    }
    float v = 128.f + scale * best * nmax;
    out[i] = (uint8_t)fminf(fmaxf(rintf(v), 0.f), 255.f);
  }
}

// ---- nested-ellipsoid phantom + Gaussian noise ---------------------------------
__global__ __launch_bounds__(256) void phantom(const GenEllipsoid* __restrict__ e, int ne, uint64_t X, uint64_t Y,
                                               uint64_t Z, uint16_t* out, float sigma, uint64_t seed) {
  const uint64_t n = X * Y * Z;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    float p[3];
    voxel_pos(i, X, Y, Z, p);
    float v = 0.f;
    for (int k = 0; k < ne; ++k) {
      const float qx = (p[0] - e[k].c[0]) / e[k].a[0], qy = (p[1] - e[k].c[1]) / e[k].a[1],
                  qz = (p[2] - e[k].c[2]) / e[k].a[2];
      if (qx * qx + qy * qy + qz * qz <= 1.f) v = e[k].value;  // later (inner) ellipsoids win
    }
    const uint64_t h = mix64(seed ^ (i * 0x9e3779b97f4a7c15ull));
    const float u1 = fmaxf(u01(h), 1e-7f), u2 = u01(mix64(h));
    const float gauss = sqrtf(-2.f * __logf(u1)) * __cosf(6.2831853f * u2);
    v = rintf(v + sigma * gauss);
    out[i] = (uint16_t)fminf(fmaxf(v, 0.f), 65535.f);
  }
}

static uint32_t blocks_for(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (uint32_t)(b > 65536 ? 65536 : b);
}

hipError_t launch_blobs(const GenBlobs& g, const uint64_t d[3], int bpv, void* out, unsigned* d_tmp, uint64_t seed,
                        hipStream_t st) {
  const uint64_t n = d[0] * d[1] * d[2];
  hipMemsetAsync(d_tmp, 0, sizeof(unsigned), st);
  hipLaunchKernelGGL(blobs_max, dim3(blocks_for(n)), dim3(256), 0, st, g, d[0], d[1], d[2], d_tmp);
  if (bpv == 1)
    hipLaunchKernelGGL(blobs_fill<uint8_t>, dim3(blocks_for(n)), dim3(256), 0, st, g, d[0], d[1], d[2], d_tmp,
                       (uint8_t*)out, 255.f, seed);
  else
    hipLaunchKernelGGL(blobs_fill<uint16_t>, dim3(blocks_for(n)), dim3(256), 0, st, g, d[0], d[1], d[2], d_tmp,
                       (uint16_t*)out, 65535.f, seed);
  return hipGetLastError();
}

hipError_t launch_fbm_shells(const GenShells& g, const uint64_t d[3], int bpv, void* out, hipStream_t st) {
  const uint64_t n = d[0] * d[1] * d[2];
  if (bpv == 1)
    hipLaunchKernelGGL(fbm_shells<uint8_t>, dim3(blocks_for(n)), dim3(256), 0, st, g, d[0], d[1], d[2], (uint8_t*)out,
                       255.f);
  else
    hipLaunchKernelGGL(fbm_shells<uint16_t>, dim3(blocks_for(n)), dim3(256), 0, st, g, d[0], d[1], d[2],
                       (uint16_t*)out, 65535.f);
  return hipGetLastError();
}

hipError_t launch_ellipsoid_sdf(const GenEllipsoid* d_e, int ne, const uint64_t d[3], void* out, hipStream_t st) {
  const uint64_t n = d[0] * d[1] * d[2];
  hipLaunchKernelGGL(ellipsoid_sdf, dim3(blocks_for(n)), dim3(256), 0, st, d_e, ne, d[0], d[1], d[2], (uint8_t*)out,
                     20.f);
  return hipGetLastError();
}

hipError_t launch_phantom(const GenEllipsoid* d_e, int ne, const uint64_t d[3], void* out, float sigma, uint64_t seed,
                          hipStream_t st) {
  const uint64_t n = d[0] * d[1] * d[2];
  hipLaunchKernelGGL(phantom, dim3(blocks_for(n)), dim3(256), 0, st, d_e, ne, d[0], d[1], d[2], (uint16_t*)out, sigma,
                     seed);
  return hipGetLastError();
}

}  // namespace tvk
