// Fixed-step brick raycaster for gfx950.
//
// One wavefront = one 8x8 pixel tile = 64 rays (one per lane); a 256-thread
// workgroup owns a 16x16 pixel group.  Workgroups are remapped so that every
// XCD (own L2) renders a contiguous band of the screen.
//
// Per tile, the visible non-empty bricks of the frame are scanned in
// compositing order 64 at a time: every lane tests one brick's conservative
// screen rectangle against the tile, a ballot yields the candidate mask, and
// the wave walks the set bits in order (wave-uniform brick index -> scalar
// loads of the 128-byte brick record).  For each brick every lane intersects
// its own ray with the brick box and marches the segment exactly as
// GLRaycaster does it per brick (GLRaycaster.cpp:383-478 with
// GLRaycaster-*-FS.glsl): fp16-rounded eye-space entry, iStepCount =
// int(L/step)+1, fixed texture-space increment, trilinear fetch from the
// brick-contiguous pool with GL_CLAMP border emulation, TF lookup (1D table in
// LDS, 2D table via L2), opacity correction, under-compositing, local
// 'a >= 0.99' exit, then (ONE_MINUS_DST_ALPHA, ONE) blending into the ray's
// accumulator.  With ert_eps > 0 a ray stops once its total transmittance is
// below ert_eps and a wave leaves the brick loop once every lane has stopped.
//
// This file is compiled with -ffp-contract=off: the float operations are the
// ones of the host planner and of the reference restatement, in the same
// order, so step counts, sample positions and TF indices are reproducible.
#include <hip/hip_runtime.h>
#include <type_traits>
#include "../tvk_types.hpp"
#include "stepnet.hpp"
#include "common.hpp"

namespace tvk {

// Trilinear interpolation of eight texels: x lerps, then y, then z, each
// lerp a + f (b - a) as one fused multiply-add.  GL leaves the filter's
// arithmetic to the implementation; the oracle uses this same form.
__device__ __forceinline__ float lerp_fma(float a, float b, float f) { return fmaf(f, b - a, a); }
__device__ __forceinline__ float tri_lerp(float c000, float c100, float c010, float c110, float c001, float c101,
                                          float c011, float c111, float f0, float f1, float f2) {
  const float c00 = lerp_fma(c000, c100, f0), c10 = lerp_fma(c010, c110, f0);
  const float c01 = lerp_fma(c001, c101, f0), c11 = lerp_fma(c011, c111, f0);
  const float c0 = lerp_fma(c00, c10, f1), c1 = lerp_fma(c01, c11, f1);
  return lerp_fma(c0, c1, f2);
}

// GL_LINEAR filter of the footprint with lower corner i0 and weights f, with
// GL_CLAMP (border colour 0) or GL_CLAMP_TO_EDGE
template <typename T>
__device__ __forceinline__ float filter_brick(const T* __restrict__ d, int nx, int ny, int nz, bool edge,
                                              const int i0[3], const float f[3]) {
  const float inv = Unorm<T>::inv();
  // per-axis tap indices (clamped for the address) and validity (border)
  int xa = i0[0], xb = i0[0] + 1, ya = i0[1], yb = i0[1] + 1, za = i0[2], zb = i0[2] + 1;
  bool vxa = xa >= 0 && xa < nx, vxb = xb >= 0 && xb < nx, vya = ya >= 0 && ya < ny;
  bool vyb = yb >= 0 && yb < ny, vza = za >= 0 && za < nz, vzb = zb >= 0 && zb < nz;
  if (edge) { vxa = vxb = vya = vyb = vza = vzb = true; }
  xa = min(max(xa, 0), nx - 1); xb = min(max(xb, 0), nx - 1); ya = min(max(ya, 0), ny - 1);
  yb = min(max(yb, 0), ny - 1); za = min(max(za, 0), nz - 1); zb = min(max(zb, 0), nz - 1);
  const uint32_t nbx = nblocks((uint32_t)nx, Blk<T>::LX), nbxy = nbx * nblocks((uint32_t)ny, 2);
  const uint32_t xA = ox<T>(xa), xB = ox<T>(xb), yA = oy<T>(ya, nbx), yB = oy<T>(yb, nbx);
  const uint32_t zA = oz<T>(za, nbxy), zB = oz<T>(zb, nbxy);
  // the clamped addresses are always inside the brick: load unconditionally and
  // select the border colour afterwards (no per-tap branch)
  // taps as stored integers (the unorm scale is applied once to the result)
  auto tap = [&](bool ok, uint32_t off) { const float v = (float)d[off]; return v * (ok ? 1.0f : 0.0f); };
  float c000 = tap(vxa && vya && vza, zA + yA + xA);
  float c100 = tap(vxb && vya && vza, zA + yA + xB);
  float c010 = tap(vxa && vyb && vza, zA + yB + xA);
  float c110 = tap(vxb && vyb && vza, zA + yB + xB);
  float c001 = tap(vxa && vya && vzb, zB + yA + xA);
  float c101 = tap(vxb && vya && vzb, zB + yA + xB);
  float c011 = tap(vxa && vyb && vzb, zB + yB + xA);
  float c111 = tap(vxb && vyb && vzb, zB + yB + xB);
  return tri_lerp(c000, c100, c010, c110, c001, c101, c011, c111, f[0], f[1], f[2]) * inv;
}

// lower corner and weights of the GL_LINEAR footprint at texture coordinate p
__device__ __forceinline__ void footprint(V3 p, int nx, int ny, int nz, int i0[3], float f[3]) {
  const int n[3] = {nx, ny, nz};
  const float pc[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float s = fminf(fmaxf(pc[a], 0.0f), 1.0f);
    const float u = s * (float)n[a] - 0.5f;
    const float fl = floorf(u);
    i0[a] = (int)fl;
    f[a] = u - fl;
  }
}

// texture3D with GL_LINEAR and GL_CLAMP (border colour 0) or GL_CLAMP_TO_EDGE
template <typename T>
__device__ __forceinline__ float sample_brick(const T* __restrict__ d, int nx, int ny, int nz, bool edge, V3 p) {
  int i0[3];
  float f[3];
  footprint(p, nx, ny, nz, i0, f);
  return filter_brick<T>(d, nx, ny, nz, edge, i0, f);
}

// Volume3D.glsl:43-53: 6-tap central difference at c +- h, h = one texel
// (1 / texture size, GLRaycaster.cpp:242-251), halved.  GL_LINEAR at s + 1/n is
// the footprint of s moved by one texel with the same weights, so on an axis
// where both taps stay inside [0, 1] the taps are the centre's footprint +-1
// texel on that axis with the centre's weights; on an axis where a tap leaves
// [0, 1] (its coordinate clamps) both taps are sampled at c +- h as stated.
__device__ __forceinline__ bool taps_unclamped(float c, float h) { return c - h >= 0.0f && c + h <= 1.0f; }

template <typename T>
__device__ __forceinline__ V3 gradient(const T* __restrict__ d, int nx, int ny, int nz, bool edge, V3 c, V3 h) {
  int i0[3];
  float f[3];
  footprint(c, nx, ny, nz, i0, f);
  const float cc[3] = {c.x, c.y, c.z}, hh[3] = {h.x, h.y, h.z};
  float plus[3], minus[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (taps_unclamped(cc[a], hh[a])) {
      int ip[3] = {i0[0], i0[1], i0[2]}, im[3] = {i0[0], i0[1], i0[2]};
      ip[a] += 1;
      im[a] -= 1;
      plus[a] = filter_brick<T>(d, nx, ny, nz, edge, ip, f);
      minus[a] = filter_brick<T>(d, nx, ny, nz, edge, im, f);
    } else {
      const V3 e = a == 0 ? mk(hh[0], 0, 0) : (a == 1 ? mk(0, hh[1], 0) : mk(0, 0, hh[2]));
      plus[a] = sample_brick<T>(d, nx, ny, nz, edge, c + e);
      minus[a] = sample_brick<T>(d, nx, ny, nz, edge, c - e);
    }
  }
  // the shader's taps: xp = +x, xm = -x, yp = -y, ym = +y, zp = +z, zm = -z
  return mk(minus[0] - plus[0], minus[1] - plus[1], minus[2] - plus[2]) / 2.0f;
}

// Centre sample plus gradient()'s six taps from one 32-voxel cross: the
// centre footprint and, per axis, the footprints one texel below and above it
// with the centre's weights (the taps of an axis whose coordinates clamp at the
// texture edge are resampled at p +- h, as gradient() does).  value() first
// loads the centre cube only, so samples whose TF entry is fully transparent
// never pay for the gradient.
// sample_brick's interpolation of stored integers, scaled to unorm once
template <typename T>
__device__ __forceinline__ float tri(float c000, float c100, float c010, float c110, float c001, float c101, float c011,
                                     float c111, float f0, float f1, float f2) {
  return tri_lerp(c000, c100, c010, c110, c001, c101, c011, c111, f0, f1, f2) * Unorm<T>::inv();
}

// FAST: every tap of the 4x4x4 neighbourhood lies inside the brick for every
// lane of the wave (decided by a ballot before the sample), so no coordinate
// clamps or border selects are needed; the values are the same.
template <typename T, bool FAST = false>
struct CrossSampler {
  const T* __restrict__ d;
  int n[3];
  bool edge;
  V3 p;
  int ic[3];
  float fc[3];
  float c[2][2][2];
  // byte offset parts and border validity of indices ic-1 .. ic+2 per axis
  uint32_t ofx[4], ofy[4], ofz[4];
  bool okx[4], oky[4], okz[4];

  __device__ __forceinline__ float V(int i, int j, int k) const {   // i,j,k in 0..3 -> index ic-1+i ...
    // offsets are clamped into the brick: unconditional load, then the border select
    // byte offsets below 2^32: a uniform base plus a 32-bit lane offset
    const float v = (float)*reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(d) + (ofz[k] + ofy[j] + ofx[i]));
    if (FAST) return v;
    // border colour 0 as a product with 0/1 (v is finite and >= 0), so that the
    // load stays unconditional and all taps are in flight at once
    return v * ((okx[i] && oky[j] && okz[k]) ? 1.0f : 0.0f);
  }

  __device__ __forceinline__ float value(const T* __restrict__ dd, int nx, int ny, int nz, bool e, V3 pos) {
    d = dd; n[0] = nx; n[1] = ny; n[2] = nz; edge = e; p = pos;
    const float pc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float s = fminf(fmaxf(pc[a], 0.0f), 1.0f), u = s * (float)n[a] - 0.5f, fl = floorf(u);
      ic[a] = (int)fl;
      fc[a] = u - fl;
    }
    const uint32_t nbx = nblocks((uint32_t)nx, Blk<T>::LX), nbxy = nbx * nblocks((uint32_t)ny, 2);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int x = ic[0] - 1 + k, y = ic[1] - 1 + k, z = ic[2] - 1 + k;
      if (FAST) {
        ofx[k] = ox<T>(x) * (uint32_t)sizeof(T);
        ofy[k] = oy<T>(y, nbx) * (uint32_t)sizeof(T);
        ofz[k] = oz<T>(z, nbxy) * (uint32_t)sizeof(T);
      } else {
        okx[k] = edge || (x >= 0 && x < nx);
        oky[k] = edge || (y >= 0 && y < ny);
        okz[k] = edge || (z >= 0 && z < nz);
        ofx[k] = ox<T>(min(max(x, 0), nx - 1)) * (uint32_t)sizeof(T);
        ofy[k] = oy<T>(min(max(y, 0), ny - 1), nbx) * (uint32_t)sizeof(T);
        ofz[k] = oz<T>(min(max(z, 0), nz - 1), nbxy) * (uint32_t)sizeof(T);
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) c[i][j][k] = V(1 + i, 1 + j, 1 + k);
    return tri<T>(c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0], c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1], fc[0],
               fc[1], fc[2]);
  }

  __device__ __forceinline__ V3 gradient(V3 h) const {
    // the index form of all six taps from the cross (gradient()'s rule); an
    // axis whose taps clamp (rare: a sample next to the texture edge) is
    // resampled at p +- h below
    const float pc[3] = {p.x, p.y, p.z}, hc[3] = {h.x, h.y, h.z};
    float xl[2][2], xr[2][2], yl[2][2], yr[2][2], zl[2][2], zr[2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        xl[j][k] = V(0, 1 + j, 1 + k); xr[j][k] = V(3, 1 + j, 1 + k);
        yl[j][k] = V(1 + j, 0, 1 + k); yr[j][k] = V(1 + j, 3, 1 + k);
        zl[j][k] = V(1 + j, 1 + k, 0); zr[j][k] = V(1 + j, 1 + k, 3);
      }
    float fxm = tri<T>(xl[0][0], c[0][0][0], xl[1][0], c[0][1][0], xl[0][1], c[0][0][1], xl[1][1], c[0][1][1],
                          fc[0], fc[1], fc[2]);
    float fxp = tri<T>(c[1][0][0], xr[0][0], c[1][1][0], xr[1][0], c[1][0][1], xr[0][1], c[1][1][1], xr[1][1],
                          fc[0], fc[1], fc[2]);
    float fym = tri<T>(yl[0][0], yl[1][0], c[0][0][0], c[1][0][0], yl[0][1], yl[1][1], c[0][0][1], c[1][0][1],
                          fc[0], fc[1], fc[2]);
    float fyp = tri<T>(c[0][1][0], c[1][1][0], yr[0][0], yr[1][0], c[0][1][1], c[1][1][1], yr[0][1], yr[1][1],
                          fc[0], fc[1], fc[2]);
    float fzm = tri<T>(zl[0][0], zl[1][0], zl[0][1], zl[1][1], c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0],
                          fc[0], fc[1], fc[2]);
    float fzp = tri<T>(c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1], zr[0][0], zr[1][0], zr[0][1], zr[1][1],
                          fc[0], fc[1], fc[2]);
    if (!(taps_unclamped(pc[0], hc[0]) && taps_unclamped(pc[1], hc[1]) && taps_unclamped(pc[2], hc[2]))) {
#pragma unroll 1
      for (int a = 0; a < 3; ++a) {
        if (taps_unclamped(pc[a], hc[a])) continue;
        const V3 e = a == 0 ? mk(hc[0], 0, 0) : (a == 1 ? mk(0, hc[1], 0) : mk(0, 0, hc[2]));
        const float tp = sample_brick<T>(d, n[0], n[1], n[2], edge, p + e);
        const float tm = sample_brick<T>(d, n[0], n[1], n[2], edge, p - e);
        if (a == 0) { fxp = tp; fxm = tm; } else if (a == 1) { fyp = tp; fym = tm; } else { fzp = tp; fzm = tm; }
      }
    }
    return mk(fxm - fxp, fym - fyp, fzm - fzp) / 2.0f;
  }
};


// sample_brick for the isosurface search: where every active lane's eight taps
// lie inside the brick (no border colour, no clamping), each x pair is one
// run2 load -- four loads instead of eight gathers -- and the weights, values
// and float operations are sample_brick's; otherwise sample_brick itself.
template <typename T>
__device__ __forceinline__ float sample_brick_iso(const T* __restrict__ d, int nx, int ny, int nz, bool edge, V3 p) {
  int i0[3];
  float f[3];
  const int n[3] = {nx, ny, nz};
  const float pc[3] = {p.x, p.y, p.z};
  bool in = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float s = fminf(fmaxf(pc[a], 0.0f), 1.0f), u = s * (float)n[a] - 0.5f, fl = floorf(u);
    i0[a] = (int)fl;
    f[a] = u - fl;
    in = in && i0[a] >= 0 && i0[a] + 1 < n[a];
  }
  if (__ballot(!in) != 0) return sample_brick<T>(d, nx, ny, nz, edge, p);
  const uint32_t nbx = nblocks((uint32_t)nx, Blk<T>::LX), nbxy = nbx * nblocks((uint32_t)ny, 2);
  const uint8_t* __restrict__ b = reinterpret_cast<const uint8_t*>(d);
  const uint32_t x0 = ox<T>(i0[0]);
  float c[2][2][2];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      uint32_t v[2];
      run2<T>(b, (x0 + oy<T>(i0[1] + j, nbx) + oz<T>(i0[2] + k, nbxy)) * (uint32_t)sizeof(T), v);
      c[0][j][k] = (float)v[0];
      c[1][j][k] = (float)v[1];
    }
  return tri_lerp(c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0], c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1],
                  f[0], f[1], f[2]) * Unorm<T>::inv();
}

// ISO_BATCH isosurface search samples at once (p, p + inc, ...): when every
// lane's 2x2x2 neighbourhoods of all of them lie inside the brick, their
// x-pair loads are issued together before any value is formed; the positions
// (accumulated one increment at a time), values and float operations are
// sample_brick_iso's.
#ifndef TVK_ISO_BATCH
#define TVK_ISO_BATCH 4
#endif
constexpr int ISO_BATCH = TVK_ISO_BATCH;
template <typename T, int N = ISO_BATCH>
__device__ __forceinline__ void sample_brick_iso_n(const T* __restrict__ d, int nx, int ny, int nz, bool edge, V3 p,
                                                   V3 inc, float (&val)[N]) {
  V3 ps[N];
  ps[0] = p;
#pragma unroll
  for (int s = 1; s < N; ++s) ps[s] = ps[s - 1] + inc;
  int i0[N][3];
  float f[N][3];
  const int n[3] = {nx, ny, nz};
  bool in = true;
#pragma unroll
  for (int s = 0; s < N; ++s) {
    const float pc[3] = {ps[s].x, ps[s].y, ps[s].z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float c = fminf(fmaxf(pc[a], 0.0f), 1.0f), u = c * (float)n[a] - 0.5f, fl = floorf(u);
      i0[s][a] = (int)fl;
      f[s][a] = u - fl;
      in = in && i0[s][a] >= 0 && i0[s][a] + 1 < n[a];
    }
  }
  if (__ballot(!in) != 0) {
#pragma unroll
    for (int s = 0; s < N; ++s) val[s] = sample_brick_iso<T>(d, nx, ny, nz, edge, ps[s]);
    return;
  }
  const uint32_t nbx = nblocks((uint32_t)nx, Blk<T>::LX), nbxy = nbx * nblocks((uint32_t)ny, 2);
  const uint8_t* __restrict__ b = reinterpret_cast<const uint8_t*>(d);
  uint32_t v[N][2][2][2];
#pragma unroll
  for (int s = 0; s < N; ++s) {
    const uint32_t x0 = ox<T>(i0[s][0]);
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k)
        run2<T>(b, (x0 + oy<T>(i0[s][1] + j, nbx) + oz<T>(i0[s][2] + k, nbxy)) * (uint32_t)sizeof(T), v[s][j][k]);
  }
#pragma unroll
  for (int s = 0; s < N; ++s) {
    float c[2][2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        c[0][j][k] = (float)v[s][j][k][0];
        c[1][j][k] = (float)v[s][j][k][1];
      }
    val[s] = tri_lerp(c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0], c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1],
                      f[s][0], f[s][1], f[s][2]) * Unorm<T>::inv();
  }
}

// Interior form (every lane's 4x4x4 neighbourhood inside the brick): the four
// central x-runs (x0-1 .. x0+2 at y0..y0+1, z0..z0+1) are one load each and
// give the centre cube and the x arms; each y or z arm (x0, x0+1) is one more
// load.  12 loads instead of 32 gathers; the voxels, weights and float
// operations are those of the general form.
template <typename T>
struct CrossSampler<T, true> {
  const T* __restrict__ d;
  int n[3];
  bool edge;
  V3 p;
  int ic[3];
  float fc[3];
  float c[2][2][2];
  float xl[2][2], xr[2][2];   // x arms at (y0+j, z0+k)
  uint32_t nbx, nbxy;

  // byte offsets of voxel (x, y, z) and of the start of its row
  __device__ __forceinline__ uint32_t vox(int x, int y, int z) const {
    return (ox<T>(x) + oy<T>(y, nbx) + oz<T>(z, nbxy)) * (uint32_t)sizeof(T);
  }
  __device__ __forceinline__ uint32_t rowo(int x, int y, int z) const {
    return (ox<T>(x & ~((1 << Blk<T>::LX) - 1)) + oy<T>(y, nbx) + oz<T>(z, nbxy)) * (uint32_t)sizeof(T);
  }
  __device__ __forceinline__ const uint8_t* base() const { return reinterpret_cast<const uint8_t*>(d); }

  __device__ __forceinline__ float value(const T* __restrict__ dd, int nx, int ny, int nz, bool e, V3 pos) {
    d = dd; n[0] = nx; n[1] = ny; n[2] = nz; edge = e; p = pos;
    const float pc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float s = fminf(fmaxf(pc[a], 0.0f), 1.0f), u = s * (float)n[a] - 0.5f, fl = floorf(u);
      ic[a] = (int)fl;
      fc[a] = u - fl;
    }
    nbx = nblocks((uint32_t)nx, Blk<T>::LX);
    nbxy = nbx * nblocks((uint32_t)ny, 2);
    const int xs = ic[0] - 1, sx = xs & ((1 << Blk<T>::LX) - 1);
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        uint32_t v[4];
        run4<T>(base(), rowo(xs, ic[1] + j, ic[2] + k), sx, v);
        xl[j][k] = (float)v[0];
        c[0][j][k] = (float)v[1];
        c[1][j][k] = (float)v[2];
        xr[j][k] = (float)v[3];
      }
    return tri<T>(c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0], c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1],
                  fc[0], fc[1], fc[2]);
  }

  // every tap of an interior sample is unclamped (1 <= floor(u) and
  // floor(u) + 2 < n on every axis put p +- 1/n inside [0, 1]): the index form
  __device__ __forceinline__ V3 gradient(V3) const {
    float yl[2][2], yr[2][2], zl[2][2], zr[2][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      uint32_t a[2], b[2];
      run2<T>(base(), vox(ic[0], ic[1] - 1, ic[2] + k), a);
      run2<T>(base(), vox(ic[0], ic[1] + 2, ic[2] + k), b);
      yl[0][k] = (float)a[0]; yl[1][k] = (float)a[1];
      yr[0][k] = (float)b[0]; yr[1][k] = (float)b[1];
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      uint32_t a[2], b[2];
      run2<T>(base(), vox(ic[0], ic[1] + j, ic[2] - 1), a);
      run2<T>(base(), vox(ic[0], ic[1] + j, ic[2] + 2), b);
      zl[0][j] = (float)a[0]; zl[1][j] = (float)a[1];
      zr[0][j] = (float)b[0]; zr[1][j] = (float)b[1];
    }
    const float fxm = tri<T>(xl[0][0], c[0][0][0], xl[1][0], c[0][1][0], xl[0][1], c[0][0][1], xl[1][1], c[0][1][1],
                             fc[0], fc[1], fc[2]);
    const float fxp = tri<T>(c[1][0][0], xr[0][0], c[1][1][0], xr[1][0], c[1][0][1], xr[0][1], c[1][1][1], xr[1][1],
                             fc[0], fc[1], fc[2]);
    const float fym = tri<T>(yl[0][0], yl[1][0], c[0][0][0], c[1][0][0], yl[0][1], yl[1][1], c[0][0][1], c[1][0][1],
                             fc[0], fc[1], fc[2]);
    const float fyp = tri<T>(c[0][1][0], c[1][1][0], yr[0][0], yr[1][0], c[0][1][1], c[1][1][1], yr[0][1], yr[1][1],
                             fc[0], fc[1], fc[2]);
    const float fzm = tri<T>(zl[0][0], zl[1][0], zl[0][1], zl[1][1], c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0],
                             fc[0], fc[1], fc[2]);
    const float fzp = tri<T>(c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1], zr[0][0], zr[1][0], zr[0][1], zr[1][1],
                             fc[0], fc[1], fc[2]);
    return mk(fxm - fxp, fym - fyp, fzm - fzp) / 2.0f;
  }
};

// normalize() with the hardware reciprocal square root (about 1 ulp; used
// for shading directions only, never for positions, step counts or TF indices)
__device__ __forceinline__ V3 normalize_rsq(V3 a) { return a * __builtin_amdgcn_rsqf(dot(a, a)); }

// lighting.glsl:33-50 (eye at the origin).  view and n are unit vectors (or
// n = 0), so reflect(view, n) is one up to rounding and its normalize() is
// left out: the specular term moves by about 1e-6 relative.
__device__ __forceinline__ V3 lighting(V3 pos, V3 n, V3 amb, V3 dif, V3 spc, V3 L) {
  V3 view = normalize_rsq(mk(0, 0, 0) - pos);
  V3 R = reflect(view, n);
  float dd = fmaxf(fabsf(dot(n, L)), 0.0f);
  float ss = pow8(fmaxf(dot(R, L), 0.0f));
  V3 c = amb + dif * dd + spc * ss;
  return mk(clamp01(c.x), clamp01(c.y), clamp01(c.z));
}

// Compose-FS.glsl Lighting()
__device__ __forceinline__ V3 iso_lighting(V3 pos, V3 n, V3 amb, V3 dif, V3 spc, V3 L) {
  n.z = fabsf(n.z);
  V3 view = normalize_rsq(mk(0, 0, 0) - pos);
  V3 R = normalize_rsq(reflect(view, n));
  float dd = fmaxf(fabsf(dot(n, L * -1.0f)), 0.0f);
  float ss = pow8(fmaxf(dot(R, L), 0.0f));
  V3 c = amb + dif * dd + spc * ss;
  return mk(clamp01(c.x), clamp01(c.y), clamp01(c.z));
}

__device__ __forceinline__ V3 xform_point(V3 p, const float* M) {
  // (p,1) * M, rows of 4
  return mk(p.x * M[0] + p.y * M[4] + p.z * M[8] + 1.0f * M[12],
            p.x * M[1] + p.y * M[5] + p.z * M[9] + 1.0f * M[13],
            p.x * M[2] + p.y * M[6] + p.z * M[10] + 1.0f * M[14]);
}
__device__ __forceinline__ V3 xform_dir(V3 p, const float* M) {
  return mk(p.x * M[0] + p.y * M[4] + p.z * M[8] + 0.0f * M[12],
            p.x * M[1] + p.y * M[5] + p.z * M[9] + 0.0f * M[13],
            p.x * M[2] + p.y * M[6] + p.z * M[10] + 0.0f * M[14]);
}
__device__ __forceinline__ V3 xform_e2t(V3 p, const float* E) {
  // E holds rows 0..3 of the eye->texture matrix, 3 columns each
  return mk(p.x * E[0] + p.y * E[3] + p.z * E[6] + 1.0f * E[9],
            p.x * E[1] + p.y * E[4] + p.z * E[7] + 1.0f * E[10],
            p.x * E[2] + p.y * E[5] + p.z * E[8] + 1.0f * E[11]);
}

__device__ __forceinline__ V3 normal_eye(V3 g, const float* nm, V3 ds) {
  V3 s = g * ds;
  V3 n = mk(nm[0] * s.x + nm[1] * s.y + nm[2] * s.z, nm[3] * s.x + nm[4] * s.y + nm[5] * s.z,
            nm[6] * s.x + nm[7] * s.y + nm[8] * s.z);
  const float l2 = dot(n, n);
  if (l2 > 0.0f) n = n * __builtin_amdgcn_rsqf(l2);
  return n;
}

// One wave renders one tile of tw x th pixels (tw * th == 64, lane = x + tw * y)
// with its top-left pixel at (tx0, ty0).  tf1 is the 1D TF as float4 (in LDS
// or global memory), list the view's brick list, touched (COUNT only) the
// view's touched-brick row.
// LEARN: 0 fixed step; 1 step proposal by the network (LearnConst); 2 fixed step
// that also records segment samples (SegOut) with the segment's RGBA at step
// multipliers 1, 2, 4 and 8.  sn_lds: the wave's network scratch (LEARN > 0).
// cols: the wave's 64-word LDS copy of the 2D TF's non-zero-alpha column mask
// (MODE 1).
template <int MODE, bool LIT, typename T, bool COUNT, int LEARN>
__device__ __forceinline__ void march_tile(const ShadeConst& S, const float4* __restrict__ tf1, const ViewConst* VC,
                                           const FrameBrick* __restrict__ list, float* __restrict__ out, int tx0,
                                           int ty0, int tw, int th, unsigned long long* counter, uint32_t* touched,
                                           const LearnConst* LC, const SegOut* SO, uint16_t* sn_lds,
                                           uint32_t view_id, uint64_t* cols, bool acc = false,
                                           float* __restrict__ depth_buf = nullptr,
                                           float4* __restrict__ hitpos_buf = nullptr) {
  const int lane = threadIdx.x & 63;
  const int px = tx0 + lane % tw, py = ty0 + lane / tw;
  const int W = (int)VC->W, H = (int)VC->H;
  const bool inside = px < W && py < H;
  const int rx0 = tx0, ry0 = ty0, rx1 = tx0 + tw - 1, ry1 = ty0 + th - 1;

  // eye-space ray through the pixel centre (eye at the origin)
  const float* ip = VC->inv_proj;
  const float n0 = ((float)px + 0.5f) * 2.0f / (float)W - 1.0f;
  const float n1 = ((float)py + 0.5f) * 2.0f / (float)H - 1.0f;
  float e[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) e[c] = n0 * ip[c] + n1 * ip[4 + c] + -1.0f * ip[8 + c] + 1.0f * ip[12 + c];
  const V3 d = mk(e[0] / e[3], e[1] / e[3], e[2] / e[3]);
  const V3 o_m = xform_point(mk(0, 0, 0), VC->inv_mv);
  const V3 d_m = xform_dir(d, VC->inv_mv);

  const V3 amb = mk(S.amb[0], S.amb[1], S.amb[2]), dif = mk(S.dif[0], S.dif[1], S.dif[2]);
  const V3 spc = mk(S.spc[0], S.spc[1], S.spc[2]), ldir = mk(S.ldir[0], S.ldir[1], S.ldir[2]);
  const V3 dsc = mk(S.dscale[0], S.dscale[1], S.dscale[2]);
  // the view's normal matrix in scalar registers: read once, not per sample
  float nm[9];
#pragma unroll
  for (int i = 0; i < 9; ++i)
    nm[i] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, VC->nm[i])));
  if constexpr (MODE == 1) {
    cols[lane] = S.tf2_cols[lane];
    sn_wave_sync();
  }
  // the view's model-space clip plane (uniform: scalar registers)
  const bool clip_on = VC->clip_on != 0u;
  float clip[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    clip[i] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, VC->clip[i])));
  const float eps = S.ert_eps;
  const bool edge = S.clamp_edge != 0;

  // acc: this launch continues the frame of an earlier launch over the bricks
  // before these in compositing order (a budgeted pool renders a frame in
  // parts): the ray state is the pixel's value so far (and, for the
  // isosurface, its depth), so the parts compose to the one-launch frame
  const size_t pix = (size_t)VC->out_offset + (size_t)py * (size_t)W + (size_t)px;
  float dst0 = 0.f, dst1 = 0.f, dst2 = 0.f, dst3 = 0.f;
  if (MODE != 2 && acc && inside) {
    const float4 p0 = reinterpret_cast<const float4*>(out)[pix];
    dst0 = p0.x; dst1 = p0.y; dst2 = p0.z; dst3 = p0.w;
  }
  bool done = !inside;
  uint32_t ns = 0;
  // isosurface state
  float best_depth = (MODE == 2 && acc && inside) ? depth_buf[pix] : 1.0f, hit_param = 0.0f;
  bool have_hit = false;
  V3 hit_pos = mk(0, 0, 0), hit_tex = mk(0, 0, 0);
  int hit_brick = 0;
  // ray parameter of this launch's best hit: with early termination on, a brick
  // whose segment starts beyond it (with a margin far above the fp16 rounding
  // of the entry point) cannot give a nearer hit -- every hit of a segment lies
  // on it, from its entry on -- and is skipped: the same frame, fewer samples
  float best_t = INFINITY;

  const int nlist = (int)VC->nlist;
  for (int base = 0; base < nlist; base += 64) {
    const int j = base + lane;
    bool cand = false;
    if (j < nlist) {
      const int16_t* r = list[j].rect;
      cand = !(r[2] < rx0 || r[0] > rx1 || r[3] < ry0 || r[1] > ry1);
    }
    uint64_t mask = __ballot(cand);
    while (mask) {
      const int k = __ffsll((unsigned long long)mask) - 1;
      mask &= mask - 1;
      const FrameBrick* fb = &list[__builtin_amdgcn_readfirstlane(base + k)];
      if (MODE != 2 && eps > 0.0f && (1.0f - dst3) < eps) done = true;
      if (LEARN == 0 && done) continue;
      // slab test in model space (t along the eye-space direction d)
      float tn = -INFINITY, tf = INFINITY;
      bool ok = true;
      {
        const float bmn[3] = {fb->bmin[0], fb->bmin[1], fb->bmin[2]};
        const float bmx[3] = {fb->bmax[0], fb->bmax[1], fb->bmax[2]};
        const float om[3] = {o_m.x, o_m.y, o_m.z}, dm[3] = {d_m.x, d_m.y, d_m.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if (dm[a] == 0.0f) {
            if (om[a] < bmn[a] || om[a] > bmx[a]) ok = false;
            continue;
          }
          float inv = 1.0f / dm[a];
          float t0 = (bmn[a] - om[a]) * inv, t1 = (bmx[a] - om[a]) * inv;
          tn = fmaxf(tn, fminf(t0, t1));
          tf = fminf(tf, fmaxf(t0, t1));
        }
      }
      if (clip_on) {
        // the brick box cut by the clip plane and closed by its cap (Clipper::BoxPlane,
        // Clipper.cpp:106-131): the ray keeps the part with n . p + d <= 0
        const float a = clip[0] * d_m.x + clip[1] * d_m.y + clip[2] * d_m.z;
        const float c = clip[0] * o_m.x + clip[1] * o_m.y + clip[2] * o_m.z + clip[3];
        if (a > 0.0f) tf = fminf(tf, -c / a);
        else if (a < 0.0f) tn = fmaxf(tn, -c / a);
        else if (c > 0.0f) ok = false;
      }
      ok = ok && tf > tn && tf > 0.0f;
      const float te = fmaxf(tn, 1.0f);  // near-plane entry when the front face is clipped
      const bool active = !done && ok && tf > te && !(MODE == 2 && eps > 0.0f && te > best_t * (1.0f + 0x1p-8f));
      if (COUNT && touched) {
        // every ray that reaches a brick takes at least one sample there (the
        // record is optional: the allocation-free render counts samples only)
        if (__ballot(active) != 0 && lane == __builtin_amdgcn_readfirstlane(lane)) touched[fb->toc] = 1u;
      }
      if (LEARN == 0 && !active) continue;
      V3 entry = d * (active ? te : 1.0f);
      const V3 exitp = d * (active ? tf : 2.0f);
      entry = mk(fp16_round(entry.x), fp16_round(entry.y), fp16_round(entry.z));
      const float* E = fb->e2t;
      const V3 entryTex = xform_e2t(entry, E), exitTex = xform_e2t(exitp, E);
      const V3 dir = exitp - entry;
      const float L = length(dir);
      const float rstep = fb->ray_step;
      const int steps = (int)(L / rstep) + 1;
      const V3 incTex = (exitTex - entryTex) / (L / rstep);
      const T* __restrict__ bd = reinterpret_cast<const T*>(S.pool) + fb->offset;
      const int nx = (int)fb->n[0], ny = (int)fb->n[1], nz = (int)fb->n[2];
      const V3 vstep = mk(fb->vstep[0], fb->vstep[1], fb->vstep[2]);
      const float sscale = fb->step_scale;
      // The fixed-step indices whose 4x4x4 neighbourhood lies inside the brick,
      // decided once per segment instead of per sample.  Sample i sits at
      // u_a = (entryTex_a + i * incTex_a) * n_a - 0.5 up to the drift of the
      // accumulated position: each addition of a coordinate below 1 rounds by at
      // most 2^-24, so after i < steps additions the drift is below steps * 2^-24
      // in texture space, steps * n_a * 2^-24 voxels; the margin is twice that
      // plus 1/16 voxel.  The interior form needs 1 <= floor(u_a) and
      // floor(u_a) + 2 < n_a on every axis.  The range only selects between two
      // forms that give the same values, so a conservative one is enough.  A
      // learned march samples a subset of these indices, after no more than
      // steps additions either.
      int in_lo = 0, in_hi = -1;
      if (MODE != 2) {
        float lo = 0.0f, hi = (float)(steps - 1);
        const float ec[3] = {entryTex.x, entryTex.y, entryTex.z}, ic3[3] = {incTex.x, incTex.y, incTex.z};
        const int nn3[3] = {nx, ny, nz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float na = (float)nn3[a];
          const float u0 = ec[a] * na - 0.5f, du = ic3[a] * na;
          const float margin = 0.0625f + (float)steps * na * 0x1p-23f;
          const float ulo = 1.0f + margin, uhi = na - 2.0f - margin;   // floor(u) in [1, n - 3]
          if (!(ulo <= uhi) || !(fabsf(u0) < 0x1p24f) || !(fabsf(du) < 0x1p24f)) {
            // no interior on this axis, or a position or increment that is not
            // finite (a zero-length segment divides by zero): the general form
            hi = -1.0f;
          } else if (du == 0.0f) {
            if (!(u0 >= ulo && u0 <= uhi)) hi = -1.0f;
          } else {
            const float t1 = (ulo - u0) / du, t2 = (uhi - u0) / du;
            lo = fmaxf(lo, fminf(t1, t2));
            hi = fminf(hi, fmaxf(t1, t2));
          }
        }
        if (hi >= lo) { in_lo = (int)ceilf(lo); in_hi = (int)floorf(hi); }
      }

      if (MODE == 2) {
        // GLRaycaster-ISO-FS.glsl:56-105 + RefineIsosurface.glsl:37-52
        // ISO_BATCH samples per iteration, their loads in flight together; the
        // samples, their order and the count are the one-by-one search's (values
        // past the segment's last step are formed and not used)
        V3 pos = entryTex;
        bool found = false;
        for (int i = 0; i < steps && !found; i += ISO_BATCH) {
          float v[ISO_BATCH];
          sample_brick_iso_n<T>(bd, nx, ny, nz, edge, pos, incTex, v);
#pragma unroll
          for (int s = 0; s < ISO_BATCH; ++s) {
            if (i + s >= steps) break;
            ++ns;
            if (v[s] >= S.iso_n) { found = true; break; }
            pos = pos + incTex;
          }
        }
        if (!found) continue;
        V3 rd = incTex / 2.0f;
        pos = pos - rd;
        for (int kk = 0; kk < S.refine_steps; ++kk) {
          rd = rd / 2.0f;
          if (sample_brick_iso<T>(bd, nx, ny, nz, edge, pos) >= S.iso_n) pos = pos - rd; else pos = pos + rd;
        }
        const float param = length(pos - entryTex) / length(exitTex - entryTex);
        const V3 he = entry * (1.0f - param) + exitp * param;
        const float zf = VC->zfar, zn = VC->znear;
        const float pxp = zf / (zf - zn), pyp = (zf * zn / (zn - zf));
        float depth = pxp + (pyp / -he.z);
        depth = fminf(fmaxf(depth, 0.0f), 1.0f);
        if (depth < best_depth) {
          best_depth = depth;
          have_hit = true;
          best_t = te + param * (tf - te);
          hit_pos = he;
          hit_param = param;
          hit_tex = pos;   // the normal is taken here, after the brick loop
          hit_brick = base + k;
        }
        continue;
      }

      // March nsamp samples from (pos, cur), the fixed-step index i0 of the
      // segment, with the step multiplied by 2^k (k = 0: GLRaycaster's step;
      // the opacity-correction exponent scales with it), under-compositing into
      // (c0..c3).  Returns false once the segment ends: 'a >= 0.99', or early
      // termination of the ray whose transmittance before the brick is tray.
      const V3 rayInc = dir / L * rstep;
      auto run = [&](int k, int nsamp, int i0, V3& pos, V3& cur, float& c0, float& c1, float& c2, float& c3,
                     float tray) -> bool {
        const float mlt = (float)(1 << k);
        const float ssc = k == 0 ? sscale : sscale * mlt;
        V3 pos_s = pos;   // the position shade() samples
        // 2D TF, gradient and shading of the sample at pos_s (the cross sampler's two forms)
        auto shade = [&](auto fast_c, float& s0, float& s1, float& s2, float& s3) {
          CrossSampler<T, decltype(fast_c)::value> cs;
          const float val = cs.value(bd, nx, ny, nz, edge, pos_s);
          const float sx = fminf(fmaxf(val * S.tf_scale, 0.0f), 1.0f);
          const int xi = min((int)floorf(sx * (float)S.tf_w), S.tf_w - 1);
          if (((cols[xi >> 6] >> (xi & 63)) & 1ull) == 0) {
            // every row of this value column is transparent: alpha 0 whatever the gradient
            s0 = s1 = s2 = s3 = 0.0f;
          } else {
            const V3 gg = cs.gradient(vstep);
            const float gm = sqrt_rn(dot(gg, gg));   // length(gg), bit for bit
            const float sy = fminf(fmaxf(1.0f - gm * S.grad_scale, 0.0f), 1.0f);
            const int yi = min((int)floorf(sy * (float)S.tf_h), S.tf_h - 1);
            const float4 tv = S.tf2f[(size_t)yi * (size_t)S.tf_w + (size_t)xi];   // byte/255, precomputed
            s0 = tv.x; s1 = tv.y; s2 = tv.z; s3 = tv.w;
            if (LIT) {
              const V3 nn = normal_eye(gg, nm, dsc);
              const V3 lc = lighting(cur, nn, amb, dif * mk(s0, s1, s2), spc, ldir);
              const float tt = 1.0f - s3;
              s3 = 1.0f - opacity_pow(tt, ssc);
              s0 = clamp01(lc.x); s1 = clamp01(lc.y); s2 = clamp01(lc.z); s3 = clamp01(s3);
            } else {
              const float tt = 1.0f - s3;
              s3 = 1.0f - opacity_pow(tt, ssc);
            }
          }
        };
        // one sample at pos (the interior form when fast_c is true): shade,
        // UnderCompositing (Compositing.glsl:33-38), advance; false once the
        // segment ends ('a >= 0.99' or early ray termination)
        // pre: the 1D sample's value, formed by the caller from the 2x2x2 taps
        // (sample_brick_iso_n); the 2D modes sample here
        auto step = [&](auto fast_c, float pre = 0.0f) -> bool {
          ++ns;
          float s0, s1, s2, s3;
          if constexpr (MODE == 0) {
            CrossSampler<T, decltype(fast_c)::value> cs;
            const float val = pre;
            const float cc = fminf(fmaxf(val * S.tf_scale, 0.0f), 1.0f);
            const int ti = min((int)floorf(cc * (float)S.tf_n), S.tf_n - 1);
            const float4 tv = tf1[ti];
            s0 = tv.x; s1 = tv.y; s2 = tv.z;
            if (LIT && tv.w != 0.0f) {
              // a transparent sample composites to nothing: only the others
              // take the gradient (the batched march loads their cross here)
              (void)cs.value(bd, nx, ny, nz, edge, pos);
              const V3 gg = cs.gradient(vstep);
              const V3 nn = normal_eye(gg, nm, dsc);
              const V3 lc = lighting(cur, nn, amb, dif * mk(s0, s1, s2), spc, ldir);
              s0 = lc.x; s1 = lc.y; s2 = lc.z;
            }
            s3 = 1.0f - opacity_pow(1.0f - tv.w, ssc);
          } else {
            pos_s = pos;
            shade(fast_c, s0, s1, s2, s3);
          }
          const float wgt = (1.0f - c3) * s3;
          c0 += s0 * wgt;
          c1 += s1 * wgt;
          c2 += s2 * wgt;
          c3 += s3 * (1.0f - c3);
          // x * 2^k is exact, so the fused form equals pos + (incTex * 2^k)
          // and adds no registers for the scaled increments
          if (LIT) cur = mk(fmaf(rayInc.x, mlt, cur.x), fmaf(rayInc.y, mlt, cur.y), fmaf(rayInc.z, mlt, cur.z));
          if (c3 >= 0.99f) return false;
          if (eps > 0.0f && tray * (1.0f - c3) < eps) return false;
          pos = mk(fmaf(incTex.x, mlt, pos.x), fmaf(incTex.y, mlt, pos.y), fmaf(incTex.z, mlt, pos.z));
          return true;
        };
        if constexpr (MODE != 0 || LIT) {
          // Each lane marches its samples before the segment's interior range
          // in the general form, the interior range in the interior form, and
          // the rest in the general form again, so that only the few samples
          // near a brick face pay for clamped taps whatever step the other
          // lanes of the wave are at.  Sample j of this run has the fixed-step
          // index i0 + j 2^k: the run's interior samples are those j with
          // in_lo <= i0 + j 2^k <= in_hi, one range as the interior indices are
          // one.  Every lane executing the interior loop is inside its range.
          const int j_lo = max(in_lo - i0 > 0 ? (in_lo - i0 + (1 << k) - 1) >> k : 0, 0);
          const int j_hi = min(in_hi - i0 >= 0 ? (in_hi - i0) >> k : -1, nsamp - 1);
          const int h_end = j_hi >= j_lo ? j_lo : nsamp;
          int i = 0;
          if constexpr (MODE == 0) {
            // lit 1D: each phase takes ISO_BATCH samples' values at once (as
            // below); a batch that runs past the phase's end discards the rest.
            // One step body (not unrolled) takes the values in turn, shifted
            // down, so that the shading's registers are not held four times.
            // The chunked march takes two at a time: four cost it 8-12 spilled
            // VGPRs (two: 6-10) against 1-4 sampling one at a time.
            const V3 inc = mk(incTex.x * mlt, incTex.y * mlt, incTex.z * mlt);
            constexpr int NB = LEARN == 0 ? ISO_BATCH : 2;
            auto phase = [&](auto fast_c, int end) -> bool {
              while (i < end) {
                float v[NB];
                sample_brick_iso_n<T, NB>(bd, nx, ny, nz, edge, pos, inc, v);
                const int e = min(end, i + NB);
#pragma nounroll
                for (; i < e; ++i) {
                  if (!step(fast_c, v[0])) return false;
#pragma unroll
                  for (int s = 0; s + 1 < NB; ++s) v[s] = v[s + 1];
                }
              }
              return true;
            };
            return phase(std::false_type{}, h_end) && phase(std::true_type{}, j_hi + 1) &&
                   phase(std::false_type{}, nsamp);
          }
          for (; i < h_end; ++i)
            if (!step(std::false_type{})) return false;
          for (; i <= j_hi; ++i)
            if (!step(std::true_type{})) return false;
          for (; i < nsamp; ++i)
            if (!step(std::false_type{})) return false;
        } else {
          // ISO_BATCH samples' loads in flight together (positions accumulated
          // one increment at a time, as step() advances them: x * 2^k is exact),
          // then composited in order; values past a termination are not used.
          // Eight at a time measured slower on the 48-request batch (69.7
          // against 61.7 ms: 167 against 110 VGPRs, the same 3 waves).
          const V3 inc = mk(incTex.x * mlt, incTex.y * mlt, incTex.z * mlt);
          for (int i = 0; i < nsamp; i += ISO_BATCH) {
            float v[ISO_BATCH];
            sample_brick_iso_n<T>(bd, nx, ny, nz, edge, pos, inc, v);
#pragma unroll
            for (int s = 0; s < ISO_BATCH; ++s) {
              if (i + s >= nsamp) break;
              if (!step(std::false_type{}, v[s])) return false;
            }
          }
        }
        return true;
      };

      float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f;
      if constexpr (LEARN == 0) {
        V3 pos = entryTex, cur = entry;
        run(0, steps, 0, pos, cur, c0, c1, c2, c3, 1.0f - dst3);
      } else {
        // Learned step: the segment is marched in chunks of SN_CHUNK fixed
        // steps.  Per chunk every lane bounds its values with the brick's cell
        // grid (probes every SN_PROBE steps), forms the features, and the
        // network proposes a multiplier and a skip flag; the wave marches the
        // chunk with the smallest multiplier any of its marching lanes asks
        // for.  A chunk whose contribution bound (transmittance before it x
        // the most opacity its value range allows) is at most skip_bound is
        // skipped without the network.  LEARN 2 marches the fixed step and
        // records chunks with their RGBA at multipliers 1, 2, 4 and 8.
        const float tray = 1.0f - dst3;
        V3 pos = entryTex, cur = entry;
        bool live = active;
        V3 dv = mk(incTex.x * (float)nx, incTex.y * (float)ny, incTex.z * (float)nz);
        const float dl = length(dv);   // the fixed step in stored voxels
        dv = dl > 0.0f ? dv * (1.0f / dl) : mk(0, 0, 0);
        // A sample within 3 voxels (per axis) of a probe interpolates voxels
        // within 4 of the probe's voxel, all inside the probe cell's dilated
        // range: probes at most SN_PROBE steps apart then bound every value of
        // the chunk when SN_PROBE steps span at most 6 voxels.  Without overlap
        // (and GL_CLAMP) the border colour 0 can enter the interpolation.
        const bool strict = (float)SN_PROBE * dl <= 6.0f;
        // the probe spacing: as many steps as span just under 6 voxels (at most
        // a chunk), SN_PROBE where that bound does not hold
        const int pstep = strict ? max(SN_PROBE, (int)fminf((float)SN_CHUNK, 5.99f / dl)) : SN_PROBE;
        const bool border0 = !edge && S.overlap == 0;
        // the network inputs that stay fixed over the segment, put(index, value)
        auto seg_features = [&](auto&& put) {
          put(9, LIT ? 1.0f : 0.0f);
          put(10, MODE == 1 ? 1.0f : 0.0f);
          put(11, fabsf(dv.x));
          put(12, fabsf(dv.y));
          put(13, fabsf(dv.z));
          put(17, sscale);
          put(20, -entry.z * 0.25f);
          put(21, dl);
          put(22, fb->tfs[3]);
          put(23, 1.0f);
        };
        // This lane's X row of the network in LDS (bf16; 24..31 zero): the
        // segment's constant inputs are stored once, the others per chunk.  A
        // row only enters its own ray's outputs, so rows of lanes that do not
        // ask are left as they are.
        uint16_t* const xrow = sn_lds + lane * SN_INP;
        if constexpr (LEARN == 1) {
          seg_features([&](int q, float v) { xrow[q] = bf16_bits(v); });
#pragma unroll
          for (int q = SN_IN; q < SN_INP; q += 2) *reinterpret_cast<uint32_t*>(xrow + q) = 0u;
        }
        const int ncx = (int)cell_count((uint32_t)nx), ncy = (int)cell_count((uint32_t)ny);
        const uint32_t* __restrict__ cg = S.cells + fb->cells;
        const int tn = S.tfst_n;
        const float* __restrict__ ts = S.tfst;
        const float vmaxs = (float)Unorm<T>::max_value();
        // the lane's chunk at i0: its steps, the transmittance before it, its
        // value range from the cell grid, the TF statistics over that range
        // and the contribution bound T0 (1 - (1 - amax)^(s nf))
        int nf = 1, lo = 0, hi = 0;
        float T0 = 0.0f, vl = 0.0f, vh = 0.0f, cnt = 1.0f, amax = 0.0f, smax = 0.0f, obound = 0.0f;
        auto bound_chunk = [&](int i0) {
          nf = min(SN_CHUNK, steps - i0);
          T0 = tray * (1.0f - c3);
          uint32_t vlo = 0xffffu, vhi = 0u;
          // probes at steps 0, pstep, 2 pstep, ... and the chunk's last step
          for (int q = 0;; q += pstep) {
            const int qs = min(q, nf - 1);
            const V3 p = pos + incTex * (float)qs;
            const float pc[3] = {p.x, p.y, p.z};
            const int nn3[3] = {nx, ny, nz};
            int ci[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const float u = fminf(fmaxf(pc[a], 0.0f), 1.0f) * (float)nn3[a] - 0.5f;
              ci[a] = min(max((int)floorf(u), 0), nn3[a] - 1) >> CELL_LW;
            }
            const uint32_t w = cg[((uint32_t)ci[2] * (uint32_t)ncy + (uint32_t)ci[1]) * (uint32_t)ncx + (uint32_t)ci[0]];
            vlo = min(vlo, w & 0xffffu);
            vhi = max(vhi, w >> 16);
            if (qs == nf - 1) break;
          }
          if (border0) vlo = 0u;
          // TF entries of the range (one stored value of slack on either side)
          auto entry_of = [&](float v) {
            const float cc = fminf(fmaxf(v * Unorm<T>::inv() * S.tf_scale, 0.0f), 1.0f);
            return min((int)floorf(cc * (float)tn), tn - 1);
          };
          vl = fmaxf((float)vlo - 1.0f, 0.0f);
          vh = fminf((float)vhi + 1.0f, vmaxs);
          lo = entry_of(vl);
          hi = max(entry_of(vh), lo);
          cnt = (float)(hi - lo + 1);
          const int kk = 31 - __clz(hi - lo + 1);
          amax = fmaxf(ts[tfst_amax(tn, kk) + lo], ts[tfst_amax(tn, kk) + hi - (1 << kk) + 1]);
          smax = 0.0f;
          if (hi > lo) {
            const int k2 = 31 - __clz(hi - lo);
            const int r = tfst_smax(tn, S.tfst_levels, k2);
            smax = fmaxf(ts[r + lo], ts[r + hi - (1 << k2)]);
          }
          obound = 1.0f - powf(1.0f - fminf(amax, 1.0f), sscale * (float)nf);
        };
        // the network inputs that change per chunk, put(index, value)
        auto chunk_features = [&](auto&& put) {
          const float x0 = fminf(fmaxf(vl * Unorm<T>::inv() * S.tf_scale, 0.0f), 1.0f);
          const float x1 = fminf(fmaxf(vh * Unorm<T>::inv() * S.tf_scale, 0.0f), 1.0f);
          put(0, x0);
          put(1, x1);
          put(2, x1 - x0);
          put(3, (ts[tfst_pa(tn) + hi + 1] - ts[tfst_pa(tn) + lo]) / cnt);
          put(4, amax);
          put(5, ts[tfst_ptv(tn) + hi] - ts[tfst_ptv(tn) + lo]);
          put(6, ts[tfst_prgb(tn) + hi] - ts[tfst_prgb(tn) + lo]);
          put(7, smax);
          put(8, (ts[tfst_pnz(tn) + hi + 1] - ts[tfst_pnz(tn) + lo]) / cnt);
          put(14, (float)nf * (1.0f / (float)SN_CHUNK));
          put(15, T0);
          put(16, -log2f(fmaxf(T0, 1e-6f)) * (1.0f / 20.0f));
          put(18, obound);
          put(19, T0 * obound);
        };
        // a skipped chunk: the fixed march's positions, bit for bit
        auto skip_steps = [&]() {
          for (int i = 0; i < nf; ++i) {
            if (LIT) cur = cur + rayInc;
            pos = pos + incTex;
          }
        };
        if constexpr (LEARN == 1) {
          // Each lane first goes by the chunks it skips exactly (bound at most
          // skip_bound) on its own, so a wave marches together only chunks
          // that some of its lanes must march; then the network (or the
          // fallback, or fixed_k) decides for the lanes that stopped at such a
          // chunk, the wave takes the smallest multiplier of its marching
          // lanes, and each lane marches or skips its own chunk.
          int i0 = 0;
          for (;;) {
            bool has = live && i0 < steps;
            while (has) {
              bound_chunk(i0);
              if (!(strict && T0 * obound <= LC->skip_bound)) break;
              skip_steps();
              i0 += SN_CHUNK;
              has = i0 < steps;
            }
            if (__ballot(has) == 0) break;
            bool skip = false;
            int k = 0;
            // a lane whose bound (or variation risk) exceeds its fallback marches
            // the fixed step and so sets the wave's multiplier to 1: then the
            // network is not asked
            bool forced = has && T0 * obound > LC->fallback;
            if (LC->fallback_tv < INFINITY && has && !forced) {
              const float tva = ts[tfst_ptv(tn) + hi] - ts[tfst_ptv(tn) + lo];
              const float tvc = ts[tfst_prgb(tn) + hi] - ts[tfst_prgb(tn) + lo];
              forced = T0 * (obound * tvc + fminf(1.0f, (float)nf * sscale * tva)) > LC->fallback_tv;
            }
            const bool ask = has && !forced;
            if (LC->fixed_k >= 0) {
              k = forced ? 0 : LC->fixed_k;
            } else if (__ballot(forced) == 0 && __ballot(ask) != 0) {
              chunk_features([&](int q, float v) { xrow[q] = bf16_bits(v); });
              float y0, y1;
              sn_eval_wave_x(LC->net, sn_lds, y0, y1);
              if (ask) {
                k = max(0, min(LC->max_log2, (int)floorf(3.0f * y0 + 0.5f - LC->margin)));
                skip = y1 > LC->skip_logit;
              }
            }
            // one multiplier for the wave: the smallest one a marching lane asks for
            const bool marching = has && !skip;
            int kw = LC->max_log2;
            for (int q = LC->max_log2 - 1; q >= 0; --q)
              if (__ballot(marching && k <= q) != 0) kw = q;
            if (has) {
              if (skip) skip_steps();
              else live = run(kw, (nf + (1 << kw) - 1) >> kw, i0, pos, cur, c0, c1, c2, c3, tray);
              i0 += SN_CHUNK;
            }
          }
        } else {
          // LEARN 2: the fixed march, chunk by chunk across the wave, recording
          // hash-selected chunks with their RGBA at multipliers 1, 2, 4, 8
          for (int i0 = 0;; i0 += SN_CHUNK) {
            const bool has = live && i0 < steps;
            if (__ballot(has) == 0) break;
            if (has) bound_chunk(i0);
            const bool skip = has && strict && T0 * obound <= LC->skip_bound;
            if (has && !skip) {
              uint32_t h = (uint32_t)(px * 73856093) ^ (uint32_t)(py * 19349663) ^ (fb->toc * 83492791u) ^
                           (view_id * 2654435761u) ^ ((uint32_t)i0 * 2246822519u) ^ SO->seed;
              h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
              if (h < SO->keep) {
                const unsigned int slot = atomicAdd(SO->count, 1u);
                if (slot < SO->cap) {
                  SegSample* o = SO->out + slot;
                  seg_features([&](int q, float v) { o->f[q] = v; });
                  chunk_features([&](int q, float v) { o->f[q] = v; });
                  o->t_before = T0;
                  o->steps = (float)nf;
#pragma unroll 1
                  for (int mi = 0; mi < 4; ++mi) {
                    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
                    V3 pp = pos, cc = cur;
                    run(mi, (nf + (1 << mi) - 1) >> mi, i0, pp, cc, a0, a1, a2, a3, 1.0f);
                    o->c[mi][0] = a0; o->c[mi][1] = a1; o->c[mi][2] = a2; o->c[mi][3] = a3;
                  }
                }
              }
            }
            if (has) {
              if (skip) skip_steps();
              else live = run(0, nf, i0, pos, cur, c0, c1, c2, c3, tray);
            }
          }
        }
        if (!active) continue;
      }
      // blend (ONE_MINUS_DST_ALPHA, ONE) into the RGBA32F target
      const float t = 1.0f - dst3;
      dst0 = c0 * t + dst0;
      dst1 = c1 * t + dst1;
      dst2 = c2 * t + dst2;
      dst3 = c3 * t + dst3;
    }
    if (__ballot(!done) == 0) break;  // every ray of the tile has terminated
  }

  if (COUNT) {
    uint32_t s = ns;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0 && s) atomicAdd(counter, (unsigned long long)s);
  }
  if (!inside) return;
  float4 res;
  if (MODE == 2 && acc && !have_hit) return;   // an earlier part's hit (or none) stands
  if (MODE == 2) {
    // Normal of the winning hit: per-lane brick, computed once all bricks are
    // done (inside the brick loop the per-brick values of lanes that take
    // different branches were kept in shared scalar state by the compiler).
    V3 hit_n = mk(0, 0, 0);
    if (have_hit) {
      const FrameBrick* hb = &list[hit_brick];
      const T* __restrict__ hd = reinterpret_cast<const T*>(S.pool) + hb->offset;
      hit_n = normal_eye(gradient<T>(hd, (int)hb->n[0], (int)hb->n[1], (int)hb->n[2], edge, hit_tex,
                                     mk(hb->vstep[0], hb->vstep[1], hb->vstep[2])),
                         nm, dsc);
    }
    if (have_hit && hit_param != 0.0f) {
      const V3 isoc = mk(S.iso_col[0], S.iso_col[1], S.iso_col[2]);
      const V3 c = iso_lighting(hit_pos, hit_n, amb, dif * isoc, spc, ldir);
      res = make_float4(c.x, c.y, c.z, 1.0f);
    } else {
      res = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  } else {
    res = make_float4(dst0, dst1, dst2, dst3);
  }
  reinterpret_cast<float4*>(out)[pix] = res;
  if (MODE == 2 && depth_buf) depth_buf[pix] = best_depth;
  // the iso hit target of GLRaycaster-ISO-FS.glsl:95-96 (eye-space hit, segment
  // parameter; cleared to 0), read back by GLRenderer::Pick
  if (MODE == 2 && hitpos_buf)
    hitpos_buf[pix] = have_hit ? make_float4(hit_pos.x, hit_pos.y, hit_pos.z, hit_param) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Multi-view launch: one 256-thread workgroup = 16x16 pixels = 4 waves of 8x8.
// Without work queues each workgroup renders one 16x16 group: consecutive
// workgroups go round-robin over the 8 XCDs, and each XCD runs a contiguous
// range of groups, in the host's order (heaviest first) when there is one.
// With work queues (A.queues, cleared to 0 before the launch) the grid is
// persistent: every wave pulls 8x8 tiles -- item t of XCD x's queue is
// quarter t & 3 of group order[xcd_start[x] + (t >> 2)] (or that index) --
// first from its own XCD's queue, then from the others in turn, and exits when
// all 8 are drained (every wave reaches that exit).  A wave that draws a light
// tile comes back for another at once instead of waiting for its workgroup.
template <int MODE, bool LIT, typename T, bool COUNT, int LEARN>
__device__ __forceinline__ void march_group(const MarchArgs& A) {
  extern __shared__ float4 s_tf[];  // 1D TF (tf_n entries, MODE 0 only)
  __shared__ __align__(16) uint16_t s_sn[LEARN > 0 ? 4 * SN_WAVE_LDS : 8];   // step network scratch, one slice per wave
  __shared__ uint64_t s_cols[MODE == 1 ? 4 * 64 : 1];                          // 2D TF column masks, one slice per wave

  const uint32_t lin = blockIdx.x, xcd = lin & 7u, j = lin >> 3;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // renders quarter q (x = q & 1, y = q >> 1) of 16x16 group g with this wave
  auto tile = [&](uint32_t g, int q) {
    int v = 0;
    while (v + 1 < A.nviews && g >= A.views[v + 1].group_offset) ++v;
    const ViewConst* VC = &A.views[v];
    const uint32_t gl = g - VC->group_offset;
    const uint32_t gx = gl % VC->groups_x, gy = gl / VC->groups_x;
    const int tx = (int)gx * 2 + (q & 1), ty = (int)gy * 2 + (q >> 1);
    march_tile<MODE, LIT, T, COUNT, LEARN>(A.S, s_tf, VC, A.list + VC->list_offset, A.out, tx * 8, ty * 8, 8, 8,
                                           A.counter, (COUNT && A.touched) ? A.touched + (size_t)v * A.ntoc : nullptr,
                                           &A.learn, &A.seg, s_sn + (LEARN > 0 ? wave * SN_WAVE_LDS : 0), (uint32_t)v,
                                           s_cols + (MODE == 1 ? wave * 64 : 0), A.accumulate != 0, A.depth, A.hitpos);
  };

  // without queues: this workgroup's group (those past the range leave
  // before the barrier, the whole workgroup at once)
  uint32_t g0 = 0;
  if (A.queues == nullptr) {
    if (A.order) {
      if (j >= A.xcd_count[xcd]) return;
      g0 = A.order[A.xcd_start[xcd] + j];
    } else {
      g0 = xcd * (A.groups_padded >> 3) + j;
      if (g0 >= A.groups_total) return;
    }
  }
  if (MODE == 0) {
    for (int i = threadIdx.x; i < A.S.tf_n; i += 256) s_tf[i] = A.S.tf1[i];
    __syncthreads();   // the only workgroup barrier: the waves run independently after it
  }
  uint32_t k = 0;   // queues tried so far, starting with this XCD's own
  for (int it = 0;; ++it) {
    uint32_t g;
    int q;
    if (A.queues == nullptr) {
      if (it > 0) return;
      g = g0;
      q = wave;
    } else {
      uint32_t t = 0, x = 0;
      for (; k < 8; ++k) {
        x = (xcd + k) & 7u;
        uint32_t tt = 0;
        if (lane == 0) tt = atomicAdd(&A.queues[x], 1u);
        tt = (uint32_t)__builtin_amdgcn_readfirstlane((int)tt);
        if (tt < A.xcd_count[x] * 4u) { t = tt; break; }   // else drained (the counter grows past it: harmless)
      }
      if (k == 8) return;
      const uint32_t gi = A.xcd_start[x] + (t >> 2);
      g = A.order ? (uint32_t)__builtin_amdgcn_readfirstlane((int)A.order[gi]) : gi;
      q = (int)(t & 3u);
    }
    tile(g, q);
  }
}

template <int MODE, bool LIT, typename T, bool COUNT, int LEARN>
__global__ __launch_bounds__(256) void march_kernel(MarchArgs A) {
  march_group<MODE, LIT, T, COUNT, LEARN>(A);
}

// The learned march (LEARN 1): the network evaluation raises the register
// count; a cap of TVK_LEARN_WAVES waves per SIMD trades a few spilled
// registers for occupancy.
#ifndef TVK_LEARN_WAVES
#define TVK_LEARN_WAVES 3
#endif
template <int MODE, bool LIT, typename T, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TVK_LEARN_WAVES))) void march_kernel_learn(
    MarchArgs A) {
  march_group<MODE, LIT, T, COUNT, 1>(A);
}

// Persistent batch launch: every wave pulls tiles from the per-XCD queues until
// all are drained.  Requests in one launch share MODE, LIT and T.  A cap of 3
// waves per SIMD lets the lit variants keep their registers (at 4 the lit 2D
// one spills 2 VGPRs and the learned one 41; measured 44.1 against 44.4 ms on
// the 64 GiB LOD batch).  Capping only the learned variants at 3 lets the
// fixed-step ones fit 4 waves with no spill, but that measured slower too
// (64.5 against 61.6 ms on the 48-request batch, 41.5 against 40.4 ms on the
// 64 GiB LOD batch).
template <int MODE, bool LIT, typename T, bool COUNT, int LEARN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void batch_kernel(
    BatchArgs B, const BatchRequest* __restrict__ req) {
  __shared__ __align__(16) uint16_t s_sn[LEARN > 0 ? 4 * SN_WAVE_LDS : 8];
  __shared__ uint64_t s_cols[MODE == 1 ? 4 * 64 : 1];
  const uint32_t xcd = blockIdx.x & 7u;
  const uint32_t chunk = (B.tiles_total + 7u) >> 3;
  const int lane = threadIdx.x & 63;
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t q = (xcd + k) & 7u;
    const uint32_t lo = q * chunk, hi = min(B.tiles_total, lo + chunk);
    if (lo >= hi) continue;
    for (;;) {
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(&B.queues[q], 1u);
      t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
      t += lo;
      if (t >= hi) break;   // this chunk is drained (the counter keeps growing past it: harmless)
      // request of tile t: last request whose tile_offset <= t
      int r0 = 0, r1 = B.nreq - 1;
      while (r0 < r1) {
        const int m = (r0 + r1 + 1) >> 1;
        if (req[m].tile_offset <= t) r0 = m; else r1 = m - 1;
      }
      const BatchRequest* R = &req[__builtin_amdgcn_readfirstlane(r0)];
      const uint32_t lt = t - R->tile_offset;
      const uint32_t ty = lt / R->tiles_x, tx = lt - ty * R->tiles_x;
      march_tile<MODE, LIT, T, COUNT, LEARN>(R->S, R->S.tf1, &R->V, B.list + R->V.list_offset, B.out,
                                             (int)(tx * B.tile_w), (int)(ty * B.tile_h), (int)B.tile_w,
                                             (int)B.tile_h, B.counter, (COUNT && B.touched) ? B.touched + R->touched_offset : nullptr,
                                             &B.learn, nullptr, s_sn + (LEARN > 0 ? (threadIdx.x >> 6) * SN_WAVE_LDS : 0),
                                             0u, s_cols + (MODE == 1 ? (threadIdx.x >> 6) * 64 : 0));
    }
  }
}

__global__ void fill_u32(unsigned int* p, unsigned int n, unsigned int v) {
  const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(unsigned int* p, unsigned int n, unsigned int v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32, dim3((n + 255u) / 256u), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

// host-side launchers (extern "C++" within the library)

// Persistent grid of a kernel: the workgroups the device holds at once (its
// occupancy at this LDS size times the CU count), a multiple of 8 XCDs.
static dim3 persistent_grid(const void* kernel, size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds) != hipSuccess || per_cu <= 0) per_cu = 1;
  const uint32_t n = (uint32_t)per_cu * (uint32_t)cus;
  return dim3((n + 7u) & ~7u);
}

template <typename K>
static void launch_k(K kernel, const MarchArgs& a, dim3 grid, size_t lds, hipStream_t st) {
  if (a.queues) grid = persistent_grid(reinterpret_cast<const void*>(kernel), lds);
  hipLaunchKernelGGL(kernel, grid, dim3(256), lds, st, a);
}

template <int M, bool L, typename T>
static void launch_one(const MarchArgs& a, dim3 grid, size_t lds, hipStream_t st, int learn) {
  if (learn == 2 && M != 2)
    launch_k(march_kernel<M, L, T, false, (M != 2 ? 2 : 0)>, a, grid, lds, st);
  else if (learn == 1 && M != 2) {
    if (a.counter)
      launch_k(march_kernel_learn<M, L, T, true>, a, grid, lds, st);
    else
      launch_k(march_kernel_learn<M, L, T, false>, a, grid, lds, st);
  } else if (a.counter)
    launch_k(march_kernel<M, L, T, true, 0>, a, grid, lds, st);
  else
    launch_k(march_kernel<M, L, T, false, 0>, a, grid, lds, st);
}

// learn: 0 fixed step, 1 step proposal (a.learn), 2 fixed step recording segment samples (a.seg)
hipError_t launch_march(const MarchArgs& a, int bpv, hipStream_t st, int learn) {
  dim3 grid(a.groups_padded);
  const size_t lds = a.mode == 0 ? (size_t)a.S.tf_n * sizeof(float4) : 0;
  const bool lit = a.lighting != 0 && a.mode != 2;
  const bool u16 = bpv == 2;
  if (learn != 0 && a.mode == 2) return hipErrorInvalidValue;
  switch (a.mode) {
    case 0:
      if (lit) { if (u16) launch_one<0, true, uint16_t>(a, grid, lds, st, learn); else launch_one<0, true, uint8_t>(a, grid, lds, st, learn); }
      else { if (u16) launch_one<0, false, uint16_t>(a, grid, lds, st, learn); else launch_one<0, false, uint8_t>(a, grid, lds, st, learn); }
      break;
    case 1:
      if (lit) { if (u16) launch_one<1, true, uint16_t>(a, grid, lds, st, learn); else launch_one<1, true, uint8_t>(a, grid, lds, st, learn); }
      else { if (u16) launch_one<1, false, uint16_t>(a, grid, lds, st, learn); else launch_one<1, false, uint8_t>(a, grid, lds, st, learn); }
      break;
    case 2:
      if (u16) launch_one<2, false, uint16_t>(a, grid, lds, st, 0); else launch_one<2, false, uint8_t>(a, grid, lds, st, 0);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int M, bool L, typename T>
static void launch_batch_one(const BatchArgs& b, int groups, hipStream_t st, int learn) {
  if (learn == 1 && M != 2) {
    if (b.counter)
      hipLaunchKernelGGL((batch_kernel<M, L, T, true, (M != 2 ? 1 : 0)>), dim3(groups), dim3(256), 0, st, b, b.req);
    else
      hipLaunchKernelGGL((batch_kernel<M, L, T, false, (M != 2 ? 1 : 0)>), dim3(groups), dim3(256), 0, st, b, b.req);
  } else if (b.counter)
    hipLaunchKernelGGL((batch_kernel<M, L, T, true, 0>), dim3(groups), dim3(256), 0, st, b, b.req);
  else
    hipLaunchKernelGGL((batch_kernel<M, L, T, false, 0>), dim3(groups), dim3(256), 0, st, b, b.req);
}

// occupancy of the exact instance launch_batch_one picks
template <int M, bool L, typename T>
static hipError_t occ_one(int* n, bool count, int learn) {
  if (learn == 1 && M != 2)
    return count ? hipOccupancyMaxActiveBlocksPerMultiprocessor(n, batch_kernel<M, L, T, true, (M != 2 ? 1 : 0)>, 256, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(n, batch_kernel<M, L, T, false, (M != 2 ? 1 : 0)>, 256, 0);
  return count ? hipOccupancyMaxActiveBlocksPerMultiprocessor(n, batch_kernel<M, L, T, true, 0>, 256, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(n, batch_kernel<M, L, T, false, 0>, 256, 0);
}

// workgroups of the persistent kernel that fit on one CU at once
int batch_occupancy(int mode, bool lit, int bpv, bool c, int l, int* n) {
  const bool u16 = bpv == 2;
  lit = lit && mode != 2;
  if (mode == 2) l = 0;
  if (mode == 0) return lit ? (u16 ? occ_one<0, true, uint16_t>(n, c, l) : occ_one<0, true, uint8_t>(n, c, l))
                            : (u16 ? occ_one<0, false, uint16_t>(n, c, l) : occ_one<0, false, uint8_t>(n, c, l));
  if (mode == 1) return lit ? (u16 ? occ_one<1, true, uint16_t>(n, c, l) : occ_one<1, true, uint8_t>(n, c, l))
                            : (u16 ? occ_one<1, false, uint16_t>(n, c, l) : occ_one<1, false, uint8_t>(n, c, l));
  return u16 ? occ_one<2, false, uint16_t>(n, c, 0) : occ_one<2, false, uint8_t>(n, c, 0);
}

// Clears the 8 queue counters with a kernel (graph-capturable) and launches the
// persistent kernel with `groups` workgroups.
hipError_t launch_batch(const BatchArgs& b, int mode, bool lit, int bpv, int groups, hipStream_t st, int learn) {
  hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, st, b.queues, 8u, 0u);
  if (b.tiles_total == 0 || groups <= 0) return hipGetLastError();
  const bool u16 = bpv == 2;
  lit = lit && mode != 2;
  if (learn != 0 && mode == 2) learn = 0;
  switch (mode) {
    case 0:
      if (lit) { if (u16) launch_batch_one<0, true, uint16_t>(b, groups, st, learn); else launch_batch_one<0, true, uint8_t>(b, groups, st, learn); }
      else { if (u16) launch_batch_one<0, false, uint16_t>(b, groups, st, learn); else launch_batch_one<0, false, uint8_t>(b, groups, st, learn); }
      break;
    case 1:
      if (lit) { if (u16) launch_batch_one<1, true, uint16_t>(b, groups, st, learn); else launch_batch_one<1, true, uint8_t>(b, groups, st, learn); }
      else { if (u16) launch_batch_one<1, false, uint16_t>(b, groups, st, learn); else launch_batch_one<1, false, uint8_t>(b, groups, st, learn); }
      break;
    case 2:
      if (u16) launch_batch_one<2, false, uint16_t>(b, groups, st, 0); else launch_batch_one<2, false, uint8_t>(b, groups, st, 0);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tvk
