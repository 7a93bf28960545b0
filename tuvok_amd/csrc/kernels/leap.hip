// GridLeaper-style single-pass traversal for gfx950: one ray per lane, 8x8
// tiles per wave, rays walk the brick grid continuously in normalised volume
// coordinates with a level of detail chosen per brick from the depth
// (GLGridLeaper-blend.glsl:65-228 with the volume-pool functions of
// GLVolumePool.cpp:484-656).  Empty bricks are leapt over (child-empty bricks
// at coarser levels give longer leaps); a non-resident brick is reported in
// the missing table and the next coarser resident level is used, and the ray
// records where it first left the requested resolution so that a later pass
// resumes from there once the brick is paged in.
//
// A brick is sampled in its own voxel coordinates in the blocked HBM pool (the
// reference samples the same point through a 3D texture atlas).  The float
// operations are those of the oracle in the same order (-ffp-contract=off).
#include <hip/hip_runtime.h>
#include <algorithm>
#include "../tvk_types.hpp"
#include "common.hpp"

namespace tvk {
namespace {

enum : uint32_t { BI_MISSING = 0, BI_CHILD_EMPTY = 1, BI_EMPTY = 2, BI_FLAG_COUNT = 3 };

// trilinear blend of eight scaled taps c[x][y][z] with weights (1-f, f)
__device__ __forceinline__ float leap_tri(float c000, float c100, float c010, float c110, float c001, float c101,
                                          float c011, float c111, float f0, float f1, float f2) {
  const float c00 = c000 * (1.0f - f0) + c100 * f0;
  const float c10 = c010 * (1.0f - f0) + c110 * f0;
  const float c01 = c001 * (1.0f - f0) + c101 * f0;
  const float c11 = c011 * (1.0f - f0) + c111 * f0;
  const float c0 = c00 * (1.0f - f1) + c10 * f1;
  const float c1 = c01 * (1.0f - f1) + c11 * f1;
  return c0 * (1.0f - f2) + c1 * f2;
}

// texture3D (GL_LINEAR, GL_CLAMP border 0) at voxel position b of a stored brick
template <typename T>
__device__ __forceinline__ float sample_vox(const T* __restrict__ d, int nx, int ny, int nz, V3 b) {
  const float inv = Unorm<T>::inv();
  const int n[3] = {nx, ny, nz};
  const float bc[3] = {b.x, b.y, b.z};
  int i0[3];
  float f[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float s = fminf(fmaxf(bc[a], 0.0f), (float)n[a]);
    const float u = s - 0.5f, fl = floorf(u);
    i0[a] = (int)fl;
    f[a] = u - fl;
  }
  const int xa = i0[0], xb = i0[0] + 1, ya = i0[1], yb = i0[1] + 1, za = i0[2], zb = i0[2] + 1;
  const bool vxa = xa >= 0, vxb = xb < nx, vya = ya >= 0, vyb = yb < ny, vza = za >= 0, vzb = zb < nz;
  const uint32_t nbx = nblocks((uint32_t)nx, Blk<T>::LX), nbxy = nbx * nblocks((uint32_t)ny, 2);
  const uint32_t xA = ox<T>(min(max(xa, 0), nx - 1)), xB = ox<T>(min(max(xb, 0), nx - 1));
  const uint32_t yA = oy<T>(min(max(ya, 0), ny - 1), nbx), yB = oy<T>(min(max(yb, 0), ny - 1), nbx);
  const uint32_t zA = oz<T>(min(max(za, 0), nz - 1), nbxy), zB = oz<T>(min(max(zb, 0), nz - 1), nbxy);
  auto tap = [&](bool ok, uint32_t off) { const float v = (float)d[off] * inv; return ok ? v : 0.0f; };
  const float c000 = tap(vxa && vya && vza, zA + yA + xA), c100 = tap(vxb && vya && vza, zA + yA + xB);
  const float c010 = tap(vxa && vyb && vza, zA + yB + xA), c110 = tap(vxb && vyb && vza, zA + yB + xB);
  const float c001 = tap(vxa && vya && vzb, zB + yA + xA), c101 = tap(vxb && vya && vzb, zB + yA + xB);
  const float c011 = tap(vxa && vyb && vzb, zB + yB + xA), c111 = tap(vxb && vyb && vzb, zB + yB + xB);
  return leap_tri(c000, c100, c010, c110, c001, c101, c011, c111, f[0], f[1], f[2]);
}

// sample_vox where every active lane's eight taps lie inside its brick: each
// x pair is one run2 load (four loads instead of eight gathers), the taps,
// weights and float operations are sample_vox's; otherwise sample_vox itself.
template <typename T>
__device__ __forceinline__ float sample_vox_fast(const T* __restrict__ d, int nx, int ny, int nz, V3 b) {
  const int n[3] = {nx, ny, nz};
  const float bc[3] = {b.x, b.y, b.z};
  int i0[3];
  float f[3];
  bool in = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float s = fminf(fmaxf(bc[a], 0.0f), (float)n[a]), u = s - 0.5f, fl = floorf(u);
    i0[a] = (int)fl;
    f[a] = u - fl;
    in = in && i0[a] >= 0 && i0[a] + 1 < n[a];
  }
  if (__ballot(!in) != 0) return sample_vox<T>(d, nx, ny, nz, b);
  const float inv = Unorm<T>::inv();
  const uint32_t nbx = nblocks((uint32_t)nx, Blk<T>::LX), nbxy = nbx * nblocks((uint32_t)ny, 2);
  const uint8_t* __restrict__ base = reinterpret_cast<const uint8_t*>(d);
  float c[2][2][2];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      uint32_t v[2];
      run2<T>(base, (ox<T>(i0[0]) + oy<T>(i0[1] + j, nbx) + oz<T>(i0[2] + k, nbxy)) * (uint32_t)sizeof(T), v);
      c[0][j][k] = (float)v[0] * inv;
      c[1][j][k] = (float)v[1] * inv;
    }
  return leap_tri(c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0], c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1],
                  f[0], f[1], f[2]);
}

// sample_vox at b and gradient_vox's six taps around it from one 32-voxel
// cross, bit for bit: valid where every tap is inside the brick and each
// tap's index is the centre's +-1 (coords() tests it per lane).  The four
// central x-runs give the centre cube and the x arms; the y and z arms are
// loaded only by gradient().
template <typename T>
struct LeapCross {
  int ic[3];
  float fc[3], fm[3], fq[3];
  float c[2][2][2], xl[2][2], xr[2][2];   // scaled taps, [x][y][z] / [y][z]
  const uint8_t* base;
  uint32_t nbx, nbxy;

  __device__ __forceinline__ bool coords(int nx, int ny, int nz, V3 b) {
    const int n[3] = {nx, ny, nz};
    const float bc[3] = {b.x, b.y, b.z};
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float s = fminf(fmaxf(bc[a], 0.0f), (float)n[a]), u = s - 0.5f, fl = floorf(u);
      ic[a] = (int)fl; fc[a] = u - fl;
      s = fminf(fmaxf(bc[a] + -1.0f, 0.0f), (float)n[a]); u = s - 0.5f; fl = floorf(u);
      const int im = (int)fl;
      fm[a] = u - fl;
      s = fminf(fmaxf(bc[a] + 1.0f, 0.0f), (float)n[a]); u = s - 0.5f; fl = floorf(u);
      const int iq = (int)fl;
      fq[a] = u - fl;
      ok = ok && ic[a] >= 1 && ic[a] + 2 < n[a] && im == ic[a] - 1 && iq == ic[a] + 1;
    }
    return ok;
  }
  __device__ __forceinline__ uint32_t vox(int x, int y, int z) const {
    return (ox<T>(x) + oy<T>(y, nbx) + oz<T>(z, nbxy)) * (uint32_t)sizeof(T);
  }
  __device__ __forceinline__ float value(const T* __restrict__ d, int nx, int ny) {
    base = reinterpret_cast<const uint8_t*>(d);
    nbx = nblocks((uint32_t)nx, Blk<T>::LX);
    nbxy = nbx * nblocks((uint32_t)ny, 2);
    const float inv = Unorm<T>::inv();
    const int xs = ic[0] - 1;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        uint32_t v[4];
        run4<T>(base, vox(xs & ~((1 << Blk<T>::LX) - 1), ic[1] + j, ic[2] + k), xs & ((1 << Blk<T>::LX) - 1), v);
        xl[j][k] = (float)v[0] * inv;
        c[0][j][k] = (float)v[1] * inv;
        c[1][j][k] = (float)v[2] * inv;
        xr[j][k] = (float)v[3] * inv;
      }
    return leap_tri(c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0], c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1],
                    fc[0], fc[1], fc[2]);
  }
  __device__ __forceinline__ V3 gradient() const {
    const float inv = Unorm<T>::inv();
    float yl[2][2], yr[2][2], zl[2][2], zr[2][2];   // [x][z] / [x][y]
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      uint32_t a[2], b[2];
      run2<T>(base, vox(ic[0], ic[1] - 1, ic[2] + k), a);
      run2<T>(base, vox(ic[0], ic[1] + 2, ic[2] + k), b);
      yl[0][k] = (float)a[0] * inv; yl[1][k] = (float)a[1] * inv;
      yr[0][k] = (float)b[0] * inv; yr[1][k] = (float)b[1] * inv;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      uint32_t a[2], b[2];
      run2<T>(base, vox(ic[0], ic[1] + j, ic[2] - 1), a);
      run2<T>(base, vox(ic[0], ic[1] + j, ic[2] + 2), b);
      zl[0][j] = (float)a[0] * inv; zl[1][j] = (float)a[1] * inv;
      zr[0][j] = (float)b[0] * inv; zr[1][j] = (float)b[1] * inv;
    }
    // gradient_vox's taps: xp = +x, xm = -x, yp = -y, ym = +y, zp = +z, zm = -z
    const float xp = leap_tri(c[1][0][0], xr[0][0], c[1][1][0], xr[1][0], c[1][0][1], xr[0][1], c[1][1][1], xr[1][1],
                              fq[0], fc[1], fc[2]);
    const float xm = leap_tri(xl[0][0], c[0][0][0], xl[1][0], c[0][1][0], xl[0][1], c[0][0][1], xl[1][1], c[0][1][1],
                              fm[0], fc[1], fc[2]);
    const float yp = leap_tri(yl[0][0], yl[1][0], c[0][0][0], c[1][0][0], yl[0][1], yl[1][1], c[0][0][1], c[1][0][1],
                              fc[0], fm[1], fc[2]);
    const float ym = leap_tri(c[0][1][0], c[1][1][0], yr[0][0], yr[1][0], c[0][1][1], c[1][1][1], yr[0][1], yr[1][1],
                              fc[0], fq[1], fc[2]);
    const float zp = leap_tri(c[0][0][1], c[1][0][1], c[0][1][1], c[1][1][1], zr[0][0], zr[1][0], zr[0][1], zr[1][1],
                              fc[0], fc[1], fq[2]);
    const float zm = leap_tri(zl[0][0], zl[1][0], zl[0][1], zl[1][1], c[0][0][0], c[1][0][0], c[0][1][0], c[1][1][0],
                              fc[0], fc[1], fm[2]);
    return mk(xm - xp, yp - ym, zm - zp) / 2.0f;
  }
};

// GLGridLeaper-GradientTools.glsl ComputeGradient with one-voxel taps
template <typename T>
__device__ __forceinline__ V3 gradient_vox(const T* __restrict__ d, int nx, int ny, int nz, V3 c) {
  const float xp = sample_vox<T>(d, nx, ny, nz, c + mk(+1.0f, 0, 0));
  const float xm = sample_vox<T>(d, nx, ny, nz, c + mk(-1.0f, 0, 0));
  const float yp = sample_vox<T>(d, nx, ny, nz, c + mk(0, -1.0f, 0));
  const float ym = sample_vox<T>(d, nx, ny, nz, c + mk(0, +1.0f, 0));
  const float zp = sample_vox<T>(d, nx, ny, nz, c + mk(0, 0, +1.0f));
  const float zm = sample_vox<T>(d, nx, ny, nz, c + mk(0, 0, -1.0f));
  return mk(xm - xp, yp - ym, zm - zp) / 2.0f;
}

// lighting.glsl:33-42 with the model-space eye
__device__ __forceinline__ V3 lighting_eye(V3 eye, V3 pos, V3 n, V3 amb, V3 dif, V3 spc, V3 L) {
  const V3 view = normalize(eye - pos);
  const V3 R = normalize(reflect(view, n));
  const float dd = fmaxf(fabsf(dot(n, L)), 0.0f);
  const float ss = pow8(fmaxf(dot(R, L), 0.0f));
  const V3 c = amb + dif * dd + spc * ss;
  return mk(clamp01(c.x), clamp01(c.y), clamp01(c.z));
}

// Compose-FS.glsl Lighting(): eye-space hit position and normal, eye at the origin
__device__ __forceinline__ V3 compose_lighting(V3 pos, V3 n, V3 amb, V3 dif, V3 spc, V3 L) {
  n.z = fabsf(n.z);
  const V3 view = normalize(mk(0, 0, 0) - pos);
  const V3 R = normalize(reflect(view, n));
  const float dd = fmaxf(fabsf(dot(n, L * -1.0f)), 0.0f);
  const float ss = pow8(fmaxf(dot(R, L), 0.0f));
  const V3 c = amb + dif * dd + spc * ss;
  return mk(clamp01(c.x), clamp01(c.y), clamp01(c.z));
}

__device__ __forceinline__ V3 xf_point(V3 p, const float* M) {
  return mk(p.x * M[0] + p.y * M[4] + p.z * M[8] + 1.0f * M[12], p.x * M[1] + p.y * M[5] + p.z * M[9] + 1.0f * M[13],
            p.x * M[2] + p.y * M[6] + p.z * M[10] + 1.0f * M[14]);
}
__device__ __forceinline__ V3 xf_dir(V3 p, const float* M) {
  return mk(p.x * M[0] + p.y * M[4] + p.z * M[8] + 0.0f * M[12], p.x * M[1] + p.y * M[5] + p.z * M[9] + 0.0f * M[13],
            p.x * M[2] + p.y * M[6] + p.z * M[10] + 0.0f * M[14]);
}

__device__ __forceinline__ uint32_t brick_id(const LeapArgs& A, uint32_t x, uint32_t y, uint32_t z, uint32_t l) {
  return A.lodoff[l] + x + y * A.ilayout[l][0] + z * A.ilayout[l][0] * A.ilayout[l][1];
}

// ReportMissingBrick: GLHashTable's insertion (GLHashTable.cpp:140-181).  The
// brick's 4D index over the finest level's brick layout, plus one, is stored at
// its value modulo the table size, or, when another brick holds that entry, at
// the next of `rehash` consecutive positions; a request that finds no free
// entry is dropped (the host pages what the table holds, and rays needing the
// dropped brick report it again in the next pass).
__device__ __forceinline__ void report_missing(const LeapArgs& A, const uint32_t bc[4]) {
  const uint32_t lx = A.ilayout[0][0], ly = A.ilayout[0][1], lz = A.ilayout[0][2];
  const uint32_t v = 1u + bc[0] + bc[1] * lx + bc[2] * lx * ly + bc[3] * lx * ly * lz;
  uint32_t r = 0;
  do {   // at least one probe, as the reference's do-while
    const uint32_t h = (v + r) % A.table_size;
    const uint32_t old = atomicCAS(&A.table[h], 0u, v);
    if (old == 0u || old == v) return;
  } while (++r < A.rehash);
}

// GLVolumePool.cpp:653-656, the floor of the logarithm read from the exponent
__device__ __forceinline__ uint32_t compute_lod(const LeapArgs& A, float depth) {
  const float x = A.lod_factor * (-depth) / A.l0err;
  const uint32_t u = __float_as_uint(x);
  const int e = (int)((u >> 23) & 0xffu) - 127;
  const uint32_t lod = (x >= 1.0f && e > 0) ? (uint32_t)e : 0u;
  return min((uint32_t)(A.nlod - 1), lod);
}

template <int MODE, bool LIT, typename T>
__device__ __forceinline__ void leap_tile(const LeapArgs& A, uint32_t tile, int lane) {
  const int tx = (int)(tile % A.tiles_x), ty = (int)(tile / A.tiles_x);
  const int px = tx * 8 + (lane & 7), py = ty * 8 + (lane >> 3);
  if (px >= A.W || py >= A.H) return;   // no wave-level operations below
  const size_t pix = (size_t)py * (size_t)A.W + (size_t)px;
  const ShadeConst& S = A.S;

  // pixel ray through the near plane
  const float* ip = A.inv_proj;
  const float n0 = ((float)px + 0.5f) * 2.0f / (float)A.W - 1.0f;
  const float n1 = ((float)py + 0.5f) * 2.0f / (float)A.H - 1.0f;
  float e[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) e[c] = n0 * ip[c] + n1 * ip[4 + c] + -1.0f * ip[8 + c] + 1.0f * ip[12 + c];
  const V3 d = mk(e[0] / e[3], e[1] / e[3], e[2] / e[3]);
  const V3 o_n = xf_point(mk(0, 0, 0), A.emm), d_n = xf_dir(d, A.emm);
  float tn = -INFINITY, tfar = INFINITY;
  bool hit = true;
  {
    const float on[3] = {o_n.x, o_n.y, o_n.z}, dn[3] = {d_n.x, d_n.y, d_n.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (dn[a] == 0.0f) {
        if (on[a] < 0.0f || on[a] > 1.0f) hit = false;
        continue;
      }
      const float inv = 1.0f / dn[a];
      const float t0 = (0.0f - on[a]) * inv, t1 = (1.0f - on[a]) * inv;
      tn = fmaxf(tn, fminf(t0, t1));
      tfar = fminf(tfar, fmaxf(t0, t1));
    }
  }
  // ray start: the box front face (or the near plane) in the first pass, the
  // resume state afterwards
  V3 entryN;
  float entryDepth;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  bool live;
  if (A.first) {
    const float te = fmaxf(tn, 1.0f);
    live = hit && tfar > tn && tfar > te;
    const V3 pe = d * te;
    entryN = xf_point(pe, A.emm);
    entryDepth = pe.z;
  } else {
    const float4 rp = A.resume[2 * pix], rc = A.resume[2 * pix + 1];
    entryN = mk(rp.x, rp.y, rp.z);
    entryDepth = rp.w;
    live = rp.w != 1000.0f;
    acc[0] = rc.x; acc[1] = rc.y; acc[2] = rc.z; acc[3] = rc.w;
  }
  float4 rpos = make_float4(entryN.x, entryN.y, entryN.z, entryDepth);
  float4 rcol = make_float4(acc[0], acc[1], acc[2], acc[3]);
  uint32_t ns = 0;
  bool optimal = true;
  // isosurface hit (MODE 2): brick voxel position, brick transform and data
  bool have_hit = false;
  V3 hit_b = mk(0, 0, 0), hit_tr = mk(0, 0, 0), hit_sc = mk(1, 1, 1);
  const T* hit_d = nullptr;
  int hit_n3[3] = {1, 1, 1};
  if (live) {
    const V3 pe = d * tfar;
    const V3 exitN = xf_point(pe, A.emm);
    const float exitDepth = pe.z;
    const V3 direction = exitN - entryN;
    const float rayLength = length(direction);
    const V3 vsz = mk(A.vsz[0], A.vsz[1], A.vsz[2]);
    const V3 dv = normalize(direction * vsz);
    const V3 step = dv / (2.0f * A.srm);
    const float stepSize = length(step);
    const float voxelSize = 0.125f / 2000.0f;
    const float oc = 1.0f / A.srm;
    const V3 dsc = mk(S.dscale[0], S.dscale[1], S.dscale[2]);
    const V3 amb = mk(S.amb[0], S.amb[1], S.amb[2]), dif = mk(S.dif[0], S.dif[1], S.dif[2]);
    const V3 spc = mk(S.spc[0], S.spc[1], S.spc[2]);
    const V3 meye = mk(A.meye[0], A.meye[1], A.meye[2]), mlight = mk(A.mlight[0], A.mlight[1], A.mlight[2]);
    const V3 inner = mk(A.inner[0], A.inner[1], A.inner[2]);
    const float ov = (float)A.overlap;
    float tt = 0.0f;
    V3 cur = entryN;
    uint32_t last[4] = {0u, 0u, 0u, 9999u};
    bool done = false;
    if (rayLength > voxelSize) {
      for (int j = 0; j < 100 && !done; ++j) {
        const float depth = entryDepth * (1.0f - tt) + exitDepth * tt;
        uint32_t lod = compute_lod(A, depth);
        const V3 ne = mk(fminf(fmaxf(cur.x, 0.0f), 1.0f), fminf(fmaxf(cur.y, 0.0f), 1.0f),
                         fminf(fmaxf(cur.z, 0.0f), 1.0f));
        bool found_req = true;
        uint32_t bc[4];
        auto coords = [&](uint32_t l) {
          bc[0] = (uint32_t)(ne.x * A.layout[l][0]);
          bc[1] = (uint32_t)(ne.y * A.layout[l][1]);
          bc[2] = (uint32_t)(ne.z * A.layout[l][2]);
          bc[3] = l;
        };
        coords(lod);
        uint32_t info = A.meta[brick_id(A, bc[0], bc[1], bc[2], lod)];
        if (info == BI_MISSING) {
          report_missing(A, bc);
          found_req = false;
          do {
            ++lod;
            coords(lod);
            info = A.meta[brick_id(A, bc[0], bc[1], bc[2], lod)];
            // BS_RequestAll, the reference's default strategy: coarser missing bricks too
            if (info == BI_MISSING) report_missing(A, bc);
          } while (info == BI_MISSING && lod + 1 < (uint32_t)A.nlod);
        }
        const bool bEmpty = info <= BI_EMPTY;
        if (bEmpty) {
          for (uint32_t lr = lod + 1; lr < (uint32_t)(A.nlod - 1); ++lr) {
            const uint32_t s0 = bc[0], s1 = bc[1], s2 = bc[2], s3 = bc[3];
            coords(lr);
            const uint32_t li = A.meta[brick_id(A, bc[0], bc[1], bc[2], lr)];
            if (li == BI_CHILD_EMPTY) {
              info = li;
              lod = lr;
            } else {
              bc[0] = s0; bc[1] = s1; bc[2] = s2; bc[3] = s3;
              break;
            }
          }
        }
        const V3 c0 = mk((float)bc[0] / A.layout[lod][0], (float)bc[1] / A.layout[lod][1],
                         (float)bc[2] / A.layout[lod][2]);
        const V3 c1 = mk((float)(bc[0] + 1) / A.layout[lod][0], (float)(bc[1] + 1) / A.layout[lod][1],
                         (float)(bc[2] + 1) / A.layout[lod][2]);
        V3 bexit;
        {
          const V3 div = mk(1.0f / direction.x, 1.0f / direction.y, 1.0f / direction.z);
          const float tix = ((div.x >= 0.0f ? c1.x : c0.x) - ne.x) * div.x;
          const float tiy = ((div.y >= 0.0f ? c1.y : c0.y) - ne.y) * div.y;
          const float tiz = ((div.z >= 0.0f ? c1.z : c0.z) - ne.z) * div.z;
          bexit = ne + direction * fminf(fminf(tix, tiy), tiz);
        }
        if (!found_req && optimal) {
          optimal = false;
          rpos = make_float4(cur.x, cur.y, cur.z, depth);
          rcol = make_float4(acc[0], acc[1], acc[2], acc[3]);
        }
        if (!bEmpty && (last[0] != bc[0] || last[1] != bc[1] || last[2] != bc[2] || last[3] != bc[3])) {
          const GatherBrick& gb = A.bricks[info - BI_FLAG_COUNT];
          const T* __restrict__ bd = reinterpret_cast<const T*>(S.pool) + gb.offset;
          const int nx = (int)gb.n[0], ny = (int)gb.n[1], nz = (int)gb.n[2];
          const V3 sc = mk(inner.x / (c1.x - c0.x), inner.y / (c1.y - c0.y), inner.z / (c1.z - c0.z));
          const V3 tr = mk(ov, ov, ov) - c0 * sc;
          const V3 pin = ne * sc + tr, pout = bexit * sc + tr;
          int steps = (int)ceilf(length(pout - pin) / stepSize);
          steps = min(steps, (int)ceilf(length((exitN - cur) * sc) / stepSize));
          V3 b = pin;
          for (int i = 0; i < steps; ++i) {
            ++ns;
            // value (and, for gradients, the cross) from the 32-voxel cross when
            // every active lane's neighbourhood lies inside its brick
            LeapCross<T> X;
            const bool cross = MODE != 2 && (LIT || MODE == 1) && __ballot(!X.coords(nx, ny, nz, b)) == 0;
            const float data = cross ? X.value(bd, nx, ny) : sample_vox_fast<T>(bd, nx, ny, nz, b);
            auto grad = [&]() { return cross ? X.gradient() : gradient_vox<T>(bd, nx, ny, nz, b); };
            if constexpr (MODE == 2) {
              // GLGridLeaper-iso.glsl: first sample at or above the isovalue,
              // then RefineIsosurface (GLGridLeaper-Method-iso.glsl) in brick voxels
              if (data >= S.iso_n) {
                V3 rd = step / 2.0f;
                V3 bb = b - rd;
                for (int k = 0; k < S.refine_steps; ++k) {
                  rd = rd / 2.0f;
                  if (sample_vox_fast<T>(bd, nx, ny, nz, bb) >= S.iso_n) bb = bb - rd; else bb = bb + rd;
                }
                // the normal is taken once the brick loop is over
                hit_b = bb; hit_tr = tr; hit_sc = sc; hit_d = bd; hit_n3[0] = nx; hit_n3[1] = ny; hit_n3[2] = nz;
                have_hit = true;
                done = true;
                break;
              }
              b = b + step;
              continue;
            }
            float c[4];
            if (MODE == 0) {
              const float cc = fminf(fmaxf(data * S.tf_scale, 0.0f), 1.0f);
              const int ti = min((int)floorf(cc * (float)S.tf_n), S.tf_n - 1);
              const float4 tv = S.tf1[ti];
              c[0] = tv.x; c[1] = tv.y; c[2] = tv.z; c[3] = tv.w;
              if (LIT) {
                V3 g = grad() * dsc;
                const float l = length(g);
                if (l > 0.0f) g = mk(g.x / l, g.y / l, g.z / l);
                const V3 lc = lighting_eye(meye, (b - tr) / sc, g, amb, mk(c[0], c[1], c[2]) * dif, spc, mlight);
                c[0] = lc.x; c[1] = lc.y; c[2] = lc.z;
              }
            } else {
              const float sx = fminf(fmaxf(data * S.tf_scale, 0.0f), 1.0f);
              const int xi = min((int)floorf(sx * (float)S.tf_w), S.tf_w - 1);
              if (((S.tf2_cols[xi >> 6] >> (xi & 63)) & 1ull) == 0) {
                // every row of this value column is transparent: alpha 0 whatever
                // the gradient, so the sample composites to nothing
                c[0] = c[1] = c[2] = c[3] = 0.0f;
              } else {
                const V3 g = grad();
                const float gm = length(g);
                const float sy = fminf(fmaxf(1.0f - gm * S.grad_scale, 0.0f), 1.0f);
                const int yi = min((int)floorf(sy * (float)S.tf_h), S.tf_h - 1);
                const float4 tv = S.tf2f[(size_t)yi * (size_t)S.tf_w + (size_t)xi];
                c[0] = tv.x; c[1] = tv.y; c[2] = tv.z; c[3] = tv.w;
                if (LIT) {
                  const V3 nrm = dsc * (gm > 0.0f ? g / gm : g);
                  const V3 lc = lighting_eye(meye, (b - tr) / sc, nrm, amb, mk(c[0], c[1], c[2]) * dif, spc, mlight);
                  c[0] = lc.x; c[1] = lc.y; c[2] = lc.z;
                }
              }
            }
            c[3] = 1.0f - opacity_pow(1.0f - c[3], oc);
            // UnderCompositing (Compositing.glsl:33-38)
            acc[0] += c[0] * (1.0f - acc[3]) * c[3];
            acc[1] += c[1] * (1.0f - acc[3]) * c[3];
            acc[2] += c[2] * (1.0f - acc[3]) * c[3];
            acc[3] += c[3] * (1.0f - acc[3]);
            if (acc[3] > 0.99f) { done = true; break; }
            b = b + step;
          }
          if (!done) cur = (b - tr) / sc;
        } else {
          cur = bexit + direction * voxelSize / rayLength;
        }
        if (done) break;
        last[0] = bc[0]; last[1] = bc[1]; last[2] = bc[2]; last[3] = bc[3];
        tt = length(entryN - bexit) / rayLength;
        if (tt > 0.9999f) break;
      }
    }
    // TerminateRay
    if (optimal) {
      // isosurface: 499 + rayHitPos.a, i.e. 501 for a (white) hit, 1000 without one
      rpos.w = (MODE == 2 && have_hit) ? 501.0f : 1000.0f;
      rcol = make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
  } else {
    rpos.w = 1000.0f;
  }
  if constexpr (MODE == 2) {
    // rayHitPos = mModelToEye * hit, rayHitNormal = mModelViewIT * ComputeNormal
    // (GLGridLeaper-GradientTools.glsl), then Compose-FS.glsl shading
    float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
    if (have_hit) {
      const V3 he = xf_point((hit_b - hit_tr) / hit_sc, A.mte);
      V3 g = gradient_vox<T>(hit_d, hit_n3[0], hit_n3[1], hit_n3[2], hit_b) * mk(S.dscale[0], S.dscale[1], S.dscale[2]);
      const float l = length(g);
      if (l > 0.0f) g = mk(g.x / l, g.y / l, g.z / l);
      const float* N = A.nit;
      const V3 ne = mk(g.x * N[0] + g.y * N[3] + g.z * N[6], g.x * N[1] + g.y * N[4] + g.z * N[7],
                       g.x * N[2] + g.y * N[5] + g.z * N[8]);
      const V3 isoc = mk(S.iso_col[0], S.iso_col[1], S.iso_col[2]);
      const V3 c = compose_lighting(he, ne, mk(S.amb[0], S.amb[1], S.amb[2]),
                                    mk(S.dif[0], S.dif[1], S.dif[2]) * isoc, mk(S.spc[0], S.spc[1], S.spc[2]),
                                    mk(S.ldir[0], S.ldir[1], S.ldir[2]));
      res = make_float4(c.x, c.y, c.z, 1.0f);
    }
    reinterpret_cast<float4*>(A.out)[pix] = res;
  } else {
    reinterpret_cast<float4*>(A.out)[pix] = make_float4(acc[0], acc[1], acc[2], acc[3]);
  }
  A.resume[2 * pix] = rpos;
  A.resume[2 * pix + 1] = rcol;
  if (A.counter && ns) atomicAdd(A.counter, (unsigned long long)ns);
}

// one 8x8 tile per wave of the grid
template <int MODE, bool LIT, typename T>
__global__ __launch_bounds__(256) void leap_kernel(LeapArgs A) {
  leap_tile<MODE, LIT, T>(A, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

// persistent grid: every wave pulls 8x8 tiles from A.queue until tiles
// 0 .. ntiles-1 are drained (every wave reaches that exit; the counter keeps
// growing past ntiles: harmless)
template <int MODE, bool LIT, typename T>
__global__ __launch_bounds__(256) void leap_kernel_queue(LeapArgs A) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(A.queue, 1u);
    t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
    if (t >= A.ntiles) break;
    leap_tile<MODE, LIT, T>(A, t, lane);
  }
}

template <int M, bool L, typename T>
void launch_leap_one(const LeapArgs& a, uint32_t blocks, hipStream_t st) {
  if (a.queue) {
    // persistent grid: the workgroups the device holds at once
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(leap_kernel_queue<M, L, T>), 256, 0) != hipSuccess || per_cu <= 0)
      per_cu = 1;
    blocks = std::min(blocks, (uint32_t)per_cu * (uint32_t)cus);
    hipLaunchKernelGGL((leap_kernel_queue<M, L, T>), dim3(blocks), dim3(256), 0, st, a);
  } else {
    hipLaunchKernelGGL((leap_kernel<M, L, T>), dim3(blocks), dim3(256), 0, st, a);
  }
}

}  // namespace

hipError_t launch_leap(const LeapArgs& a, int bpv, hipStream_t st) {
  const uint32_t tiles_y = ((uint32_t)a.H + 7u) / 8u;
  const uint32_t tiles = a.tiles_x * tiles_y;
  const uint32_t blocks = (tiles + 3u) / 4u;
  if (blocks == 0) return hipSuccess;
  const bool u16 = bpv == 2, lit = a.lighting != 0;
  if (a.mode == 0) {
    if (lit) { if (u16) launch_leap_one<0, true, uint16_t>(a, blocks, st); else launch_leap_one<0, true, uint8_t>(a, blocks, st); }
    else { if (u16) launch_leap_one<0, false, uint16_t>(a, blocks, st); else launch_leap_one<0, false, uint8_t>(a, blocks, st); }
  } else if (a.mode == 1) {
    if (lit) { if (u16) launch_leap_one<1, true, uint16_t>(a, blocks, st); else launch_leap_one<1, true, uint8_t>(a, blocks, st); }
    else { if (u16) launch_leap_one<1, false, uint16_t>(a, blocks, st); else launch_leap_one<1, false, uint8_t>(a, blocks, st); }
  } else if (a.mode == 2) {
    if (u16) launch_leap_one<2, false, uint16_t>(a, blocks, st); else launch_leap_one<2, false, uint8_t>(a, blocks, st);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tvk
