// Device-side bricking of a linear volume into the ExtendedOctree brick layout.
//
// The reference builds LOD 0 brick by brick (ExtendedOctreeConverter.cpp:288-462,
// :1381-1414), the coarser levels by 2x box filtering of the previous level's
// bricks (ExtendedOctreeConverter.inc:1-398) followed by FillOverlap
// (ExtendedOctreeConverter.cpp:1203-1372), and per-brick min/max over every
// stored voxel (ExtendedOctreeConverter.inc:408-444).  For even brick cores
// these steps have a closed form, which is what runs here as bandwidth-bound
// gathers:
//   LOD k+1 voxel g = floor(mean of the in-range voxels of LOD k at 2g + {0,1}^3)
//   brick voxel i   = LOD voxel (coords*core - overlap + i), 0 outside the
//                     volume (clamped to the edge with bClampToEdge)
// Bit-exactness against the procedural restatement is checked by the tests.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>
#include <algorithm>
#include "../tvk_types.hpp"

namespace tvk {

template <typename T> __device__ __forceinline__ int lx_of() { return sizeof(T) == 1 ? 3 : 2; }

// decode a storage index of the blocked brick layout into voxel coordinates;
// the second half of a row decodes to the next block's voxels (callers clamp
// coordinates to the brick, which gives the padding its repeated last voxel)
__device__ __forceinline__ void unswz(uint64_t i, uint32_t nbx, uint32_t nby, int lx, uint32_t& x, uint32_t& y,
                                      uint32_t& z) {
  const uint64_t blk = i >> (lx + 5);
  const uint32_t r = (uint32_t)(i & ((1u << (lx + 5)) - 1u));
  x = ((uint32_t)(blk % nbx) << lx) + (r & ((2u << lx) - 1u));
  const uint64_t q = blk / nbx;
  y = (uint32_t)(q % nby) << 2 | ((r >> (lx + 1)) & 3u);
  z = (uint32_t)(q / nby) << 2 | (r >> (lx + 3));
}

// Second half of every row := the voxels it repeats (the next block's row, or
// the clamped last voxel), after the overlap fill changed some of them.
template <typename T>
__global__ __launch_bounds__(256) void extend_rows(T* __restrict__ pool, const GatherBrick* __restrict__ bricks) {
  const GatherBrick b = bricks[blockIdx.y];
  const int lx = lx_of<T>();
  const uint32_t nbx = nblocks(b.n[0], lx), nby = nblocks(b.n[1], 2);
  const uint64_t half = padded_count(b.n, lx) >> 1;   // one half-row of every row
  for (uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x; j < half; j += (uint64_t)gridDim.x * 256) {
    const uint64_t i = ((j >> lx) << (lx + 1)) + (1u << lx) + (j & ((1u << lx) - 1u));   // an extension slot
    uint32_t x, y, z;
    unswz(i, nbx, nby, lx, x, y, z);
    x = min(x, b.n[0] - 1); y = min(y, b.n[1] - 1); z = min(z, b.n[2] - 1);
    pool[b.offset + i] = pool[b.offset + swz(x, y, z, nbx, nby, lx)];
  }
}

// LOD brick gather through LDS.  A work item is one z/y block row of a brick
// and a chunk of GATHER_CB blocks along x: its 16 x-rows (4 y by 4 z, clamped
// to the brick: padding rows repeat the last one) are read from the linear
// volume with consecutive lanes on consecutive voxels, then written as the
// blocked layout's 16-byte rows (own W voxels + the W they repeat, the last
// block's clamped to the brick's last voxel) with consecutive lanes on
// consecutive rows.  When mm is given, the brick's min/max over the own voxels
// (every stored voxel is the own voxel of exactly one row; padding repeats real
// ones) is folded into mm[2 * brick] / mm[2 * brick + 1] (init_minmax first).
// A workgroup is one (brick, slot) pair: slot s takes the items s, s + gx, ...
// of its brick, each item's row loads issued together (all in flight), the
// pool rows written with streaming stores.  Workgroups are dealt to the 8 XCDs in turn, so the linear
// workgroup id is split into (XCD, sequence) and each XCD gets a contiguous run
// of the order (brick row, slot, brick within the row of row_len bricks along
// x): the bricks of one row run side by side on one XCD, and the 128-byte lines
// that neighbouring bricks' row segments share are read through one L2.
// grid: 8 * per_xcd workgroups; those past the nbrk bricks return at once.
constexpr int GATHER_CB = 64;
template <typename T>
__global__ __launch_bounds__(256) void gather_rows(const T* __restrict__ lin, uint64_t X, uint64_t Y, uint64_t Z,
                                                   const GatherBrick* __restrict__ bricks, T* __restrict__ pool,
                                                   uint32_t core_x, uint32_t core_y, uint32_t core_z, uint32_t ov,
                                                   int clamp, uint32_t* __restrict__ mm, uint32_t nbrk, uint32_t gx,
                                                   uint32_t row_len, uint32_t per_xcd) {
  constexpr int lx = sizeof(T) == 1 ? 3 : 2, W = 1 << lx;
  constexpr int L = GATHER_CB * W + W;          // voxels of one LDS row (8-byte multiple)
  __shared__ __align__(16) T row[16][L];
  const uint32_t v = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
  const uint32_t bc = v % row_len, slot = (v / row_len) % gx, brow = v / (row_len * gx);
  const uint64_t bi = (uint64_t)brow * row_len + bc;
  if (bi >= nbrk) return;   // the whole workgroup, before any barrier
  const GatherBrick b = bricks[bi];
  const uint32_t nbx = nblocks(b.n[0], lx), nby = nblocks(b.n[1], 2), nbz = nblocks(b.n[2], 2);
  const uint32_t nch = (nbx + GATHER_CB - 1) / GATHER_CB;
  const uint32_t items = nch * nby * nbz;
  const int64_t ox0 = (int64_t)b.c[0] * core_x - ov, oy0 = (int64_t)b.c[1] * core_y - ov;
  const int64_t oz0 = (int64_t)b.c[2] * core_z - ov;
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  uint32_t mn = 0xffffffffu, mx = 0u;
  for (uint32_t it = slot; it < items; it += gx) {
    const uint32_t ch = it % nch, yb = (it / nch) % nby, zb = it / (nch * nby);
    const uint32_t bx0 = ch * GATHER_CB, nb = min((uint32_t)GATHER_CB, nbx - bx0);
    const uint32_t x0 = bx0 << lx, nl = (nb << lx) + W;   // voxels loaded per row
    __syncthreads();   // the previous item's rows are consumed
    // wave wv loads rows 4wv .. 4wv+3 (y = r & 3, z = r >> 2 of the block row),
    // NJ = ceil(nl / 64) voxels per lane and row.  Every load is issued before
    // the first LDS write, from addresses clamped into the volume (the border
    // value is selected afterwards), so that a wave keeps its loads in flight
    // together instead of waiting for each in turn.
    const int32_t ox = (int32_t)ox0, xmax = (int32_t)X - 1;
    auto stage = [&](auto njc) {
      constexpr int NJ = decltype(njc)::value;
      T vals[4][NJ];
      uint32_t ok = 0u;   // bit k * NJ + j: the voxel lies in the volume
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t r = (uint32_t)(wv * 4 + k);
        const uint32_t y = min((yb << 2) | (r & 3u), b.n[1] - 1), z = min((zb << 2) | (r >> 2), b.n[2] - 1);
        const int64_t gy = oy0 + y, gz = oz0 + z;
        const bool rin = gy >= 0 && gz >= 0 && gy < (int64_t)Y && gz < (int64_t)Z;
        const int64_t cy = gy < 0 ? 0 : (gy >= (int64_t)Y ? (int64_t)Y - 1 : gy);
        const int64_t cz = gz < 0 ? 0 : (gz >= (int64_t)Z ? (int64_t)Z - 1 : gz);
        // a row of the linear volume is addressed from one (wave-uniform) base
        // with 32-bit voxel offsets (X is at most 2^31)
        const T* __restrict__ src = lin + X * ((uint64_t)cy + Y * (uint64_t)cz);
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const uint32_t xi = (uint32_t)ln + 64u * (uint32_t)j;
          const int32_t gxv = ox + (int32_t)min(x0 + xi, b.n[0] - 1);
          vals[k][j] = src[min(max(gxv, 0), xmax)];
          if (rin && gxv >= 0 && gxv <= xmax) ok |= 1u << (k * NJ + j);
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const uint32_t xi = (uint32_t)ln + 64u * (uint32_t)j;
          if (xi < nl) row[wv * 4 + k][xi] = (clamp || ((ok >> (k * NJ + j)) & 1u)) ? vals[k][j] : (T)0;
        }
    };
    static_assert((L + 63) / 64 <= 9, "rows of at most 9 x 64 voxels");
    switch ((nl + 63u) / 64u) {
      case 1: stage(std::integral_constant<int, 1>{}); break;
      case 2: stage(std::integral_constant<int, 2>{}); break;
      case 3: stage(std::integral_constant<int, 3>{}); break;
      case 4: stage(std::integral_constant<int, 4>{}); break;
      case 5: stage(std::integral_constant<int, 5>{}); break;
      case 6: stage(std::integral_constant<int, 6>{}); break;
      case 7: stage(std::integral_constant<int, 7>{}); break;
      case 8: stage(std::integral_constant<int, 8>{}); break;
      default: stage(std::integral_constant<int, 9>{}); break;
    }
    __syncthreads();
    // consecutive lanes on consecutive 16-byte rows: own W voxels + the W repeated
    const uint64_t blk0 = ((uint64_t)zb * nby + yb) * nbx + bx0;
    for (uint32_t t = threadIdx.x; t < nb * 16u; t += 256) {
      const uint32_t bl = t >> 4, r = t & 15u;
      const uint2 own = *reinterpret_cast<const uint2*>(&row[r][bl << lx]);
      const uint2 rep = *reinterpret_cast<const uint2*>(&row[r][(bl << lx) + W]);
      const uint32_t ww[2] = {own.x, own.y};
#pragma unroll
      for (int e = 0; e < W; ++e) {
        const uint32_t v =
            (ww[e * (int)sizeof(T) / 4] >> (8 * ((e * (int)sizeof(T)) & 3))) & ((1u << (8 * sizeof(T))) - 1u);
        mn = min(mn, v);
        mx = max(mx, v);
      }
      // streaming store: the pool is not read again by this launch
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      const u32x4 q = {own.x, own.y, rep.x, rep.y};
      __builtin_nontemporal_store(q, reinterpret_cast<u32x4*>(pool + b.offset + ((blk0 + bl) * 16u + r) * (2 * W)));
    }
  }
  if (mm) {
    for (int o = 32; o > 0; o >>= 1) {
      mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, 64));
      mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMin(mm + 2 * bi, mn);
      atomicMax(mm + 2 * bi + 1, mx);
    }
  }
}

__global__ void init_minmax(uint32_t* mm, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { mm[2 * i] = 0xffffffffu; mm[2 * i + 1] = 0u; }
}

// blocked pool -> reference layout (x fastest, unpadded) at ref offsets
template <typename T>
__global__ __launch_bounds__(256) void unswizzle_bricks(const T* __restrict__ pool, const GatherBrick* __restrict__ bricks,
                                                        const uint64_t* __restrict__ ref_off, T* __restrict__ out) {
  const GatherBrick b = bricks[blockIdx.y];
  const int lx = lx_of<T>();
  const uint32_t nbx = nblocks(b.n[0], lx), nby = nblocks(b.n[1], 2);
  const uint64_t n = (uint64_t)b.n[0] * b.n[1] * b.n[2];
  const uint64_t ro = ref_off[blockIdx.y];
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint32_t x = (uint32_t)(i % b.n[0]);
    const uint64_t r = i / b.n[0];
    const uint32_t y = (uint32_t)(r % b.n[1]), z = (uint32_t)(r / b.n[1]);
    out[ro + i] = pool[b.offset + swz(x, y, z, nbx, nby, lx)];
  }
}

// reference layout (x fastest, unpadded, at src_off[brick]) -> blocked pool at
// bricks[].offset: every element of the padded layout, the repeated half-rows
// included, is the brick voxel its index decodes to (clamped to the brick)
template <typename T>
__global__ __launch_bounds__(256) void swizzle_bricks(const T* __restrict__ src, const GatherBrick* __restrict__ bricks,
                                                      const uint64_t* __restrict__ src_off, T* __restrict__ pool) {
  const GatherBrick b = bricks[blockIdx.y];
  const int lx = lx_of<T>();
  const uint32_t nbx = nblocks(b.n[0], lx), nby = nblocks(b.n[1], 2);
  const uint64_t n = padded_count(b.n, lx);
  const uint64_t so = src_off[blockIdx.y];
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    uint32_t x, y, z;
    unswz(i, nbx, nby, lx, x, y, z);
    x = min(x, b.n[0] - 1); y = min(y, b.n[1] - 1); z = min(z, b.n[2] - 1);
    pool[b.offset + i] = src[so + x + (uint64_t)b.n[0] * (y + (uint64_t)b.n[1] * z)];
  }
}

// The LOD-0 volume from the pool (UVFDataset::Export of level 0): voxel (x,y,z)
// is the core voxel of the brick that owns it.  bricks: the LOD-0 bricks in TOC
// order (nbx x nby x nbz of them).
template <typename T>
__global__ __launch_bounds__(256) void flatten_lod0(const T* __restrict__ pool, const GatherBrick* __restrict__ bricks,
                                                    uint32_t nbx, uint32_t nby, uint32_t cx, uint32_t cy, uint32_t cz,
                                                    uint32_t ov, uint64_t X, uint64_t Y, uint64_t Z, T* __restrict__ out) {
  const int lx = lx_of<T>();
  const uint64_t n = X * Y * Z;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint32_t x = (uint32_t)(i % X);
    const uint64_t r = i / X;
    const uint32_t y = (uint32_t)(r % Y), z = (uint32_t)(r / Y);
    const uint32_t bx = x / cx, by = y / cy, bz = z / cz;
    const GatherBrick b = bricks[((uint64_t)bz * nby + by) * nbx + bx];
    const uint32_t u = x - bx * cx + ov, v = y - by * cy + ov, w = z - bz * cz + ov;
    out[i] = pool[b.offset + swz(u, v, w, nblocks(b.n[0], lx), nblocks(b.n[1], 2), lx)];
  }
}

// UVFDataset::CropData (IO/uvfDataset.h:252-297): voxels whose normalised
// coordinate (x/X - 0.5, y/Y - 0.5, z/Z - 0.5) satisfies PLANE::clip,
// (n . p) + w >= 0, are set to 0.
template <typename T>
__global__ __launch_bounds__(256) void crop_plane(T* __restrict__ vol, uint64_t X, uint64_t Y, uint64_t Z, float px,
                                                  float py, float pz, float pw) {
  const uint64_t n = X * Y * Z;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t x = i % X, r = i / X, y = r % Y, z = r / Y;
    const float vx = (float)x / (float)X - 0.5f, vy = (float)y / (float)Y - 0.5f, vz = (float)z / (float)Z - 0.5f;
    // separate roundings, as the reference's float expression (no fused multiply-add)
    const float d = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(px, vx), __fmul_rn(py, vy)), __fmul_rn(pz, vz)), pw);
    if (d >= 0.0f) vol[i] = (T)0;
  }
}

// VolumeTools::Filter (VolumeTools.h:150-258) over the taps of one 2x2x2 group
// that exist, in the reference's order (x-major: p0 = (0,0,0), p1 = +z, p2 = +y,
// p3 = +y+z, p4 = +x, ...; the odd planes and lines of ExtendedOctreeConverter.inc:
// 60-248 keep the same relative order): the mean in double precision truncated
// to T, or the reference's median -- of 8: a sorting network over 4, two
// insertions and a clamp of the 7th tap; of 4: the median of the first three;
// of 2 and 1: the first tap.
template <typename T> __device__ __forceinline__ void order2(T& a, T& b) {
  if (a > b) { const T t = a; a = b; b = t; }
}
template <typename T> __device__ __forceinline__ void insert4(T& a, T& b, T& c, T& d, T p) {
  if (p > c) {
    order2(d, p);
  } else if (p < b) {
    d = c; c = b; b = p;
    order2(a, b);
  } else {
    d = c; c = p;
  }
}
template <typename T> __device__ __forceinline__ T filter_taps(const T* t, int n, bool median) {
  if (!median) {
    uint32_t sum = 0;
    for (int i = 0; i < n; ++i) sum += t[i];
    return (T)(sum / (uint32_t)n);   // double(sum) / n truncated: n is a power of two
  }
  if (n == 8) {
    T a = t[0], b = t[1], c = t[2], d = t[3];
    order2(a, b); order2(c, d); order2(a, c); order2(b, d); order2(b, c);
    insert4(a, b, c, d, t[4]);
    insert4(a, b, c, d, t[5]);
    return max(min(d, t[6]), c);
  }
  if (n == 4) {
    T a = t[0], b = t[1], c = t[2];
    order2(a, b); order2(b, c);
    return max(a, b);
  }
  return t[0];
}

// 2x filter of a level volume (mean or median); an axis of size 1 is not halved
// (ExtendedOctree.cpp:201-213).  With even brick cores the reference's
// brick-to-brick filtering (ExtendedOctreeConverter.inc) equals this
// whole-level form: the 2x2x2 groups never straddle a brick.  Grid: x along
// blockIdx.x * 256 + threadIdx.x, destination rows (y, z) along blockIdx.y in
// steps of gridDim.y, so no index is divided per voxel.  A voxel whose eight
// taps all exist loads them together (no per-tap branch); the taps keep the
// x-major order (dx, then dy, then dz) the filter sees in either form.
template <typename T>
__global__ __launch_bounds__(256) void downsample(const T* __restrict__ src, uint64_t SX, uint64_t SY, uint64_t SZ,
                                                  T* __restrict__ dst, uint64_t DX, uint64_t DY, uint64_t DZ,
                                                  int median) {
  const uint64_t fx = SX > 1 ? 2 : 1, fy = SY > 1 ? 2 : 1, fz = SZ > 1 ? 2 : 1;
  const uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (x >= DX) return;
  const uint64_t rows = DY * DZ;
  for (uint64_t r = blockIdx.y; r < rows; r += gridDim.y) {
    const uint64_t y = r % DY, z = r / DY;   // once per row and thread, not per voxel
    T taps[8];
    int cnt = 0;
    if (fx == 2 && fy == 2 && fz == 2 && 2 * x + 1 < SX && 2 * y + 1 < SY && 2 * z + 1 < SZ) {
      const T* __restrict__ s0 = src + 2 * x + SX * (2 * y + SY * (2 * z));
      const uint64_t oy = SX, oz = SX * SY;
      taps[0] = s0[0]; taps[1] = s0[oz]; taps[2] = s0[oy]; taps[3] = s0[oy + oz];
      taps[4] = s0[1]; taps[5] = s0[1 + oz]; taps[6] = s0[1 + oy]; taps[7] = s0[1 + oy + oz];
      cnt = 8;
    } else {
      for (uint64_t dx = 0; dx < fx; ++dx) {
        const uint64_t xx = x * fx + dx;
        if (xx >= SX) continue;
        for (uint64_t dy = 0; dy < fy; ++dy) {
          const uint64_t yy = y * fy + dy;
          if (yy >= SY) continue;
          for (uint64_t dz = 0; dz < fz; ++dz) {
            const uint64_t zz = z * fz + dz;
            if (zz >= SZ) continue;
            taps[cnt++] = src[xx + SX * (yy + SY * zz)];
          }
        }
      }
    }
    dst[x + DX * r] = filter_taps(taps, cnt, median != 0);
  }
}

// The reference's brick-to-brick filtering for odd brick cores
// (ExtendedOctreeConverter::DownsampleBrick, ExtendedOctreeConverter.inc:250-356):
// coarse brick B is cleared, then child 2B + d (d in {0,1}^3, if it exists)
// filters its core into the coarse core at offset d * ceil(core / 2): child
// voxel groups 2j, 2j+1 (only 2j in the last, odd plane) give coarse voxel
// j + offset, so the second child's last voxel lands in the first overlap
// voxel.  Positions no child writes stay 0.  The overlaps are FillOverlap's
// (fill_overlap_wave with the pre-fill snapshot) afterwards.
template <typename T>
__global__ __launch_bounds__(256) void downsample_bricks(T* __restrict__ pool, const GatherBrick* __restrict__ coarse,
                                                         const GatherBrick* __restrict__ fine, uint32_t fbx,
                                                         uint32_t fby, uint32_t fbz, uint32_t ov, uint32_t sx,
                                                         uint32_t sy, uint32_t sz, int median) {
  const GatherBrick B = coarse[blockIdx.y];
  const int lx = lx_of<T>();
  const uint32_t nbx = nblocks(B.n[0], lx), nby = nblocks(B.n[1], 2);
  const uint64_t n = padded_count(B.n, lx);
  const uint32_t split[3] = {sx, sy, sz}, fcount[3] = {fbx, fby, fbz};
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    uint32_t p[3];
    unswz(i, nbx, nby, lx, p[0], p[1], p[2]);
    p[0] = min(p[0], B.n[0] - 1); p[1] = min(p[1], B.n[1] - 1); p[2] = min(p[2], B.n[2] - 1);
    bool written = true;
    uint32_t child[3], lj[3], ev[3];
    for (int a = 0; a < 3 && written; ++a) {
      if (p[a] < ov) { written = false; break; }
      const uint32_t j = p[a] - ov;
      uint32_t d = 0, l = j;
      if (j >= split[a]) { d = 1; l = j - split[a]; }
      child[a] = 2 * B.c[a] + d;
      if (child[a] >= fcount[a]) { written = false; break; }
      lj[a] = l;
    }
    T v = 0;
    if (written) {
      const GatherBrick F = fine[(child[2] * fby + child[1]) * fbx + child[0]];
      uint32_t ntap[3];
      for (int a = 0; a < 3 && written; ++a) {
        ev[a] = (F.n[a] - 2 * ov) / 2;                     // evenSize of the child's core
        const uint32_t contrib = ev[a] + (F.n[a] & 1u);    // groups, the odd plane included
        if (lj[a] >= contrib) { written = false; break; }
        ntap[a] = lj[a] < ev[a] ? 2u : 1u;
      }
      if (written) {
        const uint32_t fnbx = nblocks(F.n[0], lx), fnby = nblocks(F.n[1], 2);
        T taps[8];
        int cnt = 0;
        for (uint32_t dx = 0; dx < ntap[0]; ++dx)
          for (uint32_t dy = 0; dy < ntap[1]; ++dy)
            for (uint32_t dz = 0; dz < ntap[2]; ++dz)
              taps[cnt++] = pool[F.offset + swz(2 * lj[0] + dx + ov, 2 * lj[1] + dy + ov, 2 * lj[2] + dz + ov, fnbx,
                                                fnby, lx)];
        v = filter_taps(taps, cnt, median != 0);
      }
    }
    pool[B.offset + i] = v;
  }
}

// per-brick min/max over all stored voxels: one thread per 16-byte row (its own
// W voxels; the repeated half and the padding hold copies of stored voxels),
// folded into out[2 * brick] / out[2 * brick + 1] (initialised by init_minmax).
// grid: x = chunks of the brick's rows, y = brick.
template <typename T>
__global__ __launch_bounds__(256) void brick_minmax(const T* __restrict__ pool, const GatherBrick* __restrict__ bricks,
                                                    uint32_t* __restrict__ out) {
  constexpr int lx = sizeof(T) == 1 ? 3 : 2, W = 1 << lx;
  __shared__ uint32_t s_mm[2][4];
  const GatherBrick b = bricks[blockIdx.y];
  const uint64_t rows = padded_count(b.n, lx) >> (lx + 1);
  uint32_t mn = 0xffffffffu, mx = 0u;
  auto fold = [&](uint2 w) {
    const uint32_t ww[2] = {w.x, w.y};
#pragma unroll
    for (int e = 0; e < W; ++e) {
      const uint32_t v = (ww[e * (int)sizeof(T) / 4] >> (8 * ((e * (int)sizeof(T)) & 3))) & ((1u << (8 * sizeof(T))) - 1u);
      mn = min(mn, v);
      mx = max(mx, v);
    }
  };
  // four rows per lane and pass, their loads issued together
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  for (; j + 3 * stride < rows; j += 4 * stride) {
    uint2 w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = *reinterpret_cast<const uint2*>(pool + b.offset + (j + q * stride) * (2 * W));
#pragma unroll
    for (int q = 0; q < 4; ++q) fold(w[q]);
  }
  for (; j < rows; j += stride) fold(*reinterpret_cast<const uint2*>(pool + b.offset + j * (2 * W)));   // the own half
  for (int o = 32; o > 0; o >>= 1) {
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, 64));
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, 64));
  }
  // one atomic pair per workgroup: the brick's counters are shared by all of its workgroups
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { s_mm[0][wv] = mn; s_mm[1][wv] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    mn = min(min(s_mm[0][0], s_mm[0][1]), min(s_mm[0][2], s_mm[0][3]));
    mx = max(max(s_mm[1][0], s_mm[1][1]), max(s_mm[1][2], s_mm[1][3]));
    atomicMin(out + 2 * blockIdx.y, mn);
    atomicMax(out + 2 * blockIdx.y + 1, mx);
  }
}

// Cell grid of every brick: one wave per cell reduces the stored voxels of the
// cell grown by CELL_DIL on every side (clipped to the brick) to (max << 16) | min.
template <typename T>
__global__ __launch_bounds__(256) void cell_minmax(const T* __restrict__ pool, const GatherBrick* __restrict__ bricks,
                                                   const uint32_t* __restrict__ base, uint32_t* __restrict__ out) {
  const GatherBrick b = bricks[blockIdx.y];
  const uint32_t ncx = cell_count(b.n[0]), ncy = cell_count(b.n[1]), ncz = cell_count(b.n[2]);
  const uint32_t cell = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (cell >= ncx * ncy * ncz) return;   // whole waves leave
  const int lane = threadIdx.x & 63;
  const uint32_t cx = cell % ncx, cy = (cell / ncx) % ncy, cz = cell / (ncx * ncy);
  const int x0 = max((int)(cx << CELL_LW) - CELL_DIL, 0), x1 = min((int)((cx + 1) << CELL_LW) + CELL_DIL, (int)b.n[0]);
  const int y0 = max((int)(cy << CELL_LW) - CELL_DIL, 0), y1 = min((int)((cy + 1) << CELL_LW) + CELL_DIL, (int)b.n[1]);
  const int z0 = max((int)(cz << CELL_LW) - CELL_DIL, 0), z1 = min((int)((cz + 1) << CELL_LW) + CELL_DIL, (int)b.n[2]);
  const int rx = x1 - x0, ry = y1 - y0, cnt = rx * ry * (z1 - z0);
  const int lx = lx_of<T>();
  const uint32_t nbx = nblocks(b.n[0], lx), nby = nblocks(b.n[1], 2);
  uint32_t mn = 0xffffu, mx = 0u;
  for (int i = lane; i < cnt; i += 64) {
    const int x = x0 + i % rx, t = i / rx, y = y0 + t % ry, z = z0 + t / ry;
    const uint32_t v = pool[b.offset + swz((uint32_t)x, (uint32_t)y, (uint32_t)z, nbx, nby, lx)];
    mn = min(mn, v);
    mx = max(mx, v);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn = min(mn, (uint32_t)__shfl_xor((int)mn, o, 64));
    mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, 64));
  }
  if (lane == 0) out[base[blockIdx.y] + cell] = (mx << 16) | mn;
}

// MaxGradientMagnitude over interior voxels: |(f(x-1)-f(x+1))/(2 norm), ...|
// with zero (or clamped) values outside the volume, as the brick overlap holds them.
template <typename T>
__global__ __launch_bounds__(256) void max_gradient(const T* __restrict__ lin, uint64_t X, uint64_t Y, uint64_t Z,
                                                    int clamp, double norm, double* __restrict__ out) {
  const uint64_t n = X * Y * Z;
  double best = 0.0;
  auto at = [&](int64_t x, int64_t y, int64_t z) -> double {
    if (clamp) {
      x = x < 0 ? 0 : (x >= (int64_t)X ? (int64_t)X - 1 : x);
      y = y < 0 ? 0 : (y >= (int64_t)Y ? (int64_t)Y - 1 : y);
      z = z < 0 ? 0 : (z >= (int64_t)Z ? (int64_t)Z - 1 : z);
    } else if (x < 0 || y < 0 || z < 0 || x >= (int64_t)X || y >= (int64_t)Y || z >= (int64_t)Z) {
      return 0.0;
    }
    return (double)lin[(uint64_t)x + X * ((uint64_t)y + Y * (uint64_t)z)];
  };
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const int64_t x = (int64_t)(i % X), r = (int64_t)(i / X), y = r % (int64_t)Y, z = r / (int64_t)Y;
    const double gx = (at(x - 1, y, z) - at(x + 1, y, z)) / (norm * 2);
    const double gy = (at(x, y - 1, z) - at(x, y + 1, z)) / (norm * 2);
    const double gz = (at(x, y, z - 1) - at(x, y, z + 1)) / (norm * 2);
    const double l = sqrt(gx * gx + gy * gy + gz * gz);
    best = l > best ? l : best;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(best, o, 64);
    best = t > best ? t : best;
  }
  __shared__ double sb[4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sb[w] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) best = sb[k] > best ? sb[k] : best;
    out[blockIdx.x] = best;
  }
}

// Histogram2DDataBlock::ComputeTemplate's first pass (Histogram2DDataBlock.cpp:349-395,
// ComputeGradient :326-346): over the core voxels of each brick (overlap ov
// <= x < size - ov), the gradient (f(x-1)-f(x+1))/(2 norm) per axis from the
// brick's own stored neighbours, and the largest magnitude.  grid: (gx, bricks);
// out[brick * gx + block] receives the block's maximum.
template <typename T>
__global__ __launch_bounds__(256) void brick_max_gradient(const T* __restrict__ pool,
                                                          const GatherBrick* __restrict__ bricks, uint32_t ov,
                                                          double norm, double* __restrict__ out) {
  const GatherBrick b = bricks[blockIdx.y];
  const int lx = lx_of<T>();
  const uint32_t nbx = nblocks(b.n[0], lx), nby = nblocks(b.n[1], 2);
  const uint32_t cx = b.n[0] - 2 * ov, cy = b.n[1] - 2 * ov, cz = b.n[2] - 2 * ov;
  const uint64_t n = (uint64_t)cx * cy * cz;
  const T* __restrict__ p = pool + b.offset;
  auto at = [&](uint32_t x, uint32_t y, uint32_t z) -> double { return (double)p[swz(x, y, z, nbx, nby, lx)]; };
  double best = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint32_t x = ov + (uint32_t)(i % cx);
    const uint64_t r = i / cx;
    const uint32_t y = ov + (uint32_t)(r % cy), z = ov + (uint32_t)(r / cy);
    const double gx = (at(x - 1, y, z) - at(x + 1, y, z)) / (norm * 2);
    const double gy = (at(x, y - 1, z) - at(x, y + 1, z)) / (norm * 2);
    const double gz = (at(x, y, z - 1) - at(x, y, z + 1)) / (norm * 2);
    const double l = sqrt(gx * gx + gy * gy + gz * gz);
    best = l > best ? l : best;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(best, o, 64);
    best = t > best ? t : best;
  }
  __shared__ double sb[4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sb[w] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) best = sb[k] > best ? sb[k] : best;
    out[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = best;
  }
}

// FillOverlap (ExtendedOctreeConverter.cpp:1203-1372) processes the bricks of a
// coarse level in z,y,x order and fills each brick's overlap with ten region
// copies from its neighbours, in a fixed order; the later neighbours are copied
// before their own overlap exists (all zero at that point without clamping).
// A brick only reads the final state of its left/top/front neighbours and the
// pre-fill state of the others, so the bricks with equal x+y+z can run at once:
// one launch per such wave reproduces the sequential result exactly.  The pool
// holds the closed-form bricks on entry; the cores never change.
// snap (may be null): the level's bricks before the fill, at the pool's offsets
// minus snap_base; the later neighbours are read there (any voxel), and a voxel
// no copy covers keeps its value.  Without it the pre-fill overlaps are taken as
// 0 (the even-core case: the closed-form bricks hold neighbour data there).
template <typename T>
__global__ __launch_bounds__(256) void fill_overlap_wave(T* __restrict__ pool, const GatherBrick* __restrict__ lod_bricks,
                                                         uint32_t bx, uint32_t by, uint32_t bz, uint32_t ov,
                                                         const uint32_t* __restrict__ wave,
                                                         const T* __restrict__ snap, uint64_t snap_base) {
  const uint32_t me = wave[blockIdx.y];
  const GatherBrick B = lod_bricks[me];
  const uint32_t x = B.c[0], y = B.c[1], z = B.c[2];
  const int tx = (int)B.n[0], ty = (int)B.n[1], tz = (int)B.n[2];
  const int o = (int)ov;
  const int lx = lx_of<T>();
  const uint32_t tbx = nblocks(B.n[0], lx), tby = nblocks(B.n[1], 2);
  const uint64_t sx_ = 2ull * o * ty * tz, sy_ = 2ull * o * tx * tz, sz_ = 2ull * o * tx * ty;
  const uint64_t total = sx_ + sy_ + sz_;
  const bool hasR = x + 1 < bx, hasB = y + 1 < by, hasK = z + 1 < bz, hasL = x > 0, hasT = y > 0, hasF = z > 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (uint64_t)gridDim.x * 256) {
    int vx, vy, vz;
    if (i < sx_) {             // x slabs: 2*o x ty x tz
      const uint64_t a = i % (2ull * o), r = i / (2ull * o);
      vx = a < (uint64_t)o ? (int)a : tx - 2 * o + (int)a;
      vy = (int)(r % ty); vz = (int)(r / ty);
    } else if (i < sx_ + sy_) {  // y slabs: tx x 2*o x tz
      const uint64_t j = i - sx_;
      vx = (int)(j % tx);
      const uint64_t r = j / tx, a = r % (2ull * o);
      vy = a < (uint64_t)o ? (int)a : ty - 2 * o + (int)a;
      vz = (int)(r / (2ull * o));
    } else {                   // z slabs: tx x ty x 2*o
      const uint64_t j = i - sx_ - sy_;
      vx = (int)(j % tx);
      const uint64_t r = j / tx;
      vy = (int)(r % ty);
      const uint64_t a = r / ty;
      vz = a < (uint64_t)o ? (int)a : tz - 2 * o + (int)a;
    }
    const bool xh = vx >= tx - o, yh = vy >= ty - o, zh = vz >= tz - o, xl = vx < o, yl = vy < o, zl = vz < o;
    // the ten copies in the reference's order; the last one covering the voxel wins
    int sel = -1;
    if (hasR && xh) sel = 0;
    if (hasB && yh) sel = 1;
    if (hasK && zh) sel = 2;
    if (hasL && xl) sel = 3;
    if (hasT && yl) sel = 4;
    if (hasF && zl) sel = 5;
    if (hasB && hasR && xh && yh) sel = 6;
    if (hasR && hasK && xh && zh) sel = 7;
    if (hasB && hasK && yh && zh) sel = 8;
    if (hasR && hasB && hasK && xh && yh && zh) sel = 9;
    if (sel < 0 && snap) continue;   // no copy: the pre-fill value stays
    T val = 0;
    if (sel >= 0) {
      const int dxs[10] = {1, 0, 0, -1, 0, 0, 1, 1, 0, 1};
      const int dys[10] = {0, 1, 0, 0, -1, 0, 1, 0, 1, 1};
      const int dzs[10] = {0, 0, 1, 0, 0, -1, 0, 1, 1, 1};
      const int dx = dxs[sel], dy = dys[sel], dz = dzs[sel];
      const uint32_t nidx = (x + dx) + bx * ((y + dy) + by * (z + dz));
      const GatherBrick N = lod_bricks[nidx];
      const int nx = (int)N.n[0], ny = (int)N.n[1], nz = (int)N.n[2];
      const int px = dx > 0 ? o + vx - (tx - o) : (dx < 0 ? nx - 2 * o + vx : vx);
      const int py = dy > 0 ? o + vy - (ty - o) : (dy < 0 ? ny - 2 * o + vy : vy);
      const int pz = dz > 0 ? o + vz - (tz - o) : (dz < 0 ? nz - 2 * o + vz : vz);
      const bool earlier = dx < 0 || dy < 0 || dz < 0;   // already filled: final state
      const bool core = px >= o && px < nx - o && py >= o && py < ny - o && pz >= o && pz < nz - o;
      const uint64_t at = N.offset + swz((uint32_t)px, (uint32_t)py, (uint32_t)pz, nblocks(N.n[0], lx), nblocks(N.n[1], 2), lx);
      if (earlier) val = pool[at];
      else if (snap) val = snap[at - snap_base];
      else if (core) val = pool[at];
    }
    const uint64_t dst = B.offset + swz((uint32_t)vx, (uint32_t)vy, (uint32_t)vz, tbx, tby, lx);
    pool[dst] = val;
    // padding voxels beyond the brick's last voxel repeat it
    const bool ex = vx == tx - 1 && (tx & ((1 << lx) - 1)), ey = vy == ty - 1 && (ty & 3), ez = vz == tz - 1 && (tz & 3);
    if (ex || ey || ez) {
      const int xe = ex ? (int)(tbx << lx) : vx + 1, ye = ey ? (int)(tby << 2) : vy + 1;
      const int ze = ez ? (int)(nblocks(B.n[2], 2) << 2) : vz + 1;
      for (int zz = vz; zz < ze; ++zz)
        for (int yy = vy; yy < ye; ++yy)
          for (int xx = vx; xx < xe; ++xx)
            pool[B.offset + swz((uint32_t)xx, (uint32_t)yy, (uint32_t)zz, tbx, tby, lx)] = val;
    }
  }
}

hipError_t launch_fill_overlap(void* pool, int bpv, const GatherBrick* d_lod_bricks, const uint32_t bricks[3],
                               uint32_t ov, const uint32_t* d_wave, uint32_t wave_n, uint32_t max_shell,
                               hipStream_t st, const void* snap, uint64_t snap_base) {
  uint32_t gx = (max_shell + 255) / 256;
  if (gx > 1024) gx = 1024;
  for (uint32_t first = 0; first < wave_n; first += 65535) {
    const uint32_t cnt = wave_n - first < 65535 ? wave_n - first : 65535;
    if (bpv == 1)
      hipLaunchKernelGGL(fill_overlap_wave<uint8_t>, dim3(gx, cnt), dim3(256), 0, st, (uint8_t*)pool, d_lod_bricks,
                         bricks[0], bricks[1], bricks[2], ov, d_wave + first, (const uint8_t*)snap, snap_base);
    else
      hipLaunchKernelGGL(fill_overlap_wave<uint16_t>, dim3(gx, cnt), dim3(256), 0, st, (uint16_t*)pool,
                         d_lod_bricks, bricks[0], bricks[1], bricks[2], ov, d_wave + first, (const uint16_t*)snap,
                         snap_base);
  }
  return hipGetLastError();
}

hipError_t launch_downsample_bricks(void* pool, int bpv, const GatherBrick* d_coarse, uint32_t nb,
                                    const GatherBrick* d_fine, const uint32_t fine_bricks[3], uint32_t ov,
                                    const uint32_t split[3], uint64_t max_elems, int median, hipStream_t st) {
  uint64_t gx = (max_elems + 255) / 256;
  if (gx > 2048) gx = 2048;
  if (gx == 0) gx = 1;
  for (uint32_t first = 0; first < nb; first += 65535) {
    const uint32_t cnt = nb - first < 65535 ? nb - first : 65535;
    if (bpv == 1)
      hipLaunchKernelGGL(downsample_bricks<uint8_t>, dim3((uint32_t)gx, cnt), dim3(256), 0, st, (uint8_t*)pool,
                         d_coarse + first, d_fine, fine_bricks[0], fine_bricks[1], fine_bricks[2], ov, split[0],
                         split[1], split[2], median);
    else
      hipLaunchKernelGGL(downsample_bricks<uint16_t>, dim3((uint32_t)gx, cnt), dim3(256), 0, st, (uint16_t*)pool,
                         d_coarse + first, d_fine, fine_bricks[0], fine_bricks[1], fine_bricks[2], ov, split[0],
                         split[1], split[2], median);
  }
  return hipGetLastError();
}

hipError_t launch_gather(const void* lin, int bpv, const uint64_t dims[3], const GatherBrick* d_bricks, uint32_t nb,
                         uint32_t row_len, uint32_t max_vox, void* pool, const uint32_t core[3], uint32_t ov, int clamp,
                         uint32_t* mm, hipStream_t st) {
  // max_vox: the largest padded element count; one work item per 16 x 16-byte
  // rows of GATHER_CB blocks.  Slots per brick: about 64 items per workgroup
  // (long-lived workgroups: 1.50 ms for the 1024^3 uint16 volume's 512 bricks
  // against 3.8 ms at 4 items), but at least 4096 workgroups in all while a
  // slot still has an item.
  const uint32_t item_vox = 16u * 2u * GATHER_CB * (bpv == 1 ? 8u : 4u);
  const uint32_t gx_items = max_vox / item_vox + 1u;
  uint32_t gx = max_vox / (16u * item_vox) + 1u;
  gx = std::max(gx, std::min(gx_items, (4096u + std::max(nb, 1u) - 1u) / std::max(nb, 1u)));
  if (gx > 4096) gx = 4096;
  if (row_len == 0) row_len = 1;
  // bricks per launch: at most 2^23 workgroups (the grid's work-item count
  // stays below 2^32), whole brick rows where a row fits
  uint32_t per = std::max(1u, (1u << 23) / gx);
  if (per >= row_len) per -= per % row_len;
  if (mm) hipLaunchKernelGGL(init_minmax, dim3((nb + 255) / 256), dim3(256), 0, st, mm, nb);
  for (uint32_t first = 0; first < nb; first += per) {
    const uint32_t cnt = std::min(nb - first, per), rl = std::min(row_len, cnt);
    const uint64_t wgs = (uint64_t)((cnt + rl - 1) / rl) * rl * gx;   // rows * row_len * slots
    const uint32_t per_xcd = (uint32_t)((wgs + 7) / 8);
    const dim3 grid(8u * per_xcd);
    if (bpv == 1)
      hipLaunchKernelGGL(gather_rows<uint8_t>, grid, dim3(256), 0, st, (const uint8_t*)lin, dims[0], dims[1],
                         dims[2], d_bricks + first, (uint8_t*)pool, core[0], core[1], core[2], ov, clamp,
                         mm ? mm + 2 * first : nullptr, cnt, gx, rl, per_xcd);
    else
      hipLaunchKernelGGL(gather_rows<uint16_t>, grid, dim3(256), 0, st, (const uint16_t*)lin, dims[0], dims[1],
                         dims[2], d_bricks + first, (uint16_t*)pool, core[0], core[1], core[2], ov, clamp,
                         mm ? mm + 2 * first : nullptr, cnt, gx, rl, per_xcd);
  }
  return hipGetLastError();
}

hipError_t launch_downsample(const void* src, const uint64_t sd[3], void* dst, const uint64_t dd[3], int bpv,
                             hipStream_t st, int median) {
  const uint64_t rows = dd[1] * dd[2];
  if (dd[0] == 0 || rows == 0) return hipSuccess;
  const uint64_t bx = (dd[0] + 255) / 256;
  // about 2^18 workgroups in all; each takes rows blockIdx.y, + gridDim.y, ...
  const uint64_t by = std::max<uint64_t>(1, std::min<uint64_t>({rows, 65535, ((uint64_t)1 << 18) / bx}));
  const dim3 grid((uint32_t)bx, (uint32_t)by);
  if (bpv == 1)
    hipLaunchKernelGGL(downsample<uint8_t>, grid, dim3(256), 0, st, (const uint8_t*)src, sd[0], sd[1], sd[2],
                       (uint8_t*)dst, dd[0], dd[1], dd[2], median);
  else
    hipLaunchKernelGGL(downsample<uint16_t>, grid, dim3(256), 0, st, (const uint16_t*)src, sd[0], sd[1], sd[2],
                       (uint16_t*)dst, dd[0], dd[1], dd[2], median);
  return hipGetLastError();
}

hipError_t launch_unswizzle(const void* pool, int bpv, const GatherBrick* d_bricks, const uint64_t* d_ref_off,
                            uint32_t nb, uint32_t max_vox, void* out, hipStream_t st) {
  uint32_t gx = (max_vox + 255) / 256;
  if (gx > 2048) gx = 2048;
  for (uint32_t first = 0; first < nb; first += 65535) {
    const uint32_t cnt = nb - first < 65535 ? nb - first : 65535;
    if (bpv == 1)
      hipLaunchKernelGGL(unswizzle_bricks<uint8_t>, dim3(gx, cnt), dim3(256), 0, st, (const uint8_t*)pool,
                         d_bricks + first, d_ref_off + first, (uint8_t*)out);
    else
      hipLaunchKernelGGL(unswizzle_bricks<uint16_t>, dim3(gx, cnt), dim3(256), 0, st, (const uint16_t*)pool,
                         d_bricks + first, d_ref_off + first, (uint16_t*)out);
  }
  return hipGetLastError();
}

hipError_t launch_swizzle(const void* src, int bpv, const GatherBrick* d_bricks, const uint64_t* d_src_off, uint32_t nb,
                          uint64_t max_elems, void* pool, hipStream_t st) {
  uint64_t gx = (max_elems + 255) / 256;
  if (gx > 2048) gx = 2048;
  if (gx == 0) gx = 1;
  for (uint32_t first = 0; first < nb; first += 65535) {
    const uint32_t cnt = nb - first < 65535 ? nb - first : 65535;
    if (bpv == 1)
      hipLaunchKernelGGL(swizzle_bricks<uint8_t>, dim3((uint32_t)gx, cnt), dim3(256), 0, st, (const uint8_t*)src,
                         d_bricks + first, d_src_off + first, (uint8_t*)pool);
    else
      hipLaunchKernelGGL(swizzle_bricks<uint16_t>, dim3((uint32_t)gx, cnt), dim3(256), 0, st, (const uint16_t*)src,
                         d_bricks + first, d_src_off + first, (uint16_t*)pool);
  }
  return hipGetLastError();
}

// (unsigned char) of (255 v - fMin) / (fMax - fMin), GLVolumeListElem::CreateTexture
// (GPUMemManDataStructs.cpp:410-425).  A conversion of an out-of-range double to
// unsigned char is left undefined by C++; x86 compilers truncate to a 32-bit
// integer (0x80000000 when it does not fit or is NaN) and keep the low byte,
// which is what this returns.
__device__ __forceinline__ uint8_t quantize8(uint32_t v, double fmin, double fmax) {
  const double q = (255.0 * (double)v - fmin) / (fmax - fmin);
  if (!(q > -2147483649.0 && q < 2147483648.0)) return 0;
  return (uint8_t)(uint32_t)(int32_t)q;
}

// The renderer's compatibility textures of one brick set: brick k of dst is
// brick k of src, quantized to 8 bits (quant) and/or padded to dst's size
// (GLVolumeListElem::PadData, GPUMemManDataStructs.cpp:306-360): voxel (x,y,z)
// of the padded texture is the source voxel while every coordinate is inside
// the source brick; with dup (the border in use) one more element, row and
// slice repeat the last ones; everything else is 0.
template <typename S, typename D>
__global__ __launch_bounds__(256) void compat_bricks(const S* __restrict__ src, const GatherBrick* __restrict__ sb,
                                                     D* __restrict__ dst, const GatherBrick* __restrict__ db,
                                                     double fmin, double fmax, int quant, int dup) {
  const GatherBrick s = sb[blockIdx.y], d = db[blockIdx.y];
  const int ls = lx_of<S>(), ld = lx_of<D>();
  const uint32_t snbx = nblocks(s.n[0], ls), snby = nblocks(s.n[1], 2);
  const uint32_t dnbx = nblocks(d.n[0], ld), dnby = nblocks(d.n[1], 2);
  const uint64_t n = padded_count(d.n, ld);
  const uint32_t lim0 = s.n[0] - (dup ? 0u : 1u), lim1 = s.n[1] - (dup ? 0u : 1u), lim2 = s.n[2] - (dup ? 0u : 1u);
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    uint32_t x, y, z;
    unswz(i, dnbx, dnby, ld, x, y, z);
    x = min(x, d.n[0] - 1); y = min(y, d.n[1] - 1); z = min(z, d.n[2] - 1);
    D v = 0;
    if (x <= lim0 && y <= lim1 && z <= lim2) {
      const uint32_t u = (uint32_t)src[s.offset + swz(min(x, s.n[0] - 1), min(y, s.n[1] - 1), min(z, s.n[2] - 1),
                                                      snbx, snby, ls)];
      v = quant ? (D)quantize8(u, fmin, fmax) : (D)u;
    }
    dst[d.offset + i] = v;
  }
}

hipError_t launch_compat(const void* src, int src_bpv, const GatherBrick* d_src, void* dst, int dst_bpv,
                         const GatherBrick* d_dst, uint32_t nb, uint64_t max_elems, double fmin, double fmax,
                         int dup, hipStream_t st) {
  uint64_t gx = (max_elems + 255) / 256;
  if (gx > 2048) gx = 2048;
  if (gx == 0) gx = 1;
  for (uint32_t first = 0; first < nb; first += 65535) {
    const uint32_t cnt = nb - first < 65535 ? nb - first : 65535;
    const dim3 g((uint32_t)gx, cnt);
    if (src_bpv == 2 && dst_bpv == 1)
      hipLaunchKernelGGL((compat_bricks<uint16_t, uint8_t>), g, dim3(256), 0, st, (const uint16_t*)src, d_src + first,
                         (uint8_t*)dst, d_dst + first, fmin, fmax, 1, dup);
    else if (src_bpv == 2)
      hipLaunchKernelGGL((compat_bricks<uint16_t, uint16_t>), g, dim3(256), 0, st, (const uint16_t*)src,
                         d_src + first, (uint16_t*)dst, d_dst + first, fmin, fmax, 0, dup);
    else
      hipLaunchKernelGGL((compat_bricks<uint8_t, uint8_t>), g, dim3(256), 0, st, (const uint8_t*)src, d_src + first,
                         (uint8_t*)dst, d_dst + first, fmin, fmax, 0, dup);
  }
  return hipGetLastError();
}

hipError_t launch_flatten(const void* pool, int bpv, const GatherBrick* d_bricks, const uint32_t nb[3],
                          const uint32_t core[3], uint32_t ov, const uint64_t dims[3], void* out, hipStream_t st) {
  const uint64_t n = dims[0] * dims[1] * dims[2];
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (bpv == 1)
    hipLaunchKernelGGL(flatten_lod0<uint8_t>, dim3((uint32_t)blocks), dim3(256), 0, st, (const uint8_t*)pool, d_bricks,
                       nb[0], nb[1], core[0], core[1], core[2], ov, dims[0], dims[1], dims[2], (uint8_t*)out);
  else
    hipLaunchKernelGGL(flatten_lod0<uint16_t>, dim3((uint32_t)blocks), dim3(256), 0, st, (const uint16_t*)pool,
                       d_bricks, nb[0], nb[1], core[0], core[1], core[2], ov, dims[0], dims[1], dims[2],
                       (uint16_t*)out);
  return hipGetLastError();
}

hipError_t launch_crop(void* vol, int bpv, const uint64_t dims[3], const float pl[4], hipStream_t st) {
  const uint64_t n = dims[0] * dims[1] * dims[2];
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (bpv == 1)
    hipLaunchKernelGGL(crop_plane<uint8_t>, dim3((uint32_t)blocks), dim3(256), 0, st, (uint8_t*)vol, dims[0], dims[1],
                       dims[2], pl[0], pl[1], pl[2], pl[3]);
  else
    hipLaunchKernelGGL(crop_plane<uint16_t>, dim3((uint32_t)blocks), dim3(256), 0, st, (uint16_t*)vol, dims[0],
                       dims[1], dims[2], pl[0], pl[1], pl[2], pl[3]);
  return hipGetLastError();
}

hipError_t launch_minmax(const void* pool, int bpv, const GatherBrick* d_bricks, uint32_t nb, uint32_t max_vox,
                         uint32_t* out, hipStream_t st) {
  // about 16 rows per lane (4 passes of 4), at most 64 workgroups per brick
  uint32_t gx = (max_vox / (bpv == 1 ? 16 : 8) + 4095) / 4096;
  if (gx > 64) gx = 64;
  if (gx == 0) gx = 1;
  hipLaunchKernelGGL(init_minmax, dim3((nb + 255) / 256), dim3(256), 0, st, out, nb);
  for (uint32_t first = 0; first < nb; first += 65535u) {
    const uint32_t cnt = std::min(nb - first, 65535u);
    if (bpv == 1)
      hipLaunchKernelGGL(brick_minmax<uint8_t>, dim3(gx, cnt), dim3(256), 0, st, (const uint8_t*)pool,
                         d_bricks + first, out + 2 * first);
    else
      hipLaunchKernelGGL(brick_minmax<uint16_t>, dim3(gx, cnt), dim3(256), 0, st, (const uint16_t*)pool,
                         d_bricks + first, out + 2 * first);
  }
  return hipGetLastError();
}

hipError_t launch_extend(void* pool, int bpv, const GatherBrick* d_bricks, uint32_t nb, uint32_t max_vox,
                         hipStream_t st) {
  uint32_t gx = (uint32_t)((max_vox / 2 + 255) / 256);
  if (gx > 1024) gx = 1024;
  if (gx == 0) gx = 1;
  for (uint32_t first = 0; first < nb; first += 65535u) {
    const uint32_t cnt = std::min(nb - first, 65535u);
    if (bpv == 1)
      hipLaunchKernelGGL(extend_rows<uint8_t>, dim3(gx, cnt), dim3(256), 0, st, (uint8_t*)pool, d_bricks + first);
    else
      hipLaunchKernelGGL(extend_rows<uint16_t>, dim3(gx, cnt), dim3(256), 0, st, (uint16_t*)pool, d_bricks + first);
  }
  return hipGetLastError();
}

hipError_t launch_cells(const void* pool, int bpv, const GatherBrick* d_bricks, const uint32_t* d_base, uint32_t nb,
                        uint32_t max_cells, uint32_t* out, hipStream_t st) {
  const uint32_t gx = (max_cells + 3) / 4;
  for (uint32_t first = 0; first < nb; first += 65535u) {
    const uint32_t cnt = std::min(nb - first, 65535u);
    if (bpv == 1)
      hipLaunchKernelGGL(cell_minmax<uint8_t>, dim3(gx, cnt), dim3(256), 0, st, (const uint8_t*)pool, d_bricks + first,
                         d_base + first, out);
    else
      hipLaunchKernelGGL(cell_minmax<uint16_t>, dim3(gx, cnt), dim3(256), 0, st, (const uint16_t*)pool,
                         d_bricks + first, d_base + first, out);
  }
  return hipGetLastError();
}

hipError_t launch_brick_max_gradient(const void* pool, int bpv, const GatherBrick* d_bricks, uint32_t nb,
                                     uint32_t ov, uint32_t gx, double* out, hipStream_t st) {
  for (uint32_t first = 0; first < nb; first += 65535u) {
    const uint32_t cnt = std::min(nb - first, 65535u);
    if (bpv == 1)
      hipLaunchKernelGGL(brick_max_gradient<uint8_t>, dim3(gx, cnt), dim3(256), 0, st, (const uint8_t*)pool,
                         d_bricks + first, ov, 255.0, out + (size_t)first * gx);
    else
      hipLaunchKernelGGL(brick_max_gradient<uint16_t>, dim3(gx, cnt), dim3(256), 0, st, (const uint16_t*)pool,
                         d_bricks + first, ov, 65535.0, out + (size_t)first * gx);
  }
  return hipGetLastError();
}

hipError_t launch_max_gradient(const void* lin, int bpv, const uint64_t d[3], int clamp, double* out, uint32_t blocks,
                               hipStream_t st) {
  if (bpv == 1)
    hipLaunchKernelGGL(max_gradient<uint8_t>, dim3(blocks), dim3(256), 0, st, (const uint8_t*)lin, d[0], d[1], d[2],
                       clamp, 255.0, out);
  else
    hipLaunchKernelGGL(max_gradient<uint16_t>, dim3(blocks), dim3(256), 0, st, (const uint16_t*)lin, d[0], d[1], d[2],
                       clamp, 65535.0, out);
  return hipGetLastError();
}

}  // namespace tvk
