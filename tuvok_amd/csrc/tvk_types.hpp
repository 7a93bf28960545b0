// Structures shared by the host planner and the HIP kernels.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>
#include "kernels/stepnet.hpp"

namespace tvk {

// In HBM a brick is stored as 256-byte blocks of 8x4x4 (uint8) or 4x4x4
// (uint16) voxels, blocks in x-fastest order, each axis padded to whole blocks
// with copies of the last voxel.  Every x-row of a block holds its own voxels
// followed by the first voxels of the next block's row (the last block of a
// row repeats the brick's last voxel): any 4 consecutive voxels in x then lie
// in one 16-byte row, so the march reads a gradient cross with one load per
// row.  A ray bundle marching through a brick touches compact cache lines
// instead of one line per (y,z) row.  The reference layout (x-fastest,
// unpadded) is what the download calls return.
#ifndef TVK_HD
#define TVK_HD __host__ __device__
#endif
TVK_HD inline int block_lx(int bpv) { return bpv == 1 ? 3 : 2; }   // log2 of the block width in x
TVK_HD inline uint32_t nblocks(uint32_t n, int l) { return (n + (1u << l) - 1u) >> l; }
// storage index of voxel (x,y,z): rows of 2 << lx elements, 4 rows per z slice
TVK_HD inline uint64_t swz(uint32_t x, uint32_t y, uint32_t z, uint32_t nbx, uint32_t nby, int lx) {
  return ((((uint64_t)(z >> 2) * nby + (y >> 2)) * nbx + (x >> lx)) << (lx + 5)) + ((uint64_t)(z & 3u) << (lx + 3)) +
         ((uint64_t)(y & 3u) << (lx + 1)) + (x & ((1u << lx) - 1u));
}
TVK_HD inline uint64_t padded_count(const uint32_t n[3], int lx) {
  return (uint64_t)nblocks(n[0], lx) * nblocks(n[1], 2) * nblocks(n[2], 2) << (lx + 5);
}

// One entry of a frame's visible, non-empty brick list, in compositing order.
// Everything the march needs for the brick is here, so a wavefront reads it
// with a few scalar loads.  128 bytes.
struct FrameBrick {
  float e2t[12];        // eye -> texture affine, rows 0..3 of the row-vector 4x4 (xyz columns)
  float bmin[3];        // model-space box of the brick (GLRaycaster::RenderBox)
  float bmax[3];
  float ray_step;       // fRayStepsize
  float step_scale;     // fStepScale (opacity-correction exponent)
  float vstep[3];       // vVoxelStepsize (1 / stored voxel count)
  uint32_t n[3];        // stored voxel counts (incl. overlap)
  uint64_t offset;      // element offset of the brick in the pool
  int16_t rect[4];      // conservative pixel rectangle x0,y0,x1,y1 (inclusive)
  uint32_t toc;         // TOC index of the brick
  uint32_t cells;       // first entry of the brick's cell grid (learned mode)
  // transfer-function statistics over the brick's value range (step proposal
  // features): vmin, vmax (TF coordinates), alpha mean, alpha max, alpha total
  // variation, colour total variation, largest alpha step, fraction of the
  // range with non-zero alpha
  float tfs[8];
};
static_assert(sizeof(FrameBrick) == 160, "FrameBrick layout");

// Per-view constants.
struct ViewConst {
  float inv_proj[16];   // inverse projection (row-vector convention)
  float inv_mv[16];     // inverse modelview
  float nm[9];          // gl_NormalMatrix, applied to column vectors
  float znear, zfar;
  uint32_t list_offset; // first FrameBrick of this view
  uint32_t nlist;       // number of FrameBricks of this view
  uint32_t W, H;
  uint32_t out_offset;  // first pixel of this view in the output buffer
  uint32_t tiles_x;     // 8x8 tiles per row
  uint32_t groups_x;    // 16x16 block groups per row
  uint32_t groups;      // 16x16 block groups of the view
  uint32_t group_offset;// first 16x16 group of this view in the launch
  uint32_t clip_on;     // the clip plane cuts every brick's ray segment
  uint32_t pad;
  float clip[4];        // model-space clip plane (n, d): the ray keeps n.p + d <= 0
};

// One brick of a LOD for the bricking / min-max kernels.
struct GatherBrick {
  uint32_t c[3];        // brick coordinates
  uint32_t n[3];        // stored size
  uint64_t offset;      // element offset in the pool
};

// Shading constants of one render request (a dataset, its TF and parameters).
struct ShadeConst {
  const uint8_t* pool;          // brick pool of the dataset
  const float4* tf1;            // 1D TF as float4 (byte / 255), tf_n entries
  const float4* tf2f;           // 2D TF as float4 (byte / 255), tf_w x tf_h
  const uint64_t* tf2_cols;     // bit x: value column x of the 2D TF has a non-zero alpha (64 words)
  int tf_n, tf_w, tf_h;
  float tf_scale, grad_scale, iso_n, ert_eps;
  float dscale[3], amb[3], dif[3], spc[3], ldir[3], iso_col[3];
  int refine_steps, clamp_edge;
  // learned mode: per-entry TF statistics (layout of tf_stats_layout) and the
  // dataset's cell grid (cell_grid_dims)
  const float* tfst;
  const uint32_t* cells;
  int tfst_n, tfst_levels;
  int overlap;                  // brick overlap of the dataset
};

// Cell grid of a brick: cells of CELL_W^3 stored voxels; the (min, max) of a
// cell (low and high 16 bits) span the stored voxels within CELL_DIL of it, so
// a probe's cell bounds every value interpolated within 2 voxels of the probe.
constexpr int CELL_LW = 3, CELL_W = 1 << CELL_LW, CELL_DIL = 4;
TVK_HD inline uint32_t cell_count(uint32_t n) { return (n + CELL_W - 1) >> CELL_LW; }

// Flat TF statistics over n entries (floats): prefix sums of alpha [n+1], of
// (alpha > 0) [n+1], of |d alpha| [n], of mean |d rgb| [n]; then `levels`
// rows of n of the sparse max table of alpha and `levels` rows of the sparse
// max table of |d alpha|.
TVK_HD inline int tfst_pa(int) { return 0; }
TVK_HD inline int tfst_pnz(int n) { return n + 1; }
TVK_HD inline int tfst_ptv(int n) { return 2 * n + 2; }
TVK_HD inline int tfst_prgb(int n) { return 3 * n + 2; }
TVK_HD inline int tfst_amax(int n, int l) { return 4 * n + 2 + l * n; }
TVK_HD inline int tfst_smax(int n, int levels, int l) { return 4 * n + 2 + (levels + l) * n; }
TVK_HD inline size_t tfst_size(int n, int levels) { return (size_t)4 * n + 2 + (size_t)2 * levels * n; }

// Step proposal (learned mode): network and decision thresholds.
struct LearnConst {
  StepNetDev net;
  float margin;        // subtracted from 3*y0 before rounding to a log2 multiplier
  float skip_logit;    // skip a segment when y1 exceeds this logit
  int max_log2;        // largest log2 multiplier (multiplier 1..8)
  float fallback;      // fixed step when the chunk's contribution bound exceeds this
  // fixed step when the chunk's variation risk exceeds this: T_before * (bound *
  // colour total variation + min(1, steps * step scale * alpha total variation))
  // over the TF entries of the chunk's value range
  float fallback_tv;
  float skip_bound;    // skip a chunk whose contribution bound is at most this (< 0: never)
  int fixed_k = -1;    // >= 0: no network, every marched chunk takes multiplier 2^fixed_k
};

// The learned march decides per chunk of SN_CHUNK fixed steps; the cell grid
// is probed every SN_PROBE steps of a chunk.
#ifndef TVK_SN_CHUNK
#define TVK_SN_CHUNK 32
#endif
constexpr int SN_CHUNK = TVK_SN_CHUNK, SN_PROBE = 8;

// One training sample of the step proposal: a chunk of a ray's segment
// through one brick (at most SN_CHUNK fixed steps).
struct SegSample {
  float f[24];         // features (the network input)
  float t_before;      // transmittance of the ray before the chunk
  float steps;         // fixed-step sample count of the chunk
  float c[4][4];       // chunk RGBA (premultiplied) at multipliers 1, 2, 4, 8
  float pad[6];
};
static_assert(sizeof(SegSample) == 192, "SegSample layout");

struct SegOut {
  SegSample* out;
  unsigned int* count;
  uint32_t cap;
  uint32_t keep;       // a segment is recorded when hash < keep (of 2^32)
  uint32_t seed;
};

struct MarchArgs {
  ShadeConst S;
  LearnConst learn;
  SegOut seg;
  const FrameBrick* list;
  const ViewConst* views;
  float* out;                   // RGBA32F
  unsigned long long* counter;  // composited-sample counter or nullptr
  uint32_t* touched;            // [view * ntoc + toc] = 1 once a sample is taken in the brick (or nullptr)
  uint32_t ntoc;
  int nviews;
  int lighting, mode;
  uint32_t groups_total;        // sum of groups over the views
  uint32_t groups_padded;       // the grid: a multiple of 8 (XCD count)
  // Dispatch order (or nullptr: XCD x runs groups x * groups_padded / 8 + j in
  // turn).  With an order, XCD x runs order[xcd_start[x] + j] for j <
  // xcd_count[x] and groups_padded is 8 x the largest count.
  const uint32_t* order;
  uint32_t xcd_start[8], xcd_count[8];
  // Work queues (or nullptr): 8 counters cleared before the launch; the grid
  // is then persistent and every wave pulls 8x8 tiles, item t of XCD x being
  // quarter t & 3 of the range's group t >> 2 (xcd_start / xcd_count as above,
  // in the order when there is one).
  unsigned int* queues;
  int accumulate;               // continue the pixels' values (and depths) of an earlier launch
  float* depth;                 // isosurface: per-pixel depth of the best hit so far (or nullptr)
  float4* hitpos;               // isosurface: per-pixel eye-space hit and its segment parameter (or nullptr)
};

// One request of a batch: its shading constants, view, and place in the batch.
struct BatchRequest {
  ShadeConst S;
  ViewConst V;                  // V.list_offset / V.out_offset index the batch buffers
  uint32_t tile_offset;         // first tile of the request in the batch tile space
  uint32_t ntiles;              // tiles_x * tiles_y
  uint32_t tiles_x;             // tiles per row (tile_w wide)
  uint32_t touched_offset;      // first entry of the request's touched-brick row
};

// Persistent batch launch: waves pull tiles from 8 per-XCD queues (the tile
// space is cut into 8 contiguous chunks; a wave drains its own XCD's chunk, then
// the others in turn) and exit once every chunk is exhausted.
struct BatchArgs {
  LearnConst learn;
  const BatchRequest* req;
  const FrameBrick* list;
  float* out;
  unsigned int* queues;         // 8 counters, cleared by a fill kernel before the launch
  unsigned long long* counter;  // composited-sample counter or nullptr
  uint32_t* touched;            // per request: touched_offset + toc (or nullptr)
  int nreq;
  uint32_t tiles_total;
  uint32_t tile_w, tile_h;      // wave tile shape, tile_w * tile_h == 64
};

// GridLeaper single-pass traversal (one pass of the ray-guided march).
constexpr int LEAP_MAX_LOD = 16;
struct LeapArgs {
  ShadeConst S;                // pool, TF tables, scales, lights
  const uint32_t* meta;        // per brick ID: 0 missing, 1 child empty, 2 empty, 3 + TOC index resident
  const GatherBrick* bricks;   // per TOC brick: stored size and pool offset
  uint32_t* table;             // GLHashTable of the requested bricks: serialised 4D brick index + 1, 0 = free
  uint32_t table_size, rehash; // entries; probes per request (GLHashTable.cpp:140-181)
  float4* resume;              // 2 per pixel: (normalised position, eye z; w = 1000 finished), colour
  float* out;                  // RGBA32F accumulated colour
  unsigned long long* counter; // samples (or nullptr)
  float emm[16];               // eye -> normalised volume coordinates
  float mte[16];               // normalised volume coordinates -> eye (isosurface hit)
  float nit[9];                // inverse transpose of the modelview, row vectors (isosurface normal)
  float inv_proj[16];
  float layout[LEAP_MAX_LOD][3];     // float brick layout per level
  uint32_t ilayout[LEAP_MAX_LOD][3]; // integer brick layout per level
  uint32_t lodoff[LEAP_MAX_LOD];     // first brick ID of each level
  float inner[3];              // inner (core) brick size
  float vsz[3];                // LOD-0 volume size in voxels
  float meye[3], mlight[3];    // model-space eye and light direction
  float lod_factor, l0err, srm;
  int nlod, W, H, first, lighting, mode, overlap;
  uint32_t tiles_x;
  // work queue (or nullptr: one 8x8 tile per wave of the grid): a counter
  // cleared before the launch; the grid is then persistent and every wave
  // pulls tiles 0 .. ntiles-1 in order until they are drained
  unsigned int* queue;
  uint32_t ntiles;
};
}  // namespace tvk
