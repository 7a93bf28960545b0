// Internal state of a tvk_ctx and helpers shared by the C ABI sources.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <string>
#include <vector>
#include "tvk_hip.h"
#include "host/host.hpp"
#include "kernels/generate.hpp"
#include "tvk_types.hpp"

namespace tvk {
hipError_t launch_march(const MarchArgs& a, int bpv, hipStream_t st, int learn);
hipError_t launch_batch(const BatchArgs& b, int mode, bool lit, int bpv, int groups, hipStream_t st, int learn);
hipError_t launch_gather(const void* lin, int bpv, const uint64_t dims[3], const GatherBrick* d_bricks, uint32_t nb,
                         uint32_t row_len, uint32_t max_vox, void* pool, const uint32_t core[3], uint32_t ov, int clamp,
                         uint32_t* mm, hipStream_t st);
hipError_t launch_downsample(const void* src, const uint64_t sd[3], void* dst, const uint64_t dd[3], int bpv,
                             hipStream_t st, int median);
hipError_t launch_downsample_bricks(void* pool, int bpv, const GatherBrick* d_coarse, uint32_t nb,
                                    const GatherBrick* d_fine, const uint32_t fine_bricks[3], uint32_t ov,
                                    const uint32_t split[3], uint64_t max_elems, int median, hipStream_t st);
hipError_t launch_minmax(const void* pool, int bpv, const GatherBrick* d_bricks, uint32_t nb, uint32_t max_vox,
                         uint32_t* out, hipStream_t st);
hipError_t launch_unswizzle(const void* pool, int bpv, const GatherBrick* d_bricks, const uint64_t* d_ref_off,
                            uint32_t nb, uint32_t max_vox, void* out, hipStream_t st);
hipError_t launch_max_gradient(const void* lin, int bpv, const uint64_t d[3], int clamp, double* out, uint32_t blocks,
                               hipStream_t st);
hipError_t launch_fill_overlap(void* pool, int bpv, const GatherBrick* d_lod_bricks, const uint32_t bricks[3],
                               uint32_t ov, const uint32_t* d_wave, uint32_t wave_n, uint32_t max_shell,
                               hipStream_t st, const void* snap = nullptr, uint64_t snap_base = 0);
}  // namespace tvk


namespace tvk {
hipError_t launch_extend(void* pool, int bpv, const GatherBrick* d_bricks, uint32_t nb, uint32_t max_vox,
                         hipStream_t st);
hipError_t launch_display(const float* frames, uint32_t W, uint32_t H, uint32_t nviews, const float top[3],
                          const float bottom[3], uint8_t* out, hipStream_t st);
hipError_t launch_cells(const void* pool, int bpv, const GatherBrick* d_bricks, const uint32_t* d_base, uint32_t nb,
                        uint32_t max_cells, uint32_t* out, hipStream_t st);
hipError_t launch_brick_max_gradient(const void* pool, int bpv, const GatherBrick* d_bricks, uint32_t nb,
                                     uint32_t ov, uint32_t gx, double* out, hipStream_t st);
hipError_t launch_flatten(const void* pool, int bpv, const GatherBrick* d_bricks, const uint32_t nb[3],
                          const uint32_t core[3], uint32_t ov, const uint64_t dims[3], void* out, hipStream_t st);
hipError_t launch_crop(void* vol, int bpv, const uint64_t dims[3], const float pl[4], hipStream_t st);
hipError_t launch_compat(const void* src, int src_bpv, const GatherBrick* d_src, void* dst, int dst_bpv,
                         const GatherBrick* d_dst, uint32_t nb, uint64_t max_elems, double fmin, double fmax,
                         int dup, hipStream_t st);
hipError_t launch_swizzle(const void* src, int bpv, const GatherBrick* d_bricks, const uint64_t* d_src_off, uint32_t nb,
                          uint64_t max_elems, void* pool, hipStream_t st);

// A dataset held within a device-memory budget (GPUMemMan's budget and
// texture reuse, GLVolumePool's slots): every brick lives in pinned host
// memory in the reference layout, and the HBM pool holds nslots slots of
// slot_elems elements (the largest brick's blocked layout), paged on demand.
// Shared by the contexts that share the dataset; all calls that page must
// come from one thread.
struct PoolState {
  int bpv = 1;
  uint64_t slot_elems = 0;
  uint32_t nslots = 0;
  std::vector<int64_t> slot_brick;     // per slot: TOC brick or -1
  std::vector<uint64_t> slot_time;     // per slot: last use (raycaster) / creation time (traversal)
  std::vector<int32_t> brick_slot;     // per TOC brick: slot or -1
  uint64_t clock = 2;                  // 0: never used, 1: flagged empty (GLVolumePool)
  // host store: bricks in TOC order, reference layout, in chunks of pinned
  // memory while the host grants it, pageable memory beyond
  std::vector<void*> chunks;
  std::vector<uint8_t> chunk_pinned;   // per chunk
  bool pinned_refused = false;
  std::vector<uint8_t*> brick_host;    // per TOC brick
  std::vector<uint8_t> present;        // per TOC brick: payload handed over
  // device staging for the uploads
  uint8_t* d_stage = nullptr;
  size_t stage_bytes = 0;
  GatherBrick* d_desc = nullptr;
  uint64_t* d_srcoff = nullptr;
  size_t desc_cap = 0;
  // statistics
  uint64_t uploads = 0, evictions = 0, upload_bytes = 0;
  double upload_ms = 0.0;
  ~PoolState();
};

// Cell grid of a dataset (learned mode): per brick, the (min, max) of its cells
// (tvk_types.hpp), built on first use and shared by the contexts that share
// the dataset.
struct CellGrid {
  uint32_t* d = nullptr;
  std::vector<uint32_t> base;   // first cell of each TOC brick
  ~CellGrid() { if (d) (void)hipFree(d); }
};
}  // namespace tvk

// the opaque handle of the C ABI
struct tvk_ctx {
  using Layout = tvk::Layout;
  using ViewState = tvk::ViewState;
  using FrameBrick = tvk::FrameBrick;
  using ViewConst = tvk::ViewConst;
  using LearnConst = tvk::LearnConst;

  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // dataset
  bool has_ds = false;
  Layout L;
  int clamp = 0;
  void* d_pool = nullptr;
  std::shared_ptr<void> pool_hold;   // owns d_pool; shared by contexts that share the dataset
  std::shared_ptr<void> lin_hold;    // owns d_lin
  std::shared_ptr<tvk::CellGrid> cells;   // created with the pool, filled on first learned use
  std::shared_ptr<tvk::PoolState> pst;    // null: every brick resident at its TOC offset
  // bricks declared by tvk_set_bricks (resident pool) whose payload has not
  // arrived yet: rendering waits for all of them
  std::vector<uint8_t> ingest_missing;
  uint64_t ingest_pending = 0;
  size_t pool_bytes = 0;
  void* d_lin = nullptr;
  size_t lin_bytes = 0;
  std::vector<double> minmax;
  double range_min = 0, range_max = 0;
  // transfer functions
  std::vector<uint8_t> tf1d, tf2d;
  int tf1d_n = 0, tf2d_w = 0, tf2d_h = 0;
  uint8_t* d_tf1d = nullptr;
  uint8_t* d_tf2d = nullptr;
  float4* d_tf1f = nullptr;    // tables as the sampler returns them (byte / 255)
  float4* d_tf2f = nullptr;
  float build_ms = 0.0f;         // device time of the last bricking (build_pool)
  uint64_t* d_tf2cols = nullptr;  // 64 words: value columns of the 2D TF with a non-zero alpha
  size_t tf1d_cap = 0, tf2d_cap = 0, tf1f_cap = 0, tf2f_cap = 0;
  // view + parameters
  ViewState V;
  bool has_view = false;
  bool proj_from_fov = false;   // projection made by tvk_set_camera (re-made per request aspect)
  tvk_params P{};
  // frame resources
  FrameBrick* d_list = nullptr;
  size_t list_cap = 0;
  ViewConst* d_views = nullptr;
  size_t views_cap = 0;
  uint32_t* d_order = nullptr;   // dispatch order of the 16x16 groups
  size_t order_cap = 0;
  unsigned int* d_queues = nullptr;   // the persistent launch's 8 XCD work queues
  size_t queues_cap = 0;
  float* d_out = nullptr;
  size_t out_cap = 0;
  unsigned long long* d_counter = nullptr;
  bool count = false;
  // step proposal (learned mode): the network and its decision thresholds
  struct tvk_stepnet* net = nullptr;
  tvk::LearnConst learn{};
  bool learn_on = false;   // chunked march (network, or a fixed multiplier without one)
  // per-entry statistics of the active TF (brick features of the step proposal)
  struct TfStats {
    int n = 0;
    std::vector<double> pa, ptv, prgb, pnz;   // prefix sums of alpha, |dalpha|, |drgb|, alpha > 0
    std::vector<std::vector<float>> amax, smax;   // sparse tables: max alpha, max |dalpha|
    int levels = 0;            // rows of the device tables
    std::vector<float> flat;   // the device layout (tvk_types.hpp), kept while the upload runs
  } tfst;
  float* d_tfst = nullptr;
  size_t tfst_cap = 0;
  bool tfst_dirty = true;
  // ray-guided traversal (GridLeaper): residency, metadata, miss table, resume state
  std::vector<uint8_t> leap_resident;   // per TOC brick (empty: all resident)
  uint32_t* d_leap_meta = nullptr;
  size_t leap_meta_cap = 0;
  uint32_t* d_leap_table = nullptr;    // GLHashTable of requested bricks
  size_t leap_table_cap = 0;
  uint32_t leap_table_size = 509, leap_rehash = 10;   // RendererState defaults (MasterController.h:70-74)
  uint32_t leap_table_used = 0;       // entries of the last pass's table
  tvk::GatherBrick* d_leap_bricks = nullptr;
  size_t leap_bricks_cap = 0;
  float4* d_leap_resume = nullptr;
  size_t leap_resume_cap = 0;
  size_t leap_pixels = 0;
  tvk::F3 leap_eye = tvk::F3(0.0f, 0.0f, 1.6f);   // m_vEye
  uint32_t* d_touched = nullptr;
  size_t touched_cap = 0, touched_n = 0;
  float* d_depth = nullptr;          // isosurface depth of a frame rendered in parts
  size_t depth_cap = 0;
  float4* d_hitpos = nullptr;        // isosurface hit positions of the last frame (Pick)
  size_t hitpos_cap = 0;
  bool hitpos_valid = false;
  uint32_t last_parts = 0;           // launches of the last budgeted frame
  hipEvent_t ev[8] = {};
  std::vector<FrameBrick> h_list;
  std::vector<ViewConst> h_views;
  FrameBrick* h_list_pinned = nullptr;
  size_t h_list_pinned_cap = 0;
  ViewConst* h_views_pinned = nullptr;
  size_t h_views_pinned_cap = 0;
  std::vector<float> h_cost;
  std::vector<uint32_t> h_order;
  uint32_t* h_order_pinned = nullptr;
  size_t h_order_pinned_cap = 0;
  int last_views = 0;
  uint32_t last_w = 0, last_h = 0;
  int last_lod = 0;
  // the renderer's compatibility textures (AbstrRenderer's bUseOnlyPowerOfTwo
  // and bDownSampleTo8Bits; bDisableBorder is the clamp_to_edge parameter): a
  // second pool of padded and/or quantized bricks that rendering reads, made
  // on first use after the dataset or the flags change
  int compat_pow2 = 0, compat_8bit = 0;
  int dispatch = TVK_DISPATCH_RASTER | TVK_DISPATCH_QUEUES;   // of multi-view renders
  int lod_median = 0;   // ExtendedOctreeConverter's bComputeMedian for the next bricking
  struct CompatPool {
    void* d = nullptr;
    std::shared_ptr<void> hold;
    std::vector<uint64_t> offset;   // per TOC brick: element offset
    std::vector<uint32_t> dims;     // per TOC brick: texture size (3 each)
    int bpv = 0;
    int pow2 = 0, down8 = 0, dup = 0;
    bool valid = false;
  } cp;
  // device address of this GPU's copy of the allocation-free render's scratch
  // array (looked up once per device at tvk_create); null if the lookup failed
  void* scratch = nullptr;
  std::string scratch_err;
};


namespace tvk {

inline int fail(tvk_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
inline int hipfail(tvk_ctx* c, const char* what, hipError_t e) {
  return fail(c, e == hipErrorOutOfMemory ? TVK_ERR_NOMEM : TVK_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(c, x)                                  \
  do {                                                \
    hipError_t e_ = (x);                              \
    if (e_ != hipSuccess) return hipfail(c, #x, e_);  \
  } while (0)

template <typename T>
int ensure(tvk_ctx* c, T** p, size_t* cap, size_t n) {
  if (*cap >= n && *p) return TVK_OK;
  if (*p) { (void)hipFree(*p); *p = nullptr; *cap = 0; }
  size_t want = std::max<size_t>(n, 1);
  HIPCHK(c, hipMalloc((void**)p, want * sizeof(T)));
  *cap = want;
  return TVK_OK;
}

template <typename T>
int ensure_pinned(tvk_ctx* c, T** p, size_t* cap, size_t n) {
  if (*cap >= n && *p) return TVK_OK;
  if (*p) { (void)hipHostFree(*p); *p = nullptr; *cap = 0; }
  size_t want = std::max<size_t>(n * 2, 64);
  HIPCHK(c, hipHostMalloc((void**)p, want * sizeof(T), hipHostMallocDefault));
  *cap = want;
  return TVK_OK;
}

inline int bit_width(const tvk_ctx* c) { return c->L.bpv == 1 ? 8 : 16; }

// The compatibility textures (AbstrRenderer.h:170-174): bDownSampleTo8Bits
// applies to data that is not 8 bits already; bUseOnlyPowerOfTwo to all.
inline bool downsampled(const tvk_ctx* c) { return c->compat_8bit != 0 && c->L.bpv == 2; }
inline bool compat_active(const tvk_ctx* c) { return c->has_ds && (c->compat_pow2 != 0 || downsampled(c)); }
// bytes per voxel of the textures rendering reads
inline int render_bpv(const tvk_ctx* c) { return downsampled(c) ? 1 : c->L.bpv; }
// the compatibility textures exist for the current dataset, flags and border
inline bool compat_ready(const tvk_ctx* c) {
  return c->cp.valid && c->cp.pow2 == c->compat_pow2 && c->cp.down8 == (downsampled(c) ? 1 : 0) &&
         c->cp.dup == (c->P.clamp_to_edge ? 0 : 1);
}
inline const uint32_t* compat_dims(const tvk_ctx* c) { return compat_active(c) ? c->cp.dims.data() : nullptr; }
inline const uint64_t* compat_offsets(const tvk_ctx* c) { return compat_active(c) ? c->cp.offset.data() : nullptr; }
// (re)makes the compatibility textures when they are out of date
int ensure_compat(tvk_ctx* c);
// AbstrRenderer::GetNormalizedIsovalue (AbstrRenderer.cpp:412-424)
float normalized_isovalue(const tvk_ctx* c);

// LOD of a frame (fixed, or the screen-space-error minimum of the view)
int frame_lod(tvk_ctx* c, const ViewState& V, int* lod);
// planner parameters (mode, LOD, TF non-zero limits) of the context
int plan_params(tvk_ctx* c, int lod, PlanParams& pp);
// GLRenderer::CalculateScaling
float tf_scale_of(const tvk_ctx* c);
// TF statistics over a brick's value range (FrameBrick::tfs) and its cell grid
void brick_tf_stats(tvk_ctx* c, FrameBrick& fb);
// learned mode: device TF statistics of the active TF and the dataset's cell grid
int learn_prepare(tvk_ctx* c);
// the dataset's cell grid (built on first use)
int ensure_cells(tvk_ctx* c);
// the step network as the kernels read it (false when s is null)
bool stepnet_device(const struct tvk_stepnet* s, StepNetDev* out);
// shading constants of the context's dataset, TF and parameters
void fill_shade(const tvk_ctx* c, ShadeConst& S);

// Budgeted pools (pool.cpp): slots for the bricks of `need` (TOC indices, at most
// nslots distinct) by least recent use, paging in the ones not resident;
// slot_of[i] receives the slot of need[i].  Stream-ordered: the uploads run on
// the context stream after the work already queued there.
int pool_acquire_lru(tvk_ctx* c, const std::vector<uint32_t>& need, std::vector<int32_t>& slot_of);
// Pages the given TOC bricks into the given slots (no policy).
int pool_upload(tvk_ctx* c, const std::vector<uint32_t>& toc, const std::vector<int32_t>& slots);
// Converts the current dataset to a budgeted one (budget 0: back to all resident
// needs the payloads: TVK_ERR_UNSUPPORTED unless it is already resident).
int pool_set_budget(tvk_ctx* c, uint64_t budget_bytes);
// Copies bricks [first, first+n) in the reference layout to host memory.
int pool_download(tvk_ctx* c, size_t first, size_t n, void* host, size_t bytes);
inline bool budgeted(const tvk_ctx* c) { return c->pst != nullptr; }
// a frame of a budgeted dataset (fixed step, any mode)
int render_budget_impl(tvk_ctx* c, int nviews, const float* rots, float* host_out, tvk_stats* st);

// Device buffers of an allocation-free render (capacities in entries).
struct FrameBuffers {
  FrameBrick* list;
  size_t list_cap;
  ViewConst* views;
  size_t views_cap;
  unsigned long long* counter;
  float* out;
  size_t out_floats;
};
// Plans and renders nviews views (rots: 16 floats each, or null for the
// context's rotation); learn < 0 takes the context's step proposal.  With
// fixed set nothing is allocated and the frames go to fixed->out.
int render_impl(tvk_ctx* c, int nviews, const float* rots, float* host_out, tvk_stats* st, int learn,
                const SegOut* seg, const FrameBuffers* fixed);
// sets n words at p to v with a kernel (capturable, unlike hipMemsetAsync)
hipError_t launch_fill_u32(unsigned int* p, unsigned int n, unsigned int v, hipStream_t st);
// this device's address of the allocation-free render's scratch array (call
// with the device current; looked up once per device)
hipError_t scratch_address(int device, void** out);

}  // namespace tvk
