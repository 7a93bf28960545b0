// Brick-level ingest and the budgeted HBM brick pool.
//
// Reference interfaces:
//   tvk_set_bricks / tvk_upload_brick
//       a BrickedDataset handed over brick by brick: BrickedDataset::GetBrickMetadata
//       and MaxMinForKey (IO/BrickedDataset.cpp:52-67, IO/uvfDataset.cpp:1183-1199)
//       for the table, Dataset::GetBrick (IO/Dataset.h:93-100) for each payload
//   tvk_set_pool_budget
//       GPUMemMan's memory budget and texture reuse (Renderer/GPUMemMan/GPUMemMan.cpp:766-844,
//       :846-996): the bricks of a frame are paged into a fixed set of slots,
//       least recently used first; a frame whose bricks exceed the slots is
//       rendered in parts in compositing order (the reference draws brick by
//       brick, uploading each before its draw, GLRenderer.cpp:2663-2748)
//   the traversal's paging (GLVolumePool::UploadBricks, GLVolumePool.cpp:673-717,
//       :955-959, :1720-1789) uses the same slots with its own policy (leap_api.cpp)
#include <chrono>
#include <cstdlib>
#include "ctx.hpp"

using namespace tvk;

namespace tvk {
hipError_t launch_unswizzle(const void* pool, int bpv, const GatherBrick* d_bricks, const uint64_t* d_ref_off,
                            uint32_t nb, uint32_t max_vox, void* out, hipStream_t st);
}

namespace {

constexpr size_t kChunkBytes = size_t(1) << 30;    // pinned host store allocation unit
constexpr size_t kStageBytes = size_t(256) << 20;  // device staging of one upload batch

GatherBrick gather_of(const Layout& L, size_t t, uint64_t offset) {
  GatherBrick g;
  for (int k = 0; k < 3; ++k) { g.c[k] = (uint32_t)L.toc[t].coords[k]; g.n[k] = L.toc[t].size[k]; }
  g.offset = offset;
  return g;
}

// Host memory for every brick of L in the reference layout: pinned chunks
// (asynchronous uploads) until the host refuses more page-locked memory, then
// pageable chunks (the runtime stages those uploads itself).
int alloc_store(tvk_ctx* c, PoolState& P) {
  const Layout& L = c->L;
  P.brick_host.assign(L.toc.size(), nullptr);
  P.present.assign(L.toc.size(), 0);
  size_t i = 0;
  while (i < L.toc.size()) {
    size_t bytes = 0, j = i;
    while (j < L.toc.size()) {
      const size_t b = (size_t)L.toc[j].real * (size_t)L.bpv;
      if (j > i && bytes + b > kChunkBytes) break;
      bytes += b;
      ++j;
    }
    void* p = nullptr;
    bool pinned = false;
    if (!P.pinned_refused) {
      if (hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault) == hipSuccess) {
        pinned = true;
      } else {
        (void)hipGetLastError();   // clear the refusal
        P.pinned_refused = true;
        p = nullptr;
      }
    }
    if (!pinned) {
      p = std::malloc(std::max<size_t>(bytes, 1));
      if (!p) return fail(c, TVK_ERR_NOMEM, "host memory for the brick store");
    }
    P.chunks.push_back(p);
    P.chunk_pinned.push_back(pinned ? 1 : 0);
    size_t off = 0;
    for (size_t k = i; k < j; ++k) {
      P.brick_host[k] = (uint8_t*)p + off;
      off += (size_t)L.toc[k].real * (size_t)L.bpv;
    }
    i = j;
  }
  return TVK_OK;
}

int ensure_stage(tvk_ctx* c, PoolState& P, size_t bytes, size_t nb) {
  if (P.stage_bytes < bytes) {
    if (P.d_stage) (void)hipFree(P.d_stage);
    P.d_stage = nullptr;
    P.stage_bytes = 0;
    HIPCHK(c, hipMalloc((void**)&P.d_stage, bytes));
    P.stage_bytes = bytes;
  }
  if (P.desc_cap < nb) {
    if (P.d_desc) (void)hipFree(P.d_desc);
    if (P.d_srcoff) (void)hipFree(P.d_srcoff);
    P.d_desc = nullptr;
    P.d_srcoff = nullptr;
    P.desc_cap = 0;
    HIPCHK(c, hipMalloc((void**)&P.d_desc, nb * sizeof(GatherBrick)));
    HIPCHK(c, hipMalloc((void**)&P.d_srcoff, nb * sizeof(uint64_t)));
    P.desc_cap = nb;
  }
  return TVK_OK;
}

// Uploads host payloads (reference layout) into pool element offsets dst[i] of
// `pool`, through the device staging buffer in batches.
int upload_to(tvk_ctx* c, PoolState* P, const std::vector<const uint8_t*>& src, const std::vector<uint32_t>& toc,
              const std::vector<uint64_t>& dst, void* pool) {
  const Layout& L = c->L;
  const size_t bpv = (size_t)L.bpv;
  size_t largest = 0;
  for (uint32_t t : toc) largest = std::max(largest, (size_t)L.toc[t].real * bpv);
  PoolState tmp;
  PoolState& S = P ? *P : tmp;   // staging holder (a one-off one when there is no pool state)
  const size_t stage = std::max(largest, std::min(kStageBytes, [&] {
                                  size_t s = 0;
                                  for (uint32_t t : toc) s += (size_t)L.toc[t].real * bpv;
                                  return s;
                                }()));
  int rc = ensure_stage(c, S, stage, std::min<size_t>(toc.size(), 65535));
  if (rc) return rc;
  std::vector<GatherBrick> desc;
  std::vector<uint64_t> soff;
  size_t i = 0;
  while (i < toc.size()) {
    desc.clear();
    soff.clear();
    size_t used = 0;
    uint64_t maxe = 0;
    while (i < toc.size() && desc.size() < S.desc_cap) {
      const size_t b = (size_t)L.toc[toc[i]].real * bpv;
      if (used + b > S.stage_bytes) break;
      HIPCHK(c, hipMemcpyAsync(S.d_stage + used, src[i], b, hipMemcpyHostToDevice, c->stream));
      desc.push_back(gather_of(L, toc[i], dst[i]));
      soff.push_back(used / bpv);
      maxe = std::max<uint64_t>(maxe, L.toc[toc[i]].count);
      used += b;
      ++i;
    }
    if (desc.empty()) return fail(c, TVK_ERR_NOMEM, "a brick does not fit the upload staging buffer");
    HIPCHK(c, hipMemcpyAsync(S.d_desc, desc.data(), desc.size() * sizeof(GatherBrick), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(S.d_srcoff, soff.data(), soff.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, launch_swizzle(S.d_stage, L.bpv, S.d_desc, S.d_srcoff, (uint32_t)desc.size(), maxe, pool, c->stream));
    // the staging buffer and the descriptor vectors are reused by the next batch
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (P) {
      P->upload_bytes += used;
      P->uploads += desc.size();
    }
  }
  return TVK_OK;
}

}  // namespace

PoolState::~PoolState() {
  for (size_t i = 0; i < chunks.size(); ++i) {
    if (i < chunk_pinned.size() && chunk_pinned[i]) (void)hipHostFree(chunks[i]);
    else std::free(chunks[i]);
  }
  if (d_stage) (void)hipFree(d_stage);
  if (d_desc) (void)hipFree(d_desc);
  if (d_srcoff) (void)hipFree(d_srcoff);
}

namespace tvk {

int pool_upload(tvk_ctx* c, const std::vector<uint32_t>& toc, const std::vector<int32_t>& slots) {
  PoolState& P = *c->pst;
  std::vector<const uint8_t*> src(toc.size());
  std::vector<uint64_t> dst(toc.size());
  for (size_t i = 0; i < toc.size(); ++i) {
    const uint32_t t = toc[i];
    if (!P.present[t]) {
      const TocBrick& b = c->L.toc[t];
      return fail(c, TVK_ERR_STATE, "brick (lod " + std::to_string(b.coords[3]) + ", " + std::to_string(b.coords[0]) +
                                        "," + std::to_string(b.coords[1]) + "," + std::to_string(b.coords[2]) +
                                        ") was never uploaded");
    }
    src[i] = P.brick_host[t];
    dst[i] = (uint64_t)slots[i] * P.slot_elems;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = upload_to(c, &P, src, toc, dst, c->d_pool);
  P.upload_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc) return rc;
  for (size_t i = 0; i < toc.size(); ++i) {
    const int32_t s = slots[i];
    const int64_t old = P.slot_brick[(size_t)s];
    if (old >= 0 && old != (int64_t)toc[i]) {
      P.brick_slot[(size_t)old] = -1;
      ++P.evictions;
    }
    P.slot_brick[(size_t)s] = toc[i];
    P.brick_slot[toc[i]] = s;
  }
  return TVK_OK;
}

int pool_acquire_lru(tvk_ctx* c, const std::vector<uint32_t>& need, std::vector<int32_t>& slot_of) {
  PoolState& P = *c->pst;
  const uint64_t now = ++P.clock;
  std::vector<uint32_t> missing;
  std::vector<uint8_t> keep(P.nslots, 0);
  for (uint32_t t : need) {
    const int32_t s = P.brick_slot[t];
    if (s >= 0) {
      P.slot_time[(size_t)s] = now;
      keep[(size_t)s] = 1;
    } else if (std::find(missing.begin(), missing.end(), t) == missing.end()) {
      missing.push_back(t);
    }
  }
  if (!missing.empty()) {
    // least recently used slots first (never used ones have time 0)
    std::vector<int32_t> cand;
    for (uint32_t s = 0; s < P.nslots; ++s)
      if (!keep[s]) cand.push_back((int32_t)s);
    if (cand.size() < missing.size())
      return fail(c, TVK_ERR_NOMEM, "the frame part needs more bricks than the pool holds");
    std::stable_sort(cand.begin(), cand.end(),
                     [&](int32_t a, int32_t b) { return P.slot_time[(size_t)a] < P.slot_time[(size_t)b]; });
    cand.resize(missing.size());
    const int rc = pool_upload(c, missing, cand);
    if (rc) return rc;
    for (int32_t s : cand) P.slot_time[(size_t)s] = now;
  }
  slot_of.resize(need.size());
  for (size_t i = 0; i < need.size(); ++i) slot_of[i] = P.brick_slot[need[i]];
  return TVK_OK;
}

int pool_download(tvk_ctx* c, size_t first, size_t n, void* host, size_t bytes) {
  const PoolState& P = *c->pst;
  const Layout& L = c->L;
  size_t total = 0;
  for (size_t k = first; k < first + n; ++k) total += (size_t)L.toc[k].real * (size_t)L.bpv;
  if (bytes != total) return fail(c, TVK_ERR_ARG, "size mismatch");
  uint8_t* o = (uint8_t*)host;
  for (size_t k = first; k < first + n; ++k) {
    const size_t b = (size_t)L.toc[k].real * (size_t)L.bpv;
    if (!P.present[k]) return fail(c, TVK_ERR_STATE, "brick " + std::to_string(k) + " was never uploaded");
    std::memcpy(o, P.brick_host[k], b);
    o += b;
  }
  return TVK_OK;
}

int pool_set_budget(tvk_ctx* c, uint64_t budget) {
  Layout& L = c->L;
  const size_t bpv = (size_t)L.bpv;
  const size_t nb = L.toc.size();
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (budget == 0) {
    if (!c->pst) return TVK_OK;
    // back to every brick resident at its TOC offset
    PoolState& P = *c->pst;
    for (size_t t = 0; t < nb; ++t)
      if (!P.present[t]) return fail(c, TVK_ERR_STATE, "not every brick was uploaded");
    void* full = nullptr;
    HIPCHK(c, hipMalloc(&full, (size_t)L.pool_elems * bpv));
    std::shared_ptr<void> hold(full, [](void* p) { (void)hipFree(p); });
    std::vector<const uint8_t*> src(nb);
    std::vector<uint32_t> toc(nb);
    std::vector<uint64_t> dst(nb);
    for (size_t t = 0; t < nb; ++t) { src[t] = P.brick_host[t]; toc[t] = (uint32_t)t; dst[t] = L.toc[t].offset; }
    const int rc = upload_to(c, nullptr, src, toc, dst, full);
    if (rc) return rc;
    c->d_pool = full;
    c->pool_hold = hold;
    c->pool_bytes = (size_t)L.pool_elems * bpv;
    c->pst.reset();
    c->cells = std::make_shared<CellGrid>();
    return TVK_OK;
  }
  uint64_t slot = 0;
  for (const TocBrick& b : L.toc) slot = std::max<uint64_t>(slot, b.count);
  const uint64_t fit = budget / (slot * bpv);
  if (fit < 2) return fail(c, TVK_ERR_ARG, "the budget holds fewer than two bricks");
  auto P = std::make_shared<PoolState>();
  P->bpv = (int)bpv;
  P->slot_elems = slot;
  P->nslots = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(fit, nb), 0x7fffffff);
  if (c->pst) {
    // keep the host store, re-make the slots
    PoolState& O = *c->pst;
    P->chunks.swap(O.chunks);
    P->chunk_pinned.swap(O.chunk_pinned);
    P->pinned_refused = O.pinned_refused;
    P->brick_host.swap(O.brick_host);
    P->present.swap(O.present);
  } else {
    int rc = alloc_store(c, *P);
    if (rc) return rc;
    // every brick, from the resident pool, in batches through a device buffer
    size_t i = 0;
    const size_t batch_bytes = size_t(512) << 20;
    void* tmp = nullptr;
    size_t largest = 0;
    for (const TocBrick& b : L.toc) largest = std::max(largest, (size_t)b.real * bpv);
    const size_t tmp_bytes = std::max(largest, std::min(batch_bytes, (size_t)L.ref_elems * bpv));
    HIPCHK(c, hipMalloc(&tmp, tmp_bytes));
    std::shared_ptr<void> tmp_hold(tmp, [](void* p) { (void)hipFree(p); });
    GatherBrick* d_gb = nullptr;
    uint64_t* d_ro = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_gb, 65535 * sizeof(GatherBrick)));
    std::shared_ptr<void> gb_hold(d_gb, [](void* p) { (void)hipFree(p); });
    HIPCHK(c, hipMalloc((void**)&d_ro, 65535 * sizeof(uint64_t)));
    std::shared_ptr<void> ro_hold(d_ro, [](void* p) { (void)hipFree(p); });
    while (i < nb) {
      std::vector<GatherBrick> gb;
      std::vector<uint64_t> ro;
      size_t used = 0;
      uint32_t maxv = 0;
      const size_t first = i;
      while (i < nb && gb.size() < 65535) {
        const size_t b = (size_t)L.toc[i].real * bpv;
        if (used + b > tmp_bytes) break;
        gb.push_back(gather_of(L, i, L.toc[i].offset));
        ro.push_back(used / bpv);
        maxv = std::max<uint32_t>(maxv, (uint32_t)L.toc[i].real);
        used += b;
        ++i;
      }
      HIPCHK(c, hipMemcpyAsync(d_gb, gb.data(), gb.size() * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(d_ro, ro.data(), ro.size() * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, launch_unswizzle(c->d_pool, L.bpv, d_gb, d_ro, (uint32_t)gb.size(), maxv, tmp, c->stream));
      // bricks are consecutive in the store within a chunk; copy brick by brick
      size_t off = 0;
      for (size_t k = first; k < i; ++k) {
        const size_t b = (size_t)L.toc[k].real * bpv;
        HIPCHK(c, hipMemcpyAsync(P->brick_host[k], (uint8_t*)tmp + off, b, hipMemcpyDeviceToHost, c->stream));
        off += b;
      }
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (size_t k = first; k < i; ++k) P->present[k] = 1;
    }
  }
  void* slots = nullptr;
  HIPCHK(c, hipMalloc(&slots, (size_t)P->nslots * slot * bpv));
  P->slot_brick.assign(P->nslots, -1);
  P->slot_time.assign(P->nslots, 0);
  P->brick_slot.assign(nb, -1);
  c->pool_hold = std::shared_ptr<void>(slots, [](void* p) { (void)hipFree(p); });
  c->d_pool = slots;
  c->pool_bytes = (size_t)P->nslots * slot * bpv;
  c->pst = P;
  c->cells.reset();
  c->leap_resident.clear();
  return TVK_OK;
}

}  // namespace tvk

namespace {

// Budgeted frame: per view, the brick list in compositing order is cut into
// parts whose distinct bricks fit the slots; each part pages its bricks in and
// continues the pixels of the part before it.
int render_budgeted(tvk_ctx* c, int nviews, const float* rots, float* host_out, tvk_stats* st) {
  PoolState& P = *c->pst;
  struct Part {
    int view;
    size_t first, n;   // range of the view's list
    bool cont;
  };
  std::vector<std::vector<FrameBrick>> lists((size_t)nviews);
  std::vector<ViewConst> vcs((size_t)nviews);
  std::vector<PlannedBrick> planned;
  uint64_t pixels = 0;
  uint32_t in_list = 0, nonempty = 0;
  for (int v = 0; v < nviews; ++v) {
    ViewState V = c->V;
    if (rots) V.rot = Mat4::from(rots + 16 * v);
    int lod = 0;
    int rc = frame_lod(c, V, &lod);
    if (rc) return rc;
    PlanParams pp;
    rc = plan_params(c, lod, pp);
    if (rc) return rc;
    plan_frame(c->L, c->minmax, c->range_max, V, pp, planned);
    in_list += (uint32_t)planned.size();
    make_frame_bricks(c->L, V, pp, planned, lists[(size_t)v]);
    nonempty += (uint32_t)lists[(size_t)v].size();
    make_view_const(V, vcs[(size_t)v]);
    vcs[(size_t)v].out_offset = (uint32_t)pixels;
    vcs[(size_t)v].group_offset = 0;
    pixels += (uint64_t)V.W * V.H;
    c->last_lod = lod;
  }
  if (pixels >= (1ull << 32)) return fail(c, TVK_ERR_ARG, "too many pixels in one call");
  std::vector<Part> parts;
  for (int v = 0; v < nviews; ++v) {
    const std::vector<FrameBrick>& L = lists[(size_t)v];
    size_t first = 0;
    bool cont = false;
    do {
      std::vector<uint32_t> distinct;
      size_t k = first;
      for (; k < L.size(); ++k) {
        const uint32_t t = L[k].toc;
        if (std::find(distinct.begin(), distinct.end(), t) != distinct.end()) continue;
        if (distinct.size() == P.nslots) break;
        distinct.push_back(t);
      }
      parts.push_back(Part{v, first, k - first, cont});
      first = k;
      cont = true;
    } while (first < lists[(size_t)v].size());
  }
  const bool iso = c->P.mode == TVK_RM_ISOSURFACE;
  size_t total_list = 0;
  for (const auto& l : lists) total_list += l.size();
  int rc = ensure(c, &c->d_list, &c->list_cap, std::max<size_t>(total_list, 1));
  if (!rc) rc = ensure(c, &c->d_views, &c->views_cap, parts.size());
  if (!rc) rc = ensure(c, &c->d_out, &c->out_cap, (size_t)pixels * 4);
  if (!rc && iso) rc = ensure(c, &c->d_depth, &c->depth_cap, (size_t)pixels);
  if (!rc && iso) rc = ensure(c, &c->d_hitpos, &c->hitpos_cap, (size_t)pixels);
  if (rc) return rc;
  if (c->count) {
    const size_t nt = (size_t)nviews * c->L.toc.size();
    rc = ensure(c, &c->d_touched, &c->touched_cap, nt);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_touched, 0, nt * sizeof(uint32_t), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_counter, 0, sizeof(unsigned long long), c->stream));
    c->touched_n = nt;
  }
  // host sources of the per-part uploads stay valid until the final synchronisation
  std::vector<FrameBrick> h_list(std::max<size_t>(total_list, 1));
  std::vector<ViewConst> h_views(parts.size());
  std::vector<size_t> list_base((size_t)nviews, 0);
  for (int v = 1; v < nviews; ++v) list_base[(size_t)v] = list_base[(size_t)v - 1] + lists[(size_t)v - 1].size();
  HIPCHK(c, hipEventRecord(c->ev[6], c->stream));
  std::vector<uint32_t> need;
  std::vector<int32_t> slot_of;
  for (size_t pi = 0; pi < parts.size(); ++pi) {
    const Part& pt = parts[pi];
    std::vector<FrameBrick>& L = lists[(size_t)pt.view];
    need.clear();
    for (size_t k = pt.first; k < pt.first + pt.n; ++k) need.push_back(L[k].toc);
    rc = pool_acquire_lru(c, need, slot_of);
    if (rc) return rc;
    const size_t base = list_base[(size_t)pt.view] + pt.first;
    for (size_t k = 0; k < pt.n; ++k) {
      FrameBrick fb = L[pt.first + k];
      fb.offset = (uint64_t)slot_of[k] * P.slot_elems;
      h_list[base + k] = fb;
    }
    ViewConst vc = vcs[(size_t)pt.view];
    vc.list_offset = (uint32_t)base;
    vc.nlist = (uint32_t)pt.n;
    h_views[pi] = vc;
    if (pt.n)
      HIPCHK(c, hipMemcpyAsync(c->d_list + base, h_list.data() + base, pt.n * sizeof(FrameBrick),
                               hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_views + pi, &h_views[pi], sizeof(ViewConst), hipMemcpyHostToDevice, c->stream));
    MarchArgs a{};
    fill_shade(c, a.S);
    a.list = c->d_list;
    a.views = c->d_views + pi;
    a.mode = c->P.mode;
    a.out = c->d_out;
    a.counter = c->count ? c->d_counter : nullptr;
    a.ntoc = (uint32_t)c->L.toc.size();
    a.touched = c->count ? c->d_touched + (size_t)pt.view * c->L.toc.size() : nullptr;
    a.nviews = 1;
    a.lighting = c->P.lighting;
    a.groups_total = vc.groups;
    a.groups_padded = (vc.groups + 7u) & ~7u;
    a.accumulate = pt.cont ? 1 : 0;
    a.depth = iso ? c->d_depth : nullptr;
    a.hitpos = iso ? c->d_hitpos : nullptr;
    if (a.groups_total > 0) HIPCHK(c, launch_march(a, c->L.bpv, c->stream, 0));
  }
  HIPCHK(c, hipEventRecord(c->ev[7], c->stream));
  c->last_views = nviews;
  c->last_w = c->V.W;
  c->last_h = c->V.H;
  if (host_out)
    HIPCHK(c, hipMemcpyAsync(host_out, c->d_out, (size_t)pixels * 4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  unsigned long long ns = 0;
  if (c->count) HIPCHK(c, hipMemcpyAsync(&ns, c->d_counter, sizeof(ns), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    st->samples = ns;
    st->bricks_in_list = in_list;
    st->bricks_nonempty = nonempty;
    st->ms = ms;
  }
  c->last_parts = (uint32_t)parts.size();
  c->hitpos_valid = iso;
  return TVK_OK;
}

}  // namespace

namespace tvk {
int render_budget_impl(tvk_ctx* c, int nviews, const float* rots, float* host_out, tvk_stats* st) {
  if (compat_active(c))
    return fail(c, TVK_ERR_UNSUPPORTED, "compatibility textures need every brick resident (pool budget 0)");
  return render_budgeted(c, nviews, rots, host_out, st);
}
}  // namespace tvk

extern "C" {

int tvk_set_bricks(tvk_ctx* c, int dtype, const uint64_t dims[3], const uint32_t brick[3], uint32_t overlap,
                   int lod_count, const tvk_brick_key* keys, size_t n, uint64_t budget) {
  if (!c) return TVK_ERR_ARG;
  if (!keys) return fail(c, TVK_ERR_ARG, "null brick keys");
  if (lod_count < 0) return fail(c, TVK_ERR_ARG, "lod_count must be >= 0");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Layout fresh;
  std::string e;
  if (dtype != TVK_U8 && dtype != TVK_U16) return fail(c, TVK_ERR_ARG, "dtype must be TVK_U8 or TVK_U16");
  if (!dims || !brick) return fail(c, TVK_ERR_ARG, "null dims/brick");
  if (!fresh.build(dims, brick, overlap, dtype, lod_count, &e)) return fail(c, TVK_ERR_ARG, e);
  const size_t nb = fresh.toc.size();
  if (n != nb)
    return fail(c, TVK_ERR_ARG, "the layout has " + std::to_string(nb) + " bricks, " + std::to_string(n) + " keys given");
  std::vector<double> mm(2 * nb, 0.0);
  std::vector<uint8_t> seen(nb, 0);
  for (size_t i = 0; i < n; ++i) {
    const tvk_brick_key& k = keys[i];
    if (k.lod >= fresh.lods.size()) return fail(c, TVK_ERR_RANGE, "unknown brick (lod out of range)");
    const LodInfo& li = fresh.lods[k.lod];
    if (k.index >= li.bricks[0] * li.bricks[1] * li.bricks[2]) return fail(c, TVK_ERR_RANGE, "unknown brick");
    const size_t t = (size_t)(li.offset + k.index);
    if (seen[t]) return fail(c, TVK_ERR_ARG, "brick listed twice");
    seen[t] = 1;
    if (!(k.minmax[0] <= k.minmax[1])) return fail(c, TVK_ERR_ARG, "brick min exceeds max");
    mm[2 * t] = k.minmax[0];
    mm[2 * t + 1] = k.minmax[1];
  }
  // the new dataset replaces the current one
  c->pool_hold.reset();
  c->lin_hold.reset();
  c->cells.reset();
  c->pst.reset();
  c->d_pool = c->d_lin = nullptr;
  c->pool_bytes = c->lin_bytes = 0;
  c->has_ds = false;
  c->leap_resident.clear();
  c->cp = tvk_ctx::CompatPool{};
  c->L = std::move(fresh);
  c->clamp = 0;
  c->minmax = std::move(mm);
  c->build_ms = 0.0f;
  const LodInfo& l0 = c->L.lods[0];
  const uint64_t n0 = l0.bricks[0] * l0.bricks[1] * l0.bricks[2];
  c->range_min = c->minmax[0];
  c->range_max = c->minmax[1];
  for (uint64_t i = 1; i < n0; ++i) {   // UVFDataset::ComputeRange (uvfDataset.cpp:1120-1165)
    c->range_min = std::min(c->range_min, c->minmax[2 * i]);
    c->range_max = std::max(c->range_max, c->minmax[2 * i + 1]);
  }
  uint64_t slot = 0;
  for (const TocBrick& b : c->L.toc) slot = std::max<uint64_t>(slot, b.count);
  const size_t full = (size_t)c->L.pool_elems * (size_t)dtype;
  if (budget == 0 || budget >= full) {
    // every brick resident at its TOC offset; uploads go straight to the pool
    HIPCHK(c, hipMalloc(&c->d_pool, full));
    c->pool_hold = std::shared_ptr<void>(c->d_pool, [](void* p) { (void)hipFree(p); });
    c->pool_bytes = full;
    c->cells = std::make_shared<CellGrid>();
    c->ingest_missing.assign(nb, 1);
    c->ingest_pending = nb;
  } else {
    auto P = std::make_shared<PoolState>();
    P->bpv = dtype;
    P->slot_elems = slot;
    const uint64_t fit = budget / (slot * (uint64_t)dtype);
    if (fit < 2) return fail(c, TVK_ERR_ARG, "the budget holds fewer than two bricks");
    P->nslots = (uint32_t)std::min<uint64_t>(fit, nb);
    int rc = alloc_store(c, *P);
    if (rc) return rc;
    void* slots = nullptr;
    HIPCHK(c, hipMalloc(&slots, (size_t)P->nslots * slot * (size_t)dtype));
    c->pool_hold = std::shared_ptr<void>(slots, [](void* p) { (void)hipFree(p); });
    c->d_pool = slots;
    c->pool_bytes = (size_t)P->nslots * slot * (size_t)dtype;
    P->slot_brick.assign(P->nslots, -1);
    P->slot_time.assign(P->nslots, 0);
    P->brick_slot.assign(nb, -1);
    c->pst = P;
    c->ingest_missing.clear();
    c->ingest_pending = 0;
  }
  c->has_ds = true;
  return TVK_OK;
}

int tvk_upload_brick(tvk_ctx* c, uint32_t lod, uint64_t index, const void* data, size_t bytes) {
  if (!c || !data) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  const Layout& L = c->L;
  if (lod >= L.lods.size()) return fail(c, TVK_ERR_RANGE, "unknown brick (lod out of range)");
  const LodInfo& li = L.lods[lod];
  if (index >= li.bricks[0] * li.bricks[1] * li.bricks[2]) return fail(c, TVK_ERR_RANGE, "unknown brick");
  const uint32_t t = (uint32_t)(li.offset + index);
  const size_t want = (size_t)L.toc[t].real * (size_t)L.bpv;
  if (bytes != want)
    return fail(c, TVK_ERR_ARG, "brick payload is " + std::to_string(want) + " bytes (stored size incl. overlap)");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->pst) {
    PoolState& P = *c->pst;
    // the host copy changes: wait for queued uploads that read it
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(P.brick_host[t], data, bytes);
    P.present[t] = 1;
    const int32_t s = P.brick_slot[t];
    if (s >= 0) {   // a resident copy is stale: refresh it in place
      const int rc = pool_upload(c, std::vector<uint32_t>{t}, std::vector<int32_t>{s});
      if (rc) return rc;
    }
    return TVK_OK;
  }
  if (c->ingest_missing.size() != L.toc.size())
    return fail(c, TVK_ERR_STATE, "the dataset was not declared with tvk_set_bricks");
  const int rc = upload_to(c, nullptr, std::vector<const uint8_t*>{(const uint8_t*)data}, std::vector<uint32_t>{t},
                           std::vector<uint64_t>{L.toc[t].offset}, c->d_pool);
  if (rc) return rc;
  if (c->ingest_missing[t]) {
    c->ingest_missing[t] = 0;
    --c->ingest_pending;
  }
  c->cp = tvk_ctx::CompatPool{};   // made again from the new contents on the next render
  if (c->cells && c->cells->d) c->cells = std::make_shared<CellGrid>();   // rebuilt on next use
  return TVK_OK;
}

int tvk_set_pool_budget(tvk_ctx* c, uint64_t budget) {
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (c->ingest_pending) return fail(c, TVK_ERR_STATE, "not every brick was uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  return pool_set_budget(c, budget);
}

int tvk_get_pool_stats(tvk_ctx* c, tvk_pool_stats* o) {
  if (!c || !o) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  std::memset(o, 0, sizeof(*o));
  o->pool_bytes = c->pool_bytes;
  if (!c->pst) {
    o->slots = c->L.toc.size();
    o->resident = c->L.toc.size() - c->ingest_pending;
    return TVK_OK;
  }
  const PoolState& P = *c->pst;
  o->slots = P.nslots;
  o->slot_bytes = P.slot_elems * (uint64_t)P.bpv;
  for (int64_t b : P.slot_brick) o->resident += b >= 0 ? 1 : 0;
  o->uploads = P.uploads;
  o->evictions = P.evictions;
  o->upload_bytes = P.upload_bytes;
  o->upload_ms = (float)P.upload_ms;
  o->parts = c->last_parts;
  return TVK_OK;
}

}  // extern "C"
