// C ABI of the raycasting path (include/tvk_hip.h).
#include "ctx.hpp"

using namespace tvk;

namespace {

std::vector<float4> unorm_table(const uint8_t* rgba8, size_t n) {
  std::vector<float4> f(n);
  for (size_t i = 0; i < n; ++i)
    f[i] = make_float4((float)rgba8[4 * i] / 255.0f, (float)rgba8[4 * i + 1] / 255.0f,
                       (float)rgba8[4 * i + 2] / 255.0f, (float)rgba8[4 * i + 3] / 255.0f);
  return f;
}

void free_dataset(tvk_ctx* c) {
  c->pool_hold.reset();   // frees the pool once no context shares it
  c->lin_hold.reset();
  c->cells.reset();
  c->pst.reset();
  c->ingest_missing.clear();
  c->ingest_pending = 0;
  c->d_pool = c->d_lin = nullptr;
  c->pool_bytes = c->lin_bytes = 0;
  c->has_ds = false;
  c->minmax.clear();
  c->cp = tvk_ctx::CompatPool{};
}


// Brick the linear volume at d_lin (dims of c->L) into the pool.
int build_pool(tvk_ctx* c, const void* d_lin) {
  Layout& L = c->L;
  const int bpv = L.bpv;
  const size_t keep = L.toc.empty() ? 0 : (size_t)(L.toc.back().coords[3] + 1);
  const uint32_t core[3] = {L.brick[0] - 2 * L.overlap, L.brick[1] - 2 * L.overlap, L.brick[2] - 2 * L.overlap};
  // odd cores: the coarser levels are filtered brick to brick (the 2x2x2 groups
  // stop at each brick's core), which differs from filtering the level volume
  const bool odd = keep > 1 && (core[0] % 2 || core[1] % 2 || core[2] % 2);
  if (odd && c->clamp)
    return fail(c, TVK_ERR_UNSUPPORTED, "odd brick cores with coarser LODs need bClampToEdge off");
  if (odd && L.overlap == 0)
    return fail(c, TVK_ERR_UNSUPPORTED, "odd brick cores with coarser LODs need an overlap of at least 1");
  c->cp = tvk_ctx::CompatPool{};
  c->pool_bytes = (size_t)L.pool_elems * bpv;
  HIPCHK(c, hipMalloc(&c->d_pool, c->pool_bytes));
  c->pool_hold = std::shared_ptr<void>(c->d_pool, [](void* p) { (void)hipFree(p); });
  c->cells = std::make_shared<CellGrid>();
  // descriptors of every kept brick
  std::vector<GatherBrick> gb(L.toc.size());
  for (size_t i = 0; i < L.toc.size(); ++i) {
    for (int k = 0; k < 3; ++k) { gb[i].c[k] = (uint32_t)L.toc[i].coords[k]; gb[i].n[k] = L.toc[i].size[k]; }
    gb[i].offset = L.toc[i].offset;
  }
  GatherBrick* d_gb = nullptr;
  HIPCHK(c, hipMalloc(&d_gb, gb.size() * sizeof(GatherBrick)));
  HIPCHK(c, hipMemcpyAsync(d_gb, gb.data(), gb.size() * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream));
  void* tmp[2] = {nullptr, nullptr};
  const void* src = d_lin;
  int rc = TVK_OK;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) ev[0] = ev[1] = nullptr;
  if (ev[0]) (void)hipEventRecord(ev[0], c->stream);
  // per-brick min/max: folded into the gather; recomputed for the levels whose
  // overlaps FillOverlap rewrites afterwards
  uint32_t* d_mm = nullptr;
  {
    hipError_t e = hipMalloc(&d_mm, 2 * gb.size() * sizeof(uint32_t));
    if (e != hipSuccess) rc = hipfail(c, "hipMalloc(minmax)", e);
  }
  bool refilled = false;
  void* d_snap = nullptr;   // a level's bricks before FillOverlap (odd cores)
  for (size_t lod = 0; lod < keep && rc == TVK_OK; ++lod) {
    const LodInfo& li = L.lods[lod];
    const uint32_t nb = (uint32_t)(li.bricks[0] * li.bricks[1] * li.bricks[2]);
    uint32_t maxv = 0;
    for (uint32_t b = 0; b < nb; ++b) {
      const TocBrick& t = L.toc[li.offset + b];
      maxv = std::max<uint32_t>(maxv, (uint32_t)t.count);
    }
    if (lod > 0 && odd) {
      // ExtendedOctreeConverter::DownsampleBrick from the finer level's bricks
      const LodInfo& pl = L.lods[lod - 1];
      const uint32_t fb[3] = {(uint32_t)pl.bricks[0], (uint32_t)pl.bricks[1], (uint32_t)pl.bricks[2]};
      const uint32_t split[3] = {(core[0] + 1) / 2, (core[1] + 1) / 2, (core[2] + 1) / 2};
      hipError_t e = launch_downsample_bricks(c->d_pool, bpv, d_gb + li.offset, nb, d_gb + pl.offset, fb, L.overlap,
                                              split, maxv, c->lod_median, c->stream);
      if (e != hipSuccess) { rc = hipfail(c, "downsample_bricks", e); break; }
      refilled = true;   // min/max and repeated rows of the coarse levels are made after the fill
    } else {
      if (lod > 0) {
        const uint64_t* sd = L.lods[lod - 1].size;
        void*& dst = tmp[(lod - 1) % 2];
        if (!dst) {
          size_t bytes = (size_t)(li.size[0] * li.size[1] * li.size[2]) * bpv;
          hipError_t e = hipMalloc(&dst, bytes);
          if (e != hipSuccess) { rc = hipfail(c, "hipMalloc(lod temp)", e); break; }
        }
        hipError_t e = launch_downsample(src, sd, dst, li.size, bpv, c->stream, c->lod_median);
        if (e != hipSuccess) { rc = hipfail(c, "downsample", e); break; }
        src = dst;
      }
      hipError_t e = launch_gather(src, bpv, li.size, d_gb + li.offset, nb, (uint32_t)li.bricks[0], maxv, c->d_pool,
                                   core, L.overlap, c->clamp, d_mm + 2 * li.offset, c->stream);
      if (e != hipSuccess) { rc = hipfail(c, "gather", e); break; }
    }
    if (lod > 0 && !c->clamp && nb > 1) {
      hipError_t e = hipSuccess;
      uint64_t snap_base = 0;
      if (odd) {
        // the pre-fill state that later neighbours are read in
        const TocBrick& last = L.toc[li.offset + nb - 1];
        snap_base = L.toc[li.offset].offset;
        const size_t bytes = (size_t)(last.offset + last.count - snap_base) * bpv;
        if (d_snap) { (void)hipFree(d_snap); d_snap = nullptr; }
        e = hipMalloc(&d_snap, bytes);
        if (e == hipSuccess)
          e = hipMemcpyAsync(d_snap, (uint8_t*)c->d_pool + snap_base * bpv, bytes, hipMemcpyDeviceToDevice,
                             c->stream);
        if (e != hipSuccess) { rc = hipfail(c, "fill snapshot", e); break; }
      }
      refilled = true;
      // reproduce FillOverlap's copy order on the coarse level (see bricking.hip)
      const uint32_t bxyz[3] = {(uint32_t)li.bricks[0], (uint32_t)li.bricks[1], (uint32_t)li.bricks[2]};
      std::vector<uint32_t> order;
      std::vector<uint32_t> start;
      for (uint32_t w = 0; w + 2 < bxyz[0] + bxyz[1] + bxyz[2]; ++w) {
        start.push_back((uint32_t)order.size());
        for (uint32_t z = 0; z < bxyz[2]; ++z)
          for (uint32_t y = 0; y < bxyz[1]; ++y) {
            if (z + y > w || w - z - y >= bxyz[0]) continue;
            order.push_back((w - z - y) + bxyz[0] * (y + bxyz[1] * z));
          }
      }
      start.push_back((uint32_t)order.size());
      uint32_t* d_order = nullptr;
      if (e == hipSuccess) e = hipMalloc(&d_order, order.size() * sizeof(uint32_t));
      if (e == hipSuccess)
        e = hipMemcpyAsync(d_order, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
      uint32_t shell = 0;
      for (uint32_t b = 0; b < nb; ++b) {
        const TocBrick& t = L.toc[li.offset + b];
        const uint32_t sh = 2 * L.overlap * (t.size[1] * t.size[2] + t.size[0] * t.size[2] + t.size[0] * t.size[1]);
        shell = std::max(shell, sh);
      }
      for (size_t w = 0; w + 1 < start.size() && e == hipSuccess; ++w)
        if (start[w + 1] > start[w])
          e = launch_fill_overlap(c->d_pool, bpv, d_gb + li.offset, bxyz, L.overlap, d_order + start[w],
                                  start[w + 1] - start[w], shell, c->stream, d_snap, snap_base);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (d_order) (void)hipFree(d_order);
      if (e != hipSuccess) { rc = hipfail(c, "fill_overlap", e); break; }
    }
  }
  if (rc == TVK_OK && refilled && keep > 1) {
    // the coarse levels' rows repeat voxels FillOverlap may have rewritten:
    // refresh them (tvk_types.hpp) and their min/max
    const size_t first = L.lods[1].offset;
    uint32_t maxv = 0;
    for (size_t i = first; i < L.toc.size(); ++i) maxv = std::max<uint32_t>(maxv, (uint32_t)L.toc[i].count);
    hipError_t e = launch_extend(c->d_pool, bpv, d_gb + first, (uint32_t)(gb.size() - first), maxv, c->stream);
    if (e == hipSuccess)
      e = launch_minmax(c->d_pool, bpv, d_gb + first, (uint32_t)(gb.size() - first), maxv, d_mm + 2 * first,
                        c->stream);
    if (e != hipSuccess) rc = hipfail(c, "extend/minmax", e);
  }
  if (ev[1]) (void)hipEventRecord(ev[1], c->stream);
  std::vector<uint32_t> mm(2 * gb.size());
  if (rc == TVK_OK) {
    hipError_t e = hipMemcpyAsync(mm.data(), d_mm, mm.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hipfail(c, "minmax download", e);
  }
  (void)hipStreamSynchronize(c->stream);
  c->build_ms = 0.0f;
  if (ev[0] && ev[1] && rc == TVK_OK) (void)hipEventElapsedTime(&c->build_ms, ev[0], ev[1]);
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  if (d_mm) (void)hipFree(d_mm);
  if (d_snap) (void)hipFree(d_snap);
  (void)hipFree(d_gb);
  for (void* t : tmp) if (t) (void)hipFree(t);
  if (rc != TVK_OK) return rc;
  c->minmax.resize(mm.size());
  for (size_t i = 0; i < mm.size(); ++i) c->minmax[i] = (double)mm[i];
  // UVFDataset::ComputeRange over the LOD0 bricks (uvfDataset.cpp:1120-1165)
  const LodInfo& l0 = L.lods[0];
  const uint64_t n0 = l0.bricks[0] * l0.bricks[1] * l0.bricks[2];
  c->range_min = c->minmax[0];
  c->range_max = c->minmax[1];
  for (uint64_t i = 1; i < n0; ++i) {
    c->range_min = std::min(c->range_min, c->minmax[2 * i]);
    c->range_max = std::max(c->range_max, c->minmax[2 * i + 1]);
  }
  c->has_ds = true;
  return TVK_OK;
}

// Validates the new layout first: a rejected size leaves the current dataset as it was.
int begin_dataset(tvk_ctx* c, int dtype, const uint64_t dims[3], const uint32_t brick[3], uint32_t overlap,
                  int clamp, int max_lods) {
  if (dtype != TVK_U8 && dtype != TVK_U16) return fail(c, TVK_ERR_ARG, "dtype must be TVK_U8 or TVK_U16");
  if (!dims || !brick) return fail(c, TVK_ERR_ARG, "null dims/brick");
  Layout fresh;
  std::string e;
  if (!fresh.build(dims, brick, overlap, dtype, max_lods, &e)) return fail(c, TVK_ERR_ARG, e);
  free_dataset(c);
  c->L = std::move(fresh);
  c->clamp = clamp ? 1 : 0;
  return TVK_OK;
}

// The dataset fields of a context (the pool and linear volume are held, not copied).
struct DatasetState {
  Layout L;
  int clamp;
  void* d_pool;
  std::shared_ptr<void> pool_hold, lin_hold;
  std::shared_ptr<CellGrid> cells;
  std::shared_ptr<PoolState> pst;
  size_t pool_bytes;
  void* d_lin;
  size_t lin_bytes;
  std::vector<double> minmax;
  double range_min, range_max;
  float build_ms;
  bool has_ds;
  std::vector<uint8_t> leap_resident;
};
DatasetState save_dataset(const tvk_ctx* c) {
  return DatasetState{c->L, c->clamp, c->d_pool, c->pool_hold, c->lin_hold, c->cells, c->pst, c->pool_bytes, c->d_lin,
                      c->lin_bytes, c->minmax, c->range_min, c->range_max, c->build_ms, c->has_ds, c->leap_resident};
}
void restore_dataset(tvk_ctx* c, DatasetState&& s) {
  c->L = std::move(s.L);
  c->clamp = s.clamp;
  c->d_pool = s.d_pool;
  c->pool_hold = std::move(s.pool_hold);
  c->lin_hold = std::move(s.lin_hold);
  c->cells = std::move(s.cells);
  c->pst = std::move(s.pst);
  c->pool_bytes = s.pool_bytes;
  c->d_lin = s.d_lin;
  c->lin_bytes = s.lin_bytes;
  c->minmax = std::move(s.minmax);
  c->range_min = s.range_min;
  c->range_max = s.range_max;
  c->build_ms = s.build_ms;
  c->has_ds = s.has_ds;
  c->leap_resident = std::move(s.leap_resident);
  c->cp = tvk_ctx::CompatPool{};
}

}  // namespace

extern "C" {

int tvk_default_params(tvk_params* p) {
  if (!p) return TVK_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  // AbstrRenderer constructor defaults (AbstrRenderer.cpp:71-214)
  p->mode = TVK_RM_1DTRANS;
  p->lighting = 1;
  p->srm = 1.0f;
  p->isovalue = 0.5f;
  p->lod = 0;
  p->tf_scale = 0.0f;
  p->grad_scale = 1.0f;
  for (int i = 0; i < 3; ++i) {
    p->domain_scale[i] = 1.0f;
    p->light_ambient[i] = 1.0f * 0.1f;   // m_cAmbient.xyz * w
    p->light_diffuse[i] = 1.0f;
    p->light_specular[i] = 1.0f;
    p->iso_color[i] = 0.5f;
  }
  p->light_dir[0] = 0.0f; p->light_dir[1] = 0.0f; p->light_dir[2] = -1.0f;
  p->refine_steps = 5;
  p->clamp_to_edge = 0;
  p->ert_eps = 0.0f;
  p->screen_space_error = 1.0f;
  return TVK_OK;
}

int tvk_create(tvk_ctx** out, int device) {
  if (!out) return TVK_ERR_ARG;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) return TVK_ERR_HIP;
  if (device < 0 || device >= n) return TVK_ERR_ARG;
  tvk_ctx* c = new (std::nothrow) tvk_ctx;
  if (!c) return TVK_ERR_NOMEM;
  c->device = device;
  tvk_default_params(&c->P);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return TVK_ERR_HIP;
  }
  for (auto& ev : c->ev) (void)hipEventCreate(&ev);
  // this device's address of the allocation-free render's scratch array
  {
    const hipError_t se = scratch_address(device, &c->scratch);
    if (se != hipSuccess) c->scratch_err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(se);
  }
  if (hipMalloc(&c->d_counter, sizeof(unsigned long long)) != hipSuccess) {
    delete c;
    return TVK_ERR_NOMEM;
  }
  *out = c;
  return TVK_OK;
}

void tvk_destroy(tvk_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_dataset(c);
  if (c->d_tf1d) (void)hipFree(c->d_tf1d);
  if (c->d_tf2d) (void)hipFree(c->d_tf2d);
  if (c->d_tf2f) (void)hipFree(c->d_tf2f);
  for (void* q : {(void*)c->d_leap_meta, (void*)c->d_leap_table, (void*)c->d_leap_bricks, (void*)c->d_leap_resume})
    if (q) (void)hipFree(q);
  if (c->d_tf1f) (void)hipFree(c->d_tf1f);
  if (c->d_tf2cols) (void)hipFree(c->d_tf2cols);
  if (c->d_list) (void)hipFree(c->d_list);
  if (c->d_views) (void)hipFree(c->d_views);
  if (c->d_order) (void)hipFree(c->d_order);
  if (c->d_queues) (void)hipFree(c->d_queues);
  if (c->d_out) (void)hipFree(c->d_out);
  if (c->d_counter) (void)hipFree(c->d_counter);
  if (c->d_touched) (void)hipFree(c->d_touched);
  if (c->d_depth) (void)hipFree(c->d_depth);
  if (c->d_hitpos) (void)hipFree(c->d_hitpos);
  if (c->h_list_pinned) (void)hipHostFree(c->h_list_pinned);
  if (c->h_views_pinned) (void)hipHostFree(c->h_views_pinned);
  if (c->h_order_pinned) (void)hipHostFree(c->h_order_pinned);
  for (auto& ev : c->ev) if (ev) (void)hipEventDestroy(ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* tvk_last_error(const tvk_ctx* c) { return c ? c->err.c_str() : "null context"; }

int tvk_load_volume(tvk_ctx* c, const void* host, int dtype, const uint64_t dims[3], const uint32_t brick[3],
                    uint32_t overlap, int clamp, int max_lods, int keep_linear) {
  if (!c) return TVK_ERR_ARG;
  if (!host) return fail(c, TVK_ERR_ARG, "null volume");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = begin_dataset(c, dtype, dims, brick, overlap, clamp, max_lods);
  if (rc) return rc;
  const size_t bytes = (size_t)(dims[0] * dims[1] * dims[2]) * dtype;
  HIPCHK(c, hipMalloc(&c->d_lin, bytes));
  c->lin_hold = std::shared_ptr<void>(c->d_lin, [](void* p) { (void)hipFree(p); });
  c->lin_bytes = bytes;
  HIPCHK(c, hipMemcpyAsync(c->d_lin, host, bytes, hipMemcpyHostToDevice, c->stream));
  rc = build_pool(c, c->d_lin);
  if (!keep_linear || rc) {
    (void)hipStreamSynchronize(c->stream);
    c->lin_hold.reset();
    c->d_lin = nullptr;
    c->lin_bytes = 0;
  }
  if (rc) free_dataset(c);
  return rc;
}

int tvk_generate_volume(tvk_ctx* c, int kind, uint64_t seed, int dtype, const uint64_t dims[3],
                        const uint32_t brick[3], uint32_t overlap, int clamp, int max_lods, int keep_linear) {
  if (!c) return TVK_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  int rc = begin_dataset(c, dtype, dims, brick, overlap, clamp, max_lods);
  if (rc) return rc;
  const size_t bytes = (size_t)(dims[0] * dims[1] * dims[2]) * dtype;
  HIPCHK(c, hipMalloc(&c->d_lin, bytes));
  c->lin_hold = std::shared_ptr<void>(c->d_lin, [](void* p) { (void)hipFree(p); });
  c->lin_bytes = bytes;
  SplitMix rng(seed);
  hipError_t e = hipSuccess;
  GenEllipsoid* d_e = nullptr;
  unsigned* d_tmp = nullptr;
  if (kind == TVK_GEN_BLOBS) {
    // 12 isotropic Gaussians: centres U[0.2,0.8]^3, sigma U[0.03,0.12], amplitude U[0.3,1.0]
    GenBlobs g{};
    g.n = 12;
    for (int b = 0; b < g.n; ++b) {
      for (int k = 0; k < 3; ++k) g.c[b][k] = (float)rng.uniform(0.2, 0.8);
      double sigma = rng.uniform(0.03, 0.12);
      g.amp[b] = (float)rng.uniform(0.3, 1.0);
      g.inv2s2[b] = (float)(1.0 / (2.0 * sigma * sigma));
    }
    e = hipMalloc(&d_tmp, sizeof(unsigned));
    if (e == hipSuccess) e = launch_blobs(g, dims, dtype, c->d_lin, d_tmp, seed, c->stream);
  } else if (kind == TVK_GEN_FBM_SHELLS) {
    // 3-octave value-noise fBm + 64 hollow shells (radius 0.02-0.1, 2 voxels thick, value 60000)
    GenShells g{};
    g.n = 64;
    g.octaves = 3;
    g.base_freq = 8.0f;
    g.seed = rng.next();
    for (int s = 0; s < g.n; ++s) {
      for (int k = 0; k < 3; ++k) g.c[s][k] = (float)rng.uniform(0.15, 0.85);
      g.r[s] = (float)rng.uniform(0.02, 0.1);
    }
    g.half_thickness_vox = 1.0f;
    g.shell_value = dtype == TVK_U16 ? 60000.f : 233.f;
    e = launch_fbm_shells(g, dims, dtype, c->d_lin, c->stream);
  } else if (kind == TVK_GEN_ELLIPSOID_SDF || kind == TVK_GEN_PHANTOM) {
    const int ne = kind == TVK_GEN_ELLIPSOID_SDF ? 500 : 40;
    std::vector<GenEllipsoid> el((size_t)ne);
    for (int k = 0; k < ne; ++k) {
      GenEllipsoid& q = el[(size_t)k];
      if (kind == TVK_GEN_ELLIPSOID_SDF) {
        for (int a = 0; a < 3; ++a) { q.c[a] = (float)rng.uniform(0.1, 0.9); q.a[a] = (float)rng.uniform(0.01, 0.08); }
        q.value = 0.f;
      } else {
        // nested: each ellipsoid lies inside the previous one
        double s = 0.45 * std::pow(0.93, k);
        for (int a = 0; a < 3; ++a) {
          q.c[a] = (float)(0.5 + rng.uniform(-0.15, 0.15) * s);
          q.a[a] = (float)(s * rng.uniform(0.7, 1.0));
        }
        q.value = (float)std::floor(rng.uniform(0.0, 4000.0));
      }
      q.amin = std::min(q.a[0], std::min(q.a[1], q.a[2]));
    }
    e = hipMalloc(&d_e, el.size() * sizeof(GenEllipsoid));
    if (e == hipSuccess) e = hipMemcpyAsync(d_e, el.data(), el.size() * sizeof(GenEllipsoid), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
      if (kind == TVK_GEN_ELLIPSOID_SDF) {
        if (dtype != TVK_U8) { (void)hipFree(d_e); return fail(c, TVK_ERR_ARG, "SDF volumes are uint8"); }
        e = launch_ellipsoid_sdf(d_e, ne, dims, c->d_lin, c->stream);
      } else {
        if (dtype != TVK_U16) { (void)hipFree(d_e); return fail(c, TVK_ERR_ARG, "phantom volumes are uint16"); }
        e = launch_phantom(d_e, ne, dims, c->d_lin, 20.f, rng.next(), c->stream);
      }
    }
  } else {
    return fail(c, TVK_ERR_ARG, "unknown generator");
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_e) (void)hipFree(d_e);
  if (d_tmp) (void)hipFree(d_tmp);
  if (e != hipSuccess) { free_dataset(c); return hipfail(c, "generate", e); }
  rc = build_pool(c, c->d_lin);
  if (!keep_linear || rc) {
    (void)hipStreamSynchronize(c->stream);
    c->lin_hold.reset();
    c->d_lin = nullptr;
    c->lin_bytes = 0;
  }
  if (rc) free_dataset(c);
  return rc;
}

int tvk_rebrick(tvk_ctx* c, const uint32_t brick[3], uint32_t overlap, int max_lods) {
  if (!c || !brick) return TVK_ERR_ARG;
  if (!c->has_ds || !c->d_lin) return fail(c, TVK_ERR_STATE, "rebricking needs the linear volume (keep_linear)");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // The current dataset is held until the new pool is complete: on any failure
  // (a rejected size, an allocation) the context keeps rendering it.  The two
  // pools coexist while the new one is built.
  DatasetState old = save_dataset(c);
  const uint64_t dims[3] = {c->L.dims[0], c->L.dims[1], c->L.dims[2]};
  int rc = begin_dataset(c, c->L.bpv, dims, brick, overlap, c->clamp, max_lods);
  if (rc) return rc;   // nothing was changed
  c->lin_hold = old.lin_hold;   // the linear volume outlives the old pool
  c->d_lin = old.d_lin;
  c->lin_bytes = old.lin_bytes;
  c->leap_resident.clear();
  rc = build_pool(c, c->d_lin);
  if (rc) {
    const std::string msg = c->err;
    free_dataset(c);
    restore_dataset(c, std::move(old));
    c->err = msg;
  }
  return rc;
}

int tvk_crop_dataset(tvk_ctx* c, int keep_old, float plane_used[4]) {
  (void)keep_old;   // no file is replaced: the cropped dataset replaces this context's
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (c->ingest_pending) return fail(c, TVK_ERR_STATE, "not every brick was uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const Layout& L = c->L;
  // GLRenderer::CropDataset (GLRenderer.cpp:2953-2971): the plane without the
  // viewing transformation, ExtendedPlane::Transform(inverse(rotation * translation))
  // = PLANE * M (the row vector times inverse(M) transposed, divided by |xyz|)
  float pm[4];
  {
    const Mat4 inv = inverse(c->V.rot * c->V.trans);
    const Mat4 ii = inverse(inv);
    Mat4 it;
    for (int r = 0; r < 4; ++r)
      for (int q = 0; q < 4; ++q) it.at(r, q) = ii.at(q, r);
    row_mul(c->V.clip, it, pm);
    const float l = std::sqrt(pm[0] * pm[0] + pm[1] * pm[1] + pm[2] * pm[2]);
    for (float& v : pm) v /= l;
  }
  // UVFDataset::Crop (uvfDataset.cpp:1470-1486): scale invariance,
  // transformIT(Scaling(scale / max scale * domain / max domain)) = the row vector
  // times that matrix, divided by |xyz|
  {
    const double dmax = (double)std::max(L.dims[0], std::max(L.dims[1], L.dims[2]));
    const float smax = std::max(L.scale[0], std::max(L.scale[1], L.scale[2]));
    float sc[3];
    for (int k = 0; k < 3; ++k) sc[k] = (L.scale[k] / smax) * (float)((double)L.dims[k] / dmax);
    const float t[4] = {pm[0] * sc[0], pm[1] * sc[1], pm[2] * sc[2], pm[3]};
    const float l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int k = 0; k < 4; ++k) pm[k] = t[k] / l;
  }
  if (plane_used)
    for (int k = 0; k < 4; ++k) plane_used[k] = pm[k];
  // the LOD-0 volume: the kept linear copy, or flattened from the resident bricks
  std::shared_ptr<void> lin = c->lin_hold;
  void* d_lin = c->d_lin;
  const size_t bytes = (size_t)(L.dims[0] * L.dims[1] * L.dims[2]) * (size_t)L.bpv;
  if (!d_lin) {
    if (budgeted(c)) return fail(c, TVK_ERR_UNSUPPORTED, "cropping a budgeted dataset needs its linear volume");
    HIPCHK(c, hipMalloc(&d_lin, bytes));
    lin = std::shared_ptr<void>(d_lin, [](void* p) { (void)hipFree(p); });
    const LodInfo& l0 = L.lods[0];
    std::vector<GatherBrick> gb((size_t)(l0.bricks[0] * l0.bricks[1] * l0.bricks[2]));
    for (size_t i = 0; i < gb.size(); ++i) {
      const TocBrick& t = L.toc[l0.offset + i];
      for (int k = 0; k < 3; ++k) { gb[i].c[k] = (uint32_t)t.coords[k]; gb[i].n[k] = t.size[k]; }
      gb[i].offset = t.offset;
    }
    GatherBrick* d_gb = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_gb, gb.size() * sizeof(GatherBrick)));
    std::shared_ptr<void> gb_hold(d_gb, [](void* p) { (void)hipFree(p); });
    HIPCHK(c, hipMemcpyAsync(d_gb, gb.data(), gb.size() * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream));
    const uint32_t nb[3] = {(uint32_t)l0.bricks[0], (uint32_t)l0.bricks[1], (uint32_t)l0.bricks[2]};
    const uint32_t core[3] = {L.brick[0] - 2 * L.overlap, L.brick[1] - 2 * L.overlap, L.brick[2] - 2 * L.overlap};
    HIPCHK(c, launch_flatten(c->d_pool, L.bpv, d_gb, nb, core, L.overlap, L.dims, d_lin, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  HIPCHK(c, launch_crop(d_lin, L.bpv, L.dims, pm, c->stream));
  // RAWConverter::ConvertRAWDataset of the cropped volume with the same bricking
  const uint32_t brick[3] = {L.brick[0], L.brick[1], L.brick[2]};
  const uint32_t ov = L.overlap;
  const int lods = (int)(L.toc.back().coords[3] + 1);
  const uint64_t dims[3] = {L.dims[0], L.dims[1], L.dims[2]};
  const int bpv = L.bpv, clamp = c->clamp;
  int rc = begin_dataset(c, bpv, dims, brick, ov, clamp, lods);
  if (rc) return rc;
  c->lin_hold = lin;
  c->d_lin = d_lin;
  c->lin_bytes = bytes;
  c->leap_resident.clear();
  rc = build_pool(c, d_lin);
  if (rc) free_dataset(c);
  return rc;
}

int tvk_set_lod_filter(tvk_ctx* c, int median) {
  if (!c) return TVK_ERR_ARG;
  c->lod_median = median ? 1 : 0;
  return TVK_OK;
}

int tvk_set_dispatch_order(tvk_ctx* c, int order) {
  if (!c) return TVK_ERR_ARG;
  if ((order & ~(3 | TVK_DISPATCH_QUEUES)) != 0 || (order & 3) > TVK_DISPATCH_ROWS)
    return fail(c, TVK_ERR_ARG, "dispatch order must be TVK_DISPATCH_RASTER, _GROUPS or _ROWS, or'ed with _QUEUES");
  c->dispatch = order;
  return TVK_OK;
}

int tvk_set_compat(tvk_ctx* c, int use_only_power_of_two, int downsample_to_8bits) {
  if (!c) return TVK_ERR_ARG;
  c->compat_pow2 = use_only_power_of_two ? 1 : 0;
  c->compat_8bit = downsample_to_8bits ? 1 : 0;
  return TVK_OK;
}

int tvk_share_dataset(tvk_ctx* dst, const tvk_ctx* src) {
  if (!dst || !src || dst == src) return TVK_ERR_ARG;
  if (!src->has_ds) return fail(dst, TVK_ERR_STATE, "source context has no dataset");
  if (src->device != dst->device) return fail(dst, TVK_ERR_ARG, "contexts are on different devices");
  HIPCHK(dst, hipStreamSynchronize(src->stream));   // the source's bricking is complete
  free_dataset(dst);
  dst->L = src->L;
  dst->clamp = src->clamp;
  dst->d_pool = src->d_pool;
  dst->pool_hold = src->pool_hold;
  dst->pool_bytes = src->pool_bytes;
  dst->d_lin = src->d_lin;
  dst->lin_hold = src->lin_hold;
  dst->cells = src->cells;
  dst->pst = src->pst;
  dst->ingest_missing = src->ingest_missing;
  dst->ingest_pending = src->ingest_pending;
  dst->lin_bytes = src->lin_bytes;
  dst->minmax = src->minmax;
  dst->range_min = src->range_min;
  dst->range_max = src->range_max;
  dst->has_ds = true;
  return TVK_OK;
}

int tvk_download_volume(tvk_ctx* c, void* host, size_t bytes) {
  if (!c || !host) return TVK_ERR_ARG;
  if (!c->d_lin) return fail(c, TVK_ERR_STATE, "linear volume was not kept");
  if (bytes != c->lin_bytes) return fail(c, TVK_ERR_ARG, "size mismatch");
  HIPCHK(c, hipMemcpyAsync(host, c->d_lin, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TVK_OK;
}

int tvk_dataset_info(tvk_ctx* c, tvk_dataset_desc* o) {
  if (!c || !o) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  std::memset(o, 0, sizeof(*o));
  o->lod_count = (int)(c->L.toc.back().coords[3] + 1);
  o->brick_count = c->L.toc.size();
  o->pool_bytes = c->pool_bytes;
  o->brick_bytes = c->L.ref_elems * (uint64_t)c->L.bpv;
  o->range_min = c->range_min;
  o->range_max = c->range_max;
  o->bit_width = bit_width(c);
  for (int i = 0; i < 3; ++i) { o->dims[i] = c->L.dims[i]; o->brick_size[i] = c->L.brick[i]; }
  o->overlap = c->L.overlap;
  o->build_ms = c->build_ms;
  return TVK_OK;
}

int tvk_lod_info(tvk_ctx* c, int lod, uint64_t size[3], uint64_t bricks[3], uint64_t* offset, double aspect[3]) {
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (lod < 0 || (size_t)lod >= c->L.lods.size()) return fail(c, TVK_ERR_RANGE, "lod out of range");
  const LodInfo& l = c->L.lods[(size_t)lod];
  for (int i = 0; i < 3; ++i) {
    if (size) size[i] = l.size[i];
    if (bricks) bricks[i] = l.bricks[i];
    if (aspect) aspect[i] = l.aspect[i];
  }
  if (offset) *offset = l.offset;
  return TVK_OK;
}

int tvk_brick_info(tvk_ctx* c, uint64_t i, uint64_t coords[4], uint32_t size[3], uint64_t* off) {
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (i >= c->L.toc.size()) return fail(c, TVK_ERR_RANGE, "unknown brick");
  const TocBrick& t = c->L.toc[(size_t)i];
  for (int k = 0; k < 4; ++k) if (coords) coords[k] = t.coords[k];
  for (int k = 0; k < 3; ++k) if (size) size[k] = t.size[k];
  if (off) *off = t.ref_offset * (uint64_t)c->L.bpv;   // byte offset in the tvk_get_pool layout
  return TVK_OK;
}

int tvk_brick_md(tvk_ctx* c, int lod, uint64_t index, float center[3], float extents[3], uint32_t nvox[3]) {
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  auto it = c->L.table.find(BrickKey(0, (size_t)lod, (size_t)index));
  if (it == c->L.table.end()) return fail(c, TVK_ERR_RANGE, "unknown brick");  // BrickedDataset.cpp:85
  for (int k = 0; k < 3; ++k) {
    if (center) center[k] = it->second.center[k];
    if (extents) extents[k] = it->second.extents[k];
    if (nvox) nvox[k] = it->second.n_voxels[k];
  }
  return TVK_OK;
}

int tvk_brick_minmax(tvk_ctx* c, uint64_t i, double mm[2]) {
  if (!c || !mm) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (i >= c->L.toc.size()) return fail(c, TVK_ERR_RANGE, "unknown brick");
  mm[0] = c->minmax[2 * (size_t)i];
  mm[1] = c->minmax[2 * (size_t)i + 1];
  return TVK_OK;
}

}  // extern "C"

namespace {
// Copies bricks [first, first+n) of the TOC to the host in the reference layout.
int download_ref(tvk_ctx* c, size_t first, size_t n, void* host, size_t bytes) {
  const Layout& L = c->L;
  std::vector<GatherBrick> gb(n);
  std::vector<uint64_t> ro(n);
  uint64_t total = 0;
  uint32_t maxv = 0;
  for (size_t k = 0; k < n; ++k) {
    const TocBrick& t = L.toc[first + k];
    for (int a = 0; a < 3; ++a) { gb[k].c[a] = (uint32_t)t.coords[a]; gb[k].n[a] = t.size[a]; }
    gb[k].offset = t.offset;
    ro[k] = total;
    total += t.real;
    maxv = std::max<uint32_t>(maxv, (uint32_t)t.real);
  }
  if (bytes != total * (size_t)L.bpv) return fail(c, TVK_ERR_ARG, "size mismatch");
  GatherBrick* d_gb = nullptr;
  uint64_t* d_ro = nullptr;
  void* d_out = nullptr;
  hipError_t e = hipMalloc(&d_gb, n * sizeof(GatherBrick));
  if (e == hipSuccess) e = hipMalloc(&d_ro, n * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc(&d_out, bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(d_gb, gb.data(), n * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_ro, ro.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = launch_unswizzle(c->d_pool, L.bpv, d_gb, d_ro, (uint32_t)n, maxv, d_out, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(host, d_out, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->stream);
  if (d_gb) (void)hipFree(d_gb);
  if (d_ro) (void)hipFree(d_ro);
  if (d_out) (void)hipFree(d_out);
  if (e != hipSuccess) return hipfail(c, "brick download", e);
  return TVK_OK;
}
}  // namespace

extern "C" {

int tvk_get_brick(tvk_ctx* c, uint64_t i, void* host, size_t bytes) {
  if (!c || !host) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (i >= c->L.toc.size()) return fail(c, TVK_ERR_RANGE, "unknown brick");
  if (budgeted(c)) return pool_download(c, (size_t)i, 1, host, bytes);   // the host store holds every brick
  if (c->ingest_pending) return fail(c, TVK_ERR_STATE, "not every brick was uploaded");
  return download_ref(c, (size_t)i, 1, host, bytes);
}

int tvk_get_pool(tvk_ctx* c, void* host, size_t bytes) {
  if (!c || !host) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (budgeted(c)) return pool_download(c, 0, c->L.toc.size(), host, bytes);
  if (c->ingest_pending) return fail(c, TVK_ERR_STATE, "not every brick was uploaded");
  return download_ref(c, 0, c->L.toc.size(), host, bytes);
}

int tvk_max_gradient(tvk_ctx* c, float* out) {
  if (!c || !out) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (c->ingest_pending) return fail(c, TVK_ERR_STATE, "not every brick was uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->L.overlap >= 1) {
    // Histogram2DDataBlock::ComputeTemplate at level 0 (what RAWConverter builds,
    // RAWConverter.cpp:655-660): the LOD-0 bricks' core voxels, neighbours from the
    // bricks' own overlap; a budgeted pool pages the bricks through its slots
    const LodInfo& l0 = c->L.lods[0];
    const uint32_t nb = (uint32_t)(l0.bricks[0] * l0.bricks[1] * l0.bricks[2]);
    const uint32_t gx = 16;
    const uint32_t group = budgeted(c) ? c->pst->nslots : nb;
    GatherBrick* d_gb = nullptr;
    double* d_part = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_gb, (size_t)group * sizeof(GatherBrick)));
    std::shared_ptr<void> gb_hold(d_gb, [](void* p) { (void)hipFree(p); });
    HIPCHK(c, hipMalloc((void**)&d_part, (size_t)group * gx * sizeof(double)));
    std::shared_ptr<void> part_hold(d_part, [](void* p) { (void)hipFree(p); });
    double m = 0.0;
    std::vector<double> h((size_t)group * gx);
    for (uint32_t first = 0; first < nb; first += group) {
      const uint32_t cnt = std::min(group, nb - first);
      std::vector<GatherBrick> gb(cnt);
      std::vector<uint32_t> need(cnt);
      for (uint32_t k = 0; k < cnt; ++k) need[k] = (uint32_t)(l0.offset + first + k);
      std::vector<int32_t> slot;
      if (budgeted(c)) {
        const int rc = pool_acquire_lru(c, need, slot);
        if (rc) return rc;
      }
      for (uint32_t k = 0; k < cnt; ++k) {
        const TocBrick& t = c->L.toc[need[k]];
        for (int a = 0; a < 3; ++a) { gb[k].c[a] = (uint32_t)t.coords[a]; gb[k].n[a] = t.size[a]; }
        gb[k].offset = budgeted(c) ? (uint64_t)slot[k] * c->pst->slot_elems : t.offset;
      }
      HIPCHK(c, hipMemcpyAsync(d_gb, gb.data(), cnt * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, launch_brick_max_gradient(c->d_pool, c->L.bpv, d_gb, cnt, c->L.overlap, gx, d_part, c->stream));
      HIPCHK(c, hipMemcpyAsync(h.data(), d_part, (size_t)cnt * gx * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (size_t k = 0; k < (size_t)cnt * gx; ++k) m = std::max(m, h[k]);
    }
    *out = (float)m;
    return TVK_OK;
  }
  if (!c->d_lin) return fail(c, TVK_ERR_STATE, "without overlap the gradient needs the kept linear volume");
  const uint32_t blocks = 4096;
  double* d = nullptr;
  HIPCHK(c, hipMalloc(&d, blocks * sizeof(double)));
  std::vector<double> h(blocks);
  hipError_t e = launch_max_gradient(c->d_lin, c->L.bpv, c->L.dims, c->clamp, d, blocks, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, blocks * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return hipfail(c, "max_gradient", e);
  double m = 0;
  for (double v : h) m = std::max(m, v);
  *out = (float)m;
  return TVK_OK;
}

int tvk_set_tf1d(tvk_ctx* c, const uint8_t* rgba8, int n) {
  if (!c) return TVK_ERR_ARG;
  if (!rgba8 || n < 2 || n > 4096) return fail(c, TVK_ERR_ARG, "1D TF needs 2..4096 RGBA8 entries");
  c->tf1d.assign(rgba8, rgba8 + 4 * (size_t)n);
  c->tf1d_n = n;
  c->tfst_dirty = true;
  int rc = ensure(c, &c->d_tf1d, &c->tf1d_cap, 4 * (size_t)n);
  if (!rc) rc = ensure(c, &c->d_tf1f, &c->tf1f_cap, (size_t)n);
  if (rc) return rc;
  // texture1D of the RGBA8 table returns byte/255 (GL_NEAREST): expand once
  const std::vector<float4> f = unorm_table(rgba8, (size_t)n);
  HIPCHK(c, hipMemcpyAsync(c->d_tf1d, c->tf1d.data(), 4 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tf1f, f.data(), f.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TVK_OK;
}

int tvk_set_tf2d(tvk_ctx* c, const uint8_t* rgba8, int w, int h) {
  if (!c) return TVK_ERR_ARG;
  if (!rgba8 || w < 2 || h < 2 || w > 4096 || h > 4096) return fail(c, TVK_ERR_ARG, "bad 2D TF size");
  c->tf2d.assign(rgba8, rgba8 + 4 * (size_t)w * h);
  c->tf2d_w = w;
  c->tf2d_h = h;
  c->tfst_dirty = true;
  int rc = ensure(c, &c->d_tf2d, &c->tf2d_cap, 4 * (size_t)w * h);
  if (!rc) rc = ensure(c, &c->d_tf2f, &c->tf2f_cap, (size_t)w * h);
  if (!rc && !c->d_tf2cols) { size_t cap = 0; rc = ensure(c, &c->d_tf2cols, &cap, 64); }
  if (rc) return rc;
  // texture2D of an RGBA8 table returns byte/255 (GL_NEAREST): expand once
  const std::vector<float4> f = unorm_table(rgba8, (size_t)w * h);
  uint64_t cols[64] = {0};
  for (int x = 0; x < w; ++x)
    for (int y = 0; y < h; ++y)
      if (rgba8[4 * ((size_t)y * w + x) + 3]) { cols[x >> 6] |= 1ull << (x & 63); break; }
  HIPCHK(c, hipMemcpyAsync(c->d_tf2d, c->tf2d.data(), 4 * (size_t)w * h, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tf2f, f.data(), f.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tf2cols, cols, sizeof(cols), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TVK_OK;
}

int tvk_set_view_matrices(tvk_ctx* c, const float view[16], const float rot[16], const float trans[16],
                          const float proj[16], uint32_t w, uint32_t h, float fov, float znear, float zfar) {
  if (!c || !view || !proj) return TVK_ERR_ARG;
  if (w == 0 || h == 0 || w > 32767 || h > 32767) return fail(c, TVK_ERR_ARG, "bad viewport size");
  c->V.view = Mat4::from(view);
  c->V.proj = Mat4::from(proj);
  c->V.rot = rot ? Mat4::from(rot) : Mat4();
  c->V.trans = trans ? Mat4::from(trans) : Mat4();
  c->V.W = w;
  c->V.H = h;
  c->V.fov = fov;
  c->V.znear = znear;
  c->V.zfar = zfar;
  c->has_view = true;
  c->proj_from_fov = false;
  c->leap_eye = point_mul(F3(0, 0, 0), inverse(c->V.view));   // camera position of the view matrix
  return TVK_OK;
}

int tvk_set_camera(tvk_ctx* c, const float eye[3], const float at[3], const float up[3], float fov, float znear,
                   float zfar, uint32_t w, uint32_t h) {
  if (!c || !eye || !at || !up) return TVK_ERR_ARG;
  if (w == 0 || h == 0 || w > 32767 || h > 32767) return fail(c, TVK_ERR_ARG, "bad viewport size");
  // GLRenderer::ComputeViewAndProjection (GLRenderer.cpp:889-920)
  Mat4 v = look_at(F3(eye[0], eye[1], eye[2]), F3(at[0], at[1], at[2]), F3(up[0], up[1], up[2]));
  Mat4 p = perspective(fov, (float)w / (float)h, znear, zfar);
  c->V.view = v;
  c->V.proj = p;
  c->V.rot = Mat4();
  c->V.trans = Mat4();
  c->V.W = w;
  c->V.H = h;
  c->V.fov = fov;
  c->V.znear = znear;
  c->V.zfar = zfar;
  c->has_view = true;
  c->proj_from_fov = true;
  c->leap_eye = F3(eye[0], eye[1], eye[2]);
  return TVK_OK;
}

int tvk_set_clip_plane(tvk_ctx* c, const float plane[4], int enabled) {
  if (!c) return TVK_ERR_ARG;
  if (plane) {
    if (!(plane[0] * plane[0] + plane[1] * plane[1] + plane[2] * plane[2] > 0.0f))
      return fail(c, TVK_ERR_ARG, "clip plane normal is zero");
    for (int i = 0; i < 4; ++i) c->V.clip[i] = plane[i];
  }
  c->V.clip_on = enabled != 0;
  return TVK_OK;
}

int tvk_set_rotation(tvk_ctx* c, const float rot[16]) {
  if (!c || !rot) return TVK_ERR_ARG;
  c->V.rot = Mat4::from(rot);
  return TVK_OK;
}

int tvk_get_view_matrices(tvk_ctx* c, float view[16], float proj[16]) {
  if (!c) return TVK_ERR_ARG;
  for (int i = 0; i < 16; ++i) {
    if (view) view[i] = c->V.view.a[i];
    if (proj) proj[i] = c->V.proj.a[i];
  }
  return TVK_OK;
}

int tvk_set_params(tvk_ctx* c, const tvk_params* p) {
  if (!c || !p) return TVK_ERR_ARG;
  if (p->mode < 0 || p->mode > 2) return fail(c, TVK_ERR_ARG, "invalid render mode");
  if (!(p->srm > 0.0f)) return fail(c, TVK_ERR_ARG, "sample rate modifier must be > 0");
  if (p->refine_steps < 0 || p->refine_steps > 64) return fail(c, TVK_ERR_ARG, "refine_steps out of range");
  if (p->ert_eps < 0.0f || p->ert_eps >= 1.0f) return fail(c, TVK_ERR_ARG, "ert_eps must be in [0,1)");
  if (p->mode != c->P.mode) c->tfst_dirty = true;
  c->P = *p;
  return TVK_OK;
}

int tvk_get_params(tvk_ctx* c, tvk_params* p) {
  if (!c || !p) return TVK_ERR_ARG;
  *p = c->P;
  return TVK_OK;
}

int tvk_set_count_samples(tvk_ctx* c, int on) {
  if (!c) return TVK_ERR_ARG;
  c->count = on != 0;
  return TVK_OK;
}

}  // extern "C"

namespace tvk {

int frame_lod(tvk_ctx* c, const ViewState& V, int* lod) {
  int l = c->P.lod;
  const int kept = (int)(c->L.toc.back().coords[3] + 1);
  // ComputeMinLODForCurrentView, limited to the levels the pool holds
  if (l < 0) l = std::min(min_lod_for_view(c->L, V, c->P.screen_space_error), kept - 1);
  if (l >= kept) return fail(c, TVK_ERR_RANGE, "LOD " + std::to_string(l) + " is not in the brick pool");
  *lod = l;
  return TVK_OK;
}

int plan_params(tvk_ctx* c, int lod, PlanParams& pp) {
  pp.pow2 = c->compat_pow2 != 0;
  pp.mode = c->P.mode;
  pp.lod = lod;
  pp.srm = c->P.srm;
  pp.isovalue = c->P.isovalue;
  if (c->P.mode == TVK_RM_1DTRANS) {
    if (c->tf1d_n == 0) return fail(c, TVK_ERR_STATE, "no 1D transfer function");
    tf1d_limits(c->tf1d.data(), c->tf1d_n, pp.lim);
    pp.tf_n = c->tf1d_n;
  } else if (c->P.mode == TVK_RM_2DTRANS) {
    if (c->tf2d_w == 0) return fail(c, TVK_ERR_STATE, "no 2D transfer function");
    tf2d_limits(c->tf2d.data(), c->tf2d_w, c->tf2d_h, pp.lim);
    pp.tf_n = c->tf2d_w;
  } else {
    pp.tf_n = c->tf1d_n > 0 ? c->tf1d_n : 256;
  }
  return TVK_OK;
}

float normalized_isovalue(const tvk_ctx* c) {
  if (downsampled(c)) {
    // MathTools::lerp(iso, 0.f, float(mx), 0.f, 1.f) with mx the range maximum
    // (the TF size for an empty range): float operands, a double quotient
    const double mx = c->range_min > c->range_max ? (double)c->tf1d_n : c->range_max;
    const float fmx = (float)mx;
    return (float)(0.0f + (double)(c->P.isovalue - 0.0f) * (1.0 / (double)(fmx - 0.0f)));
  }
  return (float)((double)c->P.isovalue / (double)(1 << bit_width(c)));
}

int ensure_compat(tvk_ctx* c) {
  if (!compat_active(c) || compat_ready(c)) return TVK_OK;
  const Layout& L = c->L;
  tvk_ctx::CompatPool cp;
  cp.pow2 = c->compat_pow2;
  cp.down8 = downsampled(c) ? 1 : 0;
  cp.dup = c->P.clamp_to_edge ? 0 : 1;
  cp.bpv = render_bpv(c);
  const size_t nb = L.toc.size();
  cp.offset.resize(nb);
  cp.dims.resize(3 * nb);
  std::vector<GatherBrick> src(nb), dst(nb);
  uint64_t elems = 0, max_elems = 0;
  const int lx = block_lx(cp.bpv);
  for (size_t i = 0; i < nb; ++i) {
    for (int k = 0; k < 3; ++k) {
      const uint32_t n = L.toc[i].size[k];
      src[i].c[k] = dst[i].c[k] = (uint32_t)L.toc[i].coords[k];
      src[i].n[k] = n;
      dst[i].n[k] = cp.pow2 ? next_pow2(n) : n;
      cp.dims[3 * i + k] = dst[i].n[k];
    }
    src[i].offset = L.toc[i].offset;
    dst[i].offset = elems;
    cp.offset[i] = elems;
    const uint64_t e = padded_count(dst[i].n, lx);
    elems += e;
    max_elems = std::max(max_elems, e);
  }
  const size_t bytes = (size_t)elems * (size_t)cp.bpv;
  c->cp = tvk_ctx::CompatPool{};   // the old textures go before the new ones are made
  HIPCHK(c, hipMalloc(&cp.d, std::max<size_t>(bytes, 1)));
  cp.hold = std::shared_ptr<void>(cp.d, [](void* p) { (void)hipFree(p); });
  GatherBrick* d_gb = nullptr;
  HIPCHK(c, hipMalloc(&d_gb, 2 * nb * sizeof(GatherBrick)));
  hipError_t e = hipMemcpyAsync(d_gb, src.data(), nb * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_gb + nb, dst.data(), nb * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = launch_compat(c->d_pool, L.bpv, d_gb, cp.d, cp.bpv, d_gb + nb, (uint32_t)nb, max_elems, c->range_min,
                      c->range_max, cp.dup, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_gb);
  if (e != hipSuccess) return fail(c, TVK_ERR_HIP, std::string("compatibility textures: ") + hipGetErrorString(e));
  cp.valid = true;
  c->cp = std::move(cp);
  return TVK_OK;
}

// GLRenderer::CalculateScaling (GLRenderer.cpp:1813-1819)
float tf_scale_of(const tvk_ctx* c) {
  if (c->P.tf_scale > 0.0f) return c->P.tf_scale;
  if (downsampled(c)) return 1.0f;   // the 8-bit textures span the TF already
  const double fMaxValue = c->range_max > 0 ? c->range_max : (double)std::max(c->tf1d_n, 1);
  const uint32_t iMaxValue = (uint32_t)(1u << bit_width(c)) - 1u;
  return float(iMaxValue / fMaxValue);
}

// Per-entry alpha and colour of the active TF (a 2D TF contributes, per value
// column, its largest alpha and that row's colour), with prefix sums and
// sparse max tables so that any value range is summarised in O(1).
void build_tf_stats(tvk_ctx* c) {
  tvk_ctx::TfStats& T = c->tfst;
  const bool two = c->P.mode == TVK_RM_2DTRANS;
  const int n = two ? c->tf2d_w : c->tf1d_n;
  T = tvk_ctx::TfStats{};
  T.n = n;
  c->tfst_dirty = false;
  if (n <= 0) return;
  std::vector<float> a((size_t)n), rgb(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const uint8_t* e = nullptr;
    if (two) {
      int best = 0;
      for (int y = 0; y < c->tf2d_h; ++y)
        if (c->tf2d[4 * ((size_t)y * n + i) + 3] > c->tf2d[4 * ((size_t)best * n + i) + 3]) best = y;
      e = &c->tf2d[4 * ((size_t)best * n + i)];
    } else {
      e = &c->tf1d[4 * (size_t)i];
    }
    a[(size_t)i] = e[3] / 255.0f;
    for (int k = 0; k < 3; ++k) rgb[3 * (size_t)i + k] = e[k] / 255.0f;
  }
  T.pa.assign((size_t)n + 1, 0.0);
  T.pnz.assign((size_t)n + 1, 0.0);
  T.ptv.assign((size_t)n, 0.0);
  T.prgb.assign((size_t)n, 0.0);
  std::vector<float> slope((size_t)std::max(n - 1, 1), 0.0f);
  for (int i = 0; i < n; ++i) {
    T.pa[(size_t)i + 1] = T.pa[(size_t)i] + a[(size_t)i];
    T.pnz[(size_t)i + 1] = T.pnz[(size_t)i] + (a[(size_t)i] > 0.0f ? 1.0 : 0.0);
    if (i + 1 < n) {
      const float d = std::fabs(a[(size_t)i + 1] - a[(size_t)i]);
      float dr = 0.0f;
      for (int k = 0; k < 3; ++k) dr += std::fabs(rgb[3 * ((size_t)i + 1) + k] - rgb[3 * (size_t)i + k]);
      slope[(size_t)i] = d;
      T.ptv[(size_t)i + 1] = T.ptv[(size_t)i] + d;
      T.prgb[(size_t)i + 1] = T.prgb[(size_t)i] + dr / 3.0f;
    }
  }
  auto sparse = [](const std::vector<float>& v, std::vector<std::vector<float>>& t) {
    t.clear();
    t.push_back(v);
    for (size_t w = 1; 2 * w <= v.size(); w *= 2) {
      const std::vector<float>& p = t.back();
      std::vector<float> q(v.size() - 2 * w + 1);
      for (size_t i = 0; i < q.size(); ++i) q[i] = std::max(p[i], p[i + w]);
      t.push_back(q);
    }
  };
  sparse(a, T.amax);
  sparse(slope, T.smax);
  // device layout: prefix sums, then the sparse tables with rows of n entries
  T.levels = (int)T.amax.size();
  T.flat.assign(tfst_size(n, T.levels), 0.0f);
  for (int i = 0; i <= n; ++i) {
    T.flat[(size_t)tfst_pa(n) + i] = (float)T.pa[(size_t)i];
    T.flat[(size_t)tfst_pnz(n) + i] = (float)T.pnz[(size_t)i];
  }
  for (int i = 0; i < n; ++i) {
    T.flat[(size_t)tfst_ptv(n) + i] = (float)T.ptv[(size_t)i];
    T.flat[(size_t)tfst_prgb(n) + i] = (float)T.prgb[(size_t)i];
  }
  for (int l = 0; l < T.levels; ++l)
    std::copy(T.amax[(size_t)l].begin(), T.amax[(size_t)l].end(), T.flat.begin() + tfst_amax(n, l));
  for (int l = 0; l < (int)T.smax.size() && l < T.levels; ++l)
    std::copy(T.smax[(size_t)l].begin(), T.smax[(size_t)l].end(), T.flat.begin() + tfst_smax(n, T.levels, l));
}

float range_max(const std::vector<std::vector<float>>& t, int lo, int hi) {   // inclusive, lo <= hi
  int k = 0;
  while ((2 << k) <= hi - lo + 1) ++k;
  return std::max(t[(size_t)k][(size_t)lo], t[(size_t)k][(size_t)(hi - (1 << k) + 1)]);
}

int learn_prepare(tvk_ctx* c) {
  if (budgeted(c)) return fail(c, TVK_ERR_UNSUPPORTED, "the learned step needs every brick resident (pool budget 0)");
  if (c->tfst_dirty) {
    build_tf_stats(c);
    const std::vector<float>& F = c->tfst.flat;
    if (!F.empty()) {
      int rc = ensure(c, &c->d_tfst, &c->tfst_cap, F.size());
      if (rc) return rc;
      HIPCHK(c, hipMemcpyAsync(c->d_tfst, F.data(), F.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
  }
  if (c->tfst.n <= 0) return fail(c, TVK_ERR_STATE, "the learned step needs a transfer function");
  return ensure_cells(c);
}

int ensure_cells(tvk_ctx* c) {
  if (budgeted(c)) return fail(c, TVK_ERR_UNSUPPORTED, "the cell grid needs every brick resident (pool budget 0)");
  if (c->ingest_pending) return fail(c, TVK_ERR_STATE, "not every brick was uploaded");
  if (!c->cells) return fail(c, TVK_ERR_STATE, "no dataset");
  CellGrid& G = *c->cells;
  if (G.d) return TVK_OK;
  const Layout& L = c->L;
  std::vector<GatherBrick> gb(L.toc.size());
  G.base.assign(L.toc.size(), 0);
  uint64_t total = 0;
  uint32_t max_cells = 0;
  for (size_t i = 0; i < L.toc.size(); ++i) {
    for (int k = 0; k < 3; ++k) { gb[i].c[k] = (uint32_t)L.toc[i].coords[k]; gb[i].n[k] = L.toc[i].size[k]; }
    gb[i].offset = L.toc[i].offset;
    const uint32_t nc = cell_count(gb[i].n[0]) * cell_count(gb[i].n[1]) * cell_count(gb[i].n[2]);
    G.base[i] = (uint32_t)total;
    total += nc;
    max_cells = std::max(max_cells, nc);
  }
  if (total >= (1ull << 32)) return fail(c, TVK_ERR_UNSUPPORTED, "cell grid exceeds 2^32 cells");
  GatherBrick* d_gb = nullptr;
  uint32_t* d_base = nullptr;
  uint32_t* d_cells = nullptr;
  hipError_t e = hipMalloc(&d_cells, std::max<uint64_t>(total, 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&d_gb, gb.size() * sizeof(GatherBrick));
  if (e == hipSuccess) e = hipMalloc(&d_base, G.base.size() * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_gb, gb.data(), gb.size() * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_base, G.base.data(), G.base.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = launch_cells(c->d_pool, L.bpv, d_gb, d_base, (uint32_t)gb.size(), max_cells, d_cells, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_gb) (void)hipFree(d_gb);
  if (d_base) (void)hipFree(d_base);
  if (e != hipSuccess) {
    if (d_cells) (void)hipFree(d_cells);
    return hipfail(c, "cell grid", e);
  }
  G.d = d_cells;
  return TVK_OK;
}

void brick_tf_stats(tvk_ctx* c, FrameBrick& fb) {
  const tvk_ctx::TfStats& T = c->tfst;
  fb.cells = c->cells && !c->cells->base.empty() ? c->cells->base[fb.toc] : 0u;
  for (float& v : fb.tfs) v = 0.0f;
  if (T.n <= 0 || c->P.mode == TVK_RM_ISOSURFACE) return;
  // TF coordinate and entry of a stored value, as the march computes them
  const float inv = c->L.bpv == 1 ? 1.0f / 255.0f : 1.0f / 65535.0f;
  const float sc = tf_scale_of(c);
  auto coord = [&](double v) { return std::min(std::max((float)v * inv * sc, 0.0f), 1.0f); };
  auto entry = [&](float cc) { return std::min((int)std::floor(cc * (float)T.n), T.n - 1); };
  const float c0 = coord(c->minmax[2 * (size_t)fb.toc]), c1 = coord(c->minmax[2 * (size_t)fb.toc + 1]);
  const int lo = entry(c0), hi = entry(c1);
  const double cnt = (double)(hi - lo + 1);
  fb.tfs[0] = c0;
  fb.tfs[1] = c1;
  fb.tfs[2] = (float)((T.pa[(size_t)hi + 1] - T.pa[(size_t)lo]) / cnt);
  fb.tfs[3] = range_max(T.amax, lo, hi);
  fb.tfs[4] = (float)(T.ptv[(size_t)hi] - T.ptv[(size_t)lo]);
  fb.tfs[5] = (float)(T.prgb[(size_t)hi] - T.prgb[(size_t)lo]);
  fb.tfs[6] = hi > lo ? range_max(T.smax, lo, hi - 1) : 0.0f;
  fb.tfs[7] = (float)((T.pnz[(size_t)hi + 1] - T.pnz[(size_t)lo]) / cnt);
}

// Shading constants of the context's dataset, TF and parameters
void fill_shade(const tvk_ctx* c, ShadeConst& S) {
  S = ShadeConst{};
  S.pool = (const uint8_t*)(compat_active(c) ? c->cp.d : c->d_pool);
  S.tf1 = c->d_tf1f;
  S.tf2f = c->d_tf2f;
  S.tf2_cols = c->d_tf2cols;
  S.tf_n = c->tf1d_n;
  S.tf_w = c->tf2d_w;
  S.tf_h = c->tf2d_h;
  S.tf_scale = tf_scale_of(c);
  S.grad_scale = c->P.grad_scale;
  S.iso_n = normalized_isovalue(c);
  S.ert_eps = c->P.ert_eps;
  for (int i = 0; i < 3; ++i) {
    S.dscale[i] = c->P.domain_scale[i];
    S.amb[i] = c->P.light_ambient[i];
    S.dif[i] = c->P.light_diffuse[i];
    S.spc[i] = c->P.light_specular[i];
    S.ldir[i] = c->P.light_dir[i];
    S.iso_col[i] = c->P.iso_color[i];
  }
  S.refine_steps = c->P.refine_steps;
  S.clamp_edge = c->P.clamp_to_edge;
  S.tfst = c->d_tfst;
  S.tfst_n = c->tfst.n;
  S.tfst_levels = c->tfst.levels;
  S.cells = c->cells ? c->cells->d : nullptr;
  S.overlap = (int)c->L.overlap;
}

// The launch of an allocation-free render: the planned lists go to the
// caller's buffers (fixed), checked against their capacities.
static int launch_fixed(tvk_ctx* c, int nviews, uint64_t pixels, uint32_t groups, uint32_t in_list,
                        const FrameBuffers& fx, tvk_stats* st) {
  if (c->h_list.size() > fx.list_cap)
    return fail(c, TVK_ERR_ARG, "visible-brick list (" + std::to_string(c->h_list.size()) +
                                    " entries) exceeds the allocation-free render's limit of " +
                                    std::to_string(fx.list_cap));
  if ((size_t)nviews > fx.views_cap)
    return fail(c, TVK_ERR_ARG, std::to_string(nviews) + " views exceed the allocation-free render's limit of " +
                                    std::to_string(fx.views_cap));
  if ((size_t)pixels * 4 > fx.out_floats) return fail(c, TVK_ERR_ARG, "output buffer too small");
  if (!c->h_list.empty())
    HIPCHK(c, hipMemcpyAsync(fx.list, c->h_list.data(), c->h_list.size() * sizeof(FrameBrick), hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(fx.views, c->h_views.data(), (size_t)nviews * sizeof(ViewConst), hipMemcpyHostToDevice,
                           c->stream));
  MarchArgs a{};
  fill_shade(c, a.S);
  a.list = fx.list;
  a.views = fx.views;
  a.mode = c->P.mode;
  a.out = fx.out;
  a.counter = c->count ? fx.counter : nullptr;
  a.ntoc = (uint32_t)c->L.toc.size();
  a.touched = nullptr;
  a.nviews = nviews;
  a.lighting = c->P.lighting;
  a.groups_total = groups;
  a.groups_padded = (groups + 7u) & ~7u;
  // a kernel, not hipMemsetAsync, clears the counter (capturable into a graph)
  if (c->count) HIPCHK(c, launch_fill_u32((unsigned int*)fx.counter, 2u, 0u, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[6], c->stream));
  if (a.groups_total > 0) HIPCHK(c, launch_march(a, render_bpv(c), c->stream, 0));
  HIPCHK(c, hipEventRecord(c->ev[7], c->stream));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    unsigned long long ns = 0;
    if (c->count) HIPCHK(c, hipMemcpyAsync(&ns, fx.counter, sizeof(ns), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    st->samples = ns;
    st->bricks_in_list = in_list;
    st->bricks_nonempty = (uint32_t)c->h_list.size();
    st->ms = ms;
  }
  return TVK_OK;
}

// learn: 0 fixed step, 1 step proposal of c->net, 2 fixed step recording segment samples into *seg.
// fixed: device buffers the caller provides (allocation-free render): nothing
// is allocated, the uploads come from the context's host vectors, the sample
// counter is cleared by a kernel and the frames go to fixed->out.
int render_impl(tvk_ctx* c, int nviews, const float* rots, float* host_out, tvk_stats* st, int learn,
                const SegOut* seg, const FrameBuffers* fixed) {
  if (learn < 0) learn = (c->learn_on && c->P.mode != TVK_RM_ISOSURFACE) ? 1 : 0;
  if (learn > 0 && c->P.mode == TVK_RM_ISOSURFACE)
    return fail(c, TVK_ERR_UNSUPPORTED, "step proposal is not available in isosurface mode");
  if (fixed && learn > 0)
    return fail(c, TVK_ERR_UNSUPPORTED, "the allocation-free render marches at the fixed step (detach the proposal)");
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (!c->has_view) return fail(c, TVK_ERR_STATE, "no view");
  if (nviews < 1) return fail(c, TVK_ERR_ARG, "nviews must be >= 1");
  if (c->ingest_pending)
    return fail(c, TVK_ERR_STATE, std::to_string(c->ingest_pending) + " bricks were declared but not uploaded");
  HIPCHK(c, hipSetDevice(c->device));
  if (budgeted(c)) {
    if (fixed) return fail(c, TVK_ERR_UNSUPPORTED, "the allocation-free render needs every brick resident");
    if (learn > 0) return fail(c, TVK_ERR_UNSUPPORTED, "the learned step needs every brick resident (pool budget 0)");
    return render_budget_impl(c, nviews, rots, host_out, st);
  }
  if (compat_active(c)) {
    if (learn > 0)
      return fail(c, TVK_ERR_UNSUPPORTED, "the learned step reads the dataset's own bricks (compatibility textures set)");
    if (fixed && !compat_ready(c))
      return fail(c, TVK_ERR_STATE, "the allocation-free render needs the compatibility textures: render once first");
    int rc = ensure_compat(c);
    if (rc) return rc;
  }
  // the host vectors are the upload source of an allocation-free render: the
  // previous call's copies must have consumed them
  if (fixed) HIPCHK(c, hipStreamSynchronize(c->stream));
  if (learn > 0) {
    int rc = learn_prepare(c);
    if (rc) return rc;
  }
  c->h_list.clear();
  c->h_views.assign((size_t)nviews, ViewConst{});
  std::vector<PlannedBrick> planned;
  uint32_t groups = 0, in_list = 0;
  uint64_t pixels = 0;
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
    ViewConst& vc = c->h_views[(size_t)v];
    make_view_const(V, vc);
    vc.list_offset = (uint32_t)c->h_list.size();
    make_frame_bricks(c->L, V, pp, planned, c->h_list, compat_dims(c), compat_offsets(c));
    vc.nlist = (uint32_t)c->h_list.size() - vc.list_offset;
    if (learn > 0)
      for (size_t q = vc.list_offset; q < c->h_list.size(); ++q) brick_tf_stats(c, c->h_list[q]);
    vc.out_offset = (uint32_t)pixels;
    vc.group_offset = groups;
    groups += vc.groups;
    pixels += (uint64_t)V.W * V.H;
    c->last_lod = lod;
  }
  if (pixels >= (1ull << 32)) return fail(c, TVK_ERR_ARG, "too many pixels in one call");
  if (fixed) return launch_fixed(c, nviews, pixels, groups, in_list, *fixed, st);
  // the dispatch order (raster: none)
  const int dorder = c->dispatch & 3;
  const bool ordered = dorder != TVK_DISPATCH_RASTER && groups > 0;
  uint32_t xcd_start[8] = {}, xcd_count[8] = {};
  if (ordered) {
    c->h_cost.assign((size_t)groups, 0.0f);
    for (int v = 0; v < nviews; ++v) {
      const ViewConst& vc = c->h_views[(size_t)v];
      group_costs(vc, c->h_list.data() + vc.list_offset, c->h_cost.data() + vc.group_offset);
    }
    dispatch_order(c->h_views.data(), nviews, c->h_cost, dorder, c->h_order, xcd_start, xcd_count);
  }
  int rc = ensure(c, &c->d_list, &c->list_cap, std::max<size_t>(c->h_list.size(), 1));
  if (!rc) rc = ensure(c, &c->d_views, &c->views_cap, (size_t)nviews);
  if (!rc && ordered) rc = ensure(c, &c->d_order, &c->order_cap, (size_t)groups);
  if (!rc) rc = ensure(c, &c->d_out, &c->out_cap, (size_t)pixels * 4);
  if (!rc) rc = ensure_pinned(c, &c->h_list_pinned, &c->h_list_pinned_cap, std::max<size_t>(c->h_list.size(), 1));
  if (!rc) rc = ensure_pinned(c, &c->h_views_pinned, &c->h_views_pinned_cap, (size_t)nviews);
  if (!rc && ordered) rc = ensure_pinned(c, &c->h_order_pinned, &c->h_order_pinned_cap, (size_t)groups);
  if (rc) return rc;
  // staging through pinned memory keeps the uploads asynchronous; wait for the
  // previous frame's uploads before overwriting the staging buffers
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::copy(c->h_list.begin(), c->h_list.end(), c->h_list_pinned);
  std::copy(c->h_views.begin(), c->h_views.end(), c->h_views_pinned);
  if (ordered) std::copy(c->h_order.begin(), c->h_order.end(), c->h_order_pinned);
  if (!c->h_list.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_list, c->h_list_pinned, c->h_list.size() * sizeof(FrameBrick), hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_views, c->h_views_pinned, (size_t)nviews * sizeof(ViewConst), hipMemcpyHostToDevice,
                           c->stream));
  if (ordered)
    HIPCHK(c, hipMemcpyAsync(c->d_order, c->h_order_pinned, (size_t)groups * sizeof(uint32_t), hipMemcpyHostToDevice,
                             c->stream));
  MarchArgs a{};
  fill_shade(c, a.S);
  a.list = c->d_list;
  a.views = c->d_views;
  a.mode = c->P.mode;
  a.out = c->d_out;
  a.counter = c->count ? c->d_counter : nullptr;
  a.ntoc = (uint32_t)c->L.toc.size();
  a.touched = nullptr;
  if (c->count) {
    const size_t nt = (size_t)nviews * c->L.toc.size();
    rc = ensure(c, &c->d_touched, &c->touched_cap, nt);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_touched, 0, nt * sizeof(uint32_t), c->stream));
    c->touched_n = nt;
    a.touched = c->d_touched;
  }
  a.nviews = nviews;
  if (learn == 1) {
    a.learn = c->learn;
    if (a.learn.fixed_k < 0 && !stepnet_device(c->net, &a.learn.net))
      return fail(c, TVK_ERR_STATE, "no step network");
  }
  if (learn == 2 && seg) a.seg = *seg;
  c->hitpos_valid = false;
  if (c->P.mode == TVK_RM_ISOSURFACE) {
    rc = ensure(c, &c->d_hitpos, &c->hitpos_cap, (size_t)pixels);
    if (rc) return rc;
    a.hitpos = c->d_hitpos;
    c->hitpos_valid = true;
  }
  a.lighting = c->P.lighting;
  a.groups_total = groups;
  a.groups_padded = (groups + 7u) & ~7u;
  if (ordered) {
    a.order = c->d_order;
    uint32_t most = 0;
    for (int x = 0; x < 8; ++x) {
      a.xcd_start[x] = xcd_start[x];
      a.xcd_count[x] = xcd_count[x];
      most = std::max(most, xcd_count[x]);
    }
    a.groups_padded = 8u * most;
  }
  if ((c->dispatch & TVK_DISPATCH_QUEUES) && groups > 0) {
    // persistent launch: waves pull 8x8 tiles from the 8 XCD ranges
    if (!ordered) {
      const uint32_t per = (groups + 7u) / 8u;
      for (int x = 0; x < 8; ++x) {
        a.xcd_start[x] = std::min(groups, (uint32_t)x * per);
        a.xcd_count[x] = std::min(per, groups - a.xcd_start[x]);
      }
    }
    rc = ensure(c, &c->d_queues, &c->queues_cap, 8);
    if (rc) return rc;
    HIPCHK(c, launch_fill_u32(c->d_queues, 8u, 0u, c->stream));
    a.queues = c->d_queues;
  }
  if (c->count) HIPCHK(c, hipMemsetAsync(c->d_counter, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipEventRecord(c->ev[6], c->stream));
  if (a.groups_total > 0) HIPCHK(c, launch_march(a, render_bpv(c), c->stream, learn));
  HIPCHK(c, hipEventRecord(c->ev[7], c->stream));
  c->last_views = nviews;
  c->last_w = c->V.W;
  c->last_h = c->V.H;
  if (host_out)
    HIPCHK(c, hipMemcpyAsync(host_out, c->d_out, (size_t)pixels * 4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    unsigned long long ns = 0;
    if (c->count)
      HIPCHK(c, hipMemcpyAsync(&ns, c->d_counter, sizeof(ns), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    st->samples = ns;
    st->bricks_in_list = in_list;
    st->bricks_nonempty = (uint32_t)c->h_list.size();
    st->ms = ms;
  } else if (host_out) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return TVK_OK;
}

}  // namespace

extern "C" {

int tvk_current_lod(tvk_ctx* c, int* lod) {
  if (!c || !lod) return TVK_ERR_ARG;
  if (!c->has_ds || !c->has_view) return fail(c, TVK_ERR_STATE, "no dataset or view");
  return frame_lod(c, c->V, lod);
}

int tvk_plan(tvk_ctx* c, tvk_brick_rec* out, int cap, int* n) {
  if (!c || !n) return TVK_ERR_ARG;
  if (!c->has_ds || !c->has_view) return fail(c, TVK_ERR_STATE, "no dataset or view");
  int lod = 0;
  int rc = frame_lod(c, c->V, &lod);
  if (rc) return rc;
  PlanParams pp;
  rc = plan_params(c, lod, pp);
  if (rc) return rc;
  std::vector<PlannedBrick> list;
  plan_frame(c->L, c->minmax, c->range_max, c->V, pp, list);
  *n = (int)list.size();
  if (!out) return TVK_OK;
  if ((int)list.size() > cap) return fail(c, TVK_ERR_ARG, "output too small");
  for (size_t i = 0; i < list.size(); ++i) {
    const PlannedBrick& b = list[i];
    tvk_brick_rec& r = out[i];
    r.key_index = b.key_index;
    r.toc_index = b.toc_index;
    for (int k = 0; k < 3; ++k) {
      r.center[k] = b.center[k];
      r.extension[k] = b.extension[k];
      r.voxel_count[k] = b.voxel_count[k];
      r.tex_min[k] = b.tex_min[k];
      r.tex_max[k] = b.tex_max[k];
    }
    r.distance = b.distance;
    r.is_empty = b.is_empty;
  }
  return TVK_OK;
}

int tvk_render(tvk_ctx* c, float* host_out, tvk_stats* st) {
  if (!c) return TVK_ERR_ARG;
  return render_impl(c, 1, nullptr, host_out, st, -1, nullptr, nullptr);
}

int tvk_render_views(tvk_ctx* c, int n, const float* rots, float* host_out, tvk_stats* st) {
  if (!c || !rots) return TVK_ERR_ARG;
  return render_impl(c, n, rots, host_out, st, -1, nullptr, nullptr);
}

int tvk_set_step_proposal(tvk_ctx* c, tvk_stepnet* net, const tvk_proposal_params* p) {
  if (!c) return TVK_ERR_ARG;
  if (net && !p) return fail(c, TVK_ERR_ARG, "null proposal parameters");
  c->net = nullptr;
  c->learn = LearnConst{};
  c->learn_on = false;
  if (net || (p && p->fixed_log2 >= 0)) {
    if (p->max_log2 < 0 || p->max_log2 > 3) return fail(c, TVK_ERR_ARG, "max_log2 must be 0..3");
    if (p->fixed_log2 > p->max_log2) return fail(c, TVK_ERR_ARG, "fixed_log2 exceeds max_log2");
    c->net = net;
    c->learn_on = true;
    c->learn.fixed_k = p->fixed_log2 >= 0 ? p->fixed_log2 : -1;
    c->learn.margin = p->margin;
    const double q = p->skip_threshold;
    c->learn.skip_logit = q >= 1.0 ? INFINITY : (q <= 0.0 ? -INFINITY : (float)std::log(q / (1.0 - q)));
    c->learn.max_log2 = p->max_log2;
    c->learn.fallback = p->fallback_bound > 0.0f ? p->fallback_bound : INFINITY;
    c->learn.fallback_tv = p->fallback_tv > 0.0f ? p->fallback_tv : INFINITY;
    c->learn.skip_bound = p->skip_bound;
  }
  return TVK_OK;
}

int tvk_generate_segments(tvk_ctx* c, int nviews, const float* rots, uint32_t cap, float keep, uint32_t seed,
                          float* host_out, uint32_t* n, uint32_t* total) {
  if (!c || !rots || !host_out || !n || nviews < 1 || cap == 0) return TVK_ERR_ARG;
  if (c->P.mode == TVK_RM_ISOSURFACE) return fail(c, TVK_ERR_UNSUPPORTED, "segments need a TF render mode");
  if (!(keep > 0.0f) || keep > 1.0f) return fail(c, TVK_ERR_ARG, "keep fraction must be in (0,1]");
  HIPCHK(c, hipSetDevice(c->device));
  SegSample* d_out = nullptr;
  unsigned int* d_cnt = nullptr;
  HIPCHK(c, hipMalloc((void**)&d_out, (size_t)cap * sizeof(SegSample)));
  hipError_t e = hipMalloc((void**)&d_cnt, sizeof(unsigned int));
  if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, sizeof(unsigned int), c->stream);
  int rc = e == hipSuccess ? TVK_OK : hipfail(c, "segment buffers", e);
  if (!rc) {
    SegOut so;
    so.out = d_out;
    so.count = d_cnt;
    so.cap = cap;
    so.keep = keep >= 1.0f ? 0xffffffffu : (uint32_t)((double)keep * 4294967296.0);
    so.seed = seed;
    const float eps = c->P.ert_eps;
    c->P.ert_eps = 0.0f;   // the recorded marches run without early termination
    rc = render_impl(c, nviews, rots, nullptr, nullptr, 2, &so, nullptr);
    c->P.ert_eps = eps;
  }
  unsigned int cnt = 0;
  if (!rc) {
    e = hipMemcpyAsync(&cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (total) *total = cnt;
    cnt = std::min(cnt, cap);
    if (e == hipSuccess && cnt)
      e = hipMemcpy(host_out, d_out, (size_t)cnt * sizeof(SegSample), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hipfail(c, "segment download", e);
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d_out);
  if (d_cnt) (void)hipFree(d_cnt);
  *n = rc ? 0 : cnt;
  return rc;
}

int tvk_cell_grid(tvk_ctx* c, uint32_t* out, uint64_t cap, uint64_t* n) {
  if (!c || !n) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  HIPCHK(c, hipSetDevice(c->device));
  const int rc = ensure_cells(c);
  if (rc) return rc;
  const CellGrid& G = *c->cells;
  uint64_t total = 0;
  for (size_t i = 0; i < c->L.toc.size(); ++i) {
    const uint32_t* sz = c->L.toc[i].size;
    total += (uint64_t)cell_count(sz[0]) * cell_count(sz[1]) * cell_count(sz[2]);
  }
  *n = total;
  if (out && cap) {
    HIPCHK(c, hipMemcpy(out, G.d, std::min(cap, total) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }
  return TVK_OK;
}

int tvk_touched_bytes(tvk_ctx* c, uint64_t* bytes, uint64_t* bricks, uint64_t* unique_bytes,
                      uint64_t* unique_bricks) {
  if (!c || !bytes || !bricks || !unique_bytes || !unique_bricks) return TVK_ERR_ARG;
  if (!c->touched_n) return fail(c, TVK_ERR_STATE, "no counted render yet");
  std::vector<uint32_t> t(c->touched_n);
  HIPCHK(c, hipMemcpyAsync(t.data(), c->d_touched, t.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t nt = c->L.toc.size();
  std::vector<uint8_t> any(nt, 0);
  uint64_t b = 0, n = 0, ub = 0, un = 0;
  for (size_t i = 0; i < t.size(); ++i)
    if (t[i]) {
      const size_t k = i % nt;
      const uint64_t bb = c->L.toc[k].real * (uint64_t)c->L.bpv;
      ++n; b += bb;
      if (!any[k]) { any[k] = 1; ++un; ub += bb; }
    }
  *bytes = b;
  *bricks = n;
  *unique_bytes = ub;
  *unique_bricks = un;
  return TVK_OK;
}

int tvk_touched_mask(tvk_ctx* c, uint8_t* mask, size_t n) {
  if (!c || !mask) return TVK_ERR_ARG;
  if (!c->touched_n) return fail(c, TVK_ERR_STATE, "no counted render yet");
  const size_t nt = c->L.toc.size();
  if (n != nt) return fail(c, TVK_ERR_ARG, "mask size must equal the brick count");
  std::vector<uint32_t> t(c->touched_n);
  HIPCHK(c, hipMemcpyAsync(t.data(), c->d_touched, t.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::memset(mask, 0, n);
  for (size_t i = 0; i < t.size(); ++i)
    if (t[i]) mask[i % nt] = 1;
  return TVK_OK;
}

int tvk_display(tvk_ctx* c, const float bg_top[3], const float bg_bottom[3], uint8_t* out, size_t bytes) {
  if (!c || !bg_top || !bg_bottom || !out) return TVK_ERR_ARG;
  if (c->last_views < 1 || !c->d_out) return fail(c, TVK_ERR_STATE, "no rendered frame");
  const size_t px = (size_t)c->last_w * c->last_h * (size_t)c->last_views;
  if (bytes != px * 4) return fail(c, TVK_ERR_ARG, "output size mismatch");
  HIPCHK(c, hipSetDevice(c->device));
  uint8_t* d = nullptr;
  HIPCHK(c, hipMalloc((void**)&d, bytes));
  hipError_t e = launch_display(c->d_out, c->last_w, c->last_h, (uint32_t)c->last_views, bg_top, bg_bottom, d,
                                c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return hipfail(c, "display", e);
  return TVK_OK;
}

int tvk_pick(tvk_ctx* c, uint32_t x, uint32_t y, float out[3]) {
  if (!c || !out) return TVK_ERR_ARG;
  if (c->P.mode != TVK_RM_ISOSURFACE || !c->hitpos_valid)
    return fail(c, TVK_ERR_STATE, "Can only determine pick locations in isosurface rendering mode.");
  // ReadBackPixels(mousePos.x, m_vWinSize.y - mousePos.y) of the first view
  const uint32_t W = c->last_w, H = c->last_h;
  if (x >= W || y > H || y == 0) return fail(c, TVK_ERR_RANGE, "pick position outside the frame");
  const size_t pix = (size_t)(H - y) * W + x;
  float4 v;
  HIPCHK(c, hipMemcpyAsync(&v, c->d_hitpos + pix, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (v.w == 0.0f) return fail(c, TVK_ERR_RANGE, "No intersection.");
  out[0] = v.x; out[1] = v.y; out[2] = v.z;
  return TVK_OK;
}

int tvk_frame_device_ptr(tvk_ctx* c, void** p) {
  if (!c || !p) return TVK_ERR_ARG;
  *p = c->d_out;
  return TVK_OK;
}

int tvk_event_record(tvk_ctx* c, int slot) {
  if (!c || slot < 0 || slot >= 6) return TVK_ERR_ARG;
  HIPCHK(c, hipEventRecord(c->ev[slot], c->stream));
  return TVK_OK;
}

int tvk_event_elapsed(tvk_ctx* c, int a, int b, float* ms) {
  if (!c || !ms || a < 0 || b < 0 || a >= 8 || b >= 8) return TVK_ERR_ARG;
  HIPCHK(c, hipEventSynchronize(c->ev[b]));
  HIPCHK(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
  return TVK_OK;
}

int tvk_synchronize(tvk_ctx* c) {
  if (!c) return TVK_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TVK_OK;
}

}  // extern "C"
