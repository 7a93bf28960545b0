// C ABI of the GridLeaper-style traversal (include/tvk_hip.h, "ray-guided traversal").
//
// Host half of GLGridLeaper + GLVolumePool: brick IDs and layouts
// (GLVolumePool.cpp:89-117, :205-223, :288-291), the brick metadata with the
// visibility of the current TF (:961-1100), the table of requested bricks
// (GLHashTable.cpp:90-170), residency and paging of requested bricks (:673-717,
// :955-959, :1720-1789), and the shader uniforms (GLGridLeaper.cpp:560-574,
// :700-745; GLVolumePool.cpp:790-803).  With every brick in HBM paging marks
// bricks resident; with a budgeted pool it copies them into slots, the oldest
// upload replaced first, the coarsest brick never.
#include "ctx.hpp"

using namespace tvk;

namespace tvk {
hipError_t launch_leap(const LeapArgs& a, int bpv, hipStream_t st);
}

namespace {

enum : uint32_t { BI_MISSING = 0, BI_CHILD_EMPTY = 1, BI_EMPTY = 2, BI_FLAG_COUNT = 3 };

struct Grid {
  int nlod = 0;
  float layout[LEAP_MAX_LOD][3];
  uint32_t ilayout[LEAP_MAX_LOD][3];
  uint32_t lodoff[LEAP_MAX_LOD];
  uint64_t id(uint64_t x, uint64_t y, uint64_t z, int l) const {
    return lodoff[l] + x + y * ilayout[l][0] + z * (uint64_t)ilayout[l][0] * ilayout[l][1];
  }
};

int make_grid(tvk_ctx* c, Grid& G) {
  const Layout& L = c->L;
  G.nlod = (int)L.lods.size();
  if (G.nlod > LEAP_MAX_LOD) return fail(c, TVK_ERR_UNSUPPORTED, "too many levels for the traversal");
  if (L.toc.size() != [&] {
        uint64_t n = 0;
        for (const LodInfo& li : L.lods) n += li.bricks[0] * li.bricks[1] * li.bricks[2];
        return n;
      }())
    return fail(c, TVK_ERR_UNSUPPORTED, "the traversal needs every level of detail in the pool (max_lods = 0)");
  float inner[3];
  for (int a = 0; a < 3; ++a) inner[a] = (float)(L.brick[a] - 2 * L.overlap);
  uint64_t off = 0;
  for (int l = 0; l < G.nlod; ++l) {
    for (int a = 0; a < 3; ++a) {
      // GetFloatBrickLayout (GLVolumePool.cpp:89-108)
      float b = float(L.dims[a]) / inner[a];
      b /= float(1u << l);
      if (float(uint32_t(b)) == b) b -= b * std::numeric_limits<float>::epsilon();
      G.layout[l][a] = b;
      // GetBrickLayout (:110-117)
      const uint64_t base = (uint64_t)std::ceil(double(L.dims[a]) / inner[a]);
      G.ilayout[l][a] = (uint32_t)std::ceil(double(base) / double(1ull << l));
    }
    G.lodoff[l] = (uint32_t)off;
    const LodInfo& li = L.lods[(size_t)l];
    if (li.bricks[0] != G.ilayout[l][0] || li.bricks[1] != G.ilayout[l][1] || li.bricks[2] != G.ilayout[l][2])
      return fail(c, TVK_ERR_UNSUPPORTED, "the pool's brick grid differs from the octree's at level " +
                                              std::to_string(l) + " (sizes must halve evenly)");
    off += (uint64_t)G.ilayout[l][0] * G.ilayout[l][1] * G.ilayout[l][2];
  }
  return TVK_OK;
}

// brick metadata: GLGridLeaper::RecomputeBrickVisibility + GLVolumePool visibility
int make_meta(tvk_ctx* c, const Grid& G, std::vector<uint32_t>& meta) {
  const Layout& L = c->L;
  const bool two = c->P.mode == TVK_RM_2DTRANS, iso = c->P.mode == TVK_RM_ISOSURFACE;
  double fmin = 0.0, fmax = 0.0;
  if (iso) {
    // ContainsData for RM_ISOSURFACE (GLVolumePool.cpp:983-985): min <= isovalue <= max
    fmin = fmax = (double)c->P.isovalue;
  } else {
    const int n = two ? c->tf2d_w : c->tf1d_n;
    if (n < 2) return fail(c, TVK_ERR_STATE, "no transfer function");
    const double fMaxValue = c->range_min > c->range_max ? (double)n : c->range_max;
    const double resc = fMaxValue / double(n - 1);
    uint64_t lim[4] = {0, 0, 0, 0};
    if (two) tf2d_limits(c->tf2d.data(), c->tf2d_w, c->tf2d_h, lim);
    else tf1d_limits(c->tf1d.data(), c->tf1d_n, lim);
    fmin = double(lim[0]) * resc;
    fmax = double(lim[1]) * resc;
  }
  meta.assign(L.toc.size(), BI_MISSING);
  for (int l = 0; l < G.nlod; ++l)
    for (uint64_t z = 0; z < G.ilayout[l][2]; ++z)
      for (uint64_t y = 0; y < G.ilayout[l][1]; ++y)
        for (uint64_t x = 0; x < G.ilayout[l][0]; ++x) {
          const uint64_t i = G.id(x, y, z, l);
          const bool empty = !(fmax >= c->minmax[2 * i] && fmin <= c->minmax[2 * i + 1]);
          if (!empty) {
            const bool res = budgeted(c) ? c->pst->brick_slot[i] >= 0
                                         : (c->leap_resident.empty() || c->leap_resident[i] || l == G.nlod - 1);
            meta[i] = res ? (uint32_t)(BI_FLAG_COUNT + i) : (uint32_t)BI_MISSING;
            continue;
          }
          uint32_t m = BI_CHILD_EMPTY;   // finest-level empty bricks are child empty by definition
          if (l > 0)
            for (int k = 0; k < 8 && m == BI_CHILD_EMPTY; ++k) {
              const uint64_t cx = 2 * x + (k & 1), cy = 2 * y + ((k >> 1) & 1), cz = 2 * z + (k >> 2);
              if (cx >= G.ilayout[l - 1][0] || cy >= G.ilayout[l - 1][1] || cz >= G.ilayout[l - 1][2]) continue;
              if (meta[G.id(cx, cy, cz, l - 1)] != BI_CHILD_EMPTY) m = BI_EMPTY;
            }
          meta[i] = m;
        }
  return TVK_OK;
}

// TOC bricks of the request table's entries, in table order (GLHashTable::GetData)
std::vector<uint32_t> decode_requests(const tvk_ctx* c, const Grid& G, const std::vector<uint32_t>& table) {
  std::vector<uint32_t> out;
  const uint64_t lx = G.ilayout[0][0], ly = G.ilayout[0][1], lz = G.ilayout[0][2];
  const uint64_t vol = lx * ly * lz;
  for (uint32_t e : table) {
    if (!e) continue;
    uint64_t i = (uint64_t)e - 1;   // GLHashTable::Int2Vector
    const uint64_t w = i / vol;
    i -= w * vol;
    const uint64_t z = i / (lx * ly);
    i -= z * lx * ly;
    const uint64_t y = i / lx, x = i - y * lx;
    if ((int)w < G.nlod && x < G.ilayout[w][0] && y < G.ilayout[w][1] && z < G.ilayout[w][2])
      out.push_back((uint32_t)G.id(x, y, z, (int)w));
  }
  (void)c;
  return out;
}

int fetch_table(tvk_ctx* c, std::vector<uint32_t>& table) {
  if (!c->d_leap_table || !c->leap_table_used) return fail(c, TVK_ERR_STATE, "no pass yet");
  table.assign(c->leap_table_used, 0u);
  HIPCHK(c, hipMemcpyAsync(table.data(), c->d_leap_table, table.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TVK_OK;
}

// The coarsest brick of a budgeted pool sits in a slot that is never replaced
// (GLVolumePool::UploadFirstBrick, GLVolumePool.cpp:763-766).
int leap_pin_coarsest(tvk_ctx* c) {
  PoolState& P = *c->pst;
  const uint32_t last = (uint32_t)c->L.toc.size() - 1;
  if (P.brick_slot[last] >= 0) {
    P.slot_time[(size_t)P.brick_slot[last]] = std::numeric_limits<uint64_t>::max();
    return TVK_OK;
  }
  // the oldest slot (or a free one)
  int32_t best = 0;
  for (uint32_t s = 1; s < P.nslots; ++s)
    if (P.slot_time[s] < P.slot_time[(size_t)best]) best = (int32_t)s;
  const int rc = pool_upload(c, std::vector<uint32_t>{last}, std::vector<int32_t>{best});
  if (rc) return rc;
  P.slot_time[(size_t)best] = std::numeric_limits<uint64_t>::max();
  return TVK_OK;
}

// GLVolumePool::UploadBricks with a budgeted pool: slots sorted by upload time
// (PrepareForPaging), requested bricks written to the oldest slots in request
// order, at most every slot but the coarsest brick's, each stamped with the
// next upload time; a replaced brick becomes missing.
int leap_page_budgeted(tvk_ctx* c, const std::vector<uint32_t>& req, uint64_t max_bricks, uint64_t* paged) {
  PoolState& P = *c->pst;
  int rc = leap_pin_coarsest(c);
  if (rc) return rc;
  std::vector<uint32_t> want;
  for (uint32_t t : req)
    if (P.brick_slot[t] < 0 && std::find(want.begin(), want.end(), t) == want.end()) want.push_back(t);
  std::vector<int32_t> order(P.nslots);
  for (uint32_t s = 0; s < P.nslots; ++s) order[s] = (int32_t)s;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return P.slot_time[(size_t)a] < P.slot_time[(size_t)b]; });
  const size_t room = P.nslots - 1;
  const size_t n = (size_t)std::min<uint64_t>(std::min<uint64_t>(want.size(), room), max_bricks);
  want.resize(n);
  order.resize(n);
  if (n) {
    rc = pool_upload(c, want, order);
    if (rc) return rc;
    for (int32_t s : order) P.slot_time[(size_t)s] = P.clock++;
  }
  if (paged) *paged = n;
  return TVK_OK;
}

}  // namespace

extern "C" {

int tvk_leap_set_request_table(tvk_ctx* c, uint32_t size, uint32_t rehash) {
  if (!c) return TVK_ERR_ARG;
  if (size > (1u << 26)) return fail(c, TVK_ERR_ARG, "request table too large");
  c->leap_table_size = size;
  c->leap_rehash = rehash;
  return TVK_OK;
}

int tvk_leap_requests(tvk_ctx* c, uint32_t* entries, size_t n, size_t* size) {
  if (!c) return TVK_ERR_ARG;
  std::vector<uint32_t> t;
  const int rc = fetch_table(c, t);
  if (rc) return rc;
  if (size) *size = t.size();
  if (entries) {
    if (n < t.size()) return fail(c, TVK_ERR_ARG, "output too small");
    std::copy(t.begin(), t.end(), entries);
  }
  return TVK_OK;
}

int tvk_pool_resident(tvk_ctx* c, uint8_t* flags, size_t n) {
  if (!c || !flags) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  const size_t nb = c->L.toc.size();
  if (n != nb) return fail(c, TVK_ERR_ARG, "one flag per brick expected");
  for (size_t i = 0; i < nb; ++i) {
    if (budgeted(c)) flags[i] = c->pst->brick_slot[i] >= 0 ? 1 : 0;
    else if (!c->ingest_missing.empty() && c->ingest_missing[i]) flags[i] = 0;
    else flags[i] = (c->leap_resident.empty() || c->leap_resident[i] || i + 1 == nb) ? 1 : 0;
  }
  return TVK_OK;
}

int tvk_leap_set_resident(tvk_ctx* c, const uint8_t* flags, size_t n) {
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (flags && n != c->L.toc.size()) return fail(c, TVK_ERR_ARG, "one flag per brick expected");
  if (budgeted(c)) {
    // evict everything but the coarsest brick, then page the flagged bricks in
    PoolState& P = *c->pst;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint32_t s = 0; s < P.nslots; ++s) {
      if (P.slot_brick[s] >= 0 && (size_t)P.slot_brick[s] + 1 != c->L.toc.size()) {
        P.brick_slot[(size_t)P.slot_brick[s]] = -1;
        P.slot_brick[s] = -1;
        P.slot_time[s] = 0;
      }
    }
    // NULL: only the coarsest brick stays (a budgeted pool cannot hold every brick)
    std::vector<uint32_t> req;
    if (flags) {
      for (size_t i = 0; i + 1 < n; ++i)
        if (flags[i]) req.push_back((uint32_t)i);
    }
    if (req.size() + 1 > P.nslots) return fail(c, TVK_ERR_NOMEM, "more bricks flagged than the pool's slots hold");
    uint64_t paged = 0;
    return leap_page_budgeted(c, req, req.size(), &paged);
  }
  if (!flags) { c->leap_resident.clear(); return TVK_OK; }
  c->leap_resident.assign(flags, flags + n);
  return TVK_OK;
}

int tvk_leap_render(tvk_ctx* c, int first, float* host_out, tvk_leap_stats* st) {
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds || !c->has_view) return fail(c, TVK_ERR_STATE, "no dataset or view");
  if (compat_active(c))
    return fail(c, TVK_ERR_UNSUPPORTED, "GridLeaper reads the volume pool's bricks (compatibility textures set)");
  HIPCHK(c, hipSetDevice(c->device));
  Grid G;
  int rc = make_grid(c, G);
  if (rc) return rc;
  if (c->ingest_pending) return fail(c, TVK_ERR_STATE, "not every brick was uploaded");
  if (budgeted(c)) {
    rc = leap_pin_coarsest(c);   // before the metadata: the coarsest level is always resident
    if (rc) return rc;
  }
  std::vector<uint32_t> meta;
  rc = make_meta(c, G, meta);
  if (rc) return rc;
  const size_t nb = c->L.toc.size();
  const size_t pixels = (size_t)c->V.W * c->V.H;
  if (!first && c->leap_pixels != pixels) return fail(c, TVK_ERR_STATE, "no earlier pass at this image size");
  // the request table: HashTableSize, or 0 for the size from the largest brick
  // (GLGridLeaper::InitHashTable, GLGridLeaper.cpp:266-279)
  uint32_t tsize = c->leap_table_size;
  if (tsize == 0) {
    const double rmax = (double)c->L.brick[0] * c->L.brick[1] * c->L.brick[2] / 32768.0;
    tsize = std::max<uint32_t>(15u, (uint32_t)(509.0 / rmax));
  }
  {
    const uint64_t l0 = (uint64_t)G.ilayout[0][0] * G.ilayout[0][1] * G.ilayout[0][2];
    if (l0 * (uint64_t)G.nlod + 1 >= (1ull << 32))
      return fail(c, TVK_ERR_UNSUPPORTED, "brick indices exceed the request table's 32-bit entries");
  }
  rc = ensure(c, &c->d_leap_meta, &c->leap_meta_cap, nb);
  if (!rc) rc = ensure(c, &c->d_leap_table, &c->leap_table_cap, tsize);
  if (!rc) rc = ensure(c, &c->d_leap_bricks, &c->leap_bricks_cap, nb);
  if (!rc) rc = ensure(c, &c->d_leap_resume, &c->leap_resume_cap, 2 * pixels);
  if (!rc) rc = ensure(c, &c->d_out, &c->out_cap, 4 * pixels);
  if (rc) return rc;
  std::vector<GatherBrick> gb(nb);
  for (size_t i = 0; i < nb; ++i) {
    for (int k = 0; k < 3; ++k) { gb[i].c[k] = (uint32_t)c->L.toc[i].coords[k]; gb[i].n[k] = c->L.toc[i].size[k]; }
    // a budgeted pool: the brick's slot (bricks without one are missing in the metadata)
    gb[i].offset = budgeted(c) ? (uint64_t)std::max(c->pst->brick_slot[i], 0) * c->pst->slot_elems : c->L.toc[i].offset;
  }
  // host sources of these copies live until the synchronisation below
  HIPCHK(c, hipMemcpyAsync(c->d_leap_meta, meta.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_leap_bricks, gb.data(), nb * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_leap_table, 0, tsize * sizeof(uint32_t), c->stream));
  c->leap_table_used = tsize;
  LeapArgs a{};
  fill_shade(c, a.S);
  a.meta = c->d_leap_meta;
  a.bricks = c->d_leap_bricks;
  a.table = c->d_leap_table;
  a.table_size = tsize;
  a.rehash = c->leap_rehash;
  a.resume = c->d_leap_resume;
  a.out = c->d_out;
  a.counter = c->count ? c->d_counter : nullptr;
  // eye -> normalised volume coordinates (GLGridLeaper.cpp:560-574)
  const ViewState& V = c->V;
  const F3 dom((float)c->L.dims[0], (float)c->L.dims[1], (float)c->L.dims[2]);
  const F3 sc(c->L.scale[0], c->L.scale[1], c->L.scale[2]);
  F3 ext = dom * sc;
  ext = ext / max3(ext);
  const Mat4 emm = inverse(V.modelview()) * translation(-0.0f, -0.0f, -0.0f) *
                   scaling(1.0f / ext.x, 1.0f / ext.y, 1.0f / ext.z) * translation(0.5f, 0.5f, 0.5f);
  const Mat4 ip = inverse(V.proj);
  // mModelToEye = emm^-1 and mModelViewIT = (modelview^-1)^T as row-vector maps
  // (GLGridLeaper.cpp:720-723)
  const Mat4 mte = inverse(emm), mvi = inverse(V.modelview());
  for (int i = 0; i < 16; ++i) { a.emm[i] = emm.a[i]; a.inv_proj[i] = ip.a[i]; a.mte[i] = mte.a[i]; }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a.nit[3 * i + j] = mvi.a[4 * j + i];   // n'_j = sum_i n_i (M^-1)^T[i][j]
  for (int l = 0; l < G.nlod; ++l) {
    for (int k = 0; k < 3; ++k) { a.layout[l][k] = G.layout[l][k]; a.ilayout[l][k] = G.ilayout[l][k]; }
    a.lodoff[l] = G.lodoff[l];
  }
  for (int k = 0; k < 3; ++k) {
    a.inner[k] = (float)(c->L.brick[k] - 2 * c->L.overlap);
    a.vsz[k] = dom[k];
  }
  // CullingLOD::SetScreenParams (CullingLOD.cpp:57-67) and GLVolumePool::Enable (:790-803)
  a.lod_factor = 2.0f * std::tan(V.fov * ((3.1416f / 180.0f) / 2.0f)) * c->P.screen_space_error / float(V.H);
  a.l0err = max3(ext / dom);
  a.srm = c->P.srm;
  const F3 eye_m = point_mul(c->leap_eye, emm);
  float ld[4] = {c->P.light_dir[0], c->P.light_dir[1], c->P.light_dir[2], 0.0f}, ml[4];
  row_mul(ld, emm, ml);
  const float mll = len3(F3(ml[0], ml[1], ml[2]));
  for (int k = 0; k < 3; ++k) { a.meye[k] = eye_m[k]; a.mlight[k] = ml[k] / mll; }
  const F3 vscale = sc / min3(sc);
  for (int k = 0; k < 3; ++k) a.S.dscale[k] = 1.0f / vscale[k];   // vDomainScale = 1/vScale
  a.nlod = G.nlod;
  a.W = (int)V.W;
  a.H = (int)V.H;
  a.first = first ? 1 : 0;
  a.lighting = c->P.lighting;
  a.mode = c->P.mode;
  a.overlap = (int)c->L.overlap;
  a.tiles_x = (V.W + 7u) / 8u;
  a.ntiles = a.tiles_x * ((V.H + 7u) / 8u);
  {
    // persistent launch: waves pull 8x8 tiles from one queue (a wave whose
    // tile finishes early takes another instead of idling: 6.3 -> 5.6 ms for
    // a 1080p pass over the 1024^3 volume's pyramid)
    int rc = ensure(c, &c->d_queues, &c->queues_cap, 8);
    if (rc) return rc;
    HIPCHK(c, launch_fill_u32(c->d_queues, 1u, 0u, c->stream));
    a.queue = c->d_queues;
  }
  if (c->count) HIPCHK(c, hipMemsetAsync(c->d_counter, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipEventRecord(c->ev[6], c->stream));
  HIPCHK(c, launch_leap(a, c->L.bpv, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[7], c->stream));
  c->leap_pixels = pixels;
  c->last_views = 1;
  c->hitpos_valid = false;
  if (host_out)
    HIPCHK(c, hipMemcpyAsync(host_out, c->d_out, pixels * 4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    std::vector<uint32_t> miss(tsize);
    std::vector<float4> res(2 * pixels);
    unsigned long long ns = 0;
    HIPCHK(c, hipMemcpyAsync(miss.data(), c->d_leap_table, tsize * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(res.data(), c->d_leap_resume, res.size() * sizeof(float4), hipMemcpyDeviceToHost,
                             c->stream));
    if (c->count) HIPCHK(c, hipMemcpyAsync(&ns, c->d_counter, sizeof(ns), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint32_t m : miss) st->missing += m ? 1 : 0;
    for (size_t p = 0; p < pixels; ++p) st->unfinished += res[2 * p].w != 1000.0f ? 1 : 0;
    st->samples = ns;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    st->ms = ms;
  } else {
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return TVK_OK;
}

int tvk_leap_missing(tvk_ctx* c, uint8_t* flags, size_t n, uint64_t* count) {
  if (!c) return TVK_ERR_ARG;
  const size_t nb = c->L.toc.size();
  if (flags && n != nb) return fail(c, TVK_ERR_ARG, "one flag per brick expected");
  std::vector<uint32_t> table;
  int rc = fetch_table(c, table);
  if (rc) return rc;
  Grid G;
  rc = make_grid(c, G);
  if (rc) return rc;
  const std::vector<uint32_t> req = decode_requests(c, G, table);
  if (flags) {
    std::fill(flags, flags + nb, (uint8_t)0);
    for (uint32_t t : req) flags[t] = 1;
  }
  if (count) *count = req.size();
  return TVK_OK;
}

int tvk_leap_page_in(tvk_ctx* c, uint64_t max_bricks, uint64_t* paged) {
  if (!c) return TVK_ERR_ARG;
  const size_t nb = c->L.toc.size();
  std::vector<uint32_t> table;
  int rc = fetch_table(c, table);
  if (rc) return rc;
  Grid G;
  rc = make_grid(c, G);
  if (rc) return rc;
  const std::vector<uint32_t> req = decode_requests(c, G, table);   // table order
  HIPCHK(c, hipSetDevice(c->device));
  if (budgeted(c)) return leap_page_budgeted(c, req, max_bricks, paged);
  if (c->leap_resident.empty()) c->leap_resident.assign(nb, 1);   // everything is resident already
  uint64_t k = 0;
  for (size_t j = 0; j < req.size() && k < max_bricks; ++j)
    if (!c->leap_resident[req[j]]) { c->leap_resident[req[j]] = 1; ++k; }
  if (paged) *paged = k;
  return TVK_OK;
}

int tvk_leap_resume_state(tvk_ctx* c, float* host, size_t floats) {
  if (!c || !host) return TVK_ERR_ARG;
  if (!c->leap_pixels || floats != 8 * c->leap_pixels) return fail(c, TVK_ERR_STATE, "no pass at this size");
  HIPCHK(c, hipMemcpyAsync(host, c->d_leap_resume, floats * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TVK_OK;
}

}  // extern "C"
