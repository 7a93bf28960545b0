// C ABI of the GridLeaper-style traversal (include/tvk_hip.h, "ray-guided traversal").
//
// Host half of GLGridLeaper + GLVolumePool: brick IDs and layouts
// (GLVolumePool.cpp:89-117, :205-223, :288-291), the brick metadata with the
// visibility of the current TF (:961-1100), residency and paging of requested
// bricks (:673-717, :1720-1789: here the bricks already sit in the HBM pool and
// paging marks them resident), and the shader uniforms (GLGridLeaper.cpp:560-574,
// :700-745; GLVolumePool.cpp:790-803).
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
            const bool res = c->leap_resident.empty() || c->leap_resident[i] || l == G.nlod - 1;
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

}  // namespace

extern "C" {

int tvk_leap_set_resident(tvk_ctx* c, const uint8_t* flags, size_t n) {
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds) return fail(c, TVK_ERR_STATE, "no dataset");
  if (!flags) { c->leap_resident.clear(); return TVK_OK; }
  if (n != c->L.toc.size()) return fail(c, TVK_ERR_ARG, "one flag per brick expected");
  c->leap_resident.assign(flags, flags + n);
  return TVK_OK;
}

int tvk_leap_render(tvk_ctx* c, int first, float* host_out, tvk_leap_stats* st) {
  if (!c) return TVK_ERR_ARG;
  if (!c->has_ds || !c->has_view) return fail(c, TVK_ERR_STATE, "no dataset or view");
  HIPCHK(c, hipSetDevice(c->device));
  Grid G;
  int rc = make_grid(c, G);
  if (rc) return rc;
  std::vector<uint32_t> meta;
  rc = make_meta(c, G, meta);
  if (rc) return rc;
  const size_t nb = c->L.toc.size();
  const size_t pixels = (size_t)c->V.W * c->V.H;
  if (!first && c->leap_pixels != pixels) return fail(c, TVK_ERR_STATE, "no earlier pass at this image size");
  rc = ensure(c, &c->d_leap_meta, &c->leap_meta_cap, nb);
  if (!rc) rc = ensure(c, &c->d_leap_missing, &c->leap_missing_cap, nb);
  if (!rc) rc = ensure(c, &c->d_leap_bricks, &c->leap_bricks_cap, nb);
  if (!rc) rc = ensure(c, &c->d_leap_resume, &c->leap_resume_cap, 2 * pixels);
  if (!rc) rc = ensure(c, &c->d_out, &c->out_cap, 4 * pixels);
  if (rc) return rc;
  std::vector<GatherBrick> gb(nb);
  for (size_t i = 0; i < nb; ++i) {
    for (int k = 0; k < 3; ++k) { gb[i].c[k] = (uint32_t)c->L.toc[i].coords[k]; gb[i].n[k] = c->L.toc[i].size[k]; }
    gb[i].offset = c->L.toc[i].offset;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_leap_meta, meta.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_leap_bricks, gb.data(), nb * sizeof(GatherBrick), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_leap_missing, 0, nb * sizeof(uint32_t), c->stream));
  LeapArgs a{};
  fill_shade(c, a.S);
  a.meta = c->d_leap_meta;
  a.bricks = c->d_leap_bricks;
  a.missing = c->d_leap_missing;
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
  if (c->count) HIPCHK(c, hipMemsetAsync(c->d_counter, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipEventRecord(c->ev[6], c->stream));
  HIPCHK(c, launch_leap(a, c->L.bpv, c->stream));
  HIPCHK(c, hipEventRecord(c->ev[7], c->stream));
  c->leap_pixels = pixels;
  c->last_views = 1;
  if (host_out)
    HIPCHK(c, hipMemcpyAsync(host_out, c->d_out, pixels * 4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    std::vector<uint32_t> miss(nb);
    std::vector<float4> res(2 * pixels);
    unsigned long long ns = 0;
    HIPCHK(c, hipMemcpyAsync(miss.data(), c->d_leap_missing, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
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
  if (!c->d_leap_missing || c->leap_missing_cap < c->L.toc.size()) return fail(c, TVK_ERR_STATE, "no pass yet");
  const size_t nb = c->L.toc.size();
  if (flags && n != nb) return fail(c, TVK_ERR_ARG, "one flag per brick expected");
  std::vector<uint32_t> miss(nb);
  HIPCHK(c, hipMemcpyAsync(miss.data(), c->d_leap_missing, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  uint64_t k = 0;
  for (size_t i = 0; i < nb; ++i) {
    if (flags) flags[i] = miss[i] ? 1 : 0;
    k += miss[i] ? 1 : 0;
  }
  if (count) *count = k;
  return TVK_OK;
}

int tvk_leap_page_in(tvk_ctx* c, uint64_t max_bricks, uint64_t* paged) {
  if (!c) return TVK_ERR_ARG;
  const size_t nb = c->L.toc.size();
  std::vector<uint8_t> miss(nb);
  int rc = tvk_leap_missing(c, miss.data(), nb, nullptr);
  if (rc) return rc;
  if (c->leap_resident.empty()) c->leap_resident.assign(nb, 1);   // everything is resident already
  uint64_t k = 0;
  for (size_t i = 0; i < nb && k < max_bricks; ++i)
    if (miss[i] && !c->leap_resident[i]) { c->leap_resident[i] = 1; ++k; }
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
