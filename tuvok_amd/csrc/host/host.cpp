// Host-side planning of the raycasting path.  See host.hpp.
#include "host.hpp"
#include <algorithm>
#include <cmath>
#include <limits>

namespace tvk {

bool Layout::build(const uint64_t d[3], const uint32_t b[3], uint32_t ov, int bytes, int max_lods,
                   std::string* err) {
  for (int i = 0; i < 3; ++i) {
    dims[i] = d[i];
    brick[i] = b[i];
    if (d[i] == 0) { if (err) *err = "volume dimension is zero"; return false; }
    if (d[i] > 0x7fffffffull) { if (err) *err = "volume dimension above 2^31-1"; return false; }
    if (b[i] <= 2 * ov) { if (err) *err = "brick size must exceed 2*overlap"; return false; }
    if (b[i] > 4096) { if (err) *err = "brick size above 4096"; return false; }
  }
  // the kernels address a brick's voxels with 32-bit byte offsets
  if (padded_count(b, bytes == 1 ? 3 : 2) * (uint64_t)bytes > 0xffffffffull) {
    if (err) *err = "brick too large: its blocked layout exceeds 4 GiB";
    return false;
  }
  overlap = ov;
  bpv = bytes;
  // ExtendedOctree::ComputeMetadata: halve (ceil) every axis > 1 until 1x1x1.
  lods.clear();
  uint64_t vs[3] = {d[0], d[1], d[2]};
  double asp[3] = {1, 1, 1};
  const uint64_t usable[3] = {b[0] - 2ull * ov, b[1] - 2ull * ov, b[2] - 2ull * ov};
  for (;;) {
    LodInfo l;
    if (!lods.empty()) {
      for (int i = 0; i < 3; ++i) {
        uint64_t nv = vs[i];
        if (vs[i] > 1) {
          nv = (vs[i] + 1) / 2;
          asp[i] *= (vs[i] % 2) ? (float(vs[i]) / float(nv)) : 2;
        }
        vs[i] = nv;
      }
      double mx = std::max(asp[0], std::max(asp[1], asp[2]));
      for (int i = 0; i < 3; ++i) asp[i] /= mx;
    }
    for (int i = 0; i < 3; ++i) {
      l.size[i] = vs[i];
      l.aspect[i] = asp[i];
      l.bricks[i] = (vs[i] + usable[i] - 1) / usable[i];
    }
    l.offset = lods.empty() ? 0 : lods.back().offset + lods.back().bricks[0] * lods.back().bricks[1] * lods.back().bricks[2];
    lods.push_back(l);
    if (vs[0] <= 1 && vs[1] <= 1 && vs[2] <= 1) break;
  }
  // brick table of every level (UVFDataset::ComputeMetadataTOC, uvfDataset.cpp:242-288);
  // rehash + insertion order as BrickedDataset::NBricksHint/AddBrick.
  size_t nall = 0;
  for (auto& l : lods) nall += (size_t)(l.bricks[0] * l.bricks[1] * l.bricks[2]);
  table.clear();
  table.rehash((size_t)std::ceil(float(nall) / table.max_load_factor()));
  for (size_t j = 0; j < lods.size(); ++j) {
    const LodInfo& l = lods[j];
    F3 corner(0, 0, 0);
    BrickMeta m;
    F3 asp3((float)l.aspect[0], (float)l.aspect[1], (float)l.aspect[2]);
    for (uint64_t x = 0; x < l.bricks[0]; ++x) {
      corner.y = 0;
      for (uint64_t y = 0; y < l.bricks[1]; ++y) {
        corner.z = 0;
        for (uint64_t z = 0; z < l.bricks[2]; ++z) {
          uint64_t c[4] = {x, y, z, (uint64_t)j};
          uint32_t bs[3];
          brick_size(c, bs);
          F3 nds = F3((float)l.size[0], (float)l.size[1], (float)l.size[2]) * asp3;
          float maxv = max3(nds);
          nds = nds / maxv;
          F3 eff((float)(bs[0] - 2 * ov), (float)(bs[1] - 2 * ov), (float)(bs[2] - 2 * ov));
          m.extents = eff * asp3 / maxv;
          m.center = (corner + m.extents / 2.0f) - nds * 0.5f;
          m.n_voxels[0] = bs[0]; m.n_voxels[1] = bs[1]; m.n_voxels[2] = bs[2];
          table.insert(std::make_pair(BrickKey(0, j, (size_t)(z * l.bricks[0] * l.bricks[1] + y * l.bricks[0] + x)), m));
          corner.z += m.extents.z;
        }
        corner.y += m.extents.y;
      }
      corner.x += m.extents.x;
    }
  }
  // keep the first max_lods levels in the pool
  size_t keep = lods.size();
  if (max_lods > 0 && (size_t)max_lods < keep) keep = (size_t)max_lods;
  toc.clear();
  uint64_t off = 0, roff = 0;
  const int lx = block_lx(bytes);
  for (size_t j = 0; j < keep; ++j) {
    const LodInfo& l = lods[j];
    for (uint64_t z = 0; z < l.bricks[2]; ++z)
      for (uint64_t y = 0; y < l.bricks[1]; ++y)
        for (uint64_t x = 0; x < l.bricks[0]; ++x) {
          TocBrick t;
          t.coords[0] = x; t.coords[1] = y; t.coords[2] = z; t.coords[3] = j;
          brick_size(t.coords, t.size);
          t.real = (uint64_t)t.size[0] * t.size[1] * t.size[2];
          t.count = padded_count(t.size, lx);
          t.offset = off;
          t.ref_offset = roff;
          off += t.count;
          roff += t.real;
          toc.push_back(t);
        }
  }
  pool_elems = off;
  ref_elems = roff;
  return true;
}

void Layout::brick_size(const uint64_t c[4], uint32_t out[3]) const {
  // ExtendedOctree::ComputeBrickSize (ExtendedOctree.cpp:276-285)
  const LodInfo& l = lods[c[3]];
  for (int i = 0; i < 3; ++i) {
    uint64_t core = brick[i] - 2ull * overlap;
    bool last = c[i] + 1 >= l.bricks[i];
    uint64_t rem = l.size[i] % core;
    out[i] = (uint32_t)((last && rem) ? 2ull * overlap + rem : brick[i]);
  }
}

uint64_t Layout::toc_index(uint64_t x, uint64_t y, uint64_t z, uint64_t lod) const {
  const LodInfo& l = lods[lod];
  return l.offset + x + y * l.bricks[0] + z * l.bricks[0] * l.bricks[1];
}

void Layout::texcoords(int lod, uint64_t lin, F3& tmin, F3& tmax, bool pow2) const {
  const BrickMeta& m = table.at(BrickKey(0, (size_t)lod, (size_t)lin));
  const LodInfo& l = lods[lod];
  float ov = (float)overlap;
  F3 nv((float)m.n_voxels[0], (float)m.n_voxels[1], (float)m.n_voxels[2]);
  if (pow2) {
    // the brick's voxels at the start of a texture of the next powers of two
    F3 real((float)next_pow2(m.n_voxels[0]), (float)next_pow2(m.n_voxels[1]), (float)next_pow2(m.n_voxels[2]));
    tmin = F3(ov, ov, ov) / real;
    tmax = F3(1, 1, 1) - tmin;
    tmax = tmax - (real - nv) / real;
  } else {
    tmin = F3(ov, ov, ov) / nv;
    tmax = F3(1, 1, 1) - tmin;
  }
  tmax = tmax * F3((float)l.aspect[0], (float)l.aspect[1], (float)l.aspect[2]);
}

void tf1d_limits(const uint8_t* tf, int n, uint64_t lim[2]) {
  // TransferFunction1D::Set(rgba8) rebuilds alpha as byte/255 (TransferFunction1D.cpp:146-177);
  // ComputeNonZeroLimits (:362-371) keeps the first/last non-zero alpha index.
  lim[0] = (uint64_t)n;
  lim[1] = 0;
  for (int i = 0; i < n; ++i)
    if (tf[4 * i + 3] != 0) {
      if ((uint64_t)i < lim[0]) lim[0] = (uint64_t)i;
      lim[1] = (uint64_t)i;
    }
}

void tf2d_limits(const uint8_t* tf, int w, int h, uint64_t lim[4]) {
  // TransferFunction2D::ComputeNonZeroLimits (TransferFunction2D.cpp:378-397)
  lim[0] = (uint64_t)w; lim[1] = 0; lim[2] = (uint64_t)h; lim[3] = 0;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x)
      if (tf[4 * ((size_t)y * w + x) + 3] != 0) {
        lim[0] = std::min<uint64_t>(lim[0], (uint64_t)x);
        lim[1] = std::max<uint64_t>(lim[1], (uint64_t)x);
        lim[2] = std::min<uint64_t>(lim[2], (uint64_t)y);
        lim[3] = std::max<uint64_t>(lim[3], (uint64_t)y);
      }
}

namespace {

struct Frustum {
  float pl[6][4];
  // CullingLOD::Update (CullingLOD.cpp:89-124), model matrix = identity
  explicit Frustum(const Mat4& mvp) {
    auto m = [&](int r, int c) { return mvp.at(r - 1, c - 1); };
    const int col[6] = {1, 1, 2, 2, 3, 3};
    const float sgn[6] = {-1, 1, -1, 1, 1, -1};
    for (int i = 0; i < 6; ++i)
      for (int r = 1; r <= 4; ++r) pl[i][r - 1] = sgn[i] * m(r, col[i]) + m(r, 4);
  }
  // CullingLOD::IsVisible (CullingLOD.cpp:141-160)
  bool visible(F3 c, F3 e) const {
    F3 h = e * 0.5f;
    for (int i = 0; i < 6; ++i) {
      const float* p = pl[i];
      if (p[0] * c.x + p[1] * c.y + p[2] * c.z + p[3] <=
          -(h.x * std::fabs(p[0]) + h.y * std::fabs(p[1]) + h.z * std::fabs(p[2])))
        return false;
    }
    return true;
  }
};

F3 corrected_scale(const Layout& L, int lod) {
  // AbstrRenderer.cpp:1003-1010 / :884-889
  F3 sc(L.scale[0], L.scale[1], L.scale[2]);
  const uint64_t* d = L.lods[lod].size;
  float mx = (float)std::max(d[0], std::max(d[1], d[2]));
  F3 dc = sc * F3((float)d[0], (float)d[1], (float)d[2]) / mx;
  return sc / max3(dc);
}

// AbstrRenderer.cpp:808-841
float brick_distance(F3 c, F3 e, const Mat4& mv) {
  const float eps = 0.4999f;
  float d = std::numeric_limits<float>::max();
  for (int i = 0; i < 8; ++i) {
    F3 s((i & 4) ? +e.x : -e.x, (i & 2) ? +e.y : -e.y, (i & 1) ? +e.z : -e.z);
    F3 q = point_mul(c + s * eps, mv);
    d = std::min(d, len3(q));
  }
  return d;
}

// AbstrRenderer::Clipped (AbstrRenderer.cpp:923-948): the eight corners through
// rotation * translation (dehomogenised); clipped when every corner satisfies
// PLANE::clip, (n . p) + w >= 0 (Vectors.h:1455-1457)
bool clipped(F3 c, F3 e, const Mat4& world, const float pl[4]) {
  for (int i = 0; i < 8; ++i) {
    const F3 s((i & 4) ? +e.x : -e.x, (i & 2) ? +e.y : -e.y, (i & 1) ? +e.z : -e.z);
    const F3 q = c + s * 0.5f;
    float v[4] = {q.x, q.y, q.z, 1.0f}, o[4];
    row_mul(v, world, o);
    const F3 w(o[0] / o[3], o[1] / o[3], o[2] / o[3]);
    if (!((pl[0] * w.x + pl[1] * w.y + pl[2] * w.z) + pl[3] >= 0.0f)) return false;
  }
  return true;
}

}  // namespace

void plan_frame(const Layout& L, const std::vector<double>& minmax, double range_max, const ViewState& V,
                const PlanParams& P, std::vector<PlannedBrick>& out) {
  out.clear();
  Mat4 mv = V.modelview();
  Frustum fr(mv * V.proj);
  const Mat4 world = V.rot * V.trans;
  F3 vScale = corrected_scale(L, 0);
  F3 vScaleVis = corrected_scale(L, P.lod);
  double r = range_max / double(P.tf_n - 1);   // AbstrRenderer.cpp:960-963
  for (auto it = L.table.begin(); it != L.table.end(); ++it) {
    if (std::get<0>(it->first) != 0 || (int)std::get<1>(it->first) != P.lod) continue;
    const BrickMeta& m = it->second;
    PlannedBrick b{};
    b.key_index = std::get<2>(it->first);
    const LodInfo& l = L.lods[P.lod];
    uint64_t li = b.key_index;
    b.toc_index = L.toc_index(li % l.bricks[0], (li % (l.bricks[0] * l.bricks[1])) / l.bricks[0],
                              li / (l.bricks[0] * l.bricks[1]), (uint64_t)P.lod);
    b.extension = m.extents * vScale;
    b.center = m.center * vScale;
    for (int i = 0; i < 3; ++i) b.voxel_count[i] = m.n_voxels[i];
    if (!fr.visible(m.center * vScaleVis, m.extents * vScaleVis)) continue;
    if (V.clip_on && clipped(m.center * vScaleVis, m.extents * vScaleVis, world, V.clip)) continue;
    const double mn = minmax[2 * b.toc_index], mx = minmax[2 * b.toc_index + 1];
    bool contains;
    if (P.mode == 0 || P.mode == 1) {
      double lo = double(P.lim[0]) * r, hi = double(P.lim[1]) * r;
      contains = (hi >= mn && lo <= mx);   // uvfDataset.cpp:1213-1227
    } else {
      contains = (double(P.isovalue) <= mx);  // uvfDataset.cpp:1201-1207
    }
    b.is_empty = contains ? 0 : 1;
    if (!b.is_empty) {
      L.texcoords(P.lod, b.key_index, b.tex_min, b.tex_max, P.pow2);
      b.distance = brick_distance(b.center, b.extension, mv);
    }
    out.push_back(b);
  }
  std::sort(out.begin(), out.end(),
            [](const PlannedBrick& a, const PlannedBrick& c) { return a.distance < c.distance; });
}

int min_lod_for_view(const Layout& L, const ViewState& V, float sse) {
  F3 dom((float)L.dims[0], (float)L.dims[1], (float)L.dims[2]);
  F3 ext = dom * F3(L.scale[0], L.scale[1], L.scale[2]);
  ext = ext / max3(ext);
  // CullingLOD::SetScreenParams (CullingLOD.cpp:57-67) and GetLODLevel (:126-138)
  float lodFactor = 2.0f * std::tan(V.fov * ((3.1416f / 180.0f) / 2.0f)) * sse / float(V.H);
  float lzero = max3(ext / dom);
  float zc = point_mul(F3(0, 0, 0), V.modelview()).z;
  float fz = std::max(V.znear, -zc);
  int lvl = (int)std::floor(std::log(lodFactor * fz / lzero) / std::log(2.0f));
  int maxl = (int)L.lods.size() - 1;
  return std::min(std::max(lvl, 0), maxl);
}

void make_view_const(const ViewState& V, ViewConst& vc) {
  Mat4 mv = V.modelview();
  Mat4 imv = inverse(mv), ip = inverse(V.proj);
  for (int i = 0; i < 16; ++i) { vc.inv_mv[i] = imv.a[i]; vc.inv_proj[i] = ip.a[i]; }
  // gl_NormalMatrix: inverse of the upper 3x3 of the modelview (column vectors)
  Mat4 m3;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m3.at(r, c) = mv.at(r, c);
  Mat4 inv3 = inverse(m3);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) vc.nm[r * 3 + c] = inv3.at(r, c);
  vc.znear = V.znear;
  vc.zfar = V.zfar;
  vc.clip_on = V.clip_on ? 1u : 0u;
  if (V.clip_on) model_clip_plane(V.clip, V.rot, V.trans, vc.clip);
  else vc.clip[0] = vc.clip[1] = vc.clip[2] = vc.clip[3] = 0.0f;
  vc.W = V.W;
  vc.H = V.H;
  vc.tiles_x = (V.W + 7) / 8;
  vc.groups_x = (V.W + 15) / 16;
  vc.groups = vc.groups_x * ((V.H + 15) / 16);
}

void group_costs(const ViewConst& vc, const FrameBrick* list, float* cost) {
  const int gx = (int)vc.groups_x, gy = gx > 0 ? (int)(vc.groups / vc.groups_x) : 0;
  if (gx <= 0 || gy <= 0) return;
  // 2D difference array over the groups, then prefix sums
  std::vector<double> d((size_t)(gx + 1) * (size_t)(gy + 1), 0.0);
  auto at = [&](int x, int y) -> double& { return d[(size_t)y * (size_t)(gx + 1) + (size_t)x]; };
  double unit = 0.0;   // the largest brick weight
  for (uint32_t i = 0; i < vc.nlist; ++i) {
    const FrameBrick& f = list[i];
    if (f.rect[2] < f.rect[0] || f.rect[3] < f.rect[1] || !(f.ray_step > 0.0f)) continue;
    const int x0 = std::min(std::max((int)f.rect[0], 0) / 16, gx - 1), x1 = std::min(std::max((int)f.rect[2], 0) / 16, gx - 1);
    const int y0 = std::min(std::max((int)f.rect[1], 0) / 16, gy - 1), y1 = std::min(std::max((int)f.rect[3], 0) / 16, gy - 1);
    const float ext = std::max(f.bmax[0] - f.bmin[0], std::max(f.bmax[1] - f.bmin[1], f.bmax[2] - f.bmin[2]));
    const double w = (double)(ext / f.ray_step);
    if (!(w > 0.0) || !std::isfinite(w)) continue;
    unit = std::max(unit, w);
    at(x0, y0) += w;
    at(x1 + 1, y0) -= w;
    at(x0, y1 + 1) -= w;
    at(x1 + 1, y1 + 1) += w;
  }
  const double base = 0.05 * unit;
  // in-place 2D prefix sums: the summed weight of the rectangles covering each group
  for (int y = 0; y < gy; ++y)
    for (int x = 0; x < gx; ++x) {
      double v = at(x, y);
      if (x > 0) v += at(x - 1, y);
      if (y > 0) v += at(x, y - 1);
      if (x > 0 && y > 0) v -= at(x - 1, y - 1);
      at(x, y) = v;
      cost[(size_t)y * (size_t)gx + (size_t)x] = (float)(std::max(v, 0.0) + base);
    }
}

void dispatch_order(const ViewConst* views, int nviews, const std::vector<float>& cost, int mode,
                    std::vector<uint32_t>& order, uint32_t start[8], uint32_t count[8]) {
  // units: one group (mode 1) or one row of a view's groups (mode 2), in
  // launch order; a unit covers groups first .. first + n - 1
  struct Unit { float cost; uint32_t first, n; };
  std::vector<Unit> units;
  size_t ngroups = 0;
  for (int v = 0; v < nviews; ++v) {
    const ViewConst& vc = views[v];
    const uint32_t gx = vc.groups_x, gy = gx > 0 ? vc.groups / gx : 0;
    for (uint32_t y = 0; y < gy; ++y) {
      const uint32_t row = vc.group_offset + y * gx;
      if (mode == 2) {
        double sum = 0.0;
        for (uint32_t x = 0; x < gx; ++x) sum += cost[row + x];
        units.push_back(Unit{(float)sum, row, gx});
      } else {
        for (uint32_t x = 0; x < gx; ++x) units.push_back(Unit{cost[row + x], row + x, 1u});
      }
    }
    ngroups += (size_t)gx * gy;
  }
  order.clear();
  order.reserve(ngroups);
  // 8 contiguous ranges of about ngroups / 8 groups; each range's units by
  // decreasing cost, bucketed into 1024 levels of its largest cost (ties and
  // equal buckets in launch order)
  constexpr int NB = 1024;
  std::vector<uint32_t> bucket_n(NB + 1);
  std::vector<uint32_t> idx;
  size_t b = 0, acc = 0;
  for (int x = 0; x < 8; ++x) {
    start[x] = (uint32_t)order.size();
    const size_t b0 = b;
    const double target = (double)ngroups * (double)(x + 1) / 8.0;
    while (b < units.size() && (x == 7 || (double)acc + 0.5 * (double)units[b].n < target)) acc += units[b++].n;
    float top = 0.0f;
    for (size_t i = b0; i < b; ++i) top = std::max(top, units[i].cost);
    auto level = [&](float c) {
      const int l = top > 0.0f ? (int)((double)c / (double)top * (NB - 1) + 0.5) : 0;
      return NB - 1 - std::min(std::max(l, 0), NB - 1);   // 0: the most costly
    };
    std::fill(bucket_n.begin(), bucket_n.end(), 0u);
    for (size_t i = b0; i < b; ++i) ++bucket_n[(size_t)level(units[i].cost) + 1];
    for (int l = 0; l < NB; ++l) bucket_n[(size_t)l + 1] += bucket_n[(size_t)l];
    idx.assign(b - b0, 0u);
    for (size_t i = b0; i < b; ++i) idx[bucket_n[(size_t)level(units[i].cost)]++] = (uint32_t)i;
    for (uint32_t i : idx)
      for (uint32_t q = 0; q < units[i].n; ++q) order.push_back(units[i].first + q);
    count[x] = (uint32_t)(order.size() - start[x]);
  }
}

void make_frame_bricks(const Layout& L, const ViewState& V, const PlanParams& P,
                       const std::vector<PlannedBrick>& list, std::vector<FrameBrick>& out,
                       const uint32_t* tex_dims, const uint64_t* tex_offset) {
  Mat4 mv = V.modelview();
  Mat4 imv = inverse(mv);
  Mat4 mvp = mv * V.proj;
  const uint64_t* d0 = L.lods[0].size;
  const uint64_t* dl = L.lods[P.lod].size;
  for (const PlannedBrick& b : list) {
    if (b.is_empty) continue;   // GLRaycaster::Render3DInLoop returns for empty bricks
    FrameBrick f{};
    F3 bmin = b.center - b.extension / 2.0f, bmax = b.center + b.extension / 2.0f;
    F3 sc = (b.tex_min - b.tex_max) / (bmin - bmax);
    // GLRaycaster::ComputeEyeToTextureMatrix(p1=max, t1=max coords, p2=min, t2=min coords)
    Mat4 e2t = imv * translation(-bmax.x, -bmax.y, -bmax.z) * scaling(sc.x, sc.y, sc.z) *
               translation(b.tex_max.x, b.tex_max.y, b.tex_max.z);
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 3; ++c) f.e2t[r * 3 + c] = e2t.at(r, c);
    for (int i = 0; i < 3; ++i) { f.bmin[i] = bmin[i]; f.bmax[i] = bmax[i]; }
    // GLRaycaster::SetBrickDepShaderVars (GLRaycaster.cpp:239-257): one texel of
    // the brick's texture (padded to powers of two with bUseOnlyPowerOfTwo)
    uint32_t tn[3] = {b.voxel_count[0], b.voxel_count[1], b.voxel_count[2]};
    if (tex_dims)
      for (int i = 0; i < 3; ++i) tn[i] = tex_dims[3 * b.toc_index + i];
    F3 vox = F3(1, 1, 1) / F3((float)tn[0], (float)tn[1], (float)tn[2]);
    f.ray_step = min3(b.extension * vox * 0.5f * (1.0f / P.srm));
    F3 q = F3((float)d0[0], (float)d0[1], (float)d0[2]) / F3((float)dl[0], (float)dl[1], (float)dl[2]);
    f.step_scale = 1.0f / P.srm * max3(q);
    for (int i = 0; i < 3; ++i) { f.vstep[i] = vox[i]; f.n[i] = tn[i]; }
    const TocBrick& t = L.toc[b.toc_index];
    f.offset = tex_offset ? tex_offset[b.toc_index] : t.offset;
    f.toc = (uint32_t)b.toc_index;
    // conservative pixel rectangle of the projected box
    double xmin = 1e30, xmax = -1e30, ymin = 1e30, ymax = -1e30;
    bool behind = false;
    for (int i = 0; i < 8; ++i) {
      float v[4] = {(i & 1) ? bmax.x : bmin.x, (i & 2) ? bmax.y : bmin.y, (i & 4) ? bmax.z : bmin.z, 1.f}, o[4];
      row_mul(v, mvp, o);
      if (!(o[3] > 1e-6f)) { behind = true; break; }
      double wx = ((double)o[0] / o[3] + 1.0) * 0.5 * V.W, wy = ((double)o[1] / o[3] + 1.0) * 0.5 * V.H;
      xmin = std::min(xmin, wx); xmax = std::max(xmax, wx);
      ymin = std::min(ymin, wy); ymax = std::max(ymax, wy);
    }
    int x0 = 0, y0 = 0, x1 = (int)V.W - 1, y1 = (int)V.H - 1;
    if (!behind) {
      x0 = std::max(x0, (int)std::floor(xmin - 0.5) - 1);
      y0 = std::max(y0, (int)std::floor(ymin - 0.5) - 1);
      x1 = std::min(x1, (int)std::ceil(xmax - 0.5) + 1);
      y1 = std::min(y1, (int)std::ceil(ymax - 0.5) + 1);
    }
    if (x1 < x0 || y1 < y0) { x0 = 1; x1 = 0; y0 = 1; y1 = 0; }  // off screen: never matches
    f.rect[0] = (int16_t)std::max(-32768, std::min(32767, x0));
    f.rect[1] = (int16_t)std::max(-32768, std::min(32767, y0));
    f.rect[2] = (int16_t)std::max(-32768, std::min(32767, x1));
    f.rect[3] = (int16_t)std::max(-32768, std::min(32767, y1));
    out.push_back(f);
  }
}

}  // namespace tvk
