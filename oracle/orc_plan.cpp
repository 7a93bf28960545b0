// Oracle: transfer-function tables and frame planning.
// Test infrastructure only (see orc_common.hpp).
//   TransferFunction1D::GetByteArray      TransferFunction1D.cpp:311-330
//   TransferFunction1D::Set(rgba8)        TransferFunction1D.cpp:146-177
//   TransferFunction1D::ComputeNonZeroLimits  TransferFunction1D.cpp:362-371
//   TransferFunction2D::ComputeNonZeroLimits  TransferFunction2D.cpp:378-397
//   CullingLOD::SetScreenParams/Update/IsVisible/GetLODLevel  CullingLOD.cpp:57-160
//   AbstrRenderer::ContainsData / RegionNeedsBrick / brick_distance /
//     BuildSubFrameBrickList / ComputeMinLODForCurrentView
//                                         AbstrRenderer.cpp:789-1100
//   GLRaycaster::SetBrickDepShaderVars    GLRaycaster.cpp:239-300
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>
#include "orc_internal.hpp"

using namespace orc;

extern "C" void orc_tf1d_bytes(const float* rgba, int n, uint8_t* out, int used_range) {
  for (int i = 0; i < 4 * n; ++i)
    out[i] = (unsigned char)(std::max(0.0f, std::min(rgba[i], 1.0f)) * (unsigned char)used_range);
}

extern "C" void orc_tf1d_limits(const uint8_t* tf, int n, uint64_t lim[2]) {
  // Set(rgba8) rebuilds alpha as lerp(byte, 0, 255, 0, 1); the limit test is
  // on that float being != 0, i.e. on the byte being non-zero.
  lim[0] = (uint64_t)n;
  lim[1] = 0;
  for (int i = 0; i < n; ++i) {
    float a = 0.0f + (float(tf[4 * i + 3]) - 0.0f) * (1.0f - 0.0f) / (255.0f - 0.0f);
    if (a != 0) {
      lim[0] = std::min<uint64_t>(lim[0], (uint64_t)i);
      lim[1] = (uint64_t)i;
    }
  }
}

extern "C" void orc_tf2d_limits(const uint8_t* tf, int w, int h, uint64_t lim[4]) {
  lim[0] = (uint64_t)w; lim[1] = 0; lim[2] = (uint64_t)h; lim[3] = 0;
  size_t i = 3;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      if (tf[i] != 0) {
        lim[0] = std::min<uint64_t>(lim[0], (uint64_t)x);
        lim[1] = std::max<uint64_t>(lim[1], (uint64_t)x);
        lim[2] = std::min<uint64_t>(lim[2], (uint64_t)y);
        lim[3] = std::max<uint64_t>(lim[3], (uint64_t)y);
      }
      i += 4;
    }
}

namespace {

M4 from_arr(const float* a) { M4 m; for (int i = 0; i < 16; ++i) m.a[i] = a[i]; return m; }

struct Cull {
  M4 mvp, mv;
  float planes[6][4];
  void update(const M4& modelview, const M4& proj) {
    // CullingLOD.cpp:89-124 (model matrix = identity)
    mv = modelview;
    mvp = mul(mv, proj);
    auto m = [&](int r, int c) { return mvp.m(r - 1, c - 1); };
    float P[6][4] = {
        {-m(1, 1) + m(1, 4), -m(2, 1) + m(2, 4), -m(3, 1) + m(3, 4), -m(4, 1) + m(4, 4)},
        {m(1, 1) + m(1, 4), m(2, 1) + m(2, 4), m(3, 1) + m(3, 4), m(4, 1) + m(4, 4)},
        {-m(1, 2) + m(1, 4), -m(2, 2) + m(2, 4), -m(3, 2) + m(3, 4), -m(4, 2) + m(4, 4)},
        {m(1, 2) + m(1, 4), m(2, 2) + m(2, 4), m(3, 2) + m(3, 4), m(4, 2) + m(4, 4)},
        {m(1, 3) + m(1, 4), m(2, 3) + m(2, 4), m(3, 3) + m(3, 4), m(4, 3) + m(4, 4)},
        {-m(1, 3) + m(1, 4), -m(2, 3) + m(2, 4), -m(3, 3) + m(3, 4), -m(4, 3) + m(4, 4)}};
    for (int i = 0; i < 6; ++i)
      for (int k = 0; k < 4; ++k) planes[i][k] = P[i][k];
  }
  bool visible(const V3& c, const V3& ext) const {
    // CullingLOD.cpp:141-160
    V3 h = ext * 0.5f;
    for (int i = 0; i < 6; ++i) {
      const float* p = planes[i];
      if (p[0] * c.x + p[1] * c.y + p[2] * c.z + p[3] <=
          -(h.x * std::fabs(p[0]) + h.y * std::fabs(p[1]) + h.z * std::fabs(p[2])))
        return false;
    }
    return true;
  }
};

M4 modelview_of(const orc_view* v) {
  // GLRenderer.cpp:627: modelView = rotation * translation * view
  return mul(mul(from_arr(v->rot), from_arr(v->trans)), from_arr(v->view));
}

float brick_distance(const V3& center, const V3& ext, const M4& mv) {
  // AbstrRenderer.cpp:808-841
  const float eps = 0.4999f;
  float d = std::numeric_limits<float>::max();
  for (int i = 0; i < 8; ++i) {
    V3 s((i & 4) ? +ext.x : -ext.x, (i & 2) ? +ext.y : -ext.y, (i & 1) ? +ext.z : -ext.z);
    V3 p = center + s * eps;
    V3 q = xform_point(p, mv);
    d = std::min(d, length(q));
  }
  return d;
}

// AbstrRenderer::Clipped (AbstrRenderer.cpp:923-948): the brick's eight corners
// through rotation * translation (dehomogenised); clipped when every corner
// satisfies PLANE::clip, (n . p) + w >= 0 (Vectors.h:1455-1457).
bool clipped(const V3& c, const V3& e, const M4& world, const float pl[4]) {
  for (int i = 0; i < 8; ++i) {
    const V3 s((i & 4) ? +e.x : -e.x, (i & 2) ? +e.y : -e.y, (i & 1) ? +e.z : -e.z);
    const V3 q = c + s * 0.5f;
    float v[4] = {q.x, q.y, q.z, 1.0f}, o[4];
    vmul4(v, world, o);
    const V3 w(o[0] / o[3], o[1] / o[3], o[2] / o[3]);
    if (!((pl[0] * w.x + pl[1] * w.y + pl[2] * w.z) + pl[3] >= 0.0f)) return false;
  }
  return true;
}

}  // namespace

extern "C" int orc_plan(const orc_ds* ds, const orc_view* view, const orc_params* prm,
                        orc_brick* out, int cap) {
  const Octree& t = ds->tree;
  M4 mv = modelview_of(view);
  const M4 world = mul(from_arr(view->rot), from_arr(view->trans));
  Cull cull;
  cull.update(mv, from_arr(view->proj));

  // AbstrRenderer.cpp:999-1016 (vScale from LOD0 domain)
  auto corrected = [&](int lod) {
    V3 sc(ds->scale[0], ds->scale[1], ds->scale[2]);
    const uint64_t* d = t.lods[lod].size;
    float mx = (float)std::max(d[0], std::max(d[1], d[2]));
    V3 dc = sc * V3((float)d[0], (float)d[1], (float)d[2]) / mx;
    return sc / dc.maxVal();
  };
  V3 vScale = corrected(0);
  V3 vScaleVis = corrected(prm->lod);  // RegionNeedsBrick uses the current LOD

  double range_max = orc_ds_range_max(ds);
  int tf_n = prm->mode == 1 ? prm->tf2d_w : prm->tf1d_n;
  // AbstrRenderer.cpp:960-963 (range is always valid for TOC datasets here)
  double fMaxValue = range_max;
  double r = fMaxValue / double(tf_n - 1);
  uint64_t lim[4] = {0, 0, 0, 0};
  if (prm->mode == 0) orc_tf1d_limits(prm->tf1d, prm->tf1d_n, lim);
  else if (prm->mode == 1) orc_tf2d_limits(prm->tf2d, prm->tf2d_w, prm->tf2d_h, lim);

  std::vector<orc_brick> list;
  for (auto it = ds->table.begin(); it != ds->table.end(); ++it) {
    if (std::get<0>(it->first) != 0 || (int)std::get<1>(it->first) != prm->lod) continue;
    const BrickMD& bmd = it->second;
    orc_brick b;
    std::memset(&b, 0, sizeof(b));
    V3 ext = bmd.extents * vScale, cen = bmd.center * vScale;
    b.key_index = std::get<2>(it->first);
    uint64_t c[4];
    {
      const uint64_t* bc = t.lods[prm->lod].bricks;
      uint64_t li = b.key_index;
      c[0] = li % bc[0]; c[1] = (li % (bc[0] * bc[1])) / bc[0]; c[2] = li / (bc[0] * bc[1]); c[3] = (uint64_t)prm->lod;
    }
    b.toc_index = t.coords_to_index(c);
    for (int i = 0; i < 3; ++i) { b.center[i] = cen[i]; b.extension[i] = ext[i]; b.voxel_count[i] = bmd.n_voxels[i]; }
    // RegionNeedsBrick (AbstrRenderer.cpp:874-920)
    if (!cull.visible(bmd.center * vScaleVis, bmd.extents * vScaleVis)) continue;
    if (prm->clip_on && clipped(bmd.center * vScaleVis, bmd.extents * vScaleVis, world, prm->clip)) continue;
    // ContainsData (AbstrRenderer.cpp:953-997; uvfDataset.cpp:1201-1227)
    const double* mm = t.toc[b.toc_index].minmax;
    bool contains;
    if (prm->mode == 0 || prm->mode == 1) {
      double fMin = double(lim[0]) * r, fMax = double(lim[1]) * r;
      contains = (fMax >= mm[0] && fMin <= mm[1]);
    } else {
      contains = (double(prm->isovalue) <= mm[1]);
    }
    b.is_empty = contains ? 0 : 1;
    if (!b.is_empty) {
      V3 a, z;
      ds->texcoords(prm->lod, b.key_index, prm->pow2 != 0, a, z);
      for (int i = 0; i < 3; ++i) { b.tex_min[i] = a[i]; b.tex_max[i] = z[i]; }
      b.distance = brick_distance(cen, ext, mv);
    }
    list.push_back(b);
  }
  // AbstrRenderer.cpp:1097 (operator< on fDistance, AbstrRenderer.h:104-106)
  std::sort(list.begin(), list.end(),
            [](const orc_brick& l, const orc_brick& r) { return l.distance < r.distance; });
  if ((int)list.size() > cap) return -(int)list.size();
  for (size_t i = 0; i < list.size(); ++i) out[i] = list[i];
  return (int)list.size();
}

extern "C" void orc_cull_planes(const orc_view* view, float planes[24]) {
  Cull cull;
  cull.update(modelview_of(view), from_arr(view->proj));
  for (int i = 0; i < 6; ++i)
    for (int k = 0; k < 4; ++k) planes[4 * i + k] = cull.planes[i][k];
}

extern "C" int orc_cull_visible(const orc_view* view, const float c[3], const float e[3]) {
  Cull cull;
  cull.update(modelview_of(view), from_arr(view->proj));
  return cull.visible(V3(c[0], c[1], c[2]), V3(e[0], e[1], e[2])) ? 1 : 0;
}

extern "C" int orc_cull_lod_level(const orc_view* view, float sse, const float c[3], const float e[3],
                                  const uint32_t vox[3]) {
  // CullingLOD::SetScreenParams (CullingLOD.cpp:57-67) and GetLODLevel (:126-138)
  float lodFactor = 2.0f * std::tan(view->fov * ((3.1416f / 180.0f) / 2.0f)) * sse / float(view->height);
  V3 ext(e[0], e[1], e[2]);
  float lzero = (ext / V3((float)vox[0], (float)vox[1], (float)vox[2])).maxVal();
  float zc = xform_point(V3(c[0], c[1], c[2]), modelview_of(view)).z;
  float fz = std::max(view->znear, -zc);
  return (int)std::floor(std::log(lodFactor * fz / lzero) / std::log(2.0f));
}

extern "C" int orc_min_lod(const orc_ds* ds, const orc_view* view, float sse) {
  // AbstrRenderer.cpp:789-803: the level of the whole domain at the origin,
  // clamped to [0, levels - 1]
  const Octree& t = ds->tree;
  V3 dom((float)t.dims[0], (float)t.dims[1], (float)t.dims[2]);
  V3 sc(ds->scale[0], ds->scale[1], ds->scale[2]);
  V3 ext = dom * sc;
  ext = ext / ext.maxVal();
  const float c0[3] = {0, 0, 0}, e0[3] = {ext.x, ext.y, ext.z};
  const uint32_t vox[3] = {(uint32_t)t.dims[0], (uint32_t)t.dims[1], (uint32_t)t.dims[2]};
  int lvl = orc_cull_lod_level(view, sse, c0, e0, vox);
  int maxl = (int)t.lods.size() - 1;
  return std::min(std::max(lvl, 0), maxl);
}

extern "C" void orc_brick_shader_vars(const orc_ds* ds, const orc_params* prm, const orc_brick* b,
                                      float* ray_step, float* step_scale, float vstep[3]) {
  // GLRaycaster.cpp:239-257 (no sample-rate drop): one texel of the brick's
  // texture, padded to powers of two with bUseOnlyPowerOfTwo
  uint32_t tn[3] = {b->voxel_count[0], b->voxel_count[1], b->voxel_count[2]};
  if (prm->pow2)
    for (int i = 0; i < 3; ++i) tn[i] = next_pow2(tn[i]);
  V3 vox = V3(1, 1, 1) / V3((float)tn[0], (float)tn[1], (float)tn[2]);
  float srm = prm->srm;
  V3 ext(b->extension[0], b->extension[1], b->extension[2]);
  *ray_step = (ext * vox * 0.5f * (1.0f / srm)).minVal();
  const uint64_t* d0 = ds->tree.lods[0].size;
  const uint64_t* dl = ds->tree.lods[prm->lod].size;
  V3 q = V3((float)d0[0], (float)d0[1], (float)d0[2]) / V3((float)dl[0], (float)dl[1], (float)dl[2]);
  *step_scale = 1.0f / srm * q.maxVal();
  for (int i = 0; i < 3; ++i) vstep[i] = vox[i];
}

extern "C" void orc_look_at(const float eye[3], const float at[3], const float up[3], float out[16]) {
  M4 m = look_at(V3(eye[0], eye[1], eye[2]), V3(at[0], at[1], at[2]), V3(up[0], up[1], up[2]));
  for (int i = 0; i < 16; ++i) out[i] = m.a[i];
}
extern "C" void orc_perspective(float fovy, float aspect, float n, float f, float out[16]) {
  M4 m = perspective(fovy, aspect, n, f);
  for (int i = 0; i < 16; ++i) out[i] = m.a[i];
}
extern "C" void orc_rotation_y(double angle, float out[16]) {
  M4 m = rotation_y(angle);
  for (int i = 0; i < 16; ++i) out[i] = m.a[i];
}
