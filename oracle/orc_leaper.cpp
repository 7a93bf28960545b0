// Oracle of the GridLeaper single-pass traversal (test infrastructure only).
//
// CPU restatement of GLGridLeaper's ray-guided march through the volume pool:
//   - brick IDs, float LOD layouts and LOD offsets: GLVolumePool.cpp:89-117, :205-223, :288-291
//   - brick visibility (BI_MISSING / BI_CHILD_EMPTY / BI_EMPTY / resident): GLVolumePool.cpp:961-1100
//   - GetBrick / BrickExit / NormCoordsToPoolCoords / ComputeLOD / TransformToPoolSpace:
//     GLVolumePool.cpp:484-656 (the generated shader text)
//   - the march, miss handling, resume state and termination: GLGridLeaper-blend.glsl:65-228
//   - sampling methods: GLGridLeaper-Method-{1D,1D-L,2D,2D-L}.glsl, GLGridLeaper-GradientTools.glsl
//   - eye->model matrix, LoD factor, level-zero error, model-space lights: GLGridLeaper.cpp:560-574,
//     :700-745; GLVolumePool.cpp:790-803
// The pool atlas is not modelled: a brick is sampled in its own voxel
// coordinates (voxel b of the stored brick is texel b; the texture coordinate
// of the atlas is b/poolSize offset by the slot), which is the same point.
// Ray entry and exit are the analytic intersections with the volume box (the
// reference rasterises the box faces), clipped to the near plane.
#include <cfloat>
#include <limits>
#include <cmath>
#include <cstring>
#include <vector>
#include "orc_api.h"
#include "orc_common.hpp"
#include "orc_internal.hpp"

using namespace orc;

namespace {

M4 from_arr(const float* a) { M4 m; for (int i = 0; i < 16; ++i) m.a[i] = a[i]; return m; }

enum { BI_MISSING = 0, BI_CHILD_EMPTY = 1, BI_EMPTY = 2, BI_FLAG_COUNT = 3 };

struct Leaper {
  const Octree* t;
  const orc_params* p;
  int nlod;
  V3 inner;                        // maxInnerBrickSize
  std::vector<V3> layout;          // GetFloatBrickLayout
  std::vector<uint64_t> ilayout;   // GetBrickLayout, 3 per level
  std::vector<uint64_t> lodoff;    // m_vLoDOffsetTable
  std::vector<uint32_t> meta;      // per brick ID
  M4 emm, proj, inv_proj, mte, mvi;
  int W, H;
  float lod_factor, l0err, srm;
  V3 mlight, meye, dscale, amb, dif, spc;
  float tf_scale, grad_scale;
  float iso_n;
  V3 iso_col, ldir;

  uint64_t id(uint64_t x, uint64_t y, uint64_t z, int l) const {
    const uint64_t* b = &ilayout[3 * (size_t)l];
    return lodoff[(size_t)l] + x + y * b[0] + z * b[0] * b[1];
  }
};

float tf_alpha_limits_scale(const Octree& t, const orc_params* p, double& fmin, double& fmax) {
  // GLGridLeaper::RecomputeBrickVisibility (GLGridLeaper.cpp:642-660)
  double rmin = 1e300, rmax = -1e300;
  for (uint64_t i = 0; i < t.lods[0].bricks[0] * t.lods[0].bricks[1] * t.lods[0].bricks[2]; ++i) {
    rmin = std::min(rmin, t.toc[i].minmax[0]);
    rmax = std::max(rmax, t.toc[i].minmax[1]);
  }
  const int n = p->mode == 1 ? p->tf2d_w : p->tf1d_n;
  const double fMaxValue = rmin > rmax ? (double)n : rmax;
  const double resc = fMaxValue / double(n - 1);
  uint64_t lim[4] = {0, 0, 0, 0};
  if (p->mode == 1) orc_tf2d_limits(p->tf2d, p->tf2d_w, p->tf2d_h, lim);
  else orc_tf1d_limits(p->tf1d, p->tf1d_n, lim);
  fmin = double(lim[0]) * resc;
  fmax = double(lim[1]) * resc;
  return 0.0f;
}

bool setup(Leaper& L, const orc_ds* ds, const orc_view* v, const orc_params* p, const uint8_t* resident) {
  const Octree& t = ds->tree;
  L.t = &t;
  L.p = p;
  L.nlod = (int)t.lods.size();
  L.inner = V3((float)(t.brick[0] - 2 * t.overlap), (float)(t.brick[1] - 2 * t.overlap),
               (float)(t.brick[2] - 2 * t.overlap));
  const uint64_t vs[3] = {t.dims[0], t.dims[1], t.dims[2]};
  uint64_t off = 0;
  for (int l = 0; l < L.nlod; ++l) {
    // GetFloatBrickLayout (GLVolumePool.cpp:89-108)
    float b[3];
    for (int a = 0; a < 3; ++a) {
      b[a] = float(vs[a]) / L.inner[a];
      b[a] /= float(1u << l);
      if (float(uint32_t(b[a])) == b[a]) b[a] -= b[a] * std::numeric_limits<float>::epsilon();
    }
    L.layout.push_back(V3(b[0], b[1], b[2]));
    // GetBrickLayout (:110-117) with GetLoDSize (:83-88)
    for (int a = 0; a < 3; ++a) {
      const uint64_t base = (uint64_t)std::ceil(double(vs[a]) / L.inner[a]);
      L.ilayout.push_back((uint64_t)std::ceil(double(base) / double(1ull << l)));
    }
    L.lodoff.push_back(off);
    off += L.ilayout[3 * l] * L.ilayout[3 * l + 1] * L.ilayout[3 * l + 2];
  }
  if (off != t.toc.size()) return false;   // the pool's brick grid must be the octree's
  // visibility (GLVolumePool.cpp:961-1100): scalar range test, child-empty propagation
  double fmin = 0, fmax = 0;
  if (p->mode == 2) {
    // ContainsData for RM_ISOSURFACE (GLVolumePool.cpp:983-985): min <= isovalue <= max
    fmin = fmax = (double)p->isovalue;
  } else {
    tf_alpha_limits_scale(t, p, fmin, fmax);
  }
  std::vector<uint8_t> empty(t.toc.size(), 0);
  for (size_t i = 0; i < t.toc.size(); ++i)
    empty[i] = !(fmax >= t.toc[i].minmax[0] && fmin <= t.toc[i].minmax[1]);
  L.meta.assign(t.toc.size(), BI_MISSING);
  for (int l = 0; l < L.nlod; ++l) {
    const uint64_t* b = &L.ilayout[3 * (size_t)l];
    for (uint64_t z = 0; z < b[2]; ++z)
      for (uint64_t y = 0; y < b[1]; ++y)
        for (uint64_t x = 0; x < b[0]; ++x) {
          const uint64_t i = L.id(x, y, z, l);
          if (!empty[i]) {
            const bool res = !resident || resident[i] || l == L.nlod - 1;   // the coarsest level is always resident
            L.meta[i] = res ? (uint32_t)(BI_FLAG_COUNT + i) : (uint32_t)BI_MISSING;
            continue;
          }
          uint32_t m = BI_CHILD_EMPTY;
          if (l > 0) {
            const uint64_t* c = &L.ilayout[3 * (size_t)(l - 1)];
            // children outside the finer grid count as child empty
            for (int k = 0; k < 8 && m == BI_CHILD_EMPTY; ++k) {
              const uint64_t cx = 2 * x + (k & 1), cy = 2 * y + ((k >> 1) & 1), cz = 2 * z + (k >> 2);
              if (cx >= c[0] || cy >= c[1] || cz >= c[2]) continue;
              if (L.meta[L.id(cx, cy, cz, l - 1)] != BI_CHILD_EMPTY) m = BI_EMPTY;
            }
          }
          L.meta[i] = m;
        }
  }
  // eye -> normalised volume coordinates (GLGridLeaper.cpp:560-574)
  V3 dom((float)t.dims[0], (float)t.dims[1], (float)t.dims[2]);
  V3 sc(ds->scale[0], ds->scale[1], ds->scale[2]);
  V3 ext = dom * sc;
  ext = ext / ext.maxVal();
  const M4 mv = mul(mul(from_arr(v->rot), from_arr(v->trans)), from_arr(v->view));
  L.emm = mul(mul(mul(inverse(mv), translation(-0.0f, -0.0f, -0.0f)), scaling(1.0f / ext.x, 1.0f / ext.y, 1.0f / ext.z)),
              translation(0.5f, 0.5f, 0.5f));
  // isosurface: mModelToEye and mModelViewIT (GLGridLeaper.cpp:720-723)
  L.mte = inverse(L.emm);
  L.mvi = inverse(mv);
  L.proj = from_arr(v->proj);
  L.inv_proj = inverse(L.proj);
  L.W = (int)v->width;
  L.H = (int)v->height;
  // CullingLOD::GetLoDFactor and GLVolumePool::Enable
  L.lod_factor = 2.0f * std::tan(v->fov * ((3.1416f / 180.0f) / 2.0f)) * 1.0f / float(v->height);
  L.l0err = (ext / dom).maxVal();
  L.srm = p->srm;
  V3 vscale = sc / sc.minVal();
  L.dscale = V3(1, 1, 1) / vscale;
  L.amb = V3(p->light_ambient[0], p->light_ambient[1], p->light_ambient[2]);
  L.dif = V3(p->light_diffuse[0], p->light_diffuse[1], p->light_diffuse[2]);
  L.spc = V3(p->light_specular[0], p->light_specular[1], p->light_specular[2]);
  const V3 ld(p->light_dir[0], p->light_dir[1], p->light_dir[2]);
  const V3 ml = xform_dir(ld, L.emm);
  const float mll = length(ml);
  L.mlight = V3(ml.x / mll, ml.y / mll, ml.z / mll);   // FLOATVECTOR3::normalized
  L.tf_scale = p->tf_scale;
  L.grad_scale = p->grad_scale;
  L.iso_n = (float)((double)p->isovalue / (double)(1 << p->bit_width));   // GetNormalizedIsovalue
  L.iso_col = V3(p->iso_color[0], p->iso_color[1], p->iso_color[2]);
  L.ldir = V3(p->light_dir[0], p->light_dir[1], p->light_dir[2]);
  return true;
}

struct VoxTex {   // a stored brick, sampled in its voxel coordinates with GL_LINEAR / GL_CLAMP
  const uint8_t* data;
  int n[3];
  int bpv;
  float inv_max;
  float texel(int x, int y, int z) const {
    if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) return 0.0f;
    size_t i = (size_t)x + (size_t)n[0] * ((size_t)y + (size_t)n[1] * (size_t)z);
    unsigned v = bpv == 1 ? data[i] : ((const uint16_t*)data)[i];
    return (float)v * inv_max;
  }
  float sample(const V3& b) const {
    int i0[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
      const float s = std::min(std::max(b[a], 0.0f), (float)n[a]);
      const float u = s - 0.5f, fl = std::floor(u);
      i0[a] = (int)fl;
      f[a] = u - fl;
    }
    float c000 = texel(i0[0], i0[1], i0[2]), c100 = texel(i0[0] + 1, i0[1], i0[2]);
    float c010 = texel(i0[0], i0[1] + 1, i0[2]), c110 = texel(i0[0] + 1, i0[1] + 1, i0[2]);
    float c001 = texel(i0[0], i0[1], i0[2] + 1), c101 = texel(i0[0] + 1, i0[1], i0[2] + 1);
    float c011 = texel(i0[0], i0[1] + 1, i0[2] + 1), c111 = texel(i0[0] + 1, i0[1] + 1, i0[2] + 1);
    float c00 = c000 * (1.0f - f[0]) + c100 * f[0];
    float c10 = c010 * (1.0f - f[0]) + c110 * f[0];
    float c01 = c001 * (1.0f - f[0]) + c101 * f[0];
    float c11 = c011 * (1.0f - f[0]) + c111 * f[0];
    float c0 = c00 * (1.0f - f[1]) + c10 * f[1];
    float c1 = c01 * (1.0f - f[1]) + c11 * f[1];
    return c0 * (1.0f - f[2]) + c1 * f[2];
  }
  // GLGridLeaper-GradientTools.glsl ComputeGradient, one voxel per tap
  V3 gradient(const V3& c) const {
    float xp = sample(c + V3(+1.0f, 0, 0)), xm = sample(c + V3(-1.0f, 0, 0));
    float yp = sample(c + V3(0, -1.0f, 0)), ym = sample(c + V3(0, +1.0f, 0));
    float zp = sample(c + V3(0, 0, +1.0f)), zm = sample(c + V3(0, 0, -1.0f));
    return V3(xm - xp, yp - ym, zm - zp) / 2.0f;
  }
};

inline V3 reflect(const V3& I, const V3& N) { return I - N * (2.0f * dot(N, I)); }
inline float pow8(float x) { float x2 = x * x; float x4 = x2 * x2; return x4 * x4; }
inline float clamp01(float v) { return std::min(std::max(v, 0.0f), 1.0f); }
inline V3 normalize_rcp(const V3& a) { const float r = 1.0f / length(a); return a * r; }

// lighting.glsl:33-42 with the model-space eye
V3 lighting(const V3& eye, const V3& pos, const V3& n, const V3& amb, const V3& diff, const V3& spec, const V3& L) {
  V3 view = normalize_rcp(eye - pos);
  V3 R = normalize_rcp(reflect(view, n));
  float d = std::max(std::fabs(dot(n, L)), 0.0f);
  float s = pow8(std::max(dot(R, L), 0.0f));
  V3 c = amb + diff * d + spec * s;
  return V3(clamp01(c.x), clamp01(c.y), clamp01(c.z));
}

// Compose-FS.glsl Lighting(): eye-space position and normal, eye at the origin
V3 compose_lighting(const V3& pos, V3 n, const V3& amb, const V3& diff, const V3& spec, const V3& L) {
  n.z = std::fabs(n.z);
  V3 view = normalize_rcp(V3(0, 0, 0) - pos);
  V3 R = normalize_rcp(reflect(view, n));
  float d = std::max(std::fabs(dot(n, L * -1.0f)), 0.0f);
  float s = pow8(std::max(dot(R, L), 0.0f));
  V3 c = amb + diff * d + spec * s;
  return V3(clamp01(c.x), clamp01(c.y), clamp01(c.z));
}

void tf1(const orc_params* p, float c, float o[4]) {
  float s = std::min(std::max(c, 0.0f), 1.0f);
  int i = std::min((int)std::floor(s * (float)p->tf1d_n), p->tf1d_n - 1);
  for (int k = 0; k < 4; ++k) o[k] = (float)p->tf1d[4 * i + k] / 255.0f;
}
void tf2(const orc_params* p, float cx, float cy, float o[4]) {
  float sx = std::min(std::max(cx, 0.0f), 1.0f), sy = std::min(std::max(cy, 0.0f), 1.0f);
  int x = std::min((int)std::floor(sx * (float)p->tf2d_w), p->tf2d_w - 1);
  int y = std::min((int)std::floor(sy * (float)p->tf2d_h), p->tf2d_h - 1);
  const uint8_t* e = p->tf2d + 4 * ((size_t)y * (size_t)p->tf2d_w + (size_t)x);
  for (int k = 0; k < 4; ++k) o[k] = (float)e[k] / 255.0f;
}

// ComputeColorFromVolume of the Method shaders; b: voxel position in the brick,
// m: normalised (model) position of the sample
void color_at(const Leaper& L, const VoxTex& tx, const V3& b, const V3& m, float c[4]) {
  const orc_params* p = L.p;
  const float data = tx.sample(b);
  if (p->mode == 0) {
    tf1(p, data * L.tf_scale, c);
    if (p->lighting) {
      V3 g = tx.gradient(b) * L.dscale;   // ComputeNormal: safe normalisation by division
      const float l = length(g);
      if (l > 0.0f) g = V3(g.x / l, g.y / l, g.z / l);
      V3 lc = lighting(L.meye, m, g, L.amb, V3(c[0], c[1], c[2]) * L.dif, L.spc, L.mlight);
      c[0] = lc.x; c[1] = lc.y; c[2] = lc.z;
    }
  } else {
    const V3 g = tx.gradient(b);
    const float gm = length(g);
    tf2(p, data * L.tf_scale, 1.0f - gm * L.grad_scale, c);
    if (p->lighting) {
      const V3 nrm = L.dscale * (gm > 0.0f ? g / gm : g);
      V3 lc = lighting(L.meye, m, nrm, L.amb, V3(c[0], c[1], c[2]) * L.dif, L.spc, L.mlight);
      c[0] = lc.x; c[1] = lc.y; c[2] = lc.z;
    }
  }
}

uint32_t compute_lod(const Leaper& L, float depth) {
  // GLVolumePool.cpp:653-656: min(iMaxLOD, uint(log2(x))).  uint(log2(x)) is
  // taken as the exact floor of the logarithm, read from the binary exponent
  // (a float log2 may round across a power of two); x < 1 gives level 0.
  const float x = L.lod_factor * (-depth) / L.l0err;
  uint32_t u;
  std::memcpy(&u, &x, 4);
  const int e = (int)((u >> 23) & 0xffu) - 127;
  const uint32_t lod = (x >= 1.0f && e > 0) ? (uint32_t)e : 0u;
  return std::min((uint32_t)(L.nlod - 1), lod);
}

struct Ray {
  V3 pos;          // normalised start
  float depth;     // eye-space z of the start
  float col[4];    // start colour
  bool live;
};

// One pass of GLGridLeaper-blend.glsl for one pixel.  out: accRayColor,
// resume: rayResumePos (xyz, w) + rayResumeColor, missing: bricks reported.
void pass_pixel(const Leaper& L, int px, int py, const Ray& r, float out[4], float resume[8],
                uint8_t* missing, uint64_t& ns) {
  const Octree& t = *L.t;
  const orc_params* p = L.p;
  float acc[4] = {r.col[0], r.col[1], r.col[2], r.col[3]};
  for (int k = 0; k < 4; ++k) { resume[4 + k] = acc[k]; out[k] = p->mode == 2 ? 0.0f : acc[k]; }
  resume[0] = r.pos.x; resume[1] = r.pos.y; resume[2] = r.pos.z; resume[3] = r.depth;
  if (!r.live) { resume[3] = 1000.0f; return; }
  // exit point and depth from the box's back face along the pixel ray
  float ndc[4] = {((float)px + 0.5f) * 2.0f / (float)L.W - 1.0f, ((float)py + 0.5f) * 2.0f / (float)L.H - 1.0f,
                  -1.0f, 1.0f};
  float e4[4];
  vmul4(ndc, L.inv_proj, e4);
  const V3 d(e4[0] / e4[3], e4[1] / e4[3], e4[2] / e4[3]);
  const V3 o_n = xform_point(V3(0, 0, 0), L.emm), d_n = xform_dir(d, L.emm);
  float tn = -INFINITY, tfar = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (d_n[a] == 0.0f) continue;
    const float inv = 1.0f / d_n[a];
    const float t0 = (0.0f - o_n[a]) * inv, t1 = (1.0f - o_n[a]) * inv;
    tn = std::max(tn, std::min(t0, t1));
    tfar = std::min(tfar, std::max(t0, t1));
  }
  const V3 pe = d * tfar;
  const V3 exitN = xform_point(pe, L.emm);
  const float exitDepth = pe.z;
  const V3 entryN = r.pos;
  const float entryDepth = r.depth;
  const V3 direction = exitN - entryN;
  const float rayLength = length(direction);
  // TransformToPoolSpace: normalize(direction * volumeSize) / (2 srm), in voxels
  const V3 vsz((float)t.dims[0], (float)t.dims[1], (float)t.dims[2]);
  const V3 dv = normalize_rcp(direction * vsz);
  const V3 step = dv / (2.0f * L.srm);
  const float stepSize = length(step);
  const float voxelSize = 0.125f / 2000.0f;
  const float oc = 1.0f / L.srm;
  float tt = 0.0f;
  bool optimal = true;
  V3 cur = entryN;
  uint64_t last[4] = {0, 0, 0, 9999};
  auto terminate = [&](bool opt) {
    if (opt) {
      resume[3] = 1000.0f;
      for (int k = 0; k < 4; ++k) resume[4 + k] = acc[k];
    }
  };
  if (rayLength > voxelSize) {
    for (int j = 0; j < 100; ++j) {
      const float depth = entryDepth * (1.0f - tt) + exitDepth * tt;   // mix
      uint32_t lod = compute_lod(L, depth);
      // GetBrick
      V3 ne(std::min(std::max(cur.x, 0.0f), 1.0f), std::min(std::max(cur.y, 0.0f), 1.0f),
            std::min(std::max(cur.z, 0.0f), 1.0f));
      bool found_req = true;
      uint64_t bc[4];
      auto coords = [&](uint32_t l) {
        bc[0] = (uint64_t)(uint32_t)(ne.x * L.layout[l].x);
        bc[1] = (uint64_t)(uint32_t)(ne.y * L.layout[l].y);
        bc[2] = (uint64_t)(uint32_t)(ne.z * L.layout[l].z);
        bc[3] = l;
      };
      coords(lod);
      uint32_t info = L.meta[L.id(bc[0], bc[1], bc[2], (int)lod)];
      if (info == BI_MISSING) {
        __atomic_store_n(&missing[L.id(bc[0], bc[1], bc[2], (int)lod)], (uint8_t)1, __ATOMIC_RELAXED);
        found_req = false;
        do {
          ++lod;
          coords(lod);
          info = L.meta[L.id(bc[0], bc[1], bc[2], (int)lod)];
          // RendererState::BS_RequestAll, the default strategy (MasterController.h:70-74,
          // GLVolumePool.cpp:570-586): the coarser missing bricks are requested too
          if (info == BI_MISSING)
            __atomic_store_n(&missing[L.id(bc[0], bc[1], bc[2], (int)lod)], (uint8_t)1, __ATOMIC_RELAXED);
        } while (info == BI_MISSING);
      }
      const bool bEmpty = info <= BI_EMPTY;
      if (bEmpty) {
        for (uint32_t lr = lod + 1; lr < (uint32_t)(L.nlod - 1); ++lr) {
          uint64_t save[4] = {bc[0], bc[1], bc[2], bc[3]};
          coords(lr);
          const uint32_t li = L.meta[L.id(bc[0], bc[1], bc[2], (int)lr)];
          if (li == BI_CHILD_EMPTY) {
            info = li;
            lod = lr;
          } else {
            for (int k = 0; k < 4; ++k) bc[k] = save[k];
            break;
          }
        }
      }
      const V3 c0(float(bc[0]) / L.layout[lod].x, float(bc[1]) / L.layout[lod].y, float(bc[2]) / L.layout[lod].z);
      const V3 c1(float(bc[0] + 1) / L.layout[lod].x, float(bc[1] + 1) / L.layout[lod].y,
                  float(bc[2] + 1) / L.layout[lod].z);
      // BrickExit
      V3 bexit;
      {
        const V3 div(1.0f / direction.x, 1.0f / direction.y, 1.0f / direction.z);
        float ti[3];
        for (int a = 0; a < 3; ++a) ti[a] = ((div[a] >= 0.0f ? c1[a] : c0[a]) - ne[a]) * div[a];
        bexit = ne + direction * std::fmin(std::fmin(ti[0], ti[1]), ti[2]);
      }
      if (!found_req && optimal) {
        optimal = false;
        resume[0] = cur.x; resume[1] = cur.y; resume[2] = cur.z; resume[3] = depth;
        for (int k = 0; k < 4; ++k) resume[4 + k] = acc[k];
      }
      if (!bEmpty && (last[0] != bc[0] || last[1] != bc[1] || last[2] != bc[2] || last[3] != bc[3])) {
        const uint64_t bi = info - BI_FLAG_COUNT;
        const TocEntry& te = t.toc[bi];
        VoxTex tx;
        tx.data = t.data.data() + te.offset;
        tx.n[0] = (int)te.size[0]; tx.n[1] = (int)te.size[1]; tx.n[2] = (int)te.size[2];
        tx.bpv = (int)t.bpv;
        tx.inv_max = t.bpv == 1 ? 1.0f / 255.0f : 1.0f / 65535.0f;
        // NormCoordsToPoolCoords in brick voxels: scale = inner / corner span, origin at the core start
        const V3 sc(L.inner.x / (c1.x - c0.x), L.inner.y / (c1.y - c0.y), L.inner.z / (c1.z - c0.z));
        const float ov = (float)t.overlap;
        const V3 tr = V3(ov, ov, ov) - c0 * sc;
        const V3 pin = ne * sc + tr, pout = bexit * sc + tr;
        int steps = (int)std::ceil(length(pout - pin) / stepSize);
        steps = std::min(steps, (int)std::ceil(length((exitN - cur) * sc) / stepSize));
        V3 b = pin;
        for (int i = 0; i < steps; ++i) {
          ++ns;
          if (p->mode == 2) {
            // GLGridLeaper-iso.glsl: first sample at or above the isovalue, then
            // RefineIsosurface (GLGridLeaper-Method-iso.glsl) in brick voxels
            if (tx.sample(b) >= L.iso_n) {
              V3 rd = step / 2.0f;
              V3 bb = b - rd;
              for (int k = 0; k < p->refine_steps; ++k) {
                rd = rd / 2.0f;
                if (tx.sample(bb) >= L.iso_n) bb = bb - rd; else bb = bb + rd;
              }
              // rayHitPos = mModelToEye * hit; rayHitNormal = mModelViewIT * ComputeNormal
              const V3 he = xform_point((bb - tr) / sc, L.mte);
              V3 g = tx.gradient(bb) * L.dscale;
              const float l = length(g);
              if (l > 0.0f) g = V3(g.x / l, g.y / l, g.z / l);
              const float* MI = L.mvi.a;
              const V3 ne(g.x * MI[0] + g.y * MI[1] + g.z * MI[2], g.x * MI[4] + g.y * MI[5] + g.z * MI[6],
                          g.x * MI[8] + g.y * MI[9] + g.z * MI[10]);
              const V3 c = compose_lighting(he, ne, L.amb, L.dif * L.iso_col, L.spc, L.ldir);   // Compose-FS.glsl
              out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = 1.0f;
              if (optimal) resume[3] = 501.0f;   // TerminateRay: 499 + rayHitPos.a (1 + white)
              return;
            }
            b = b + step;
            continue;
          }
          float c[4];
          color_at(L, tx, b, (b - tr) / sc, c);
          c[3] = 1.0f - opacity_pow(1.0f - c[3], oc);   // OpacityCorrectColor
          // UnderCompositing (Compositing.glsl:33-38)
          acc[0] += c[0] * (1.0f - acc[3]) * c[3];
          acc[1] += c[1] * (1.0f - acc[3]) * c[3];
          acc[2] += c[2] * (1.0f - acc[3]) * c[3];
          acc[3] += c[3] * (1.0f - acc[3]);
          if (acc[3] > 0.99f) {
            for (int k = 0; k < 4; ++k) out[k] = acc[k];
            terminate(optimal);
            return;
          }
          b = b + step;
        }
        cur = (b - tr) / sc;
      } else {
        cur = bexit + direction * voxelSize / rayLength;
      }
      for (int k = 0; k < 4; ++k) last[k] = bc[k];
      tt = length(entryN - bexit) / rayLength;
      if (tt > 0.9999f) break;
    }
  }
  for (int k = 0; k < 4; ++k) out[k] = p->mode == 2 ? 0.0f : acc[k];
  terminate(optimal);
}

}  // namespace

extern "C" int orc_leap_pass(const orc_ds* ds, const orc_view* v, const orc_params* p, const float eye[3],
                             const uint8_t* resident, int first, const float* resume_in, float* out, float* resume_out,
                             uint8_t* missing, uint64_t* samples) {
  Leaper L;
  if (!setup(L, ds, v, p, resident)) return -1;
  L.meye = xform_point(V3(eye[0], eye[1], eye[2]), L.emm);
  std::memset(missing, 0, ds->tree.toc.size());
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
  for (int py = 0; py < L.H; ++py) {
    for (int px = 0; px < L.W; ++px) {
      const size_t k = (size_t)py * L.W + px;
      Ray r;
      for (int c = 0; c < 4; ++c) r.col[c] = 0.0f;
      if (first) {
        float ndc[4] = {((float)px + 0.5f) * 2.0f / (float)L.W - 1.0f, ((float)py + 0.5f) * 2.0f / (float)L.H - 1.0f,
                        -1.0f, 1.0f};
        float e4[4];
        vmul4(ndc, L.inv_proj, e4);
        const V3 d(e4[0] / e4[3], e4[1] / e4[3], e4[2] / e4[3]);
        const V3 o_n = xform_point(V3(0, 0, 0), L.emm), d_n = xform_dir(d, L.emm);
        float tn = -INFINITY, tfar = INFINITY;
        bool hit = true;
        for (int a = 0; a < 3; ++a) {
          if (d_n[a] == 0.0f) {
            if (o_n[a] < 0.0f || o_n[a] > 1.0f) hit = false;
            continue;
          }
          const float inv = 1.0f / d_n[a];
          const float t0 = (0.0f - o_n[a]) * inv, t1 = (1.0f - o_n[a]) * inv;
          tn = std::max(tn, std::min(t0, t1));
          tfar = std::min(tfar, std::max(t0, t1));
        }
        const float te = std::max(tn, 1.0f);
        r.live = hit && tfar > tn && tfar > te;
        const V3 pe = d * te;
        r.pos = xform_point(pe, L.emm);
        r.depth = pe.z;
      } else {
        const float* rs = resume_in + 8 * k;
        r.pos = V3(rs[0], rs[1], rs[2]);
        r.depth = rs[3];
        r.live = rs[3] != 1000.0f;
        for (int c = 0; c < 4; ++c) r.col[c] = rs[4 + c];
      }
      uint64_t ns = 0;
      pass_pixel(L, px, py, r, out + 4 * k, resume_out + 8 * k, missing, ns);
      total += ns;
    }
  }
  if (samples) *samples = total;
  return 0;
}
