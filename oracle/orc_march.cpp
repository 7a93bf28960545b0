// Oracle: GLRaycaster march restated on the CPU, plus an FP64 integrator.
// Test infrastructure only (see orc_common.hpp).
//
// GL-faithful FP32 mode.  Each visible, non-empty brick of the sorted list is
// processed in list order, exactly as GLRenderer::Render3DView
// (GLRenderer.cpp:2663-2748) drives GLRaycaster::Render3DInLoop
// (GLRaycaster.cpp:383-478):
//   * ray entry = eye-space position of the brick's front face, stored in an
//     RGBA16F target (GLRaycaster.cpp:97) -> rounded to binary16 here;
//   * ray exit = eye-space position of the back face (fp32 varying);
//   * eye->texture matrix of GLRaycaster::ComputeEyeToTextureMatrix
//     (GLRaycaster.cpp:589-612);
//   * the march of GLRaycaster-{1D,1D-light,2D,2D-light,ISO}-FS.glsl with
//     iStepCount = int(L/step)+1, fixed increment, per-brick 'a >= 0.99' exit;
//   * VRender1D opacity correction a' = 1-(1-a)^fStepScale (VRender1D.glsl:47-57);
//   * UnderCompositing (Compositing.glsl:33-38);
//   * blending into the RGBA32F target with (ONE_MINUS_DST_ALPHA, ONE)
//     (GLRenderer.cpp:151-153): dst += src * (1 - dst.a);
//   * texture fetches: unorm L8/L16 (v/255, v/65535) with GL_LINEAR and GL_CLAMP
//     (border colour 0; GPUMemManDataStructs.cpp:470-485), TF tables RGBA8 with
//     GL_NEAREST + CLAMP_TO_EDGE (GLTexture1D.h:48-51, GLTexture2D.h:50-53);
//   * gradients (Volume3D.glsl:43-60), lighting (lighting.glsl:33-50),
//     iso refinement (RefineIsosurface.glsl:37-52), iso hit depth test and
//     compose (GLRaycaster-ISO-FS.glsl:56-105, Compose-FS.glsl).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#include "orc_internal.hpp"
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {

M4 from_arr(const float* a) { M4 m; for (int i = 0; i < 16; ++i) m.a[i] = a[i]; return m; }

struct BrickTex {
  const uint8_t* data;
  int n[3];
  int bpv;
  float inv_max;
  bool clamp_edge;

  float texel(int x, int y, int z) const {
    if (clamp_edge) {
      x = std::min(std::max(x, 0), n[0] - 1);
      y = std::min(std::max(y, 0), n[1] - 1);
      z = std::min(std::max(z, 0), n[2] - 1);
    } else if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) {
      return 0.0f;  // GL_CLAMP border colour
    }
    size_t i = (size_t)x + (size_t)n[0] * ((size_t)y + (size_t)n[1] * (size_t)z);
    unsigned v = bpv == 1 ? data[i] : ((const uint16_t*)data)[i];
    return (float)v;   // stored integer; sample() applies the unorm scale once
  }
  // texture3D(texVolume, p).x with GL_LINEAR
  float sample(const V3& p) const {
    float u[3];
    int i0[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
      float s = std::min(std::max(p[a], 0.0f), 1.0f);
      u[a] = s * (float)n[a] - 0.5f;
      float fl = std::floor(u[a]);
      i0[a] = (int)fl;
      f[a] = u[a] - fl;
    }
    float c000 = texel(i0[0], i0[1], i0[2]), c100 = texel(i0[0] + 1, i0[1], i0[2]);
    float c010 = texel(i0[0], i0[1] + 1, i0[2]), c110 = texel(i0[0] + 1, i0[1] + 1, i0[2]);
    float c001 = texel(i0[0], i0[1], i0[2] + 1), c101 = texel(i0[0] + 1, i0[1], i0[2] + 1);
    float c011 = texel(i0[0], i0[1] + 1, i0[2] + 1), c111 = texel(i0[0] + 1, i0[1] + 1, i0[2] + 1);
    // GL leaves the filter arithmetic to the implementation: x, y, z lerps,
    // each a + f (b - a) as one fused multiply-add, of the stored integers,
    // scaled to unorm once (the device sampler's form)
    auto lerp = [](float a, float b, float t) { return std::fma(t, b - a, a); };
    float c00 = lerp(c000, c100, f[0]), c10 = lerp(c010, c110, f[0]);
    float c01 = lerp(c001, c101, f[0]), c11 = lerp(c011, c111, f[0]);
    float c0 = lerp(c00, c10, f[1]), c1 = lerp(c01, c11, f[1]);
    return lerp(c0, c1, f[2]) * inv_max;
  }
  // trilinear filter of the footprint with lower corner i0 and weights f
  // (sample()'s lerps and scale)
  float filter(const int i0[3], const float f[3]) const {
    float c000 = texel(i0[0], i0[1], i0[2]), c100 = texel(i0[0] + 1, i0[1], i0[2]);
    float c010 = texel(i0[0], i0[1] + 1, i0[2]), c110 = texel(i0[0] + 1, i0[1] + 1, i0[2]);
    float c001 = texel(i0[0], i0[1], i0[2] + 1), c101 = texel(i0[0] + 1, i0[1], i0[2] + 1);
    float c011 = texel(i0[0], i0[1] + 1, i0[2] + 1), c111 = texel(i0[0] + 1, i0[1] + 1, i0[2] + 1);
    auto lerp = [](float a, float b, float t) { return std::fma(t, b - a, a); };
    float c00 = lerp(c000, c100, f[0]), c10 = lerp(c010, c110, f[0]);
    float c01 = lerp(c001, c101, f[0]), c11 = lerp(c011, c111, f[0]);
    float c0 = lerp(c00, c10, f[1]), c1 = lerp(c01, c11, f[1]);
    return lerp(c0, c1, f[2]) * inv_max;
  }
  // Volume3D.glsl:43-53: six taps at c +- h, h = vVoxelStepsize = one texel
  // (1 / texture size, GLRaycaster.cpp:242-251).  GL_LINEAR at s + 1/n is the
  // footprint of s moved by one texel with the same weights ((s + 1/n) n - 1/2
  // = u + 1), so on an axis where both taps stay inside [0, 1] (no coordinate
  // clamp) the taps are taken in that form: the centre's lower corner +-1 on
  // that axis and the centre's weights.  On an axis where a tap leaves [0, 1]
  // the clamped coordinate moves the footprint differently, and both of that
  // axis's taps are sample(c +- h) as the shader states them.
  V3 gradient(const V3& c, const V3& h) const {
    int i0[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
      float s = std::min(std::max(c[a], 0.0f), 1.0f);
      float u = s * (float)n[a] - 0.5f, fl = std::floor(u);
      i0[a] = (int)fl;
      f[a] = u - fl;
    }
    float plus[3], minus[3];   // taps at c + h e_a and c - h e_a
    for (int a = 0; a < 3; ++a) {
      if (c[a] - h[a] >= 0.0f && c[a] + h[a] <= 1.0f) {
        int ip[3] = {i0[0], i0[1], i0[2]}, im[3] = {i0[0], i0[1], i0[2]};
        ip[a] += 1;
        im[a] -= 1;
        plus[a] = filter(ip, f);
        minus[a] = filter(im, f);
      } else {
        V3 e(0, 0, 0);
        e[a] = h[a];
        plus[a] = sample(c + e);
        minus[a] = sample(c - e);
      }
    }
    // the shader's names: xp = +x, xm = -x, yp = -y, ym = +y, zp = +z, zm = -z
    const float xp = plus[0], xm = minus[0], yp = minus[1], ym = plus[1], zp = plus[2], zm = minus[2];
    return V3(xm - xp, yp - ym, zm - zp) / 2.0f;
  }
};

struct Tf {
  int mode;
  const uint8_t* t1; int n1;
  const uint8_t* t2; int w2, h2;
  void rgba1(float c, float out[4]) const {  // texture1D, NEAREST, CLAMP_TO_EDGE
    float s = std::min(std::max(c, 0.0f), 1.0f);
    int i = std::min((int)std::floor(s * (float)n1), n1 - 1);
    for (int k = 0; k < 4; ++k) out[k] = (float)t1[4 * i + k] / 255.0f;
  }
  void rgba2(float cx, float cy, float out[4]) const {  // texture2D, NEAREST, CLAMP_TO_EDGE
    float sx = std::min(std::max(cx, 0.0f), 1.0f), sy = std::min(std::max(cy, 0.0f), 1.0f);
    int x = std::min((int)std::floor(sx * (float)w2), w2 - 1);
    int y = std::min((int)std::floor(sy * (float)h2), h2 - 1);
    const uint8_t* p = t2 + 4 * ((size_t)y * (size_t)w2 + (size_t)x);
    for (int k = 0; k < 4; ++k) out[k] = (float)p[k] / 255.0f;
  }
};

inline V3 reflect(const V3& I, const V3& N) { return I - N * (2.0f * dot(N, I)); }
// pow(x, 8.0) of lighting.glsl evaluated by repeated squaring
inline float pow8(float x) { float x2 = x * x; float x4 = x2 * x2; return x4 * x4; }
// 1 - pow(1 - a, s)
inline float opacity_correct(float a, float s) { return 1.0f - opacity_pow(1.0f - a, s); }
inline float clamp01(float v) { return std::min(std::max(v, 0.0f), 1.0f); }

// GLSL normalize(): one reciprocal of the length, then three products (the GLSL
// built-in is implementation defined; this is the form the HIP kernel uses)
inline V3 normalize_rcp(const V3& a) { const float r = 1.0f / length(a); return a * r; }

// lighting.glsl:33-50 with vEyePos = 0
V3 lighting(const V3& pos, const V3& n, const V3& amb, const V3& diff, const V3& spec, const V3& L) {
  V3 view = normalize_rcp(V3(0, 0, 0) - pos);
  V3 R = normalize_rcp(reflect(view, n));
  float d = std::max(std::fabs(dot(n, L)), 0.0f);
  float s = pow8(std::max(dot(R, L), 0.0f));
  V3 c = amb + diff * d + spec * s;
  return V3(clamp01(c.x), clamp01(c.y), clamp01(c.z));
}

// Compose-FS.glsl Lighting()
V3 iso_lighting(const V3& pos, V3 n, const V3& amb, const V3& diff, const V3& spec, const V3& L) {
  n.z = std::fabs(n.z);
  V3 view = normalize_rcp(V3(0, 0, 0) - pos);
  V3 R = normalize_rcp(reflect(view, n));
  float d = std::max(std::fabs(dot(n, L * -1.0f)), 0.0f);
  float s = pow8(std::max(dot(R, L), 0.0f));
  V3 c = amb + diff * d + spec * s;
  return V3(clamp01(c.x), clamp01(c.y), clamp01(c.z));
}

// (unsigned char)((255.0 * v - fMin) / (fMax - fMin)), GLVolumeListElem::CreateTexture
// (GPUMemManDataStructs.cpp:410-425).  C++ leaves the conversion of an
// out-of-range double undefined; x86 compilers truncate to a 32-bit integer
// (0x80000000 when it does not fit or is NaN) and keep the low byte.
inline uint8_t quantize8(unsigned v, double fmin, double fmax) {
  const double q = (255.0 * (double)v - fmin) / (fmax - fmin);
  if (!(q > -2147483649.0 && q < 2147483648.0)) return 0;
  return (uint8_t)(uint32_t)(int32_t)q;
}

// The texture GLVolumeListElem::CreateTexture makes of one brick (reference
// layout, x fastest): 16-bit values quantized to 8 bits (down8), then padded to
// the next power of two per axis (pow2) by PadData (GPUMemManDataStructs.cpp:
// 306-360): rows copied, the element after each row, the row after each slice
// and the slice after the last repeated unless the border is disabled, the rest
// zero.  n_out: the texture size.
std::vector<uint8_t> compat_texture(const uint8_t* src, const uint32_t size[3], int bpv, bool down8, bool pow2,
                                    bool disable_border, double fmin, double fmax, int n_out[3]) {
  const size_t nv = (size_t)size[0] * size[1] * size[2];
  std::vector<uint8_t> data;
  int ob = bpv;
  if (down8) {
    data.resize(nv);
    const uint16_t* s16 = reinterpret_cast<const uint16_t*>(src);
    for (size_t i = 0; i < nv; ++i) data[i] = quantize8(s16[i], fmin, fmax);
    ob = 1;
  } else {
    data.assign(src, src + nv * (size_t)bpv);
  }
  for (int k = 0; k < 3; ++k) n_out[k] = (int)size[k];
  if (!pow2) return data;
  const size_t P[3] = {next_pow2(size[0]), next_pow2(size[1]), next_pow2(size[2])};
  const size_t el = (size_t)ob, row_src = size[0] * el, row_dst = P[0] * el;
  std::vector<uint8_t> out(row_dst * P[1] * P[2], 0);
  size_t t = 0, s = 0;
  for (size_t z = 0; z < size[2]; ++z) {
    for (size_t y = 0; y < size[1]; ++y) {
      std::memcpy(&out[t], &data[s], row_src);
      if (!disable_border && row_dst > row_src) std::memcpy(&out[t + row_src], &out[t + row_src - el], el);
      t += row_dst;
      s += row_src;
    }
    if (P[1] > size[1]) {
      if (!disable_border) std::memcpy(&out[t], &out[t - row_dst], row_dst);
      t += (P[1] - size[1]) * row_dst;
    }
  }
  if (!disable_border && P[2] > size[2]) std::memcpy(&out[t], &out[t - P[1] * row_dst], P[1] * row_dst);
  for (int k = 0; k < 3; ++k) n_out[k] = (int)P[k];
  return out;
}

struct Entry {
  BrickTex tex;
  V3 bmin, bmax;      // model-space box of the brick
  M4 eye2tex;         // gl_TextureMatrix[0]
  float ray_step, step_scale;
  V3 vstep;
  bool empty;
};

struct Frame {
  const orc_ds* ds;
  const orc_params* p;
  M4 mv, inv_mv, inv_proj, proj;
  float nm[3][3];     // gl_NormalMatrix (column-vector convention)
  int W, H;
  float znear, zfar;
  std::vector<Entry> bricks;
  Tf tf;
  V3 amb, dif, spc, ldir, dscale, iso_col;
  bool clip_on = false;
  float clip[4] = {0, 0, 0, 0};   // model space, GLRaycaster::RenderBox
  float iso_n = 0.0f;             // AbstrRenderer::GetNormalizedIsovalue
  std::vector<std::vector<uint8_t>> compat_tex;   // the compatibility textures in use

  bool setup(const orc_ds* d, const orc_view* v, const orc_params* prm) {
    ds = d; p = prm;
    clip_on = prm->clip_on != 0;
    if (clip_on) model_clip_plane(prm->clip, from_arr(v->rot), from_arr(v->trans), clip);
    W = (int)v->width; H = (int)v->height;
    znear = v->znear; zfar = v->zfar;
    mv = mul(mul(from_arr(v->rot), from_arr(v->trans)), from_arr(v->view));
    inv_mv = inverse(mv);
    proj = from_arr(v->proj);
    inv_proj = inverse(proj);
    // gl_NormalMatrix = inverse of the upper 3x3 of the row-vector modelview,
    // applied to column vectors.
    {
      M4 m3;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m3.m(r, c) = mv.m(r, c);
      m3.m(3, 0) = m3.m(3, 1) = m3.m(3, 2) = 0; m3.m(0, 3) = m3.m(1, 3) = m3.m(2, 3) = 0; m3.m(3, 3) = 1;
      M4 inv = inverse(m3);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) nm[r][c] = inv.m(r, c);
    }
    tf.mode = prm->mode;
    tf.t1 = prm->tf1d; tf.n1 = prm->tf1d_n;
    tf.t2 = prm->tf2d; tf.w2 = prm->tf2d_w; tf.h2 = prm->tf2d_h;
    amb = V3(prm->light_ambient[0], prm->light_ambient[1], prm->light_ambient[2]);
    dif = V3(prm->light_diffuse[0], prm->light_diffuse[1], prm->light_diffuse[2]);
    spc = V3(prm->light_specular[0], prm->light_specular[1], prm->light_specular[2]);
    ldir = V3(prm->light_dir[0], prm->light_dir[1], prm->light_dir[2]);
    dscale = V3(prm->domain_scale[0], prm->domain_scale[1], prm->domain_scale[2]);
    iso_col = V3(prm->iso_color[0], prm->iso_color[1], prm->iso_color[2]);

    std::vector<orc_brick> list(4096);
    int n = orc_plan(d, v, prm, list.data(), (int)list.size());
    if (n < 0) { list.resize((size_t)(-n)); n = orc_plan(d, v, prm, list.data(), (int)list.size()); }
    list.resize((size_t)n);
    const Octree& t = d->tree;
    // dataset range (UVFDataset::GetRange over the LOD-0 bricks)
    double fmin = std::numeric_limits<double>::max(), fmax = -std::numeric_limits<double>::max();
    for (const auto& te : t.toc)
      if (te.coords[3] == 0) { fmin = std::min(fmin, te.minmax[0]); fmax = std::max(fmax, te.minmax[1]); }
    const bool down8 = prm->down8 != 0 && t.bpv == 2;
    if (down8) {
      // MathTools::lerp(m_fIsovalue, 0.f, float(mx), 0.f, 1.f), AbstrRenderer.cpp:412-424
      const double mx = fmin > fmax ? (double)prm->tf1d_n : fmax;
      const float fmx = (float)mx;
      iso_n = (float)(0.0f + (double)(prm->isovalue - 0.0f) * (1.0 / (double)(fmx - 0.0f)));
    } else {
      iso_n = (float)((double)prm->isovalue / (double)(1 << prm->bit_width));
    }
    compat_tex.reserve(list.size());
    for (const auto& b : list) {
      Entry e;
      e.empty = b.is_empty != 0;
      const TocEntry& te = t.toc[b.toc_index];
      e.tex.data = t.data.data() + te.offset;
      e.tex.n[0] = (int)te.size[0]; e.tex.n[1] = (int)te.size[1]; e.tex.n[2] = (int)te.size[2];
      e.tex.bpv = (int)t.bpv;
      if ((prm->pow2 || down8) && !e.empty) {
        const uint32_t sz[3] = {(uint32_t)te.size[0], (uint32_t)te.size[1], (uint32_t)te.size[2]};
        compat_tex.push_back(compat_texture(e.tex.data, sz, (int)t.bpv, down8, prm->pow2 != 0,
                                            prm->clamp_to_edge != 0, fmin, fmax, e.tex.n));
        e.tex.data = compat_tex.back().data();
        if (down8) e.tex.bpv = 1;
      }
      e.tex.inv_max = e.tex.bpv == 1 ? 1.0f / 255.0f : 1.0f / 65535.0f;
      e.tex.clamp_edge = prm->clamp_to_edge != 0;
      V3 c(b.center[0], b.center[1], b.center[2]), x(b.extension[0], b.extension[1], b.extension[2]);
      // GLRaycaster::RenderBox (GLRaycaster.cpp:302-345)
      e.bmin = c - x / 2.0f;
      e.bmax = c + x / 2.0f;
      V3 tmin(b.tex_min[0], b.tex_min[1], b.tex_min[2]), tmax(b.tex_max[0], b.tex_max[1], b.tex_max[2]);
      // ComputeEyeToTextureMatrix(p1=max point, t1=max coords, p2=min point, t2=min coords)
      V3 sc = (tmin - tmax) / (e.bmin - e.bmax);
      e.eye2tex = mul(mul(mul(inv_mv, translation(-e.bmax.x, -e.bmax.y, -e.bmax.z)),
                          scaling(sc.x, sc.y, sc.z)),
                      translation(tmax.x, tmax.y, tmax.z));
      float vs[3];
      orc_brick_shader_vars(d, prm, &b, &e.ray_step, &e.step_scale, vs);
      e.vstep = V3(vs[0], vs[1], vs[2]);
      bricks.push_back(e);
    }
    return true;
  }

  V3 normal_eye(const V3& g) const {
    V3 s = g * dscale;
    V3 n(nm[0][0] * s.x + nm[0][1] * s.y + nm[0][2] * s.z,
         nm[1][0] * s.x + nm[1][1] * s.y + nm[1][2] * s.z,
         nm[2][0] * s.x + nm[2][1] * s.y + nm[2][2] * s.z);
    float l = length(n);
    if (l > 0.0f) { const float r = 1.0f / l; n = n * r; }
    return n;
  }

  V3 pixel_dir(int px, int py) const {
    float ndc[4] = {((float)px + 0.5f) * 2.0f / (float)W - 1.0f, ((float)py + 0.5f) * 2.0f / (float)H - 1.0f,
                    -1.0f, 1.0f};
    float e[4];
    vmul4(ndc, inv_proj, e);
    return V3(e[0] / e[3], e[1] / e[3], e[2] / e[3]);  // a point on the near plane
  }

  // ray (eye at origin, direction d) vs model-space box; returns t range
  bool intersect(const V3& o_m, const V3& d_m, const Entry& b, float& tn, float& tf) const {
    tn = -INFINITY; tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
      if (d_m[a] == 0.0f) {
        if (o_m[a] < b.bmin[a] || o_m[a] > b.bmax[a]) return false;
        continue;
      }
      float inv = 1.0f / d_m[a];
      float t0 = (b.bmin[a] - o_m[a]) * inv, t1 = (b.bmax[a] - o_m[a]) * inv;
      tn = std::max(tn, std::min(t0, t1));
      tf = std::min(tf, std::max(t0, t1));
    }
    if (clip_on) {
      // the brick box cut by the clip plane and closed by its cap (Clipper::BoxPlane,
      // Clipper.cpp:106-131): the ray keeps the part with n . p + d <= 0
      const float a = clip[0] * d_m.x + clip[1] * d_m.y + clip[2] * d_m.z;
      const float c = clip[0] * o_m.x + clip[1] * o_m.y + clip[2] * o_m.z + clip[3];
      if (a > 0.0f) tf = std::min(tf, -c / a);
      else if (a < 0.0f) tn = std::max(tn, -c / a);
      else if (c > 0.0f) return false;
    }
    return tf > tn && tf > 0.0f;
  }

  void render_pixel(int px, int py, float out[4], uint64_t& ns) const {
    V3 d = pixel_dir(px, py);
    V3 o_m = xform_point(V3(0, 0, 0), inv_mv);
    V3 d_m = xform_dir(d, inv_mv);
    float dst[4] = {0, 0, 0, 0};
    // iso state
    float best_depth = 1.0f;
    bool have_hit = false;
    V3 hit_pos, hit_n;
    float hit_param = 0;
    const float eps = p->ert_eps;
    for (const Entry& b : bricks) {
      if (b.empty) continue;
      if (eps > 0 && p->mode != 2 && (1.0f - dst[3]) < eps) break;
      float t_in, t_out;
      if (!intersect(o_m, d_m, b, t_in, t_out)) continue;
      // near-plane entry when the front face is clipped (GLRaycaster.cpp:351-369)
      float te = std::max(t_in, 1.0f);
      if (t_out <= te) continue;
      V3 entry = d * te, exit = d * t_out;
      entry = V3(round_fp16(entry.x), round_fp16(entry.y), round_fp16(entry.z));
      V3 entryTex = xform_point(entry, b.eye2tex), exitTex = xform_point(exit, b.eye2tex);
      V3 dir = exit - entry;
      float L = length(dir);
      int steps = (int)(L / b.ray_step) + 1;
      V3 incTex = (exitTex - entryTex) / (L / b.ray_step);
      V3 pos = entryTex;
      if (p->mode == 2) {
        // GLRaycaster-ISO-FS.glsl:56-105
        const float iso = iso_n;
        bool found = false;
        for (int i = 0; i < steps; ++i) {
          ++ns;
          if (b.tex.sample(pos) >= iso) { found = true; break; }
          pos = pos + incTex;
        }
        if (!found) continue;
        // RefineIsosurface.glsl:37-52
        V3 rd = incTex / 2.0f;
        pos = pos - rd;
        for (int k = 0; k < p->refine_steps; ++k) {
          rd = rd / 2.0f;
          if (b.tex.sample(pos) >= iso) pos = pos - rd; else pos = pos + rd;
        }
        float param = length(pos - entryTex) / length(exitTex - entryTex);
        V3 he = entry * (1.0f - param) + exit * param;
        float px_ = zfar / (zfar - znear), py_ = (zfar * znear / (znear - zfar));
        float depth = px_ + (py_ / -he.z);
        depth = std::min(std::max(depth, 0.0f), 1.0f);
        if (depth < best_depth) {
          best_depth = depth;
          have_hit = true;
          hit_pos = he; hit_param = param;
          hit_n = normal_eye(b.tex.gradient(pos, b.vstep));
        }
        continue;
      }
      float col[4] = {0, 0, 0, 0};
      V3 cur = entry;
      V3 rayInc = dir / L * b.ray_step;
      for (int i = 0; i < steps; ++i) {
        ++ns;
        float s[4];
        float v = b.tex.sample(pos);
        if (p->mode == 0) {
          tf.rgba1(v * p->tf_scale, s);
          if (p->lighting) {
            V3 g = b.tex.gradient(pos, b.vstep);
            V3 n = normal_eye(g);
            V3 lc = lighting(cur, n, amb, dif * V3(s[0], s[1], s[2]), spc, ldir);
            s[0] = lc.x; s[1] = lc.y; s[2] = lc.z;
          }
          s[3] = opacity_correct(s[3], b.step_scale);
        } else {
          V3 g = b.tex.gradient(pos, b.vstep);
          float gm = length(g);
          tf.rgba2(v * p->tf_scale, 1.0f - gm * p->grad_scale, s);
          if (p->lighting) {
            V3 n = normal_eye(g);
            V3 lc = lighting(cur, n, amb, dif * V3(s[0], s[1], s[2]), spc, ldir);
            s[3] = opacity_correct(s[3], b.step_scale);
            s[0] = clamp01(lc.x); s[1] = clamp01(lc.y); s[2] = clamp01(lc.z); s[3] = clamp01(s[3]);
          } else {
            s[3] = opacity_correct(s[3], b.step_scale);
          }
        }
        // UnderCompositing
        float w = (1.0f - col[3]) * s[3];
        col[0] += s[0] * w; col[1] += s[1] * w; col[2] += s[2] * w;
        col[3] += s[3] * (1.0f - col[3]);
        cur = cur + rayInc;
        if (col[3] >= 0.99f) break;
        if (eps > 0 && (1.0f - dst[3]) * (1.0f - col[3]) < eps) break;
        pos = pos + incTex;
      }
      // blend (ONE_MINUS_DST_ALPHA, ONE)
      float t = 1.0f - dst[3];
      for (int k = 0; k < 4; ++k) dst[k] = col[k] * t + dst[k];
    }
    if (p->mode == 2) {
      if (have_hit && hit_param != 0.0f) {
        V3 c = iso_lighting(hit_pos, hit_n, amb, dif * iso_col, spc, ldir);
        out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = 1.0f;
      } else {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
      }
      return;
    }
    for (int k = 0; k < 4; ++k) out[k] = dst[k];
  }
};

}  // namespace

extern "C" int orc_render(const orc_ds* ds, const orc_view* v, const orc_params* p, const int roi[4],
                          float* out, uint64_t* n_samples, int n_threads) {
  Frame f;
  if (!f.setup(ds, v, p)) return -1;
  int x0 = roi[0], y0 = roi[1], x1 = roi[2], y1 = roi[3];
  int w = x1 - x0;
  uint64_t total = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
  for (int y = y0; y < y1; ++y) {
    uint64_t ns = 0;
    for (int x = x0; x < x1; ++x) f.render_pixel(x, y, out + 4 * ((size_t)(y - y0) * (size_t)w + (size_t)(x - x0)), ns);
    total += ns;
  }
  if (n_samples) *n_samples = total;
  return 0;
}

// ---------------------------------------------------------------------------
// FP64 integrator of the same emission-absorption model.
namespace {

struct Field64 {
  const uint8_t* vol;  // linear LOD0 volume, x fastest
  uint64_t n[3];
  int bpv;
  double inv_max;
  bool clamp_edge;
  double vox(int64_t x, int64_t y, int64_t z) const {
    if (clamp_edge) {
      x = std::min<int64_t>(std::max<int64_t>(x, 0), (int64_t)n[0] - 1);
      y = std::min<int64_t>(std::max<int64_t>(y, 0), (int64_t)n[1] - 1);
      z = std::min<int64_t>(std::max<int64_t>(z, 0), (int64_t)n[2] - 1);
    } else if (x < 0 || y < 0 || z < 0 || x >= (int64_t)n[0] || y >= (int64_t)n[1] || z >= (int64_t)n[2]) {
      return 0.0;
    }
    size_t i = (size_t)x + (size_t)n[0] * ((size_t)y + (size_t)n[1] * (size_t)z);
    unsigned v = bpv == 1 ? vol[i] : ((const uint16_t*)vol)[i];
    return (double)v * inv_max;
  }
  // u,v,w in voxel-centre coordinates (voxel i centred at i)
  double sample(double u, double v, double w) const {
    double fx = std::floor(u), fy = std::floor(v), fz = std::floor(w);
    int64_t x = (int64_t)fx, y = (int64_t)fy, z = (int64_t)fz;
    double a = u - fx, b = v - fy, c = w - fz;
    double r = 0;
    for (int k = 0; k < 8; ++k) {
      int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
      double wgt = (dx ? a : 1 - a) * (dy ? b : 1 - b) * (dz ? c : 1 - c);
      if (wgt != 0) r += wgt * vox(x + dx, y + dy, z + dz);
    }
    return r;
  }
};

}  // namespace

extern "C" int orc_render_fp64(const orc_ds* ds, const orc_view* v, const orc_params* p, const int roi[4],
                               double step_model, float* out, int n_threads) {
  const Octree& t = ds->tree;
  if (p->lod != 0) return -2;
  // reassemble the LOD0 volume from the core voxels of the LOD0 bricks
  std::vector<uint8_t> vol((size_t)(t.dims[0] * t.dims[1] * t.dims[2] * t.bpv));
  const uint64_t* bb = t.lods[0].bricks;
  for (uint64_t i = 0; i < bb[0] * bb[1] * bb[2]; ++i) {
    const TocEntry& e = t.toc[i];
    uint64_t core[3] = {t.brick[0] - 2 * t.overlap, t.brick[1] - 2 * t.overlap, t.brick[2] - 2 * t.overlap};
    for (uint64_t z = 0; z < e.size[2] - 2 * t.overlap; ++z)
      for (uint64_t y = 0; y < e.size[1] - 2 * t.overlap; ++y) {
        uint64_t gx = e.coords[0] * core[0], gy = e.coords[1] * core[1] + y, gz = e.coords[2] * core[2] + z;
        const uint8_t* src = t.data.data() + e.offset +
                             t.bpv * ((t.overlap) + (y + t.overlap) * e.size[0] + (z + t.overlap) * e.size[0] * e.size[1]);
        uint8_t* dst = vol.data() + t.bpv * (gx + gy * t.dims[0] + gz * t.dims[0] * t.dims[1]);
        std::memcpy(dst, src, (size_t)((e.size[0] - 2 * t.overlap) * t.bpv));
      }
  }
  Field64 F{vol.data(), {t.dims[0], t.dims[1], t.dims[2]}, (int)t.bpv, t.bpv == 1 ? 1.0 / 255.0 : 1.0 / 65535.0, false};
  Frame f;
  f.setup(ds, v, p);
  // volume box in model space
  float mx = (float)std::max(t.dims[0], std::max(t.dims[1], t.dims[2]));
  double ext[3] = {t.dims[0] / (double)mx * ds->scale[0], t.dims[1] / (double)mx * ds->scale[1],
                   t.dims[2] / (double)mx * ds->scale[2]};
  {
    double sm = std::max(ext[0], std::max(ext[1], ext[2]));
    for (int i = 0; i < 3; ++i) ext[i] /= sm;
  }
  double mvd[16], invd[16];
  for (int i = 0; i < 16; ++i) { mvd[i] = f.mv.a[i]; invd[i] = f.inv_mv.a[i]; }
  int x0 = roi[0], y0 = roi[1], x1 = roi[2], y1 = roi[3];
  int w = x1 - x0;
  const double iso = (double)p->isovalue / (double)(1 << p->bit_width);
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int py = y0; py < y1; ++py) {
    for (int px = x0; px < x1; ++px) {
      float* o = out + 4 * ((size_t)(py - y0) * (size_t)w + (size_t)(px - x0));
      V3 dn = f.pixel_dir(px, py);
      double d[3] = {dn.x, dn.y, dn.z};
      double om[3] = {invd[12], invd[13], invd[14]};
      double dm[3];
      for (int c = 0; c < 3; ++c) dm[c] = d[0] * invd[0 * 4 + c] + d[1] * invd[1 * 4 + c] + d[2] * invd[2 * 4 + c];
      double dl = std::sqrt(dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2]);
      for (int c = 0; c < 3; ++c) dm[c] /= dl;  // unit direction in model space
      double tn = -1e300, tf = 1e300;
      bool hit = true;
      for (int a = 0; a < 3; ++a) {
        double lo = -ext[a] / 2, hi = ext[a] / 2;
        if (dm[a] == 0) { if (om[a] < lo || om[a] > hi) hit = false; continue; }
        double t0 = (lo - om[a]) / dm[a], t1 = (hi - om[a]) / dm[a];
        tn = std::max(tn, std::min(t0, t1));
        tf = std::min(tf, std::max(t0, t1));
      }
      o[0] = o[1] = o[2] = o[3] = 0;
      if (!hit || tf <= tn || tf <= 0) continue;
      tn = std::max(tn, 0.0);
      auto to_vox = [&](double tt, double uvw[3]) {
        for (int a = 0; a < 3; ++a) {
          double pm = om[a] + dm[a] * tt;
          uvw[a] = (pm / ext[a] + 0.5) * (double)t.dims[a] - 0.5;
        }
      };
      auto grad = [&](const double u[3], double g[3]) {
        // central differences at +-1 voxel, Volume3D.glsl sign convention
        g[0] = (F.sample(u[0] - 1, u[1], u[2]) - F.sample(u[0] + 1, u[1], u[2])) / 2;
        g[1] = (F.sample(u[0], u[1] - 1, u[2]) - F.sample(u[0], u[1] + 1, u[2])) / 2;
        g[2] = (F.sample(u[0], u[1], u[2] - 1) - F.sample(u[0], u[1], u[2] + 1)) / 2;
      };
      auto eye_of = [&](double tt, double e[3]) {
        double pm[3] = {om[0] + dm[0] * tt, om[1] + dm[1] * tt, om[2] + dm[2] * tt};
        for (int c = 0; c < 3; ++c) e[c] = pm[0] * mvd[0 * 4 + c] + pm[1] * mvd[1 * 4 + c] + pm[2] * mvd[2 * 4 + c] + mvd[12 + c];
      };
      auto nrm = [&](const double g[3], double n[3]) {
        double s[3] = {g[0] * p->domain_scale[0], g[1] * p->domain_scale[1], g[2] * p->domain_scale[2]};
        for (int r = 0; r < 3; ++r) n[r] = f.nm[r][0] * s[0] + f.nm[r][1] * s[1] + f.nm[r][2] * s[2];
        double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l > 0) for (int r = 0; r < 3; ++r) n[r] /= l;
      };
      auto light = [&](const double pos[3], const double n[3], const double dif[3], double c[3], bool iso_mode) {
        double nn[3] = {n[0], n[1], n[2]};
        if (iso_mode) nn[2] = std::fabs(nn[2]);
        double pl = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
        double vd[3] = {-pos[0] / pl, -pos[1] / pl, -pos[2] / pl};
        double dn_ = nn[0] * vd[0] + nn[1] * vd[1] + nn[2] * vd[2];
        double R[3] = {vd[0] - 2 * dn_ * nn[0], vd[1] - 2 * dn_ * nn[1], vd[2] - 2 * dn_ * nn[2]};
        double rl = std::sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        if (rl > 0) for (int k = 0; k < 3; ++k) R[k] /= rl;
        double L[3] = {p->light_dir[0], p->light_dir[1], p->light_dir[2]};
        double nl = nn[0] * L[0] + nn[1] * L[1] + nn[2] * L[2];
        double dterm = std::fabs(nl);
        double rlv = std::max(R[0] * L[0] + R[1] * L[1] + R[2] * L[2], 0.0);
        double sterm = std::pow(rlv, 8.0);
        for (int k = 0; k < 3; ++k)
          c[k] = std::min(std::max(p->light_ambient[k] + dif[k] * dterm + p->light_specular[k] * sterm, 0.0), 1.0);
      };
      if (p->mode == 2) {
        double u[3], prev_t = tn;
        bool found = false;
        double th = tn;
        for (double tt = tn; tt <= tf; tt += step_model) {
          to_vox(tt, u);
          if (F.sample(u[0], u[1], u[2]) >= iso) { found = true; th = tt; break; }
          prev_t = tt;
        }
        if (!found) continue;
        double a = prev_t, b = th;
        for (int k = 0; k < 40 && th != tn; ++k) {
          double m = 0.5 * (a + b);
          to_vox(m, u);
          if (F.sample(u[0], u[1], u[2]) >= iso) b = m; else a = m;
        }
        to_vox(b, u);
        double g[3], n[3], e[3], c[3];
        grad(u, g);
        nrm(g, n);
        eye_of(b, e);
        double dif[3] = {p->light_diffuse[0] * p->iso_color[0], p->light_diffuse[1] * p->iso_color[1],
                         p->light_diffuse[2] * p->iso_color[2]};
        light(e, n, dif, c, true);
        o[0] = (float)c[0]; o[1] = (float)c[1]; o[2] = (float)c[2]; o[3] = 1.0f;
        continue;
      }
      double acc[4] = {0, 0, 0, 0};
      for (double tt = tn; tt <= tf; tt += step_model) {
        double u[3];
        to_vox(tt, u);
        double val = F.sample(u[0], u[1], u[2]);
        double g[3] = {0, 0, 0};
        float s[4];
        if (p->mode == 1 || p->lighting) grad(u, g);
        if (p->mode == 0) {
          f.tf.rgba1((float)(val * p->tf_scale), s);
        } else {
          double gm = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
          f.tf.rgba2((float)(val * p->tf_scale), (float)(1.0 - gm * p->grad_scale), s);
        }
        double rgb[3] = {s[0], s[1], s[2]};
        if (p->lighting) {
          double n[3], e[3], dif[3] = {p->light_diffuse[0] * s[0], p->light_diffuse[1] * s[1], p->light_diffuse[2] * s[2]};
          nrm(g, n);
          eye_of(tt, e);
          light(e, n, dif, rgb, false);
        }
        double a = 1.0 - std::pow(1.0 - (double)s[3], 1.0 / (double)p->srm);
        double wgt = (1.0 - acc[3]) * a;
        for (int k = 0; k < 3; ++k) acc[k] += rgb[k] * wgt;
        acc[3] += a * (1.0 - acc[3]);
      }
      for (int k = 0; k < 4; ++k) o[k] = (float)acc[k];
    }
  }
  return 0;
}

// opacity_pow for the tests that pin it against the closed form
extern "C" void orc_opacity_pow(const float* y, const float* s, int n, float* out) {
  for (int i = 0; i < n; ++i) out[i] = opacity_pow(y[i], s[i]);
}
