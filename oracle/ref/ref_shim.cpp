// C entry points over reference code compiled where it lies (test infrastructure
// only).  Built by `make -C oracle ref` into oracle/_ref/libtvkref.so from
//   Renderer/CullingLOD.cpp          (compiled unchanged)
// and the header-only pieces included below:
//   Basics/Vectors.h                 MATRIX4 product, BuildLookAt, Perspective (:936-946, :1250-1277)
//   IO/UVF/ExtendedOctree/VolumeTools.h  Filter (2, 4 and 8 taps; mean and median) (:168-262)
//   IO/Brick.h                       BrickKey, BKeyHash, BrickTable (:57-74)
// The tests compare the oracle's restatements with these functions; nothing in
// the product loads this library.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
// Vectors.h declares its camera builders under USEGL, next to GL matrix-stack
// helpers that need the GL declarations (the reference's vendored GLEW header)
#include "3rdParty/GLEW/GL/glew.h"
#include "Basics/Vectors.h"
#include "StdTuvokDefines.h"
#include "IO/UVF/ExtendedOctree/VolumeTools.h"
#include "IO/Brick.h"
// The frustum planes are a private member of CullingLOD: the class is read
// with public access so that the planes can be returned as computed.  Its
// layout is unchanged (CullingLOD.cpp itself is compiled as it is).
#define private public
#include "Renderer/CullingLOD.h"
#undef private

namespace {

FLOATMATRIX4 from(const float* a) {
  FLOATMATRIX4 m;
  for (int i = 0; i < 16; ++i) m.array[i] = a[i];
  return m;
}
void to(const FLOATMATRIX4& m, float* a) {
  for (int i = 0; i < 16; ++i) a[i] = m.array[i];
}

// CullingLOD as AbstrRenderer drives it: GLRenderer::ComputeViewAndProjection sets
// the projection and screen parameters (GLRenderer.cpp:885-887), PlanFrame the
// view matrix = the region's modelview, then Update (AbstrRenderer.cpp:1126-1127);
// the model matrix keeps its identity default.
tuvok::CullingLOD make_cull(const float mv[16], const float proj[16], float fov, float aspect, float znear,
                            float zfar, uint32_t height, float sse) {
  tuvok::CullingLOD c(sse);
  c.SetProjectionMatrix(from(proj));
  c.SetScreenParams(fov, aspect, znear, zfar, height);
  c.SetViewMatrix(from(mv));
  c.Update();
  return c;
}

template <typename T>
void filter_all(int taps, int median, const void* in, uint64_t n, void* out) {
  const T* s = static_cast<const T*>(in);
  T* d = static_cast<T*>(out);
  for (uint64_t i = 0; i < n; ++i) {
    const T* v = s + i * (uint64_t)taps;
    if (taps == 8)
      d[i] = median ? VolumeTools::Filter<T, double, true>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7])
                    : VolumeTools::Filter<T, double, false>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
    else if (taps == 4)
      d[i] = median ? VolumeTools::Filter<T, double, true>(v[0], v[1], v[2], v[3])
                    : VolumeTools::Filter<T, double, false>(v[0], v[1], v[2], v[3]);
    else
      d[i] = median ? VolumeTools::Filter<T, double, true>(v[0], v[1]) : VolumeTools::Filter<T, double, false>(v[0], v[1]);
  }
}

}  // namespace

extern "C" {

void ref_look_at(const float eye[3], const float at[3], const float up[3], float out[16]) {
  FLOATMATRIX4 m;
  m.BuildLookAt(FLOATVECTOR3(eye[0], eye[1], eye[2]), FLOATVECTOR3(at[0], at[1], at[2]),
                FLOATVECTOR3(up[0], up[1], up[2]));
  to(m, out);
}

void ref_perspective(float fovy, float aspect, float znear, float zfar, float out[16]) {
  FLOATMATRIX4 m;
  m.Perspective(fovy, aspect, znear, zfar);
  to(m, out);
}

void ref_mat_mul(const float a[16], const float b[16], float out[16]) { to(from(a) * from(b), out); }

// the six frustum planes (x, y, z, w each) of CullingLOD::Update and the LOD factor
void ref_cull_planes(const float mv[16], const float proj[16], float fov, float aspect, float znear, float zfar,
                     uint32_t height, float sse, float planes[24], float* lod_factor) {
  tuvok::CullingLOD c = make_cull(mv, proj, fov, aspect, znear, zfar, height, sse);
  for (int i = 0; i < 6; ++i) {
    planes[4 * i] = c.m_Planes[i].x;
    planes[4 * i + 1] = c.m_Planes[i].y;
    planes[4 * i + 2] = c.m_Planes[i].z;
    planes[4 * i + 3] = c.m_Planes[i].w;
  }
  *lod_factor = c.GetLoDFactor();
}

// CullingLOD::IsVisible for n boxes (center, extent: 3 floats each)
void ref_cull_visible(const float mv[16], const float proj[16], float fov, float aspect, float znear, float zfar,
                      uint32_t height, const float* center, const float* extent, int n, uint8_t* out) {
  tuvok::CullingLOD c = make_cull(mv, proj, fov, aspect, znear, zfar, height, 1.0f);
  for (int i = 0; i < n; ++i)
    out[i] = c.IsVisible(FLOATVECTOR3(center[3 * i], center[3 * i + 1], center[3 * i + 2]),
                         FLOATVECTOR3(extent[3 * i], extent[3 * i + 1], extent[3 * i + 2]))
                 ? 1
                 : 0;
}

// CullingLOD::GetLODLevel for n boxes (unclamped, as the reference returns it)
void ref_cull_lod_level(const float mv[16], const float proj[16], float fov, float aspect, float znear, float zfar,
                        uint32_t height, float sse, const float* center, const float* extent,
                        const uint32_t* voxels, int n, int* out) {
  tuvok::CullingLOD c = make_cull(mv, proj, fov, aspect, znear, zfar, height, sse);
  for (int i = 0; i < n; ++i)
    out[i] = c.GetLODLevel(FLOATVECTOR3(center[3 * i], center[3 * i + 1], center[3 * i + 2]),
                           FLOATVECTOR3(extent[3 * i], extent[3 * i + 1], extent[3 * i + 2]),
                           UINTVECTOR3(voxels[3 * i], voxels[3 * i + 1], voxels[3 * i + 2]));
}

// VolumeTools::Filter over n groups of `taps` values (taps 2, 4 or 8), bpv 1 or 2
void ref_filter(int taps, int median, int bpv, const void* in, uint64_t n, void* out) {
  if (bpv == 1) filter_all<uint8_t>(taps, median, in, n, out);
  else filter_all<uint16_t>(taps, median, in, n, out);
}

// Iteration order of a BrickTable (unordered_map with the reference's BKeyHash)
// filled the way UVFDataset::ComputeMetadataTOC fills it (uvfDataset.cpp:242-288):
// NBricksHint's rehash over all bricks (BrickedDataset.cpp:44-49), then per LOD
// keys inserted with x outermost and z innermost.  bricks: nl x 3 brick counts.
// out: (lod, index) pairs, 2 per brick.
void ref_table_order(const uint64_t* bricks, int nl, uint64_t* out) {
  tuvok::BrickTable t;
  size_t nall = 0;
  for (int j = 0; j < nl; ++j) nall += (size_t)(bricks[3 * j] * bricks[3 * j + 1] * bricks[3 * j + 2]);
  t.rehash(size_t(std::ceil(float(nall) / t.max_load_factor())));
  for (int j = 0; j < nl; ++j) {
    const uint64_t bx = bricks[3 * j], by = bricks[3 * j + 1], bz = bricks[3 * j + 2];
    for (uint64_t x = 0; x < bx; ++x)
      for (uint64_t y = 0; y < by; ++y)
        for (uint64_t z = 0; z < bz; ++z)
          t.insert(std::make_pair(tuvok::BrickKey(0, (size_t)j, (size_t)(z * bx * by + y * bx + x)), tuvok::BrickMD()));
  }
  size_t i = 0;
  for (auto it = t.begin(); it != t.end(); ++it, ++i) {
    out[2 * i] = std::get<1>(it->first);
    out[2 * i + 1] = std::get<2>(it->first);
  }
}

}  // extern "C"
