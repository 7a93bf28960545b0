// Host-side half of the raycasting path: ExtendedOctree metadata, the brick
// table of UVFDataset, transfer-function limits and the per-frame planner of
// AbstrRenderer/GLRaycaster.  Everything here is O(bricks) per frame; the
// per-voxel and per-ray work runs in the HIP kernels.
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>
#include "../tvk_types.hpp"
#include "mat.hpp"

namespace tvk {

struct LodInfo {
  uint64_t size[3];     // voxels of the LOD
  uint64_t bricks[3];   // bricks per axis
  uint64_t offset;      // first TOC index of the LOD
  double aspect[3];     // anisotropic downsampling aspect
};

struct TocBrick {
  uint64_t coords[4];   // x, y, z, lod
  uint32_t size[3];     // stored voxels incl. overlap
  uint64_t offset;      // element offset in the pool (blocked layout)
  uint64_t count;       // elements in the pool (padded to whole blocks)
  uint64_t real;        // voxels of the brick (reference layout)
  uint64_t ref_offset;  // element offset in the reference layout (TOC order, unpadded)
};

// BrickKey = <timestep, LOD, 1D index>; hash of the reference (Brick.h:63-73) so
// that the table iterates in the same order as BrickedDataset::bricks.
typedef std::tuple<size_t, size_t, size_t> BrickKey;
struct BrickKeyHash {
  size_t operator()(const BrickKey& k) const {
    size_t ts = std::hash<size_t>()(std::get<0>(k));
    size_t lod = std::hash<size_t>()(std::get<1>(k));
    size_t idx = std::hash<size_t>()(std::get<2>(k));
    size_t seed = lod;
    seed ^= idx + 0x9e3779b9 + (seed << 6) + (seed >> 2);
    seed ^= ts + 0x9e3779b9 + (seed << 6) + (seed >> 2);
    return seed;
  }
};
struct BrickMeta {
  F3 center, extents;
  uint32_t n_voxels[3];
};
typedef std::unordered_map<BrickKey, BrickMeta, BrickKeyHash> BrickTable;

struct Layout {
  uint64_t dims[3] = {0, 0, 0};
  uint32_t brick[3] = {0, 0, 0};
  uint32_t overlap = 0;
  int bpv = 1;
  float scale[3] = {1, 1, 1};
  std::vector<LodInfo> lods;
  std::vector<TocBrick> toc;
  BrickTable table;
  uint64_t pool_elems = 0;   // blocked, padded
  uint64_t ref_elems = 0;    // reference layout

  // ExtendedOctree::ComputeMetadata (ExtendedOctree.cpp:188-243); keeps the first
  // max_lods levels (0 = all, like the reference).
  bool build(const uint64_t d[3], const uint32_t b[3], uint32_t ov, int bytes, int max_lods,
             std::string* err);
  void brick_size(const uint64_t c[4], uint32_t out[3]) const;
  uint64_t toc_index(uint64_t x, uint64_t y, uint64_t z, uint64_t lod) const;
  // UVFDataset::GetTextCoords (uvfDataset.cpp:1747-1774)
  void texcoords(int lod, uint64_t lin, F3& tmin, F3& tmax, bool pow2 = false) const;
};

// MathTools::NextPow2 (MathTools.h:71): the smallest power of two >= n
inline uint32_t next_pow2(uint32_t n) {
  uint32_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

struct ViewState {
  Mat4 view, rot, trans, proj;
  uint32_t W = 0, H = 0;
  float fov = 50.f, znear = 0.01f, zfar = 1000.f;
  bool clip_on = false;        // m_bClipPlaneOn
  float clip[4] = {0, 0, 1, 0};   // m_ClipPlane.Plane(), world space
  Mat4 modelview() const { return rot * trans * view; }  // GLRenderer.cpp:627
};

struct PlanParams {
  int mode = 0;
  int lod = 0;
  float srm = 1.f;
  float isovalue = 0.f;
  uint64_t lim[4] = {0, 0, 0, 0};  // TF non-zero limits (1D: [0..1], 2D: [0..3])
  int tf_n = 256;                  // 1D TF size or 2D TF width
  bool pow2 = false;               // power-of-two textures (texture coordinates of the padded brick)
};

struct PlannedBrick {
  uint64_t key_index, toc_index;
  F3 center, extension;
  uint32_t voxel_count[3];
  F3 tex_min, tex_max;
  float distance;
  int is_empty;
};

// AbstrRenderer::PlanFrame / BuildSubFrameBrickList (AbstrRenderer.cpp:999-1212)
void plan_frame(const Layout& L, const std::vector<double>& minmax, double range_max,
                const ViewState& V, const PlanParams& P, std::vector<PlannedBrick>& out);
// ComputeMinLODForCurrentView (AbstrRenderer.cpp:789-803)
int min_lod_for_view(const Layout& L, const ViewState& V, float sse);
// Per-brick uniforms + eye->texture matrix + screen rect (GLRaycaster.cpp:239-345, :589-612)
// tex_dims/tex_offset (per TOC brick, may be null): the textures rendering
// reads when they are not the pool's bricks (power-of-two padded sizes)
void make_frame_bricks(const Layout& L, const ViewState& V, const PlanParams& P,
                       const std::vector<PlannedBrick>& list, std::vector<FrameBrick>& out,
                       const uint32_t* tex_dims = nullptr, const uint64_t* tex_offset = nullptr);
void make_view_const(const ViewState& V, ViewConst& vc);
// Work estimate of each 16x16 group of a view (vc.groups entries from cost):
// the samples a ray takes across every brick whose pixel rectangle covers the
// group (the box's longest edge over its ray step), summed, plus a twentieth of
// the largest brick's weight for the fixed cost (ray set-up, brick tests, the
// pixel write).
void group_costs(const ViewConst& vc, const FrameBrick* list, float* cost);
// The dispatch order of a launch's groups (cost: group_costs of every view at
// its group_offset): 8 contiguous ranges of about equal group count, one per
// XCD, each range's units -- single groups (mode 1) or rows of a view's groups
// (mode 2) -- in decreasing cost, groups of a row in order: order[start[x] + j]
// is XCD x's j-th group, j < count[x].
void dispatch_order(const ViewConst* views, int nviews, const std::vector<float>& cost, int mode,
                    std::vector<uint32_t>& order, uint32_t start[8], uint32_t count[8]);

// TransferFunction1D::Set(rgba8) + ComputeNonZeroLimits; TransferFunction2D::ComputeNonZeroLimits
void tf1d_limits(const uint8_t* rgba8, int n, uint64_t lim[2]);
void tf2d_limits(const uint8_t* rgba8, int w, int h, uint64_t lim[4]);

}  // namespace tvk
