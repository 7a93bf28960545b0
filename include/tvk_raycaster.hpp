// tvk_raycaster.hpp -- header-only C++ front end over the C ABI (tvk_hip.h).
//
// tvk::Raycaster mirrors the part of tuvok::AbstrRenderer (Renderer/AbstrRenderer.h:112-881)
// that drives a frame of GLRaycaster, with the same method names, so a Tuvok
// renderer subclass (an AbstrRenderer holding a tvk::Raycaster) forwards its
// virtuals one to one.
// Errors surface as tvk::Error exceptions on this (C++) side of the boundary,
// like the std::runtime_error / std::domain_error the reference throws.
#pragma once
#include <array>
#include <stdexcept>
#include <string>
#include <vector>
#include "tvk_hip.h"

namespace tvk {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

class Raycaster {
 public:
  explicit Raycaster(int device = 0) {
    int rc = tvk_create(&ctx_, device);
    if (rc != TVK_OK) throw Error(rc, "tvk_create failed");
    tvk_default_params(&p_);
    p_.lod = 0;
  }
  // AbstrRenderer(mc, bUseOnlyPowerOfTwo, bDownSampleTo8Bits, bDisableBorder)
  Raycaster(int device, bool pow2, bool to8, bool disable_border) : Raycaster(device) {
    chk(tvk_set_compat(ctx_, pow2 ? 1 : 0, to8 ? 1 : 0));
    p_.clamp_to_edge = disable_border ? 1 : 0;
    push();
  }
  ~Raycaster() { tvk_destroy(ctx_); }
  Raycaster(const Raycaster&) = delete;
  Raycaster& operator=(const Raycaster&) = delete;

  // dataset: RAWConverter::ConvertRAWDataset + UVFDataset load, in HBM
  void LoadVolume(const void* voxels, int dtype, const uint64_t dims[3], const uint32_t brick[3],
                  uint32_t overlap, bool clamp_to_edge = false, int max_lods = 0) {
    chk(tvk_load_volume(ctx_, voxels, dtype, dims, brick, overlap, clamp_to_edge, max_lods, 0));
  }

  // a BrickedDataset handed over brick by brick: the layout and every brick's
  // key and MaxMinForKey range, then each Dataset::GetBrick payload
  void SetBricks(int dtype, const uint64_t dims[3], const uint32_t brick[3], uint32_t overlap,
                 const std::vector<tvk_brick_key>& keys, int lod_count = 0, uint64_t budget_bytes = 0) {
    chk(tvk_set_bricks(ctx_, dtype, dims, brick, overlap, lod_count, keys.data(), keys.size(), budget_bytes));
  }
  void UploadBrick(uint32_t lod, uint64_t index, const void* data, size_t bytes) {
    chk(tvk_upload_brick(ctx_, lod, index, data, bytes));
  }
  // GPUMemMan's memory budget for this dataset's bricks (0: all resident)
  void SetPoolBudget(uint64_t bytes) { chk(tvk_set_pool_budget(ctx_, bytes)); }
  // ExtendedOctreeConverter's bComputeMedian for the next bricking
  void SetLODFilter(bool median) { chk(tvk_set_lod_filter(ctx_, median ? 1 : 0)); }
  // TVK_DISPATCH_RASTER / _GROUPS / _ROWS: the order multi-view renders hand
  // their pixel groups to the GPU (frames are identical in every order)
  void SetDispatchOrder(int order) { chk(tvk_set_dispatch_order(ctx_, order)); }
  // SetClipPlane + EnableClipPlane / DisableClipPlane (AbstrRenderer.h:215-219)
  void SetClipPlane(const std::array<float, 4>& plane, bool enabled = true) {
    chk(tvk_set_clip_plane(ctx_, plane.data(), enabled ? 1 : 0));
  }
  // AbstrRenderer::CropDataset (AbstrRenderer.h:480)
  std::array<float, 4> CropDataset(bool keep_old = false) {
    std::array<float, 4> used{};
    chk(tvk_crop_dataset(ctx_, keep_old ? 1 : 0, used.data()));
    return used;
  }
  // AbstrRenderer::Pick (AbstrRenderer.h:245): the isosurface hit under a window position
  std::array<float, 3> Pick(uint32_t x, uint32_t y) {
    std::array<float, 3> pos{};
    chk(tvk_pick(ctx_, x, y, pos.data()));
    return pos;
  }
  tvk_pool_stats PoolStats() {
    tvk_pool_stats s{};
    chk(tvk_get_pool_stats(ctx_, &s));
    return s;
  }
  // UVFDataset::MaxMinForKey by TOC index
  std::array<double, 2> MaxMinForKey(uint64_t toc) {
    std::array<double, 2> mm{};
    chk(tvk_brick_minmax(ctx_, toc, mm.data()));
    return mm;
  }

  // AbstrRenderer::Set1DTrans / Changed2DTrans
  void Set1DTrans(const std::vector<unsigned char>& rgba) {
    chk(tvk_set_tf1d(ctx_, rgba.data(), (int)(rgba.size() / 4)));
  }
  void Set2DTrans(const std::vector<unsigned char>& rgba, int w, int h) { chk(tvk_set_tf2d(ctx_, rgba.data(), w, h)); }

  // state setters
  void SetRendermode(int mode) { p_.mode = mode; push(); }
  void SetUseLighting(bool on) { p_.lighting = on ? 1 : 0; push(); }
  void SetSampleRateModifier(float f) { p_.srm = f; push(); }
  void SetIsoValue(float v) { p_.isovalue = v; push(); }
  void SetGradientScale(float s) { p_.grad_scale = s; push(); }
  void SetEarlyRayTermination(float eps) { p_.ert_eps = eps; push(); }

  // view: SetViewParameters + eye/at/up + Resize (GLRenderer::ComputeViewAndProjection)
  void SetViewParameters(float fov, float znear, float zfar) { fov_ = fov; zn_ = znear; zf_ = zfar; cam(); }
  void SetViewPosition(std::array<float, 3> eye, std::array<float, 3> at = {0, 0, 0},
                       std::array<float, 3> up = {0, 1, 0}) {
    eye_ = eye; at_ = at; up_ = up; cam();
  }
  void Resize(uint32_t w, uint32_t h) { w_ = w; h_ = h; cam(); }
  void SetUserMatrices(const float view[16], const float proj[16], const float rot[16], const float trans[16]) {
    chk(tvk_set_view_matrices(ctx_, view, rot, trans, proj, w_, h_, fov_, zn_, zf_));
  }
  void SetRotation(const float rot[16]) { chk(tvk_set_rotation(ctx_, rot)); }

  // AbstrRenderer::PlanFrame / BuildSubFrameBrickList
  std::vector<tvk_brick_rec> PlanFrame() {
    int n = 0;
    chk(tvk_plan(ctx_, nullptr, 0, &n));
    std::vector<tvk_brick_rec> v((size_t)n);
    chk(tvk_plan(ctx_, v.data(), n, &n));
    return v;
  }

  // GLRenderer::Paint; returns the premultiplied RGBA32F frame (GL row order)
  std::vector<float> Paint(tvk_stats* st = nullptr) {
    std::vector<float> img((size_t)w_ * h_ * 4);
    chk(tvk_render(ctx_, img.data(), st));
    return img;
  }

  tvk_ctx* handle() const { return ctx_; }

 private:
  void chk(int rc) const {
    if (rc != TVK_OK) throw Error(rc, tvk_last_error(ctx_));
  }
  void push() { chk(tvk_set_params(ctx_, &p_)); }
  void cam() {
    if (!w_ || !h_) return;
    chk(tvk_set_camera(ctx_, eye_.data(), at_.data(), up_.data(), fov_, zn_, zf_, w_, h_));
  }
  tvk_ctx* ctx_ = nullptr;
  tvk_params p_{};
  std::array<float, 3> eye_{{0.f, 0.f, 1.6f}}, at_{{0.f, 0.f, 0.f}}, up_{{0.f, 1.f, 0.f}};
  float fov_ = 50.f, zn_ = 0.01f, zf_ = 1000.f;
  uint32_t w_ = 0, h_ = 0;
};

}  // namespace tvk
