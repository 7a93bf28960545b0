/* tvk_hip.h -- C ABI of the MI355X-native Tuvok raycasting path.
 *
 * One context = one GPU + one HIP stream.  The library owns all device memory;
 * the caller owns every host buffer it passes.  Every call returns TVK_OK (0)
 * or a negative error code; tvk_last_error() gives the message.  No C++
 * exception crosses this boundary.  A context is single-threaded (like the
 * GL context that owns a Tuvok renderer).
 *
 * The entry points replace the following reference interfaces:
 *   tvk_load_volume / tvk_generate_volume
 *       RAWConverter::ConvertRAWDataset -> ExtendedOctreeConverter::Convert
 *       (IO/RAWConverter.h:79-100, IO/UVF/ExtendedOctree/ExtendedOctreeConverter.cpp:128-280)
 *       and the upload half of GPUMemMan::GetVolume (Renderer/GPUMemMan/GPUMemMan.cpp:846-996)
 *   tvk_lod_info        ExtendedOctree::GetLoDSize/GetBrickCount/GetBrickAspect
 *                       (IO/UVF/ExtendedOctree/ExtendedOctree.cpp:255-294)
 *   tvk_brick_md        BrickedDataset::GetBrickMetadata (IO/BrickedDataset.cpp:52-67),
 *                       filled as UVFDataset::ComputeMetadataTOC (IO/uvfDataset.cpp:242-288)
 *   tvk_brick_minmax    UVFDataset::MaxMinForKey (IO/uvfDataset.cpp:1183-1199)
 *   tvk_get_brick       Dataset::GetBrick (IO/Dataset.h:93-100)
 *   tvk_set_tf1d        AbstrRenderer::Set1DTrans (Renderer/AbstrRenderer.h:189)
 *   tvk_set_tf2d        AbstrRenderer::Changed2DTrans (Renderer/AbstrRenderer.h:195)
 *   tvk_set_camera      AbstrRenderer::SetViewParameters + GLRenderer::ComputeViewAndProjection
 *                       (Renderer/AbstrRenderer.h:250-261, Renderer/GL/GLRenderer.cpp:889-920)
 *   tvk_set_view_matrices  AbstrRenderer::SetUserMatrices + RenderRegion rotation/translation
 *                       (Renderer/AbstrRenderer.h:236-238, Renderer/RenderRegion.h:50-196)
 *   tvk_set_params      SetRendermode / SetUseLighting / SetSampleRateModifier / SetIsoValue
 *                       (Renderer/AbstrRenderer.h:323, :414-418)
 *   tvk_plan            AbstrRenderer::PlanFrame + BuildSubFrameBrickList
 *                       (Renderer/AbstrRenderer.cpp:999-1100, :1125-1212)
 *   tvk_render          GLRenderer::Paint -> Render3DRegion -> Render3DView with
 *                       GLRaycaster::Render3DInLoop (Renderer/GL/GLRenderer.cpp:571-667,
 *                       :2663-2748; Renderer/GL/GLRaycaster.cpp:383-478); the output is the
 *                       premultiplied RGBA32F image of GLRenderer::GetLastFBO (GLRenderer.h:105)
 */
#ifndef TVK_HIP_H
#define TVK_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct tvk_ctx tvk_ctx;

enum {
  TVK_OK = 0,
  TVK_ERR_ARG = -1,      /* invalid argument                                   */
  TVK_ERR_HIP = -2,      /* HIP runtime error                                  */
  TVK_ERR_STATE = -3,    /* call out of order (no dataset, no TF, ...)         */
  TVK_ERR_RANGE = -4,    /* index/key out of range (BrickedDataset domain_error)*/
  TVK_ERR_NOMEM = -5,    /* device allocation failed                           */
  TVK_ERR_UNSUPPORTED = -6
};

enum { TVK_U8 = 1, TVK_U16 = 2 };                          /* bytes per voxel */
enum { TVK_RM_1DTRANS = 0, TVK_RM_2DTRANS = 1, TVK_RM_ISOSURFACE = 2 }; /* ERenderMode */

/* synthetic volume generators (tvk_generate_volume) */
enum {
  TVK_GEN_BLOBS = 1,      /* sum of isotropic Gaussians + integer noise         */
  TVK_GEN_FBM_SHELLS = 2, /* 3-octave value-noise fBm + hollow spherical shells  */
  TVK_GEN_ELLIPSOID_SDF = 3, /* clamp(128 + 20*d_voxels) of random ellipsoids    */
  TVK_GEN_PHANTOM = 4     /* nested ellipsoids, piecewise constant + noise       */
};

typedef struct {
  int mode;              /* TVK_RM_*                                               */
  int lighting;          /* m_bUseLighting                                          */
  float srm;             /* m_fSampleRateModifier                                   */
  float isovalue;        /* m_fIsovalue, data units                                 */
  int lod;               /* LOD to render (<0: ComputeMinLODForCurrentView)          */
  float tf_scale;        /* fTransScale; <=0 -> computed as CalculateScaling()      */
  float grad_scale;      /* fGradientScale = 1/MaxGradientMagnitude                  */
  float domain_scale[3]; /* vDomainScale = 1/GetScale()                              */
  float light_ambient[3], light_diffuse[3], light_specular[3], light_dir[3];
  float iso_color[3];
  int refine_steps;      /* RefineIsosurface iterations (reference: 5)               */
  int clamp_to_edge;     /* 0: GL_CLAMP (border 0, bDisableBorder=false), 1: CLAMP_TO_EDGE */
  float ert_eps;         /* 0: exact reference semantics.  >0: a ray stops once its
                            total transmittance is below ert_eps (error <= ert_eps
                            per channel); decided per wavefront by ballot.         */
  float screen_space_error; /* CullingLOD m_fScreenSpaceError (for lod < 0)          */
} tvk_params;

typedef struct {
  uint64_t key_index;    /* BrickKey 1D index within the LOD                    */
  uint64_t toc_index;    /* index into the brick pool (TOC order)               */
  float center[3], extension[3];
  uint32_t voxel_count[3];
  float tex_min[3], tex_max[3];
  float distance;        /* Brick::fDistance                                    */
  int is_empty;          /* Brick::bIsEmpty                                     */
} tvk_brick_rec;

typedef struct {
  int lod_count;
  uint64_t brick_count;  /* all LODs                                            */
  uint64_t pool_bytes;   /* HBM bytes of the brick pool (blocked layout)        */
  uint64_t brick_bytes;  /* bytes of all bricks in the reference layout         */
  double range_min, range_max;   /* UVFDataset::GetRange                        */
  int bit_width;
  uint64_t dims[3];
  uint32_t brick_size[3];
  uint32_t overlap;
  float build_ms;        /* device time of the last bricking: gathers, pyramid,
                            overlap fill, row copies, min/max                    */
} tvk_dataset_desc;

typedef struct {
  uint64_t samples;      /* composited samples executed (when counting is on)   */
  uint32_t bricks_in_list, bricks_nonempty;
  float ms;              /* device time of the last tvk_render call             */
} tvk_stats;

/* ---- context ------------------------------------------------------------ */
int tvk_create(tvk_ctx** out, int device);
void tvk_destroy(tvk_ctx* ctx);
const char* tvk_last_error(const tvk_ctx* ctx);
int tvk_default_params(tvk_params* p);   /* AbstrRenderer constructor defaults */

/* ---- dataset ------------------------------------------------------------ */
/* Brick a linear volume (x fastest) into the ExtendedOctree layout in HBM:
 * brick = stored brick size incl. 2*overlap, LOD pyramid by 2x box filter
 * (mean, truncating), per-brick min/max over all stored voxels.  max_lods = 0
 * builds every level down to 1x1x1 like the reference; k > 0 keeps the first
 * k levels.  keep_linear retains the linear LOD0 volume for tvk_download_volume.
 * Bricks are at most 4096 voxels a side and their blocked device layout at most
 * 4 GiB (the kernels use 32-bit offsets inside a brick): TVK_ERR_ARG otherwise. */
int tvk_load_volume(tvk_ctx* ctx, const void* host_voxels, int dtype, const uint64_t dims[3],
                    const uint32_t brick[3], uint32_t overlap, int clamp_to_edge, int max_lods,
                    int keep_linear);
/* Same, with the linear volume generated on the device (TVK_GEN_*). */
int tvk_generate_volume(tvk_ctx* ctx, int kind, uint64_t seed, int dtype, const uint64_t dims[3],
                        const uint32_t brick[3], uint32_t overlap, int clamp_to_edge,
                        int max_lods, int keep_linear);
/* The filter of the coarser levels for the next bricking (load, generate,
 * rebrick, crop): ExtendedOctreeConverter::Convert's bComputeMedian
 * (ExtendedOctreeConverter.h:97-113; VolumeTools::Filter, VolumeTools.h:150-258):
 * 0 = the mean of each 2x2x2 group (default), 1 = the reference's median.  Odd
 * brick cores (brick size - 2*overlap) filter brick to brick as the reference
 * does (ExtendedOctreeConverter.inc:250-356) and need an overlap of at least 1
 * and bClampToEdge off when coarser levels are kept (TVK_ERR_UNSUPPORTED). */
int tvk_set_lod_filter(tvk_ctx* ctx, int median);
/* How a multi-view render (tvk_render, tvk_render_views) hands its 16x16 pixel
 * groups to the GPU; no counterpart in the reference, whose GLRaycaster draws
 * one brick at a time (GLRaycaster.cpp:408-470).  The order:
 * TVK_DISPATCH_RASTER: each XCD runs a contiguous eighth of the groups in
 * raster order.  TVK_DISPATCH_GROUPS / TVK_DISPATCH_ROWS: the same eighths,
 * each run most costly group (or row of groups) first, the cost being the
 * samples a ray takes across the non-empty bricks covering the group -- long
 * rays then start early instead of trailing the launch.  Or'ed with
 * TVK_DISPATCH_QUEUES: a persistent launch whose waves pull 8x8 tiles of those
 * eighths, in that order, from one work queue per XCD (then from the other
 * XCDs'), instead of one workgroup per group.  Default: RASTER | QUEUES.
 * Frames are identical in every setting.  tvk_render_views_noalloc always
 * runs raster order with one workgroup per group (it allocates nothing). */
#define TVK_DISPATCH_RASTER 0
#define TVK_DISPATCH_GROUPS 1
#define TVK_DISPATCH_ROWS 2
#define TVK_DISPATCH_QUEUES 4
int tvk_set_dispatch_order(tvk_ctx* ctx, int order);
/* Copy the retained linear LOD0 volume back to the host. */
/* Makes dst render src's dataset: the brick pool (and kept linear volume) in HBM
 * is shared, not copied, and stays allocated while any context holds it; dst
 * keeps its own TF, parameters and camera.  GPUMemMan shares one dataset's
 * textures between renderers the same way (GPUMemMan.cpp:846-996). */
int tvk_share_dataset(tvk_ctx* dst, const tvk_ctx* src);
/* Re-bricks the dataset from its kept linear volume with another brick size,
 * overlap and level count (ExtendedOctreeConverter::Convert with new
 * parameters, on the device); used by the brick-size autotuner. */
int tvk_rebrick(tvk_ctx* ctx, const uint32_t brick[3], uint32_t overlap, int max_lods);
int tvk_download_volume(tvk_ctx* ctx, void* host_out, size_t bytes);
/* AbstrRenderer::CropDataset (GLRenderer.cpp:2953-2971, UVFDataset::Crop,
 * uvfDataset.cpp:1470-1580): the clip plane without the viewing transformation,
 * made scale invariant, zeroes every LOD-0 voxel whose normalised coordinate
 * (x/X - 0.5, ...) satisfies (n.p) + w >= 0 (CropData, uvfDataset.h:252-297); the
 * volume is then bricked again with the dataset's brick size, overlap and level
 * count (the kept linear volume, or LOD 0 flattened from the resident bricks).  No
 * file is written (keep_old has no effect).  plane_used (may be NULL) receives the
 * normalised-coordinate plane that was applied. */
int tvk_crop_dataset(tvk_ctx* ctx, int keep_old, float plane_used[4]);
/* The renderer's compatibility flags, AbstrRenderer's constructor arguments
 * bUseOnlyPowerOfTwo and bDownSampleTo8Bits (AbstrRenderer.h:160-174;
 * bDisableBorder is the clamp_to_edge parameter).  Rendering then reads
 * textures made from the pool's bricks as GLVolumeListElem::CreateTexture makes
 * them (GPUMemManDataStructs.cpp:292-360, :396-425): 16-bit bricks quantized
 * to (unsigned char)((255 v - fMin) / (fMax - fMin)) over the dataset range
 * (x86's conversion of out-of-range values: truncation to 32 bits, low byte),
 * and/or bricks padded to the next power of two per axis (one repeated element,
 * row and slice past the data unless the border is disabled, zeros beyond),
 * with the power-of-two texture coordinates, voxel step and ray step
 * (UVFDataset::GetTextCoords, uvfDataset.cpp:1747-1774; GLRaycaster.cpp:242-256),
 * a TF scale of 1 and the 8-bit isovalue normalisation (GLRenderer.cpp:1813-1819,
 * AbstrRenderer.cpp:412-438).  The textures are made (in HBM, beside the
 * pool) on the next render after the dataset or the flags change.  GridLeaper,
 * a pool budget and the step proposal read the pool's own bricks and return
 * TVK_ERR_UNSUPPORTED while a flag is set; the allocation-free render returns
 * TVK_ERR_STATE until a render has made the textures. */
int tvk_set_compat(tvk_ctx* ctx, int use_only_power_of_two, int downsample_to_8bits);
int tvk_dataset_info(tvk_ctx* ctx, tvk_dataset_desc* out);
int tvk_lod_info(tvk_ctx* ctx, int lod, uint64_t size[3], uint64_t bricks[3], uint64_t* offset,
                 double aspect[3]);
int tvk_brick_info(tvk_ctx* ctx, uint64_t toc_index, uint64_t coords[4], uint32_t size[3],
                   uint64_t* pool_offset);
int tvk_brick_md(tvk_ctx* ctx, int lod, uint64_t index, float center[3], float extents[3],
                 uint32_t n_voxels[3]);
int tvk_brick_minmax(tvk_ctx* ctx, uint64_t toc_index, double minmax[2]);
/* Brick payloads in the reference layout (x fastest, stored size incl. overlap);
 * tvk_get_pool returns every brick in TOC order (brick_bytes in total) and
 * tvk_brick_info's pool_offset is the byte offset of a brick in it. */
int tvk_get_brick(tvk_ctx* ctx, uint64_t toc_index, void* host_out, size_t bytes);
int tvk_get_pool(tvk_ctx* ctx, void* host_out, size_t bytes);
/* ---- brick-level ingest and the budgeted pool -------------------------------
 * A dataset handed over brick by brick, as a BrickedDataset holds it: the
 * ExtendedOctree layout (dims, stored brick size incl. 2*overlap, overlap,
 * lod_count levels, 0 = every level down to 1x1x1) and, per brick, its key
 * and MaxMinForKey range (BrickedDataset::GetBrickMetadata / MaxMinForKey,
 * IO/BrickedDataset.cpp:52-67, IO/uvfDataset.cpp:1183-1199).  Every brick of
 * the layout is listed once (TVK_ERR_RANGE for a key outside it, the
 * BrickedDataset domain_error).  budget_bytes = 0 (or a budget holding the
 * whole pool): every brick resident in HBM, and rendering waits until every
 * payload was uploaded.  Otherwise the bricks stay in pinned host memory and
 * the HBM pool holds as many brick slots as the budget allows (at least 2):
 * tvk_render pages the bricks of a frame in, least recently used slot first
 * (GPUMemMan::GetVolume, Renderer/GPUMemMan/GPUMemMan.cpp:846-996), and renders
 * a frame whose bricks exceed the slots in parts, in compositing order, with
 * the same result; the traversal pages with GLVolumePool's policy.  Batches and
 * the learned step need every brick resident. */
typedef struct {
  uint32_t lod;          /* BrickKey LOD                                        */
  uint64_t index;        /* BrickKey 1D index within the LOD (x + y*bx + z*bx*by) */
  double minmax[2];      /* MaxMinForKey: min and max over the stored voxels     */
} tvk_brick_key;
int tvk_set_bricks(tvk_ctx* ctx, int dtype, const uint64_t dims[3], const uint32_t brick[3], uint32_t overlap,
                   int lod_count, const tvk_brick_key* keys, size_t n, uint64_t budget_bytes);
/* A UVF file's ExtendedOctree dataset into the pool: UVFDataset's open of a
 * version-5 file (IO/uvfDataset.cpp; UVF::Open, IO/UVF/UVF.cpp:74-305) with the
 * octree's header, TOC and bricks as ExtendedOctree::Open / GetBrickData read
 * them (IO/UVF/ExtendedOctree/ExtendedOctree.cpp:87-165, :313-360): scalar
 * 8- or 16-bit voxels, bricks uncompressed or zlib-compressed, not stored as
 * 2D atlases (TVK_ERR_UNSUPPORTED otherwise).  The bricks go through
 * tvk_set_bricks / tvk_upload_brick with budget_bytes (0: every brick in HBM),
 * with each brick's min/max taken from its stored voxels; the dataset's
 * voxel aspect becomes its scale.  info (may be NULL) describes the file. */
typedef struct {
  uint64_t file_version;   /* GlobalHeader ulFileVersion                         */
  uint32_t octree_version; /* ExtendedOctree m_iVersion                          */
  uint64_t octree_offset;  /* byte offset of the octree header in the file      */
  uint64_t dims[3];        /* m_vVolumeSize                                      */
  uint32_t brick_size[3];  /* m_iBrickSize (stored, incl. overlap)               */
  double aspect[3];        /* m_vVolumeAspect                                    */
  uint32_t overlap;
  int bit_width;
  int lod_count;
  uint64_t brick_count;
  uint64_t zlib_bricks;    /* bricks stored zlib-compressed                     */
} tvk_uvf_info;
int tvk_open_uvf(tvk_ctx* ctx, const char* path, uint64_t budget_bytes, tvk_uvf_info* info);
/* One brick's payload, Dataset::GetBrick (IO/Dataset.h:93-100): the reference
 * layout, x fastest, stored size incl. overlap (bytes must match). */
int tvk_upload_brick(tvk_ctx* ctx, uint32_t lod, uint64_t index, const void* data, size_t bytes);
/* Holds the current dataset within budget_bytes of HBM (as above; its bricks are
 * copied to pinned host memory first), or, with 0, makes every brick resident
 * again.  Contexts that shared the dataset before keep their own pool. */
int tvk_set_pool_budget(tvk_ctx* ctx, uint64_t budget_bytes);
typedef struct {
  uint64_t pool_bytes;   /* HBM bytes of the pool                                    */
  uint64_t slots;        /* brick slots (every brick when the pool is not budgeted)   */
  uint64_t slot_bytes;
  uint64_t resident;     /* slots holding a brick                                    */
  uint64_t uploads, evictions, upload_bytes;   /* since the budget was set             */
  float upload_ms;       /* host time spent in uploads                                */
  uint32_t parts;        /* launches of the last budgeted frame                       */
} tvk_pool_stats;
int tvk_get_pool_stats(tvk_ctx* ctx, tvk_pool_stats* out);

/* MaxGradientMagnitude of the dataset's 2D histogram, as RAWConverter computes it at
 * level 0 (Histogram2DDataBlock.cpp:349-395, RAWConverter.cpp:655-660): the largest
 * |(f(x-1)-f(x+1))/(2*(2^bits-1)), ...| over the core voxels of the LOD-0 bricks, with
 * the neighbours from the bricks' overlap (a dataset without overlap uses the
 * kept linear volume, zero outside).  fGradientScale is its inverse. */
int tvk_max_gradient(tvk_ctx* ctx, float* out);

/* ---- transfer functions --------------------------------------------------- */
int tvk_set_tf1d(tvk_ctx* ctx, const uint8_t* rgba8, int n);
int tvk_set_tf2d(tvk_ctx* ctx, const uint8_t* rgba8, int w, int h);

/* 2D TF from swatches, TransferFunction2D::RenderTransferFunction8Bit + GetByteArray
 * (IO/TransferFunction2D.cpp:222-241, :294-361) without Qt: the 1D TF (n1d RGBA8
 * entries, may be 0) stretched over a canvas of GetRenderSize(), every swatch -- a
 * polygon of n_points normalised (x, y) points filled with the odd-even rule, coloured
 * by a linear gradient from (gradient[0], gradient[1]) to (gradient[2], gradient[3]) or
 * a radial one centred at the first point, stops at positions 0..1 -- painted source-over
 * in order, then the canvas scaled to w x h (nearest).  out: w*h RGBA8, row 0 first.
 * Host code, no device needed.  Qt's own raster rules (pixel coverage, gradient
 * rounding) are not reproduced bit for bit: the rules used are stated in
 * csrc/host/tf2d.cpp. */
typedef struct {
  float pos;             /* GradientStop::first                                   */
  float rgba[4];         /* GradientStop::second                                  */
} tvk_gradient_stop;
typedef struct {
  int radial;            /* TFPolygon::bRadial                                     */
  uint32_t first_point, n_points;   /* into points[] (pairs)                      */
  float gradient[4];     /* TFPolygon::pGradientCoords[0].xy, [1].xy (normalised)  */
  uint32_t first_stop, n_stops;     /* into stops[]                               */
} tvk_swatch;
int tvk_rasterize_tf2d(int w, int h, const uint8_t* tf1d_rgba8, int n1d, const tvk_swatch* swatches, int nswatch,
                       const float* points, size_t npoints, const tvk_gradient_stop* stops, size_t nstops,
                       uint8_t* out);
/* Changed2DTrans with swatches: rasterizes as above and sets the table (tvk_set_tf2d) */
int tvk_set_tf2d_swatches(tvk_ctx* ctx, int w, int h, const uint8_t* tf1d_rgba8, int n1d, const tvk_swatch* swatches,
                          int nswatch, const float* points, size_t npoints, const tvk_gradient_stop* stops,
                          size_t nstops);

/* ---- view ------------------------------------------------------------------ */
/* Row-vector matrices in Basics/Vectors.h layout (array[0]=m11, array[12..14] = translation);
 * modelView = rot * trans * view (GLRenderer.cpp:627). */
int tvk_set_view_matrices(tvk_ctx* ctx, const float view[16], const float rot[16],
                          const float trans[16], const float proj[16], uint32_t width,
                          uint32_t height, float fov, float znear, float zfar);
/* BuildLookAt + Perspective (Vectors.h:1250-1277); rotation/translation = identity */
int tvk_set_camera(tvk_ctx* ctx, const float eye[3], const float at[3], const float up[3], float fov,
                   float znear, float zfar, uint32_t width, uint32_t height);
/* RenderRegion rotation (e.g. an orbit); translation stays as set */
int tvk_set_rotation(tvk_ctx* ctx, const float rot[16]);
int tvk_get_view_matrices(tvk_ctx* ctx, float view[16], float proj[16]);
/* AbstrRenderer::SetClipPlane / EnableClipPlane / DisableClipPlane
 * (Renderer/AbstrRenderer.h:215-219): plane = ExtendedPlane::Plane() in world space
 * (x, y, z, w; NULL keeps the current one), enabled = m_bClipPlaneOn.  Bricks whose
 * eight corners are all clipped ((n.p) + w >= 0 after rotation * translation) leave
 * the visible-brick list (AbstrRenderer::Clipped, AbstrRenderer.cpp:923-948), and
 * every brick's ray segment keeps the part with n.p + d <= 0 of the plane carried
 * to model space (GLRaycaster::RenderBox + Clipper::BoxPlane, GLRaycaster.cpp:328-336).
 * The ray-guided traversal has no clip plane (as GLGridLeaper). */
int tvk_set_clip_plane(tvk_ctx* ctx, const float plane[4], int enabled);

/* ---- parameters, planning, rendering -------------------------------------- */
int tvk_set_params(tvk_ctx* ctx, const tvk_params* p);
int tvk_get_params(tvk_ctx* ctx, tvk_params* p);
int tvk_current_lod(tvk_ctx* ctx, int* lod);
int tvk_plan(tvk_ctx* ctx, tvk_brick_rec* out, int cap, int* n);
/* Render one frame.  host_out (may be NULL) receives W*H RGBA32F, premultiplied,
 * row 0 = bottom (GL window order). */
int tvk_render(tvk_ctx* ctx, float* host_out, tvk_stats* stats);
/* Render n views (rotation matrices rots[n*16]) of the current dataset, TF and
 * camera in one kernel launch on the context stream; host_out (may be NULL)
 * receives n frames.  With host_out and stats both NULL (also in tvk_render) the
 * call returns once the frame is queued on the context stream, so the host plans
 * the next frame while this one runs; tvk_synchronize waits for it. */
int tvk_render_views(tvk_ctx* ctx, int n, const float* rots, float* host_out, tvk_stats* stats);
/* Enable/disable the per-frame sample counter and touched-brick record (off by
 * default: costs an atomic per wave). */
int tvk_set_count_samples(tvk_ctx* ctx, int on);
/* Stored bytes (and count) of the bricks in which the last counted render took at
 * least one sample: summed over its views, and over the union of its views (each
 * brick once). */
int tvk_touched_bytes(tvk_ctx* ctx, uint64_t* bytes, uint64_t* bricks, uint64_t* unique_bytes,
                      uint64_t* unique_bricks);
/* Per TOC brick: 1 where the last counted render (any of its views) took a
 * sample in the brick (V_eff, the bricks a batch's requests deduplicate). */
int tvk_touched_mask(tvk_ctx* ctx, uint8_t* mask, size_t n);
/* Allocation-free variant of tvk_render_views.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory while they
 * render (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc, hipFree or allocating routine: the
 * planned brick list, view constants and sample counter go to one static
 * 8 MiB device array per GPU, and the frames (n x W x H RGBA32F, premultiplied,
 * GL row order) to d_out, device memory of out_bytes bytes that the caller
 * owns.  rots may be NULL when n == 1 (the context's rotation).  Limits: 281
 * views and 52,019 visible non-empty bricks summed over the views; beyond
 * them, or with a too small d_out, TVK_ERR_ARG.  Fixed step only (a context
 * with a step proposal attached gets TVK_ERR_UNSUPPORTED); the dataset, TF and
 * view must have been set up by the allocating calls beforehand.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU race.
 * tvk_display and tvk_frame_device_ptr keep referring to the frames of
 * tvk_render / tvk_render_views. */
int tvk_render_views_noalloc(tvk_ctx* ctx, int n, const float* rots, float* d_out, size_t out_bytes,
                             tvk_stats* stats);
/* Device pointer of the last rendered frame(s) (for zero-copy consumers). */
int tvk_frame_device_ptr(tvk_ctx* ctx, void** dptr);
/* The last rendered frame(s) as the window shows them: drawn over the
 * background (a vertical ramp from bg_top to bg_bottom; equal colours: a flat
 * clear) with (SRC_ALPHA, ONE_MINUS_SRC_ALPHA) into 8 bits per channel, GL row
 * order.  Replaces GLRenderer::CopyImageToDisplayBuffer + ClearColorBuffer /
 * DrawBackGradient (GLRenderer.cpp:489-510, :1539-1579, :1640-1680).
 * out: views x H x W x 4 bytes (RGBA). */
int tvk_display(tvk_ctx* ctx, const float bg_top[3], const float bg_bottom[3], uint8_t* out, size_t bytes);

/* ---- timing on the context stream ------------------------------------------- */
int tvk_event_record(tvk_ctx* ctx, int slot);          /* slot in [0, 6); 6 and 7 bracket the last launch */
int tvk_event_elapsed(tvk_ctx* ctx, int a, int b, float* ms);
int tvk_synchronize(tvk_ctx* ctx);

/* ---- batches ---------------------------------------------------------------
 * Many render requests -- each one the dataset, TF, parameters and camera of a
 * tvk_ctx with its own view rotation and image size -- rendered by persistent
 * kernels that pull ray tiles from device work queues (one launch per kernel
 * variant: render mode x lighting x voxel type), replayed as one hipGraph per
 * frame.  No reference counterpart as one call: AbstrRenderer::Paint renders
 * its RenderRegions one after the other (GLRenderer.cpp:571-667); each request's
 * frame equals tvk_render of its context at that rotation and size.  A request
 * keeps a pointer to its context: the context must outlive the batch's use of
 * it, and is read (not modified) by tvk_batch_render. */
typedef struct tvk_batch tvk_batch;
typedef struct {
  int tile_w, tile_h;    /* pixels per wave tile; tile_w * tile_h == 64          */
  int groups_per_cu;     /* persistent workgroups per CU (0: occupancy limit)    */
  int use_graph;         /* replay the frame as a captured hipGraph              */
} tvk_tuning;
int tvk_default_tuning(tvk_tuning* t);   /* 8x8 tiles, occupancy-sized grid, graph on */
int tvk_batch_create(tvk_batch** out, int device);
void tvk_batch_destroy(tvk_batch* b);
const char* tvk_batch_last_error(const tvk_batch* b);
int tvk_batch_clear(tvk_batch* b);
/* rot: 4x4 rotation (NULL: the context's); w, h: image size (the context's
 * projection is re-made for this aspect when it came from tvk_set_camera) */
int tvk_batch_add(tvk_batch* b, tvk_ctx* ctx, const float rot[16], uint32_t w, uint32_t h, int* id);
int tvk_batch_set_tuning(tvk_batch* b, const tvk_tuning* t);
int tvk_batch_get_tuning(tvk_batch* b, tvk_tuning* t);
int tvk_batch_set_count_samples(tvk_batch* b, int on);
/* Plans every request, uploads the brick lists and request table, replays the
 * frame graph; stats->ms is the device time of the launches. */
int tvk_batch_render(tvk_batch* b, tvk_stats* stats);
/* RGBA32F premultiplied frame of request id from the last render */
int tvk_batch_frame(tvk_batch* b, int id, float* host_out);
int tvk_batch_frame_device_ptr(tvk_batch* b, int id, void** dptr);
/* Times every tile shape x grid size on the current requests (reps renders
 * each after a warm-up), keeps and returns the fastest. */
int tvk_batch_autotune(tvk_batch* b, int reps, tvk_tuning* best, float* best_ms);

/* ---- ray-guided traversal (GLGridLeaper) ----------------------------------------
 * Single-pass traversal of GLGridLeaper (GLGridLeaper.cpp, GLGridLeaper-blend.glsl,
 * GLVolumePool.cpp:484-656): each ray walks the brick grid in normalised volume
 * coordinates with a level of detail per brick from its depth, leaps over empty
 * and child-empty bricks, and falls back to the next coarser resident level
 * when a brick is not resident, reporting it (GLHashTable) and recording where
 * it left the requested resolution.  Needs every level in the pool (max_lods 0)
 * and a brick grid that halves evenly.  Modes: 1D and 2D TF, lit or not. */
typedef struct {
  uint64_t samples;      /* composited samples (when counting is on)                */
  uint64_t missing;      /* bricks reported missing by the pass                      */
  uint64_t unfinished;   /* rays that left the requested resolution (resume later)   */
  float ms;              /* device time of the pass                                   */
} tvk_leap_stats;
/* residency per TOC brick (NULL: every brick resident, or, for a budgeted pool,
 * only the coarsest brick); the coarsest level is always resident.  For a
 * budgeted pool the flagged bricks are paged in (at most the slots less one). */
int tvk_leap_set_resident(tvk_ctx* ctx, const uint8_t* flags, size_t n);
/* one pass: first = 1 starts every ray at the volume box, 0 resumes the unfinished
 * rays of the previous pass (GLGridLeaper::Continue3DDraw) */
int tvk_leap_render(tvk_ctx* ctx, int first, float* host_out, tvk_leap_stats* stats);
/* Missing bricks are requested through GLHashTable (Renderer/GL/GLHashTable.cpp:90-170):
 * a table of `size` entries (0: sized from the largest brick as
 * GLGridLeaper::InitHashTable does) holding the brick's 4D index + 1 at its value
 * modulo size or one of the next `rehash` entries; requests that find no free
 * entry are dropped for the pass.  Defaults 509 and 10 (MasterController.h:70-74).
 * Coarser missing bricks met while falling back are requested too (BS_RequestAll). */
int tvk_leap_set_request_table(tvk_ctx* ctx, uint32_t size, uint32_t rehash);
/* the last pass's table (GLHashTable::GetData before condensing); *size = entries */
int tvk_leap_requests(tvk_ctx* ctx, uint32_t* entries, size_t n, size_t* size);
/* bricks requested by the last pass (flags per TOC brick, may be NULL) */
int tvk_leap_missing(tvk_ctx* ctx, uint8_t* flags, size_t n, uint64_t* count);
/* Pages up to max_bricks requested bricks in, in table order
 * (GLVolumePool::UploadBricks, GLVolumePool.cpp:673-717, :955-959, :1720-1789).  With
 * every brick in HBM this marks them resident; with a budgeted pool it copies them
 * into the slots of the oldest uploads (at most all slots but the coarsest
 * brick's per call), and the bricks replaced become missing. */
int tvk_leap_page_in(tvk_ctx* ctx, uint64_t max_bricks, uint64_t* paged);
/* GLRenderer::Pick (GLRenderer.cpp:2856-2872): the eye-space isosurface hit of the
 * last tvk_render / tvk_render_views frame (view 0) at window position (x, y), y from
 * the top (read at row H - y).  TVK_ERR_STATE outside isosurface mode, TVK_ERR_RANGE
 * when the ray hit nothing ("No intersection."). */
int tvk_pick(tvk_ctx* ctx, uint32_t x, uint32_t y, float out[3]);
/* residency per TOC brick: in a slot (budgeted pool), or flagged resident */
int tvk_pool_resident(tvk_ctx* ctx, uint8_t* flags, size_t n);
/* resume state of the last pass: per pixel 8 floats (normalised xyz, eye z or 1000 when finished, RGBA) */
int tvk_leap_resume_state(tvk_ctx* ctx, float* host, size_t floats);

/* ---- step-proposal network ----------------------------------------------------
 * No reference counterpart (the reference marches at a fixed step).  A small
 * MLP, 24 features -> 64 -> 64 -> 2 with ReLU, BF16 weights and FP32
 * accumulation on the matrix cores, proposes per ray and brick segment a step
 * multiplier (output 0: log2(multiplier) / 3) and a skip flag (output 1: logit).
 * Parameters are an FP32 vector of tvk_stepnet_param_count() floats:
 * W1[24][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2] (row-major, in x out). */
typedef struct tvk_stepnet tvk_stepnet;
int tvk_stepnet_create(tvk_stepnet** out, int device, uint64_t seed);   /* He-uniform init */
void tvk_stepnet_destroy(tvk_stepnet* s);
const char* tvk_stepnet_last_error(const tvk_stepnet* s);
int tvk_stepnet_param_count(void);
int tvk_stepnet_set_params(tvk_stepnet* s, const float* theta, int n);  /* also resets Adam */
int tvk_stepnet_get_params(tvk_stepnet* s, float* theta, int n);
/* outputs (n x 2) of n feature rows (n x 24) */
int tvk_stepnet_infer(tvk_stepnet* s, const float* feat, int n, float* out);
/* training samples: features n x 24, targets n x 3 (t0 = log2(multiplier)/3, t1 = skip in {0,1}, weight) */
int tvk_stepnet_set_data(tvk_stepnet* s, const float* feat, const float* tgt, int n);
/* gradient of mean weighted loss over the first n samples; grad[count] = summed weighted loss */
int tvk_stepnet_grad(tvk_stepnet* s, int n, float* grad);
/* Adam over the loaded samples; losses[epochs] (may be NULL) gets each epoch's mean loss */
int tvk_stepnet_train(tvk_stepnet* s, int n, int epochs, int batch, float lr, float weight_decay, uint64_t seed,
                      float* losses);

/* Learned step on a context: with a network attached, tvk_render and
 * tvk_render_views (1D/2D TF modes) march every ray segment through a brick
 * in chunks of 32 fixed steps; per chunk the network proposes a step
 * multiplier m (step m * fRayStepsize, opacity correction exponent
 * m * fStepScale; a wavefront takes the smallest m of its rays) or a skip.
 * The contribution bound of a chunk is T_before * (1-(1-alpha_max)^n), with
 * alpha_max the largest TF alpha over the chunk's value range (from the
 * dataset's cell grid) and n its fixed-step count.  NULL detaches (fixed step). */
typedef struct {
  float margin;          /* multiplier = 2^clamp(floor(3*y0 + 0.5 - margin), 0, max_log2) */
  float skip_threshold;  /* skip when sigmoid(y1) > threshold (>= 1: never skip)          */
  int max_log2;          /* 0..3                                                          */
  float fallback_bound;  /* fixed step where the contribution bound exceeds this (<= 0: never) */
  float skip_bound;      /* skip, without the network, chunks whose bound is <= this (< 0: never;
                            0 skips the chunks whose value range is fully transparent, which
                            leaves the frame unchanged) */
  int fixed_log2;        /* >= 0: no network (net may be NULL), every chunk that is not skipped
                            takes multiplier 2^fixed_log2 (the chunked march without the
                            network, for comparison); < 0: the network decides */
  float fallback_tv;     /* fixed step where the chunk's variation risk exceeds this (<= 0: never):
                            T_before * (bound * TVc + min(1, n * fStepScale * TVa)), TVc and TVa
                            the total variation of the TF's colour and alpha over the chunk's
                            value range -- a homogeneous chunk keeps its proposal however
                            opaque it is; a forced chunk holds its wavefront to multiplier 1 */
} tvk_proposal_params;
int tvk_set_step_proposal(tvk_ctx* ctx, tvk_stepnet* net, const tvk_proposal_params* p);
/* The dataset's cell grid (built on first use): per TOC brick, cells of 8^3
 * stored voxels, x fastest, ceil(n/8) per axis; each word is (max << 16) | min
 * over the cell's stored voxels grown by 4 on every side (clipped to the
 * brick).  Writes up to cap words and the total count (*n). */
int tvk_cell_grid(tvk_ctx* ctx, uint32_t* out, uint64_t cap, uint64_t* n);
/* Training samples from the fixed-step march: renders nviews views (rots[n*16])
 * and records a hash-selected fraction `keep` of the chunks (at most cap) that
 * are not fully transparent, each as 48 floats: 24 features, transmittance
 * before the chunk, fixed-step sample count, the chunk's premultiplied RGBA at
 * multipliers 1, 2, 4, 8, 6 pad.  *n: records written; *total (may be NULL):
 * chunks selected by the hash, written or not. */
int tvk_generate_segments(tvk_ctx* ctx, int nviews, const float* rots, uint32_t cap, float keep, uint32_t seed,
                          float* out, uint32_t* n, uint32_t* total);
/* The same proposal for every 1D/2D-TF request of a batch (NULL: fixed step;
 * fixed_log2 must be < 0). */
int tvk_batch_set_step_proposal(tvk_batch* b, tvk_stepnet* net, const tvk_proposal_params* p);

#ifdef __cplusplus
}
#endif
#endif /* TVK_HIP_H */
