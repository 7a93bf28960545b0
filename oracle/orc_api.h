/* Oracle C ABI (test infrastructure only; see orc_common.hpp).
 *
 * A CPU restatement of the reference's volume-raycasting path:
 *   - ExtendedOctree bricking / LOD pyramid / per-brick min-max / brick table
 *   - TransferFunction1D/2D byte tables and non-zero limits
 *   - CullingLOD + AbstrRenderer::BuildSubFrameBrickList (visible-brick list)
 *   - GLRaycaster per-brick march (1D, 1D-lit, 2D, 2D-lit, ISO + compose),
 *     in a GL-faithful FP32 mode, plus an FP64 emission-absorption integrator.
 */
#ifndef ORC_API_H
#define ORC_API_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- a1: ExtendedOctree metadata -------------------------------------- */
/* ExtendedOctree.cpp:188-243.  Returns the LOD count (<= maxl) */
int orc_octree_meta(const uint64_t dims[3], const uint32_t brick[3], uint32_t overlap,
                    uint64_t* lod_size, uint64_t* lod_bricks, uint64_t* lod_offset,
                    double* lod_aspect, int maxl);

/* ---- a2..a6: full conversion into an in-memory dataset ----------------- */
typedef struct orc_ds orc_ds;
/* bytes_per_voxel: 1 (uint8) or 2 (uint16).  brick = stored brick size incl.
 * overlap (ExtendedOctree m_iBrickSize); clamp = bClampToEdge; median =
 * bComputeMedian; scale = dataset voxel aspect (UVF domain scale).          */
orc_ds* orc_ds_create(const void* voxels, int bytes_per_voxel, const uint64_t dims[3],
                      const uint32_t brick[3], uint32_t overlap, int clamp, int median,
                      const float scale[3]);
/* Same, but only LOD 0 is converted (the table still lists every level); with
 * voxels NULL nothing is converted (metadata and brick table only). */
orc_ds* orc_ds_create_lod0(const void* voxels, int bytes_per_voxel, const uint64_t dims[3],
                           const uint32_t brick[3], uint32_t overlap, int clamp, const float scale[3]);
void     orc_ds_destroy(orc_ds*);
int      orc_ds_lod_count(const orc_ds*);
uint64_t orc_ds_brick_count(const orc_ds*);               /* all LODs, TOC order */
void     orc_ds_brick_info(const orc_ds*, uint64_t toc_index, uint64_t coords[4],
                           uint32_t size[3], uint64_t* byte_offset, double minmax[2]);
const void* orc_ds_brick_data(const orc_ds*, uint64_t toc_index);
uint64_t orc_ds_total_bytes(const orc_ds*);
const void* orc_ds_all_data(const orc_ds*);                /* all bricks, TOC order */
/* UVFDataset::ComputeMetadataTOC (uvfDataset.cpp:242-288) */
void     orc_ds_brick_md(const orc_ds*, int lod, uint64_t lin_index, float center[3],
                         float extents[3], uint32_t n_voxels[3]);
/* UVFDataset::GetTextCoords (uvfDataset.cpp:1747-1774) */
void     orc_ds_texcoords(const orc_ds*, int lod, uint64_t lin_index, int pow2,
                          float tmin[3], float tmax[3]);
double   orc_ds_range_max(const orc_ds*);
/* iteration order of the brick table: (lod, 1D index) pairs; returns the brick count */
uint64_t orc_ds_table_order(const orc_ds*, uint64_t* out, uint64_t cap);
/* VolumeTools::Filter (VolumeTools.h:168-262) over n groups of taps (2, 4, 8) values */
void orc_filter(int taps, int median, int bpv, const void* in, uint64_t n, void* out);

/* ---- a10 / a7: transfer functions -------------------------------------- */
/* TransferFunction1D::GetByteArray (TransferFunction1D.cpp:311-330) */
void orc_tf1d_bytes(const float* rgba, int n, uint8_t* out, int used_range);
/* TransferFunction1D::Set(rgba8) + ComputeNonZeroLimits (:146-177, :362-371) */
void orc_tf1d_limits(const uint8_t* rgba8, int n, uint64_t lim[2]);
/* TransferFunction2D::ComputeNonZeroLimits (TransferFunction2D.cpp:378-397) */
void orc_tf2d_limits(const uint8_t* rgba8, int w, int h, uint64_t lim[4]);

/* ---- a8 / a9 / a12: frame planning ------------------------------------- */
typedef struct {
  float view[16];      /* m_mView (BuildLookAt), row-vector convention        */
  float rot[16];       /* RenderRegion rotation                                */
  float trans[16];     /* RenderRegion translation                             */
  float proj[16];      /* m_mProjection (Perspective)                          */
  uint32_t width, height;
  float fov, znear, zfar;
} orc_view;

typedef struct {
  int   mode;           /* 0 = 1D TF, 1 = 2D TF, 2 = isosurface               */
  int   lighting;       /* m_bUseLighting                                      */
  float srm;            /* m_fSampleRateModifier                               */
  float isovalue;       /* m_fIsovalue (data units)                            */
  int   lod;            /* m_iCurrentLOD                                       */
  int   bit_width;      /* dataset bit width (8 or 16)                         */
  const uint8_t* tf1d; int tf1d_n;          /* RGBA8, tf1d_n entries         */
  const uint8_t* tf2d; int tf2d_w, tf2d_h;  /* RGBA8, row-major, x = value   */
  float tf_scale;       /* fTransScale   (GLRenderer.cpp:1813-1819)            */
  float grad_scale;     /* fGradientScale (GLRenderer.cpp:1840-1841)           */
  float domain_scale[3];/* vDomainScale = 1/GetScale()                         */
  float light_ambient[3], light_diffuse[3], light_specular[3], light_dir[3];
  float iso_color[3];
  int   refine_steps;   /* RefineIsosurface.glsl loop count (reference: 5)     */
  int   clamp_to_edge;  /* volume wrap: 0 = GL_CLAMP (border 0), 1 = CLAMP_TO_EDGE */
  float ert_eps;        /* 0 = reference semantics; >0: stop a ray once its
                           total transmittance drops below ert_eps           */
  int   clip_on;        /* m_bClipPlaneOn                                      */
  float clip[4];        /* m_ClipPlane.Plane() (world space: x, y, z, w)      */
  int   pow2;           /* bUseOnlyPowerOfTwo: textures padded to powers of two */
  int   down8;          /* bDownSampleTo8Bits: 16-bit bricks quantized to 8 bits */
} orc_params;

typedef struct {
  uint64_t key_index;   /* 1D brick index within the LOD                       */
  uint64_t toc_index;   /* index into the TOC / brick pool                     */
  float center[3], extension[3];
  uint32_t voxel_count[3];
  float tex_min[3], tex_max[3];
  float distance;
  int   is_empty;
} orc_brick;

/* PlanFrame + BuildSubFrameBrickList (AbstrRenderer.cpp:999-1100, :1125-1212).
 * Returns the number of bricks in the list (or -needed if cap is too small). */
int orc_plan(const orc_ds*, const orc_view*, const orc_params*, orc_brick* out, int cap);
/* ComputeMinLODForCurrentView (AbstrRenderer.cpp:789-803; CullingLOD.cpp:126-138) */
int orc_min_lod(const orc_ds*, const orc_view*, float screen_space_error);
/* CullingLOD::Update's six planes (x, y, z, w each; CullingLOD.cpp:89-124), IsVisible
 * (:141-160) and GetLODLevel (:126-138, unclamped) for the view's modelview */
void orc_cull_planes(const orc_view*, float planes[24]);
int orc_cull_visible(const orc_view*, const float center[3], const float extent[3]);
int orc_cull_lod_level(const orc_view*, float screen_space_error, const float center[3], const float extent[3],
                       const uint32_t voxels[3]);

/* ---- a13 / a14: the march ---------------------------------------------- */
/* Renders pixels [x0,x1) x [y0,y1) of the frame into out_rgba (RGBA32F,
 * premultiplied, row-major over the ROI, row 0 = window row y0 (GL bottom-up)).
 * mode/lighting select GLRaycaster-{1D,1D-light,2D,2D-light,ISO}-FS.          */
int orc_render(const orc_ds*, const orc_view*, const orc_params*, const int roi[4],
               float* out_rgba, uint64_t* n_samples, int n_threads);
/* FP64 emission-absorption integral of the same ray model (no fp16 entry, no
 * per-brick restart, continuous trilinear field, no early termination).      */
int orc_render_fp64(const orc_ds*, const orc_view*, const orc_params*, const int roi[4],
                    double step_model, float* out_rgba, int n_threads);
/* the opacity correction's pow: out[i] = pow(y[i], s[i]) as the march evaluates it */
void orc_opacity_pow(const float* y, const float* s, int n, float* out);

/* camera matrices (Vectors.h:1250-1277), row-vector layout */
void orc_look_at(const float eye[3], const float at[3], const float up[3], float out[16]);
void orc_perspective(float fovy, float aspect, float znear, float zfar, float out[16]);
void orc_rotation_y(double angle, float out[16]);

/* fp16 helpers (exposed for tests) */
float orc_round_fp16(float f);

/* GLRaycaster::SetBrickDepShaderVars (GLRaycaster.cpp:239-300) */
void orc_brick_shader_vars(const orc_ds*, const orc_params*, const orc_brick*,
                           float* ray_step, float* step_scale, float voxel_step[3]);

/* ---- step-proposal network (no reference counterpart) -------------------- */
/* FP64 forward/backward of the 24-64-64-2 MLP; bf16 = 1 rounds to bfloat16 at
 * the points the HIP kernels do.  theta: orc_sn_param_count() floats. */
int orc_sn_param_count(void);
void orc_sn_forward(const float* theta, const float* x, int n, int bf16, double* y);
void orc_sn_grad(const float* theta, const float* x, const float* tgt, int n, int bf16, double* grad);

/* ---- a15: GridLeaper single-pass traversal -------------------------------- */
/* One pass of GLGridLeaper-blend.glsl over the whole frame.  eye: m_vEye (the
 * renderer's camera position).  resident: per TOC brick (NULL: all); the
 * coarsest level is always resident.  first: rays start at the volume box
 * (else at resume_in).  out: W*H RGBA (accRayColor); resume_out: W*H*8
 * (rayResumePos xyzw, rayResumeColor; w = 1000 marks a finished ray);
 * missing: per TOC brick, 1 where a brick was reported missing.  Returns -1 if
 * the pool's brick grid is not the octree's (needs power-of-two-friendly sizes). */
int orc_leap_pass(const orc_ds* ds, const orc_view* v, const orc_params* p, const float eye[3],
                  const uint8_t* resident, int first, const float* resume_in, float* out, float* resume_out,
                  uint8_t* missing, uint64_t* samples);

#ifdef __cplusplus
}
#endif
#endif
