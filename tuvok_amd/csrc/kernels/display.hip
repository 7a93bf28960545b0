// Display step of a rendered frame: GLRenderer::CopyImageToDisplayBuffer
// (GLRenderer.cpp:1539-1579) draws the premultiplied RGBA32F image over the
// cleared window with glBlendFunc(GL_SRC_ALPHA, GL_ONE_MINUS_SRC_ALPHA) through
// Transfer-FS.glsl (a plain copy), after ClearColorBuffer (:489-510) filled the
// window with the background: one colour, or DrawBackGradient's (:1640-1680)
// vertical ramp from the top colour to the bottom colour (alpha 0).  The window
// stores 8 bits per channel.  The premultiplied colour is multiplied by alpha
// once more, as the blend function does.
//
// Compiled with -ffp-contract=off: the float operations are the ones written.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>

namespace tvk {

__device__ __forceinline__ uint32_t unorm8(float x) {
  x = fminf(fmaxf(x, 0.0f), 1.0f);
  return (uint32_t)floorf(x * 255.0f + 0.5f);
}

// frames: nviews planes of W x H RGBA32F in GL row order (row 0 = bottom)
__global__ __launch_bounds__(256) void display_over_background(const float4* __restrict__ frames, uint32_t W,
                                                               uint32_t H, uint32_t nviews, float3 top, float3 bottom,
                                                               uint32_t* __restrict__ out) {
  const uint64_t n = (uint64_t)W * H * nviews;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint32_t row = (uint32_t)((i / W) % H);
    // the ramp at the pixel centre: bottom colour at NDC y = -1, top colour at +1
    const float t = ((float)row + 0.5f) / (float)H;
    const float bx = bottom.x + t * (top.x - bottom.x), by = bottom.y + t * (top.y - bottom.y),
                bz = bottom.z + t * (top.z - bottom.z);
    const float4 s = frames[i];
    const float k = 1.0f - s.w;
    const uint32_t r = unorm8(s.x * s.w + bx * k), g = unorm8(s.y * s.w + by * k), b = unorm8(s.z * s.w + bz * k);
    const uint32_t a = unorm8(s.w * s.w);   // background alpha 0
    out[i] = r | (g << 8) | (b << 16) | (a << 24);
  }
}

hipError_t launch_display(const float* frames, uint32_t W, uint32_t H, uint32_t nviews, const float top[3],
                          const float bottom[3], uint8_t* out, hipStream_t st) {
  const uint64_t n = (uint64_t)W * H * nviews;
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(display_over_background, dim3(blocks), dim3(256), 0, st, reinterpret_cast<const float4*>(frames),
                     W, H, nviews, make_float3(top[0], top[1], top[2]), make_float3(bottom[0], bottom[1], bottom[2]),
                     reinterpret_cast<uint32_t*>(out));
  return hipGetLastError();
}

}  // namespace tvk
