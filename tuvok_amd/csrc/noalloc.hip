// Allocation-free rendering (tvk_render_views_noalloc): the frame's planned
// brick list, view constants and sample counter live in one static device
// array instead of buffers the library allocates, and the caller passes the
// output frames in device memory.  For callers that cannot allocate device
// memory while they render (HIP graph capture, tight serving loops).
//
// Every GPU holds its own copy of the array, at its own address, until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: the name is a host-side stand-in, not a device address.  The address
// is looked up once per device (scratch_address, called by tvk_create with the
// device current) and every use goes through it.  All calls on one GPU share
// that GPU's array: calls that overlap on one GPU (two contexts on two streams)
// race.
#include <mutex>
#include "ctx.hpp"

namespace {

// Layout of the scratch array: the 64-bit sample counter, the view constants
// from byte 256, the visible-brick list from byte 64 KiB.
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr size_t kViewsAt = 256;
constexpr size_t kListAt = size_t(64) << 10;
constexpr size_t kMaxViews = (kListAt - kViewsAt) / sizeof(tvk::ViewConst);
constexpr size_t kMaxList = (kScratchBytes - kListAt) / sizeof(tvk::FrameBrick);
static_assert(kListAt % alignof(tvk::FrameBrick) == 0 && kViewsAt % alignof(tvk::ViewConst) == 0, "scratch layout");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
std::vector<void*> g_addr;   // per device: that device's address of `scratch`

}  // namespace

namespace tvk {

hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if ((size_t)device >= g_addr.size()) g_addr.resize((size_t)device + 1, nullptr);
  if (!g_addr[(size_t)device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_addr[(size_t)device] = p;
  }
  *out = g_addr[(size_t)device];
  return hipSuccess;
}

}  // namespace tvk

extern "C" int tvk_render_views_noalloc(tvk_ctx* c, int n, const float* rots, float* d_out, size_t out_bytes,
                                        tvk_stats* st) {
  using namespace tvk;
  if (!c || !d_out) return TVK_ERR_ARG;
  if (!c->scratch) return fail(c, TVK_ERR_STATE, "allocation-free render unavailable: " + c->scratch_err);
  if (n < 1) return fail(c, TVK_ERR_ARG, "nviews must be >= 1");
  if (!rots && n != 1) return fail(c, TVK_ERR_ARG, "several views need their rotations");
  unsigned char* base = (unsigned char*)c->scratch;
  FrameBuffers fx;
  fx.counter = (unsigned long long*)base;
  fx.views = (ViewConst*)(base + kViewsAt);
  fx.views_cap = kMaxViews;
  fx.list = (FrameBrick*)(base + kListAt);
  fx.list_cap = kMaxList;
  fx.out = d_out;
  fx.out_floats = out_bytes / sizeof(float);
  return render_impl(c, n, rots, nullptr, st, -1, nullptr, &fx);
}
