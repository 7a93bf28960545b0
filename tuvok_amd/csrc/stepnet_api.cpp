// C ABI of the step-proposal network (include/tvk_hip.h, "step network").
#include <random>
#include "ctx.hpp"
#include "kernels/stepnet.hpp"

using namespace tvk;

namespace tvk {
hipError_t launch_sn_pack(const float* th, uint16_t* w1t, uint16_t* w2t, uint16_t* w2, uint16_t* w3t, float* b1,
                          float* b2, float* b3, hipStream_t st);
hipError_t launch_sn_infer(const StepNetDev& N, const float* feat, int n, float* out, hipStream_t st);
hipError_t launch_sn_grad(const StepNetDev& N, const float* W3f, const float* feat, const float* tgt,
                          const uint32_t* idx, int n, float inv_batch, float* part, float* grad, hipStream_t st);
hipError_t launch_sn_adam(float* th, float* m, float* v, const float* grad, float lr, float b1, float b2, float eps,
                          int step, float wd, hipStream_t st);
}  // namespace tvk

struct tvk_stepnet {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float* d_theta = nullptr;   // fp32 master parameters
  float* d_m = nullptr;       // Adam moments
  float* d_v = nullptr;
  float* d_grad = nullptr;    // SN_NPARAM + 1 (loss)
  uint16_t* d_w = nullptr;    // packed bf16: w1t | w2t | w2 | w3t
  float* d_b = nullptr;       // b1 | b2 | b3
  float* d_part = nullptr;
  size_t part_cap = 0;
  // training data
  float* d_feat = nullptr;
  float* d_tgt = nullptr;
  uint32_t* d_idx = nullptr;
  size_t data_cap = 0;
  float* d_io = nullptr;      // inference in/out staging
  size_t io_cap = 0;
  int adam_step = 0;
  StepNetDev dev() const {
    StepNetDev N;
    N.w1t = d_w;
    N.w2t = d_w + SN_HID * SN_INP;
    N.w2 = N.w2t + SN_HID * SN_HID;
    N.w3t = N.w2 + SN_HID * SN_HID;
    N.b1 = d_b;
    N.b2 = d_b + SN_HID;
    N.b3 = d_b + 2 * SN_HID;
    return N;
  }
};

namespace tvk {
// the network as the march kernel reads it (nullptr when none is attached)
bool stepnet_device(const tvk_stepnet* s, StepNetDev* out) {
  if (!s) return false;
  *out = s->dev();
  return true;
}
}  // namespace tvk

namespace {

int sfail(tvk_stepnet* s, int code, const std::string& m) {
  if (s) s->err = m;
  return code;
}
#define SCHK(s, x)                                                                                     \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess)                                                                              \
      return sfail(s, e_ == hipErrorOutOfMemory ? TVK_ERR_NOMEM : TVK_ERR_HIP,                         \
                   std::string(#x) + ": " + hipGetErrorString(e_));                                    \
  } while (0)

template <typename T>
int sensure(tvk_stepnet* s, T** p, size_t* cap, size_t n) {
  if (*cap >= n && *p) return TVK_OK;
  if (*p) { (void)hipFree(*p); *p = nullptr; *cap = 0; }
  SCHK(s, hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
  *cap = std::max<size_t>(n, 1);
  return TVK_OK;
}

int repack(tvk_stepnet* s) {
  StepNetDev N = s->dev();
  SCHK(s, launch_sn_pack(s->d_theta, const_cast<uint16_t*>(N.w1t), const_cast<uint16_t*>(N.w2t),
                         const_cast<uint16_t*>(N.w2), const_cast<uint16_t*>(N.w3t), const_cast<float*>(N.b1),
                         const_cast<float*>(N.b2), const_cast<float*>(N.b3), s->stream));
  return TVK_OK;
}

// He-uniform initialisation of the weights, zero biases (seeded, on the host)
void init_params(std::vector<float>& th, uint64_t seed) {
  th.assign(SN_NPARAM, 0.0f);
  std::mt19937_64 g(seed);
  auto fill = [&](int off, int fan_in, int n) {
    const double a = std::sqrt(6.0 / fan_in);
    std::uniform_real_distribution<double> u(-a, a);
    for (int i = 0; i < n; ++i) th[off + i] = (float)u(g);
  };
  fill(SN_OFF_W1, SN_IN, SN_IN * SN_HID);
  fill(SN_OFF_W2, SN_HID, SN_HID * SN_HID);
  fill(SN_OFF_W3, SN_HID, SN_HID * SN_OUT);
}

}  // namespace

extern "C" {

int tvk_stepnet_create(tvk_stepnet** out, int device, uint64_t seed) {
  if (!out) return TVK_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TVK_ERR_HIP;
  if (device < 0 || device >= n) return TVK_ERR_ARG;
  tvk_stepnet* s = new (std::nothrow) tvk_stepnet;
  if (!s) return TVK_ERR_NOMEM;
  s->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_theta, SN_NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_m, SN_NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_v, SN_NPARAM * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_grad, (SN_NPARAM + 1) * sizeof(float));
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->d_w, (SN_HID * SN_INP + 2 * SN_HID * SN_HID + 32 * SN_HID) * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_b, (2 * SN_HID + SN_OUT) * sizeof(float));
  if (e != hipSuccess) {
    tvk_stepnet_destroy(s);
    return TVK_ERR_HIP;
  }
  std::vector<float> th;
  init_params(th, seed);
  int rc = tvk_stepnet_set_params(s, th.data(), SN_NPARAM);
  if (rc) {
    tvk_stepnet_destroy(s);
    return rc;
  }
  *out = s;
  return TVK_OK;
}

void tvk_stepnet_destroy(tvk_stepnet* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void* p : {(void*)s->d_theta, (void*)s->d_m, (void*)s->d_v, (void*)s->d_grad, (void*)s->d_w, (void*)s->d_b,
                  (void*)s->d_part, (void*)s->d_feat, (void*)s->d_tgt, (void*)s->d_idx, (void*)s->d_io})
    if (p) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

const char* tvk_stepnet_last_error(const tvk_stepnet* s) { return s ? s->err.c_str() : "null network"; }

int tvk_stepnet_param_count(void) { return SN_NPARAM; }

int tvk_stepnet_set_params(tvk_stepnet* s, const float* theta, int n) {
  if (!s || !theta) return TVK_ERR_ARG;
  if (n != SN_NPARAM) return sfail(s, TVK_ERR_ARG, "expected " + std::to_string(SN_NPARAM) + " parameters");
  SCHK(s, hipSetDevice(s->device));
  SCHK(s, hipMemcpyAsync(s->d_theta, theta, SN_NPARAM * sizeof(float), hipMemcpyHostToDevice, s->stream));
  SCHK(s, hipMemsetAsync(s->d_m, 0, SN_NPARAM * sizeof(float), s->stream));
  SCHK(s, hipMemsetAsync(s->d_v, 0, SN_NPARAM * sizeof(float), s->stream));
  s->adam_step = 0;
  int rc = repack(s);
  if (rc) return rc;
  SCHK(s, hipStreamSynchronize(s->stream));
  return TVK_OK;
}

int tvk_stepnet_get_params(tvk_stepnet* s, float* theta, int n) {
  if (!s || !theta) return TVK_ERR_ARG;
  if (n != SN_NPARAM) return sfail(s, TVK_ERR_ARG, "expected " + std::to_string(SN_NPARAM) + " parameters");
  SCHK(s, hipSetDevice(s->device));
  SCHK(s, hipMemcpyAsync(theta, s->d_theta, SN_NPARAM * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  SCHK(s, hipStreamSynchronize(s->stream));
  return TVK_OK;
}

int tvk_stepnet_infer(tvk_stepnet* s, const float* feat, int n, float* out) {
  if (!s || !feat || !out || n < 0) return TVK_ERR_ARG;
  if (n == 0) return TVK_OK;
  SCHK(s, hipSetDevice(s->device));
  int rc = sensure(s, &s->d_io, &s->io_cap, (size_t)n * (SN_IN + 2));
  if (rc) return rc;
  float* d_in = s->d_io;
  float* d_out = s->d_io + (size_t)n * SN_IN;
  SCHK(s, hipMemcpyAsync(d_in, feat, (size_t)n * SN_IN * sizeof(float), hipMemcpyHostToDevice, s->stream));
  SCHK(s, launch_sn_infer(s->dev(), d_in, n, d_out, s->stream));
  SCHK(s, hipMemcpyAsync(out, d_out, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  SCHK(s, hipStreamSynchronize(s->stream));
  return TVK_OK;
}

// loads n samples (features n x 24, targets n x 3: t0, t1, weight) for training
int tvk_stepnet_set_data(tvk_stepnet* s, const float* feat, const float* tgt, int n) {
  if (!s || !feat || !tgt || n <= 0) return TVK_ERR_ARG;
  SCHK(s, hipSetDevice(s->device));
  if (s->data_cap < (size_t)n) {
    for (void* p : {(void*)s->d_feat, (void*)s->d_tgt, (void*)s->d_idx}) if (p) (void)hipFree(p);
    s->d_feat = s->d_tgt = nullptr;
    s->d_idx = nullptr;
    s->data_cap = 0;
    SCHK(s, hipMalloc((void**)&s->d_feat, (size_t)n * SN_IN * sizeof(float)));
    SCHK(s, hipMalloc((void**)&s->d_tgt, (size_t)n * 3 * sizeof(float)));
    SCHK(s, hipMalloc((void**)&s->d_idx, (size_t)n * sizeof(uint32_t)));
    s->data_cap = (size_t)n;
  }
  SCHK(s, hipMemcpyAsync(s->d_feat, feat, (size_t)n * SN_IN * sizeof(float), hipMemcpyHostToDevice, s->stream));
  SCHK(s, hipMemcpyAsync(s->d_tgt, tgt, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, s->stream));
  SCHK(s, hipStreamSynchronize(s->stream));
  return TVK_OK;
}

// Gradient (and summed weighted loss in grad[SN_NPARAM]) of the first n loaded
// samples with inv_batch = 1/n: the FP32 result of the BF16 MFMA kernels.
int tvk_stepnet_grad(tvk_stepnet* s, int n, float* grad) {
  if (!s || !grad || n <= 0) return TVK_ERR_ARG;
  if ((size_t)n > s->data_cap) return sfail(s, TVK_ERR_ARG, "fewer samples loaded");
  SCHK(s, hipSetDevice(s->device));
  const size_t tiles = ((size_t)n + 63) / 64;
  int rc = sensure(s, &s->d_part, &s->part_cap, tiles * (SN_NPARAM + 1));
  if (rc) return rc;
  SCHK(s, launch_sn_grad(s->dev(), s->d_theta + SN_OFF_W3, s->d_feat, s->d_tgt, nullptr, n, 1.0f / (float)n,
                         s->d_part, s->d_grad, s->stream));
  SCHK(s, hipMemcpyAsync(grad, s->d_grad, (SN_NPARAM + 1) * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  SCHK(s, hipStreamSynchronize(s->stream));
  return TVK_OK;
}

// Adam over the loaded samples: `epochs` passes of shuffled minibatches.
// losses (may be NULL) receives the mean weighted loss of each epoch.
int tvk_stepnet_train(tvk_stepnet* s, int n, int epochs, int batch, float lr, float weight_decay, uint64_t seed,
                      float* losses) {
  if (!s || n <= 0 || epochs <= 0 || batch < 64) return TVK_ERR_ARG;
  if ((size_t)n > s->data_cap) return sfail(s, TVK_ERR_ARG, "fewer samples loaded");
  SCHK(s, hipSetDevice(s->device));
  const size_t tiles = ((size_t)batch + 63) / 64;
  int rc = sensure(s, &s->d_part, &s->part_cap, tiles * (SN_NPARAM + 1));
  if (rc) return rc;
  std::vector<uint32_t> perm((size_t)n);
  for (int i = 0; i < n; ++i) perm[(size_t)i] = (uint32_t)i;
  std::mt19937_64 g(seed);
  std::vector<float> gl(SN_NPARAM + 1);
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(perm.begin(), perm.end(), g);
    SCHK(s, hipMemcpyAsync(s->d_idx, perm.data(), perm.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream));
    double lsum = 0.0;
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int nb = std::min(batch, n - b0);
      SCHK(s, launch_sn_grad(s->dev(), s->d_theta + SN_OFF_W3, s->d_feat, s->d_tgt, s->d_idx + b0, nb,
                             1.0f / (float)nb, s->d_part, s->d_grad, s->stream));
      ++s->adam_step;
      SCHK(s, launch_sn_adam(s->d_theta, s->d_m, s->d_v, s->d_grad, lr, 0.9f, 0.999f, 1e-8f, s->adam_step,
                             weight_decay, s->stream));
      rc = repack(s);
      if (rc) return rc;
      if (losses) {
        float l = 0.0f;
        SCHK(s, hipMemcpyAsync(&l, s->d_grad + SN_NPARAM, sizeof(float), hipMemcpyDeviceToHost, s->stream));
        SCHK(s, hipStreamSynchronize(s->stream));
        lsum += l;
      }
    }
    if (losses) losses[ep] = (float)(lsum / n);
  }
  SCHK(s, hipStreamSynchronize(s->stream));
  return TVK_OK;
}

}  // extern "C"
