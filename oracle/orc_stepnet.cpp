// Oracle of the step-proposal network (test infrastructure only).
//
// The network has no counterpart in the reference; this is the FP64 statement
// of what the HIP kernels compute: 24 -> 64 -> 64 -> 2, ReLU hidden layers,
// loss w * ((y0 - t0)^2 / 2 + BCE(sigmoid(y1), t1)), mean over the batch.
// With bf16 = 1 the values the kernels round to bfloat16 (inputs, weights,
// hidden activations, output and hidden gradients) are rounded here at the
// same points (round to nearest even); with bf16 = 0 everything is exact FP64,
// which is what the finite-difference test of the gradient uses.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "orc_api.h"

namespace {

constexpr int NI = 24, NH = 64, NO = 2;
constexpr int OW1 = 0, OB1 = OW1 + NI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH * NO;
constexpr int NP = OB3 + NO;

// float -> bfloat16 -> float, round to nearest even (finite values)
double rbf(double v) {
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return (double)f;
}

struct Net {
  std::vector<double> w1, b1, w2, b2, w3, b3;
  Net(const float* th, bool bf) {
    auto q = [&](double v) { return bf ? rbf(v) : v; };
    w1.resize(NI * NH); w2.resize(NH * NH); w3.resize(NH * NO);
    for (int i = 0; i < NI * NH; ++i) w1[i] = q(th[OW1 + i]);
    for (int i = 0; i < NH * NH; ++i) w2[i] = q(th[OW2 + i]);
    for (int i = 0; i < NH * NO; ++i) w3[i] = q(th[OW3 + i]);
    b1.assign(th + OB1, th + OB1 + NH);
    b2.assign(th + OB2, th + OB2 + NH);
    b3.assign(th + OB3, th + OB3 + NO);
  }
};

struct Fwd {
  double x[NI], h1[NH], h2[NH], y[NO];
};

void forward(const Net& N, const float* xin, bool bf, Fwd& F) {
  auto q = [&](double v) { return bf ? rbf(v) : v; };
  for (int f = 0; f < NI; ++f) F.x[f] = q(xin[f]);
  for (int o = 0; o < NH; ++o) {
    double z = 0;
    for (int f = 0; f < NI; ++f) z += F.x[f] * N.w1[f * NH + o];
    F.h1[o] = q(std::max(z + N.b1[o], 0.0));
  }
  for (int o = 0; o < NH; ++o) {
    double z = 0;
    for (int i = 0; i < NH; ++i) z += F.h1[i] * N.w2[i * NH + o];
    F.h2[o] = q(std::max(z + N.b2[o], 0.0));
  }
  for (int o = 0; o < NO; ++o) {
    double z = 0;
    for (int i = 0; i < NH; ++i) z += F.h2[i] * N.w3[i * NO + o];
    F.y[o] = z + N.b3[o];
  }
}

}  // namespace

extern "C" {

int orc_sn_param_count(void) { return NP; }

void orc_sn_forward(const float* theta, const float* x, int n, int bf16, double* y) {
  const Net N(theta, bf16 != 0);
  Fwd F;
  for (int s = 0; s < n; ++s) {
    forward(N, x + (size_t)s * NI, bf16 != 0, F);
    y[2 * (size_t)s] = F.y[0];
    y[2 * (size_t)s + 1] = F.y[1];
  }
}

// grad[NP] = d(mean weighted loss)/d theta, grad[NP] = summed weighted loss
void orc_sn_grad(const float* theta, const float* x, const float* tgt, int n, int bf16, double* grad) {
  const bool bf = bf16 != 0;
  auto q = [&](double v) { return bf ? rbf(v) : v; };
  const Net N(theta, bf);
  std::memset(grad, 0, (NP + 1) * sizeof(double));
  Fwd F;
  for (int s = 0; s < n; ++s) {
    forward(N, x + (size_t)s * NI, bf, F);
    const double t0 = tgt[3 * (size_t)s], t1 = tgt[3 * (size_t)s + 1], w = tgt[3 * (size_t)s + 2];
    const double e = F.y[0] - t0, y1 = F.y[1];
    const double sg = 1.0 / (1.0 + std::exp(-y1));
    grad[NP] += w * (0.5 * e * e + std::max(y1, 0.0) - y1 * t1 + std::log1p(std::exp(-std::fabs(y1))));
    const double dy[NO] = {q(w * e / n), q(w * (sg - t1) / n)};
    double dh2[NH], dh1[NH];
    for (int h = 0; h < NH; ++h)
      dh2[h] = F.h2[h] != 0 ? q(dy[0] * N.w3[h * NO] + dy[1] * N.w3[h * NO + 1]) : 0.0;
    for (int i = 0; i < NH; ++i) {
      double g = 0;
      for (int o = 0; o < NH; ++o) g += N.w2[i * NH + o] * dh2[o];
      dh1[i] = F.h1[i] != 0 ? q(g) : 0.0;
    }
    for (int h = 0; h < NH; ++h)
      for (int o = 0; o < NO; ++o) grad[OW3 + h * NO + o] += F.h2[h] * dy[o];
    for (int o = 0; o < NO; ++o) grad[OB3 + o] += dy[o];
    for (int i = 0; i < NH; ++i)
      for (int o = 0; o < NH; ++o) grad[OW2 + i * NH + o] += F.h1[i] * dh2[o];
    for (int o = 0; o < NH; ++o) grad[OB2 + o] += dh2[o];
    for (int f = 0; f < NI; ++f)
      for (int o = 0; o < NH; ++o) grad[OW1 + f * NH + o] += F.x[f] * dh1[o];
    for (int o = 0; o < NH; ++o) grad[OB1 + o] += dh1[o];
  }
}

}  // extern "C"
