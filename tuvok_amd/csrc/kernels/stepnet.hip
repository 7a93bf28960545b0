// Training and stand-alone inference kernels of the step-proposal network
// (stepnet.hpp): BF16 MFMA products with FP32 accumulation.
//
// Training step on one 64-sample tile (one wave per workgroup):
//   z1 = W1^T x + b1, h1 = bf16(relu(z1)); z2 = W2^T h1 + b2, h2 = bf16(relu(z2));
//   y = W3^T h2 + b3;  loss = w * (y0 - t0)^2 / 2 + w * BCE(sigmoid(y1), t1)
//   dy = bf16(w * (y0 - t0) / B, w * (sigmoid(y1) - t1) / B)
//   dh2 = bf16((W3 dy) * [h2 > 0]);  dh1 = bf16((W2 dh2) * [h1 > 0])
//   dW3 = h2 dy^T, dW2 = h1 dh2^T, dW1 = x dh1^T (sums over the tile), db = sums of dy, dh2, dh1
// Each workgroup writes its partial gradient; a second kernel sums the partials
// in a fixed order (deterministic) and applies Adam to the FP32 master weights,
// then the bf16 copies are re-packed.
#include <hip/hip_runtime.h>
#include "stepnet.hpp"

namespace tvk {

// ---- packing -----------------------------------------------------------------
__global__ void sn_pack(const float* __restrict__ th, uint16_t* w1t, uint16_t* w2t, uint16_t* w2, uint16_t* w3t,
                        float* b1, float* b2, float* b3) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < SN_HID * SN_INP) {   // w1t[o][f]
    const int o = t / SN_INP, f = t % SN_INP;
    w1t[t] = f < SN_IN ? bf16_bits(th[SN_OFF_W1 + f * SN_HID + o]) : (uint16_t)0;
  }
  if (t < SN_HID * SN_HID) {
    const int a = t / SN_HID, b = t % SN_HID;
    w2t[t] = bf16_bits(th[SN_OFF_W2 + b * SN_HID + a]);   // [o=a][i=b]
    w2[t] = bf16_bits(th[SN_OFF_W2 + a * SN_HID + b]);    // [i=a][o=b]
  }
  if (t < 32 * SN_HID) {   // w3t[o][i]
    const int o = t / SN_HID, i = t % SN_HID;
    w3t[t] = o < SN_OUT ? bf16_bits(th[SN_OFF_W3 + i * SN_OUT + o]) : (uint16_t)0;
  }
  if (t < SN_HID) { b1[t] = th[SN_OFF_B1 + t]; b2[t] = th[SN_OFF_B2 + t]; }
  if (t < SN_OUT) b3[t] = th[SN_OFF_B3 + t];
}

// ---- inference over a feature array (tests, calibration) --------------------
__global__ __launch_bounds__(256) void sn_infer(StepNetDev N, const float* __restrict__ feat, int n,
                                                float* __restrict__ out) {
  __shared__ __align__(16) uint16_t lds[4 * SN_WAVE_LDS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int base = (blockIdx.x * 4 + wave) * 64;
  if (base >= n) return;   // uniform per wave; no workgroup barrier below
  const int s = min(base + lane, n - 1);
  float f[SN_IN];
#pragma unroll
  for (int k = 0; k < SN_IN; ++k) f[k] = feat[(size_t)s * SN_IN + k];
  float y0, y1;
  sn_eval_wave(N, f, lds + wave * SN_WAVE_LDS, y0, y1);
  if (base + lane < n) {
    out[2 * (size_t)(base + lane)] = y0;
    out[2 * (size_t)(base + lane) + 1] = y1;
  }
}

// ---- training: forward + backward of one 64-sample tile per workgroup --------
constexpr int LDF = SN_INP + 8;   // padded row strides (16-byte multiples)
constexpr int LDT = 64 + 8;

__global__ __launch_bounds__(64) void sn_train_tile(StepNetDev N, const float* __restrict__ W3f,
                                                    const float* __restrict__ feat, const float* __restrict__ tgt,
                                                    const uint32_t* __restrict__ idx, int n, float inv_batch,
                                                    float* __restrict__ part) {
  // every operand row starts on a 16-byte boundary (LDF, LDT multiples of 8)
  __shared__ __align__(16) uint16_t X[64 * LDF];      // [ray][f]
  __shared__ __align__(16) uint16_t XT[32 * LDT];     // [f][ray]
  __shared__ __align__(16) uint16_t H1[64 * LDT];
  __shared__ __align__(16) uint16_t H1T[64 * LDT];
  __shared__ __align__(16) uint16_t H2[64 * LDT];
  __shared__ __align__(16) uint16_t H2T[64 * LDT];
  __shared__ __align__(16) uint16_t D2[64 * LDT];
  __shared__ __align__(16) uint16_t D2T[64 * LDT];
  __shared__ __align__(16) uint16_t D1T[64 * LDT];
  __shared__ __align__(16) uint16_t DYT[32 * LDT];    // [o][ray], rows >= 2 zero
  __shared__ float Y[64][2];
  const int lane = threadIdx.x;
  const int r = lane & 31;
  const int t = blockIdx.x * 64 + lane;
  const bool valid = t < n;
  const uint32_t s = idx ? idx[min(t, n - 1)] : (uint32_t)min(t, n - 1);
  // stage x (bf16) in both orientations
#pragma unroll
  for (int k = 0; k < SN_INP; ++k) {
    const uint16_t b = (k < SN_IN && valid) ? bf16_bits(feat[(size_t)s * SN_IN + k]) : (uint16_t)0;
    X[lane * LDF + k] = b;
    XT[k * LDT + lane] = b;
  }
  for (int k = lane; k < 32 * LDT; k += 64) DYT[k] = 0;
  __syncthreads();
  // layer 1: h1^T[o][ray]
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const sn_f32x16 acc = sn_tile_nt<SN_INP>(N.w1t, SN_INP, X, LDF, 32 * mt, 32 * nt, lane);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int o = 32 * mt + sn_row(i, lane), ray = 32 * nt + r;
        const uint16_t h = bf16_bits(fmaxf(acc[i] + N.b1[o], 0.0f));
        H1[ray * LDT + o] = h;
        H1T[o * LDT + ray] = h;
      }
    }
  __syncthreads();
  // layer 2
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const sn_f32x16 acc = sn_tile_nt<SN_HID>(N.w2t, SN_HID, H1, LDT, 32 * mt, 32 * nt, lane);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int o = 32 * mt + sn_row(i, lane), ray = 32 * nt + r;
        const uint16_t h = bf16_bits(fmaxf(acc[i] + N.b2[o], 0.0f));
        H2[ray * LDT + o] = h;
        H2T[o * LDT + ray] = h;
      }
    }
  __syncthreads();
  // layer 3: outputs of column ray sit in elements 0, 1 of lanes 0..31
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    const sn_f32x16 acc = sn_tile_nt<SN_HID>(N.w3t, SN_HID, H2, LDT, 0, 32 * nt, lane);
    if (lane < 32) {
      Y[32 * nt + r][0] = acc[0] + N.b3[0];
      Y[32 * nt + r][1] = acc[1] + N.b3[1];
    }
  }
  __syncthreads();
  // loss and output gradient of this lane's sample
  float loss = 0.0f, dy0 = 0.0f, dy1 = 0.0f;
  if (valid) {
    const float y0 = Y[lane][0], y1 = Y[lane][1];
    const float t0 = tgt[(size_t)s * 3 + 0], t1 = tgt[(size_t)s * 3 + 1], w = tgt[(size_t)s * 3 + 2];
    const float e = y0 - t0;
    // BCE with logits: max(y,0) - y t + log(1 + exp(-|y|))
    const float bce = fmaxf(y1, 0.0f) - y1 * t1 + log1pf(expf(-fabsf(y1)));
    loss = w * (0.5f * e * e + bce);
    const float sg = 1.0f / (1.0f + expf(-y1));
    dy0 = w * e * inv_batch;
    dy1 = w * (sg - t1) * inv_batch;
  }
  const uint16_t d0 = bf16_bits(dy0), d1 = bf16_bits(dy1);
  DYT[0 * LDT + lane] = d0;
  DYT[1 * LDT + lane] = d1;
  // dh2 = (W3 dy) * [h2 > 0]   (fp32 W3 rounded to bf16 like the forward copy)
  const float e0 = bf16_value(d0), e1 = bf16_value(d1);
  for (int h = 0; h < SN_HID; ++h) {
    const float w30 = bf16_value(bf16_bits(W3f[h * SN_OUT + 0])), w31 = bf16_value(bf16_bits(W3f[h * SN_OUT + 1]));
    const float g = H2[lane * LDT + h] != 0 ? e0 * w30 + e1 * w31 : 0.0f;
    const uint16_t gb = bf16_bits(g);
    D2[lane * LDT + h] = gb;
    D2T[h * LDT + lane] = gb;
  }
  __syncthreads();
  float* P = part + (size_t)blockIdx.x * (SN_NPARAM + 1);
  // dW3^T[o][i] = sum_ray dy[ray][o] h2[ray][i]: A = DYT [o][ray], Bt = H2T [i][ray]
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    const sn_f32x16 acc = sn_tile_nt<64>(DYT, LDT, H2T, LDT, 0, 32 * nt, lane);
    if (lane < 32) {
      const int i = 32 * nt + r;
      P[SN_OFF_W3 + i * SN_OUT + 0] = acc[0];
      P[SN_OFF_W3 + i * SN_OUT + 1] = acc[1];
    }
  }
  // dW2^T[o][i] = sum_ray dh2[ray][o] h1[ray][i]: A = D2T [o][ray], Bt = H1T [i][ray]
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const sn_f32x16 acc = sn_tile_nt<64>(D2T, LDT, H1T, LDT, 32 * mt, 32 * nt, lane);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int o = 32 * mt + sn_row(k, lane), i = 32 * nt + r;
        P[SN_OFF_W2 + i * SN_HID + o] = acc[k];
      }
    }
  // dh1^T[i][ray] = sum_o W2[i][o] dh2[ray][o] * [h1 > 0]: A = W2 [i][o], Bt = D2 [ray][o]
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const sn_f32x16 acc = sn_tile_nt<64>(N.w2, SN_HID, D2, LDT, 32 * mt, 32 * nt, lane);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int i = 32 * mt + sn_row(k, lane), ray = 32 * nt + r;
        D1T[i * LDT + ray] = H1[ray * LDT + i] != 0 ? bf16_bits(acc[k]) : (uint16_t)0;
      }
    }
  __syncthreads();
  // dW1^T[o][f] = sum_ray dh1[ray][o] x[ray][f]: A = D1T [o][ray], Bt = XT [f][ray]
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    const sn_f32x16 acc = sn_tile_nt<64>(D1T, LDT, XT, LDT, 32 * mt, 0, lane);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int o = 32 * mt + sn_row(k, lane), f = r;
      if (f < SN_IN) P[SN_OFF_W1 + f * SN_HID + o] = acc[k];
    }
  }
  // bias gradients: sums over the tile of the bf16 gradients
  {
    float g1 = 0.0f, g2 = 0.0f;
    for (int j = 0; j < 64; ++j) {
      g1 += bf16_value(D1T[lane * LDT + j]);
      g2 += bf16_value(D2T[lane * LDT + j]);
    }
    P[SN_OFF_B1 + lane] = g1;
    P[SN_OFF_B2 + lane] = g2;
    if (lane < SN_OUT) {
      float g3 = 0.0f;
      for (int j = 0; j < 64; ++j) g3 += bf16_value(DYT[lane * LDT + j]);
      P[SN_OFF_B3 + lane] = g3;
    }
  }
  // loss of the tile (fixed-order tree over the wave)
  float l = loss;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if (lane == 0) P[SN_NPARAM] = l;
}

// Sums the partials of `tiles` workgroups (fixed order) into grad[SN_NPARAM + 1].
__global__ void sn_reduce(const float* __restrict__ part, int tiles, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > SN_NPARAM) return;
  float s = 0.0f;
  for (int t = 0; t < tiles; ++t) s += part[(size_t)t * (SN_NPARAM + 1) + p];
  grad[p] = s;
}

// Adam on the fp32 master parameters (bias-corrected, step >= 1).
__global__ void sn_adam(float* __restrict__ th, float* __restrict__ m, float* __restrict__ v,
                        const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2,
                        float wd) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= SN_NPARAM) return;
  const float g = grad[p] + wd * th[p];
  const float mm = b1 * m[p] + (1.0f - b1) * g;
  const float vv = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  th[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ---- launchers ------------------------------------------------------------------
hipError_t launch_sn_pack(const float* th, uint16_t* w1t, uint16_t* w2t, uint16_t* w2, uint16_t* w3t, float* b1,
                          float* b2, float* b3, hipStream_t st) {
  hipLaunchKernelGGL(sn_pack, dim3((SN_HID * SN_HID + 255) / 256), dim3(256), 0, st, th, w1t, w2t, w2, w3t, b1, b2,
                     b3);
  return hipGetLastError();
}

hipError_t launch_sn_infer(const StepNetDev& N, const float* feat, int n, float* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int blocks = (n + 255) / 256;
  hipLaunchKernelGGL(sn_infer, dim3(blocks), dim3(256), 0, st, N, feat, n, out);
  return hipGetLastError();
}

// partial gradients of ceil(n/64) tiles into part[tiles][SN_NPARAM + 1], then summed into grad
hipError_t launch_sn_grad(const StepNetDev& N, const float* W3f, const float* feat, const float* tgt,
                          const uint32_t* idx, int n, float inv_batch, float* part, float* grad, hipStream_t st) {
  if (n <= 0) return hipErrorInvalidValue;
  const int tiles = (n + 63) / 64;
  hipLaunchKernelGGL(sn_train_tile, dim3(tiles), dim3(64), 0, st, N, W3f, feat, tgt, idx, n, inv_batch, part);
  hipLaunchKernelGGL(sn_reduce, dim3((SN_NPARAM + 1 + 255) / 256), dim3(256), 0, st, part, tiles, grad);
  return hipGetLastError();
}

hipError_t launch_sn_adam(float* th, float* m, float* v, const float* grad, float lr, float b1, float b2, float eps,
                          int step, float wd, hipStream_t st) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(sn_adam, dim3((SN_NPARAM + 255) / 256), dim3(256), 0, st, th, m, v, grad, lr, b1, b2, eps, c1,
                     c2, wd);
  return hipGetLastError();
}

}  // namespace tvk
