// Step-proposal network: 24 features -> 64 -> 64 -> 2 (ReLU hidden layers),
// BF16 weights and activations, FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16).  Output 0 is the proposed log2 step multiplier
// divided by 3 (multiplier 1..8), output 1 the logit of "skip this segment".
//
// Operand layouts of one 32x32x16 MFMA (lane l = 0..63, r = l % 32, h = l / 32):
//   A (32 x 16):  lane l holds A[r][8h .. 8h+7]
//   B (16 x 32):  lane l holds B[8h .. 8h+7][r]
//   D (32 x 32):  lane l, element i (0..15) is D[8*(i/4) + 4h + i%4][r]
// Every product below is written "NT": C = A * Bt^T with both A and Bt stored
// row-major with K contiguous, so each operand is one 16-byte load per lane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tvk {

constexpr int SN_IN = 24;     // features
constexpr int SN_INP = 32;    // features padded to the MFMA K step
constexpr int SN_HID = 64;
constexpr int SN_OUT = 2;
// fp32 parameter vector: W1[24][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2]
constexpr int SN_OFF_W1 = 0;
constexpr int SN_OFF_B1 = SN_OFF_W1 + SN_IN * SN_HID;
constexpr int SN_OFF_W2 = SN_OFF_B1 + SN_HID;
constexpr int SN_OFF_B2 = SN_OFF_W2 + SN_HID * SN_HID;
constexpr int SN_OFF_W3 = SN_OFF_B2 + SN_HID;
constexpr int SN_OFF_B3 = SN_OFF_W3 + SN_HID * SN_OUT;
constexpr int SN_NPARAM = SN_OFF_B3 + SN_OUT;   // 5890

// Device copy of the network as the kernels read it (bf16 weights, K-contiguous).
struct StepNetDev {
  const uint16_t* w1t;   // [64][32]  W1[f][o] at [o][f], f >= 24 zero
  const uint16_t* w2t;   // [64][64]  W2[i][o] at [o][i]
  const uint16_t* w2;    // [64][64]  W2[i][o] at [i][o]  (backward)
  const uint16_t* w3t;   // [32][64]  W3[i][o] at [o][i], o >= 2 zero
  const float* b1;       // [64]
  const float* b2;       // [64]
  const float* b3;       // [2]
};

typedef __bf16 sn_bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short sn_u16x8 __attribute__((ext_vector_type(8)));
typedef float sn_f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16 bits, round to nearest even (finite values)
__host__ __device__ inline uint16_t bf16_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_value(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

__device__ inline sn_f32x16 sn_mfma(sn_u16x8 a, sn_u16x8 b, sn_f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(sn_bf16x8, a), __builtin_bit_cast(sn_bf16x8, b),
                                                 c, 0, 0, 0);
}

// one 32x32 tile of C = A * Bt^T, K a multiple of 16; rows m0.., n0..
template <int K>
__device__ inline sn_f32x16 sn_tile_nt(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, int m0, int n0,
                                       int lane) {
  sn_f32x16 acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
  const int r = lane & 31, kh = (lane >> 5) * 8;
#pragma unroll
  for (int ks = 0; ks < K / 16; ++ks) {
    const sn_u16x8 a = *reinterpret_cast<const sn_u16x8*>(A + (size_t)(m0 + r) * lda + ks * 16 + kh);
    const sn_u16x8 b = *reinterpret_cast<const sn_u16x8*>(Bt + (size_t)(n0 + r) * ldb + ks * 16 + kh);
    acc = sn_mfma(a, b, acc);
  }
  return acc;
}

// row of D element i for this lane (the column is lane % 32)
__device__ inline int sn_row(int i, int lane) { return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3); }

// orders LDS writes and reads of the lanes of one wave
__device__ inline void sn_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS scratch per wave for sn_eval_wave: X [64][32] and H1 [32][64] bf16
constexpr int SN_WAVE_LDS = 64 * SN_INP + 32 * SN_HID;

// Evaluates the network for the 64 rays of a wave (lane = ray) whose inputs
// are already in the X rows of lds (bf16, row = lane, features 24..31 zero).
// lds: SN_WAVE_LDS uint16 per wave.  Layers 1 and 2 are MFMA products over 32
// rays at a time (hidden unit x ray tiles); layer 3 (64 -> 2) is folded into
// the layer-2 epilogue: every lane dots its 16 bf16-rounded hidden units of a
// tile with the bf16 W3 columns, and the two lanes that hold the same ray add
// their partial sums.  Loops are kept rolled so that the evaluation adds few
// registers to the caller's.
__device__ inline void sn_eval_wave_x(const StepNetDev& N, uint16_t* lds, float& y0, float& y1) {
  const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
  uint16_t* xs = lds;                   // X [64][32]: ray, feature
  uint16_t* h1 = lds + 64 * SN_INP;     // H1 [32][64]: ray of the half, hidden unit
  sn_wave_sync();
  float o0 = 0.0f, o1 = 0.0f;
#pragma unroll 1
  for (int half = 0; half < 2; ++half) {
#pragma unroll 1
    for (int mt = 0; mt < 2; ++mt) {
      const sn_f32x16 acc = sn_tile_nt<SN_INP>(N.w1t, SN_INP, xs + 32 * half * SN_INP, SN_INP, 32 * mt, 0, lane);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int o = 32 * mt + 8 * q + 4 * h;   // rows o .. o+3 of element 4q .. 4q+3
        const float4 b = *reinterpret_cast<const float4*>(N.b1 + o);
        const float bb[4] = {b.x, b.y, b.z, b.w};
        uint16_t hv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) hv[j] = bf16_bits(fmaxf(acc[4 * q + j] + bb[j], 0.0f));
        uint2 w;
        w.x = (uint32_t)hv[0] | ((uint32_t)hv[1] << 16);
        w.y = (uint32_t)hv[2] | ((uint32_t)hv[3] << 16);
        *reinterpret_cast<uint2*>(h1 + r * SN_HID + o) = w;
      }
    }
    sn_wave_sync();
    float p0 = 0.0f, p1 = 0.0f;
#pragma unroll 1
    for (int mt = 0; mt < 2; ++mt) {
      const sn_f32x16 acc = sn_tile_nt<SN_HID>(N.w2t, SN_HID, h1, SN_HID, 32 * mt, 0, lane);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int o = 32 * mt + 8 * q + 4 * h;
        const float4 b = *reinterpret_cast<const float4*>(N.b2 + o);
        const uint2 wa = *reinterpret_cast<const uint2*>(N.w3t + o);            // W3[o..o+3][0]
        const uint2 wb = *reinterpret_cast<const uint2*>(N.w3t + SN_HID + o);   // W3[o..o+3][1]
        const float bb[4] = {b.x, b.y, b.z, b.w};
        const uint32_t a2[4] = {wa.x & 0xffffu, wa.x >> 16, wa.y & 0xffffu, wa.y >> 16};
        const uint32_t b2[4] = {wb.x & 0xffffu, wb.x >> 16, wb.y & 0xffffu, wb.y >> 16};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float v = bf16_value(bf16_bits(fmaxf(acc[4 * q + j] + bb[j], 0.0f)));
          p0 += v * bf16_value((uint16_t)a2[j]);
          p1 += v * bf16_value((uint16_t)b2[j]);
        }
      }
    }
    // lanes l and l ^ 32 hold the two halves of the hidden units of ray r
    p0 += __shfl_xor(p0, 32, 64);
    p1 += __shfl_xor(p1, 32, 64);
    if (h == half) { o0 = p0 + N.b3[0]; o1 = p1 + N.b3[1]; }
    sn_wave_sync();   // H1 is rewritten by the next half
  }
  y0 = o0;
  y1 = o1;
}

// The same from this lane's features f (fp32, rounded to bf16 here).
__device__ inline void sn_eval_wave(const StepNetDev& N, const float* f, uint16_t* lds, float& y0, float& y1) {
  const int lane = threadIdx.x & 63;
  sn_u16x8 v[SN_INP / 8];
#pragma unroll
  for (int k = 0; k < SN_INP; ++k) v[k >> 3][k & 7] = k < SN_IN ? bf16_bits(f[k]) : (uint16_t)0;
#pragma unroll
  for (int q = 0; q < SN_INP / 8; ++q) *reinterpret_cast<sn_u16x8*>(lds + lane * SN_INP + 8 * q) = v[q];
  sn_eval_wave_x(N, lds, y0, y1);
}

}  // namespace tvk
