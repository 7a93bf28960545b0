// Oracle: CPU restatement of Tuvok's volume-raycasting path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (tuvok_amd/, include/)
// links, loads or calls this code; it is the checker that the HIP path is
// compared against in tests/, __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py.
//
// Math conventions follow Basics/Vectors.h of the reference:
//   * MATRIX4 stores m11..m44 row-major in array[16] (array[0]=m11, array[1]=m12, ...)
//   * vectors are ROW vectors: v' = v * M  (Vectors.h:434)
//   * A*B is the ordinary row-major product (Vectors.h:936-946), so v*(A*B)
//     applies A first.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

namespace orc {

struct V3 {
  float x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(float a, float b, float c) : x(a), y(b), z(c) {}
  V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator*(const V3& o) const { return V3(x * o.x, y * o.y, z * o.z); }
  V3 operator/(const V3& o) const { return V3(x / o.x, y / o.y, z / o.z); }
  V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
  V3 operator/(float s) const { return V3(x / s, y / s, z / s); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  float maxVal() const { return std::max(x, std::max(y, z)); }
  float minVal() const { return std::min(x, std::min(y, z)); }
};
inline float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// MathTools::NextPow2 (MathTools.h:71; returns n itself when it is a power of two)
inline uint32_t next_pow2(uint32_t n) {
  uint32_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

// GLSL pow(y, s) of the opacity correction 1 - (1 - a)^s (VRender1D.glsl:47-57,
// OpacityCorrectColor of GLGridLeaper-blend.glsl; the built-in's precision is
// implementation defined).  Exponents that are powers of two from 1/4 to 16 --
// all that a sample-rate modifier of 2^k and a level's 2^k domain ratio give --
// are evaluated by square roots and repeated squaring, any other by powf.  The
// HIP kernels take the same steps, so the two agree bit for bit.
inline float opacity_pow(float y, float s) {
  if (s == 1.0f) return y;
  if (s == 0.5f) return std::sqrt(y);
  if (s == 0.25f) return std::sqrt(std::sqrt(y));
  if (s == 2.0f || s == 4.0f || s == 8.0f || s == 16.0f) {
    float r = y * y;
    for (float e = 4.0f; e <= s; e *= 2.0f) r = r * r;
    return r;
  }
  return std::pow(y, s);
}
inline float length(const V3& a) { return std::sqrt(dot(a, a)); }
inline V3 cross(const V3& a, const V3& b) {  // Vectors.h operator% (cross product)
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline V3 normalize(const V3& a) { float l = length(a); return a / l; }

struct M4 {
  float a[16];
  M4() { for (int i = 0; i < 16; ++i) a[i] = (i % 5 == 0) ? 1.f : 0.f; }
  float& m(int r, int c) { return a[r * 4 + c]; }
  float m(int r, int c) const { return a[r * 4 + c]; }
};

// Vectors.h:936-946
inline M4 mul(const M4& A, const M4& B) {
  M4 R;
  for (int x = 0; x < 16; x += 4)
    for (int y = 0; y < 4; ++y)
      R.a[x + y] = A.a[0 + x] * B.a[0 + y] + A.a[1 + x] * B.a[4 + y] +
                   A.a[2 + x] * B.a[8 + y] + A.a[3 + x] * B.a[12 + y];
  return R;
}

// row-vector * matrix, Vectors.h:434-439
inline void vmul4(const float v[4], const M4& M, float out[4]) {
  for (int c = 0; c < 4; ++c)
    out[c] = v[0] * M.m(0, c) + v[1] * M.m(1, c) + v[2] * M.m(2, c) + v[3] * M.m(3, c);
}
inline V3 xform_point(const V3& p, const M4& M) {  // (p,1)*M, xyz (no dehomogenisation)
  float v[4] = {p.x, p.y, p.z, 1.f}, o[4];
  vmul4(v, M, o);
  return V3(o[0], o[1], o[2]);
}
inline V3 xform_dir(const V3& d, const M4& M) {  // (d,0)*M
  float v[4] = {d.x, d.y, d.z, 0.f}, o[4];
  vmul4(v, M, o);
  return V3(o[0], o[1], o[2]);
}

// General 4x4 inverse (Vectors.h:1042 computes the same adjugate/determinant
// quotient; evaluation order differs, which only moves the last ulp).
inline M4 inverse(const M4& M) {
  double s[16];
  for (int i = 0; i < 16; ++i) s[i] = M.a[i];
  double inv[16];
  inv[0] = s[5] * s[10] * s[15] - s[5] * s[11] * s[14] - s[9] * s[6] * s[15] + s[9] * s[7] * s[14] + s[13] * s[6] * s[11] - s[13] * s[7] * s[10];
  inv[4] = -s[4] * s[10] * s[15] + s[4] * s[11] * s[14] + s[8] * s[6] * s[15] - s[8] * s[7] * s[14] - s[12] * s[6] * s[11] + s[12] * s[7] * s[10];
  inv[8] = s[4] * s[9] * s[15] - s[4] * s[11] * s[13] - s[8] * s[5] * s[15] + s[8] * s[7] * s[13] + s[12] * s[5] * s[11] - s[12] * s[7] * s[9];
  inv[12] = -s[4] * s[9] * s[14] + s[4] * s[10] * s[13] + s[8] * s[5] * s[14] - s[8] * s[6] * s[13] - s[12] * s[5] * s[10] + s[12] * s[6] * s[9];
  inv[1] = -s[1] * s[10] * s[15] + s[1] * s[11] * s[14] + s[9] * s[2] * s[15] - s[9] * s[3] * s[14] - s[13] * s[2] * s[11] + s[13] * s[3] * s[10];
  inv[5] = s[0] * s[10] * s[15] - s[0] * s[11] * s[14] - s[8] * s[2] * s[15] + s[8] * s[3] * s[14] + s[12] * s[2] * s[11] - s[12] * s[3] * s[10];
  inv[9] = -s[0] * s[9] * s[15] + s[0] * s[11] * s[13] + s[8] * s[1] * s[15] - s[8] * s[3] * s[13] - s[12] * s[1] * s[11] + s[12] * s[3] * s[9];
  inv[13] = s[0] * s[9] * s[14] - s[0] * s[10] * s[13] - s[8] * s[1] * s[14] + s[8] * s[2] * s[13] + s[12] * s[1] * s[10] - s[12] * s[2] * s[9];
  inv[2] = s[1] * s[6] * s[15] - s[1] * s[7] * s[14] - s[5] * s[2] * s[15] + s[5] * s[3] * s[14] + s[13] * s[2] * s[7] - s[13] * s[3] * s[6];
  inv[6] = -s[0] * s[6] * s[15] + s[0] * s[7] * s[14] + s[4] * s[2] * s[15] - s[4] * s[3] * s[14] - s[12] * s[2] * s[7] + s[12] * s[3] * s[6];
  inv[10] = s[0] * s[5] * s[15] - s[0] * s[7] * s[13] - s[4] * s[1] * s[15] + s[4] * s[3] * s[13] + s[12] * s[1] * s[7] - s[12] * s[3] * s[5];
  inv[14] = -s[0] * s[5] * s[14] + s[0] * s[6] * s[13] + s[4] * s[1] * s[14] - s[4] * s[2] * s[13] - s[12] * s[1] * s[6] + s[12] * s[2] * s[5];
  inv[3] = -s[1] * s[6] * s[11] + s[1] * s[7] * s[10] + s[5] * s[2] * s[11] - s[5] * s[3] * s[10] - s[9] * s[2] * s[7] + s[9] * s[3] * s[6];
  inv[7] = s[0] * s[6] * s[11] - s[0] * s[7] * s[10] - s[4] * s[2] * s[11] + s[4] * s[3] * s[10] + s[8] * s[2] * s[7] - s[8] * s[3] * s[6];
  inv[11] = -s[0] * s[5] * s[11] + s[0] * s[7] * s[9] + s[4] * s[1] * s[11] - s[4] * s[3] * s[9] - s[8] * s[1] * s[7] + s[8] * s[3] * s[5];
  inv[15] = s[0] * s[5] * s[10] - s[0] * s[6] * s[9] - s[4] * s[1] * s[10] + s[4] * s[2] * s[9] + s[8] * s[1] * s[6] - s[8] * s[2] * s[5];
  double det = s[0] * inv[0] + s[1] * inv[4] + s[2] * inv[8] + s[3] * inv[12];
  M4 R;
  for (int i = 0; i < 16; ++i) R.a[i] = (float)(inv[i] / det);
  return R;
}

// The clip plane in model space as GLRaycaster::RenderBox forms it
// (GLRaycaster.cpp:328-336): m_ClipPlane.Plane() * inverse(rotation * translation),
// where PLANE * M is PLANE::transform (Vectors.h:1460-1485): the row vector times
// the transposed inverse of M, divided by the length of its xyz; then the
// normal normalised again and d kept.  Points with n.p + d <= 0 stay
// (Clipper::TriPlane keeps fa <= 0, Clipper.cpp:78-104).
inline void model_clip_plane(const float plane[4], const M4& rot, const M4& trans, float out[4]) {
  const M4 inv = inverse(mul(rot, trans));
  const M4 ii = inverse(inv);
  M4 it;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) it.m(r, c) = ii.m(c, r);
  float t[4];
  vmul4(plane, it, t);
  const float l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  for (float& v : t) v /= l;
  const float l2 = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  out[0] = t[0] / l2; out[1] = t[1] / l2; out[2] = t[2] / l2; out[3] = t[3];
}

// Vectors.h:970-975
inline M4 translation(float x, float y, float z) {
  M4 R; R.m(3, 0) = x; R.m(3, 1) = y; R.m(3, 2) = z; return R;
}
// Vectors.h:977-982
inline M4 scaling(float x, float y, float z) {
  M4 R; R.m(0, 0) = x; R.m(1, 1) = y; R.m(2, 2) = z; return R;
}
// Vectors.h:994-1002
inline M4 rotation_y(double angle) {
  M4 R; float c = (float)std::cos(angle), s = (float)std::sin(angle);
  R.m(0, 0) = c; R.m(0, 2) = -s; R.m(2, 0) = s; R.m(2, 2) = c; return R;
}
// Vectors.h:1250-1265
inline M4 look_at(const V3& eye, const V3& at, const V3& up) {
  V3 F = at - eye, U = up, S = cross(F, U);
  U = cross(S, F);
  F = normalize(F); U = normalize(U); S = normalize(S);
  M4 R;
  R.a[0] = S.x;  R.a[4] = S.y;  R.a[8] = S.z;   R.a[12] = -dot(S, eye);
  R.a[1] = U.x;  R.a[5] = U.y;  R.a[9] = U.z;   R.a[13] = -dot(U, eye);
  R.a[2] = -F.x; R.a[6] = -F.y; R.a[10] = -F.z; R.a[14] = dot(F, eye);
  R.a[3] = 0;    R.a[7] = 0;    R.a[11] = 0;    R.a[15] = 1;
  return R;
}
// Vectors.h:1267-1277
inline M4 perspective(float fovy, float aspect, float n, float f) {
  fovy = fovy * (float)(3.14159265358979323846 / 180.0);
  float cotan = (float)(1.0 / std::tan((double)fovy / 2.0));
  M4 R;
  for (int i = 0; i < 16; ++i) R.a[i] = 0;
  R.a[0] = cotan / aspect;
  R.a[5] = cotan;
  R.a[10] = -(f + n) / (f - n);
  R.a[14] = -2.f * (f * n) / (f - n);
  R.a[11] = -1.f;
  return R;
}

// IEEE binary16 round-to-nearest-even, used to model the RGBA16F ray-entry
// render target of GLRaycaster.cpp:97.
inline uint16_t float_to_half_bits(float f) {
  uint32_t x; std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t e = (x >> 23) & 0xffu;
  uint32_t mant = x & 0x7fffffu;
  if (e == 0xffu) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
  int32_t exp = (int32_t)e - 127 + 15;
  if (exp >= 31) return (uint16_t)(sign | 0x7c00u);
  if (exp <= 0) {
    if (exp < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - exp);
    uint32_t hm = mant >> shift;
    uint32_t rem = mant & ((1u << shift) - 1u);
    uint32_t halfway = 1u << (shift - 1u);
    if (rem > halfway || (rem == halfway && (hm & 1u))) hm++;
    return (uint16_t)(sign | hm);
  }
  uint32_t h = sign | ((uint32_t)exp << 10) | (mant >> 13);
  uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)h;
}
inline float half_bits_to_float(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) { x = sign; }
    else {
      int ee = -1;
      do { ++ee; m <<= 1; } while ((m & 0x400u) == 0);
      m &= 0x3ffu;
      x = sign | ((uint32_t)(127 - 15 - ee) << 23) | (m << 13);
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e - 15 + 127) << 23) | (m << 13);
  }
  float f; std::memcpy(&f, &x, 4); return f;
}
inline float round_fp16(float f) { return half_bits_to_float(float_to_half_bits(f)); }

}  // namespace orc
