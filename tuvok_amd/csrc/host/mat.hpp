// Row-vector 3D math with the conventions of the reference's Basics/Vectors.h:
// MATRIX4 keeps m11..m44 row-major in array[16]; vectors multiply from the left
// (v' = v * M, Vectors.h:434); A*B applies A first (Vectors.h:936).
// The float operation order here is the one the HIP kernels repeat, so host
// and device agree bit-for-bit on every derived quantity.
#pragma once
#include <cmath>
#include <cstdint>

namespace tvk {

struct F3 {
  float x = 0, y = 0, z = 0;
  F3() = default;
  F3(float a, float b, float c) : x(a), y(b), z(c) {}
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline F3 operator+(F3 a, F3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, F3 b) { return F3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline F3 operator/(F3 a, F3 b) { return F3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline F3 operator*(F3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
inline F3 operator/(F3 a, float s) { return F3(a.x / s, a.y / s, a.z / s); }
inline float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len3(F3 a) { return std::sqrt(dot3(a, a)); }
inline F3 cross3(F3 a, F3 b) { return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline F3 norm3(F3 a) { return a / len3(a); }
inline float max3(F3 a) { return a.x > a.y ? (a.x > a.z ? a.x : a.z) : (a.y > a.z ? a.y : a.z); }
inline float min3(F3 a) { return a.x < a.y ? (a.x < a.z ? a.x : a.z) : (a.y < a.z ? a.y : a.z); }

struct Mat4 {
  float a[16];
  Mat4() { identity(); }
  void identity() { for (int i = 0; i < 16; ++i) a[i] = (i % 5 == 0) ? 1.f : 0.f; }
  float& at(int r, int c) { return a[r * 4 + c]; }
  float at(int r, int c) const { return a[r * 4 + c]; }
  static Mat4 from(const float* p) { Mat4 m; for (int i = 0; i < 16; ++i) m.a[i] = p[i]; return m; }
};

inline Mat4 operator*(const Mat4& A, const Mat4& B) {
  Mat4 R;
  for (int r = 0; r < 16; r += 4)
    for (int c = 0; c < 4; ++c)
      R.a[r + c] = A.a[r] * B.a[c] + A.a[r + 1] * B.a[4 + c] + A.a[r + 2] * B.a[8 + c] + A.a[r + 3] * B.a[12 + c];
  return R;
}

// (v0,v1,v2,v3) * M
inline void row_mul(const float v[4], const Mat4& M, float o[4]) {
  for (int c = 0; c < 4; ++c)
    o[c] = v[0] * M.at(0, c) + v[1] * M.at(1, c) + v[2] * M.at(2, c) + v[3] * M.at(3, c);
}
inline F3 point_mul(F3 p, const Mat4& M) {
  float v[4] = {p.x, p.y, p.z, 1.f}, o[4];
  row_mul(v, M, o);
  return F3(o[0], o[1], o[2]);
}

// Inverse by the adjugate in double precision, rounded to float.
inline Mat4 inverse(const Mat4& M) {
  double m[16], v[16];
  for (int i = 0; i < 16; ++i) m[i] = M.a[i];
  v[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  v[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  v[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  v[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  v[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  v[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  v[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  v[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  v[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  v[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  v[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  v[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  v[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  v[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  v[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  v[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double det = m[0] * v[0] + m[1] * v[4] + m[2] * v[8] + m[3] * v[12];
  Mat4 R;
  for (int i = 0; i < 16; ++i) R.a[i] = (float)(v[i] / det);
  return R;
}

// The clip plane in model space as GLRaycaster::RenderBox forms it
// (GLRaycaster.cpp:328-336): m_ClipPlane.Plane() * inverse(rotation * translation),
// PLANE * M being PLANE::transform (Vectors.h:1460-1485) -- the row vector times
// the transposed inverse of M, divided by the length of its xyz -- then the
// normal normalised again and d kept.  Points with n.p + d <= 0 stay
// (Clipper::TriPlane keeps fa <= 0, Clipper.cpp:78-104).
inline void model_clip_plane(const float plane[4], const Mat4& rot, const Mat4& trans, float out[4]) {
  const Mat4 inv = inverse(rot * trans);
  const Mat4 ii = inverse(inv);
  Mat4 it;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) it.at(r, c) = ii.at(c, r);
  float t[4];
  row_mul(plane, it, t);
  const float l = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  for (float& v : t) v /= l;
  const float l2 = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  out[0] = t[0] / l2; out[1] = t[1] / l2; out[2] = t[2] / l2; out[3] = t[3];
}

inline Mat4 translation(float x, float y, float z) { Mat4 R; R.at(3, 0) = x; R.at(3, 1) = y; R.at(3, 2) = z; return R; }
inline Mat4 scaling(float x, float y, float z) { Mat4 R; R.at(0, 0) = x; R.at(1, 1) = y; R.at(2, 2) = z; return R; }
inline Mat4 rotation_y(double angle) {
  Mat4 R;
  float c = (float)std::cos(angle), s = (float)std::sin(angle);
  R.at(0, 0) = c; R.at(0, 2) = -s; R.at(2, 0) = s; R.at(2, 2) = c;
  return R;
}

// MATRIX4::BuildLookAt (Vectors.h:1250-1265)
inline Mat4 look_at(F3 eye, F3 at, F3 up) {
  F3 F = at - eye, U = up, S = cross3(F, U);
  U = cross3(S, F);
  F = norm3(F); U = norm3(U); S = norm3(S);
  Mat4 R;
  R.a[0] = S.x;  R.a[4] = S.y;  R.a[8] = S.z;   R.a[12] = -dot3(S, eye);
  R.a[1] = U.x;  R.a[5] = U.y;  R.a[9] = U.z;   R.a[13] = -dot3(U, eye);
  R.a[2] = -F.x; R.a[6] = -F.y; R.a[10] = -F.z; R.a[14] = dot3(F, eye);
  R.a[3] = 0;    R.a[7] = 0;    R.a[11] = 0;    R.a[15] = 1;
  return R;
}

// MATRIX4::Perspective (Vectors.h:1267-1277)
inline Mat4 perspective(float fovy, float aspect, float n, float f) {
  fovy = fovy * (float)(3.14159265358979323846 / 180.0);
  float cotan = (float)(1.0 / std::tan((double)fovy / 2.0));
  Mat4 R;
  for (int i = 0; i < 16; ++i) R.a[i] = 0.f;
  R.a[0] = cotan / aspect;
  R.a[5] = cotan;
  R.a[10] = -(f + n) / (f - n);
  R.a[14] = -2.f * (f * n) / (f - n);
  R.a[11] = -1.f;
  return R;
}

}  // namespace tvk
