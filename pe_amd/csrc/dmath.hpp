// FP64 small-vector algebra for the HCSITS time step, host and device.
//
// Each operation keeps pe's evaluation order (pe/math/Vector3.h, Matrix3x3.h,
// Matrix2x2.h, Quaternion.h) so that, compiled with -ffp-contract=off, results
// are bit-identical to the reference wherever libm is not involved.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define PE_HD __host__ __device__ __forceinline__

namespace pe_amd {

struct d3 {
  double x, y, z;
};
PE_HD d3 mk3(double a, double b, double c) { d3 r; r.x = a; r.y = b; r.z = c; return r; }
PE_HD d3 operator+(d3 a, d3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
PE_HD d3 operator-(d3 a, d3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
PE_HD d3 operator-(d3 a) { return mk3(-a.x, -a.y, -a.z); }
PE_HD d3 operator*(d3 a, double s) { return mk3(a.x * s, a.y * s, a.z * s); }
PE_HD d3 operator*(double s, d3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
// trans(a) * b  (Vector3.h:2096-2099)
PE_HD double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// a % b  (Vector3.h:2339-2345)
PE_HD d3 cross3(d3 a, d3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Vector3::length (Vector3.h:952-959); sqrt is correctly rounded on host and device
PE_HD double len3(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector3::normalize / getNormalized (Vector3.h:994-1039): reciprocal, then scale
PE_HD d3 nrm3(d3 a) {
  const double len = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  if (len == 0.0) return a;
  const double il = 1.0 / len;
  return mk3(a.x * il, a.y * il, a.z * il);
}
PE_HD double comp(d3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Row-major 3x3.
struct m33 {
  double a[9];
};
PE_HD m33 mat_cols(d3 c1, d3 c2, d3 c3) {  // Matrix3x3.h:351-358
  m33 m;
  m.a[0] = c1.x; m.a[1] = c2.x; m.a[2] = c3.x;
  m.a[3] = c1.y; m.a[4] = c2.y; m.a[5] = c3.y;
  m.a[6] = c1.z; m.a[7] = c2.z; m.a[8] = c3.z;
  return m;
}
PE_HD m33 mat_mul(const m33& l, const m33& r) {  // Matrix3x3.h:2241-2253
  m33 m;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      m.a[3 * i + j] = l.a[3 * i] * r.a[j] + l.a[3 * i + 1] * r.a[3 + j] + l.a[3 * i + 2] * r.a[6 + j];
  return m;
}
PE_HD d3 mat_vec(const m33& l, d3 r) {  // Matrix3x3.h:2172-2178
  return mk3(l.a[0] * r.x + l.a[1] * r.y + l.a[2] * r.z, l.a[3] * r.x + l.a[4] * r.y + l.a[5] * r.z,
             l.a[6] * r.x + l.a[7] * r.y + l.a[8] * r.z);
}
PE_HD m33 mat_trans(const m33& m) {  // Matrix3x3.h:1857-1860
  m33 r;
  r.a[0] = m.a[0]; r.a[1] = m.a[3]; r.a[2] = m.a[6];
  r.a[3] = m.a[1]; r.a[4] = m.a[4]; r.a[5] = m.a[7];
  r.a[6] = m.a[2]; r.a[7] = m.a[5]; r.a[8] = m.a[8];
  return r;
}
// trans(C) * v without forming the transpose: rows of trans(C) are the columns of C
PE_HD d3 mat_tvec(const m33& m, d3 v) {
  return mk3(m.a[0] * v.x + m.a[3] * v.y + m.a[6] * v.z, m.a[1] * v.x + m.a[4] * v.y + m.a[7] * v.z,
             m.a[2] * v.x + m.a[5] * v.y + m.a[8] * v.z);
}
// [v]x * M  (vec % mat, Matrix3x3.h:2417-2430)
PE_HD m33 skew_mul(d3 v, const m33& m) {
  m33 r;
  r.a[0] = v.y * m.a[6] - v.z * m.a[3]; r.a[1] = v.y * m.a[7] - v.z * m.a[4]; r.a[2] = v.y * m.a[8] - v.z * m.a[5];
  r.a[3] = v.z * m.a[0] - v.x * m.a[6]; r.a[4] = v.z * m.a[1] - v.x * m.a[7]; r.a[5] = v.z * m.a[2] - v.x * m.a[8];
  r.a[6] = v.x * m.a[3] - v.y * m.a[0]; r.a[7] = v.x * m.a[4] - v.y * m.a[1]; r.a[8] = v.x * m.a[5] - v.y * m.a[2];
  return r;
}
// M * [v]x  (mat % vec, Matrix3x3.h:2355-2367)
PE_HD m33 mul_skew(const m33& m, d3 v) {
  m33 r;
  r.a[0] = m.a[1] * v.z - m.a[2] * v.y; r.a[1] = m.a[2] * v.x - m.a[0] * v.z; r.a[2] = m.a[0] * v.y - m.a[1] * v.x;
  r.a[3] = m.a[4] * v.z - m.a[5] * v.y; r.a[4] = m.a[5] * v.x - m.a[3] * v.z; r.a[5] = m.a[3] * v.y - m.a[4] * v.x;
  r.a[6] = m.a[7] * v.z - m.a[8] * v.y; r.a[7] = m.a[8] * v.x - m.a[6] * v.z; r.a[8] = m.a[6] * v.y - m.a[7] * v.x;
  return r;
}
// Cofactor inverse (Matrix3x3::getInverse, Matrix3x3.h:1084-1104)
PE_HD m33 mat_inv(const m33& m) {
  const double* v = m.a;
  double det = v[0] * ((v[4] * v[8]) - (v[7] * v[5])) + v[1] * ((v[5] * v[6]) - (v[8] * v[3])) +
               v[2] * ((v[3] * v[7]) - (v[4] * v[6]));
  det = 1.0 / det;
  m33 r;
  r.a[0] = det * ((v[4] * v[8]) - (v[5] * v[7]));
  r.a[1] = det * ((v[7] * v[2]) - (v[8] * v[1]));
  r.a[2] = det * ((v[1] * v[5]) - (v[2] * v[4]));
  r.a[3] = det * ((v[5] * v[6]) - (v[3] * v[8]));
  r.a[4] = det * ((v[8] * v[0]) - (v[6] * v[2]));
  r.a[5] = det * ((v[2] * v[3]) - (v[0] * v[5]));
  r.a[6] = det * ((v[3] * v[7]) - (v[4] * v[6]));
  r.a[7] = det * ((v[6] * v[1]) - (v[7] * v[0]));
  r.a[8] = det * ((v[0] * v[4]) - (v[1] * v[3]));
  return r;
}

// Quaternion (w, x, y, z).
struct qd {
  double v[4];
};
// Quaternion(axis, angle) (Quaternion.h:274-294); Trig supplies sin and cos.
template <class Trig>
PE_HD qd quat_axis_angle(d3 axis, double angle) {
  qd q;
  if (fabs(angle) < 1e-15) { q.v[0] = 1; q.v[1] = 0; q.v[2] = 0; q.v[3] = 0; return q; }
  const double sina = Trig::sin(angle * 0.5);
  const double cosa = Trig::cos(angle * 0.5);
  axis = nrm3(axis);
  q.v[0] = cosa; q.v[1] = sina * axis.x; q.v[2] = sina * axis.y; q.v[3] = sina * axis.z;
  return q;
}
// Quaternion product with conditional renormalisation (Quaternion.h:1146-1160)
PE_HD qd quat_mul(const qd& l, const qd& r) {
  const double a = l.v[0] * r.v[0] - l.v[1] * r.v[1] - l.v[2] * r.v[2] - l.v[3] * r.v[3];
  const double i = l.v[0] * r.v[1] + l.v[1] * r.v[0] + l.v[2] * r.v[3] - l.v[3] * r.v[2];
  const double j = l.v[0] * r.v[2] + l.v[2] * r.v[0] + l.v[3] * r.v[1] - l.v[1] * r.v[3];
  const double k = l.v[0] * r.v[3] + l.v[3] * r.v[0] + l.v[1] * r.v[2] - l.v[2] * r.v[1];
  const double len2 = a * a + i * i + j * j + k * k;
  qd q;
  if (fabs(len2 - 1.0) < 1e-8) {
    q.v[0] = a; q.v[1] = i; q.v[2] = j; q.v[3] = k;
  } else {
    const double il = 1.0 / sqrt(len2);
    q.v[0] = a * il; q.v[1] = i * il; q.v[2] = j * il; q.v[3] = k * il;
  }
  return q;
}
// Quaternion::toRotationMatrix (Quaternion.h:591-602)
PE_HD m33 quat_to_R(const qd& q) {
  const double* v = q.v;
  m33 m;
  m.a[0] = 1.0 - 2.0 * v[2] * v[2] - 2.0 * v[3] * v[3];
  m.a[1] = 2.0 * (v[1] * v[2] - v[0] * v[3]);
  m.a[2] = 2.0 * (v[1] * v[3] + v[0] * v[2]);
  m.a[3] = 2.0 * (v[1] * v[2] + v[0] * v[3]);
  m.a[4] = 1.0 - 2.0 * v[1] * v[1] - 2.0 * v[3] * v[3];
  m.a[5] = 2.0 * (v[2] * v[3] - v[0] * v[1]);
  m.a[6] = 2.0 * (v[1] * v[3] - v[0] * v[2]);
  m.a[7] = 2.0 * (v[2] * v[3] + v[0] * v[1]);
  m.a[8] = 1.0 - 2.0 * v[1] * v[1] - 2.0 * v[2] * v[2];
  return m;
}
// Quaternion::rotate (Quaternion.h:744-756)
PE_HD d3 quat_rotate(const qd& q, d3 x) {
  const double* v = q.v;
  const double w = v[1] * x.x + v[2] * x.y + v[3] * x.z;
  const double a = v[0] * x.x - v[3] * x.y + v[2] * x.z;
  const double b = v[0] * x.y - v[1] * x.z + v[3] * x.x;
  const double c = v[0] * x.z - v[2] * x.x + v[1] * x.y;
  return mk3(v[0] * a + v[1] * w + v[2] * c - v[3] * b, v[0] * b + v[2] * w + v[3] * a - v[1] * c,
             v[0] * c + v[3] * w + v[1] * b - v[2] * a);
}

}  // namespace pe_amd
