// Face reconstruction along one pencil, per field, from the primitive values of the cells around
// a face.  Face f lies between cells f-1 and f; "ql" is the state on its low side (cell f-1's
// right edge) and "qr" the state on its high side (cell f's left edge), the caller's convention
// in grid/grid.cpp:229-239.
//
// PPM (reconstruct.cpp:22-58, called with its ql/qr arguments swapped at :94): theta = 2
// minmod-theta slopes, the interface value a_{i+1/2} = (q_i + q_{i+1})/2 + (s_i - s_{i+1})/6
// (plain average for a low-order field), then the Colella-Woodward monotonisation of each cell's
// pair of edges (skipped for a smooth field).  A face needs the six cells f-3 .. f+2.
//
// The FP64 functions keep the reference's operation order (build without FP contraction); the
// FP32 ones are the mixed form's.
#pragma once
#include <hip/hip_runtime.h>

#include "euler_phys.h"

namespace amrh {

namespace x64 {

// a_{i+1/2} from cells i, i+1 and their slopes (reconstruct.cpp:31-37)
__device__ __forceinline__ double ppm_iface(double qa, double qb, double sa, double sb, bool low_order) {
  double v = (qa + qb) * 0.5;
  if (!low_order) v += (sa - sb) * (1.0 / 6.0);
  return v;
}
// monotonise cell value q with edges A (left) and B (right) (reconstruct.cpp:39-53)
__device__ __forceinline__ void ppm_mono(double q, double& A, double& B) {
  const double tmp1 = B - A;
  const double tmp2 = B + A;
  if ((B - q) * (q - A) <= 0.0) {
    B = A = q;
  } else if (tmp1 * (q - 0.5 * tmp2) > (1.0 / 6.0) * tmp1 * tmp1) {
    A = 3.0 * q - 2.0 * B;
  } else if (-(1.0 / 6.0) * tmp1 * tmp1 > tmp1 * (q - 0.5 * tmp2)) {
    B = 3.0 * q - 2.0 * A;
  }
}
// q[0..5] = cells f-3 .. f+2
__device__ __forceinline__ void ppm_face(const double* q, double& ql, double& qr, bool low_order, bool smooth) {
  const double s1 = minmod_theta(q[2] - q[1], q[1] - q[0], 2.0);  // cell f-2
  const double s2 = minmod_theta(q[3] - q[2], q[2] - q[1], 2.0);  // cell f-1
  const double s3 = minmod_theta(q[4] - q[3], q[3] - q[2], 2.0);  // cell f
  const double s4 = minmod_theta(q[5] - q[4], q[4] - q[3], 2.0);  // cell f+1
  double a1 = ppm_iface(q[1], q[2], s1, s2, low_order);           // between f-2 and f-1
  double a2 = ppm_iface(q[2], q[3], s2, s3, low_order);           // face f itself
  double a3 = ppm_iface(q[3], q[4], s3, s4, low_order);           // between f and f+1
  double a2r = a2;
  if (!smooth) {
    ppm_mono(q[2], a1, a2);   // cell f-1: edges (a1, a2) -> its right edge is ql
    ppm_mono(q[3], a2r, a3);  // cell f: edges (a2, a3) -> its left edge is qr
  }
  ql = a2;
  qr = a2r;
}
// PPM along a whole pencil: q[0 .. N+5] are cells -3 .. N+2, and ql[f], qr[f] for the faces
// f = 0 .. N are the values ppm_face gives each of them.  Every slope, interface value and
// monotonisation is formed once and shared by the faces that use it (for N = 8: 12 slopes, 11
// interfaces and 10 monotonisations for the pencil's 9 faces, against 36, 27 and 18 face by face);
// each is the same function of the same inputs, so the values are ppm_face's bit for bit.
template <int N>
__device__ __forceinline__ void ppm_pencil(const double* q, double* ql, double* qr, bool low_order, bool smooth) {
  double s[N + 4], a[N + 3];
#pragma unroll
  for (int c = 0; c < N + 4; c++) s[c] = minmod_theta(q[c + 2] - q[c + 1], q[c + 1] - q[c], 2.0);  // cell c-2
#pragma unroll
  for (int c = 0; c < N + 3; c++) a[c] = ppm_iface(q[c + 1], q[c + 2], s[c], s[c + 1], low_order);  // cells c-2 | c-1
#pragma unroll
  for (int c = 0; c < N + 2; c++) {  // cell c-1, edges a[c] (left) and a[c+1] (right)
    double A = a[c], B = a[c + 1];
    if (!smooth) ppm_mono(q[c + 2], A, B);
    if (c >= 1) qr[c - 1] = A;  // the high side of face c-1
    if (c <= N) ql[c] = B;      // the low side of face c
  }
}
// minmod-theta: q[0..3] = cells f-2 .. f+1 (reconstruct.cpp:60-71)
__device__ __forceinline__ void minmod_face(const double* q, double& ql, double& qr, double theta) {
  const double sl = minmod_theta(q[2] - q[1], q[1] - q[0], theta);
  const double sr = minmod_theta(q[3] - q[2], q[2] - q[1], theta);
  ql = q[1] + sl / 2.0;
  qr = q[2] - sr / 2.0;
}

}  // namespace x64

namespace f32 {

__device__ __forceinline__ float ppm_iface(float qa, float qb, float sa, float sb, bool low_order) {
#pragma clang fp contract(fast)
  float v = 0.5f * (qa + qb);
  if (!low_order) v += (1.0f / 6.0f) * (sa - sb);
  return v;
}
__device__ __forceinline__ void ppm_mono(float q, float& A, float& B) {
#pragma clang fp contract(fast)
  const float tmp1 = B - A;
  const float c = tmp1 * (q - 0.5f * (B + A));
  const float t6 = (1.0f / 6.0f) * tmp1 * tmp1;
  if ((B - q) * (q - A) <= 0.0f) {
    B = A = q;
  } else if (c > t6) {
    A = 3.0f * q - 2.0f * B;
  } else if (-t6 > c) {
    B = 3.0f * q - 2.0f * A;
  }
}
__device__ __forceinline__ void ppm_face(const float* q, float& ql, float& qr, bool low_order, bool smooth) {
#pragma clang fp contract(fast)
  const float s1 = minmod_theta(q[2] - q[1], q[1] - q[0], 2.0f);
  const float s2 = minmod_theta(q[3] - q[2], q[2] - q[1], 2.0f);
  const float s3 = minmod_theta(q[4] - q[3], q[3] - q[2], 2.0f);
  const float s4 = minmod_theta(q[5] - q[4], q[4] - q[3], 2.0f);
  float a1 = ppm_iface(q[1], q[2], s1, s2, low_order);
  float a2 = ppm_iface(q[2], q[3], s2, s3, low_order);
  float a3 = ppm_iface(q[3], q[4], s3, s4, low_order);
  float a2r = a2;
  if (!smooth) {
    ppm_mono(q[2], a1, a2);
    ppm_mono(q[3], a2r, a3);
  }
  ql = a2;
  qr = a2r;
}
// ppm_pencil (above) in FP32
template <int N>
__device__ __forceinline__ void ppm_pencil(const float* q, float* ql, float* qr, bool low_order, bool smooth) {
  float s[N + 4], a[N + 3];
#pragma unroll
  for (int c = 0; c < N + 4; c++) s[c] = minmod_theta(q[c + 2] - q[c + 1], q[c + 1] - q[c], 2.0f);
#pragma unroll
  for (int c = 0; c < N + 3; c++) a[c] = ppm_iface(q[c + 1], q[c + 2], s[c], s[c + 1], low_order);
#pragma unroll
  for (int c = 0; c < N + 2; c++) {
    float A = a[c], B = a[c + 1];
    if (!smooth) ppm_mono(q[c + 2], A, B);
    if (c >= 1) qr[c - 1] = A;
    if (c <= N) ql[c] = B;
  }
}
__device__ __forceinline__ void minmod_face(const float* q, float& ql, float& qr, float theta) {
#pragma clang fp contract(fast)
  const float sl = minmod_theta(q[2] - q[1], q[1] - q[0], theta);
  const float sr = minmod_theta(q[3] - q[2], q[2] - q[1], theta);
  ql = q[1] + 0.5f * sl;
  qr = q[2] - 0.5f * sr;
}
// Two adjacent faces f and f+1 of one pencil, each with the values ppm_face gives it, from the
// seven cells f-3 .. f+3: the slopes of cells f-2 .. f+2, the interface values between f-2 and
// f+2 and the monotonisation of cell f (whose left edge is face f's qr and right edge face f+1's
// ql) are formed once for both faces (5 slopes, 4 interfaces, 3 monotonisations instead of 8, 6, 4)
__device__ __forceinline__ void ppm_face_pair(const float* q, float& ql0, float& qr0, float& ql1, float& qr1,
                                              bool low_order, bool smooth) {
#pragma clang fp contract(fast)
  const float s1 = minmod_theta(q[2] - q[1], q[1] - q[0], 2.0f);
  const float s2 = minmod_theta(q[3] - q[2], q[2] - q[1], 2.0f);
  const float s3 = minmod_theta(q[4] - q[3], q[3] - q[2], 2.0f);
  const float s4 = minmod_theta(q[5] - q[4], q[4] - q[3], 2.0f);
  const float s5 = minmod_theta(q[6] - q[5], q[5] - q[4], 2.0f);
  float a1 = ppm_iface(q[1], q[2], s1, s2, low_order);
  float a2 = ppm_iface(q[2], q[3], s2, s3, low_order);  // face f
  float a3 = ppm_iface(q[3], q[4], s3, s4, low_order);  // face f+1
  float a4 = ppm_iface(q[4], q[5], s4, s5, low_order);
  float a2r = a2, a3r = a3;
  if (!smooth) {
    ppm_mono(q[2], a1, a2);   // cell f-1: right edge = face f's ql
    ppm_mono(q[3], a2r, a3);  // cell f: left edge = face f's qr, right edge = face f+1's ql
    ppm_mono(q[4], a3r, a4);  // cell f+1: left edge = face f+1's qr
  }
  ql0 = a2;
  qr0 = a2r;
  ql1 = a3;
  qr1 = a3r;
}
// the same for minmod-theta from the five cells f-2 .. f+2 (the slope of cell f serves both faces)
__device__ __forceinline__ void minmod_face_pair(const float* q, float& ql0, float& qr0, float& ql1, float& qr1,
                                                 float theta) {
#pragma clang fp contract(fast)
  const float s1 = minmod_theta(q[2] - q[1], q[1] - q[0], theta);
  const float s2 = minmod_theta(q[3] - q[2], q[2] - q[1], theta);
  const float s3 = minmod_theta(q[4] - q[3], q[3] - q[2], theta);
  ql0 = q[1] + 0.5f * s1;
  qr0 = q[2] - 0.5f * s2;
  ql1 = q[2] + 0.5f * s2;
  qr1 = q[3] - 0.5f * s3;
}

}  // namespace f32
}  // namespace amrh
