// Ideal-gas ("EULER") physics on the device, two forms.
//
// Exact64: FP64, every operation in the reference's order (euler/state.cpp:148-258,
//   reconstruct.cpp:60-71, real.h:20-70, grid/grid.cpp:238-239), for a build without FP
//   contraction: matches the FP64 CPU path bit for bit except where libm/OCML pow differ by an ulp.
// Mixed: reconstruction, fluxes and wavespeeds in FP32, starting from primitives that were formed
//   in FP64.  Pressure is taken from the reconstructed internal energy directly (the reference
//   forms E = eps + ek and then E - ek, which is the same number in exact arithmetic but cancels
//   catastrophically in FP32 at high Mach number); the dual-energy test is the same inequality.
//   Divergence accumulation, the RK combination and the floor stay FP64 in the kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace amrh {

enum { RHO = 0, SX = 1, SY = 2, SZ = 3, ET = 4, TAU = 5, NFE = 6 };

struct PhysParams {
  double gamma, rho_floor, ei_floor, theta, cfl;
};

// ------------------------------------------------------------------------------------------
// Exact FP64
// ------------------------------------------------------------------------------------------
namespace x64 {

__device__ __forceinline__ double rmax(double a, double b) { return a > b ? a : b; }   // real.h:20
__device__ __forceinline__ double rmin(double a, double b) { return a > b ? b : a; }   // real.h:38
__device__ __forceinline__ double sgn(double a) { return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0); }
__device__ __forceinline__ double minmod2(double a, double b) {                         // real.h:60
  return 0.5 * (sgn(a) + sgn(b)) * rmin(fabs(a), fabs(b));
}
__device__ __forceinline__ double minmod_theta(double a, double b, double th) {         // real.h:64
  return minmod2(th * a, minmod2(th * b, 0.5 * (a + b)));
}
__device__ __forceinline__ double ek(const double* u) {                                // state.cpp:148
  return 0.5 * (u[SX] * u[SX] + u[SY] * u[SY] + u[SZ] * u[SZ]) / u[RHO];
}
__device__ __forceinline__ double pg(const PhysParams& P, const double* u) {           // state.cpp:156
  const double ei0 = u[ET] - ek(u);
  if (ei0 < 0.001 * u[ET]) return (P.gamma - 1.0) * pow(u[TAU], P.gamma);
  return (P.gamma - 1.0) * ei0;
}
__device__ __forceinline__ void to_prim(double* u) {                                   // state.cpp:246
  u[ET] -= ek(u);
  u[SX] /= u[RHO];
  u[SY] /= u[RHO];
  u[SZ] /= u[RHO];
}
__device__ __forceinline__ void to_con(double* u) {                                    // state.cpp:253
  u[SX] *= u[RHO];
  u[SY] *= u[RHO];
  u[SZ] *= u[RHO];
  u[ET] += ek(u);
}
__device__ __forceinline__ void floor_state(const PhysParams& P, double* u) {          // state.cpp:86
  u[RHO] = rmax(u[RHO], P.rho_floor);
  const double ei0 = u[ET] - ek(u);
  if (ei0 > 0.1 * u[ET]) u[TAU] = pow(rmax(ei0, P.ei_floor), 1.0 / P.gamma);
}
// slope of the middle cell b from neighbours a, c (reconstruct.cpp:81-85)
__device__ __forceinline__ void slope(const double* a, const double* b, const double* c, double* s,
                                      double th) {
#pragma unroll
  for (int l = 0; l < NFE; l++) s[l] = minmod_theta(c[l] - b[l], b[l] - a[l], th);
}
// KT face flux in direction D from the primitive states and slopes of the two adjacent cells
// (reconstruct.cpp:86-89, grid/grid.cpp:233-239, state.cpp:171-220).  eig = max(eL, eR).
// KT face flux from the two reconstructed primitive face states (converted in place)
template <int D>
__device__ __forceinline__ void kt_flux(const PhysParams& P, double* ql, double* qr, double* F, double& eig) {
  to_con(ql);
  to_con(qr);
  const double vl = ql[SX + D] / ql[RHO], vr = qr[SX + D] / qr[RHO];
  const double pgl = pg(P, ql), pgr = pg(P, qr);
  const double el = fabs(vl) + sqrt(P.gamma * pgl / ql[RHO]);
  const double er = fabs(vr) + sqrt(P.gamma * pgr / qr[RHO]);
  const double a = rmax(el, er);
  eig = a;
#pragma unroll
  for (int l = 0; l < NFE; l++) {
    double fl = ql[l] * vl, fr = qr[l] * vr;
    if (l == SX + D) {
      fl += pgl;
      fr += pgr;
    }
    if (l == ET) {
      fl += vl * pgl;
      fr += vr * pgr;
    }
    F[l] = ((fl + fr) - (qr[l] - ql[l]) * a) * 0.5;
  }
}
template <int D>
__device__ __forceinline__ void kt_face(const PhysParams& P, const double* pl, const double* sl,
                                        const double* pr, const double* sr, double* F, double& eig) {
  double ql[NFE], qr[NFE];
#pragma unroll
  for (int l = 0; l < NFE; l++) {
    ql[l] = pl[l] + sl[l] * 0.5;
    qr[l] = pr[l] - sr[l] * 0.5;
  }
  kt_flux<D>(P, ql, qr, F, eig);
}
// max(|v_d| + c_s) of the two reconstructed primitive face states (converted in place)
template <int D>
__device__ __forceinline__ double eig_q(const PhysParams& P, double* ql, double* qr) {
  to_con(ql);
  to_con(qr);
  const double el = fabs(ql[SX + D] / ql[RHO]) + sqrt(P.gamma * pg(P, ql) / ql[RHO]);
  const double er = fabs(qr[SX + D] / qr[RHO]) + sqrt(P.gamma * pg(P, qr) / qr[RHO]);
  return rmax(el, er);
}
template <int D>
__device__ __forceinline__ double face_eig(const PhysParams& P, const double* pl, const double* sl,
                                           const double* pr, const double* sr) {
  double ql[NFE], qr[NFE];
#pragma unroll
  for (int l = 0; l < NFE; l++) {
    ql[l] = pl[l] + sl[l] * 0.5;
    qr[l] = pr[l] - sr[l] * 0.5;
  }
  return eig_q<D>(P, ql, qr);
}

}  // namespace x64

// ------------------------------------------------------------------------------------------
// Mixed: FP32 reconstruction / flux / wavespeed
// ------------------------------------------------------------------------------------------
namespace f32 {

struct PhysF {
  float gamma, gm1, theta;
};

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// x^y for x > 0 via the hardware log2/exp2 (about 1-2 ulp in FP32)
__device__ __forceinline__ float fast_pow(float x, float y) {
#pragma clang fp contract(fast)
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// minmod(theta*a, theta*b, (a+b)/2): nonzero only when all three agree in sign, and then
// sign(a) * theta * min(|a|, |b|, |a+b| / (2 theta)) -- one multiply by theta instead of two
__device__ __forceinline__ float minmod_theta(float a, float b, float th) {
#pragma clang fp contract(fast)
  const float m = th * fminf(fminf(fabsf(a), fabsf(b)), fabsf((a + b) * (0.5f / th)));
  return (a * b > 0.0f) ? copysignf(m, a) : 0.0f;
}
__device__ __forceinline__ void slope(const float* a, const float* b, const float* c, float* s, float th) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int l = 0; l < NFE; l++) s[l] = minmod_theta(c[l] - b[l], b[l] - a[l], th);
}

// total energy, pressure and signal speed |v_d| + c_s of one reconstructed primitive face state
// (the dual-energy branch is a real branch, so whole wavefronts that do not need it skip the
// transcendental pair)
template <int D>
__device__ __forceinline__ void face_side(const PhysF& P, const float* q, float& E, float& p, float& e) {
#pragma clang fp contract(fast)
  const float v2 = q[SX] * q[SX] + q[SY] * q[SY] + q[SZ] * q[SZ];
  E = q[ET] + 0.5f * q[RHO] * v2;
  p = P.gm1 * q[ET];
  if (__builtin_expect(q[ET] < 0.001f * E, 0)) p = P.gm1 * fast_pow(q[TAU], P.gamma);
  e = fabsf(q[SX + D]) + fast_sqrt(P.gamma * p * fast_rcp(q[RHO]));
}

// KT face flux from the two reconstructed primitive face states.  The combination
// F = (f_L + f_R)/2 - a (U_R - U_L)/2 with f = U v_d + pressure terms is regrouped as
// F = U_L (v_L + a)/2 + U_R (v_R - a)/2 + (pressure terms)/2, and U_L (v_L + a)/2 for the momenta as
// (rho_L (v_L + a)/2) v_L: two operations per field instead of six (same value in exact arithmetic)
template <int D>
__device__ __forceinline__ void kt_flux(const PhysF& P, const float* ql, const float* qr, float* F, float& eig) {
#pragma clang fp contract(fast)
  float EL, pL, eL, ER, pR, eR;
  face_side<D>(P, ql, EL, pL, eL);
  face_side<D>(P, qr, ER, pR, eR);
  const float a = fmaxf(eL, eR);
  eig = a;
  const float vL = ql[SX + D], vR = qr[SX + D];
  const float cL = 0.5f * (vL + a), cR = 0.5f * (vR - a);
  const float wL = ql[RHO] * cL, wR = qr[RHO] * cR;
  F[RHO] = wL + wR;
  F[SX] = wL * ql[SX] + wR * qr[SX];
  F[SY] = wL * ql[SY] + wR * qr[SY];
  F[SZ] = wL * ql[SZ] + wR * qr[SZ];
  F[SX + D] += 0.5f * (pL + pR);
  F[ET] = EL * cL + ER * cR + 0.5f * (vL * pL + vR * pR);
  F[TAU] = ql[TAU] * cL + qr[TAU] * cR;
}
template <int D>
__device__ __forceinline__ void kt_face(const PhysF& P, const float* pl, const float* sl, const float* pr,
                                        const float* sr, float* F, float& eig) {
#pragma clang fp contract(fast)
  float ql[NFE], qr[NFE];
#pragma unroll
  for (int l = 0; l < NFE; l++) {
    ql[l] = pl[l] + 0.5f * sl[l];
    qr[l] = pr[l] - 0.5f * sr[l];
  }
  kt_flux<D>(P, ql, qr, F, eig);
}
template <int D>
__device__ __forceinline__ float eig_q(const PhysF& P, const float* ql, const float* qr) {
#pragma clang fp contract(fast)
  float E, p, el, er;
  face_side<D>(P, ql, E, p, el);
  face_side<D>(P, qr, E, p, er);
  return fmaxf(el, er);
}

template <int D>
__device__ __forceinline__ float face_eig(const PhysF& P, const float* pl, const float* sl, const float* pr,
                                          const float* sr) {
#pragma clang fp contract(fast)
  float E, p, el, er, q[NFE];
#pragma unroll
  for (int l = 0; l < NFE; l++) q[l] = pl[l] + 0.5f * sl[l];
  face_side<D>(P, q, E, p, el);
#pragma unroll
  for (int l = 0; l < NFE; l++) q[l] = pr[l] - 0.5f * sr[l];
  face_side<D>(P, q, E, p, er);
  return fmaxf(el, er);
}

// FP64 helpers of the mixed form's conversion and update (not bit-matched to the reference)
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);  // ~2^-26 relative
  const double e = fma(-x, r, 1.0);
  return fma(r, e, r);  // one Newton step: ~2^-52
}

// libm pow, out of line: the kernels that inline pow_inv_gamma would otherwise hold pow's FP64
// polynomial constants in registers for the whole block (they spilled them), although this path
// runs only for gamma != 5/3
__device__ __attribute__((noinline)) inline double pow_out_of_line(double x, double y) { return pow(x, y); }

// x^(1/gamma) for x >= 1e-20: FP32 estimate, then for gamma = 5/3 one FP64 Newton step on
// y^5 = x^3 (relative error ~1e-13); other gamma fall back to pow.
__device__ __forceinline__ double pow_inv_gamma(double x, double gamma) {
  if (gamma != 5.0 / 3.0) return pow_out_of_line(x, 1.0 / gamma);
  const float xf = (float)x;
  const double y0 = (double)__builtin_amdgcn_exp2f(0.6f * __builtin_amdgcn_logf(xf));
  const double y2 = y0 * y0;
  const double x3 = x * x * x;
  return fma(0.8, y0, 0.2 * x3 * rcp64(y2 * y2));
}

}  // namespace f32
}  // namespace amrh
