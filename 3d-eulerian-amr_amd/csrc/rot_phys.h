// Rotating-frame binary-star state (the reference's ZTWD State, binary/state_ztwd.cpp) on the
// device, FP64 in the reference's operation order (build without FP contraction).
//
// Fields (binary/state_ztwd.cpp:15-28): rho, rho_acc, rho_don, s_R (cylindrical radial momentum),
// l_z (inertial-frame angular momentum), s_z, pot (rho * effective potential), E, tau (internal
// energy tracer), T, three species densities.  Velocities are in the frame rotating at Omega;
// positions X are cell or face centres.  The EOS is the Helmholtz one or an ideal gas
// p = rho T (1 + zbar)/abar, both behind one interface on the eos_t record.
#pragma once
#include <hip/hip_runtime.h>

#include "helmholtz.h"

namespace amrh {
namespace rot {

enum { D0 = 0, DACC = 1, DDON = 2, SR = 3, LZ = 4, SZ = 5, POT = 6, ET = 7, TAU = 8, TT = 9, SPEC = 10, NF = 13 };
constexpr int NSPEC = 3;
enum { EOS_IDEAL = 0, EOS_HELMHOLTZ = 1 };
constexpr double TMIN_IDEAL = 1.0e-20;  // the ideal gas is unit-free: its T floor only keeps T > 0

struct RotParams {
  double Omega, Omega0, fR, ftheta;  // Binary::Omega, Binary::Omega0, State::fR, State::ftheta
  double rho_floor;                  // binary/state_ztwd.cpp:31
  double gamma;                      // ideal-gas EOS
  double theta;                      // minmod theta
  double abar[NSPEC];                // binary/state_ztwd.cpp:28
};

__device__ __forceinline__ double rmax(double a, double b) { return a > b ? a : b; }

__device__ __forceinline__ double Rc(const double* X) { return sqrt(X[0] * X[0] + X[1] * X[1]); }
__device__ __forceinline__ double st(const RotParams& c, const double* u, const double* X) {
  return u[LZ] / Rc(X) - u[D0] * Rc(X) * c.Omega;
}
__device__ __forceinline__ double vx(const RotParams& c, const double* u, const double* X) {
  return (X[0] * u[SR] - X[1] * st(c, u, X)) / Rc(X) / u[D0];
}
__device__ __forceinline__ double vy(const RotParams& c, const double* u, const double* X) {
  return (X[1] * u[SR] + X[0] * st(c, u, X)) / Rc(X) / u[D0];
}
__device__ __forceinline__ double vz(const double* u) { return u[SZ] / u[D0]; }
__device__ __forceinline__ double dvx(const RotParams& c, const double* X) { return -X[1] * (c.Omega - c.Omega0); }
__device__ __forceinline__ double dvy(const RotParams& c, const double* X) { return X[0] * (c.Omega - c.Omega0); }
__device__ __forceinline__ double ek(const RotParams& c, const double* u, const double* X) {
  const double t = st(c, u, X);
  return 0.5 * (u[SR] * u[SR] + t * t + u[SZ] * u[SZ]) / u[D0];
}
// binary/state_ztwd.cpp:140-156 (ZTWD branch)
__device__ __forceinline__ double ei(const RotParams& c, const double* u, const double* X) {
  const double e = u[ET] - ek(c, u, X);
  if (e < u[ET] * 0.001) return u[TAU];
  return e;
}
// A() (binary/state_ztwd.cpp:33-41)
__device__ __forceinline__ double Abar(const RotParams& c, const double* u) {
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < NSPEC; i++) sum += u[SPEC + i] / c.abar[i];
  sum = u[D0] / sum;
  return sum;
}

template <int EOS>
__device__ __forceinline__ void eos_eval(const helm::HelmDev& H, const RotParams& c, helm::Eos& q) {
  if constexpr (EOS == EOS_HELMHOLTZ) {
    helm::eos(H, q);
  } else {
    const double T = q.T > TMIN_IDEAL ? q.T : TMIN_IDEAL;
    const double mu_inv = (1.0 + q.zbar) / q.abar;
    q.p = q.rho * T * mu_inv;
    q.cv = mu_inv / (c.gamma - 1.0);
    q.e = q.cv * T;
    q.cs = sqrt(c.gamma * q.p / q.rho);
    q.s = 0.0;
    q.dsdt = 0.0;
  }
}

// helmholtz_compute_T (helmholtz/helmeos.cpp:436-486); the ideal gas runs the same iteration.
// SEP: the Helmholtz iteration on the separable FP64 e and c_v (the mixed form)
template <int EOS, bool SEP = false>
__device__ __forceinline__ void eos_compute_T(const helm::HelmDev& H, const RotParams& c, helm::Eos& q) {
  if constexpr (EOS == EOS_HELMHOLTZ && SEP) {
    helm::compute_T_sep(H, q);
  } else if constexpr (EOS == EOS_HELMHOLTZ) {
    helm::compute_T(H, q);
  } else {
    double T0 = q.T;
    const double e0 = q.e;
    q.T = TMIN_IDEAL;
    eos_eval<EOS>(H, c, q);
    q.T = T0;
    double a = 1.0;
    if (q.e >= e0) {
      q.T = TMIN_IDEAL;
      return;
    }
    int iters = 0;
    double f;
    do {
      T0 = q.T > TMIN_IDEAL ? q.T : TMIN_IDEAL;
      eos_eval<EOS>(H, c, q);
      f = q.e - e0;
      const double dT = -f / q.cv;
      q.T += a * dT;
      const double lo = (T0 + TMIN_IDEAL) / 2.0;
      q.T = q.T > lo ? q.T : lo;
      iters++;
      if (iters % 10 == 0) a *= 0.9;
      if (q.T < TMIN_IDEAL) {
        atomicOr(H.err, helm::ERR_TMIN);
        return;
      }
    } while (fabs(f / e0) >= H.toler && iters < 100);
    eos_eval<EOS>(H, c, q);
    if (iters >= 100) atomicOr(H.err, helm::ERR_NEWTON);
  }
}

// State::compute_eos (binary/state_ztwd.cpp:168-176): pressure and sound speed of a state
// the same from its inputs: density, temperature and Abar of the state
template <int EOS>
__device__ __forceinline__ void eos_p_cs_in(const helm::HelmDev& H, const RotParams& c, double rho, double T,
                                            double abar, double& p, double& cs) {
  helm::Eos q{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  q.rho = rho;
  q.T = T;
  q.abar = abar;
  q.zbar = abar / 2.0;
  eos_eval<EOS>(H, c, q);
  p = q.p;
  cs = q.cs;
}
template <int EOS>
__device__ __forceinline__ void eos_p_cs(const helm::HelmDev& H, const RotParams& c, const double* u, double& p,
                                         double& cs) {
  eos_p_cs_in<EOS>(H, c, u[D0], u[TT], Abar(c, u), p, cs);
}

// binary/state_ztwd.cpp:581-599
__device__ __forceinline__ void to_prim(const RotParams& c, double* u, const double* X) {
#pragma unroll
  for (int i = 0; i < NSPEC; i++) u[SPEC + i] /= u[D0];
  u[DDON] /= u[D0];
  u[DACC] /= u[D0];
  u[ET] -= ek(c, u, X);
  u[SR] /= u[D0];
  u[LZ] /= u[D0] * Rc(X);
  u[LZ] -= c.Omega * Rc(X);
  u[SZ] /= u[D0];
  u[POT] /= u[D0];
}

// binary/state_ztwd.cpp:556-579
__device__ __forceinline__ void to_con(const RotParams& c, double* u, const double* X) {
  u[SZ] *= u[D0];
  u[POT] *= u[D0];
  u[SR] *= u[D0];
  u[LZ] += c.Omega * Rc(X);
  u[LZ] *= u[D0] * Rc(X);
  u[ET] += ek(c, u, X);
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < NSPEC; i++) sum += u[SPEC + i];
#pragma unroll
  for (int i = 0; i < NSPEC; i++) u[SPEC + i] /= sum;
  sum = 0.0;
  sum += u[DDON];
  sum += u[DACC];
  u[DDON] /= sum;
  u[DACC] /= sum;
#pragma unroll
  for (int i = 0; i < NSPEC; i++) u[SPEC + i] *= u[D0];
  u[DDON] *= u[D0];
  u[DACC] *= u[D0];
}

// binary/state_ztwd.cpp:360-396 with the eigenvalue of :176-209; returns |v_d| + c_s
// the flux of state u given its pressure p and sound speed cs (State::x/y/z_flux after
// compute_eos); returns the wavespeed |v_d| + cs
template <int D>
__device__ __forceinline__ double flux_pc(const RotParams& c, const double* u, const double* X, double p, double cs,
                                          double* F) {
  const double v = D == 0 ? vx(c, u, X) : (D == 1 ? vy(c, u, X) : vz(u));
#pragma unroll
  for (int i = 0; i < NF; i++) F[i] = v * u[i];
  if constexpr (D == 0) {
    F[SR] += X[0] * p / Rc(X);
    F[LZ] -= X[1] * p;
    F[ET] += (v + dvx(c, X)) * p;
  } else if constexpr (D == 1) {
    F[SR] += X[1] * p / Rc(X);
    F[LZ] += X[0] * p;
    F[ET] += (v + dvy(c, X)) * p;
  } else {
    F[SZ] += p;
    F[ET] += v * p;
  }
  return fabs(v) + cs;
}
template <int EOS, int D>
__device__ __forceinline__ double flux(const helm::HelmDev& H, const RotParams& c, const double* u, const double* X,
                                       double* F) {
  double p, cs;
  eos_p_cs<EOS>(H, c, u, p, cs);
  return flux_pc<D>(c, u, X, p, cs, F);
}

// binary/state_ztwd.cpp:482-554 (SEP: the mixed form's Newton solve, eos_compute_T)
template <int EOS, bool SEP = false>
__device__ __forceinline__ void floor_state(const helm::HelmDev& H, const RotParams& c, double* u, const double* X) {
  if (u[D0] < c.rho_floor) {
    u[D0] = c.rho_floor;
    u[DACC] = c.rho_floor * 0.5;
    u[DDON] = c.rho_floor * 0.5;
    u[SPEC + 0] = c.rho_floor;
#pragma unroll
    for (int i = 1; i < NSPEC; i++) u[SPEC + i] = 0.0;
  } else {
#pragma unroll
    for (int i = 0; i < NSPEC; i++) u[SPEC + i] = rmax(u[SPEC + i], 0.0);
    u[DDON] = rmax(u[DDON], c.rho_floor / 2.0);
    u[DACC] = rmax(u[DACC], c.rho_floor / 2.0);
  }
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < NSPEC; i++) sum += u[SPEC + i];
  if (sum != 0.0) {
#pragma unroll
    for (int i = 0; i < NSPEC; i++) u[SPEC + i] *= u[D0] / sum;
  } else {
    u[SPEC] = u[D0];
#pragma unroll
    for (int i = 1; i < NSPEC; i++) u[SPEC + i] = 0.0;
  }
  sum = 0.0;
  sum += u[DDON];
  sum += u[DACC];
  if (sum != 0.0) {
    u[DDON] *= u[D0] / sum;
    u[DACC] *= u[D0] / sum;
  } else {
    u[DDON] = u[D0] / 2.0;
    u[DACC] = u[D0] / 2.0;
  }
  const double d = u[D0];
  const double ue = ei(c, u, X);
  helm::Eos q{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  q.abar = Abar(c, u);
  q.zbar = Abar(c, u) / 2.0;
  q.rho = d;
  q.e = ue / d;
  q.T = u[TT];
  eos_compute_T<EOS, SEP>(H, c, q);
  u[TT] = q.T;
  if (ue > 0.1 * u[ET]) u[TAU] = ue;
}

// binary/state_ztwd.cpp:398-443 on node face nf (XL..ZU); the boundary object's face is the
// inverted one, so node face XL takes the XU branch (v_x > 0 is inflow there)
__device__ __forceinline__ void enforce_outflow(const RotParams& c, double* u, const double* X, int nf) {
  u[ET] = ei(c, u, X);
  double v0;
  if (nf == 0 || nf == 1) {
    const double v = vx(c, u, X);
    if (nf == 0 ? v > 0.0 : v < 0.0) {
      v0 = vy(c, u, X);
      u[LZ] = (X[0] * (v0 + X[0] * c.Omega)) * u[D0];
      u[SR] = X[1] * v0 * u[D0] / Rc(X);
    }
  } else if (nf == 2 || nf == 3) {
    const double v = vy(c, u, X);
    if (nf == 2 ? v > 0.0 : v < 0.0) {
      v0 = vx(c, u, X);
      u[LZ] = (-X[1] * (v0 - X[1] * c.Omega)) * u[D0];
      u[SR] = X[0] * v0 * u[D0] / Rc(X);
    }
  } else if (nf == 4) {
    if (u[SZ] > 0.0) u[SZ] = 0.0;
  } else {
    if (u[SZ] < 0.0) u[SZ] = 0.0;
  }
  const double t = st(c, u, X);
  u[ET] += 0.5 * (u[SR] * u[SR] + t * t + u[SZ] * u[SZ]) / u[D0];
}

// ------------------------------------------------------------------------------------------
// The mixed form's face states: reconstructed primitives back to conserved, EOS pressure and
// sound speed (the Helmholtz one by helm::p_cs32), and the directional flux in FP32.  Same formulas
// as above; not bit-matched to the reference.
// ------------------------------------------------------------------------------------------
namespace sp {

struct RotF {
  float Omega, Omega0, gamma, gm1inv, rho_floor;
  float abar_inv[NSPEC];
};

__host__ __device__ inline RotF make(const RotParams& c) {
  RotF f;
  f.Omega = (float)c.Omega;
  f.Omega0 = (float)c.Omega0;
  f.gamma = (float)c.gamma;
  f.gm1inv = (float)(1.0 / (c.gamma - 1.0));
  f.rho_floor = (float)c.rho_floor;
  for (int i = 0; i < NSPEC; i++) f.abar_inv[i] = (float)(1.0 / c.abar[i]);
  return f;
}

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float Rc(const float* X) { return __builtin_amdgcn_sqrtf(X[0] * X[0] + X[1] * X[1]); }
__device__ __forceinline__ float st(const RotF& c, const float* u, float R) { return u[LZ] * rcp(R) - u[D0] * R * c.Omega; }

__device__ __forceinline__ void to_con(const RotF& c, float* u, const float* X) {
#pragma clang fp contract(fast)
  const float R = Rc(X);
  u[SZ] *= u[D0];
  u[POT] *= u[D0];
  u[SR] *= u[D0];
  u[LZ] = (u[LZ] + c.Omega * R) * (u[D0] * R);
  const float t = st(c, u, R);
  u[ET] += 0.5f * (u[SR] * u[SR] + t * t + u[SZ] * u[SZ]) * rcp(u[D0]);
  const float s1 = rcp(u[SPEC] + u[SPEC + 1] + u[SPEC + 2]) * u[D0];
#pragma unroll
  for (int i = 0; i < NSPEC; i++) u[SPEC + i] *= s1;
  const float s2 = rcp(u[DDON] + u[DACC]) * u[D0];
  u[DDON] *= s2;
  u[DACC] *= s2;
}

// velocity along d of a conserved state
template <int D>
__device__ __forceinline__ float vel(const RotF& c, const float* u, const float* X, float R) {
#pragma clang fp contract(fast)
  const float ir = rcp(R * u[D0]);
  if constexpr (D == 0) return (X[0] * u[SR] - X[1] * st(c, u, R)) * ir;
  if constexpr (D == 1) return (X[1] * u[SR] + X[0] * st(c, u, R)) * ir;
  return u[SZ] * rcp(u[D0]);
}

template <int EOS>
__device__ __forceinline__ void p_cs(const helm::HelmDev& H, const RotF& c, const float* u, float& p, float& cs) {
#pragma clang fp contract(fast)
  const float abar = u[D0] * rcp(u[SPEC] * c.abar_inv[0] + u[SPEC + 1] * c.abar_inv[1] + u[SPEC + 2] * c.abar_inv[2]);
  const float zbar = 0.5f * abar;
  if constexpr (EOS == EOS_HELMHOLTZ) {
    helm::p_cs32(H, u[D0], u[TT], abar, zbar, p, cs);
  } else {
    const float T = u[TT] > (float)TMIN_IDEAL ? u[TT] : (float)TMIN_IDEAL;
    p = u[D0] * T * ((1.0f + zbar) * rcp(abar));
    cs = __builtin_amdgcn_sqrtf(c.gamma * p * rcp(u[D0]));
  }
}

// flux of a conserved face state along D; returns |v_d| + c_s
template <int EOS, int D>
__device__ __forceinline__ float flux(const helm::HelmDev& H, const RotF& c, const float* u, const float* X, float* F) {
#pragma clang fp contract(fast)
  float p, cs;
  p_cs<EOS>(H, c, u, p, cs);
  const float R = Rc(X);
  const float v = vel<D>(c, u, X, R);
#pragma unroll
  for (int i = 0; i < NF; i++) F[i] = v * u[i];
  if constexpr (D == 0) {
    F[SR] += X[0] * p * rcp(R);
    F[LZ] -= X[1] * p;
    F[ET] += (v - X[1] * (c.Omega - c.Omega0)) * p;
  } else if constexpr (D == 1) {
    F[SR] += X[1] * p * rcp(R);
    F[LZ] += X[0] * p;
    F[ET] += (v + X[0] * (c.Omega - c.Omega0)) * p;
  } else {
    F[SZ] += p;
    F[ET] += v * p;
  }
  return fabsf(v) + cs;
}

}  // namespace sp
}  // namespace rot
}  // namespace amrh
