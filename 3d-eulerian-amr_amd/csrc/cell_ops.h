// Per-cell operations of the ideal-gas state shared by the sweep and block kernels: the compute
// type of each precision form, conserved -> primitive conversion and the floor of the mixed form.
#pragma once
#include <hip/hip_runtime.h>

#include "euler_phys.h"

namespace amrh {

template <bool MIXED>
struct Prec {
  using T = double;
};
template <>
struct Prec<true> {
  using T = float;
};

// conserved FP64 -> primitive in the compute type
template <bool MIXED>
__device__ __forceinline__ void make_prim(const double* u, typename Prec<MIXED>::T* p) {
  if constexpr (!MIXED) {
#pragma unroll
    for (int l = 0; l < NFE; l++) p[l] = u[l];
    x64::to_prim(p);
  } else {
    const double inv = f32::rcp64(u[RHO]);
    const double vx = u[SX] * inv, vy = u[SY] * inv, vz = u[SZ] * inv;
    const double ekin = 0.5 * (u[SX] * vx + u[SY] * vy + u[SZ] * vz);
    p[RHO] = (float)u[RHO];
    p[SX] = (float)vx;
    p[SY] = (float)vy;
    p[SZ] = (float)vz;
    p[ET] = (float)(u[ET] - ekin);
    p[TAU] = (float)u[TAU];
  }
}

// State::floor of the mixed form (euler/state.cpp:86-92): same tests in FP64, the entropy tracer
// from pow_inv_gamma instead of libm pow
__device__ __forceinline__ void floor_mixed(const PhysParams& P, double* u) {
  u[RHO] = x64::rmax(u[RHO], P.rho_floor);
  const double ek = 0.5 * (u[SX] * u[SX] + u[SY] * u[SY] + u[SZ] * u[SZ]) * f32::rcp64(u[RHO]);
  const double ei0 = u[ET] - ek;
  if (ei0 > 0.1 * u[ET]) u[TAU] = f32::pow_inv_gamma(x64::rmax(ei0, P.ei_floor), P.gamma);
}

}  // namespace amrh
