// Per-cell operations of the ideal-gas state shared by the sweep and block kernels: the compute
// type of each precision form, conserved -> primitive conversion and the floor of the mixed form.
#pragma once
#include <hip/hip_runtime.h>

#include "amrh_layout.h"
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

// outflow ghost of boundary cell u across node face `face` (the boundary object's face is the
// node's inverted: a lower node face removes v > 0; the z test uses sz, euler/state.cpp:45-84)
__device__ __forceinline__ void outflow_fix(double* u, int face) {
  const int d = face >> 1;
  const bool upper = face & 1;
  const double v = (d == 2) ? u[SZ] : u[SX + d] / u[RHO];
  const bool inflow = upper ? (v < 0.0) : (v > 0.0);
  if (inflow) {
    const double s = u[SX + d];
    u[ET] = u[ET] - 0.5 * s * s / u[RHO];
    u[SX + d] = 0.0;
  }
}

constexpr int kMirror = 6;

// a fetched boundary cell -> the ghost it stands for: none (-1), the outflow fix of a node face,
// or the z mirror (grid/grid_phys_bound.cpp:27-54, euler/state.cpp:94-96)
__device__ __forceinline__ void ghost_post(double* u, int post) {
  if (post == kMirror)
    u[SZ] = -u[SZ];
  else if (post >= 0)
    outflow_fix(u, post);
}

// Self-ghosting outflow blocks: the interior element ghost-slab element e (>= N^3 of the
// halo-slab record) copies, and how it becomes the ghost (post: the node face to fix, or kMirror)
template <int N, int G>
__device__ __forceinline__ int outflow_source(int e, bool zr, int& post) {
  constexpr int N3 = N * N * N, S = G * N * N;
  const int s = (e - N3) / S, r = (e - N3) % S;
  int i, j, k;
  post = s;
  if (s < 2) {
    j = (r / G) % N;
    k = r / (G * N);
    i = s == XL ? 0 : N - 1;
  } else if (s < 4) {
    i = r % N;
    k = r / (G * N);
    j = s == YL ? 0 : N - 1;
  } else {
    i = r % N;
    j = (r / N) % N;
    const int t = r / (N * N);
    if (s == ZL && zr) {
      k = G - 1 - t;  // padded z = t holds padded 2G-1-t
      post = kMirror;
    } else {
      k = s == ZL ? 0 : N - 1;
    }
  }
  return (k * N + j) * N + i;
}

}  // namespace amrh
