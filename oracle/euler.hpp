// ORACLE -- test infrastructure only.  Nothing in the product path may include, link or call
// this file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, and only
// as the checker.
//
// FP64 restatement of the ideal-gas ("EULER") per-cell physics of 3d-eulerian-amr, operation by
// operation, so that an FP64 GPU kernel compiled without FP contraction can be compared with it
// bit for bit (except for libm pow, which may differ by one ulp).  Field order and constants:
//   euler/state.cpp:9-17   rho, sx, sy, sz, E, tau;  gamma = 5/3, rho_floor = ei_floor = 1e-20
//   euler/euler_defs.h:20  MINMOD_THETA = 1.3
#pragma once
#include <cmath>

namespace orc {

enum { RHO = 0, SX = 1, SY = 2, SZ = 3, ET = 4, TAU = 5, NF_EULER = 6 };

struct Phys {
  double gamma;      // euler/state.cpp:9
  double rho_floor;  // euler/state.cpp:10
  double ei_floor;   // euler/state.cpp:11
  double theta;      // euler/euler_defs.h:20
  double cfl;        // euler/euler_defs.h:46 GRID_CFL_FACTOR
};

// real.h:20-22 / :38-40 -- note the strict comparisons (they decide which operand a tie returns)
inline double rmax(double a, double b) { return a > b ? a : b; }
inline double rmin(double a, double b) { return a > b ? b : a; }
// real.h:50-58
inline double sgn(double a) { return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0); }
// real.h:60-62
inline double minmod2(double a, double b) { return 0.5 * (sgn(a) + sgn(b)) * rmin(std::fabs(a), std::fabs(b)); }
// real.h:68-70
inline double minmod3(double a, double b, double c) { return minmod2(a, minmod2(b, c)); }
// real.h:64-66, vector.h:193-201
inline double minmod_theta(double a, double b, double th) { return minmod3(th * a, th * b, 0.5 * (a + b)); }

// euler/state.cpp:148-150  ek = 0.5*(sx^2+sy^2+sz^2)/rho, evaluated left to right
inline double ek(const double* u) {
  return 0.5 * (u[SX] * u[SX] + u[SY] * u[SY] + u[SZ] * u[SZ]) / u[RHO];
}

// euler/state.cpp:156-163  dual-energy pressure
inline double pg(const Phys& P, const double* u) {
  const double ei0 = u[ET] - ek(u);
  if (ei0 < 0.001 * u[ET]) {
    return (P.gamma - 1.0) * std::pow(u[TAU], P.gamma);
  } else {
    return (P.gamma - 1.0) * ei0;
  }
}

// euler/state.cpp:165-169
inline double cs(const Phys& P, const double* u) { return std::sqrt(P.gamma * pg(P, u) / u[RHO]); }

// euler/state.cpp:171-181  |v_d| + c_s, d = 0,1,2 for x,y,z
inline double max_abs_eigen(const Phys& P, const double* u, int d) {
  return std::fabs(u[SX + d] / u[RHO]) + cs(P, u);
}

// euler/state.cpp:183-220  F = U*v_d; F[s_d] += p; F[E] += v_d*p
inline void flux(const Phys& P, const double* u, int d, double* f, int nf) {
  const double v = u[SX + d] / u[RHO];
  const double p = pg(P, u);
  for (int l = 0; l < nf; l++) f[l] = u[l] * v;
  f[SX + d] += p;
  f[ET] += v * p;
}

// euler/state.cpp:246-251  (ek evaluated before any field is changed)
inline void to_prim(double* u) {
  u[ET] -= ek(u);
  u[SX] /= u[RHO];
  u[SY] /= u[RHO];
  u[SZ] /= u[RHO];
}

// euler/state.cpp:253-258  (ek evaluated after the momenta are restored)
inline void to_con(double* u) {
  u[SX] *= u[RHO];
  u[SY] *= u[RHO];
  u[SZ] *= u[RHO];
  u[ET] += ek(u);
}

// euler/state.cpp:86-92
inline void floor_state(const Phys& P, double* u) {
  u[RHO] = rmax(u[RHO], P.rho_floor);
  const double ei0 = u[ET] - ek(u);
  if (ei0 > 0.1 * u[ET]) {
    u[TAU] = std::pow(rmax(ei0, P.ei_floor), 1.0 / P.gamma);
  }
}

// euler/state.cpp:45-84.  `face` is the face of the boundary object (the node's own face
// inverted, grid/grid_node.cpp:1383-1391 and :284-290): XL=0 XU=1 YL=2 YU=3 ZL=4 ZU=5
// (oct_node/oct_face.h:6-13).  Inflow momentum through the boundary is removed together with
// its kinetic energy.
inline void enforce_outflow(double* u, int face) {
  const int d = face >> 1;
  const bool upper = (face & 1) != 0;
  const double v = (d == 2) ? u[SZ] : u[SX + d] / u[RHO];  // ZU/ZL test sz itself (state.cpp:71,77)
  const bool inflow = upper ? (v > 0.0) : (v < 0.0);
  if (inflow) {
    const double s = u[SX + d];
    u[ET] = u[ET] - 0.5 * s * s / u[RHO];
    u[SX + d] = 0.0;
  }
}

}  // namespace orc
