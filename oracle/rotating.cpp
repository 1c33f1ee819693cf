// ORACLE -- test infrastructure only (tests/, bench.py CPU baseline).  Never linked by the product.
//
// FP64 restatement of the rotating-frame binary-star state of 3d-eulerian-amr (the ZTWD State,
// binary/state_ztwd.cpp) and of the substep around it on a batch of independent 8^3 blocks:
//   fields (binary/state_ztwd.cpp:15-28)   rho, rho_acc, rho_don, s_R, l_z, s_z, pot, E, tau, T,
//                                          3 species densities
//   to_prim / to_con (:560-599), st, vx, vy, vz, dvx, dvy, ek, ei (:121-166, :241-267),
//   A() / Z() (:33-44), x/y/z_flux (:360-396), max_abs_*_eigen (:176-209),
//   source (:347-358), gravity_source / internal_energy_source (:269-288), floor (:482-554),
//   enforce_outflow (:398-443) with GridPhysBound's face positions (grid/grid_phys_bound.cpp:27-54)
//   Grid::compute_*_flux with SAVE_RECONSTRUCTIONS (grid/grid.cpp:215-302, binary_defs.h:12),
//   sum_*_difs (:304-344), add_difs (:406-420), Binary::add_difs (binary/binary.cpp:205-254),
//   Poisson::compute_forces (poisson/poisson.cpp:223-247), Binary::pot_to_grid
//   (binary/binary.cpp:825-849) after every substep (BinaryDriver::sub_step, binary_driver.cpp:346-382,
//   with the Poisson solve replaced by a given potential).
// The EOS is the Helmholtz one (helmholtz.cpp) or an ideal gas p = rho T (1 + zbar)/abar,
// e = p/((gamma-1) rho), the same Newton solve for T.
// Positions: block b's interior spans [x0, x0 + n h) per axis; face p of the padded pencil lies at
// x0 + (p - g) h and the cell centre half a cell above -- the reference's
// (offset + p) h - GRID_DIM - bw h0 - origin (grid/grid.cpp:500-526) for dyadic blocks.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "recon.hpp"

namespace orc {
int helm_eos_state(double* r);
int helm_compute_T_state(double* r, double toler);
}  // namespace orc

using namespace orc;

namespace {

enum { D0 = 0, DACC = 1, DDON = 2, SR = 3, LZ = 4, SZ = 5, POT = 6, ETR = 7, TAUR = 8, TT = 9, SPEC = 10, NFR = 13 };
constexpr int NSPEC = 3;
constexpr double TMIN_IDEAL = 1.0e-20;  // the ideal gas is unit-free: its T floor only keeps T > 0

struct RC {
  int n, g, G, recon, zref, eos;
  long G3;
  double theta, rho_floor, Omega, Omega0, fR, ftheta, gamma, cfl;
  double abar[NSPEC];
  double toler;  // Newton tolerance (helmholtz/helmeos.cpp:449 uses 1e-9)
  explicit RC(const orc_rot_cfg* c)
      : n(c->n), g(c->g), G(c->n + 2 * c->g), recon(c->recon), zref(c->z_reflect), eos(c->eos),
        G3((long)G * G * G), theta(c->theta), rho_floor(c->rho_floor), Omega(c->Omega), Omega0(c->Omega0),
        fR(c->fR), ftheta(c->ftheta), gamma(c->gamma), cfl(c->cfl),
        toler(c->newton_tol > 0.0 ? c->newton_tol : 1.0e-9) {
    for (int i = 0; i < NSPEC; i++) abar[i] = c->abar[i];
  }
  long at(int i, int j, int k) const { return ((long)k * G + j) * G + i; }
};

struct X3 {
  double x[3];
};

inline double Rc(const X3& X) { return std::sqrt(X.x[0] * X.x[0] + X.x[1] * X.x[1]); }
inline double st(const RC& c, const double* u, const X3& X) { return u[LZ] / Rc(X) - u[D0] * Rc(X) * c.Omega; }
inline double vx(const RC& c, const double* u, const X3& X) {
  return (X.x[0] * u[SR] - X.x[1] * st(c, u, X)) / Rc(X) / u[D0];
}
inline double vy(const RC& c, const double* u, const X3& X) {
  return (X.x[1] * u[SR] + X.x[0] * st(c, u, X)) / Rc(X) / u[D0];
}
inline double vz(const double* u) { return u[SZ] / u[D0]; }
inline double dvx(const RC& c, const X3& X) { return -X.x[1] * (c.Omega - c.Omega0); }
inline double dvy(const RC& c, const X3& X) { return X.x[0] * (c.Omega - c.Omega0); }
inline double ekr(const RC& c, const double* u, const X3& X) {
  const double t = st(c, u, X);
  return 0.5 * (u[SR] * u[SR] + t * t + u[SZ] * u[SZ]) / u[D0];
}
inline double eir(const RC& c, const double* u, const X3& X) {
  const double e = u[ETR] - ekr(c, u, X);
  if (e < u[ETR] * 0.001) return u[TAUR];
  return e;
}
inline double Abar(const RC& c, const double* u) {
  double sum = 0.0;
  for (int i = 0; i < NSPEC; i++) sum += u[SPEC + i] / c.abar[i];
  sum = u[D0] / sum;
  return sum;
}

// the ideal-gas EOS on an eos_t record (same record and unit-free)
int ideal_eos(const RC& c, double* q) {
  // q: p, e, cs, cv, abar, zbar, rho, T, s, dsdt
  const double rho = q[6], T = q[7] > TMIN_IDEAL ? q[7] : TMIN_IDEAL;
  const double mu_inv = (1.0 + q[5]) / q[4];
  q[0] = rho * T * mu_inv;
  q[3] = mu_inv / (c.gamma - 1.0);
  q[1] = q[3] * T;
  q[2] = std::sqrt(c.gamma * q[0] / rho);
  q[8] = 0.0;
  q[9] = 0.0;
  return 0;
}

// helmholtz_compute_T's iteration (helmholtz/helmeos.cpp:436-486) on the ideal gas
int ideal_compute_T(const RC& c, double* q) {
  double T0 = q[7];
  const double e0 = q[1];
  q[7] = TMIN_IDEAL;
  ideal_eos(c, q);
  q[7] = T0;
  double a = 1.0;
  if (q[1] >= e0) {
    q[7] = TMIN_IDEAL;
    return 0;
  }
  int iters = 0;
  double f;
  do {
    T0 = q[7] > TMIN_IDEAL ? q[7] : TMIN_IDEAL;
    ideal_eos(c, q);
    f = q[1] - e0;
    const double dT = -f / q[3];
    q[7] += a * dT;
    const double lo = (T0 + TMIN_IDEAL) / 2.0;
    q[7] = q[7] > lo ? q[7] : lo;
    iters++;
    if (iters % 10 == 0) a *= 0.9;
    if (q[7] < TMIN_IDEAL) return 128;
  } while (std::fabs(f / e0) >= c.toler && iters < 100);
  ideal_eos(c, q);
  return iters >= 100 ? 64 : 0;
}

// State::compute_eos (binary/state_ztwd.cpp:168-176): p and cs of a conserved state
int eos_of(const RC& c, const double* u, double* p, double* cs) {
  double q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  q[6] = u[D0];
  q[7] = u[TT];
  q[4] = Abar(c, u);
  q[5] = Abar(c, u) / 2.0;
  const int e = c.eos == 1 ? helm_eos_state(q) : ideal_eos(c, q);
  *p = q[0];
  if (cs) *cs = q[2];
  return e;
}

// binary/state_ztwd.cpp:581-599
void to_prim(const RC& c, double* u, const X3& X) {
  for (int i = 0; i < NSPEC; i++) u[SPEC + i] /= u[D0];
  u[DDON] /= u[D0];
  u[DACC] /= u[D0];
  u[ETR] -= ekr(c, u, X);
  u[SR] /= u[D0];
  u[LZ] /= u[D0] * Rc(X);
  u[LZ] -= c.Omega * Rc(X);
  u[SZ] /= u[D0];
  u[POT] /= u[D0];
}

// binary/state_ztwd.cpp:556-579
void to_con(const RC& c, double* u, const X3& X) {
  u[SZ] *= u[D0];
  u[POT] *= u[D0];
  u[SR] *= u[D0];
  u[LZ] += c.Omega * Rc(X);
  u[LZ] *= u[D0] * Rc(X);
  u[ETR] += ekr(c, u, X);
  double sum = 0.0;
  for (int i = 0; i < NSPEC; i++) sum += u[SPEC + i];
  for (int i = 0; i < NSPEC; i++) u[SPEC + i] /= sum;
  sum = 0.0;
  sum += u[DDON];
  sum += u[DACC];
  u[DDON] /= sum;
  u[DACC] /= sum;
  for (int i = 0; i < NSPEC; i++) u[SPEC + i] *= u[D0];
  u[DDON] *= u[D0];
  u[DACC] *= u[D0];
}

// binary/state_ztwd.cpp:360-396 (+ :176-209 for the eigenvalue); returns |v_d| + c_s
double flux(const RC& c, const double* u, const X3& X, int d, double* F, int* err) {
  double p, cs;
  *err |= eos_of(c, u, &p, &cs);
  const double v = d == 0 ? vx(c, u, X) : (d == 1 ? vy(c, u, X) : vz(u));
  for (int i = 0; i < NFR; i++) F[i] = v * u[i];
  if (d == 0) {
    F[SR] += X.x[0] * p / Rc(X);
    F[LZ] -= X.x[1] * p;
    F[ETR] += (v + dvx(c, X)) * p;
  } else if (d == 1) {
    F[SR] += X.x[1] * p / Rc(X);
    F[LZ] += X.x[0] * p;
    F[ETR] += (v + dvy(c, X)) * p;
  } else {
    F[SZ] += p;
    F[ETR] += v * p;
  }
  return std::fabs(v) + cs;
}

// binary/state_ztwd.cpp:482-554
int floor_state(const RC& c, double* u, const X3& X) {
  if (u[D0] < c.rho_floor) {
    u[D0] = c.rho_floor;
    u[DACC] = c.rho_floor * 0.5;
    u[DDON] = c.rho_floor * 0.5;
    u[SPEC + 0] = c.rho_floor;
    for (int i = 1; i < NSPEC; i++) u[SPEC + i] = 0.0;
  } else {
    for (int i = 0; i < NSPEC; i++) u[SPEC + i] = rmax(u[SPEC + i], 0.0);
    u[DDON] = rmax(u[DDON], c.rho_floor / 2.0);
    u[DACC] = rmax(u[DACC], c.rho_floor / 2.0);
  }
  double sum = 0.0;
  for (int i = 0; i < NSPEC; i++) sum += u[SPEC + i];
  if (sum != 0.0) {
    for (int i = 0; i < NSPEC; i++) u[SPEC + i] *= u[D0] / sum;
  } else {
    u[SPEC] = u[D0];
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
  const double ue = eir(c, u, X);
  double q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  q[4] = Abar(c, u);
  q[5] = Abar(c, u) / 2.0;
  q[6] = d;
  q[1] = ue / d;
  q[7] = u[TT];
  const int err = c.eos == 1 ? helm_compute_T_state(q, c.toler) : ideal_compute_T(c, q);
  u[TT] = q[7];
  if (ue > 0.1 * u[ETR]) u[TAUR] = ue;
  return err;
}

// binary/state_ztwd.cpp:398-443 on node face `nf` (0 XL .. 5 ZU of the node); the boundary
// object's face is the inverted one, whose "upper" branch tests the lower node face
void enforce_outflow(const RC& c, double* u, const X3& X, int nf) {
  u[ETR] = eir(c, u, X);
  double v0;
  switch (nf) {
    case 0:  // bound face XU
      if (vx(c, u, X) > 0.0) {
        v0 = vy(c, u, X);
        u[LZ] = (X.x[0] * (v0 + X.x[0] * c.Omega)) * u[D0];
        u[SR] = X.x[1] * v0 * u[D0] / Rc(X);
      }
      break;
    case 1:  // XL
      if (vx(c, u, X) < 0.0) {
        v0 = vy(c, u, X);
        u[LZ] = (X.x[0] * (v0 + X.x[0] * c.Omega)) * u[D0];
        u[SR] = X.x[1] * v0 * u[D0] / Rc(X);
      }
      break;
    case 2:  // YU
      if (vy(c, u, X) > 0.0) {
        v0 = vx(c, u, X);
        u[LZ] = (-X.x[1] * (v0 - X.x[1] * c.Omega)) * u[D0];
        u[SR] = X.x[0] * v0 * u[D0] / Rc(X);
      }
      break;
    case 3:  // YL
      if (vy(c, u, X) < 0.0) {
        v0 = vx(c, u, X);
        u[LZ] = (-X.x[1] * (v0 - X.x[1] * c.Omega)) * u[D0];
        u[SR] = X.x[0] * v0 * u[D0] / Rc(X);
      }
      break;
    case 4:  // ZU
      if (u[SZ] > 0.0) u[SZ] = 0.0;
      break;
    case 5:  // ZL
      if (u[SZ] < 0.0) u[SZ] = 0.0;
      break;
  }
  const double t = st(c, u, X);
  u[ETR] += 0.5 * (u[SR] * u[SR] + t * t + u[SZ] * u[SZ]) / u[D0];
}

inline X3 pos(const RC& c, const double* x0, double h, int pi, int pj, int pk, int face_dir) {
  // cell centre of padded cell (pi, pj, pk); with face_dir = d, the low face of that cell in d
  X3 X;
  const int p[3] = {pi, pj, pk};
  for (int d = 0; d < 3; d++) {
    const double xf = x0[d] + (double)(p[d] - c.g) * h;
    X.x[d] = d == face_dir ? xf : xf + 0.5 * h;
  }
  return X;
}

inline long pencil_cell(const RC& c, int d, int a, int b, int p) {
  if (d == 0) return c.at(p, a, b);
  if (d == 1) return c.at(a, p, b);
  return c.at(a, b, p);
}

// Grid::compute_*_flux with saved face states (grid/grid.cpp:215-302), for one block and
// direction: F[f*G3 + cell] for faces g..G-g, vf[cell] = v_d of (ql + qr)/2 at that face
int block_flux(const RC& c, const double* U, const double* x0, double h, int d, double* F, double* vf,
               double* eig_max) {
  int err = 0;
  const unsigned pot_mask = 1u << POT;
  for (int b = c.g; b < c.G - c.g; b++)
    for (int a = c.g; a < c.G - c.g; a++) {
      double q0[MAXG][MAXF], ql[MAXG][MAXF], qr[MAXG][MAXF];
      for (int p = 0; p < c.G; p++) {
        const long e = pencil_cell(c, d, a, b, p);
        for (int l = 0; l < NFR; l++) q0[p][l] = U[l * c.G3 + e];
        int ijk[3];
        if (d == 0) ijk[0] = p, ijk[1] = a, ijk[2] = b;
        else if (d == 1) ijk[0] = a, ijk[1] = p, ijk[2] = b;
        else ijk[0] = a, ijk[1] = b, ijk[2] = p;
        to_prim(c, q0[p], pos(c, x0, h, ijk[0], ijk[1], ijk[2], -1));
      }
      if (c.recon == 1)
        recon_ppm(NFR, c.G, q0, ql, qr, pot_mask, pot_mask);
      else
        recon_minmod(NFR, c.theta, c.G, q0, ql, qr);
      for (int p = c.g; p < c.G - c.g + 1; p++) {
        int ijk[3];
        if (d == 0) ijk[0] = p, ijk[1] = a, ijk[2] = b;
        else if (d == 1) ijk[0] = a, ijk[1] = p, ijk[2] = b;
        else ijk[0] = a, ijk[1] = b, ijk[2] = p;
        const X3 X = pos(c, x0, h, ijk[0], ijk[1], ijk[2], d);
        to_con(c, ql[p], X);
        to_con(c, qr[p], X);
        double uf[NFR];
        for (int l = 0; l < NFR; l++) uf[l] = (ql[p][l] + qr[p][l]) * 0.5;
        double fl[NFR], fr[NFR];
        const double el = flux(c, ql[p], X, d, fl, &err);
        const double er = flux(c, qr[p], X, d, fr, &err);
        const double amax = rmax(el, er);
        if (eig_max) *eig_max = rmax(*eig_max, amax);
        const long e = pencil_cell(c, d, a, b, p);
        if (F) {
          for (int l = 0; l < NFR; l++) F[l * c.G3 + e] = ((fl[l] + fr[l]) - (qr[p][l] - ql[p][l]) * amax) * 0.5;
          vf[e] = d == 0 ? vx(c, uf, X) : (d == 1 ? vy(c, uf, X) : vz(uf));
        }
      }
    }
  return err;
}

inline double phi_at(const RC& c, const double* phi, int i, int j, int k) {
  // interior-relative cell (i, j, k) in [-1, n]^3 of the (n+2)^3 potential
  const int P = c.n + 2;
  return phi[((long)(k + 1) * P + (j + 1)) * P + (i + 1)];
}

// Grid::sum_{x,y,z}_difs (grid/grid.cpp:304-344) for direction d from the face fluxes F
void block_sum_difs(const RC& c, const double* F, double* D, int d, double h) {
  const double hinv = 1.0 / h;
  const long step = d == 0 ? 1 : (d == 1 ? c.G : (long)c.G * c.G);
  for (int k = c.g; k < c.G - c.g; k++)
    for (int j = c.g; j < c.G - c.g; j++)
      for (int i = c.g; i < c.G - c.g; i++) {
        const long e = c.at(i, j, k);
        for (int l = 0; l < NFR; l++) D[l * c.G3 + e] -= (F[l * c.G3 + e + step] - F[l * c.G3 + e]) * hinv;
      }
}

// Binary::add_difs (binary/binary.cpp:205-254) and Grid::add_difs (grid/grid.cpp:406-420) with
// the floor, then Binary::pot_to_grid (binary/binary.cpp:825-849).  The cell force is the mean
// of its two face forces: from the given Poisson face forces (force[d], (n+2)^3 each, face i
// between cells i-1 and i), or Poisson::compute_forces of phi (poisson/poisson.cpp:223-247).
int block_update(const RC& c, double* U, const double* U0, const double* D, const double* vf, const double* phi,
                 const double* const* force, const double* x0, double h, double dt, double beta) {
  int err = 0;
  const int P = c.n + 2;
  for (int k = c.g; k < c.G - c.g; k++)
    for (int j = c.g; j < c.G - c.g; j++)
      for (int i = c.g; i < c.G - c.g; i++) {
        const long e = c.at(i, j, k);
        const X3 X = pos(c, x0, h, i, j, k, -1);
        double s[NFR], Dl[NFR], sv[NFR];
        for (int l = 0; l < NFR; l++) {
          s[l] = U[l * c.G3 + e];
          Dl[l] = D[l * c.G3 + e];
        }
        const int ii = i - c.g, jj = j - c.g, kk = k - c.g;
        const double pc = phi_at(c, phi, ii, jj, kk);
        double gf[3];
        if (force) {
          auto fat = [&](int d, int a, int b2, int c2) { return force[d][((long)c2 * P + b2) * P + a]; };
          gf[0] = 0.5 * (fat(0, ii + 1, jj + 1, kk + 1) + fat(0, ii + 2, jj + 1, kk + 1));
          gf[1] = 0.5 * (fat(1, ii + 1, jj + 1, kk + 1) + fat(1, ii + 1, jj + 2, kk + 1));
          gf[2] = 0.5 * (fat(2, ii + 1, jj + 1, kk + 1) + fat(2, ii + 1, jj + 1, kk + 2));
        } else {
          const double fxl = -(pc - phi_at(c, phi, ii - 1, jj, kk)) / h;
          const double fxh = -(phi_at(c, phi, ii + 1, jj, kk) - pc) / h;
          const double fyl = -(pc - phi_at(c, phi, ii, jj - 1, kk)) / h;
          const double fyh = -(phi_at(c, phi, ii, jj + 1, kk) - pc) / h;
          const double fzl = -(pc - phi_at(c, phi, ii, jj, kk - 1)) / h;
          const double fzh = -(phi_at(c, phi, ii, jj, kk + 1) - pc) / h;
          gf[0] = 0.5 * (fxl + fxh);
          gf[1] = 0.5 * (fyl + fyh);
          gf[2] = 0.5 * (fzl + fzh);
        }
        // gravity_source (binary/state_ztwd.cpp:277-288)
        for (int l = 0; l < NFR; l++) sv[l] = 0.0;
        const double dd = s[D0];
        sv[SR] += dd * (X.x[0] * gf[0] + X.x[1] * gf[1]) / Rc(X);
        sv[LZ] += dd * (X.x[0] * gf[1] - X.x[1] * gf[0]);
        sv[SZ] += dd * gf[2];
        sv[ETR] += Dl[POT];
        sv[ETR] -= Dl[D0] * s[POT] / dd;
        sv[ETR] += s[D0] * (dvx(c, X) * gf[0] + dvy(c, X) * gf[1]);
        for (int l = 0; l < NFR; l++) Dl[l] += sv[l];
        // -p div u from the saved face states (binary.cpp:231-239)
        const double* vX = vf;
        const double* vY = vf + c.G3;
        const double* vZ = vf + 2 * c.G3;
        double du = 0.0;
        du += vX[c.at(i + 1, j, k)];
        du += vY[c.at(i, j + 1, k)];
        du += vZ[c.at(i, j, k + 1)];
        du -= vX[e];
        du -= vY[e];
        du -= vZ[e];
        du /= h;
        double p;
        err |= eos_of(c, s, &p, nullptr);
        for (int l = 0; l < NFR; l++) sv[l] = 0.0;
        sv[TAUR] = -p * du;
        for (int l = 0; l < NFR; l++) Dl[l] += sv[l];
        Dl[POT] = 0.0;
        Dl[TT] = 0.0;
        // Grid::add_difs (grid/grid.cpp:406-420) with State::source (binary/state_ztwd.cpp:347-358)
        for (int l = 0; l < NFR; l++) sv[l] = 0.0;
        {
          const double R = Rc(X);
          sv[SR] += (p + std::pow(s[LZ] / R, 2) / s[D0]) / R;
          const double th = std::atan2(X.x[1], X.x[0]);
          sv[SR] += s[D0] * c.fR * std::cos(th - c.ftheta);
          sv[LZ] -= s[D0] * c.fR * std::sin(th - c.ftheta) * R;
          sv[ETR] += s[SR] * c.fR * std::cos(th - c.ftheta);
          sv[ETR] -= st(c, s, X) * c.fR * std::sin(th - c.ftheta) * R;
        }
        for (int l = 0; l < NFR; l++) Dl[l] += sv[l];
        double u[NFR];
        for (int l = 0; l < NFR; l++) u[l] = (s[l] + Dl[l] * dt) * beta + U0[l * c.G3 + e] * (1.0 - beta);
        err |= floor_state(c, u, X);
        // Binary::pot_to_grid (binary/binary.cpp:825-849)
        const double R2 = X.x[0] * X.x[0] + X.x[1] * X.x[1];
        u[POT] = u[D0] * (pc - 0.5 * R2 * c.Omega0 * c.Omega0);
        for (int l = 0; l < NFR; l++) U[l * c.G3 + e] = u[l];
      }
  return err;
}

int block_substep(const RC& c, double* U, const double* U0, const double* phi, const double* x0, double h,
                  double dt, double beta) {
  int err = 0;
  std::vector<double> F((size_t)NFR * c.G3), D((size_t)NFR * c.G3, 0.0), vf((size_t)3 * c.G3, 0.0);
  for (int d = 0; d < 3; d++) {
    err |= block_flux(c, U, x0, h, d, F.data(), vf.data() + (size_t)d * c.G3, nullptr);
    block_sum_difs(c, F.data(), D.data(), d, h);
  }
  err |= block_update(c, U, U0, D.data(), vf.data(), phi, nullptr, x0, h, dt, beta);
  return err;
}

// GridPhysBound (grid/grid_phys_bound.cpp:27-54, :60-169): a ghost copies the boundary cell next
// to it and applies enforce_outflow at the boundary face; the lower z face mirrors with z_reflect
void face_outflow_ghosts(const RC& c, double* U, const double* x0, double h, int nf) {
  const int G = c.G, g = c.g;
  {
    const int d = nf >> 1, up = nf & 1;
    for (int b = g; b < G - g; b++)
      for (int a = g; a < G - g; a++)
        for (int t = 0; t < g; t++) {
          const int pg = up ? G - g + t : t;          // ghost position along d
          const int pb = up ? G - g - 1 : g;          // boundary cell
          const bool mirror = c.zref && nf == 4;
          const int ps = mirror ? 2 * g - 1 - t : pb;  // mirrored source (grid_phys_bound.cpp:122-136)
          const long eg = pencil_cell(c, d, a, b, pg), es = pencil_cell(c, d, a, b, ps);
          double u[NFR];
          for (int l = 0; l < NFR; l++) u[l] = U[l * c.G3 + es];
          if (mirror) {
            u[SZ] = -u[SZ];
          } else {
            int ijk[3];
            const int pf = up ? pb + 1 : pb;  // the boundary face
            if (d == 0) ijk[0] = pf, ijk[1] = a, ijk[2] = b;
            else if (d == 1) ijk[0] = a, ijk[1] = pf, ijk[2] = b;
            else ijk[0] = a, ijk[1] = b, ijk[2] = pf;
            enforce_outflow(c, u, pos(c, x0, h, ijk[0], ijk[1], ijk[2], d), nf);
          }
          for (int l = 0; l < NFR; l++) U[l * c.G3 + eg] = u[l];
        }
  }
}

void block_outflow_ghosts(const RC& c, double* U, const double* x0, double h) {
  for (int nf = 0; nf < 6; nf++) face_outflow_ghosts(c, U, x0, h, nf);
}

}  // namespace

// ---- the rotating-frame state on the octree (BinaryDriver::sub_step, binary_driver.cpp:346-382) ----
// GridNode::enforce_boundaries (grid/grid_node.cpp:59-101) with GridAMRBound
// (grid/grid_amr_bound.cpp:27-78) and GridPhysBound (grid/grid_phys_bound.cpp:27-169);
// compute_*_flux children first, adjust_*_flux (grid/grid_node.cpp:1076-1274), sum_*_difs with
// sum_*_flow_off at the root (grid/grid_node.cpp:515-551), Binary::add_difs + Grid::add_difs
// (binary/binary.cpp:205-254, grid/grid_node.cpp:337-357) with the Poisson face forces, and
// inject_from_children with the rotating floor at the parent cell (grid/grid_node.cpp:1035-1074).
struct orc_rot_tree {
  struct RNode {
    int level = 0, parent = -1, cidx = -1, child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    int loc[3] = {0, 0, 0}, off[3] = {0, 0, 0};
    double h = 0.0, x0[3] = {0, 0, 0};
    std::vector<double> U, U0, F, vf, D, phi, force;  // F: the face fluxes of x, y, z
    double DFO[NFR] = {0}, FO[NFR] = {0}, FO0[NFR] = {0};
  };
  RC c;
  double h0 = 0.0, grid_dim = 1.0, origin[3] = {0.0, 0.0, 0.0};
  bool use_forces = false;
  std::vector<RNode> nodes;
  // low interior face of node nd, Grid::xf(bw) (grid/grid.cpp:512-524): x and y subtract the
  // origin; z subtracts GRID_DIM but not the origin, or under Z_REFLECT the origin but not GRID_DIM
  void set_x0(RNode& nd) const {
    for (int d = 0; d < 3; d++) {
      double x = double(nd.off[d] + c.g) * nd.h;
      if (d < 2)
        x = x - grid_dim - c.g * h0 - origin[d];
      else if (c.zref)
        x = x - c.g * h0 - origin[2];
      else
        x = x - grid_dim - c.g * h0;
      nd.x0[d] = x;
    }
  }
  explicit orc_rot_tree(const orc_rot_cfg* cfg) : c(cfg) {}
  int sibling(int b, int face) const {
    const RNode& nd = nodes[b];
    const int d = face >> 1;
    int l[3] = {nd.loc[0], nd.loc[1], nd.loc[2]};
    l[d] += (face & 1) ? 1 : -1;
    if (l[d] < 0 || l[d] >= (1 << nd.level)) return -1;
    for (int q = 0; q < (int)nodes.size(); q++)
      if (nodes[q].level == nd.level && nodes[q].loc[0] == l[0] && nodes[q].loc[1] == l[1] && nodes[q].loc[2] == l[2])
        return q;
    return -1;
  }
  bool phys(int b, int face) const {  // oct_node/oct_node.cpp:86-110
    const RNode& nd = nodes[b];
    const int d = face >> 1;
    return (face & 1) ? nd.loc[d] == (1 << nd.level) - 1 : nd.loc[d] == 0;
  }
  void postorder(int i, std::vector<int>& out) const {
    for (int q = 0; q < 8; q++)
      if (nodes[i].child[q] >= 0) postorder(nodes[i].child[q], out);
    out.push_back(i);
  }
  void preorder(int i, std::vector<int>& out) const {
    out.push_back(i);
    for (int q = 0; q < 8; q++)
      if (nodes[i].child[q] >= 0) preorder(nodes[i].child[q], out);
  }
  std::vector<int> level_order() const {
    std::vector<int> ord;
    int maxl = 0;
    for (auto& nd : nodes) maxl = nd.level > maxl ? nd.level : maxl;
    for (int l = 0; l <= maxl; l++)
      for (int i = 0; i < (int)nodes.size(); i++)
        if (nodes[i].level == l) ord.push_back(i);
    return ord;
  }
};

namespace {

using RT = orc_rot_tree;

// GridAMRBound::operator() for ghost (i, j, k) of node nd across a face of direction d: the
// parent's cell with a quarter of the minmod slope along d, sign by the cell's parity
void rot_amr_ghost(const RT* t, const RT::RNode& nd, int d, int i, int j, int k, double* u) {
  const RC& c = t->c;
  const RT::RNode& P = t->nodes[nd.parent];
  const int idx[3] = {i, j, k};
  int i1[3], par[3];
  for (int e = 0; e < 3; e++) {
    const int f = nd.off[e] + idx[e] - 2 * P.off[e];
    par[e] = f % 2;
    i1[e] = f / 2;
  }
  int p1[3] = {i1[0], i1[1], i1[2]}, m1[3] = {i1[0], i1[1], i1[2]};
  p1[d] += 1;
  m1[d] -= 1;
  const long e0 = c.at(i1[0], i1[1], i1[2]), ep = c.at(p1[0], p1[1], p1[2]), em = c.at(m1[0], m1[1], m1[2]);
  for (int l = 0; l < NFR; l++) {
    const double uc = P.U[l * c.G3 + e0];
    double m = minmod2(P.U[l * c.G3 + ep] - uc, uc - P.U[l * c.G3 + em]);
    if (par[d] == 1) m = 0.0 - m;
    u[l] = uc - m * 0.25;
  }
}

void rot_node_ghosts(RT* t, int ni) {
  const RC& c = t->c;
  RT::RNode& nd = t->nodes[ni];
  for (int face = 0; face < 6; face++) {
    const int d = face >> 1;
    const bool upper = face & 1;
    const int s = t->sibling(ni, face);
    if (s < 0 && t->phys(ni, face)) {
      face_outflow_ghosts(c, nd.U.data(), nd.x0, nd.h, face);
      continue;
    }
    for (int b = c.g; b < c.G - c.g; b++)
      for (int a = c.g; a < c.G - c.g; a++)
        for (int tt = 0; tt < c.g; tt++) {
          const int gp = upper ? (c.G - c.g + tt) : tt;
          const long eg = pencil_cell(c, d, a, b, gp);
          double u[NFR];
          if (s >= 0) {
            // the sibling's interior cell at the same distance from the shared face
            const int sp = upper ? (gp - c.G + 2 * c.g) : (c.G - 2 * c.g + gp);
            const RT::RNode& S = t->nodes[s];
            const long es = pencil_cell(c, d, a, b, sp);
            for (int l = 0; l < NFR; l++) u[l] = S.U[l * c.G3 + es];
          } else {
            int ijk[3];
            if (d == 0) ijk[0] = gp, ijk[1] = a, ijk[2] = b;
            else if (d == 1) ijk[0] = a, ijk[1] = gp, ijk[2] = b;
            else ijk[0] = a, ijk[1] = b, ijk[2] = gp;
            rot_amr_ghost(t, nd, d, ijk[0], ijk[1], ijk[2], u);
          }
          for (int l = 0; l < NFR; l++) nd.U[l * c.G3 + eg] = u[l];
        }
  }
}

// GridNode::adjust_{x,y,z}_flux for one node: where exactly one side of a face plane (the node's
// low face, its middle plane, its high face) is refined, the node's face fluxes there become the
// means of the 2x2 fine face fluxes
void rot_node_adjust_flux(RT* t, int ni, int d) {
  const RC& c = t->c;
  RT::RNode& nd = t->nodes[ni];
  auto child_of = [&](int node, int cx, int cy, int cz) -> int {
    if (node < 0) return -1;
    return t->nodes[node].child[cx | (cy << 1) | (cz << 2)];
  };
  const int sL = t->sibling(ni, 2 * d), sU = t->sibling(ni, 2 * d + 1);
  for (int l = 0; l < 3; l++)
    for (int ck = 0; ck < 2; ck++)
      for (int cj = 0; cj < 2; cj++) {
        auto ci = [&](int node, int nbit) {
          if (d == 0) return child_of(node, nbit, cj, ck);
          if (d == 1) return child_of(node, cj, nbit, ck);
          return child_of(node, cj, ck, nbit);
        };
        int cl, cr, i0;
        if (l == 0) { cr = ci(ni, 0); cl = ci(sL, 1); i0 = c.g; }
        else if (l == 1) { cl = ci(ni, 0); cr = ci(ni, 1); i0 = c.G / 2; }
        else { cr = ci(sU, 0); cl = ci(ni, 1); i0 = c.G - c.g; }
        int ch, i;
        if (cr < 0 && cl >= 0) { ch = cl; i = c.G - c.g; }
        else if (cl < 0 && cr >= 0) { ch = cr; i = c.g; }
        else continue;
        const RT::RNode& C = t->nodes[ch];
        for (int k = c.g; k <= c.G - c.g - 1; k += 2)
          for (int j = c.g; j <= c.G - c.g - 1; j += 2) {
            const int k0 = (c.g + k) / 2 + ck * (c.G / 2 - c.g);
            const int j0 = (c.g + j) / 2 + cj * (c.G / 2 - c.g);
            for (int f = 0; f < NFR; f++) {
              const double* CF = C.F.data() + ((size_t)d * NFR + f) * c.G3;
              double v = CF[pencil_cell(c, d, j, k, i)];
              v += CF[pencil_cell(c, d, j + 1, k, i)];
              v += CF[pencil_cell(c, d, j, k + 1, i)];
              v += CF[pencil_cell(c, d, j + 1, k + 1, i)];
              v *= 0.25;
              nd.F[((size_t)d * NFR + f) * c.G3 + pencil_cell(c, d, j0, k0, i0)] = v;
            }
          }
      }
}

// Grid::sum_{x,y,z}_flow_off (grid/grid.cpp:350-404)
void rot_flow_off(RT* t, RT::RNode& nd, int d) {
  const RC& c = t->c;
  for (int k = c.g; k < c.G - c.g; k++)
    for (int j = c.g; j < c.G - c.g; j++)
      for (int f = 0; f < NFR; f++) {
        const double* F = nd.F.data() + ((size_t)d * NFR + f) * c.G3;
        double tmp;
        if (d == 2 && c.zref)
          tmp = F[pencil_cell(c, d, j, k, c.G - c.g)] * nd.h * nd.h;
        else
          tmp = (F[pencil_cell(c, d, j, k, c.G - c.g)] - F[pencil_cell(c, d, j, k, c.g)]) * nd.h * nd.h;
        nd.DFO[f] += tmp;
      }
}

// GridNode::inject_from_children for one (parent, child) pair, floored at the parent cell
int rot_inject_child(RT* t, int pi, int q) {
  const RC& c = t->c;
  RT::RNode& P = t->nodes[pi];
  const RT::RNode& C = t->nodes[P.child[q]];
  const int cx = q & 1, cy = (q >> 1) & 1, cz = (q >> 2) & 1;
  int err = 0;
  for (int k = c.g; k <= c.G - c.g - 1; k += 2)
    for (int j = c.g; j <= c.G - c.g - 1; j += 2)
      for (int i = c.g; i < c.G - c.g; i += 2) {
        const int k0 = (c.g + k) / 2 + cz * (c.G / 2 - c.g);
        const int j0 = (c.g + j) / 2 + cy * (c.G / 2 - c.g);
        const int i0 = (c.g + i) / 2 + cx * (c.G / 2 - c.g);
        double u[NFR];
        for (int l = 0; l < NFR; l++) {
          const double* f = C.U.data() + l * c.G3;
          double s = 0.0;
          s += f[c.at(i + 0, j + 0, k + 0)];
          s += f[c.at(i + 1, j + 0, k + 0)];
          s += f[c.at(i + 0, j + 1, k + 0)];
          s += f[c.at(i + 1, j + 1, k + 0)];
          s += f[c.at(i + 0, j + 0, k + 1)];
          s += f[c.at(i + 1, j + 0, k + 1)];
          s += f[c.at(i + 0, j + 1, k + 1)];
          s += f[c.at(i + 1, j + 1, k + 1)];
          s *= 0.125;
          u[l] = s;
        }
        err |= floor_state(c, u, pos(c, P.x0, P.h, i0, j0, k0, -1));
        for (int l = 0; l < NFR; l++) P.U[l * c.G3 + c.at(i0, j0, k0)] = u[l];
      }
  return err;
}

// GridNode::inject_from_parent (grid/grid_node.cpp:193-233) of the rotating state: a new child's
// interior from its parent's cells (ghosts filled) with minmod slopes per axis, each child cell
// floored at its own position (Binary::inject_from_parent, binary/binary.cpp:852-855, adds only
// the potential, which the solver owns)
int rot_prolong_child(RT* t, int ni) {
  const RC& c = t->c;
  RT::RNode& C = t->nodes[ni];
  const RT::RNode& P = t->nodes[C.parent];
  const int q = C.cidx;
  const int cx = q & 1, cy = (q >> 1) & 1, cz = (q >> 2) & 1;
  int err = 0;
  auto Pu = [&](int l, int i, int j, int k) { return P.U[l * c.G3 + c.at(i, j, k)]; };
  for (int k = c.g; k <= c.G - c.g - 1; k += 2)
    for (int j = c.g; j <= c.G - c.g - 1; j += 2) {
      const int k0 = (c.g + k) / 2 + cz * (c.G / 2 - c.g);
      const int j0 = (c.g + j) / 2 + cy * (c.G / 2 - c.g);
      int i0 = c.g + cx * (c.G / 2 - c.g);
      for (int i = c.g; i < c.G - c.g; i += 2, i0++) {
        double v[8][NFR];
        for (int l = 0; l < NFR; l++) {
          const double u = Pu(l, i0, j0, k0);
          const double s1 = minmod2(Pu(l, i0 + 1, j0, k0) - u, u - Pu(l, i0 - 1, j0, k0));
          const double s2 = minmod2(Pu(l, i0, j0 + 1, k0) - u, u - Pu(l, i0, j0 - 1, k0));
          const double s3 = minmod2(Pu(l, i0, j0, k0 + 1) - u, u - Pu(l, i0, j0, k0 - 1));
          v[0][l] = u - (s1 + s2 + s3) * 0.25;
          v[1][l] = u + (s1 - s2 - s3) * 0.25;
          v[2][l] = u - (s1 - s2 + s3) * 0.25;
          v[3][l] = u + (s1 + s2 - s3) * 0.25;
          v[4][l] = u - (s1 + s2 - s3) * 0.25;
          v[5][l] = u + (s1 - s2 + s3) * 0.25;
          v[6][l] = u - (s1 - s2 - s3) * 0.25;
          v[7][l] = u + (s1 + s2 + s3) * 0.25;
        }
        for (int m = 0; m < 8; m++) {
          const int ii = i + (m & 1), jj = j + ((m >> 1) & 1), kk = k + (m >> 2);
          err |= floor_state(c, v[m], pos(c, C.x0, C.h, ii, jj, kk, -1));
          for (int l = 0; l < NFR; l++) C.U[l * c.G3 + c.at(ii, jj, kk)] = v[m][l];
        }
      }
    }
  return err;
}

}  // namespace

extern "C" {

int orc_rot_tree_prolong(orc_rot_tree* t, int node) {
  if (node <= 0 || node >= (int)t->nodes.size()) return -1;
  return rot_prolong_child(t, node);
}

orc_rot_tree* orc_rot_tree_create(const orc_rot_cfg* cfg, double grid_dim, int nn, const int* parent,
                                  const int* cidx) {
  RC c(cfg);
  if (c.G > MAXG || nn < 1 || parent[0] != -1) return nullptr;
  orc_rot_tree* t = new orc_rot_tree(cfg);
  t->grid_dim = grid_dim;
  t->h0 = 2.0 * grid_dim / double(c.n);  // grid/grid.cpp:70
  const int P = c.n + 2;
  t->nodes.resize(nn);
  for (int b = 0; b < nn; b++) {
    RT::RNode& nd = t->nodes[b];
    if (b > 0) {
      const int p = parent[b], q = cidx[b];
      if (p < 0 || p >= b || q < 0 || q > 7 || t->nodes[p].child[q] >= 0) {
        delete t;
        return nullptr;
      }
      RT::RNode& Pn = t->nodes[p];
      Pn.child[q] = b;
      nd.parent = p;
      nd.cidx = q;
      nd.level = Pn.level + 1;
      for (int d = 0; d < 3; d++) {
        const int cv = (q >> d) & 1;
        nd.loc[d] = 2 * Pn.loc[d] + cv;                       // oct_node/oct_node.cpp:177
        nd.off[d] = 2 * Pn.off[d] + c.g + cv * (c.G - 2 * c.g);  // grid/grid_node.cpp:276
      }
      nd.h = Pn.h * 0.5;
    } else {
      nd.h = t->h0;
    }
    t->set_x0(nd);
    nd.U.assign((size_t)NFR * c.G3, 0.0);
    nd.U0 = nd.U;
    nd.D = nd.U;
    nd.F.assign((size_t)3 * NFR * c.G3, 0.0);
    nd.vf.assign((size_t)3 * c.G3, 0.0);
    nd.phi.assign((size_t)P * P * P, 0.0);
    nd.force.assign((size_t)3 * P * P * P, 0.0);
  }
  return t;
}

void orc_rot_tree_destroy(orc_rot_tree* t) { delete t; }

double* orc_rot_tree_U(orc_rot_tree* t, int node) { return t->nodes[node].U.data(); }
double* orc_rot_tree_phi(orc_rot_tree* t, int node) { return t->nodes[node].phi.data(); }
double* orc_rot_tree_force(orc_rot_tree* t, int node) { return t->nodes[node].force.data(); }
const double* orc_rot_tree_F(const orc_rot_tree* t, int node, int d) {
  return t->nodes[node].F.data() + (size_t)d * NFR * t->c.G3;
}

int orc_rot_tree_node_x0(const orc_rot_tree* t, int node, double* x0, double* h) {
  if (node < 0 || node >= (int)t->nodes.size()) return 1;
  for (int d = 0; d < 3; d++) x0[d] = t->nodes[node].x0[d];
  *h = t->nodes[node].h;
  return 0;
}

void orc_rot_tree_use_forces(orc_rot_tree* t, int on) { t->use_forces = on != 0; }

// GridNode::set_origin (grid/grid_node.cpp:20-26): every node takes the origin
void orc_rot_tree_set_origin(orc_rot_tree* t, const double* origin) {
  for (int d = 0; d < 3; d++) t->origin[d] = origin[d];
  for (auto& nd : t->nodes) t->set_x0(nd);
}

int orc_rot_tree_fill_ghosts(orc_rot_tree* t) {
  // level order: a node reads same-level interiors, its own interior and its parent's cells with
  // the parent's face ghosts, which the reference's pre-order recursion has filled
  for (int ni : t->level_order()) rot_node_ghosts(t, ni);
  return 0;
}

int orc_rot_tree_store(orc_rot_tree* t) {
  for (auto& nd : t->nodes) {
    nd.U0 = nd.U;
    for (int l = 0; l < NFR; l++) nd.FO0[l] = nd.FO[l];
  }
  return 0;
}

// enforce_boundaries, the fluxes with reflux, the flux divergence, Binary::add_difs (sources, RK
// combination, floor) and the potential energy from each node's phi; no restriction (the driver
// solves for phi and restricts afterwards).  Returns the OR of the floor's EOS flags.
int orc_rot_tree_substep(orc_rot_tree* t, double dt, double beta) {
  const RC& c = t->c;
  orc_rot_tree_fill_ghosts(t);
  const int nn = (int)t->nodes.size();
  for (auto& nd : t->nodes) {
    std::fill(nd.D.begin(), nd.D.end(), 0.0);
    for (int l = 0; l < NFR; l++) nd.DFO[l] = 0.0;
  }
  std::vector<int> po;
  t->postorder(0, po);
  int err = 0;
  for (int d = 0; d < 3; d++) {
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int ni = 0; ni < nn; ni++) {
      RT::RNode& nd = t->nodes[ni];
      err |= block_flux(c, nd.U.data(), nd.x0, nd.h, d, nd.F.data() + (size_t)d * NFR * c.G3,
                        nd.vf.data() + (size_t)d * c.G3, nullptr);
    }
    for (int ni : po) rot_node_adjust_flux(t, ni, d);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ni = 0; ni < nn; ni++)
      block_sum_difs(c, t->nodes[ni].F.data() + (size_t)d * NFR * c.G3, t->nodes[ni].D.data(), d, t->nodes[ni].h);
    rot_flow_off(t, t->nodes[0], d);
  }
  const int P = c.n + 2;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
  for (int ni = 0; ni < nn; ni++) {
    RT::RNode& nd = t->nodes[ni];
    const double* fp[3] = {nd.force.data(), nd.force.data() + (size_t)P * P * P, nd.force.data() + (size_t)2 * P * P * P};
    err |= block_update(c, nd.U.data(), nd.U0.data(), nd.D.data(), nd.vf.data(), nd.phi.data(),
                        t->use_forces ? fp : nullptr, nd.x0, nd.h, dt, beta);
  }
  RT::RNode& r = t->nodes[0];
  for (int l = 0; l < NFR; l++) r.FO[l] = (r.FO[l] + r.DFO[l] * dt) * beta + r.FO0[l] * (1.0 - beta);
  return err;
}

// Binary::pot_to_grid (binary/binary.cpp:825-849) from each node's phi
int orc_rot_tree_pot_to_grid(orc_rot_tree* t) {
  const RC& c = t->c;
  for (auto& nd : t->nodes)
    for (int k = c.g; k < c.G - c.g; k++)
      for (int j = c.g; j < c.G - c.g; j++)
        for (int i = c.g; i < c.G - c.g; i++) {
          const X3 X = pos(c, nd.x0, nd.h, i, j, k, -1);
          const long e = c.at(i, j, k);
          const double d = nd.U[D0 * c.G3 + e];
          const double R2 = X.x[0] * X.x[0] + X.x[1] * X.x[1];
          nd.U[POT * c.G3 + e] = d * (phi_at(c, nd.phi.data(), i - c.g, j - c.g, k - c.g) - 0.5 * R2 * c.Omega0 * c.Omega0);
        }
  return 0;
}

int orc_rot_tree_inject(orc_rot_tree* t) {
  std::vector<int> po;
  t->postorder(0, po);
  int err = 0;
  for (int ni : po)
    for (int q = 0; q < 8; q++)
      if (t->nodes[ni].child[q] >= 0) err |= rot_inject_child(t, ni, q);
  return err;
}

// GridNode::max_dt (grid/grid_node.cpp:163-174) over Grid::max_dt (grid/grid.cpp:422-483)
static double rot_max_dt_rec(const orc_rot_tree* t, const std::vector<double>& eig, int ni) {
  const auto& nd = t->nodes[ni];
  double dtinv = 0.0;
  for (int q = 0; q < 8; q++)
    if (nd.child[q] >= 0) {
      const double cdtinv = 1.0 / rot_max_dt_rec(t, eig, nd.child[q]);
      dtinv = rmax(cdtinv, dtinv);
    }
  double own_inv = eig[ni] / nd.h;
  if (own_inv == 0.0) own_inv = 1.0e-10;
  const double own = (1.0 / own_inv) * t->c.cfl;
  dtinv = rmax(1.0 / own, dtinv);
  return 1.0 / dtinv;
}

double orc_rot_tree_max_dt(orc_rot_tree* t) {
  const int nn = (int)t->nodes.size();
  std::vector<double> eig(nn, 0.0);
  orc_rot_tree_fill_ghosts(t);
#pragma omp parallel for schedule(dynamic, 1)
  for (int ni = 0; ni < nn; ni++) {
    double m = 0.0;
    for (int d = 0; d < 3; d++)
      block_flux(t->c, t->nodes[ni].U.data(), t->nodes[ni].x0, t->nodes[ni].h, d, nullptr, nullptr, &m);
    eig[ni] = m;
  }
  return rot_max_dt_rec(t, eig, 0);
}

int orc_rot_tree_flow_off(const orc_rot_tree* t, double* fo) {
  for (int l = 0; l < NFR; l++) fo[l] = t->nodes[0].FO[l];
  return 0;
}

// State sum over the uncovered cells, weighted by h^3 (grid/grid_node.cpp:1489-1523)
int orc_rot_tree_state_sum(const orc_rot_tree* t, double* sum) {
  const RC& c = t->c;
  for (int l = 0; l < NFR; l++) sum[l] = 0.0;
  std::vector<int> pre;
  t->preorder(0, pre);
  for (int ni : pre) {
    const auto& nd = t->nodes[ni];
    const double h3 = std::pow(nd.h, 3);
    for (int l = 0; l < NFR; l++) {
      double a0 = 0.0;
      for (int k = c.g; k < c.G - c.g; k++)
        for (int j = c.g; j < c.G - c.g; j++)
          for (int i = c.g; i < c.G - c.g; i++) {
            const int q = (2 * i / c.G) | ((2 * j / c.G) << 1) | ((2 * k / c.G) << 2);
            if (nd.child[q] >= 0) continue;
            a0 += nd.U[l * c.G3 + c.at(i, j, k)] * h3;
          }
      sum[l] += a0;
    }
  }
  return 0;
}

int orc_rot_batch_substep(const orc_rot_cfg* cfg, int nb, double* U, const double* U0, const double* phi,
                          const double* x0, const double* h, double dt, double beta, int* flags) {
  RC c(cfg);
  if (c.G > MAXG || c.n + 2 * c.g != c.G) return 1;
  const int P = c.n + 2;
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    const int e = block_substep(c, U + (long)b * NFR * c.G3, U0 + (long)b * NFR * c.G3,
                                phi + (long)b * P * P * P, x0 + 3L * b, h[b], dt, beta);
    if (flags) flags[b] = e;
  }
  return 0;
}

int orc_rot_batch_max_eig(const orc_rot_cfg* cfg, int nb, const double* U, const double* x0, const double* h,
                          double* eig) {
  RC c(cfg);
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) {
    double m = 0.0;
    for (int d = 0; d < 3; d++) block_flux(c, U + (long)b * NFR * c.G3, x0 + 3L * b, h[b], d, nullptr, nullptr, &m);
    eig[b] = m;
  }
  return 0;
}

int orc_rot_batch_outflow_ghosts(const orc_rot_cfg* cfg, int nb, double* U, const double* x0, const double* h) {
  RC c(cfg);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; b++) block_outflow_ghosts(c, U + (long)b * NFR * c.G3, x0 + 3L * b, h[b]);
  return 0;
}

// single-state access for the tests: op 0 to_prim, 1 to_con, 2..4 flux in x/y/z (F out, returns
// the eigenvalue in F[13]), 5 floor, 6 outflow on node face F[0]
int orc_rot_state(const orc_rot_cfg* cfg, int op, double* u, const double* X, double* F) {
  RC c(cfg);
  X3 P{{X[0], X[1], X[2]}};
  int err = 0;
  switch (op) {
    case 0: to_prim(c, u, P); break;
    case 1: to_con(c, u, P); break;
    case 2: case 3: case 4: F[NFR] = flux(c, u, P, op - 2, F, &err); break;
    case 5: err = floor_state(c, u, P); break;
    case 6: enforce_outflow(c, u, P, (int)F[0]); break;
    default: return -1;
  }
  return err;
}

}  // extern "C"
