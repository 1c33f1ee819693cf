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
#include <cstring>
#include <vector>

#include "oracle.h"
#include "recon.hpp"

namespace orc {
int helm_eos_state(double* r);
int helm_compute_T_state(double* r);
}  // namespace orc

using namespace orc;

namespace {

enum { D0 = 0, DACC = 1, DDON = 2, SR = 3, LZ = 4, SZ = 5, POT = 6, ETR = 7, TAUR = 8, TT = 9, SPEC = 10, NFR = 13 };
constexpr int NSPEC = 3;
constexpr double TMIN_IDEAL = 1.0e-20;  // the ideal gas is unit-free: its T floor only keeps T > 0

struct RC {
  int n, g, G, recon, zref, eos;
  long G3;
  double theta, rho_floor, Omega, Omega0, fR, ftheta, gamma;
  double abar[NSPEC];
  explicit RC(const orc_rot_cfg* c)
      : n(c->n), g(c->g), G(c->n + 2 * c->g), recon(c->recon), zref(c->z_reflect), eos(c->eos),
        G3((long)G * G * G), theta(c->theta), rho_floor(c->rho_floor), Omega(c->Omega), Omega0(c->Omega0),
        fR(c->fR), ftheta(c->ftheta), gamma(c->gamma) {
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
  } while (std::fabs(f / e0) >= 1.0e-9 && iters < 100);
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
  const int err = c.eos == 1 ? helm_compute_T_state(q) : ideal_compute_T(c, q);
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

int block_substep(const RC& c, double* U, const double* U0, const double* phi, const double* x0, double h,
                  double dt, double beta) {
  int err = 0;
  std::vector<double> F((size_t)NFR * c.G3), D((size_t)NFR * c.G3, 0.0), vf((size_t)3 * c.G3, 0.0);
  const double hinv = 1.0 / h;
  for (int d = 0; d < 3; d++) {
    err |= block_flux(c, U, x0, h, d, F.data(), vf.data() + (size_t)d * c.G3, nullptr);
    const long step = d == 0 ? 1 : (d == 1 ? c.G : (long)c.G * c.G);
    for (int k = c.g; k < c.G - c.g; k++)
      for (int j = c.g; j < c.G - c.g; j++)
        for (int i = c.g; i < c.G - c.g; i++) {
          const long e = c.at(i, j, k);
          for (int l = 0; l < NFR; l++) D[l * c.G3 + e] -= (F[l * c.G3 + e + step] - F[l * c.G3 + e]) * hinv;
        }
  }
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
        // Binary::add_difs (binary/binary.cpp:205-254) with the forces of Poisson::compute_forces
        const int ii = i - c.g, jj = j - c.g, kk = k - c.g;
        const double pc = phi_at(c, phi, ii, jj, kk);
        double gf[3];
        {
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
        const double* vX = vf.data();
        const double* vY = vf.data() + c.G3;
        const double* vZ = vf.data() + 2 * c.G3;
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

// GridPhysBound (grid/grid_phys_bound.cpp:27-54, :60-169): a ghost copies the boundary cell next
// to it and applies enforce_outflow at the boundary face; the lower z face mirrors with z_reflect
void block_outflow_ghosts(const RC& c, double* U, const double* x0, double h) {
  const int G = c.G, g = c.g;
  for (int nf = 0; nf < 6; nf++) {
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

}  // namespace

extern "C" {

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
