// ORACLE -- test infrastructure only (tests/, bench.py CPU baseline).  Never linked by the product.
//
// FP64 restatement of the Helmholtz EOS of helmholtz/helmeos.cpp:
//   read_helm_table (:68-131)      table nodes, spacings and their inverses (from rows in memory
//                                  instead of the helmholtz.table file)
//   helmholtz_set_cgs_units (:133-142), eos_to_code / eos_from_code (:144-164)
//   helmholtz_eos (:164-421)       ions analytic + biquintic Hermite electron-positron free energy
//   helmholtz_compute_T (:436-486) damped Newton for T from e, energy form
// The table keeps the reference's layout (f[i][j], i density, j temperature), so comparing with
// the device path also checks the device's per-cell repacking.  Where the reference aborts the
// oracle returns a flag (same codes as the device).
#include <cmath>
#include <vector>

#include "oracle.h"

namespace {

constexpr int IMAX = 1081, JMAX = 401;
constexpr double TMIN = 1.0e+3;

// helmholtz/const.h:4-13
constexpr double pi = 3.1415926535897932384e0;
constexpr double hpl = 6.6260689633e-27;
constexpr double avo = 6.0221417930e23;
constexpr double kerg = 1.380650424e-16;
constexpr double amu = 1.66053878283e-24;

struct Table {
  std::vector<double> f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt;  // [IMAX*JMAX], (i, j) at i*JMAX + j
  double t[JMAX], d[IMAX];
  double dt_sav[JMAX], dt2_sav[JMAX], dti_sav[JMAX], dt2i_sav[JMAX];
  double dd_sav[IMAX], dd2_sav[IMAX], ddi_sav[IMAX], dd2i_sav[IMAX];
  double tlo, tstpi, dlo, dstpi;
  double cm = 1.0, g = 1.0, s = 1.0, K = 1.0, erg = 1.0, dyne = 1.0;
  bool ready = false;
};
Table T_;

double mx(double a, double b) { return a > b ? a : b; }
double mn(double a, double b) { return a < b ? a : b; }

inline double psi0(double z) { return z * z * z * (z * (-6.0 * z + 15.0) - 10.0) + 1.0; }
inline double dpsi0(double z) { return z * z * (z * (-30.0 * z + 60.0) - 30.0); }
inline double ddpsi0(double z) { return z * (z * (-120.0 * z + 180.0) - 60.0); }
inline double psi1(double z) { return z * (z * z * (z * (-3.0 * z + 8.0) - 6.0) + 1.0); }
inline double dpsi1(double z) { return z * z * (z * (-15.0 * z + 32.0) - 18.0) + 1.0; }
inline double ddpsi1(double z) { return z * (z * (-60.0 * z + 96.0) - 36.0); }
inline double psi2(double z) { return 0.5 * z * z * (z * (z * (-z + 3.0) - 3.0) + 1.0); }
inline double dpsi2(double z) { return 0.5 * z * (z * (z * (-5.0 * z + 12.0) - 9.0) + 2.0); }
inline double ddpsi2(double z) { return 0.5 * (z * (z * (-20.0 * z + 36.0) - 18.0) + 2.0); }

// h5 (:29-47): sum over the 36 products in the reference's term order
double h5(const double* fi, const double wt[6], const double wd[6]) {
  // wt = (w0t, w1t, w2t, w0mt, w1mt, w2mt), wd likewise; term k pairs a density weight with a
  // temperature weight as listed at :29-47
  static const int td[36][2] = {
      {0, 0}, {3, 0}, {0, 3}, {3, 3}, {0, 1}, {3, 1}, {0, 4}, {3, 4}, {0, 2}, {3, 2}, {0, 5}, {3, 5},
      {1, 0}, {4, 0}, {1, 3}, {4, 3}, {2, 0}, {5, 0}, {2, 3}, {5, 3}, {1, 1}, {4, 1}, {1, 4}, {4, 4},
      {2, 1}, {5, 1}, {2, 4}, {5, 4}, {1, 2}, {4, 2}, {1, 5}, {4, 5}, {2, 2}, {5, 2}, {2, 5}, {5, 5}};
  double r = 0.0;
  for (int k = 0; k < 36; k++) {
    const double term = fi[k] * wd[td[k][0]] * wt[td[k][1]];
    r = (k == 0) ? term : r + term;
  }
  return r;
}

struct St {
  double p, e, cs, cv, abar, zbar, rho, T, s, dsdt;
};

void from_code(St& q) {
  const Table& H = T_;
  q.p *= H.dyne / std::pow(H.cm, 2);
  q.e *= H.erg / H.g;
  q.T *= H.K;
  q.cs *= H.cm / H.s;
  q.cv *= H.erg / H.g / H.K;
  q.rho *= H.g / std::pow(H.cm, 3);
  q.s *= H.erg / H.K / H.g;
  q.dsdt *= H.erg / H.K / H.K / H.g;
}

void to_code(St& q) {
  const Table& H = T_;
  q.p /= H.dyne / std::pow(H.cm, 2);
  q.e /= H.erg / H.g;
  q.T /= H.K;
  q.cs /= H.cm / H.s;
  q.cv /= H.erg / H.g / H.K;
  q.rho /= H.g / std::pow(H.cm, 3);
  q.s /= H.erg / H.K / H.g;
  q.dsdt /= H.erg / H.K / H.K / H.g;
}

// helmholtz_eos (:164-421); returns 0 or an error flag
int eos(St& q) {
  const Table& H = T_;
  from_code(q);
  const double den = q.rho;
  const double temp = mx(q.T, TMIN);
  const double abar = q.abar, zbar = q.zbar;
  const double ytot1 = 1.0 / abar;
  const double ye = mx(1.0 - 16, ytot1 * zbar);  // the reference's literal, :177
  const double deni = 1.0 / den;
  const double kt = kerg * temp;
  const double xni = avo * ytot1 * den;
  const double dxnidd = avo * ytot1;
  const double pion = xni * kt;
  const double dpiondd = dxnidd * kt;
  const double dpiondt = xni * kerg;
  const double eion = 1.5 * pion * deni;
  const double deiondt = 1.5 * dpiondt * deni;
  double x = abar * abar * std::sqrt(abar) * deni / avo;
  double s = ((2.0 * pi * amu * kerg) / (hpl * hpl)) * temp;
  const double z = x * s * std::sqrt(s);
  const double y = std::log(z);
  const double tempi = 1.0 / temp;
  const double sion = (pion * deni + eion) * tempi + kerg * avo * ytot1 * y;
  const double dsiondt =
      (dpiondt * deni + deiondt) * tempi - (pion * deni + eion) * tempi * tempi + 1.5 * kerg * avo * tempi * ytot1;
  const double din = ye * den;
  if ((temp > H.t[JMAX - 1]) || (temp < H.t[0]) || (din > H.d[IMAX - 1]) || (din < H.d[0])) return 8;
  int jat = int((std::log10(temp) - H.tlo) * H.tstpi) + 1;
  jat = (int)mx(1, mn(jat, JMAX - 1));
  int iat = int((std::log10(din) - H.dlo) * H.dstpi) + 1;
  iat = (int)mx(1, mn(iat, IMAX - 1));
  --jat;
  --iat;
  const std::vector<double>* tabs[9] = {&H.f, &H.ft, &H.ftt, &H.fd, &H.fdd, &H.fdt, &H.fddt, &H.fdtt, &H.fddtt};
  double fi[36];
  for (int qn = 0; qn < 9; qn++) {
    const std::vector<double>& A = *tabs[qn];
    fi[4 * qn + 0] = A[(size_t)iat * JMAX + jat];
    fi[4 * qn + 1] = A[(size_t)(iat + 1) * JMAX + jat];
    fi[4 * qn + 2] = A[(size_t)iat * JMAX + jat + 1];
    fi[4 * qn + 3] = A[(size_t)(iat + 1) * JMAX + jat + 1];
  }
  const double xt = mx((temp - H.t[jat]) * H.dti_sav[jat], 0.0);
  const double xd = mx((din - H.d[iat]) * H.ddi_sav[iat], 0.0);
  const double mxt = 1.0 - xt, mxd = 1.0 - xd;
  const double w_t[6] = {psi0(xt), psi1(xt) * H.dt_sav[jat], psi2(xt) * H.dt2_sav[jat],
                         psi0(mxt), -psi1(mxt) * H.dt_sav[jat], psi2(mxt) * H.dt2_sav[jat]};
  const double w_d[6] = {psi0(xd), psi1(xd) * H.dd_sav[iat], psi2(xd) * H.dd2_sav[iat],
                         psi0(mxd), -psi1(mxd) * H.dd_sav[iat], psi2(mxd) * H.dd2_sav[iat]};
  const double dw_t[6] = {dpsi0(xt) * H.dti_sav[jat], dpsi1(xt), dpsi2(xt) * H.dt_sav[jat],
                          -dpsi0(mxt) * H.dti_sav[jat], dpsi1(mxt), -dpsi2(mxt) * H.dt_sav[jat]};
  const double dw_d[6] = {dpsi0(xd) * H.ddi_sav[iat], dpsi1(xd), dpsi2(xd) * H.dd_sav[iat],
                          -dpsi0(mxd) * H.ddi_sav[iat], dpsi1(mxd), -dpsi2(mxd) * H.dd_sav[iat]};
  const double ddw_t[6] = {ddpsi0(xt) * H.dt2i_sav[jat], ddpsi1(xt) * H.dti_sav[jat], ddpsi2(xt),
                           ddpsi0(mxt) * H.dt2i_sav[jat], -ddpsi1(mxt) * H.dti_sav[jat], ddpsi2(mxt)};
  const double ddw_d[6] = {ddpsi0(xd) * H.dd2i_sav[iat], ddpsi1(xd) * H.ddi_sav[iat], ddpsi2(xd),
                           ddpsi0(mxd) * H.dd2i_sav[iat], -ddpsi1(mxd) * H.ddi_sav[iat], ddpsi2(mxd)};
  const double free = h5(fi, w_t, w_d);
  const double df_d = h5(fi, w_t, dw_d);
  const double df_t = h5(fi, dw_t, w_d);
  const double df_dd = h5(fi, w_t, ddw_d);
  const double df_tt = h5(fi, ddw_t, w_d);
  const double df_dt = h5(fi, dw_t, dw_d);
  x = din * din;
  const double pele = x * df_d;
  const double dpepdd = ye * (x * df_dd + 2.0 * din * df_d);
  const double dpepdt = x * df_dt;
  const double sele = -df_t * ye;
  const double dsepdt = -df_tt * ye;
  const double eele = ye * free + temp * sele;
  const double deepdt = temp * dsepdt;
  q.p = pele + pion;
  q.e = eele + eion;
  q.s = sele + sion;
  q.dsdt = dsiondt + dsepdt;
  q.cv = deiondt + deepdt;
  const double dpgasdd = dpiondd + dpepdd;
  const double dpgasdt = dpiondt + dpepdt;
  x = (q.p * deni * deni) * (dpgasdt / q.cv) + dpgasdd;
  int err = 0;
  if (q.cv < 0.0) {
    err = 16;
    q.cs = NAN;
  } else if (x < 0.0) {
    err = 32;
    q.cs = 0.0;
  } else {
    q.cs = std::sqrt(x);
  }
  to_code(q);
  return err;
}

// helmholtz_compute_T, energy form (:436-486); toler is the reference's 1e-9 unless a caller sets it
int compute_T(St& q, double toler = 1.0e-9) {
  double T0 = q.T;
  const double e0 = q.e;
  q.T = TMIN;
  int err = eos(q);
  q.T = T0;
  double a = 1.0;
  if (q.e >= e0) {
    q.T = TMIN;
  } else {
    int iters = 0;
    double f;
    do {
      T0 = mx(q.T, TMIN);
      err |= eos(q);
      f = q.e - e0;
      const double dT = -f / q.cv;
      q.T += a * dT;
      q.T = mx(q.T, (T0 + TMIN) / 2.0);
      iters++;
      if (iters % 10 == 0) a *= 0.9;
      if (q.T < TMIN) return err | 128;
    } while (std::fabs(f / e0) >= toler && iters < 100);
    err |= eos(q);
    if (iters >= 100) err |= 64;
  }
  if (q.T < TMIN) err |= 128;
  return err;
}

}  // namespace

namespace orc {
// one eos_t state (10 doubles) for the rotating-state restatement
int helm_eos_state(double* r) {
  St q{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9]};
  const int e = eos(q);
  const double o[10] = {q.p, q.e, q.cs, q.cv, q.abar, q.zbar, q.rho, q.T, q.s, q.dsdt};
  for (int c = 0; c < 10; c++) r[c] = o[c];
  return e;
}
int helm_compute_T_state(double* r, double toler) {
  St q{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9]};
  const int e = compute_T(q, toler);
  const double o[10] = {q.p, q.e, q.cs, q.cv, q.abar, q.zbar, q.rho, q.T, q.s, q.dsdt};
  for (int c = 0; c < 10; c++) r[c] = o[c];
  return e;
}
}  // namespace orc

extern "C" {

void orc_helm_table(const double* rows) {
  Table& H = T_;
  const size_t nn = (size_t)IMAX * JMAX;
  std::vector<double>* tabs[9] = {&H.f, &H.fd, &H.ft, &H.fdd, &H.ftt, &H.fdt, &H.fddt, &H.fdtt, &H.fddtt};
  for (auto* v : tabs) v->assign(nn, 0.0);
  H.tlo = 3.0;
  const double thi = 13.0;
  const double tstp = (thi - H.tlo) / double(JMAX - 1);
  H.tstpi = 1.0 / tstp;
  H.dlo = -12.0;
  const double dhi = 15.0;
  const double dstp = (dhi - H.dlo) / double(IMAX - 1);
  H.dstpi = 1.0 / dstp;
  for (int j = 0; j < JMAX; j++) {
    H.t[j] = std::pow(10.0, H.tlo + j * tstp);
    for (int i = 0; i < IMAX; i++) {
      H.d[i] = std::pow(10.0, H.dlo + i * dstp);
      const double* r = rows + ((size_t)j * IMAX + i) * 9;
      for (int c = 0; c < 9; c++) (*tabs[c])[(size_t)i * JMAX + j] = r[c];
    }
  }
  for (int j = 0; j < JMAX - 1; j++) {
    const double dth = H.t[j + 1] - H.t[j];
    const double dt2 = dth * dth;
    H.dt_sav[j] = dth;
    H.dt2_sav[j] = dt2;
    H.dti_sav[j] = 1.0 / dth;
    H.dt2i_sav[j] = 1.0 / dt2;
  }
  for (int i = 0; i < IMAX - 1; i++) {
    const double dd = H.d[i + 1] - H.d[i];
    const double dd2 = dd * dd;
    H.dd_sav[i] = dd;
    H.dd2_sav[i] = dd2;
    H.ddi_sav[i] = 1.0 / dd;
    H.dd2i_sav[i] = 1.0 / dd2;
  }
  H.ready = true;
}

void orc_helm_set_units(double cm, double g, double s, double K) {
  Table& H = T_;
  H.cm = cm;
  H.g = g;
  H.s = s;
  H.K = K;
  H.erg = g * cm * cm / s / s;
  H.dyne = g * cm / s / s;
}

int orc_helm_eos(int n, double* st, int* flags) {
  if (!T_.ready) return 1;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++) {
    St q;
    double* r = st + (size_t)k * 10;
    q = St{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9]};
    const int e = eos(q);
    const double o[10] = {q.p, q.e, q.cs, q.cv, q.abar, q.zbar, q.rho, q.T, q.s, q.dsdt};
    for (int c = 0; c < 10; c++) r[c] = o[c];
    if (flags) flags[k] = e;
  }
  return 0;
}

int orc_helm_compute_T(int n, double* st, int* flags) {
  if (!T_.ready) return 1;
#pragma omp parallel for schedule(dynamic, 64)
  for (int k = 0; k < n; k++) {
    double* r = st + (size_t)k * 10;
    St q{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9]};
    const int e = compute_T(q);
    const double o[10] = {q.p, q.e, q.cs, q.cv, q.abar, q.zbar, q.rho, q.T, q.s, q.dsdt};
    for (int c = 0; c < 10; c++) r[c] = o[c];
    if (flags) flags[k] = e;
  }
  return 0;
}

}  // extern "C"
