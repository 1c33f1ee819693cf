// Helmholtz free-energy EOS (ions analytic, electron-positron gas from a biquintic Hermite
// interpolation of a tabulated free energy) and its Newton solve for T, on the device, FP64.
//
// Every operation follows helmholtz/helmeos.cpp:164-421 (helmholtz_eos) and :436-486
// (helmholtz_compute_T, energy form) in the reference's order, for a build without FP
// contraction, so the pressure, energy, sound speed and heat capacity are bit-identical to the
// FP64 CPU path.  The table is the reference's (IMAX x JMAX nodes over log10(ye*rho) in [-12, 15]
// and log10(T) in [3, 13], helmholtz/helmeos.cpp:7-8, :78-86) repacked per interpolation cell:
// the 36 coefficients fi[0..35] that one evaluation reads (:218-253) lie contiguous, so a call
// reads one 288-byte record instead of 36 scattered values.
#pragma once
#include <hip/hip_runtime.h>

namespace amrh {
namespace helm {

constexpr int IMAX = 1081;  // density nodes (helmholtz/helmeos.cpp:7)
constexpr int JMAX = 401;   // temperature nodes (:8)
constexpr double TMIN = 1.0e+3;  // HELMHOLTZ_TMIN (:10)

// physical constants (helmholtz/const.h:4-13), cgs
constexpr double pi = 3.1415926535897932384e0;
constexpr double h_planck = 6.6260689633e-27;
constexpr double avo = 6.0221417930e23;
constexpr double kerg = 1.380650424e-16;
constexpr double amu = 1.66053878283e-24;

// device error flags (ORed into HelmDev::err)
enum { ERR_RANGE = 8, ERR_CV = 16, ERR_CS = 32, ERR_NEWTON = 64, ERR_TMIN = 128 };

struct HelmDev {
  const double* cell;  // [(JMAX-1)*(IMAX-1)][36], cell (iat, jat) at jat*(IMAX-1) + iat
  const double* t;     // [JMAX] node temperatures
  const double* d;     // [IMAX] node densities (ye*rho)
  const double* dt_sav;
  const double* dt2_sav;
  const double* dti_sav;
  const double* dt2i_sav;
  const double* dd_sav;
  const double* dd2_sav;
  const double* ddi_sav;
  const double* dd2i_sav;
  double tlo, tstpi, dlo, dstpi;
  // unit factors of helmholtz_set_cgs_units (:133-142) as eos_from_code / eos_to_code apply them
  double f_p, f_e, f_T, f_cs, f_cv, f_rho, f_s, f_dsdt;
  int* err;
};

// eos_t (helmholtz/helmholtz.h:4-6)
struct Eos {
  double p, e, cs, cv, abar, zbar, rho, T, s, dsdt;
};

__device__ __forceinline__ double hmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double hmin(double a, double b) { return a < b ? a : b; }

// the basis functions of :14-59
__device__ __forceinline__ double psi0(double z) { return z * z * z * (z * (-6.0 * z + 15.0) - 10.0) + 1.0; }
__device__ __forceinline__ double dpsi0(double z) { return z * z * (z * (-30.0 * z + 60.0) - 30.0); }
__device__ __forceinline__ double ddpsi0(double z) { return z * (z * (-120.0 * z + 180.0) - 60.0); }
__device__ __forceinline__ double psi1(double z) { return z * (z * z * (z * (-3.0 * z + 8.0) - 6.0) + 1.0); }
__device__ __forceinline__ double dpsi1(double z) { return z * z * (z * (-15.0 * z + 32.0) - 18.0) + 1.0; }
__device__ __forceinline__ double ddpsi1(double z) { return z * (z * (-60.0 * z + 96.0) - 36.0); }
__device__ __forceinline__ double psi2(double z) { return 0.5 * z * z * (z * (z * (-z + 3.0) - 3.0) + 1.0); }
__device__ __forceinline__ double dpsi2(double z) { return 0.5 * z * (z * (z * (-5.0 * z + 12.0) - 9.0) + 2.0); }
__device__ __forceinline__ double ddpsi2(double z) { return 0.5 * (z * (z * (-20.0 * z + 36.0) - 18.0) + 2.0); }

// biquintic sum h5 (:29-47): the left-to-right sum of 36 products, each fi * wd * wt
__device__ __forceinline__ double h5(const double* fi, double w0t, double w1t, double w2t, double w0mt, double w1mt,
                                     double w2mt, double w0d, double w1d, double w2d, double w0md, double w1md,
                                     double w2md) {
  double r = fi[0] * w0d * w0t;
  r = r + fi[1] * w0md * w0t;
  r = r + fi[2] * w0d * w0mt;
  r = r + fi[3] * w0md * w0mt;
  r = r + fi[4] * w0d * w1t;
  r = r + fi[5] * w0md * w1t;
  r = r + fi[6] * w0d * w1mt;
  r = r + fi[7] * w0md * w1mt;
  r = r + fi[8] * w0d * w2t;
  r = r + fi[9] * w0md * w2t;
  r = r + fi[10] * w0d * w2mt;
  r = r + fi[11] * w0md * w2mt;
  r = r + fi[12] * w1d * w0t;
  r = r + fi[13] * w1md * w0t;
  r = r + fi[14] * w1d * w0mt;
  r = r + fi[15] * w1md * w0mt;
  r = r + fi[16] * w2d * w0t;
  r = r + fi[17] * w2md * w0t;
  r = r + fi[18] * w2d * w0mt;
  r = r + fi[19] * w2md * w0mt;
  r = r + fi[20] * w1d * w1t;
  r = r + fi[21] * w1md * w1t;
  r = r + fi[22] * w1d * w1mt;
  r = r + fi[23] * w1md * w1mt;
  r = r + fi[24] * w2d * w1t;
  r = r + fi[25] * w2md * w1t;
  r = r + fi[26] * w2d * w1mt;
  r = r + fi[27] * w2md * w1mt;
  r = r + fi[28] * w1d * w2t;
  r = r + fi[29] * w1md * w2t;
  r = r + fi[30] * w1d * w2mt;
  r = r + fi[31] * w1md * w2mt;
  r = r + fi[32] * w2d * w2t;
  r = r + fi[33] * w2md * w2t;
  r = r + fi[34] * w2d * w2mt;
  r = r + fi[35] * w2md * w2mt;
  return r;
}

// helmholtz_eos (:164-421).  On a state out of the table's range, or with a negative heat
// capacity or squared sound speed, the reference aborts; here an error flag is raised and the
// outputs are NaN.
__device__ inline void eos(const HelmDev& H, Eos& q) {
  // eos_from_code (:155-164): only rho and T feed the evaluation; the rest are overwritten
  q.p *= H.f_p;
  q.e *= H.f_e;
  q.T *= H.f_T;
  q.cs *= H.f_cs;
  q.cv *= H.f_cv;
  q.rho *= H.f_rho;
  q.s *= H.f_s;
  q.dsdt *= H.f_dsdt;

  const double den = q.rho;
  const double temp = hmax(q.T, TMIN);
  const double abar = q.abar;
  const double zbar = q.zbar;
  const double ytot1 = 1.0 / abar;
  // the reference's floor literal "1.0d - 16" evaluates to -15.0 (:177)
  const double ye = hmax(-15.0, ytot1 * zbar);

  const double deni = 1.0 / den;
  const double kt = kerg * temp;

  // ions (:186-210)
  const double xni = avo * ytot1 * den;
  const double dxnidd = avo * ytot1;
  const double pion = xni * kt;
  const double dpiondd = dxnidd * kt;
  const double dpiondt = xni * kerg;
  const double eion = 1.5 * pion * deni;
  const double deiondt = 1.5 * dpiondt * deni;
  double x = abar * abar * sqrt(abar) * deni / avo;
  double s = ((2.0 * pi * amu * kerg) / (h_planck * h_planck)) * temp;
  const double z = x * s * sqrt(s);
  const double y = log(z);
  const double tempi = 1.0 / temp;
  const double sion = (pion * deni + eion) * tempi + kerg * avo * ytot1 * y;
  const double dsiondt = (dpiondt * deni + deiondt) * tempi - (pion * deni + eion) * tempi * tempi +
                         1.5 * kerg * avo * tempi * ytot1;

  // electron-positron gas: enter the table with ye*den (:213-226)
  const double din = ye * den;
  if ((temp > H.t[JMAX - 1]) || (temp < H.t[0]) || (din > H.d[IMAX - 1]) || (din < H.d[0])) {
    atomicOr(H.err, ERR_RANGE);
    q.p = q.e = q.s = q.dsdt = q.cv = q.cs = __builtin_nan("");
    q.rho /= H.f_rho;
    q.T /= H.f_T;
    return;
  }
  int jat = int((log10(temp) - H.tlo) * H.tstpi) + 1;
  jat = jat > JMAX - 1 ? JMAX - 1 : jat;
  jat = jat < 1 ? 1 : jat;
  int iat = int((log10(din) - H.dlo) * H.dstpi) + 1;
  iat = iat > IMAX - 1 ? IMAX - 1 : iat;
  iat = iat < 1 ? 1 : iat;
  --jat;
  --iat;

  double fi[36];
  {
    const double2* src = reinterpret_cast<const double2*>(H.cell + ((long)jat * (IMAX - 1) + iat) * 36);
#pragma unroll
    for (int r = 0; r < 18; r++) {
      const double2 v = src[r];
      fi[2 * r] = v.x;
      fi[2 * r + 1] = v.y;
    }
  }

  // differences and basis functions (:255-321)
  const double xt = hmax((temp - H.t[jat]) * H.dti_sav[jat], 0.0);
  const double xd = hmax((din - H.d[iat]) * H.ddi_sav[iat], 0.0);
  const double mxt = 1.0 - xt;
  const double mxd = 1.0 - xd;
  const double dts = H.dt_sav[jat], dt2s = H.dt2_sav[jat], dtis = H.dti_sav[jat], dt2is = H.dt2i_sav[jat];
  const double dds = H.dd_sav[iat], dd2s = H.dd2_sav[iat], ddis = H.ddi_sav[iat], dd2is = H.dd2i_sav[iat];

  const double si0t = psi0(xt);
  const double si1t = psi1(xt) * dts;
  const double si2t = psi2(xt) * dt2s;
  const double si0mt = psi0(mxt);
  const double si1mt = -psi1(mxt) * dts;
  const double si2mt = psi2(mxt) * dt2s;
  const double si0d = psi0(xd);
  const double si1d = psi1(xd) * dds;
  const double si2d = psi2(xd) * dd2s;
  const double si0md = psi0(mxd);
  const double si1md = -psi1(mxd) * dds;
  const double si2md = psi2(mxd) * dd2s;

  const double dsi0t = dpsi0(xt) * dtis;
  const double dsi1t = dpsi1(xt);
  const double dsi2t = dpsi2(xt) * dts;
  const double dsi0mt = -dpsi0(mxt) * dtis;
  const double dsi1mt = dpsi1(mxt);
  const double dsi2mt = -dpsi2(mxt) * dts;
  const double dsi0d = dpsi0(xd) * ddis;
  const double dsi1d = dpsi1(xd);
  const double dsi2d = dpsi2(xd) * dds;
  const double dsi0md = -dpsi0(mxd) * ddis;
  const double dsi1md = dpsi1(mxd);
  const double dsi2md = -dpsi2(mxd) * dds;

  const double ddsi0t = ddpsi0(xt) * dt2is;
  const double ddsi1t = ddpsi1(xt) * dtis;
  const double ddsi2t = ddpsi2(xt);
  const double ddsi0mt = ddpsi0(mxt) * dt2is;
  const double ddsi1mt = -ddpsi1(mxt) * dtis;
  const double ddsi2mt = ddpsi2(mxt);
  const double ddsi0d = ddpsi0(xd) * dd2is;
  const double ddsi1d = ddpsi1(xd) * ddis;
  const double ddsi2d = ddpsi2(xd);
  const double ddsi0md = ddpsi0(mxd) * dd2is;
  const double ddsi1md = -ddpsi1(mxd) * ddis;
  const double ddsi2md = ddpsi2(mxd);

  // the free energy and its derivatives (:323-351)
  const double free = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, si0d, si1d, si2d, si0md, si1md, si2md);
  const double df_d = h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md);
  const double df_t = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, si0d, si1d, si2d, si0md, si1md, si2md);
  const double df_dd =
      h5(fi, si0t, si1t, si2t, si0mt, si1mt, si2mt, ddsi0d, ddsi1d, ddsi2d, ddsi0md, ddsi1md, ddsi2md);
  const double df_tt =
      h5(fi, ddsi0t, ddsi1t, ddsi2t, ddsi0mt, ddsi1mt, ddsi2mt, si0d, si1d, si2d, si0md, si1md, si2md);
  const double df_dt = h5(fi, dsi0t, dsi1t, dsi2t, dsi0mt, dsi1mt, dsi2mt, dsi0d, dsi1d, dsi2d, dsi0md, dsi1md, dsi2md);

  // electron-positron thermodynamics (:380-396); the cubic weights of :353-378 feed nothing here
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
  if (q.cv < 0.0) {
    atomicOr(H.err, ERR_CV);
    q.cs = __builtin_nan("");
  } else if (x < 0.0) {
    atomicOr(H.err, ERR_CS);
    q.cs = 0.0;
  } else {
    q.cs = sqrt(x);
  }
  // eos_to_code (:144-153)
  q.p /= H.f_p;
  q.e /= H.f_e;
  q.T /= H.f_T;
  q.cs /= H.f_cs;
  q.cv /= H.f_cv;
  q.rho /= H.f_rho;
  q.s /= H.f_s;
  q.dsdt /= H.f_dsdt;
}

// helmholtz_compute_T, energy form (:436-486): damped Newton on e(T) = e0 from the state's T.
// Where the reference aborts (T below TMIN, no convergence in 100 iterations) a flag is raised.
__device__ inline void compute_T(const HelmDev& H, Eos& q) {
  double T0 = q.T;
  const double e0 = q.e;
  q.T = TMIN;
  eos(H, q);
  q.T = T0;
  const double toler = 1.0e-9;
  double a = 1.0;
  if (q.e >= e0) {
    q.T = TMIN;
  } else {
    int iters = 0;
    double f;
    do {
      T0 = hmax(q.T, TMIN);
      eos(H, q);
      f = q.e - e0;
      const double dT = -f / q.cv;
      q.T += a * dT;
      q.T = hmax(q.T, (T0 + TMIN) / 2.0);
      iters++;
      if (iters % 10 == 0) a *= 0.9;
      if (q.T < TMIN) {
        atomicOr(H.err, ERR_TMIN);
        return;
      }
    } while (fabs(f / e0) >= toler && iters < 100);
    eos(H, q);
    if (iters >= 100) atomicOr(H.err, ERR_NEWTON);
  }
  if (q.T < TMIN) atomicOr(H.err, ERR_TMIN);
}

}  // namespace helm
}  // namespace amrh
