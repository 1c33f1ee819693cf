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
  double toler;  // relative tolerance of compute_T's Newton solve (the reference's 1e-9, :449)
  // the FP32 form's coefficients (p_cs32): [(JMAX-1)*(IMAX-1)][CELL32], see p_cs32
  const float* cell32;
};

// floats per cell of the FP32 coefficients: 36, then the cell's power-of-two exponent, then padding
constexpr int CELL32 = 40;

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

// h5's term k is fi[k] * (density weight kH5D[k]) * (temperature weight kH5T[k]), the weights
// indexed {w0, w0m, w1, w1m, w2, w2m}
constexpr int kH5D[36] = {0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 2, 3, 2, 3, 4, 5,
                          4, 5, 2, 3, 2, 3, 4, 5, 4, 5, 2, 3, 2, 3, 4, 5, 4, 5};
constexpr int kH5T[36] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 0, 0, 1, 1, 0, 0,
                          1, 1, 2, 2, 3, 3, 2, 2, 3, 3, 4, 4, 5, 5, 4, 4, 5, 5};

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

  // the free energy and its derivatives (:323-351): the six biquintic sums h5 forms, in one pass
  // over the cell's 36 coefficients (each sum in h5's term order, so the same values); the
  // coefficients stream in as they are used instead of all 36 being held at once
  const double TW[3][6] = {{si0t, si0mt, si1t, si1mt, si2t, si2mt},
                           {dsi0t, dsi0mt, dsi1t, dsi1mt, dsi2t, dsi2mt},
                           {ddsi0t, ddsi0mt, ddsi1t, ddsi1mt, ddsi2t, ddsi2mt}};
  const double DW[3][6] = {{si0d, si0md, si1d, si1md, si2d, si2md},
                           {dsi0d, dsi0md, dsi1d, dsi1md, dsi2d, dsi2md},
                           {ddsi0d, ddsi0md, ddsi1d, ddsi1md, ddsi2d, ddsi2md}};
  // the sums' (temperature, density) weight orders: free, df_d, df_t, df_dd, df_tt, df_dt
  constexpr int kT[6] = {0, 0, 1, 0, 2, 1};
  constexpr int kD[6] = {0, 1, 0, 2, 0, 1};
  double r[6];
  {
    const double2* src = reinterpret_cast<const double2*>(H.cell + ((long)jat * (IMAX - 1) + iat) * 36);
#pragma unroll
    for (int k2 = 0; k2 < 18; k2++) {
      const double2 v = src[k2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int k = 2 * k2 + h;
        const double f = h ? v.y : v.x;
#pragma unroll
        for (int m = 0; m < 6; m++) {
          const double t = f * DW[kD[m]][kH5D[k]] * TW[kT[m]][kH5T[k]];
          r[m] = k == 0 ? t : r[m] + t;
        }
      }
    }
  }
  const double free = r[0], df_d = r[1], df_t = r[2], df_dd = r[3], df_tt = r[4], df_dt = r[5];

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

// ------------------------------------------------------------------------------------------------
// Pressure and sound speed in FP32, for the mixed form's face states (not bit-matched to the
// reference).  The same biquintic Hermite interpolant of the same table, written per cell in the
// monomial basis, f = sum_{m,n<6} a_mn xt^m xd^n (xt, xd in [0,1] across the cell): the host
// converts each cell's 36 Hermite coefficients in FP64.  Derivatives taken by Horner sums in this
// basis do not cancel, where the Hermite weights' second derivatives cancel terms some 10^3 times
// larger than the result -- which FP32 coefficients could not carry.  a_00 (the only coefficient a
// constant in f reaches) is stored as 0: only derivatives are formed, so the degenerate part of f
// leaves no large value to subtract.  The cell is normalised by a power of two (slot 36 holds its
// exponent) so that no table overflows FP32; the scale, the node spacings and the final p, c_s
// assembly are FP64.  Only the four sums p and c_s need (f_d, f_dd, f_dt, f_tt) are formed.
// ------------------------------------------------------------------------------------------------

// pressure and sound speed (code units) of a state given in code units; same range and sign
// checks as eos() (the flags are raised, p and cs are NaN)
__device__ inline void p_cs32(const HelmDev& H, float rho_c, float T_c, float abar_f, float zbar_f, float& p_out,
                              float& cs_out) {
  const double den = (double)rho_c * H.f_rho;
  const double temp = hmax((double)T_c * H.f_T, TMIN);
  const double abar = (double)abar_f, zbar = (double)zbar_f;
  const double ytot1 = 1.0 / abar;
  const double ye = hmax(-15.0, ytot1 * zbar);
  const double din = ye * den;
  if ((temp > H.t[JMAX - 1]) || (temp < H.t[0]) || (din > H.d[IMAX - 1]) || (din < H.d[0]) || !(den > 0.0)) {
    atomicOr(H.err, ERR_RANGE);
    p_out = cs_out = __builtin_nanf("");
    return;
  }
  // the cell: a hardware log2 estimate, then settled on the exact node values
  constexpr float LOG10_2 = 0.30102999566398120f;
  int jat = (int)(((float)__builtin_amdgcn_logf((float)temp) * LOG10_2 - (float)H.tlo) * (float)H.tstpi);
  int iat = (int)(((float)__builtin_amdgcn_logf((float)din) * LOG10_2 - (float)H.dlo) * (float)H.dstpi);
  jat = jat < 0 ? 0 : (jat > JMAX - 2 ? JMAX - 2 : jat);
  iat = iat < 0 ? 0 : (iat > IMAX - 2 ? IMAX - 2 : iat);
  if (jat > 0 && temp < H.t[jat]) jat--;
  if (jat < JMAX - 2 && temp >= H.t[jat + 1]) jat++;
  if (iat > 0 && din < H.d[iat]) iat--;
  if (iat < IMAX - 2 && din >= H.d[iat + 1]) iat++;
  const double dtis = H.dti_sav[jat], ddis = H.ddi_sav[iat];
  const float xt = (float)hmax((temp - H.t[jat]) * dtis, 0.0);
  const float xd = (float)hmax((din - H.d[iat]) * ddis, 0.0);

  // a[6 m + n]: the cell's monomial coefficients of xt^m xd^n
  float a[36];
  const float4* src = reinterpret_cast<const float4*>(H.cell32 + ((long)jat * (IMAX - 1) + iat) * CELL32);
#pragma unroll
  for (int r = 0; r < 9; r++) {
    const float4 v = src[r];
    a[4 * r] = v.x;
    a[4 * r + 1] = v.y;
    a[4 * r + 2] = v.z;
    a[4 * r + 3] = v.w;
  }
  const float ex = H.cell32[((long)jat * (IMAX - 1) + iat) * CELL32 + 36];
  float s01, s11, s02, s20;  // (T, d) derivative orders, times dt^T dd^d
  {
#pragma clang fp contract(fast)
    float p0[6], p1[6], p2[6];  // per T power m: sum_n a_mn xd^n and its first two xd-derivatives
#pragma unroll
    for (int m = 0; m < 6; m++) {
      const float* r = a + 6 * m;
      p0[m] = ((((r[5] * xd + r[4]) * xd + r[3]) * xd + r[2]) * xd + r[1]) * xd + r[0];
      p1[m] = (((5.0f * r[5] * xd + 4.0f * r[4]) * xd + 3.0f * r[3]) * xd + 2.0f * r[2]) * xd + r[1];
      p2[m] = ((20.0f * r[5] * xd + 12.0f * r[4]) * xd + 6.0f * r[3]) * xd + 2.0f * r[2];
    }
    s01 = ((((p1[5] * xt + p1[4]) * xt + p1[3]) * xt + p1[2]) * xt + p1[1]) * xt + p1[0];
    s02 = ((((p2[5] * xt + p2[4]) * xt + p2[3]) * xt + p2[2]) * xt + p2[1]) * xt + p2[0];
    s11 = (((5.0f * p1[5] * xt + 4.0f * p1[4]) * xt + 3.0f * p1[3]) * xt + 2.0f * p1[2]) * xt + p1[1];
    s20 = ((20.0f * p0[5] * xt + 12.0f * p0[4]) * xt + 6.0f * p0[3]) * xt + 2.0f * p0[2];
  }
  const double sc = ldexp(1.0, (int)ex);
  const double df_d = (double)s01 * sc * ddis;
  const double df_dd = (double)s02 * sc * H.dd2i_sav[iat];
  const double df_dt = (double)s11 * sc * ddis * dtis;
  const double df_tt = (double)s20 * sc * H.dt2i_sav[jat];

  // ions and electrons as in eos() (:186-210, :380-410), no entropy
  const double deni = 1.0 / den;
  const double xni = avo * ytot1 * den;
  const double pion = xni * kerg * temp;
  const double dpiondd = avo * ytot1 * kerg * temp;
  const double dpiondt = xni * kerg;
  const double deiondt = 1.5 * dpiondt * deni;
  const double x2 = din * din;
  const double pele = x2 * df_d;
  const double dpepdd = ye * (x2 * df_dd + 2.0 * din * df_d);
  const double dpepdt = x2 * df_dt;
  const double deepdt = -temp * df_tt * ye;
  const double p = pele + pion;
  const double cv = deiondt + deepdt;
  const double x = (p * deni * deni) * ((dpiondt + dpepdt) / cv) + (dpiondd + dpepdd);
  if (cv < 0.0) {
    atomicOr(H.err, ERR_CV);
    p_out = cs_out = __builtin_nanf("");
    return;
  }
  p_out = (float)(p / H.f_p);
  if (x < 0.0) {
    atomicOr(H.err, ERR_CS);
    cs_out = 0.0f;
  } else {
    cs_out = (float)(sqrt(x) / H.f_cs);
  }
}

// ------------------------------------------------------------------------------------------------
// Specific internal energy and heat capacity in FP64 (the mixed form's Newton solve for T; not
// bit-matched to the reference): the same interpolant, summed separably -- free, f_t and f_tt share
// one density contraction, 54 products instead of 3 x 36 in the reference's order -- and nothing
// but e and c_v (the Newton iteration reads no other output).  Returns the range flags as eos().
// ------------------------------------------------------------------------------------------------
__device__ inline bool e_cv_sep(const HelmDev& H, Eos& q) {
  const double den = q.rho * H.f_rho;
  const double temp = hmax(q.T * H.f_T, TMIN);
  const double ytot1 = 1.0 / q.abar;
  const double ye = hmax(-15.0, ytot1 * q.zbar);
  const double din = ye * den;
  if ((temp > H.t[JMAX - 1]) || (temp < H.t[0]) || (din > H.d[IMAX - 1]) || (din < H.d[0])) {
    atomicOr(H.err, ERR_RANGE);
    q.e = q.cv = __builtin_nan("");
    return false;
  }
  int jat = int((log10(temp) - H.tlo) * H.tstpi) + 1;
  jat = jat > JMAX - 1 ? JMAX - 1 : jat;
  jat = jat < 1 ? 1 : jat;
  int iat = int((log10(din) - H.dlo) * H.dstpi) + 1;
  iat = iat > IMAX - 1 ? IMAX - 1 : iat;
  iat = iat < 1 ? 1 : iat;
  --jat;
  --iat;
  const double xt = hmax((temp - H.t[jat]) * H.dti_sav[jat], 0.0);
  const double xd = hmax((din - H.d[iat]) * H.ddi_sav[iat], 0.0);
  const double mxt = 1.0 - xt, mxd = 1.0 - xd;
  const double dts = H.dt_sav[jat], dt2s = H.dt2_sav[jat], dtis = H.dti_sav[jat], dt2is = H.dt2i_sav[jat];
  const double dds = H.dd_sav[iat], dd2s = H.dd2_sav[iat];
  // density weights of the value (index of fi within a quantity's corner pair: lower, upper node)
  const double wd[6] = {psi0(xd), psi0(mxd), psi1(xd) * dds, -psi1(mxd) * dds, psi2(xd) * dd2s, psi2(mxd) * dd2s};
  // G[b][st]: the density contraction of the quantities of T-derivative order b at T node st;
  // the cell's coefficients: quantity q (f, ft, ftt, fd, fdd, fdt, fddt, fdtt, fddtt), corners
  // (i,j), (i+1,j), (i,j+1), (i+1,j+1) at 4 q + corner
  const double* c = H.cell + ((long)jat * (IMAX - 1) + iat) * 36;
  auto contract = [&](int q0, int q1, int q2, int st) {
    const double2 a = *reinterpret_cast<const double2*>(c + 4 * q0 + 2 * st);
    const double2 b = *reinterpret_cast<const double2*>(c + 4 * q1 + 2 * st);
    const double2 e = *reinterpret_cast<const double2*>(c + 4 * q2 + 2 * st);
    return a.x * wd[0] + a.y * wd[1] + b.x * wd[2] + b.y * wd[3] + e.x * wd[4] + e.y * wd[5];
  };
  // T-derivative order 0: (f, fd, fdd); 1: (ft, fdt, fddt); 2: (ftt, fdtt, fddtt)
  const double g00 = contract(0, 3, 4, 0), g01 = contract(0, 3, 4, 1);
  const double g10 = contract(1, 5, 6, 0), g11 = contract(1, 5, 6, 1);
  const double g20 = contract(2, 7, 8, 0), g21 = contract(2, 7, 8, 1);
  const double free = g00 * psi0(xt) + g01 * psi0(mxt) + g10 * (psi1(xt) * dts) + g11 * (-psi1(mxt) * dts) +
                      g20 * (psi2(xt) * dt2s) + g21 * (psi2(mxt) * dt2s);
  const double df_t = g00 * (dpsi0(xt) * dtis) + g01 * (-dpsi0(mxt) * dtis) + g10 * dpsi1(xt) + g11 * dpsi1(mxt) +
                      g20 * (dpsi2(xt) * dts) + g21 * (-dpsi2(mxt) * dts);
  const double df_tt = g00 * (ddpsi0(xt) * dt2is) + g01 * (ddpsi0(mxt) * dt2is) + g10 * (ddpsi1(xt) * dtis) +
                       g11 * (-ddpsi1(mxt) * dtis) + g20 * ddpsi2(xt) + g21 * ddpsi2(mxt);
  // ions (:186-197) and electrons (:380-391): e = eion + ye (free - T f_t), cv = deiondt - T f_tt ye
  const double kt = kerg * temp;
  const double eion = 1.5 * avo * ytot1 * kt;
  const double deiondt = 1.5 * avo * ytot1 * kerg;
  q.e = (eion + ye * (free - temp * df_t)) / H.f_e;
  q.cv = (deiondt - temp * df_tt * ye) / H.f_cv;
  return true;
}

// helmholtz_compute_T's iteration (:436-486) on e_cv_sep, FP64: the same start at T, TMIN test,
// damping schedule, floor and tolerance; the final full evaluation is skipped (it changes no T)
__device__ inline void compute_T_sep(const HelmDev& H, Eos& q) {
  double T0 = q.T;
  const double e0 = q.e;
  q.T = TMIN;  // in code units, as the reference sets eos->T (:443)
  if (!e_cv_sep(H, q)) return;
  q.T = T0;
  const double tmin = TMIN;
  double a = 1.0;
  if (q.e >= e0) {
    q.T = tmin;
    return;
  }
  int iters = 0;
  double f;
  do {
    T0 = hmax(q.T, tmin);
    if (!e_cv_sep(H, q)) return;
    f = q.e - e0;
    const double dT = -f / q.cv;
    q.T += a * dT;
    q.T = hmax(q.T, (T0 + tmin) / 2.0);
    iters++;
    if (iters % 10 == 0) a *= 0.9;
    if (q.T < tmin) {
      atomicOr(H.err, ERR_TMIN);
      return;
    }
  } while (fabs(f / e0) >= H.toler && iters < 100);
  if (iters >= 100) atomicOr(H.err, ERR_NEWTON);
}

// helmholtz_compute_T, energy form (:436-486): damped Newton on e(T) = e0 from the state's T.
// Where the reference aborts (T below TMIN, no convergence in 100 iterations) a flag is raised.
__device__ inline void compute_T(const HelmDev& H, Eos& q) {
  double T0 = q.T;
  const double e0 = q.e;
  q.T = TMIN;
  eos(H, q);
  q.T = T0;
  const double toler = H.toler;
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
