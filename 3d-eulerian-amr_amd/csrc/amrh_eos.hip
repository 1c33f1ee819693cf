// C ABI of the Helmholtz EOS: the device table and batch evaluation of helmholtz_eos /
// helmholtz_compute_T (helmholtz/helmeos.cpp).  The kernels of the rotating-frame state use the
// same device table through amrh_eos_device().
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "amrh_ctx.h"
#include "amrh_eos.h"
#include "amrhydro.h"
#include "helmholtz.h"

using namespace amrh;

struct amrh_eos {
  helm::HelmDev H{};
  double* cell = nullptr;
  float* cell32 = nullptr;  // the FP32 form's coefficients (helm::p_cs32)
  double* small = nullptr;  // t, d and the eight delta arrays
  int* err = nullptr;
  double cm = 1.0, g = 1.0, s = 1.0, K = 1.0;
};

namespace {

using helm::IMAX;
using helm::JMAX;

void set_factors(amrh_eos* E) {
  // helmholtz_set_cgs_units (helmholtz/helmeos.cpp:133-142) and the factors eos_from_code applies
  const double cm = E->cm, g = E->g, s = E->s, K = E->K;
  const double erg = g * cm * cm / s / s;
  const double dyne = g * cm / s / s;
  E->H.f_p = dyne / pow(cm, 2);
  E->H.f_e = erg / g;
  E->H.f_T = K;
  E->H.f_cs = cm / s;
  E->H.f_cv = erg / g / K;
  E->H.f_rho = g / pow(cm, 3);
  E->H.f_s = erg / K / g;
  E->H.f_dsdt = erg / K / K / g;
}

// mode 0: helmholtz_eos; 1: helmholtz_compute_T; 2: the mixed form's FP32 face EOS (p and cs only)
__global__ void eos_kernel(helm::HelmDev H, helm::Eos* st, int n, int mode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  helm::Eos q = st[i];
  if (mode == 1) {
    helm::compute_T(H, q);
  } else if (mode == 2) {
    float p, cs;
    helm::p_cs32(H, (float)q.rho, (float)q.T, (float)q.abar, (float)q.zbar, p, cs);
    q.p = p;
    q.cs = cs;
  } else {
    helm::eos(H, q);
  }
  st[i] = q;
}

int eos_batch(amrh_eos* E, int n, void* st_dev, void* stream, int mode) {
  if (!E || !st_dev) return fail(AMRH_EINVAL, "null argument");
  if (n < 0) return fail(AMRH_EINVAL, "negative count");
  if (n == 0) return AMRH_OK;
  const unsigned grid = (unsigned)((n + 255) / 256);
  eos_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(E->H, (helm::Eos*)st_dev, n, mode);
  HIPCHK(hipGetLastError());
  return AMRH_OK;
}

}  // namespace

const helm::HelmDev* amrh::eos_device(const amrh_eos* E) { return E ? &E->H : nullptr; }

extern "C" {

int amrh_eos_create(const double* rows, amrh_eos** out) {
  if (!rows || !out) return fail(AMRH_EINVAL, "null argument");
  *out = nullptr;
  // node coordinates and deltas exactly as read_helm_table (helmholtz/helmeos.cpp:68-131)
  const double tlo = 3.0, thi = 13.0, dlo = -12.0, dhi = 15.0;
  const double tstp = (thi - tlo) / double(JMAX - 1), tstpi = 1.0 / tstp;
  const double dstp = (dhi - dlo) / double(IMAX - 1), dstpi = 1.0 / dstp;
  std::vector<double> small((size_t)JMAX * 5 + (size_t)IMAX * 5 + 2, 0.0);
  double* t = small.data();
  double* dt_sav = t + JMAX;
  double* dt2_sav = dt_sav + JMAX;
  double* dti_sav = dt2_sav + JMAX;
  double* dt2i_sav = dti_sav + JMAX;
  double* d = dt2i_sav + JMAX;
  double* dd_sav = d + IMAX;
  double* dd2_sav = dd_sav + IMAX;
  double* ddi_sav = dd2_sav + IMAX;
  double* dd2i_sav = ddi_sav + IMAX;
  for (int j = 0; j < JMAX; j++) t[j] = pow(10.0, tlo + j * tstp);
  for (int i = 0; i < IMAX; i++) d[i] = pow(10.0, dlo + i * dstp);
  for (int j = 0; j < JMAX - 1; j++) {
    const double dth = t[j + 1] - t[j];
    const double dt2 = dth * dth;
    dt_sav[j] = dth;
    dt2_sav[j] = dt2;
    dti_sav[j] = 1.0 / dth;
    dt2i_sav[j] = 1.0 / dt2;
  }
  for (int i = 0; i < IMAX - 1; i++) {
    const double dd = d[i + 1] - d[i];
    const double dd2 = dd * dd;
    dd_sav[i] = dd;
    dd2_sav[i] = dd2;
    ddi_sav[i] = 1.0 / dd;
    dd2i_sav[i] = 1.0 / dd2;
  }
  // per interpolation cell, the 36 coefficients in the order the biquintic sum reads them
  // (helmholtz/helmeos.cpp:218-253): quantity q in (f, ft, ftt, fd, fdd, fdt, fddt, fdtt, fddtt),
  // corners (i,j), (i+1,j), (i,j+1), (i+1,j+1).  The file's columns are
  // f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt (:91-92).
  static const int col_of_q[9] = {0, 2, 4, 1, 3, 5, 6, 7, 8};
  const size_t ncell = (size_t)(IMAX - 1) * (JMAX - 1);
  std::vector<double> cell(ncell * 36);
  for (int j = 0; j < JMAX - 1; j++)
    for (int i = 0; i < IMAX - 1; i++) {
      double* c = &cell[((size_t)j * (IMAX - 1) + i) * 36];
      for (int q = 0; q < 9; q++) {
        const int col = col_of_q[q];
        c[4 * q + 0] = rows[((size_t)j * IMAX + i) * 9 + col];
        c[4 * q + 1] = rows[((size_t)j * IMAX + i + 1) * 9 + col];
        c[4 * q + 2] = rows[((size_t)(j + 1) * IMAX + i) * 9 + col];
        c[4 * q + 3] = rows[((size_t)(j + 1) * IMAX + i + 1) * 9 + col];
      }
    }
  // the FP32 form (helm::p_cs32): per cell, the interpolant in monomials a_mn xt^m xd^n.  With the
  // Hermite data F[T][D] (T = 2 b + s_t, D = 2 a + s_d: the coefficient of derivative orders (a in d,
  // b in T) at node sides (s_d, s_t), times dd^a dt^b) and M[m][2 o + s] the x^m coefficient of the
  // weight of order o at side s -- psi_o(x) for s = 0, (-1)^o psi_o(1 - x) for s = 1
  // (helmholtz/helmeos.cpp:14-59, :255-321) -- the cell is a = M F M^T.  a_00 is stored as 0 (only
  // derivatives are formed); the cell is normalised by 2^E, E in slot 36.
  static const int q_of[3][3] = {{0, 1, 2}, {3, 5, 7}, {4, 6, 8}};  // [a][b] -> q
  double M[6][6] = {};
  {
    const double psi[3][6] = {{1.0, 0.0, 0.0, -10.0, 15.0, -6.0},  // psi0 (:14-15)
                              {0.0, 1.0, 0.0, -6.0, 8.0, -3.0},     // psi1 (:20-21)
                              {0.0, 0.0, 0.5, -1.5, 1.5, -0.5}};    // psi2 (:26-27)
    const double binom[6][6] = {{1}, {1, 1}, {1, 2, 1}, {1, 3, 3, 1}, {1, 4, 6, 4, 1}, {1, 5, 10, 10, 5, 1}};
    for (int o = 0; o < 3; o++) {
      for (int m = 0; m < 6; m++) M[m][2 * o] = psi[o][m];
      // (-1)^o psi_o(1 - x) = (-1)^o sum_k c_k sum_m C(k, m) (-x)^m
      for (int k = 0; k < 6; k++)
        for (int m = 0; m <= k; m++) M[m][2 * o + 1] += ((o & 1) ? -1.0 : 1.0) * psi[o][k] * binom[k][m] * ((m & 1) ? -1.0 : 1.0);
    }
  }
  std::vector<float> cell32(ncell * helm::CELL32, 0.0f);
  for (int j = 0; j < JMAX - 1; j++)
    for (int i = 0; i < IMAX - 1; i++) {
      const size_t ci = (size_t)j * (IMAX - 1) + i;
      const double* c = &cell[ci * 36];
      double F[6][6];
      for (int b = 0; b < 3; b++)
        for (int st = 0; st < 2; st++)
          for (int a = 0; a < 3; a++)
            for (int sd = 0; sd < 2; sd++)
              F[2 * b + st][2 * a + sd] = c[4 * q_of[a][b] + sd + 2 * st] * std::pow(dt_sav[j], b) * std::pow(dd_sav[i], a);
      double FM[6][6], A[6][6], mx = 0.0;
      for (int t = 0; t < 6; t++)
        for (int n = 0; n < 6; n++) {
          double acc = 0.0;
          for (int d = 0; d < 6; d++) acc += F[t][d] * M[n][d];
          FM[t][n] = acc;
        }
      for (int m = 0; m < 6; m++)
        for (int n = 0; n < 6; n++) {
          double acc = 0.0;
          for (int t = 0; t < 6; t++) acc += M[m][t] * FM[t][n];
          A[m][n] = (m == 0 && n == 0) ? 0.0 : acc;
          mx = std::fmax(mx, std::fabs(A[m][n]));
        }
      int ex = 0;
      if (mx > 0.0 && std::isfinite(mx)) (void)std::frexp(mx, &ex);
      float* o = &cell32[ci * helm::CELL32];
      for (int m = 0; m < 6; m++)
        for (int n = 0; n < 6; n++) o[6 * m + n] = (float)std::ldexp(A[m][n], -ex);
      o[36] = (float)ex;
    }
  amrh_eos* E = new amrh_eos;
  hipError_t e = hipMalloc(&E->cell, sizeof(double) * cell.size());
  if (e == hipSuccess) e = hipMalloc(&E->cell32, sizeof(float) * cell32.size());
  if (e == hipSuccess)
    e = hipMemcpy(E->cell32, cell32.data(), sizeof(float) * cell32.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&E->small, sizeof(double) * small.size());
  if (e == hipSuccess) e = hipMalloc(&E->err, sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(E->cell, cell.data(), sizeof(double) * cell.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(E->small, small.data(), sizeof(double) * small.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(E->err, 0, sizeof(int));
  if (e != hipSuccess) {
    amrh_eos_destroy(E);
    return fail(AMRH_EHIP, "EOS table allocation failed: %s", hipGetErrorString(e));
  }
  helm::HelmDev& H = E->H;
  H.cell = E->cell;
  H.cell32 = E->cell32;
  const long off_dt = JMAX, off_d = 5L * JMAX;
  H.t = E->small;
  H.dt_sav = E->small + off_dt;
  H.dt2_sav = H.dt_sav + JMAX;
  H.dti_sav = H.dt2_sav + JMAX;
  H.dt2i_sav = H.dti_sav + JMAX;
  H.d = E->small + off_d;
  H.dd_sav = H.d + IMAX;
  H.dd2_sav = H.dd_sav + IMAX;
  H.ddi_sav = H.dd2_sav + IMAX;
  H.dd2i_sav = H.ddi_sav + IMAX;
  H.tlo = tlo;
  H.tstpi = tstpi;
  H.dlo = dlo;
  H.dstpi = dstpi;
  H.err = E->err;
  H.toler = 1.0e-9;  // helmholtz/helmeos.cpp:449
  set_factors(E);
  *out = E;
  return AMRH_OK;
}

int amrh_eos_load(const char* path, amrh_eos** out) {
  if (!path || !out) return fail(AMRH_EINVAL, "null argument");
  FILE* fp = fopen(path, "rt");
  if (!fp) return fail(AMRH_EINVAL, "cannot open EOS table %s", path);
  std::vector<double> rows((size_t)IMAX * JMAX * 9);
  for (size_t r = 0; r < (size_t)IMAX * JMAX; r++) {
    double* v = &rows[r * 9];
    if (fscanf(fp, "%le %le %le %le %le %le %le %le %le\n", v, v + 1, v + 2, v + 3, v + 4, v + 5, v + 6, v + 7,
               v + 8) != 9) {
      fclose(fp);
      return fail(AMRH_EINVAL, "EOS table %s: row %zu unreadable (need %d x %d rows of 9 values)", path, r, JMAX,
                  IMAX);
    }
  }
  fclose(fp);
  return amrh_eos_create(rows.data(), out);
}

void amrh_eos_destroy(amrh_eos* E) {
  if (!E) return;
  if (E->cell) (void)hipFree(E->cell);
  if (E->cell32) (void)hipFree(E->cell32);
  if (E->small) (void)hipFree(E->small);
  if (E->err) (void)hipFree(E->err);
  delete E;
}

int amrh_eos_set_units(amrh_eos* E, double cm, double g, double s, double K) {
  if (!E) return fail(AMRH_EINVAL, "null EOS");
  if (!(cm > 0.0 && g > 0.0 && s > 0.0 && K > 0.0)) return fail(AMRH_EINVAL, "unit factors must be positive");
  E->cm = cm;
  E->g = g;
  E->s = s;
  E->K = K;
  set_factors(E);
  return AMRH_OK;
}

int amrh_eos_eval(amrh_eos* E, int n, void* st_dev, void* stream) { return eos_batch(E, n, st_dev, stream, 0); }

int amrh_eos_compute_T(amrh_eos* E, int n, void* st_dev, void* stream) { return eos_batch(E, n, st_dev, stream, 1); }

int amrh_eos_eval_f32(amrh_eos* E, int n, void* st_dev, void* stream) { return eos_batch(E, n, st_dev, stream, 2); }

int amrh_eos_error(amrh_eos* E, int* flags) {
  if (!E || !flags) return fail(AMRH_EINVAL, "null argument");
  int v = 0;
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(&v, E->err, sizeof v, hipMemcpyDeviceToHost));
  HIPCHK(hipMemset(E->err, 0, sizeof(int)));
  *flags = v;
  return AMRH_OK;
}

}  // extern "C"
