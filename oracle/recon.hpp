// ORACLE -- test infrastructure only (see oracle.cpp).  Face reconstruction along one pencil,
// shared by the ideal-gas and rotating-state restatements.
#pragma once
#include "euler.hpp"

namespace orc {

constexpr int MAXG = 40;  // longest pencil supported
constexpr int MAXF = 16;  // most fields supported

// reconstruct.cpp:79-90  (slopes on 1..len-2, faces 2..len-2)
inline void recon_minmod(int nf, double theta, int len, double q0[][MAXF], double ql[][MAXF], double qr[][MAXF]) {
  static thread_local double slope[MAXG][MAXF];
  for (int i = 1; i < len - 1; i++) {
    for (int l = 0; l < nf; l++) {
      const double up = q0[i + 1][l] - q0[i][l];
      const double um = q0[i][l] - q0[i - 1][l];
      slope[i][l] = minmod_theta(up, um, theta);
    }
  }
  for (int i = 2; i < len - 1; i++) {
    for (int l = 0; l < nf; l++) {
      ql[i][l] = q0[i - 1][l] + slope[i - 1][l] / 2.0;
      qr[i][l] = q0[i][l] - slope[i][l] / 2.0;
    }
  }
}

// reconstruct.cpp:22-58 as called at :94 -- the routine's "ql" argument receives the caller's qr
// and vice versa.  Here A is the routine's ql (= caller qr) and B its qr (= caller ql).
// low_order / smooth: per-field bit masks (all clear for the ideal gas, euler/state.h:37-42; the
// potential field for the rotating state, binary/state_ztwd.cpp:50-56).
inline void recon_ppm(int nf, int len, double q0[][MAXF], double qlc[][MAXF], double qrc[][MAXF],
                      const unsigned low_order_mask, const unsigned smooth_mask) {
  static thread_local double s[MAXG][MAXF];
  double (*A)[MAXF] = qrc;
  double (*B)[MAXF] = qlc;
  for (int i = 1; i < len - 1; i++)
    for (int l = 0; l < nf; l++) {
      const double up = q0[i + 1][l] - q0[i][l];
      const double um = q0[i][l] - q0[i - 1][l];
      s[i][l] = minmod_theta(up, um, 2.0);
    }
  for (int i = 1; i < len - 2; i++) {
    for (int l = 0; l < nf; l++) {
      B[i][l] = (q0[i][l] + q0[i + 1][l]) * 0.5;
      if (!((low_order_mask >> l) & 1u)) B[i][l] += (s[i][l] - s[i + 1][l]) * (1.0 / 6.0);
    }
    // A[1] = B[0] reads a value the reference never initialises; it is never used for a face
    // the caller keeps, so it is left as is here.
    if (i > 1)
      for (int l = 0; l < nf; l++) A[i][l] = B[i - 1][l];
  }
  for (int i = 2; i < len - 2; i++) {
    for (int l = 0; l < nf; l++) {
      if ((smooth_mask >> l) & 1u) continue;
      const double tmp1 = B[i][l] - A[i][l];
      const double tmp2 = B[i][l] + A[i][l];
      if ((B[i][l] - q0[i][l]) * (q0[i][l] - A[i][l]) <= 0.0) {
        B[i][l] = A[i][l] = q0[i][l];
      } else if (tmp1 * (q0[i][l] - 0.5 * tmp2) > (1.0 / 6.0) * tmp1 * tmp1) {
        A[i][l] = 3.0 * q0[i][l] - 2.0 * B[i][l];
      } else if (-(1.0 / 6.0) * tmp1 * tmp1 > tmp1 * (q0[i][l] - 0.5 * tmp2)) {
        B[i][l] = 3.0 * q0[i][l] - 2.0 * A[i][l];
      }
    }
  }
  for (int i = len - 3; i > 2; i--)
    for (int l = 0; l < nf; l++) B[i][l] = B[i - 1][l];
}

}  // namespace orc
