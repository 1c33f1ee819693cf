// ORACLE support -- test infrastructure only.  Compiled in this container against the
// reference's own reconstruct.cpp / euler/state.cpp (see ref/Makefile); never shipped, never
// run on the GPU box.  It lets the tests feed one pencil of conserved states through the
// reference's Reconstruct and State exactly as Grid::compute_x_flux does (grid/grid.cpp:225-241),
// and compare the oracle's restatement with it bit for bit.
#include "reconstruct.h"
#include "state.h"

static const Reconstruct& rec() {
  static Reconstruct r;
  return r;
}

extern "C" {

int ref_gnx() { return GNX; }
int ref_bw() { return BW; }
int ref_nf() { return STATE_NF; }

// q: GNX x NF conserved (AoS).  F: GNX x NF, eig: 2*GNX, written for faces BW..GNX-BW.
void ref_pencil_flux(int dir, const double* q, double* F, double* eig) {
  State q0[GNX], ql[GNX], qr[GNX];
  _3Vec x = 0.0;
  for (int i = 0; i < GNX; i++) {
    for (int l = 0; l < STATE_NF; l++) q0[i][l] = q[i * STATE_NF + l];
    q0[i].to_prim(x);
  }
  rec()(q0, ql, qr);
  for (int i = BW; i < GNX - BW + 1; i++) {
    ql[i].to_con(x);
    qr[i].to_con(x);
    Real eL, eR;
    Vector<Real, STATE_NF> fl, fr;
    if (dir == 0) {
      eL = ql[i].max_abs_x_eigen(x);
      eR = qr[i].max_abs_x_eigen(x);
      fl = ql[i].x_flux(x);
      fr = qr[i].x_flux(x);
    } else if (dir == 1) {
      eL = ql[i].max_abs_y_eigen(x);
      eR = qr[i].max_abs_y_eigen(x);
      fl = ql[i].y_flux(x);
      fr = qr[i].y_flux(x);
    } else {
      eL = ql[i].max_abs_z_eigen(x);
      eR = qr[i].max_abs_z_eigen(x);
      fl = ql[i].z_flux(x);
      fr = qr[i].z_flux(x);
    }
    const Real a = max(eL, eR);
    const Vector<Real, STATE_NF> f = ((fl + fr) - (qr[i] - ql[i]) * a) * 0.5;
    for (int l = 0; l < STATE_NF; l++) F[i * STATE_NF + l] = f[l];
    eig[2 * i] = eL;
    eig[2 * i + 1] = eR;
  }
}

// reconstructed face states after to_con, faces BW..GNX-BW
void ref_pencil_faces(const double* q, double* qlo, double* qro) {
  State q0[GNX], ql[GNX], qr[GNX];
  _3Vec x = 0.0;
  for (int i = 0; i < GNX; i++) {
    for (int l = 0; l < STATE_NF; l++) q0[i][l] = q[i * STATE_NF + l];
    q0[i].to_prim(x);
  }
  rec()(q0, ql, qr);
  for (int i = BW; i < GNX - BW + 1; i++) {
    ql[i].to_con(x);
    qr[i].to_con(x);
    for (int l = 0; l < STATE_NF; l++) {
      qlo[i * STATE_NF + l] = ql[i][l];
      qro[i * STATE_NF + l] = qr[i][l];
    }
  }
}

void ref_floor(double* u) {
  State s;
  _3Vec x = 0.0;
  for (int l = 0; l < STATE_NF; l++) s[l] = u[l];
  s.floor(x);
  for (int l = 0; l < STATE_NF; l++) u[l] = s[l];
}

void ref_outflow(int face, double* u) {
  State s;
  _3Vec x = 0.0;
  for (int l = 0; l < STATE_NF; l++) s[l] = u[l];
  const OctFace f = face;
  s.enforce_outflow(x, f);
  for (int l = 0; l < STATE_NF; l++) u[l] = s[l];
}

double ref_pg(const double* u) {
  State s;
  for (int l = 0; l < STATE_NF; l++) s[l] = u[l];
  return s.pg();
}

}  // extern "C"
