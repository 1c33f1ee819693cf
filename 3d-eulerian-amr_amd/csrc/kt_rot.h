// Arguments and launchers of the rotating-frame state's kernels (kt_rot.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "helmholtz.h"
#include "rot_phys.h"

namespace amrh {

struct RotArgs {
  const double* Uin;      // current stage records (13 fields)
  double* Uout;           // output records
  const double* U0;       // RK base records (read when beta != 1)
  const double* phi;      // [nb][(n+2)^3] gravitational potential with one ghost layer
  const double* x0;       // [nb][3] low corner of each block's interior
  const double* h;        // [nb] cell width
  const double* dt_dev;   // device dt or nullptr
  double dt, beta;
  int nblocks;
  int* err;
  rot::RotParams c;
  helm::HelmDev H;        // the Helmholtz table (err points at the context's flags)
  double* block_eig;      // CFL scan: per-block max eigenvalue (nullptr = skip)
  unsigned long long* gmax;  // CFL scan: max over blocks of eig/h as ordered bits (nullptr = skip)
};

hipError_t launch_rot_substep(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s);
hipError_t launch_rot_maxeig(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s);
hipError_t launch_rot_ghosts(int g, double* U, const double* x0, const double* h, int nblocks, int z_reflect,
                             const rot::RotParams& c, hipStream_t s);

}  // namespace amrh
