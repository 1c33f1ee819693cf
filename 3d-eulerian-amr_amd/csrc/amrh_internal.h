// Launchers shared between the kernel translation units and the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "kt_kernels.h"

namespace amrh {

// grid <= 0: one workgroup per block; otherwise a persistent grid of that many workgroups.
// variant (flat batches of the minmod sweep): 0 the default register budget, 1..3 the kernel held to
// 2, 3, 4 waves per SIMD
constexpr int kSweepVariants = 4;
// the most timed stages one RK step has (amrh_step_timed)
constexpr int kMaxStepLaunches = 16;
// mode (the sweep kernel's flat batches): 0 the substep; 2 the first substep of an RK step with dt
// deferred (fluxes, max wavespeed into a.gmax, divergence into a.dbuf); 3 the next substep, whose
// input is that first stage formed from U^n = a.Uin and a.dbuf with the device dt (a.U0 = a.Uin)
hipError_t launch_substep(int n, const SubstepArgs& a, bool mixed, int grid, hipStream_t s, int variant = 0,
                          int mode = 0);
hipError_t launch_maxeig(int n, const MaxEigArgs& m, bool mixed, int grid, hipStream_t s);
hipError_t launch_outflow_ghosts(int n, int g, double* U, int nblocks, int z_reflect, hipStream_t s);
hipError_t launch_pack(int n, int g, const double* pad, double* rec, long nrec_elems, bool to_rec, hipStream_t s);

// the block kernel (kt_block.hip): n = 8 with PPM (g = 3) or minmod (g = 2)
bool block_scheme_supported(int n, int g, int recon);
hipError_t launch_block_substep(int n, int g, int recon, const SubstepArgs& a, bool mixed, int grid, hipStream_t s);
hipError_t launch_block_maxeig(int n, int g, int recon, const MaxEigArgs& m, bool mixed, int grid, hipStream_t s);
// the deferred-dt step of a flat outflow batch on the block kernel: the first substep with D and the
// max face speed (a.dbuf, a.gmax), then the elementwise first stage from a.Uin = U^n into a.Uout
hipError_t launch_block_deferred(int n, int g, int recon, const SubstepArgs& a, bool mixed, int grid,
                                 hipStream_t s);
hipError_t launch_block_first_stage(int n, int g, const SubstepArgs& a, bool mixed, hipStream_t s);
hipError_t launch_dt_from_gmax(unsigned long long* gmax, double cfl, double scale, double growth, double* dt,
                               hipStream_t s);
hipError_t launch_reset_u64(unsigned long long* p, hipStream_t s);

}  // namespace amrh
