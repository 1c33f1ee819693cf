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
  // Poisson face forces (nullptr: phi's differences): fx of every block, then fy, then fz, each
  // block (n+2)^3 with face i between cells i-1 and i; nforce = blocks per component
  const double* force;
  long nforce;
  // octree (nodes != nullptr): the launch's blocks are nodes[0 .. nblocks)
  const int* nodes;
  const int* slot;      // reflux override slot per [node][dir][plane 0, n/2, n][quadrant] or -1
  double* ovr;          // override face fluxes [slot][n/2][n/2][13]: read, and written by exports
  const int* exp_slot;  // [node][face][2]: the slots this node's boundary-face 2x2 means go to, or -1
  double* ledger;       // the root's boundary face fluxes [dir][low, high][n][n][13] (nullptr: none)
  // octree substep: per [node][face] (type, source) of the ghost exchange (amr_kernels.h); a face
  // with a same-level neighbour is read from the neighbour's cells (nullptr: every ghost from the
  // record's slabs)
  const int* ghost;
  // an isolated outflow block (flat batches): each ghost is formed from the boundary cell it copies
  // where it is staged (grid/grid_phys_bound.cpp:27-54); the ghost slabs are not read
  int outflow;
  int z_reflect;
  // the deferred-dt step (flat outflow batches): the first substep's D per interior cell,
  // [block][field][n^3] (written by the deferred substep, read by rot_stage_kernel)
  double* dbuf;
};

hipError_t launch_rot_substep(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s);
hipError_t launch_rot_maxeig(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s);
// the deferred-dt step of a flat outflow batch: the first substep with D and the max face speed
// (a.dbuf, a.gmax; a.gmax null: D only), then the elementwise stage from a.Uin (with a.beta,
// a.U0 and a.dt_dev) into a.Uout
hipError_t launch_rot_deferred(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s);
hipError_t launch_rot_stage(int g, int eos, bool mixed, const RotArgs& a, hipStream_t s);
// the same first substep over one octree level (a.nodes, with the reflux tables a.slot, a.ovr,
// a.exp_slot, a.ledger, a.ghost as for the substep)
hipError_t launch_rot_tree_deferred(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s);
hipError_t launch_rot_ghosts(int g, double* U, const double* x0, const double* h, int nblocks, int z_reflect,
                             const rot::RotParams& c, hipStream_t s);
// the octree's ghost fill of one level (ghost/rel: the tables of amr_kernels.h); covered (per node,
// bit c for child c) != nullptr skips the same-level-neighbour faces of leaves
hipError_t launch_rot_tree_ghosts(int g, double* U, const int* nodes, int count, const int* ghost, const int* rel,
                                  const double* x0, const double* h, int z_reflect, const rot::RotParams& c, int* err,
                                  hipStream_t s, const unsigned char* covered = nullptr);
// inject_from_children for one parent level: pairs (parent, child << 24 | child node)
hipError_t launch_rot_restrict(int g, int eos, double* U, const int2* pairs, int count, const double* x0,
                               const double* h, const rot::RotParams& c, const helm::HelmDev& H, hipStream_t s);
// inject_from_parent of new nodes (node, parent, child index), floored at the child cells
hipError_t launch_rot_prolong(int g, int eos, double* U, const int3* nodes, int count, const double* x0,
                              const double* h, const rot::RotParams& c, const helm::HelmDev& H, hipStream_t s);
// Binary::pot_to_grid on every block: pot = rho (phi - R^2 Omega0^2 / 2)
hipError_t launch_rot_pot_to_grid(int g, double* U, const double* phi, const double* x0, const double* h,
                                  int nblocks, const rot::RotParams& c, hipStream_t s);

}  // namespace amrh
