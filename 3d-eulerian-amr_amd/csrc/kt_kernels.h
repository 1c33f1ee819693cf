// Fused Kurganov-Tadmor substep kernel and its CFL twin.
//
// One workgroup per block, one thread per (i, j) column, N*N threads.  The block is swept plane by
// plane in z:
//   * each thread keeps its column's primitive states for planes k, k+1, k+2 in registers, the
//     z-slope of plane k and the z-face flux at k-1/2, so every z-face is evaluated once;
//   * the primitives of plane k, with their x/y ghost cells, are staged in LDS, and the (N+1)*N
//     x-faces and N*(N+1) y-faces of the plane are spread over all threads and written to LDS;
//   * the cell (i, j, k) then accumulates D = -(dFx + dFy + dFz)/h in that order
//     (grid/grid.cpp:304-344), forms U = (U + D*dt)*beta + U0*(1-beta), floors it
//     (grid/grid.cpp:406-420) and stores it.
// Each cell's primitives are formed once per substep (the reference redoes to_prim per sweep;
// the value is the same).  HBM traffic per block is the algorithmic count: interior + face ghost
// slabs read once, interior written once, U0 interior read when beta != 1.
#pragma once
#include "amrh_layout.h"
#include "euler_phys.h"

namespace amrh {

struct SubstepArgs {
  const double* Uin;       // current stage records
  double* Uout;            // output records (may alias Uin)
  const double* U0;        // RK base records (interior read when has_u0)
  const double* h;         // per-block cell width
  const double* dt_dev;    // device dt or nullptr
  double dt, beta;
  int nblocks;
  int block0;              // first block handled by this launch
  PhysParams P;
  int* err;
  // octree coupling (all null / -1 for a flat batch)
  const int* slot;         // [nb][3 dir][3 plane][4 quadrant] reflux override slot or -1
  const double* ovr;       // override face fluxes [slot][N/2][N/2][NF] (fine-face averages)
  int ledger_block;        // block whose domain-boundary fluxes feed the flow-off ledger
  double* ledger;          // [3 dir][2 side][N][N][NF] boundary face fluxes of that block
  // octree, level by level (the block kernel): the launch's blocks are nodes[0 .. nblocks), and
  // each writes the 2x2 means of its boundary faces to the coarse slots exp_slot[node][face][2]
  const int* nodes;
  const int* exp_slot;
  double* ovr_out;         // = ovr, written
  // isolated outflow blocks: self-ghosting (each ghost formed from the boundary cell it copies
  // where it is read; the ghost slabs are neither read nor written), z-mirrored below with z_reflect
  int outflow;
  int z_reflect;
  // an RK step with deferred dt (flat outflow batches, the sweep kernel): the divergence of the
  // first substep, interior cells only [block][NF][N^3] (written by the deferred first substep,
  // read by the next one), and the global max wavespeed/h that the deferred substep reduces into
  double* dbuf;
  unsigned long long* gmax;
};

struct MaxEigArgs {
  const double* U;
  const double* h;
  double* block_eig;           // per block max eig (nullptr = skip)
  unsigned long long* gmax;    // global max of eig/h as ordered bits (nullptr = skip)
  int nblocks;
  PhysParams P;
  int outflow;    // self-ghosting outflow blocks (as SubstepArgs::outflow; the sweep kernel)
  int z_reflect;
};

}  // namespace amrh
