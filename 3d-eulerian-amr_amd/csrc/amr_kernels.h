// Octree coupling kernels (amr_kernels.hip) and the tables they read.
#pragma once
#include <hip/hip_runtime.h>

#include "euler_phys.h"

namespace amrh {

// ghost source of one face of one block: ghost[(b*6 + face)*2 + {0,1}] = {type, source block}
enum GhostType { GHOST_SIB = 0, GHOST_COARSE = 1, GHOST_PHYS = 2 };

// one fine boundary face plane whose 2x2 means replace a coarse block's faces
struct ExportItem {
  int fine;  // fine block
  int dir;   // 0, 1, 2
  int side;  // 0: the fine block's lower face (plane 0), 1: its upper face (plane N)
  int slot;  // override slot written (owned by the coarse block)
};

hipError_t launch_tree_ghosts(int n, int g, double* U, const int* nodes, int count, const int* ghost, const int* rel,
                              int z_reflect, int* err, hipStream_t s);
hipError_t launch_reflux_export(int n, bool mixed, const double* U, const ExportItem* items, int count,
                                const int* slot, double* ovr, const PhysParams& P, hipStream_t s);
hipError_t launch_restrict(int n, int g, double* U, const int2* pairs, int count, const PhysParams& P,
                           hipStream_t s);
// the root's flow-off ledger of nf fields: FO = (FO + dFO dt) beta + FO0 (1 - beta), fo = [FO | FO0]
hipError_t launch_ledger(int n, int nf, const double* ledger, double h, const double* dt_dev, double dt, double beta,
                         int z_reflect, double* fo, hipStream_t s);
hipError_t launch_ledger_store(double* fo, int nf, hipStream_t s);     // FO0 = FO
hipError_t launch_ledger_restore(double* fo, int nf, hipStream_t s);   // FO = FO0

}  // namespace amrh
