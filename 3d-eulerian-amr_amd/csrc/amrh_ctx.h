// The context behind the C ABI, shared by the translation units that implement it.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <vector>

#include "amr_kernels.h"
#include "amrh_layout.h"
#include "amrhydro.h"
#include "euler_phys.h"
#include "kt_kernels.h"
#include "kt_rot.h"

namespace amrh {

// Octree topology and the device tables derived from it (amrh_tree.hip).
struct TreeState {
  int nn = 0, maxlevel = 0;
  std::vector<int> parent, cidx, level;
  std::vector<std::array<int, 3>> loc, off;
  std::vector<std::array<int, 8>> child;
  std::vector<double> h;
  // host copies of the tables
  std::vector<int> level_nodes, level_start, level_count;
  // the rotating kernel's launch over a level with more nodes than this: a grid of this many
  // workgroups striding over the nodes (four per CU); smaller levels: one workgroup per node
  int level_grid = 1024;
  std::vector<int> item_start, item_count;  // export items per fine level
  std::vector<int> pair_start, pair_count;  // restriction pairs per parent level
  int nslots = 0;
  // device tables
  int* ghost_dev = nullptr;
  int* rel_dev = nullptr;
  int* level_nodes_dev = nullptr;
  int* slot_dev = nullptr;
  int* exp_dev = nullptr;  // [node][face][2]: the override slots its boundary-face means go to, or -1
  double* ovr_dev = nullptr;
  ExportItem* items_dev = nullptr;
  int2* pairs_dev = nullptr;
  double* ledger_dev = nullptr;
  double* fo_dev = nullptr;  // [FO | FO0]
  unsigned char* covered_dev = nullptr;  // per node: bit c set when child c exists
  void free_dev();
};

}  // namespace amrh

struct amrh_ctx {
  amrh_config cfg;
  int nb = 0;
  amrh::Layout L{8, 2, 6};
  // the state buffers: two, and a third for the rotating state on an octree, whose substep reads
  // the cells across same-level faces from the neighbours' inputs and so never updates in place
  double* buf[3] = {nullptr, nullptr, nullptr};
  int cur = 0;     // buffer holding the current stage
  int u0buf = -1;  // buffer holding the RK base, -1 = none stored
  double* h_dev = nullptr;
  double* eig_dev = nullptr;
  unsigned long long* gmax_dev = nullptr;
  int* err_dev = nullptr;
  double* pad_stage = nullptr;  // padded SoA staging for upload/download
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  bool have_h = false, have_u = false;
  amrh::TreeState* tree = nullptr;
  double origin[3] = {0.0, 0.0, 0.0};  // Grid::origin of every node of the tree (grid/grid.h:24)
  // AMRH_PHYS_ROTATING
  double* x0_dev = nullptr;    // [nb][3] block origins
  double* phi_dev = nullptr;   // [nb][(n+2)^3] potential
  const double* force_dev = nullptr;  // Poisson face forces fx | fy | fz, [3][nb][(n+2)^3] (nullptr: phi's)
  double* force_own = nullptr;        // the context's copy of host-given forces
  const amrh_eos* eos = nullptr;
  bool have_x0 = false, have_phi = false;
  double* diag_part = nullptr;  // state_sum/max/min partials (allocated on first use)
  int grid_substep = 0;  // persistent grid of the substep kernel (0: one workgroup per block)
  int grid_maxeig = 0;   // same for the CFL scan
  int variant = 0;       // register-budget variant of the flat minmod substep (launch_substep)
  float tune_ms_default = 0.f, tune_ms_best = 0.f;  // the autotuner's timings of the substep
  // the deferred-dt RK step of self-ghosting flat batches (amrh_graph.hip): the first substep's
  // divergence [nb][nf][n^3] (allocated on first use) and the launch shapes of its two kernels
  // (index 0: the deferred first substep, 1: the substep that follows it)
  double* dbuf = nullptr;
  int schedule = AMRH_SCHEDULE_DEFERRED;  // where deferred steps apply; amrh_autotune picks the faster
  int grid_deferred[3] = {0, 0, 0};
  int variant_deferred[3] = {0, 0, 0};
  float tune_ms_deferred[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};  // [kernel][default, tuned]
  // the rotating state's later substeps of a deferred step as a D-only substep and the stage pass
  // (-1: by the EOS, the Helmholtz EOS split; amrh_autotune measures it)
  int split_later = -1;
  // captured RK steps, one graph per buffer a step starts from (amrh_graph.hip), and the buffer
  // each leaves the new state in
  hipGraphExec_t step_exec[3] = {nullptr, nullptr, nullptr};
  int step_exec_end[3] = {0, 0, 0};
  int step_exec_rk = 0;
  const double* step_exec_dt = nullptr;
};

namespace amrh {

// a state buffer that holds neither the current stage nor the RK base (-1: none; with two buffers
// and the base stored apart from the current stage, a substep updates in place)
inline int free_buf(const amrh_ctx* c) {
  for (int q = 0; q < 3; q++)
    if (c->buf[q] && q != c->cur && q != c->u0buf) return q;
  return -1;
}

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
PhysParams params(const amrh_config& c);
int substep_core(amrh_ctx* c, double dt, double beta, const double* dt_dev, bool tree);
// the substep / CFL-scan kernels of the context's scheme (the column-sweep kernel for minmod,
// the block kernel for PPM); a block-kernel substep in outflow mode is followed by the ghost fill
// variant: the flat minmod substep's register-budget variant; -1 = the context's
hipError_t ctx_substep_launch(const amrh_ctx* c, const SubstepArgs& a, int grid, int variant = -1);
hipError_t ctx_maxeig_launch(const amrh_ctx* c, const MaxEigArgs& m, int grid);
// the physics' extra inputs are in place (0) or an AMRH_ESTATE failure
int ctx_ready(const amrh_ctx* c);
// ghost fill of isolated outflow blocks for the context's physics
hipError_t ctx_outflow_ghosts(const amrh_ctx* c, double* U);
// a flat batch of isolated outflow blocks: its kernels form every ghost from the boundary cell it
// copies where they read it, so the ghost slabs in memory are not maintained
bool self_ghosting(const amrh_ctx* c);
// the substep arguments of a flat batch (no octree), input = the current stage
SubstepArgs flat_args(const amrh_ctx* c);
// the rotating kernels' arguments common to every launch: potential, forces, positions, constants
RotArgs ctx_rot_args(const amrh_ctx* c);
// forget the captured step graphs (their kernel arguments hold pointers and tables that a call
// is about to replace); frees them
void drop_graphs(amrh_ctx* c);
// the low interior face of every node of the tree, Grid::xf(bw) with the context's origin, to the
// device (amrh_tree.hip)
int tree_upload_x0(amrh_ctx* c);
// amrh_tree_substep_dev; fill = false skips the ghost fill (the caller filled the current state)
int tree_substep_dev(amrh_ctx* c, const double* dt_dev, double beta, bool fill);
// the rotating state's deferred-dt step on the octree: the ghost fill and the deferred first
// substep of every level (D, the CFL maximum, reflux and ledger faces), then -- dt known, U^n
// stored -- the first stage of every node and that substep's ledger (amrh_tree.hip)
int tree_deferred_first(amrh_ctx* c);
int tree_first_stage(amrh_ctx* c, const double* dt_dev);
// state_sum / max / min reductions (amrh_diag.hip): the checks, the size of the per-block partials,
// and the two passes (no allocation; asynchronous on the context's stream)
int diag_check(const amrh_ctx* c);
long diag_part_elems(const amrh_ctx* c);
int diag_launch(const amrh_ctx* c, double* part, double* out, bool sums_layout);

}  // namespace amrh

#define HIPCHK(expr)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) return ::amrh::fail(AMRH_EHIP, "%s: %s", #expr, hipGetErrorString(e_));       \
  } while (0)
