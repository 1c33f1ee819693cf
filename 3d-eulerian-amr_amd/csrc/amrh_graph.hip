// Launch-shape autotuner and hipGraph capture of whole RK steps.
//
// Autotuner: the sweep kernels walk the batch with a grid stride, so the grid can be one
// workgroup per block or a persistent grid of k workgroups per CU.  amrh_autotune times every
// candidate on the context's own data with HIP events and keeps the fastest for each kernel.
//
// Graphs: a step (CFL scan -> device dt, store, rk substeps) launches a fixed kernel sequence whose
// buffer pointers depend only on which ping-pong buffer holds the state when it starts, so two
// graphs -- one per parity -- replay any number of steps with no host work but one launch each.
// dt is read from device memory by every kernel, so replays need no re-capture.
#include <hip/hip_runtime.h>

#include <vector>

#include "amrh_ctx.h"
#include "amrh_internal.h"

using namespace amrh;

namespace {

int n_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 256;
  return cus > 0 ? cus : 256;
}

// time `reps` launches of fn on the context's stream, in ms per launch
template <typename F>
int time_launches(amrh_ctx* c, int reps, F fn, float* ms) {
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  if (int rc = fn()) return rc;  // warm-up
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; r++)
    if (int rc = fn()) return rc;
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return AMRH_OK;
}

const double kBetas[3][3] = {{1.0, 0.0, 0.0}, {1.0, 0.5, 0.0}, {1.0, 0.25, 2.0 / 3.0}};

// one RK step with the device dt control (scale 1, growth MAXDTINC = 1.25), as launched
int step_once(amrh_ctx* c, int rk, double* dt_dev) {
  if (c->tree) {
    if (int rc = amrh_tree_fill_ghosts(c)) return rc;
  }
  if (int rc = amrh_dt_device(c, 1.0, 1.25, dt_dev)) return rc;
  if (int rc = amrh_store(c)) return rc;
  for (int s = 0; s < rk; s++) {
    const double beta = kBetas[rk - 1][s];
    // the first substep's ghosts are the step's own (the CFL scan and store change no cell)
    const int rc = c->tree ? tree_substep_dev(c, dt_dev, beta, s > 0) : amrh_substep_dev(c, dt_dev, beta);
    if (rc) return rc;
    // the rotating state's substep leaves the restriction to the driver, after the potential
    if (c->tree && c->cfg.physics == AMRH_PHYS_ROTATING)
      if (int rc2 = amrh_tree_inject(c)) return rc2;
  }
  return AMRH_OK;
}

}  // namespace

void amrh::drop_graphs(amrh_ctx* c) {
  for (int p = 0; p < 2; p++)
    if (c->step_exec[p]) {
      (void)hipGraphExecDestroy(c->step_exec[p]);
      c->step_exec[p] = nullptr;
    }
  c->step_exec_rk = 0;
  c->step_exec_dt = nullptr;
}

extern "C" {

int amrh_autotune(amrh_ctx* c, int* grid_substep, int* grid_maxeig) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->have_u || !c->have_h) return fail(AMRH_ESTATE, "state or cell widths not set");
  if (int rc = ctx_ready(c)) return rc;
  const int other = 1 - c->cur;
  if (c->u0buf == other) return fail(AMRH_ESTATE, "autotune writes the spare buffer: call it before amrh_store");
  const int cus = n_cus();
  std::vector<int> cand = {0};
  for (int k : {1, 2, 3, 4, 6, 8, 12, 16})
    if (k * cus < c->nb) cand.push_back(k * cus);
  // the register-budget variants exist for the flat minmod sweep of the ideal gas
  const bool variants = c->cfg.physics == AMRH_PHYS_EULER && c->cfg.recon == AMRH_RECON_MINMOD;
  const int nvar = variants ? kSweepVariants : 1;
  SubstepArgs a{};
  a.Uin = c->buf[c->cur];
  a.Uout = c->buf[other];
  a.U0 = c->buf[c->cur];
  a.h = c->h_dev;
  a.dt = 0.0;  // timing only: the update is U + 0*D (beta = 1)
  a.beta = 1.0;
  a.nblocks = c->nb;
  a.P = params(c->cfg);
  a.err = c->err_dev;
  a.ledger_block = -1;
  a.outflow = c->cfg.boundary == AMRH_BOUNDARY_OUTFLOW && !c->tree;
  a.z_reflect = c->cfg.z_reflect;
  MaxEigArgs m{};
  m.U = c->buf[c->cur];
  m.h = c->h_dev;
  m.block_eig = c->eig_dev;
  m.nblocks = c->nb;
  m.P = params(c->cfg);
  float best_s = 1e30f, best_m = 1e30f, def_s = 0.f;
  int gs = 0, gm = 0, vs = 0;
  for (int v = 0; v < nvar; v++)
    for (int g : cand) {
      float ms = 0.f;
      if (int rc = time_launches(c, 3, [&]() -> int {
            HIPCHK(ctx_substep_launch(c, a, g, v));
            return AMRH_OK;
          }, &ms))
        return rc;
      if (v == 0 && g == 0) def_s = ms;
      if (ms < best_s) {
        best_s = ms;
        gs = g;
        vs = v;
      }
    }
  for (int g : cand) {
    float ms = 0.f;
    if (int rc = time_launches(c, 3, [&]() -> int {
          HIPCHK(ctx_maxeig_launch(c, m, g));
          return AMRH_OK;
        }, &ms))
      return rc;
    if (ms < best_m) { best_m = ms; gm = g; }
  }
  c->grid_substep = gs;
  c->grid_maxeig = gm;
  c->variant = vs;
  c->tune_ms_default = def_s;
  c->tune_ms_best = best_s;
  drop_graphs(c);
  if (grid_substep) *grid_substep = gs;
  if (grid_maxeig) *grid_maxeig = gm;
  return AMRH_OK;
}

int amrh_launch_info(const amrh_ctx* c, int* grid_substep, int* grid_maxeig, int* variant, double* ms_default,
                     double* ms_tuned) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (grid_substep) *grid_substep = c->grid_substep;
  if (grid_maxeig) *grid_maxeig = c->grid_maxeig;
  if (variant) *variant = c->variant;
  if (ms_default) *ms_default = c->tune_ms_default;
  if (ms_tuned) *ms_tuned = c->tune_ms_best;
  return AMRH_OK;
}

int amrh_set_variant(amrh_ctx* c, int variant) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (variant < 0 || variant >= kSweepVariants) return fail(AMRH_EINVAL, "variant must be 0..%d", kSweepVariants - 1);
  c->variant = variant;
  drop_graphs(c);
  return AMRH_OK;
}

int amrh_set_launch(amrh_ctx* c, int grid_substep, int grid_maxeig) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (grid_substep < 0 || grid_maxeig < 0) return fail(AMRH_EINVAL, "negative grid");
  c->grid_substep = grid_substep;
  c->grid_maxeig = grid_maxeig;
  drop_graphs(c);
  return AMRH_OK;
}

int amrh_run_steps(amrh_ctx* c, int rk, int nsteps, double* dt_dev, int use_graph) {
  if (!c || !dt_dev) return fail(AMRH_EINVAL, "null argument");
  if (rk < 1 || rk > 3 || nsteps < 0) return fail(AMRH_EINVAL, "rk must be 1..3, nsteps >= 0");
  if (!c->tree && c->cfg.boundary != AMRH_BOUNDARY_OUTFLOW)
    return fail(AMRH_ESTATE, "a flat batch runs whole steps only with AMRH_BOUNDARY_OUTFLOW");
  if (!c->have_u || !c->have_h) return fail(AMRH_ESTATE, "state or cell widths not set");
  if (!use_graph) {
    for (int s = 0; s < nsteps; s++)
      if (int rc = step_once(c, rk, dt_dev)) return rc;
    return AMRH_OK;
  }
  if (c->step_exec_rk != rk || c->step_exec_dt != dt_dev) drop_graphs(c);
  for (int s = 0; s < nsteps; s++) {
    const int p = c->cur;
    if (!c->step_exec[p]) {
      // capture this parity's step; the host-side buffer bookkeeping runs during capture exactly
      // as it does when the step is launched directly
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rc = step_once(c, rk, dt_dev);
      hipGraph_t got = nullptr;
      const hipError_t ec = hipStreamEndCapture(c->stream, &got);
      if (rc) {
        if (got) (void)hipGraphDestroy(got);
        return rc;
      }
      if (ec != hipSuccess) return fail(AMRH_EHIP, "graph capture: %s", hipGetErrorString(ec));
      g = got;
      const hipError_t ei = hipGraphInstantiate(&c->step_exec[p], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) return fail(AMRH_EHIP, "graph instantiate: %s", hipGetErrorString(ei));
      c->step_exec_rk = rk;
      c->step_exec_dt = dt_dev;
      // capture did not run the step: restore the state it started from, then launch it
      c->cur = p;
    }
    HIPCHK(hipGraphLaunch(c->step_exec[p], c->stream));
    c->u0buf = p;
    c->cur = 1 - p;
  }
  return AMRH_OK;
}

}  // extern "C"
