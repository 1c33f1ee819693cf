// Launch-shape autotuner and hipGraph capture of whole RK steps.
//
// Autotuner: the sweep kernels walk the batch with a grid stride, so the grid can be one
// workgroup per block or a persistent grid of k workgroups per CU.  amrh_autotune times every
// candidate on the context's own data with HIP events and keeps the fastest for each kernel.
//
// Graphs: a step (CFL scan -> device dt, store, rk substeps) launches a fixed kernel sequence whose
// buffer pointers depend only on which state buffer holds the state when it starts, so one graph
// per starting buffer (two, or three for the rotating octree's three buffers) replays any number
// of steps with no host work but one launch each; the capture records the buffer a step ends in.
// dt is read from device memory by every kernel, so replays need no re-capture.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// the median time of `reps` (at most 8) back-to-back launches of fn on the context's stream, in
// ms, after one warm-up launch (events between the launches: one slow launch does not decide)
template <typename F>
int time_launches(amrh_ctx* c, int reps, F fn, float* ms) {
  constexpr int kMax = 8;
  reps = reps < 1 ? 1 : (reps > kMax ? kMax : reps);
  hipEvent_t ev[kMax + 1];
  int made = 0;
  int rc = AMRH_OK;
  for (; made <= reps; made++)
    if (hipEventCreate(&ev[made]) != hipSuccess) {
      rc = fail(AMRH_EHIP, "hipEventCreate failed");
      break;
    }
  if (rc == AMRH_OK) rc = fn();  // warm-up
  for (int r = 0; rc == AMRH_OK && r < reps; r++) {
    if (hipEventRecord(ev[r], c->stream) != hipSuccess) rc = fail(AMRH_EHIP, "hipEventRecord failed");
    if (rc == AMRH_OK) rc = fn();
  }
  if (rc == AMRH_OK && (hipEventRecord(ev[reps], c->stream) != hipSuccess || hipEventSynchronize(ev[reps]) != hipSuccess))
    rc = fail(AMRH_EHIP, "timing event failed");
  float t[kMax];
  for (int r = 0; rc == AMRH_OK && r < reps; r++)
    if (hipEventElapsedTime(&t[r], ev[r], ev[r + 1]) != hipSuccess) rc = fail(AMRH_EHIP, "hipEventElapsedTime failed");
  for (int i = 0; i < made; i++) (void)hipEventDestroy(ev[i]);
  if (rc != AMRH_OK) return rc;
  std::sort(t, t + reps);
  *ms = t[reps / 2];
  return AMRH_OK;
}

const double kBetas[3][3] = {{1.0, 0.0, 0.0}, {1.0, 0.5, 0.0}, {1.0, 0.25, 2.0 / 3.0}};

// HIP events between the launches of one step, for amrh_step_timed (null: not timed)
struct Marks {
  hipEvent_t ev[kMaxStepLaunches + 1] = {};
  int kind[kMaxStepLaunches] = {};
  int n = 0;     // launches marked
  bool open = false;
};

int mark_begin(amrh_ctx* c, Marks* mk) {
  if (!mk) return AMRH_OK;
  if (mk->n >= kMaxStepLaunches) return fail(AMRH_EINVAL, "more than %d launches in a step", kMaxStepLaunches);
  if (!mk->open) {
    HIPCHK(hipEventRecord(mk->ev[mk->n], c->stream));
    mk->open = true;
  }
  return AMRH_OK;
}

// the launch just enqueued ends at this event, and the next one starts there
int mark_end(amrh_ctx* c, Marks* mk, int kind) {
  if (!mk) return AMRH_OK;
  if (mk->n >= kMaxStepLaunches) return fail(AMRH_EINVAL, "more than %d stages in a step", kMaxStepLaunches);
  mk->kind[mk->n] = kind;
  mk->n++;
  HIPCHK(hipEventRecord(mk->ev[mk->n], c->stream));
  return AMRH_OK;
}

int ensure_dbuf(amrh_ctx* c) {
  if (c->dbuf) return AMRH_OK;
  const size_t n3 = (size_t)c->cfg.n * c->cfg.n * c->cfg.n;
  HIPCHK(hipMalloc(&c->dbuf, sizeof(double) * n3 * c->cfg.nf * c->nb));
  return AMRH_OK;
}

// The RK step of a self-ghosting flat batch with its dt deferred: the first substep forms the
// fluxes of U^n, reduces their KT face speeds -- the same numbers Grid::max_dt's scan forms,
// grid/grid.cpp:422-483 -- into the CFL maximum, and stores the divergence D instead of U1;
// the dt kernel then applies HydroDriver's control (dt = cfl/max, capped at growth times the last
// dt, euler/hydro_driver.cpp:107-113); the second substep forms U1 = floor(U^n + D dt) for every
// cell it reads, with U^n as its RK base.  The buffers end as after scan + store + rk substeps
// (U^n in the RK base, the state in the other buffer), and every cell has the same bits.
// The rotating state's first stage needs the floor's Newton solve for T; formed for every staged
// cell it would cost the substep three times over, so an elementwise pass forms it once per cell
// after the dt kernel and the remaining substeps are plain.  The block kernel (PPM) stages each
// block once per substep, so its first stage is the same elementwise pass.
bool sweep_deferred(const amrh_ctx* c) {
  return self_ghosting(c) && c->cfg.physics == AMRH_PHYS_EULER && c->cfg.recon == AMRH_RECON_MINMOD;
}
bool rot_deferred(const amrh_ctx* c) { return self_ghosting(c) && c->cfg.physics == AMRH_PHYS_ROTATING; }
bool block_deferred(const amrh_ctx* c) {
  return self_ghosting(c) && c->cfg.physics == AMRH_PHYS_EULER && c->cfg.recon != AMRH_RECON_MINMOD;
}
// the deferred step with an elementwise first-stage pass
bool pass_deferred(const amrh_ctx* c) { return rot_deferred(c) || block_deferred(c); }
// the rotating state on the octree: the deferred substep level by level, the pass over every node
bool tree_deferred(const amrh_ctx* c) { return c->tree && c->cfg.physics == AMRH_PHYS_ROTATING; }
bool deferred_capable(const amrh_ctx* c) { return sweep_deferred(c) || pass_deferred(c) || tree_deferred(c); }

// the deferred substep of a pass_deferred batch: fluxes of Uin, the CFL maximum into gmax_dev
// (cfl = false, the rotating state's split later substeps: not formed), D
hipError_t pass_first_launch(const amrh_ctx* c, double* Uin, int grid, bool cfl = true) {
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  if (rot_deferred(c)) {
    RotArgs r = ctx_rot_args(c);
    r.Uin = Uin;
    r.dbuf = c->dbuf;
    r.gmax = cfl ? c->gmax_dev : nullptr;
    r.outflow = 1;
    r.z_reflect = c->cfg.z_reflect;
    r.beta = 1.0;
    return launch_rot_deferred(c->cfg.g, c->cfg.recon, c->cfg.eos, mixed, r, grid, c->stream);
  }
  SubstepArgs a = flat_args(c);
  a.Uin = Uin;
  a.dbuf = c->dbuf;
  a.gmax = c->gmax_dev;
  a.beta = 1.0;
  return launch_block_deferred(c->cfg.n, c->cfg.g, c->cfg.recon, a, mixed, grid, c->stream);
}

// its stage pass: Uout = floor((Uin + D dt) 1) for every interior cell (the rotating state's
// later substeps: floor((Uin + D dt) beta + U0 (1 - beta)))
hipError_t pass_stage_launch(const amrh_ctx* c, double* Uin, double* Uout, const double* dt_dev, double beta = 1.0,
                             double* U0 = nullptr) {
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  if (rot_deferred(c)) {
    RotArgs r = ctx_rot_args(c);
    r.Uin = Uin;
    r.Uout = Uout;
    r.U0 = U0;
    r.dbuf = c->dbuf;
    r.dt_dev = dt_dev;
    r.beta = beta;
    return launch_rot_stage(c->cfg.g, c->cfg.eos, mixed, r, c->stream);
  }
  if (beta != 1.0) return hipErrorInvalidValue;
  SubstepArgs a = flat_args(c);
  a.Uin = Uin;
  a.Uout = Uout;
  a.dbuf = c->dbuf;
  a.dt_dev = dt_dev;
  a.beta = 1.0;
  return launch_block_first_stage(c->cfg.n, c->cfg.g, a, mixed, c->stream);
}
bool deferred_ok(const amrh_ctx* c, int rk) {
  return rk >= 2 && deferred_capable(c) && c->dbuf && c->schedule == AMRH_SCHEDULE_DEFERRED;
}

// the deferred step of the rotating state and of the block kernel: the deferred substep, dt, the
// first-stage pass, then plain substeps with U^n as their RK base
int step_deferred_pass(amrh_ctx* c, int rk, double* dt_dev, Marks* mk) {
  if (int rc = ctx_ready(c)) return rc;
  const int p = c->cur, q = 1 - p;
  if (int rc = mark_begin(c, mk)) return rc;
  HIPCHK(pass_first_launch(c, c->buf[p], c->grid_deferred[0]));
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_FIRST)) return rc;
  HIPCHK(launch_dt_from_gmax(c->gmax_dev, c->cfg.cfl, 1.0, 1.25, dt_dev, c->stream));
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_DT)) return rc;
  c->u0buf = p;  // the store: U^n stays where it is as the RK base
  HIPCHK(pass_stage_launch(c, c->buf[p], c->buf[q], dt_dev));
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_FIRST_STAGE)) return rc;
  c->cur = q;
  const bool split = rot_deferred(c) && (c->split_later >= 0 ? c->split_later != 0 : c->cfg.eos == AMRH_EOS_HELMHOLTZ);
  for (int s = 1; s < rk; s++) {
    if (split) {
      // the substep as D (no CFL maximum) and the stage pass, in place as the substep would be
      HIPCHK(pass_first_launch(c, c->buf[c->cur], c->grid_deferred[2], false));
      HIPCHK(pass_stage_launch(c, c->buf[c->cur], c->buf[c->cur], dt_dev, kBetas[rk - 1][s], c->buf[c->u0buf]));
    } else if (int rc = amrh_substep_dev(c, dt_dev, kBetas[rk - 1][s])) {
      return rc;
    }
    if (int rc = mark_end(c, mk, AMRH_LAUNCH_SUBSTEP)) return rc;
  }
  return AMRH_OK;
}

// the octree's (rotating state): ghost fill + deferred substep of every level, dt, the store, the
// first stage of every node and its ledger, the restriction, then the plain substeps (each with
// its ghost fill and followed by the restriction) -- the bits of ghost fill + scan + store + substeps
int step_deferred_tree(amrh_ctx* c, int rk, double* dt_dev, Marks* mk) {
  if (int rc = mark_begin(c, mk)) return rc;
  if (int rc = tree_deferred_first(c)) return rc;
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_FIRST)) return rc;
  HIPCHK(launch_dt_from_gmax(c->gmax_dev, c->cfg.cfl, 1.0, 1.25, dt_dev, c->stream));
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_DT)) return rc;
  if (int rc = amrh_store(c)) return rc;
  if (int rc = tree_first_stage(c, dt_dev)) return rc;
  if (int rc = amrh_tree_inject(c)) return rc;
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_FIRST_STAGE)) return rc;
  for (int s = 1; s < rk; s++) {
    if (int rc = tree_substep_dev(c, dt_dev, kBetas[rk - 1][s], true)) return rc;
    if (int rc = amrh_tree_inject(c)) return rc;
    if (int rc = mark_end(c, mk, AMRH_LAUNCH_SUBSTEP)) return rc;
  }
  return AMRH_OK;
}

int step_deferred(amrh_ctx* c, int rk, double* dt_dev, Marks* mk) {
  if (c->tree) return step_deferred_tree(c, rk, dt_dev, mk);
  if (pass_deferred(c)) return step_deferred_pass(c, rk, dt_dev, mk);
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  const int p = c->cur, q = 1 - p;
  SubstepArgs a = flat_args(c);
  a.U0 = c->buf[p];
  a.dbuf = c->dbuf;
  a.gmax = c->gmax_dev;
  a.beta = 1.0;
  if (int rc = mark_begin(c, mk)) return rc;
  HIPCHK(launch_substep(c->cfg.n, a, mixed, c->grid_deferred[0], c->stream, c->variant_deferred[0], 2));
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_FIRST)) return rc;
  HIPCHK(launch_dt_from_gmax(c->gmax_dev, c->cfg.cfl, 1.0, 1.25, dt_dev, c->stream));
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_DT)) return rc;
  c->u0buf = p;  // the store: U^n stays where it is as the RK base
  a.Uout = c->buf[q];
  a.dt_dev = dt_dev;
  a.beta = kBetas[rk - 1][1];
  HIPCHK(launch_substep(c->cfg.n, a, mixed, c->grid_deferred[1], c->stream, c->variant_deferred[1], 3));
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_SECOND)) return rc;
  c->cur = q;
  for (int s = 2; s < rk; s++) {
    if (int rc = amrh_substep_dev(c, dt_dev, kBetas[rk - 1][s])) return rc;
    if (int rc = mark_end(c, mk, AMRH_LAUNCH_SUBSTEP)) return rc;
  }
  return AMRH_OK;
}

// one RK step with the device dt control (scale 1, growth MAXDTINC = 1.25), as launched
int step_once(amrh_ctx* c, int rk, double* dt_dev, Marks* mk = nullptr) {
  if (deferred_ok(c, rk)) return step_deferred(c, rk, dt_dev, mk);
  if (int rc = mark_begin(c, mk)) return rc;
  if (c->tree) {
    if (int rc = amrh_tree_fill_ghosts(c)) return rc;
    if (int rc = mark_end(c, mk, AMRH_LAUNCH_GHOSTS)) return rc;
  }
  if (int rc = amrh_dt_device(c, 1.0, 1.25, dt_dev)) return rc;  // the scan and the dt kernel
  if (int rc = mark_end(c, mk, AMRH_LAUNCH_SCAN)) return rc;
  if (int rc = amrh_store(c)) return rc;
  for (int s = 0; s < rk; s++) {
    const double beta = kBetas[rk - 1][s];
    // the first substep's ghosts are the step's own (the CFL scan and store change no cell)
    const int rc = c->tree ? tree_substep_dev(c, dt_dev, beta, s > 0) : amrh_substep_dev(c, dt_dev, beta);
    if (rc) return rc;
    // the rotating state's substep leaves the restriction to the driver, after the potential
    if (c->tree && c->cfg.physics == AMRH_PHYS_ROTATING)
      if (int rc2 = amrh_tree_inject(c)) return rc2;
    if (int rc = mark_end(c, mk, AMRH_LAUNCH_SUBSTEP)) return rc;
  }
  return AMRH_OK;
}

}  // namespace

void amrh::drop_graphs(amrh_ctx* c) {
  for (int p = 0; p < 3; p++)
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
  const int other = free_buf(c);
  if (other < 0) return fail(AMRH_ESTATE, "autotune writes the spare buffer: call it before amrh_store");
  const int cus = n_cus();
  std::vector<int> cand = {0};
  for (int k : {1, 2, 3, 4, 6, 8, 12, 16})
    if (k * cus < c->nb) cand.push_back(k * cus);
  // the register-budget variants exist for the flat minmod sweep of the ideal gas
  const bool variants = c->cfg.physics == AMRH_PHYS_EULER && c->cfg.recon == AMRH_RECON_MINMOD;
  const int nvar = variants ? kSweepVariants : 1;
  // The substep is timed as every substep of a step but the first runs: beta != 1, updating its
  // input buffer in place, with the RK base U0 in another buffer (timed with U0 = its input, the
  // base's reads hit the L2 and the substep measured 10% faster than it runs in a step).  The spare
  // buffer takes a copy of the state, which dt = 0 and beta = 0.25 leave as it is (u/4 + 3u/4).
  const size_t state_bytes = sizeof(double) * (size_t)c->nb * (size_t)c->L.block();
  HIPCHK(hipMemcpyAsync(c->buf[other], c->buf[c->cur], state_bytes, hipMemcpyDeviceToDevice, c->stream));
  SubstepArgs a{};
  a.Uin = c->buf[other];
  a.Uout = c->buf[other];
  a.U0 = c->buf[c->cur];
  a.h = c->h_dev;
  a.dt = 0.0;
  a.beta = 0.25;
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
  m.outflow = self_ghosting(c) ? 1 : 0;
  m.z_reflect = c->cfg.z_reflect;
  float best_s = 1e30f, best_m = 1e30f, def_s = 0.f;
  int gs = 0, gm = 0, vs = 0;
  for (int v = 0; v < nvar; v++)
    for (int g : cand) {
      float ms = 0.f;
      if (int rc = time_launches(c, 5, [&]() -> int {
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
    if (int rc = time_launches(c, 5, [&]() -> int {
          HIPCHK(ctx_maxeig_launch(c, m, g));
          return AMRH_OK;
        }, &ms))
      return rc;
    if (ms < best_m) { best_m = ms; gm = g; }
  }
  // the beta = 1 substep at the pick (the first substep of a plain step)
  float ms_b1 = 0.f;
  {
    SubstepArgs a1 = a;
    a1.beta = 1.0;
    a1.Uin = c->buf[c->cur];  // a step's first substep writes the other buffer
    a1.U0 = c->buf[c->cur];
    if (int rc = time_launches(c, 5, [&]() -> int {
          HIPCHK(ctx_substep_launch(c, a1, gs, vs));
          return AMRH_OK;
        }, &ms_b1))
      return rc;
  }
  // the two kernels of the deferred-dt step (self-ghosting flat batches): the first substep, then
  // the one that forms its input from U^n and D; timed with dt = 0 (eig_dev[0] as the device dt),
  // writing only D, the CFL word (cleared after) and the spare buffer
  int gd[3] = {0, 0, 0}, vd[3] = {0, 0, 0};
  float msd[3][2] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
  if (pass_deferred(c)) {
    // the rotating state's and the block kernel's: the deferred substep over the grids, and the
    // first-stage pass (kernel 1)
    if (int rc = ensure_dbuf(c)) return rc;
    HIPCHK(hipMemsetAsync(c->eig_dev, 0, sizeof(double), c->stream));
    float best = 1e30f;
    for (int g : cand) {
      float ms = 0.f;
      if (int rc = time_launches(c, 5, [&]() -> int {
            HIPCHK(pass_first_launch(c, c->buf[c->cur], g));
            return AMRH_OK;
          }, &ms))
        return rc;
      if (g == 0) msd[0][0] = ms;
      if (ms < best) {
        best = ms;
        gd[0] = g;
      }
    }
    msd[0][1] = best;
    HIPCHK(launch_reset_u64(c->gmax_dev, c->stream));
    if (int rc = time_launches(c, 5, [&]() -> int {
          HIPCHK(pass_stage_launch(c, c->buf[c->cur], c->buf[other], c->eig_dev));
          return AMRH_OK;
        }, &msd[1][1]))
      return rc;
    msd[1][0] = msd[1][1];
    // an RK step is the scan and a beta = 1 substep, or the deferred substep and the first-stage
    // pass, then the same beta != 1 substeps (the dt kernel is common)
    c->schedule = msd[0][1] + msd[1][1] < best_m + ms_b1 ? AMRH_SCHEDULE_DEFERRED : AMRH_SCHEDULE_PLAIN;
    if (rot_deferred(c)) {
      // the later substeps: the substep (kernel 2's default time) against D + the stage pass
      float ms_d = 0.f, ms_st = 0.f;
      if (int rc = time_launches(c, 5, [&]() -> int {
            HIPCHK(pass_first_launch(c, c->buf[c->cur], gd[0], false));
            return AMRH_OK;
          }, &ms_d))
        return rc;
      if (int rc = time_launches(c, 5, [&]() -> int {
            HIPCHK(pass_stage_launch(c, c->buf[c->cur], c->buf[other], c->eig_dev, 0.25, c->buf[c->cur]));
            return AMRH_OK;
          }, &ms_st))
        return rc;
      gd[2] = gd[0];
      msd[2][0] = best_s;
      msd[2][1] = ms_d + ms_st;
      vd[2] = msd[2][1] < best_s ? 1 : 0;
      c->split_later = vd[2];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  } else if (sweep_deferred(c)) {
    if (int rc = ensure_dbuf(c)) return rc;
    const bool mixed = c->cfg.precision == AMRH_MIXED;
    HIPCHK(hipMemsetAsync(c->eig_dev, 0, sizeof(double), c->stream));
    for (int mode = 2; mode <= 3; mode++) {
      SubstepArgs d = flat_args(c);
      d.U0 = d.Uin;
      d.dbuf = c->dbuf;
      d.gmax = c->gmax_dev;
      d.beta = mode == 2 ? 1.0 : 0.25;
      if (mode == 3) {
        d.Uout = c->buf[other];
        d.dt_dev = c->eig_dev;
      }
      const int kx = mode - 2;
      float best = 1e30f;
      for (int v = 0; v < kSweepVariants; v++)
        for (int g : cand) {
          float ms = 0.f;
          if (int rc = time_launches(c, 5, [&]() -> int {
                HIPCHK(launch_substep(c->cfg.n, d, mixed, g, c->stream, v, mode));
                return AMRH_OK;
              }, &ms))
            return rc;
          if (v == 0 && g == 0) msd[kx][0] = ms;
          if (ms < best) {
            best = ms;
            gd[kx] = g;
            vd[kx] = v;
          }
        }
      msd[kx][1] = best;
    }
    HIPCHK(launch_reset_u64(c->gmax_dev, c->stream));
    // the schedule: an RK3 step is the scan, a beta = 1 and two beta != 1 substeps, or the two
    // deferred kernels and one beta != 1 substep (the dt kernel is common)
    const float plain = best_m + ms_b1 + 2.f * best_s, deferred = msd[0][1] + msd[1][1] + best_s;
    c->schedule = deferred < plain ? AMRH_SCHEDULE_DEFERRED : AMRH_SCHEDULE_PLAIN;
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->grid_substep = gs;
  c->grid_maxeig = gm;
  c->variant = vs;
  c->tune_ms_default = def_s;
  c->tune_ms_best = best_s;
  for (int kx = 0; kx < 3; kx++) {
    c->grid_deferred[kx] = gd[kx];
    c->variant_deferred[kx] = vd[kx];
    c->tune_ms_deferred[kx][0] = msd[kx][0];
    c->tune_ms_deferred[kx][1] = msd[kx][1];
  }
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

int amrh_deferred_launch_info(const amrh_ctx* c, int kernel, int* grid, int* variant, double* ms_default,
                              double* ms_tuned) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (kernel < 0 || kernel > 2)
    return fail(AMRH_EINVAL, "kernel must be 0 (the deferred first substep), 1 (the second kernel of the step) or 2 "
                             "(the rotating state's later substeps)");
  if (grid) *grid = c->grid_deferred[kernel];
  if (variant) *variant = c->variant_deferred[kernel];
  if (ms_default) *ms_default = c->tune_ms_deferred[kernel][0];
  if (ms_tuned) *ms_tuned = c->tune_ms_deferred[kernel][1];
  return AMRH_OK;
}

int amrh_set_schedule(amrh_ctx* c, int schedule) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (schedule != AMRH_SCHEDULE_PLAIN && schedule != AMRH_SCHEDULE_DEFERRED)
    return fail(AMRH_EINVAL, "schedule must be AMRH_SCHEDULE_PLAIN or AMRH_SCHEDULE_DEFERRED");
  c->schedule = schedule;
  drop_graphs(c);
  return AMRH_OK;
}

int amrh_get_schedule(const amrh_ctx* c, int* schedule) {
  if (!c || !schedule) return fail(AMRH_EINVAL, "null argument");
  *schedule = deferred_capable(c) ? c->schedule : AMRH_SCHEDULE_PLAIN;
  return AMRH_OK;
}

int amrh_run_steps(amrh_ctx* c, int rk, int nsteps, double* dt_dev, int use_graph) {
  if (!c || !dt_dev) return fail(AMRH_EINVAL, "null argument");
  if (rk < 1 || rk > 3 || nsteps < 0) return fail(AMRH_EINVAL, "rk must be 1..3, nsteps >= 0");
  if (!c->tree && c->cfg.boundary != AMRH_BOUNDARY_OUTFLOW)
    return fail(AMRH_ESTATE, "a flat batch runs whole steps only with AMRH_BOUNDARY_OUTFLOW");
  if (!c->have_u || !c->have_h) return fail(AMRH_ESTATE, "state or cell widths not set");
  if (int rc = ctx_ready(c)) return rc;
  if (rk >= 2 && deferred_capable(c))
    if (int rc = ensure_dbuf(c)) return rc;  // before any capture
  if (!use_graph) {
    for (int s = 0; s < nsteps; s++)
      if (int rc = step_once(c, rk, dt_dev)) return rc;
    return AMRH_OK;
  }
  if (c->step_exec_rk != rk || c->step_exec_dt != dt_dev) drop_graphs(c);
  for (int s = 0; s < nsteps; s++) {
    const int p = c->cur;
    if (!c->step_exec[p]) {
      // capture the step from this buffer; the host-side buffer bookkeeping runs during capture exactly
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
      c->step_exec_end[p] = c->cur;
      // capture did not run the step: restore the state it started from, then launch it
      c->cur = p;
    }
    HIPCHK(hipGraphLaunch(c->step_exec[p], c->stream));
    c->u0buf = p;
    c->cur = c->step_exec_end[p];
  }
  return AMRH_OK;
}

int amrh_step_timed(amrh_ctx* c, int rk, double* dt_dev, float* ms, int* kinds, int max_stages, int* nstages) {
  if (!c || !dt_dev || !ms || !kinds || !nstages) return fail(AMRH_EINVAL, "null argument");
  if (rk < 1 || rk > 3) return fail(AMRH_EINVAL, "rk must be 1..3");
  if (max_stages < 1) return fail(AMRH_EINVAL, "max_stages must be positive");
  if (!c->tree && c->cfg.boundary != AMRH_BOUNDARY_OUTFLOW)
    return fail(AMRH_ESTATE, "a flat batch runs whole steps only with AMRH_BOUNDARY_OUTFLOW");
  if (!c->have_u || !c->have_h) return fail(AMRH_ESTATE, "state or cell widths not set");
  if (int rc = ctx_ready(c)) return rc;
  if (rk >= 2 && deferred_capable(c))
    if (int rc = ensure_dbuf(c)) return rc;
  Marks mk;
  int made = 0, rc = AMRH_OK;
  for (; made <= kMaxStepLaunches && rc == AMRH_OK; made++) {
    const hipError_t e = hipEventCreate(&mk.ev[made]);
    if (e != hipSuccess) rc = fail(AMRH_EHIP, "hipEventCreate: %s", hipGetErrorString(e));
  }
  if (rc == AMRH_OK) rc = step_once(c, rk, dt_dev, &mk);
  if (rc == AMRH_OK && mk.n > max_stages) rc = fail(AMRH_EINVAL, "the step has %d stages, more than max_stages", mk.n);
  if (rc == AMRH_OK) {
    const hipError_t e = hipEventSynchronize(mk.ev[mk.n]);
    if (e != hipSuccess) rc = fail(AMRH_EHIP, "hipEventSynchronize: %s", hipGetErrorString(e));
  }
  for (int k = 0; rc == AMRH_OK && k < mk.n; k++) {
    const hipError_t e = hipEventElapsedTime(&ms[k], mk.ev[k], mk.ev[k + 1]);
    if (e != hipSuccess) rc = fail(AMRH_EHIP, "hipEventElapsedTime: %s", hipGetErrorString(e));
    kinds[k] = mk.kind[k];
  }
  if (rc == AMRH_OK) *nstages = mk.n;
  for (int k = 0; k < made; k++)
    if (mk.ev[k]) (void)hipEventDestroy(mk.ev[k]);
  return rc;
}

}  // extern "C"
