// C ABI of amrhydro (include/amrhydro.h): context, buffers, launches.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "amrh_ctx.h"
#include "amrh_eos.h"
#include "amrh_internal.h"
#include "amrh_layout.h"
#include "kt_rot.h"
#include "amrhydro.h"

using namespace amrh;

namespace amrh {

namespace {
thread_local std::string g_err;
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

PhysParams params(const amrh_config& c) {
  PhysParams P;
  P.gamma = c.gamma;
  P.rho_floor = c.rho_floor;
  P.ei_floor = c.ei_floor;
  P.theta = c.theta;
  P.cfl = c.cfl;
  return P;
}

}  // namespace amrh

namespace {

long padded_elems(const amrh_ctx* c) {
  const long P = c->cfg.n + 2 * c->cfg.g;
  return (long)c->nb * c->cfg.nf * P * P * P;
}
long record_elems(const amrh_ctx* c) { return (long)c->nb * c->L.block(); }

int check_cfg(const amrh_config* c) {
  if (!c) return fail(AMRH_EINVAL, "null config");
  if (c->physics == AMRH_PHYS_EULER) {
    if (c->nf != 6) return fail(AMRH_EINVAL, "EULER physics has 6 fields, got %d", c->nf);
  } else if (c->physics == AMRH_PHYS_ROTATING) {
    if (c->nf != rot::NF) return fail(AMRH_EINVAL, "ROTATING physics has %d fields, got %d", rot::NF, c->nf);
    if (c->n != 8) return fail(AMRH_EINVAL, "ROTATING physics runs on n = 8 blocks, got n = %d", c->n);
    if (c->eos != AMRH_EOS_IDEAL && c->eos != AMRH_EOS_HELMHOLTZ) return fail(AMRH_EINVAL, "eos %d unknown", c->eos);
    for (int i = 0; i < 3; i++)
      if (!(c->abar[i] > 0.0)) return fail(AMRH_EINVAL, "abar[%d] must be positive", i);
    if (!(c->rho_floor > 0.0)) return fail(AMRH_EINVAL, "rho_floor must be positive");
    if (!(c->newton_tol >= 0.0 && c->newton_tol < 1.0)) return fail(AMRH_EINVAL, "newton_tol must be in [0, 1)");
  } else {
    return fail(AMRH_EINVAL, "physics %d not supported", c->physics);
  }
  if (c->n != 8 && c->n != 16) return fail(AMRH_EINVAL, "n must be 8 or 16, got %d", c->n);
  if (c->recon == AMRH_RECON_MINMOD) {
    if (c->g != 2) return fail(AMRH_EINVAL, "minmod reconstruction needs g = 2, got %d", c->g);
  } else if (c->recon == AMRH_RECON_PPM) {
    if (c->g != 3) return fail(AMRH_EINVAL, "PPM reconstruction needs g = 3, got %d", c->g);
    if (c->n != 8) return fail(AMRH_EINVAL, "PPM reconstruction runs on n = 8 blocks, got n = %d", c->n);
  } else {
    return fail(AMRH_EINVAL, "recon %d unknown", c->recon);
  }
  if (c->precision != AMRH_FP64 && c->precision != AMRH_MIXED)
    return fail(AMRH_EINVAL, "precision %d unknown", c->precision);
  if (!(c->gamma > 1.0)) return fail(AMRH_EINVAL, "gamma must exceed 1");
  if (c->boundary != AMRH_BOUNDARY_EXTERNAL && c->boundary != AMRH_BOUNDARY_OUTFLOW)
    return fail(AMRH_EINVAL, "boundary %d unknown", c->boundary);
  return AMRH_OK;
}

}  // namespace

extern "C" {

const char* amrh_last_error(void) { return g_err.c_str(); }

int amrh_create(const amrh_config* cfg, int nblocks, amrh_ctx** out) {
  if (!out) return fail(AMRH_EINVAL, "null out");
  *out = nullptr;
  if (int rc = check_cfg(cfg)) return rc;
  if (nblocks <= 0) return fail(AMRH_EINVAL, "nblocks must be positive");
  amrh_ctx* c = new amrh_ctx;
  c->cfg = *cfg;
  c->nb = nblocks;
  c->L = Layout{cfg->n, cfg->g, cfg->nf};
  const size_t rec_bytes = sizeof(double) * (size_t)record_elems(c);
  hipError_t e = hipSuccess;
  for (int q = 0; q < 2 && e == hipSuccess; q++) e = hipMalloc(&c->buf[q], rec_bytes);
  if (e == hipSuccess) e = hipMalloc(&c->h_dev, sizeof(double) * nblocks);
  if (e == hipSuccess) e = hipMalloc(&c->eig_dev, sizeof(double) * nblocks);
  if (e == hipSuccess) e = hipMalloc(&c->gmax_dev, sizeof(unsigned long long) * 2);
  if (e == hipSuccess) e = hipMalloc(&c->err_dev, sizeof(int));
  if (e == hipSuccess && cfg->physics == AMRH_PHYS_ROTATING) {
    const size_t P = (size_t)cfg->n + 2;
    e = hipMalloc(&c->x0_dev, sizeof(double) * 3 * nblocks);
    if (e == hipSuccess) e = hipMalloc(&c->phi_dev, sizeof(double) * P * P * P * nblocks);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemset(c->err_dev, 0, sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->gmax_dev, 0, sizeof(unsigned long long) * 2);
  if (e == hipSuccess) e = hipMemset(c->buf[0], 0, rec_bytes);
  if (e == hipSuccess) e = hipMemset(c->buf[1], 0, rec_bytes);
  if (e != hipSuccess) {
    amrh_destroy(c);
    return fail(AMRH_EHIP, "allocation failed: %s", hipGetErrorString(e));
  }
  c->stream = c->own_stream;
  *out = c;
  return AMRH_OK;
}

void amrh_destroy(amrh_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  for (int q = 0; q < 3; q++)
    if (c->buf[q]) (void)hipFree(c->buf[q]);
  if (c->h_dev) (void)hipFree(c->h_dev);
  if (c->eig_dev) (void)hipFree(c->eig_dev);
  if (c->gmax_dev) (void)hipFree(c->gmax_dev);
  if (c->err_dev) (void)hipFree(c->err_dev);
  if (c->pad_stage) (void)hipFree(c->pad_stage);
  if (c->x0_dev) (void)hipFree(c->x0_dev);
  if (c->diag_part) (void)hipFree(c->diag_part);
  if (c->phi_dev) (void)hipFree(c->phi_dev);
  if (c->force_own) (void)hipFree(c->force_own);
  if (c->dbuf) (void)hipFree(c->dbuf);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  if (c->tree) {
    c->tree->free_dev();
    delete c->tree;
  }
  delete c;
}

int amrh_set_stream(amrh_ctx* c, void* stream) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  c->stream = stream ? (hipStream_t)stream : c->own_stream;
  return AMRH_OK;
}

int amrh_sync(amrh_ctx* c) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRH_OK;
}

int amrh_set_block_h(amrh_ctx* c, const double* h) {
  if (!c || !h) return fail(AMRH_EINVAL, "null argument");
  for (int b = 0; b < c->nb; b++)
    if (!(h[b] > 0.0)) return fail(AMRH_EINVAL, "h[%d] = %g is not positive", b, h[b]);
  HIPCHK(hipMemcpyAsync(c->h_dev, h, sizeof(double) * c->nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_h = true;
  return AMRH_OK;
}

static int ensure_stage(amrh_ctx* c) {
  if (!c->pad_stage) HIPCHK(hipMalloc(&c->pad_stage, sizeof(double) * (size_t)padded_elems(c)));
  return AMRH_OK;
}

int amrh_upload_device(amrh_ctx* c, const void* pad) {
  if (!c || !pad) return fail(AMRH_EINVAL, "null argument");
  HIPCHK(launch_pack(c->cfg.n, c->cfg.g, (const double*)pad, c->buf[c->cur], record_elems(c), true, c->stream));
  c->have_u = true;
  c->u0buf = -1;
  return AMRH_OK;
}

int amrh_upload(amrh_ctx* c, const double* U) {
  if (!c || !U) return fail(AMRH_EINVAL, "null argument");
  if (int rc = ensure_stage(c)) return rc;
  HIPCHK(hipMemcpyAsync(c->pad_stage, U, sizeof(double) * (size_t)padded_elems(c), hipMemcpyHostToDevice,
                        c->stream));
  int rc = amrh_upload_device(c, c->pad_stage);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRH_OK;
}

int amrh_download_device(amrh_ctx* c, void* pad) {
  if (!c || !pad) return fail(AMRH_EINVAL, "null argument");
  if (!c->have_u) return fail(AMRH_ESTATE, "no state uploaded");
  // a self-ghosting batch does not maintain its ghost slabs: form them for the copy (the rotating
  // state's need the block origins)
  if (self_ghosting(c) && (c->cfg.physics != AMRH_PHYS_ROTATING || c->have_x0))
    HIPCHK(ctx_outflow_ghosts(c, c->buf[c->cur]));
  HIPCHK(hipMemsetAsync(pad, 0, sizeof(double) * (size_t)padded_elems(c), c->stream));
  HIPCHK(launch_pack(c->cfg.n, c->cfg.g, (const double*)pad, c->buf[c->cur], record_elems(c), false, c->stream));
  return AMRH_OK;
}

int amrh_download(amrh_ctx* c, double* U) {
  if (!c || !U) return fail(AMRH_EINVAL, "null argument");
  if (int rc = ensure_stage(c)) return rc;
  if (int rc = amrh_download_device(c, c->pad_stage)) return rc;
  HIPCHK(hipMemcpyAsync(U, c->pad_stage, sizeof(double) * (size_t)padded_elems(c), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRH_OK;
}

int amrh_store(amrh_ctx* c) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->have_u) return fail(AMRH_ESTATE, "no state uploaded");
  c->u0buf = c->cur;
  if (c->tree) HIPCHK(launch_ledger_store(c->tree->fo_dev, c->cfg.nf, c->stream));  // FO0 = FO
  return AMRH_OK;
}

int amrh_restore(amrh_ctx* c) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->have_u) return fail(AMRH_ESTATE, "no state uploaded");
  if (c->u0buf < 0) return fail(AMRH_ESTATE, "restore needs amrh_store() first");
  // U = U0: the RK base buffer becomes the current stage again (it stays the RK base, so the next
  // substep writes the other buffer and U0 survives for another restore)
  c->cur = c->u0buf;
  if (c->tree) HIPCHK(launch_ledger_restore(c->tree->fo_dev, c->cfg.nf, c->stream));  // FO = FO0
  return AMRH_OK;
}

int amrh_fill_outflow_ghosts(amrh_ctx* c) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->have_u) return fail(AMRH_ESTATE, "no state uploaded");
  if (int rc = amrh::ctx_ready(c)) return rc;
  HIPCHK(ctx_outflow_ghosts(c, c->buf[c->cur]));
  return AMRH_OK;
}

int amrh_set_block_origin(amrh_ctx* c, const double* x0) {
  if (!c || !x0) return fail(AMRH_EINVAL, "null argument");
  if (c->tree) return fail(AMRH_EINVAL, "an octree's positions follow from amrh_tree_set");
  if (!c->x0_dev) {
    drop_graphs(c);  // a new device array
    HIPCHK(hipMalloc(&c->x0_dev, sizeof(double) * 3 * c->nb));
  }
  HIPCHK(hipMemcpyAsync(c->x0_dev, x0, sizeof(double) * 3 * c->nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_x0 = true;
  return AMRH_OK;
}

static size_t phi_bytes(const amrh_ctx* c) {
  const size_t P = (size_t)c->cfg.n + 2;
  return sizeof(double) * P * P * P * c->nb;
}

int amrh_set_potential(amrh_ctx* c, const double* phi) {
  if (!c || !phi) return fail(AMRH_EINVAL, "null argument");
  if (c->cfg.physics != AMRH_PHYS_ROTATING) return fail(AMRH_EINVAL, "the potential belongs to ROTATING physics");
  HIPCHK(hipMemcpyAsync(c->phi_dev, phi, phi_bytes(c), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_phi = true;
  return AMRH_OK;
}

int amrh_set_potential_device(amrh_ctx* c, const void* phi) {
  if (!c || !phi) return fail(AMRH_EINVAL, "null argument");
  if (c->cfg.physics != AMRH_PHYS_ROTATING) return fail(AMRH_EINVAL, "the potential belongs to ROTATING physics");
  HIPCHK(hipMemcpyAsync(c->phi_dev, phi, phi_bytes(c), hipMemcpyDeviceToDevice, c->stream));
  c->have_phi = true;
  return AMRH_OK;
}

int amrh_set_forces(amrh_ctx* c, const double* f) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (c->cfg.physics != AMRH_PHYS_ROTATING) return fail(AMRH_EINVAL, "the forces belong to ROTATING physics");
  drop_graphs(c);  // captured substeps hold the force pointer
  if (!f) {
    c->force_dev = nullptr;
    return AMRH_OK;
  }
  if (!c->force_own) HIPCHK(hipMalloc(&c->force_own, 3 * phi_bytes(c)));
  HIPCHK(hipMemcpyAsync(c->force_own, f, 3 * phi_bytes(c), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->force_dev = c->force_own;
  return AMRH_OK;
}

int amrh_set_forces_device(amrh_ctx* c, const void* f) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (c->cfg.physics != AMRH_PHYS_ROTATING) return fail(AMRH_EINVAL, "the forces belong to ROTATING physics");
  if (c->force_dev != (const double*)f) drop_graphs(c);  // captured substeps hold the force pointer
  c->force_dev = (const double*)f;
  return AMRH_OK;
}

int amrh_pot_to_grid(amrh_ctx* c) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (c->cfg.physics != AMRH_PHYS_ROTATING) return fail(AMRH_EINVAL, "pot belongs to ROTATING physics");
  if (!c->have_u) return fail(AMRH_ESTATE, "no state uploaded");
  if (int rc = ctx_ready(c)) return rc;
  HIPCHK(launch_rot_pot_to_grid(c->cfg.g, c->buf[c->cur], c->phi_dev, c->x0_dev, c->h_dev, c->nb,
                                ctx_rot_args(c).c, c->stream));
  return AMRH_OK;
}

int amrh_attach_eos(amrh_ctx* c, amrh_eos* eos) {
  if (!c || !eos) return fail(AMRH_EINVAL, "null argument");
  if (c->cfg.physics != AMRH_PHYS_ROTATING || c->cfg.eos != AMRH_EOS_HELMHOLTZ)
    return fail(AMRH_EINVAL, "an EOS table belongs to ROTATING physics with eos = AMRH_EOS_HELMHOLTZ");
  if (c->eos != eos) drop_graphs(c);  // captured substeps hold the table's pointers
  c->eos = eos;
  return AMRH_OK;
}

}  // extern "C"

namespace {
rot::RotParams rot_params(const amrh_config& g) {
  rot::RotParams r{};
  r.Omega = g.omega;
  r.Omega0 = g.omega0;
  r.fR = g.f_r;
  r.ftheta = g.f_theta;
  r.rho_floor = g.rho_floor;
  r.gamma = g.gamma;
  r.theta = g.theta;
  for (int i = 0; i < 3; i++) r.abar[i] = g.abar[i];
  return r;
}

}  // namespace

RotArgs amrh::ctx_rot_args(const amrh_ctx* c) {
  RotArgs r{};
  r.phi = c->phi_dev;
  r.force = c->force_dev;
  r.nforce = c->nb;
  r.x0 = c->x0_dev;
  r.h = c->h_dev;
  r.nblocks = c->nb;
  r.err = c->err_dev;
  r.c = rot_params(c->cfg);
  if (c->eos) r.H = *eos_device(c->eos);
  r.H.err = c->err_dev;
  r.H.toler = c->cfg.newton_tol > 0.0 ? c->cfg.newton_tol : 1.0e-9;
  return r;
}

namespace {
RotArgs rot_args(const amrh_ctx* c) { return amrh::ctx_rot_args(c); }
}  // namespace

int amrh::ctx_ready(const amrh_ctx* c) {
  if (c->cfg.physics != AMRH_PHYS_ROTATING) return AMRH_OK;
  if (!c->have_x0) return fail(AMRH_ESTATE, "ROTATING physics needs amrh_set_block_origin");
  if (!c->have_phi) return fail(AMRH_ESTATE, "ROTATING physics needs amrh_set_potential");
  if (c->cfg.eos == AMRH_EOS_HELMHOLTZ && !c->eos) return fail(AMRH_ESTATE, "eos = HELMHOLTZ needs amrh_attach_eos");
  return AMRH_OK;
}

bool amrh::self_ghosting(const amrh_ctx* c) { return !c->tree && c->cfg.boundary == AMRH_BOUNDARY_OUTFLOW; }

SubstepArgs amrh::flat_args(const amrh_ctx* c) {
  SubstepArgs a{};
  a.Uin = c->buf[c->cur];
  a.h = c->h_dev;
  a.nblocks = c->nb;
  a.block0 = 0;
  a.P = params(c->cfg);
  a.err = c->err_dev;
  a.ledger_block = -1;
  a.outflow = (!c->tree && c->cfg.boundary == AMRH_BOUNDARY_OUTFLOW) ? 1 : 0;
  a.z_reflect = c->cfg.z_reflect;
  return a;
}

hipError_t amrh::ctx_outflow_ghosts(const amrh_ctx* c, double* U) {
  if (c->cfg.physics == AMRH_PHYS_ROTATING)
    return launch_rot_ghosts(c->cfg.g, U, c->x0_dev, c->h_dev, c->nb, c->cfg.z_reflect, rot_params(c->cfg), c->stream);
  return launch_outflow_ghosts(c->cfg.n, c->cfg.g, U, c->nb, c->cfg.z_reflect, c->stream);
}

hipError_t amrh::ctx_substep_launch(const amrh_ctx* c, const SubstepArgs& a, int grid, int variant) {
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  if (c->cfg.physics == AMRH_PHYS_ROTATING) {
    RotArgs r = rot_args(c);
    r.Uin = a.Uin;
    r.Uout = a.Uout;
    r.U0 = a.U0;
    r.dt_dev = a.dt_dev;
    r.dt = a.dt;
    r.beta = a.beta;
    r.outflow = a.outflow;
    r.z_reflect = a.z_reflect;
    return launch_rot_substep(c->cfg.g, c->cfg.recon, c->cfg.eos, mixed, r, grid, c->stream);
  }
  if (c->cfg.recon == AMRH_RECON_MINMOD)
    return launch_substep(c->cfg.n, a, mixed, grid, c->stream, variant >= 0 ? variant : c->variant);
  return launch_block_substep(c->cfg.n, c->cfg.g, c->cfg.recon, a, mixed, grid, c->stream);
}

hipError_t amrh::ctx_maxeig_launch(const amrh_ctx* c, const MaxEigArgs& m, int grid) {
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  if (c->cfg.physics == AMRH_PHYS_ROTATING) {
    RotArgs r = rot_args(c);
    r.Uin = m.U;
    r.block_eig = m.block_eig;
    r.gmax = m.gmax;
    r.outflow = m.outflow;
    r.z_reflect = m.z_reflect;
    return launch_rot_maxeig(c->cfg.g, c->cfg.recon, c->cfg.eos, mixed, r, grid, c->stream);
  }
  if (c->cfg.recon == AMRH_RECON_MINMOD) return launch_maxeig(c->cfg.n, m, mixed, grid, c->stream);
  return launch_block_maxeig(c->cfg.n, c->cfg.g, c->cfg.recon, m, mixed, grid, c->stream);
}

// The fused substep over every block; with `tree` the reflux overrides and the root's flow-off
// ledger of the octree are applied (the callers have exported the fine faces first).
int amrh::substep_core(amrh_ctx* c, double dt, double beta, const double* dt_dev, bool tree) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->have_u || !c->have_h) return fail(AMRH_ESTATE, "state or cell widths not set");
  if (int rc = ctx_ready(c)) return rc;
  const bool need_u0 = beta != 1.0;
  if (need_u0 && c->u0buf < 0) return fail(AMRH_ESTATE, "beta != 1 needs amrh_store() first");
  SubstepArgs a{};
  a.Uin = c->buf[c->cur];
  int out = c->cur;
  if (c->u0buf == c->cur) out = free_buf(c);  // never overwrite the RK base
  if (out < 0) return fail(AMRH_ESTATE, "no free state buffer");
  a.Uout = c->buf[out];
  a.U0 = c->u0buf >= 0 ? c->buf[c->u0buf] : c->buf[c->cur];
  a.h = c->h_dev;
  a.dt_dev = dt_dev;
  a.dt = dt;
  a.beta = beta;
  a.nblocks = c->nb;
  a.block0 = 0;
  a.P = params(c->cfg);
  a.err = c->err_dev;
  a.slot = nullptr;
  a.ovr = nullptr;
  a.ledger_block = -1;
  a.ledger = nullptr;
  a.outflow = (!tree && c->cfg.boundary == AMRH_BOUNDARY_OUTFLOW) ? 1 : 0;
  a.z_reflect = c->cfg.z_reflect;
  if (tree && c->tree) {
    a.slot = c->tree->slot_dev;
    a.ovr = c->tree->ovr_dev;
    a.ledger_block = 0;
    a.ledger = c->tree->ledger_dev;
  }
  HIPCHK(ctx_substep_launch(c, a, c->grid_substep));
  c->cur = out;
  return AMRH_OK;
}

extern "C" {

int amrh_substep(amrh_ctx* c, double dt, double beta) { return substep_core(c, dt, beta, nullptr, false); }

int amrh_substep_dev(amrh_ctx* c, const double* dt_dev, double beta) {
  if (!dt_dev) return fail(AMRH_EINVAL, "null dt_dev");
  return substep_core(c, 0.0, beta, dt_dev, false);
}

static int maxeig_impl(amrh_ctx* c, bool per_block, bool global) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->have_u || !c->have_h) return fail(AMRH_ESTATE, "state or cell widths not set");
  if (int rc = ctx_ready(c)) return rc;
  MaxEigArgs m{};
  m.U = c->buf[c->cur];
  m.h = c->h_dev;
  m.block_eig = per_block ? c->eig_dev : nullptr;
  m.gmax = global ? c->gmax_dev : nullptr;
  m.nblocks = c->nb;
  m.P = params(c->cfg);
  m.outflow = self_ghosting(c) ? 1 : 0;
  m.z_reflect = c->cfg.z_reflect;
  HIPCHK(ctx_maxeig_launch(c, m, c->grid_maxeig));
  return AMRH_OK;
}

int amrh_block_max_eig(amrh_ctx* c, double* out) {
  if (!out) return fail(AMRH_EINVAL, "null out");
  if (int rc = maxeig_impl(c, true, false)) return rc;
  HIPCHK(hipMemcpyAsync(out, c->eig_dev, sizeof(double) * c->nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRH_OK;
}

int amrh_max_speed(amrh_ctx* c, double* out) {
  if (!out) return fail(AMRH_EINVAL, "null out");
  if (int rc = maxeig_impl(c, false, true)) return rc;
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, c->gmax_dev, sizeof bits, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(launch_reset_u64(c->gmax_dev, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::memcpy(out, &bits, sizeof(double));
  return AMRH_OK;
}

int amrh_dt_device(amrh_ctx* c, double scale, double growth, double* dt_dev) {
  if (!dt_dev) return fail(AMRH_EINVAL, "null dt_dev");
  if (int rc = maxeig_impl(c, false, true)) return rc;
  HIPCHK(launch_dt_from_gmax(c->gmax_dev, c->cfg.cfl, scale, growth, dt_dev, c->stream));
  return AMRH_OK;
}

int amrh_error(amrh_ctx* c, int* flags) {
  if (!c || !flags) return fail(AMRH_EINVAL, "null argument");
  int v = 0;
  HIPCHK(hipMemcpyAsync(&v, c->err_dev, sizeof v, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemsetAsync(c->err_dev, 0, sizeof(int), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *flags = v;
  return AMRH_OK;
}

double amrh_substep_bytes(const amrh_ctx* c, int beta_is_one) {
  if (!c) return 0.0;
  const double n = c->cfg.n, g = c->cfg.g, nf = c->cfg.nf;
  const double n3 = n * n * n;
  const double read_u = n3 + 6.0 * g * n * n;
  double per_block = 8.0 * nf * (read_u + n3 + (beta_is_one ? 0.0 : n3));
  // the rotating state also reads the potential over the interior and its six face halos
  if (c->cfg.physics == AMRH_PHYS_ROTATING) per_block += 8.0 * (n3 + 6.0 * n * n);
  return per_block * c->nb;
}

double amrh_ghost_write_bytes(const amrh_ctx* c) {
  // self-ghosting batches write no ghosts; the others' ghost fill writes every face slab
  if (!c || c->cfg.boundary != AMRH_BOUNDARY_OUTFLOW || self_ghosting(c)) return 0.0;
  const double n = c->cfg.n, g = c->cfg.g, nf = c->cfg.nf;
  return 8.0 * nf * 6.0 * g * n * n * c->nb;
}

}  // extern "C"
