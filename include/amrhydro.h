/* amrhydro -- MI355X (gfx950) hydro right-hand side of an octree-AMR Eulerian code.
 *
 * C ABI between host code (the tree / driver) and the HIP kernels.  Plain pointers and sizes; no
 * exceptions; every entry point returns 0 on success or an AMRH_E* code (amrh_last_error() gives
 * the message).  One context = one device stream; calls on a context are asynchronous with
 * respect to the host except where stated, and a context must not be used from two host threads
 * at once.  The device owns the state buffers; the host owns the tree.
 *
 * Each entry point replaces a method of the reference's Grid / GridNode interface
 * (grid/grid.h:81-96, grid/grid_node.h:47-80), applied to every block of the batch at once.
 *
 * Host-side array layout ("padded SoA"): per block, per field, a GNX^3 cube, GNX = n + 2g, x
 * fastest: element (b, f, i, j, k) at ((b*nf + f)*GNX + k)*GNX*GNX + j*GNX + i.  It is the
 * reference's Array3d<State,GNX,GNX,GNX> (array3d.h:11) with the fields split out.  Only the
 * interior and the six face ghost slabs are meaningful (the reference never reads edge or corner
 * ghosts, grid/grid_node.cpp:59-112); downloads write zeros there.
 */
#ifndef AMRHYDRO_H
#define AMRHYDRO_H

#ifdef __cplusplus
extern "C" {
#endif

#define AMRH_OK 0
#define AMRH_EINVAL 1   /* bad argument */
#define AMRH_EHIP 2     /* HIP runtime error */
#define AMRH_ESTATE 3   /* call out of order (e.g. no tree, no upload) */
#define AMRH_EDEVICE 4  /* device-side error flag raised (see amrh_error) */

#define AMRH_PHYS_EULER 0    /* ideal gas, fields rho sx sy sz E tau (euler/state.cpp:9-17) */
#define AMRH_PHYS_ROTATING 1 /* rotating-frame binary-star state (binary/state_ztwd.cpp:15-28): 13 fields
                                rho, rho_acc, rho_don, s_R, l_z, s_z, pot, E, tau, T, 3 species;
                                n = 8; a given gravitational potential (amrh_set_potential) and
                                block positions (amrh_set_block_origin) */

/* equation of state of AMRH_PHYS_ROTATING (amrh_config.eos) */
#define AMRH_EOS_IDEAL 0      /* p = rho T (1 + zbar)/abar, e = p/((gamma-1) rho) */
#define AMRH_EOS_HELMHOLTZ 1  /* helmholtz/helmeos.cpp on a table attached with amrh_attach_eos */

#define AMRH_RECON_MINMOD 0 /* minmod-theta slopes (reconstruct.cpp:60-71); g = 2 */
#define AMRH_RECON_PPM 1    /* piecewise parabolic, theta = 2 slopes + monotonisation
                               (reconstruct.cpp:22-58 as dispatched at :92-102); g = 3, n = 8 */

#define AMRH_FP64 0  /* FP64 end to end; matches the FP64 CPU path to roundoff */
#define AMRH_MIXED 1 /* FP32 reconstruction / fluxes / wavespeeds, FP64 accumulation + RK + floor */

/* ghost layers of a flat batch (amrh_config.boundary) */
#define AMRH_BOUNDARY_EXTERNAL 0 /* the caller fills the ghost slabs before every substep */
#define AMRH_BOUNDARY_OUTFLOW 1  /* every block is an isolated domain with outflow faces
                                    (z-mirror below with z_reflect): the kernels form each ghost
                                    from the boundary cell it copies where they read it, and the
                                    ghost slabs in memory are not maintained (amrh_download and
                                    amrh_fill_outflow_ghosts form them) */

/* device error flags (amrh_error) */
#define AMRH_DERR_NEG_DENSITY 1 /* a face or cell state with rho <= 0 reached the flux */

typedef struct {
  int n;         /* interior cells per block edge; 8 or 16 (euler/euler_defs.h:27) */
  int g;         /* ghost width: 2 for minmod (euler/euler_defs.h:25), 3 for PPM (binary/binary_defs.h:30) */
  int nf;        /* fields; 6 for AMRH_PHYS_EULER (euler/state.h:16) */
  int physics;   /* AMRH_PHYS_* */
  int recon;     /* AMRH_RECON_* */
  int precision; /* AMRH_FP64 or AMRH_MIXED */
  int z_reflect; /* lower z boundary reflects (euler/euler_defs.h:21) */
  double gamma;      /* 5/3 (euler/state.cpp:9) */
  double rho_floor;  /* 1e-20 (euler/state.cpp:10) */
  double ei_floor;   /* 1e-20 (euler/state.cpp:11) */
  double theta;      /* 1.3 (euler/euler_defs.h:20) */
  double cfl;        /* 0.4 (euler/euler_defs.h:46) */
  double grid_dim;   /* half edge of the root domain (grid/grid.cpp:70,512-526) */
  int boundary;      /* AMRH_BOUNDARY_* (flat batches; the octree path fills its own ghosts) */
  /* AMRH_PHYS_ROTATING only */
  int eos;           /* AMRH_EOS_* */
  double omega;      /* frame angular velocity Binary::Omega (binary/binary.h:43) */
  double omega0;     /* Binary::Omega0, the potential's reference frame (binary/binary.cpp:24) */
  double f_r;        /* State::fR, centre-of-mass correction force (binary/binary_driver.cpp:194) */
  double f_theta;    /* State::ftheta (binary/binary_driver.cpp:195) */
  double abar[3];    /* species mass numbers (binary/state_ztwd.cpp:28) */
  double newton_tol; /* relative tolerance |e(T) - e|/e of the floor's Newton solve for T; 0 = the
                        reference's 1e-9 (helmholtz/helmeos.cpp:449) */
} amrh_config;

typedef struct amrh_ctx amrh_ctx;

/* Create a context for `nblocks` blocks on the current HIP device.  Allocates two state buffers
 * (the current stage and the RK base U0; amrh_tree_set adds a third for the rotating state),
 * per-block metadata and reduction scratch.
 * Replaces the storage of Grid (grid/grid.h:26-32) for every node of the tree. */
int amrh_create(const amrh_config* cfg, int nblocks, amrh_ctx** out);
void amrh_destroy(amrh_ctx* ctx);
const char* amrh_last_error(void);

/* Run all further work of this context on `stream` (a hipStream_t); NULL = the context's own. */
int amrh_set_stream(amrh_ctx* ctx, void* stream);
/* Block until all work of the context has finished; returns a pending HIP error if any. */
int amrh_sync(amrh_ctx* ctx);

/* Per-block cell width h (Grid::set_dx, grid/grid.cpp:536-539).  Host array of nblocks. */
int amrh_set_block_h(amrh_ctx* ctx, const double* h);

/* Copy the padded SoA state of every block host->device (the whole current stage).
 * Replaces direct writes into Grid::U (grid/grid.cpp:153-159). */
int amrh_upload(amrh_ctx* ctx, const double* U_padded);
/* Same, from a device pointer holding the padded SoA layout. */
int amrh_upload_device(amrh_ctx* ctx, const void* U_padded_dev);
/* Copy the current stage device->host, padded SoA (interior + face ghost slabs). */
int amrh_download(amrh_ctx* ctx, double* U_padded);
int amrh_download_device(amrh_ctx* ctx, void* U_padded_dev);

/* U0 = U for every block (Grid::store, grid/grid.cpp:173-185).  No copy is made: the current
 * buffer becomes the RK base and the next substep writes the other buffer. */
int amrh_store(amrh_ctx* ctx);
/* U = U0 for every block and, on an octree, the root's flow-off ledger FO = FO0 (Grid::restore,
 * grid/grid.cpp:187-199, recursed by GridNode::restore, grid/grid_node.cpp:486-493).  No copy is
 * made: the RK base buffer becomes the current stage again and stays the RK base, so a restored
 * step can be redone and restored again -- BinaryDriver::step's dt rejection (REDO,
 * binary/binary_driver.cpp:414-423).  AMRH_ESTATE before the first amrh_store. */
int amrh_restore(amrh_ctx* ctx);

/* Ghost fill of isolated blocks whose six neighbours are physical outflow boundaries
 * (GridPhysBound, grid/grid_phys_bound.cpp:27-54 with State::enforce_outflow,
 * euler/state.cpp:45-84; z-mirror on the lower z face when z_reflect). */
int amrh_fill_outflow_ghosts(amrh_ctx* ctx);

/* One RK substep over every block, fused: clear_difs, compute_{x,y,z}_flux, sum_{x,y,z}_difs,
 * add_difs with floor (grid/grid.cpp:201-344, 406-420):
 *   U <- (U + dt*D(U))*beta + U0*(1-beta),  D = -sum_d (F_{d,i+1}-F_{d,i})/h.
 * With AMRH_BOUNDARY_EXTERNAL the ghost slabs of the current stage must be filled. */
int amrh_substep(amrh_ctx* ctx, double dt, double beta);
/* Same, with dt read from device memory when the kernel runs (e.g. written by amrh_dt_device, or
 * replayed from a captured graph); dt_dev points to one double in device memory. */
int amrh_substep_dev(amrh_ctx* ctx, const double* dt_dev, double beta);

/* Grid::max_dt scan (grid/grid.cpp:422-473): per block, the maximum over all interior faces of
 * the three sweeps of |v_d| + c_s of both reconstructed face states.  Writes nblocks doubles to
 * the host array `max_eig` (synchronous). */
int amrh_block_max_eig(amrh_ctx* ctx, double* max_eig);
/* Device reduction of max_b(max_eig_b / h_b) (a wavefront + atomic max over the bit pattern of
 * non-negative doubles, order independent).  Synchronous; writes one double. */
int amrh_max_speed(amrh_ctx* ctx, double* max_eig_over_h);
/* dt = scale * cfl / max_b(max_eig_b/h_b) computed on the device into dt_dev[0] (no host round
 * trip).  With growth > 0 the result is also capped at growth * dt_dev[0] (the previous dt).  The
 * reference's control is scale = MAXINITDT = 1e-2, growth = 0 at t = 0, then scale = 1,
 * growth = MAXDTINC = 1.25 (euler/hydro_driver.cpp:107-113, euler/euler_defs.h:45,57). */
int amrh_dt_device(amrh_ctx* ctx, double scale, double growth, double* dt_dev);

/* Device error flags raised since the last call (AMRH_DERR_*); clears them.  Synchronous. */
int amrh_error(amrh_ctx* ctx, int* flags);

/* Bytes of HBM one substep over the batch moves by the algorithmic count: the interior and its six
 * face ghost slabs read, the interior written, U0's interior read when beta != 1; for the rotating
 * state also the potential over the interior and its six face halos.  amrh_ghost_write_bytes: the
 * ghost slabs a separate ghost refresh would write per substep; 0 for AMRH_BOUNDARY_OUTFLOW
 * batches, whose kernels form their ghosts where they read them and write none. */
double amrh_substep_bytes(const amrh_ctx* ctx, int beta_is_one);
double amrh_ghost_write_bytes(const amrh_ctx* ctx);

/* ---- launch shape and whole steps (no counterpart in the reference: its loops are OpenMP) -----
 * The sweep kernels walk the batch with a grid stride.  amrh_autotune times one workgroup per
 * block against persistent grids of k workgroups per CU on the context's own state and keeps the
 * fastest for the substep (timed as a beta != 1 substep, the form of all but a step's first) and
 * for the CFL scan; each time is the median of five launches (writes the spare state buffer: call
 * it before amrh_store).  amrh_set_launch sets them by hand (0 = one workgroup per block). */
int amrh_autotune(amrh_ctx* ctx, int* grid_substep, int* grid_maxeig);
int amrh_set_launch(amrh_ctx* ctx, int grid_substep, int grid_maxeig);
/* The flat minmod substep of the ideal gas also comes in register-budget variants -- 0 the
 * default, 1..3 the kernel held to 2, 3, 4 waves per SIMD (256, 168, 128 registers) -- and
 * amrh_autotune times every variant with every grid.  amrh_launch_info returns the choice and the
 * autotuner's substep times of the default launch and of the pick (ms; 0 before amrh_autotune);
 * amrh_set_variant sets the variant by hand.  Results do not depend on the launch shape. */
int amrh_launch_info(const amrh_ctx* ctx, int* grid_substep, int* grid_maxeig, int* variant, double* ms_default,
                     double* ms_tuned);
int amrh_set_variant(amrh_ctx* ctx, int variant);
/* HydroDriver::step repeated nsteps times (euler/hydro_driver.cpp:171-194, regrid excluded) with
 * the reference's dt control on the device: dt_dev[0] holds the previous dt and becomes
 * min(MAXDTINC*dt, CFL dt).  For the octree (after amrh_tree_set) each step fills ghosts, scans
 * the CFL speed, stores U0 and runs rk tree substeps; for a flat batch it needs
 * boundary = AMRH_BOUNDARY_OUTFLOW.  With use_graph each step is one launch of a captured
 * hipGraph (one graph per state buffer a step starts from, captured on first use; a new dt_dev,
 * rk or launch shape re-captures).  Asynchronous. */
int amrh_run_steps(amrh_ctx* ctx, int rk, int nsteps, double* dt_dev, int use_graph);
/* In amrh_run_steps and amrh_step_timed, an RK2/RK3 step of a flat AMRH_BOUNDARY_OUTFLOW batch runs
 * with its dt deferred: the first substep forms the fluxes of U^n, reduces their KT face speeds --
 * the numbers Grid::max_dt's scan forms (grid/grid.cpp:422-483) -- to the CFL maximum and stores
 * the divergence D (sources included); the dt kernel applies the control above; the minmod ideal
 * gas's second substep forms U1 = floor(U^n + D dt) for every cell it reads; on the PPM block
 * kernel and for the rotating state (floor with the Newton T, pot reset) an elementwise pass forms
 * U1 before plain substeps.  The rotating state's octree steps run the same way: the deferred
 * substep level by level, finest first, with its reflux overrides, exports and ledger faces (none
 * depends on dt), then dt, the store, the first stage of every node, the ledger and the
 * restriction; its first ghost fill is the substep's (same-level faces read from the neighbours).
 * No separate CFL scan, no store: the same bits as amrh_dt_device + amrh_store +
 * amrh_substep_dev.  With the Helmholtz EOS the rotating state's later substeps run as a substep
 * that stores D and the same elementwise pass (its Newton solve for T outside the flux kernel),
 * again the same bits.  amrh_autotune also picks the launch shapes of those kernels, and measures
 * whether the later substeps split; amrh_deferred_launch_info returns them (kernel 0: the deferred
 * first substep, 1: the second substep or the first-stage pass, 2: the rotating state's later
 * substeps -- variant 1 if split, ms_default the substep's time, ms_tuned the split form's), as
 * amrh_launch_info does. */
int amrh_deferred_launch_info(const amrh_ctx* ctx, int kernel, int* grid, int* variant, double* ms_default,
                              double* ms_tuned);
/* Whether such steps run deferred (the default) or as scan + store + substeps: amrh_autotune keeps
 * the faster of the two by the kernel times it measured; amrh_set_schedule sets it by hand (the
 * results are the same bits); amrh_get_schedule reports AMRH_SCHEDULE_PLAIN for a batch that
 * cannot run deferred steps. */
#define AMRH_SCHEDULE_PLAIN 0
#define AMRH_SCHEDULE_DEFERRED 1
int amrh_set_schedule(amrh_ctx* ctx, int schedule);
int amrh_get_schedule(const amrh_ctx* ctx, int* schedule);
/* Stages of one timed step (amrh_step_timed) */
#define AMRH_LAUNCH_SUBSTEP 0 /* a substep (with its ghost fill where one follows it) */
#define AMRH_LAUNCH_SCAN 1    /* the CFL scan and the dt kernel */
#define AMRH_LAUNCH_FIRST 2   /* the deferred first substep (fluxes, CFL maximum, D) */
#define AMRH_LAUNCH_DT 3      /* the dt kernel of a deferred step */
#define AMRH_LAUNCH_SECOND 4  /* the substep after it (first stage formed from U^n and D) */
#define AMRH_LAUNCH_GHOSTS 5  /* the octree's ghost fill at the start of a step */
#define AMRH_LAUNCH_FIRST_STAGE 6 /* the elementwise first-stage pass of a deferred step */
/* One step of amrh_run_steps (use_graph = 0) with HIP events recorded on the context's stream
 * between its stages: ms[s] is stage s's duration and kinds[s] its AMRH_LAUNCH_* kind, for
 * s < *nstages <= max_stages.  Synchronizes.  A measurement aid; no counterpart in the reference. */
int amrh_step_timed(amrh_ctx* ctx, int rk, double* dt_dev, float* ms, int* kinds, int max_stages, int* nstages);

/* ---- octree coupling (GridNode recursions, grid/grid_node.cpp) -------------------------------
 * Block b of the context is node b of the tree.  Node 0 is the root (parent[0] = -1); every other
 * node has a parent with a smaller index and child[b] in 0..7 (bit0 x, bit1 y, bit2 z;
 * oct_node/child_index.cpp:66-76).  Locations, offsets (grid/grid_node.cpp:276), cell widths
 * (h0 = 2*grid_dim/n, halved per level; grid/grid.cpp:70, grid/grid_node.cpp:279) and the ghost,
 * reflux and restriction tables are derived here.  Replaces OctNode::create_child's wiring
 * (oct_node/oct_node.cpp:172-240) and GridNode::create_child's boundary objects
 * (grid/grid_node.cpp:271-304) for a tree built on the host.  Runs the ideal gas with minmod
 * (g = 2) or PPM (n = 8, g = 3) and the rotating-frame state (n = 8; minmod g = 2 or PPM g = 3),
 * whose node origins (Grid::xf, grid/grid.cpp:512-526) follow from the tree.  For the rotating
 * state it allocates a third state buffer: its octree substep reads the cells across same-level
 * faces from the neighbours' input, so it never updates in place. */
int amrh_tree_set(amrh_ctx* ctx, const int* parent, const int* child);
/* GridNode::set_origin (grid/grid_node.cpp:20-26): the origin subtracted from every node's cell and
 * face positions, Grid::xf = (offset + i) h - GRID_DIM - bw h0 - origin (grid/grid.cpp:512-524); as
 * there, z subtracts origin[2] only when the lower z boundary reflects.  BinaryDriver moves it to
 * the centre of mass (binary/binary_driver.cpp:544-551).  Default 0; kept across amrh_tree_set,
 * carried into a repacked tree and through checkpoints.  Synchronous. */
int amrh_tree_set_origin(amrh_ctx* ctx, const double* origin);
int amrh_tree_get_origin(const amrh_ctx* ctx, double* origin);
/* GridNode::enforce_boundaries (grid/grid_node.cpp:59-112), level by level: same-level copy,
 * coarse->fine interpolation (grid/grid_amr_bound.cpp:27-78), physical boundary
 * (grid/grid_phys_bound.cpp:27-54). */
int amrh_tree_fill_ghosts(amrh_ctx* ctx);
/* HydroDriver::sub_step + inject_from_children (euler/hydro_driver.cpp:171-214): ghost fill,
 * fluxes with reflux (grid/grid_node.cpp:1076-1274), divergence, add_difs on every node, the root's
 * flow-off ledger (grid/grid.cpp:350-404, 419), then restriction (grid/grid_node.cpp:1035-1074).
 * The rotating-frame state runs BinaryDriver::sub_step's hydro part (binary/binary_driver.cpp:
 * 346-372): the same steps with Binary::add_difs's sources (binary/binary.cpp:205-254) from the
 * potential or the Poisson face forces and pot reset from the potential, level by level finest
 * first, and stops before the restriction: the driver solves for phi, calls amrh_pot_to_grid and
 * then amrh_tree_inject (BinaryDriver::solve, binary/binary_driver.cpp:296-344), or calls
 * amrh_tree_inject at once for a fixed potential. */
int amrh_tree_substep(amrh_ctx* ctx, double dt, double beta);
int amrh_tree_substep_dev(amrh_ctx* ctx, const double* dt_dev, double beta);
/* GridNode::inject_from_children (grid/grid_node.cpp:1035-1074). */
int amrh_tree_inject(amrh_ctx* ctx);
/* GridNode::max_dt (grid/grid_node.cpp:163-174): per-node Grid::max_dt from the device scan,
 * combined on the host in the reference's recursion.  Synchronous. */
int amrh_tree_max_dt(amrh_ctx* ctx, double* dt);
/* The root's flow-off ledger FO (Grid::flow_off, grid/grid.cpp:346-348): nf doubles. */
int amrh_tree_flow_off(amrh_ctx* ctx, double* fo);

/* ---- regrid and checkpoints (octree) ------------------------------------------------------------
 * After the host changes the tree (GridNode::check_for_refine, grid/grid_node.cpp:957-1033), a
 * context for the new tree (amrh_create + amrh_tree_set) is filled from the old one: node b takes
 * the record of src node src_index[b], or, with src_index[b] = -1 (a node created by refinement),
 * is filled from its parent by GridNode::inject_from_parent (grid/grid_node.cpp:193-233: minmod
 * slopes per axis, then the floor), level by level; the root's flow-off ledger carries over; ghosts
 * are filled at the end.  Nodes absent from the new tree are dropped (their parents already hold
 * the restricted values).  The rotating-frame state is prolonged the same way, with the species
 * renormalisation and the Newton T of its floor at the child's cells; its potential for the new
 * nodes comes from amrh_poisson_phi_from_state (Binary::inject_from_parent,
 * binary/binary.cpp:852-855).  The new tree takes the old one's origin.  Synchronous. */
int amrh_tree_repack(amrh_ctx* dst, amrh_ctx* src, const int* src_index);
/* The density test of GridNode::set_refine_flags (grid/grid_node.cpp:647-667): flags[b] bit c is
 * set when an interior cell of node b's octant c has rho > min_density.  The flag propagation,
 * proper nesting and mirroring rules (:631-956) stay with the host's tree.  Synchronous. */
int amrh_tree_density_flags(amrh_ctx* ctx, double min_density, unsigned char* flags);
/* GridNode::write (grid/grid_node.cpp:29-43, grid/grid.cpp:104-124) of the whole tree to `path`, in
 * the reference's byte layout: depth first, per node h, h0, time, offsets and origins, the ledger
 * FO (the root's; zero elsewhere), the padded state (x slowest, field fastest; edge and corner
 * ghosts as zeros), child flags, refine_alpha and age (ages: one per node, or NULL for -1). */
int amrh_tree_checkpoint_write(amrh_ctx* ctx, const char* path, double time, double refine_alpha,
                               const int* ages);
/* GridNode::read (grid/grid_node.cpp:45-57): a new context (cfg gives n, g, nf and the constants)
 * holding the tree and state of the checkpoint; its node count, the root's time and refine_alpha
 * are returned. */
int amrh_tree_checkpoint_read(const char* path, const amrh_config* cfg, amrh_ctx** out, int* nblocks,
                              double* time, double* refine_alpha);
/* The same with a caller-owned block of bytes before the root's record: Binary::write /
 * Binary::read (binary/binary.cpp:70-93, :175-203) put the binary's parameters there (a bool, 15
 * doubles: l1x, h, M1, fill_factor, n, R1, q, a, K1, K2, phi0_1, phi0_2, Omega, phase, Omega0, and
 * two ints: accretor and donor type; 129 bytes).  The library writes the bytes as given and reads
 * them back into root_prefix (NULL: skipped). */
int amrh_tree_checkpoint_write_ex(amrh_ctx* ctx, const char* path, double time, double refine_alpha,
                                  const int* ages, const void* root_prefix, long prefix_bytes);
int amrh_tree_checkpoint_read_ex(const char* path, const amrh_config* cfg, amrh_ctx** out, int* nblocks,
                                 double* time, double* refine_alpha, void* root_prefix, long prefix_bytes);
/* The mesh GridNode::output hands to Silo (grid/grid_node.cpp:1315-1356, :1422-1487): nodes in
 * depth-first order; per node the (GNX+1)^3 cell-corner points xf(i), yf(j), zf(k), x fastest; one
 * hexahedral zone per interior cell not covered by a child, in k, j, i order, with its 8 corner
 * indices in the reference's order; per zone the state's fields.  Counts first: npoints =
 * nodes * (GNX+1)^3, nzones = uncovered interior cells. */
int amrh_tree_output_counts(amrh_ctx* ctx, long* npoints, long* nzones);
/* x, y, z: npoints doubles each; nodelist: 8 * nzones ints; values: nf * nzones doubles, field
 * major.  Any pointer may be NULL to skip that part.  Synchronous. */
int amrh_tree_output_mesh(amrh_ctx* ctx, double* x, double* y, double* z, int* nodelist, double* values);
/* The same mesh and zone fields as a legacy binary VTK unstructured grid (readable where the
 * reference's Silo files are read); the fields are named as State::field_name
 * (euler/state.cpp:24-42).  Synchronous. */
int amrh_tree_output_vtk(amrh_ctx* ctx, const char* path, double time);

/* ---- diagnostics (device reductions) ---------------------------------------------------------
 * GridNode::state_sum (grid/grid_node.cpp:1489-1527) over the uncovered interior cells of every
 * block (a cell is covered when its octant has a child, :1355-1362): sums[q*nf + f] for q = 0..3
 * is sum(U_f h^3), sum(U_f h^3 x), sum(U_f h^3 y), sum(U_f h^3 z) at cell centres.  A flat batch
 * needs amrh_set_block_origin first; an octree's positions follow from amrh_tree_set.  The block
 * sums are combined in block order, so the result does not depend on the launch.  Synchronous. */
int amrh_state_sum(amrh_ctx* ctx, double* sums);
/* GridNode::state_max / state_min (grid/grid_node.cpp:1529-1589): per field over the same cells. */
int amrh_state_max(amrh_ctx* ctx, double* out);
int amrh_state_min(amrh_ctx* ctx, double* out);
/* amrh_state_sum for callers that cannot allocate device memory (work captured into a HIP graph,
 * per-step loops that forbid hipMalloc).  No counterpart in the reference.  The same two reductions
 * and the same result, written to sums_dev (4*nf doubles in device memory, amrh_state_sum's
 * order) asynchronously on the context's stream; the call allocates nothing and copies nothing to
 * the host.  The per-block partials live in a fixed array of the library, one per GPU, of
 * AMRH_STATE_SUM_DEV_MAX_PARTIALS doubles (24 MiB): nblocks * nf * 6 may not exceed it (AMRH_EINVAL
 * otherwise; e.g. at most 87,381 blocks of 6 fields or 40,329 of 13).  All calls on one GPU share
 * that array, so calls that overlap in time on one GPU race. */
#define AMRH_STATE_SUM_DEV_MAX_PARTIALS 3145728L
int amrh_state_sum_dev(amrh_ctx* ctx, double* sums_dev);

/* ---- block positions, rotating-frame state (AMRH_PHYS_ROTATING) --------------------------------
 * Block b's interior spans [x0, x0 + n h) on each axis (x0: nblocks * 3 doubles, host): the cell
 * and face positions of Grid::X / Xfx (grid/grid.cpp:13-45, :500-526), used by the rotating state
 * and by amrh_state_sum.  Flat batches only.  Synchronous. */
int amrh_set_block_origin(amrh_ctx* ctx, const double* x0);
/* The gravitational potential phi of Poisson::compute_forces (poisson/poisson.cpp:223-247): per
 * block (n+2)^3 doubles, x fastest, cell (i, j, k) of the interior at ((k+1)(n+2) + j+1)(n+2) + i+1
 * (one ghost layer).  It stays fixed across substeps: after every substep the pot field is reset to
 * rho (phi - R^2 omega0^2 / 2), Binary::pot_to_grid (binary/binary.cpp:825-849).  Host array,
 * synchronous; the _device form takes a device pointer in the same layout. */
int amrh_set_potential(amrh_ctx* ctx, const double* phi);
int amrh_set_potential_device(amrh_ctx* ctx, const void* phi_dev);
/* The gravitational force of Binary::add_difs (binary/binary.cpp:225-227) from Poisson's face
 * forces instead of phi's differences: fx of every block, then fy, then fz (3 * nblocks * (n+2)^3
 * doubles), each in phi's layout, fx(i) on the face between cells i-1 and i; the cell's force is
 * the mean of its two faces'.  Where a fine node meets a coarse one these carry the coarse-fine
 * averaging of Poisson::adjust_*_force (poisson/poisson.cpp:474-635).  Host array (copied,
 * synchronous) or a device pointer kept by reference (amrh_poisson_forces_device); NULL returns
 * to phi's differences. */
int amrh_set_forces(amrh_ctx* ctx, const double* f);
int amrh_set_forces_device(amrh_ctx* ctx, const void* f_dev);
/* Binary::pot_to_grid (binary/binary.cpp:825-849): pot = rho (phi - R^2 omega0^2 / 2) on every
 * block's interior from the current potential.  Asynchronous. */
int amrh_pot_to_grid(amrh_ctx* ctx);

/* ---- Helmholtz EOS (helmholtz/helmeos.cpp) -------------------------------------------------
 * The electron-positron free-energy table and the EOS built on it, on the device, FP64, in the
 * reference's operation order: pressure, energy, sound speed and heat capacity are bit-identical
 * to the FP64 CPU path (entropy to an ulp of log).  A state is the reference's eos_t
 * (helmholtz/helmholtz.h:4-6): 10 doubles p, e, cs, cv, abar, zbar, rho, T, s, dsdt. */
typedef struct amrh_eos amrh_eos;

/* Device table from `rows`: IMAX*JMAX = 1081*401 records of 9 doubles (f, fd, ft, fdd, ftt, fdt,
 * fddt, fdtt, fddtt), temperature outer, density inner -- the layout of the helmholtz.table file
 * that read_helm_table parses (helmholtz/helmeos.cpp:68-131); node coordinates and spacings are
 * formed as there.  The device keeps, per interpolation cell, the 36 coefficients one evaluation
 * reads, contiguous (125 MB).  Synchronous. */
int amrh_eos_create(const double* rows, amrh_eos** out);
/* Same, parsing a helmholtz.table text file. */
int amrh_eos_load(const char* path, amrh_eos** out);
void amrh_eos_destroy(amrh_eos* eos);
/* helmholtz_set_cgs_units (helmholtz/helmeos.cpp:133-142): code units in cgs; default all 1. */
int amrh_eos_set_units(amrh_eos* eos, double cm, double g, double s, double K);
/* helmholtz_eos (helmholtz/helmeos.cpp:164-421) on n states in device memory (st_dev: n * 10
 * doubles), in place, on `stream` (a hipStream_t; NULL = the null stream).  Asynchronous. */
int amrh_eos_eval(amrh_eos* eos, int n, void* st_dev, void* stream);
/* helmholtz_compute_T, energy form (helmholtz/helmeos.cpp:436-486): damped Newton for the T at
 * which e(rho, T) equals the state's e, starting from the state's T; in place.  Asynchronous. */
int amrh_eos_compute_T(amrh_eos* eos, int n, void* st_dev, void* stream);
/* The mixed form's face-state EOS (what a mixed rotating substep evaluates at its faces, the
 * reference's helmholtz_eos as called from binary/state_ztwd.cpp:360-396): p and cs from rho, T,
 * abar, zbar rounded to FP32, by the FP32 monomial form of each interpolation cell, written to the
 * states' p and cs; the other fields are left as they are.  Asynchronous. */
int amrh_eos_eval_f32(amrh_eos* eos, int n, void* st_dev, void* stream);
/* Device error flags raised since the last call, where the reference aborts: 8 state outside
 * the table, 16 negative heat capacity, 32 negative squared sound speed, 64 Newton did not
 * converge in 100 iterations, 128 T fell below TMIN.  Synchronises the device; clears them. */
int amrh_eos_error(amrh_eos* eos, int* flags);
/* The EOS table used by a context with eos = AMRH_EOS_HELMHOLTZ (State::compute_eos,
 * binary/state_ztwd.cpp:168-176); the table must outlive the context.  Its error flags are raised
 * in the context's (amrh_error). */
int amrh_attach_eos(amrh_ctx* ctx, amrh_eos* eos);

/* ---- Self-gravity: the multigrid Poisson solver on the octree (poisson/) --------------------
 * Poisson : GridNode (poisson/poisson.h:14-69) over every node of a tree given as (parent, child
 * index) arrays, parent before child, node 0 the root -- the same arrays as amrh_tree_set.  Per
 * node the solver keeps (n+2)^3 cubes with one ghost layer, x fastest (PNX = GNX - 2 BW + 2,
 * poisson/poisson.h:4): the fields below.  phi's layout is amrh_set_potential's, so
 * amrh_set_potential_device(rotating_ctx, amrh_poisson_phi_device(p)) hands the solved potential
 * to a rotating-state batch whose blocks are the tree's nodes. */
typedef struct amrh_poisson amrh_poisson;
enum {
  AMRH_PSN_PHI = 0,   /* the potential */
  AMRH_PSN_PHI0 = 1,  /* store_phi's copy */
  AMRH_PSN_DPHI = 2,  /* the V-cycle's correction; after a cycle, the residual of solution_error */
  AMRH_PSN_DPHI1 = 3, /* the V-cycle's right-hand side */
  AMRH_PSN_S = 4,     /* the source 4 pi G rho (poisson/poisson.cpp:956-972) */
  AMRH_PSN_FX = 5,    /* face forces of compute_forces: fx(i) on the face between i-1 and i */
  AMRH_PSN_FY = 6,
  AMRH_PSN_FZ = 7
};
/* Builds the tree as GridNode::create_child does (offsets with ghost width bw, grid/grid_node.cpp:
 * 270-299) and the neighbour objects of each face: a node, a PoissonPhysBound table (reflecting on
 * the low z face when z_reflect, poisson/poisson_phys_bound.cpp:8-60) or a PoissonAMRBound on the
 * parent.  n = 8, 12 or 16.  stream: a hipStream_t, or NULL for a private one.  Fields start at 0. */
int amrh_poisson_create(int n, int bw, double grid_dim, int z_reflect, int nnodes, const int* parent,
                        const int* child, void* stream, amrh_poisson** out);
void amrh_poisson_destroy(amrh_poisson* p);
/* Host copies of one field for every node (nnodes * (n+2)^3 doubles).  Synchronous. */
int amrh_poisson_upload(amrh_poisson* p, int field, const double* host);
int amrh_poisson_download(amrh_poisson* p, int field, double* host);
/* Device pointer to phi (nnodes * (n+2)^3 doubles), valid while p lives. */
const void* amrh_poisson_phi_device(const amrh_poisson* p);
/* Device pointer to the face forces fx, fy, fz (3 * nnodes * (n+2)^3 doubles, amrh_set_forces's
 * layout), valid while p lives. */
const void* amrh_poisson_forces_device(const amrh_poisson* p);
/* Poisson::set_source (poisson/poisson.cpp:956-972) from a hydro batch whose blocks are the same
 * nodes (same n): S = 4 pi G rho on the interior (State::poisson_source,
 * binary/state_ztwd.cpp:219-222).  Synchronous. */
int amrh_poisson_source_from_state(amrh_poisson* p, const amrh_ctx* hydro, double G);
/* Binary::pot_from_grid (binary/binary.cpp:857-884) from a rotating-state context whose blocks are
 * the same nodes: phi = pot / rho + R^2 omega0^2 / 2 on every interior cell, then, once
 * amrh_poisson_prepare has made the boundary tables, the ghosts of every level
 * (enforce_phi_boundaries).  Used for the first
 * solve and after a regrid, where Binary::inject_from_parent recomputes a new node's phi this way
 * (binary/binary.cpp:852-855).  Synchronous. */
int amrh_poisson_phi_from_state(amrh_poisson* p, const amrh_ctx* hydro);
/* Binary::compute_source_sums and compute_physical_boundaries (binary/binary.cpp:938-1019): per
 * node and octant the source's monopole, and the boundary phi of every physical face as the
 * tree's sum of -alpha m / r (alpha = 2 with z reflection).  Synchronous. */
int amrh_poisson_prepare(amrh_poisson* p);
/* The boundary phi tables as [nnodes][6][(n+2)^2] (index [a][c] over the face's other two axes in
 * x, y, z order); entries of faces that are not physical are ignored / returned as 0.
 * write = 1 replaces the tables (a caller's own boundary condition).  Synchronous. */
int amrh_poisson_boundary_tables(amrh_poisson* p, double* tab, int write);
/* Poisson::phi_from_children (poisson/poisson.cpp:707-743).  Synchronous. */
int amrh_poisson_phi_from_children(amrh_poisson* p);
/* enforce_phi_boundaries for every level, root first (poisson/poisson.cpp:745-902). */
int amrh_poisson_phi_boundaries(amrh_poisson* p);
/* compute_forces with the coarse-fine averaging of adjust_{x,y,z}_force (poisson/poisson.cpp:
 * 223-247, 474-635) and solution_error (:637-668) into *error (may be NULL).  Synchronous. */
int amrh_poisson_forces(amrh_poisson* p, double* error);
/* Poisson::vcycle (poisson/poisson.cpp:1069-1074) with the source already set: vcycle0, forces,
 * and the error sum |div f - S| h^3 over unrefined cells.  Needs the boundary tables. */
int amrh_poisson_vcycle(amrh_poisson* p, double* error);
/* BinaryDriver::solve (binary/binary_driver.cpp:294-344): prepare (unless prepare = 0, keeping
 * tables set by amrh_poisson_boundary_tables), forces, then V-cycles until the error is at most
 * toler; AMRH_ESTATE after max_iter cycles (the reference aborts).  The V-cycles start from the
 * phi the solver holds. */
int amrh_poisson_solve(amrh_poisson* p, double toler, int max_iter, int prepare, int* iters, double* error);
/* store_phi / restore_phi / restore_half_phi (poisson/poisson.cpp:144-221) around the RK stages of
 * BinaryDriver::sub_step (binary/binary_driver.cpp:362-373): mode 0, 1, 2. */
int amrh_poisson_store_phi(amrh_poisson* p, int mode);

#ifdef __cplusplus
}
#endif
#endif
