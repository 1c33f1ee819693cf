/* ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py CPU baseline).
 *
 * C API of the FP64 CPU restatement of the 3d-eulerian-amr hydro right-hand side.  Every array
 * is a "padded SoA block": per block, per field, a GNX^3 cube (GNX = n + 2g), x fastest, i.e.
 * element (f, i, j, k) of block b sits at ((b*nf + f)*GNX + k)*GNX*GNX + j*GNX + i.  This is the
 * reference's Array3d<State,GNX,GNX,GNX> (array3d.h:11, x fastest) with fields split out. */
#ifndef AMRH_ORACLE_H
#define AMRH_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n;          /* interior cells per edge (euler/euler_defs.h:27: GNX = 8 + 2*BW) */
  int g;          /* ghost width BW (euler/euler_defs.h:25) */
  int nf;         /* fields (euler/state.h:16 STATE_NF = 6) */
  int recon;      /* 0 = minmod theta (reconstruct.cpp:60-71), 1 = PPM (reconstruct.cpp:22-58,94) */
  int z_reflect;  /* euler/euler_defs.h:21 Z_REFLECT */
  double gamma, rho_floor, ei_floor, theta, cfl, grid_dim;
} orc_cfg;

/* --- per pencil (for pinning against the reference's own State/Reconstruct) ---------------- */
/* q: len x nf conserved states (AoS).  Writes face fluxes F[i] (face between cells i-1 and i) and
 * the two face eigenvalues eig[2*i], eig[2*i+1] for faces g..len-g; other entries untouched. */
int orc_pencil_flux(const orc_cfg* c, int dir, int len, const double* q, double* F, double* eig);
/* reconstructed face states (conserved, after to_con) ql[i], qr[i] for faces g..len-g */
int orc_pencil_faces(const orc_cfg* c, int len, const double* q, double* ql, double* qr);
void orc_floor(const orc_cfg* c, double* u);
void orc_outflow(const orc_cfg* c, int face, double* u);

/* --- flat batch of independent blocks (each with its own cell width h[b]) ------------------ */
/* One RK substep over every block: clear_difs, x/y/z flux + divergence, add_difs (a9-a11).
 * U: padded blocks (ghosts must be filled); U0: padded blocks (interior read).  Updates the
 * interior of U in place.  If D_out != NULL it receives the divergence (padded layout). */
int orc_batch_substep(const orc_cfg* c, int nb, double* U, const double* U0, const double* h,
                      double dt, double beta, double* D_out);
/* Per block: max over all interior faces of the three directions of |v|+cs (grid/grid.cpp:422-473,
 * before the division by h). */
int orc_batch_max_eig(const orc_cfg* c, int nb, const double* U, double* eig);
/* Grid::max_dt of one block with width h: 0.4*(1/(max_eig/h)) (grid/grid.cpp:474-481) */
double orc_block_dt(const orc_cfg* c, double max_eig, double h);
/* Ghost fill of isolated blocks whose six neighbours are all outflow boundaries
 * (grid/grid_phys_bound.cpp:27-54 with euler/state.cpp:45-84); z lower face reflects if z_reflect. */
int orc_batch_outflow_ghosts(const orc_cfg* c, int nb, double* U);
/* n RK steps of order rk (euler/hydro_driver.cpp:171-194) on a flat batch with outflow ghost
 * refresh before every substep; dt fixed.  U0 scratch of the same size as U. */
int orc_batch_rk_steps(const orc_cfg* c, int nb, double* U, double* U0, const double* h, double dt,
                       int rk, int nsteps);

/* --- reference-shaped restatement: the same arithmetic in the reference's structure (AoS states
 * with a virtual-table pointer, 56 B each; passes over nodes, each an OpenMP parallel-for inside
 * one node).  A CPU baseline and its parity check; flat batches of lone outflow nodes. */
void* orc_refshape_create(const orc_cfg* c, int nb, const double* U, const double* h);
void orc_refshape_destroy(void* r);
int orc_refshape_download(const void* r, double* U);
/* max over nodes of (max face eigenvalue / h): Grid::max_dt before the CFL factor */
double orc_refshape_max_eig(void* r);
/* one RK step of order rk with a fixed dt: store, and per substep the ghost fill and sub_step;
 * the ghosts are refilled at the end */
int orc_refshape_rk_step(void* r, double dt, int rk);

/* --- octree (GridNode recursion, grid/grid_node.cpp) -------------------------------------- */
typedef struct orc_tree orc_tree;
orc_tree* orc_tree_create(const orc_cfg* c);
void orc_tree_destroy(orc_tree* t);
/* add the child `child` (bit0 x, bit1 y, bit2 z; oct_node/child_index.cpp:66-76) of node `parent`;
 * returns the new node's index.  Node 0 is the root. */
int orc_tree_add_child(orc_tree* t, int parent, int child);
int orc_tree_nnodes(const orc_tree* t);
/* level, location[3], offset[3] (grid/grid_node.cpp:276), h */
int orc_tree_node_info(const orc_tree* t, int node, int* level, int* loc, int* off, double* h);
double* orc_tree_node_U(orc_tree* t, int node); /* padded nf x GNX^3, writable */
int orc_tree_enforce_boundaries(orc_tree* t);   /* grid/grid_node.cpp:59-112 */
int orc_tree_inject_from_children(orc_tree* t); /* grid/grid_node.cpp:1035-1074 */
int orc_tree_store(orc_tree* t);                /* grid/grid_node.cpp:495-515 */
double orc_tree_max_dt(const orc_tree* t);      /* grid/grid_node.cpp:163-174 */
/* HydroDriver::sub_step (euler/hydro_driver.cpp:200-214) followed by inject_from_children */
int orc_tree_substep(orc_tree* t, double dt, double beta);
/* sum over uncovered cells of U*h^3 (grid/grid_node.cpp:1489-1529, a0 only) and the root ledger */
int orc_tree_state_sum(const orc_tree* t, double* sum);
int orc_tree_flow_off(const orc_tree* t, double* fo);
/* GridNode::inject_from_parent (grid/grid_node.cpp:193-233) into node from its parent (whose
 * ghosts must be filled) */
int orc_tree_prolong(orc_tree* t, int node);
/* face flux array of a node for the last computed direction (debug) */
const double* orc_tree_node_F(const orc_tree* t, int node);

/* ---- Helmholtz EOS (helmholtz/helmeos.cpp), helmholtz.cpp --------------------------------
 * rows: 1081*401 records of 9 doubles in the helmholtz.table layout (temperature outer).
 * States: 10 doubles each in eos_t order (p, e, cs, cv, abar, zbar, rho, T, s, dsdt), in place;
 * flags[k] gets the error bits where the reference aborts (8 range, 16 cv < 0, 32 cs^2 < 0,
 * 64 Newton did not converge, 128 T < TMIN). */
void orc_helm_table(const double* rows);
void orc_helm_set_units(double cm, double g, double s, double K);
int orc_helm_eos(int n, double* st, int* flags);
int orc_helm_compute_T(int n, double* st, int* flags);

/* ---- rotating-frame binary-star state (binary/state_ztwd.cpp), rotating.cpp ----------------
 * 13 fields: rho, rho_acc, rho_don, s_R, l_z, s_z, pot, E, tau, T, 3 species.  Blocks are padded
 * SoA (nb, 13, G, G, G), G = n + 2g; phi (nb, n+2, n+2, n+2) is the given gravitational potential
 * with one ghost layer; x0 (nb, 3) the low corner of each block's interior, h (nb) its cell width. */
typedef struct {
  int n, g, recon, z_reflect;
  int eos;            /* 0 ideal gas p = rho T (1 + zbar)/abar, 1 Helmholtz (orc_helm_table first) */
  double theta;       /* minmod theta (PPM uses 2) */
  double rho_floor;   /* binary/state_ztwd.cpp:31 */
  double Omega, Omega0, fR, ftheta;  /* Binary::Omega, Omega0; State::fR, ftheta */
  double gamma;       /* ideal-gas EOS only */
  double cfl;
  double abar[3];     /* species mass numbers (binary/state_ztwd.cpp:28) */
  double newton_tol;  /* relative tolerance of the T Newton solve; 0 = the reference's 1e-9
                         (helmholtz/helmeos.cpp:449) */
} orc_rot_cfg;

int orc_rot_batch_substep(const orc_rot_cfg* c, int nb, double* U, const double* U0, const double* phi,
                          const double* x0, const double* h, double dt, double beta, int* flags);
int orc_rot_batch_max_eig(const orc_rot_cfg* c, int nb, const double* U, const double* x0, const double* h,
                          double* eig);
int orc_rot_batch_outflow_ghosts(const orc_rot_cfg* c, int nb, double* U, const double* x0, const double* h);
int orc_rot_state(const orc_rot_cfg* c, int op, double* u, const double* X, double* F);

/* The rotating-frame state on the octree: nodes as (parent, child index) arrays, parent first, node
 * 0 the root; per node U (13, G, G, G), phi (n+2)^3 and the Poisson face forces fx, fy, fz
 * (3 x (n+2)^3, face i between cells i-1 and i), read instead of phi's differences when
 * use_forces is set. */
typedef struct orc_rot_tree orc_rot_tree;
orc_rot_tree* orc_rot_tree_create(const orc_rot_cfg* c, double grid_dim, int nn, const int* parent,
                                  const int* cidx);
void orc_rot_tree_destroy(orc_rot_tree* t);
double* orc_rot_tree_U(orc_rot_tree* t, int node);
double* orc_rot_tree_phi(orc_rot_tree* t, int node);
double* orc_rot_tree_force(orc_rot_tree* t, int node);
const double* orc_rot_tree_F(const orc_rot_tree* t, int node, int d);
int orc_rot_tree_node_x0(const orc_rot_tree* t, int node, double* x0, double* h);
void orc_rot_tree_use_forces(orc_rot_tree* t, int on);
void orc_rot_tree_set_origin(orc_rot_tree* t, const double* origin);
int orc_rot_tree_fill_ghosts(orc_rot_tree* t);
int orc_rot_tree_store(orc_rot_tree* t);
int orc_rot_tree_substep(orc_rot_tree* t, double dt, double beta);
int orc_rot_tree_pot_to_grid(orc_rot_tree* t);
int orc_rot_tree_inject(orc_rot_tree* t);
/* GridNode::inject_from_parent (grid/grid_node.cpp:193-233) into node from its parent (ghosts
 * filled), floored at the child's cells; returns the floor's EOS flags */
int orc_rot_tree_prolong(orc_rot_tree* t, int node);
double orc_rot_tree_max_dt(orc_rot_tree* t);
int orc_rot_tree_flow_off(const orc_rot_tree* t, double* fo);
int orc_rot_tree_state_sum(const orc_rot_tree* t, double* sum);

/* ---- self-gravity: the multigrid Poisson solver on the octree (poisson/), poisson.cpp ------
 * Nodes are given as (parent, child index) arrays, parent before child, node 0 the root.  Every
 * per-node array is (n+2)^3 doubles, x fastest, one ghost layer (Poisson's PNX = GNX - 2 BW + 2).
 * field: 0 phi, 1 S (source, 4 pi G rho), 2 fx, 3 fy, 4 fz, 5 dphi, 6 dphi1, 7 phi0. */
typedef struct orc_poisson orc_poisson;
orc_poisson* orc_poisson_create(int n, int bw, double grid_dim, int z_reflect, int nn, const int* parent,
                                const int* cidx);
void orc_poisson_destroy(orc_poisson* p);
double* orc_poisson_field(orc_poisson* p, int node, int field);
/* source sums and the physical-boundary phi tables (binary/binary.cpp:938-1019) */
int orc_poisson_prepare(orc_poisson* p);
double* orc_poisson_boundary_table(orc_poisson* p, int node, int face); /* after prepare; P x P */
double orc_poisson_boundary_phi(const orc_poisson* p, double x, double y, double z);
double orc_poisson_cell_centre(const orc_poisson* p, int node, int axis, int i);
int orc_poisson_phi_from_children(orc_poisson* p);
int orc_poisson_enforce_phi(orc_poisson* p);              /* every level, root first */
int orc_poisson_enforce_phi_level(orc_poisson* p, int level);
int orc_poisson_forces(orc_poisson* p);                   /* compute_forces with adjust_force */
double orc_poisson_error(orc_poisson* p);                 /* solution_error */
double orc_poisson_vcycle(orc_poisson* p);                /* vcycle0 + forces + error */
/* BinaryDriver::solve: vcycles until the error is at most toler (1 if max_iter is exceeded) */
int orc_poisson_solve(orc_poisson* p, double toler, int max_iter, int prepare, int* iters, double* error);
int orc_poisson_store_phi(orc_poisson* p, int mode); /* 0 store, 1 restore, 2 restore_half */

#ifdef __cplusplus
}
#endif
#endif
