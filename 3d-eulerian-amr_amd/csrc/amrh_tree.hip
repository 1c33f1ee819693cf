// Octree topology on the host and the device tables the coupling kernels read.
//
// The reference wires OctNode siblings pointer by pointer as children are created
// (oct_node/oct_node.cpp:172-240) and puts GridAMRBound / GridPhysBound objects where a same-level
// neighbour is missing (grid/grid_node.cpp:271-304).  Here a node is identified by (level,
// location); a face's ghost source is the same-level node at location +/- 1 if it exists, the
// physical boundary if the location is at the domain edge (oct_node/oct_node.cpp:86-110), else the
// parent (coarse->fine interpolation).  Offsets and widths follow grid/grid_node.cpp:276-279.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

#include "amrh_ctx.h"
#include "amrh_internal.h"

using namespace amrh;

void TreeState::free_dev() {
  for (void* p : {(void*)ghost_dev, (void*)rel_dev, (void*)level_nodes_dev, (void*)slot_dev, (void*)ovr_dev,
                  (void*)items_dev, (void*)pairs_dev, (void*)ledger_dev, (void*)fo_dev, (void*)covered_dev,
                  (void*)exp_dev})
    if (p) (void)hipFree(p);
  ghost_dev = rel_dev = level_nodes_dev = slot_dev = exp_dev = nullptr;
  ovr_dev = ledger_dev = fo_dev = nullptr;
  covered_dev = nullptr;
  items_dev = nullptr;
  pairs_dev = nullptr;
}

namespace {

template <typename T>
hipError_t upload_vec(T** dst, const std::vector<T>& v) {
  const size_t bytes = sizeof(T) * std::max<size_t>(v.size(), 1);
  hipError_t e = hipMalloc(dst, bytes);
  if (e != hipSuccess) return e;
  if (!v.empty()) e = hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  return e;
}

int build(amrh_ctx* c, const int* parent, const int* child) {
  const int n = c->cfg.n, g = c->cfg.g, G = n + 2 * g, nf = c->cfg.nf;
  TreeState* t = new TreeState;
  t->nn = c->nb;
  {
    // four workgroups per CU: on the 18,752-node finest level of the binary's octree 2-3% faster
    // than one workgroup per node (the flat rotating batch: 7%)
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      t->level_grid = 4 * cus;
  }
  t->parent.assign(t->nn, -1);
  t->cidx.assign(t->nn, -1);
  t->level.assign(t->nn, 0);
  t->loc.assign(t->nn, {0, 0, 0});
  t->off.assign(t->nn, {0, 0, 0});
  t->child.assign(t->nn, {-1, -1, -1, -1, -1, -1, -1, -1});
  t->h.assign(t->nn, 0.0);
  std::map<std::tuple<int, int, int, int>, int> by_loc;
  t->h[0] = 2.0 * c->cfg.grid_dim / double(n);  // grid/grid.cpp:70
  by_loc[std::make_tuple(0, 0, 0, 0)] = 0;
  if (parent[0] != -1) {
    delete t;
    return fail(AMRH_EINVAL, "node 0 must be the root (parent -1)");
  }
  for (int b = 1; b < t->nn; b++) {
    const int p = parent[b], ci = child[b];
    if (p < 0 || p >= b || ci < 0 || ci > 7 || t->child[p][ci] >= 0) {
      delete t;
      return fail(AMRH_EINVAL, "node %d: bad parent %d / child index %d", b, p, ci);
    }
    t->parent[b] = p;
    t->cidx[b] = ci;
    t->child[p][ci] = b;
    t->level[b] = t->level[p] + 1;
    const int cv[3] = {ci & 1, (ci >> 1) & 1, (ci >> 2) & 1};
    for (int d = 0; d < 3; d++) {
      t->loc[b][d] = 2 * t->loc[p][d] + cv[d];
      t->off[b][d] = 2 * t->off[p][d] + g + cv[d] * (G - 2 * g);
    }
    t->h[b] = t->h[p] * 0.5;
    by_loc[std::make_tuple(t->level[b], t->loc[b][0], t->loc[b][1], t->loc[b][2])] = b;
    t->maxlevel = std::max(t->maxlevel, t->level[b]);
  }
  auto sibling = [&](int b, int face) {
    const int d = face >> 1;
    std::array<int, 3> l = t->loc[b];
    l[d] += (face & 1) ? 1 : -1;
    if (l[d] < 0 || l[d] >= (1 << t->level[b])) return -1;
    auto it = by_loc.find(std::make_tuple(t->level[b], l[0], l[1], l[2]));
    return it == by_loc.end() ? -1 : it->second;
  };
  auto phys = [&](int b, int face) {
    const int d = face >> 1;
    return (face & 1) ? t->loc[b][d] == (1 << t->level[b]) - 1 : t->loc[b][d] == 0;
  };
  // ghost sources
  std::vector<int> ghost(t->nn * 12), rel(t->nn * 3, 0);
  for (int b = 0; b < t->nn; b++) {
    for (int f = 0; f < 6; f++) {
      const int s = sibling(b, f);
      int type, src;
      if (s >= 0) {
        type = GHOST_SIB;
        src = s;
      } else if (phys(b, f)) {
        type = GHOST_PHYS;
        src = b;
      } else {
        type = GHOST_COARSE;
        src = t->parent[b];
        if (src < 0) {
          delete t;
          return fail(AMRH_EINVAL, "node %d face %d has no neighbour and no parent", b, f);
        }
      }
      ghost[(b * 6 + f) * 2] = type;
      ghost[(b * 6 + f) * 2 + 1] = src;
    }
    if (t->parent[b] >= 0)
      for (int d = 0; d < 3; d++) rel[b * 3 + d] = t->off[b][d] - 2 * t->off[t->parent[b]][d];
  }
  // nodes per level
  t->level_start.assign(t->maxlevel + 1, 0);
  t->level_count.assign(t->maxlevel + 1, 0);
  for (int l = 0; l <= t->maxlevel; l++) {
    t->level_start[l] = (int)t->level_nodes.size();
    for (int b = 0; b < t->nn; b++)
      if (t->level[b] == l) t->level_nodes.push_back(b);
    t->level_count[l] = (int)t->level_nodes.size() - t->level_start[l];
  }
  // reflux slots and export items (grid/grid_node.cpp:1076-1274)
  std::vector<int> slot(t->nn * 36, -1);
  std::vector<std::vector<ExportItem>> by_level(t->maxlevel + 1);
  auto child_of = [&](int node, int c3x, int c3y, int c3z) {
    if (node < 0) return -1;
    return t->child[node][c3x | (c3y << 1) | (c3z << 2)];
  };
  for (int b = 0; b < t->nn; b++)
    for (int d = 0; d < 3; d++) {
      const int sL = sibling(b, 2 * d), sU = sibling(b, 2 * d + 1);
      for (int pl = 0; pl < 3; pl++)
        for (int ck = 0; ck < 2; ck++)
          for (int cj = 0; cj < 2; cj++) {
            auto ci = [&](int node, int nb) {
              if (d == 0) return child_of(node, nb, cj, ck);
              if (d == 1) return child_of(node, cj, nb, ck);
              return child_of(node, cj, ck, nb);
            };
            int cl, cr;
            if (pl == 0) { cr = ci(b, 0); cl = ci(sL, 1); }
            else if (pl == 1) { cl = ci(b, 0); cr = ci(b, 1); }
            else { cr = ci(sU, 0); cl = ci(b, 1); }
            int fine, side;
            if (cr < 0 && cl >= 0) { fine = cl; side = 1; }
            else if (cl < 0 && cr >= 0) { fine = cr; side = 0; }
            else continue;
            const int s = t->nslots++;
            slot[((b * 3 + d) * 3 + pl) * 4 + cj + 2 * ck] = s;
            by_level[t->level[fine]].push_back(ExportItem{fine, d, side, s});
          }
    }
  // per node and face: the slots its boundary-face means go to -- the parent's and, across a
  // parent's face, the parent's neighbour's (both adjust their face there)
  std::vector<int> exp_slot(t->nn * 12, -1);
  for (int l = 0; l <= t->maxlevel; l++)
    for (const ExportItem& it : by_level[l]) {
      int* e2 = &exp_slot[(it.fine * 6 + 2 * it.dir + it.side) * 2];
      if (e2[0] < 0) {
        e2[0] = it.slot;
      } else if (e2[1] < 0) {
        e2[1] = it.slot;
      } else {
        delete t;
        return fail(AMRH_EINVAL, "node %d face %d: more than two coarse faces take its fluxes", it.fine,
                    2 * it.dir + it.side);
      }
    }
  std::vector<ExportItem> items;
  t->item_start.assign(t->maxlevel + 1, 0);
  t->item_count.assign(t->maxlevel + 1, 0);
  for (int l = 0; l <= t->maxlevel; l++) {
    t->item_start[l] = (int)items.size();
    items.insert(items.end(), by_level[l].begin(), by_level[l].end());
    t->item_count[l] = (int)by_level[l].size();
  }
  // restriction pairs per parent level
  std::vector<int2> pairs;
  t->pair_start.assign(t->maxlevel + 1, 0);
  t->pair_count.assign(t->maxlevel + 1, 0);
  for (int l = 0; l <= t->maxlevel; l++) {
    t->pair_start[l] = (int)pairs.size();
    for (int b = 0; b < t->nn; b++)
      if (t->level[b] == l)
        for (int ci = 0; ci < 8; ci++)
          if (t->child[b][ci] >= 0) pairs.push_back(make_int2(b, (ci << 24) | t->child[b][ci]));
    t->pair_count[l] = (int)pairs.size() - t->pair_start[l];
  }
  hipError_t e = upload_vec(&t->ghost_dev, ghost);
  if (e == hipSuccess) e = upload_vec(&t->rel_dev, rel);
  if (e == hipSuccess) e = upload_vec(&t->level_nodes_dev, t->level_nodes);
  if (e == hipSuccess) e = upload_vec(&t->slot_dev, slot);
  if (e == hipSuccess) e = upload_vec(&t->items_dev, items);
  if (e == hipSuccess) e = upload_vec(&t->exp_dev, exp_slot);
  if (e == hipSuccess) e = upload_vec(&t->pairs_dev, pairs);
  std::vector<unsigned char> covered(t->nn, 0);
  for (int b = 0; b < t->nn; b++)
    for (int ci = 0; ci < 8; ci++)
      if (t->child[b][ci] >= 0) covered[b] |= (unsigned char)(1u << ci);
  if (e == hipSuccess) e = upload_vec(&t->covered_dev, covered);
  const long half = n / 2;
  if (e == hipSuccess) e = hipMalloc(&t->ovr_dev, sizeof(double) * std::max<long>(1, t->nslots) * half * half * nf);
  if (e == hipSuccess) e = hipMalloc(&t->ledger_dev, sizeof(double) * 3 * 2 * n * n * nf);
  if (e == hipSuccess) e = hipMalloc(&t->fo_dev, sizeof(double) * 2 * nf);
  if (e == hipSuccess) e = hipMemset(t->fo_dev, 0, sizeof(double) * 2 * nf);
  if (e == hipSuccess) e = hipMemset(t->ledger_dev, 0, sizeof(double) * 3 * 2 * n * n * nf);
  if (e != hipSuccess) {
    t->free_dev();
    delete t;
    return fail(AMRH_EHIP, "tree tables: %s", hipGetErrorString(e));
  }
  if (c->tree) {
    c->tree->free_dev();
    delete c->tree;
  }
  c->tree = t;
  return AMRH_OK;
}

int level_grid(const TreeState* t, int count) { return count > t->level_grid ? t->level_grid : 0; }

// partial: the rotating state's leaves keep their same-level-neighbour faces unfilled (its octree
// substep reads those cells from the neighbours); every other ghost is filled
int fill_ghosts(amrh_ctx* c, bool partial = false) {
  TreeState* t = c->tree;
  const bool rot = c->cfg.physics == AMRH_PHYS_ROTATING;
  const rot::RotParams rp = rot ? ctx_rot_args(c).c : rot::RotParams{};
  for (int l = 0; l <= t->maxlevel; l++) {
    const int* nodes = t->level_nodes_dev + t->level_start[l];
    if (rot)
      HIPCHK(launch_rot_tree_ghosts(c->cfg.g, c->buf[c->cur], nodes, t->level_count[l], t->ghost_dev, t->rel_dev,
                                    c->x0_dev, c->h_dev, c->cfg.z_reflect, rp, c->err_dev, c->stream,
                                    partial ? t->covered_dev : nullptr));
    else
      HIPCHK(launch_tree_ghosts(c->cfg.n, c->cfg.g, c->buf[c->cur], nodes, t->level_count[l], t->ghost_dev, t->rel_dev,
                                c->cfg.z_reflect, c->err_dev, c->stream));
  }
  return AMRH_OK;
}

int inject(amrh_ctx* c) {
  TreeState* t = c->tree;
  if (c->cfg.physics == AMRH_PHYS_ROTATING) {
    if (c->cfg.eos == AMRH_EOS_HELMHOLTZ && !c->eos) return fail(AMRH_ESTATE, "eos = HELMHOLTZ needs amrh_attach_eos");
    const RotArgs a = ctx_rot_args(c);
    for (int l = t->maxlevel - 1; l >= 0; l--)
      HIPCHK(launch_rot_restrict(c->cfg.g, c->cfg.eos, c->buf[c->cur], t->pairs_dev + t->pair_start[l],
                                 t->pair_count[l], c->x0_dev, c->h_dev, a.c, a.H, c->stream));
    return AMRH_OK;
  }
  const PhysParams P = params(c->cfg);
  for (int l = t->maxlevel - 1; l >= 0; l--)
    HIPCHK(launch_restrict(c->cfg.n, c->cfg.g, c->buf[c->cur], t->pairs_dev + t->pair_start[l], t->pair_count[l],
                           P, c->stream));
  return AMRH_OK;
}

// BinaryDriver::sub_step's hydro part (binary/binary_driver.cpp:346-372) on the octree: ghost fill,
// then the rotating substep level by level, finest first, each level applying the reflux means its
// finer neighbours exported and exporting its own boundary faces' means to its coarser neighbours;
// the root's ledger; no restriction (the driver restricts after the potential is in place)
int rot_tree_substep(amrh_ctx* c, double dt, double beta, const double* dt_dev, bool fill) {
  TreeState* t = c->tree;
  if (int rc = ctx_ready(c)) return rc;
  if (beta != 1.0 && c->u0buf < 0) return fail(AMRH_ESTATE, "beta != 1 needs amrh_store() first");
  if (fill)
    if (int rc = fill_ghosts(c, true)) return rc;
  RotArgs a = ctx_rot_args(c);
  a.Uin = c->buf[c->cur];
  // never in place: a node reads the cells across its same-level faces from the neighbours' input,
  // and a neighbour of its level may finish before it starts; nor over the RK base
  const int out = free_buf(c);
  if (out < 0) return fail(AMRH_ESTATE, "the rotating octree needs its third state buffer (amrh_tree_set)");
  a.Uout = c->buf[out];
  a.U0 = c->u0buf >= 0 ? c->buf[c->u0buf] : c->buf[c->cur];
  a.dt_dev = dt_dev;
  a.dt = dt;
  a.beta = beta;
  a.slot = t->slot_dev;
  a.ovr = t->ovr_dev;
  a.exp_slot = t->exp_dev;
  a.ledger = t->ledger_dev;
  a.ghost = t->ghost_dev;
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  for (int l = t->maxlevel; l >= 0; l--) {
    a.nodes = t->level_nodes_dev + t->level_start[l];
    a.nblocks = t->level_count[l];
    HIPCHK(launch_rot_substep(c->cfg.g, c->cfg.recon, c->cfg.eos, mixed, a, level_grid(t, a.nblocks), c->stream));
  }
  c->cur = out;
  HIPCHK(launch_ledger(c->cfg.n, c->cfg.nf, t->ledger_dev, t->h[0], dt_dev, dt, beta, c->cfg.z_reflect, t->fo_dev,
                       c->stream));
  return AMRH_OK;
}

// The ideal gas with PPM: the block kernel level by level, finest first, applying and exporting the
// reflux means as the rotating kernel does; the ledger; then the restriction
int block_tree_substep(amrh_ctx* c, double dt, double beta, const double* dt_dev, bool fill) {
  TreeState* t = c->tree;
  if (beta != 1.0 && c->u0buf < 0) return fail(AMRH_ESTATE, "beta != 1 needs amrh_store() first");
  if (fill)
    if (int rc = fill_ghosts(c)) return rc;
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
  a.P = params(c->cfg);
  a.err = c->err_dev;
  a.slot = t->slot_dev;
  a.ovr = t->ovr_dev;
  a.ovr_out = t->ovr_dev;
  a.exp_slot = t->exp_dev;
  a.ledger_block = 0;
  a.ledger = t->ledger_dev;
  a.z_reflect = c->cfg.z_reflect;
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  for (int l = t->maxlevel; l >= 0; l--) {
    a.nodes = t->level_nodes_dev + t->level_start[l];
    a.nblocks = t->level_count[l];
    HIPCHK(launch_block_substep(c->cfg.n, c->cfg.g, c->cfg.recon, a, mixed, 0, c->stream));
  }
  c->cur = out;
  HIPCHK(launch_ledger(c->cfg.n, c->cfg.nf, t->ledger_dev, t->h[0], dt_dev, dt, beta, c->cfg.z_reflect, t->fo_dev,
                       c->stream));
  return inject(c);
}

// fill = false: the caller filled the ghosts of the current state itself (and changed no cell since)
int tree_substep(amrh_ctx* c, double dt, double beta, const double* dt_dev, bool fill) {
  TreeState* t = c->tree;
  if (c->cfg.physics == AMRH_PHYS_ROTATING) return rot_tree_substep(c, dt, beta, dt_dev, fill);
  if (c->cfg.recon == AMRH_RECON_PPM) return block_tree_substep(c, dt, beta, dt_dev, fill);
  if (fill)
    if (int rc = fill_ghosts(c)) return rc;
  const PhysParams P = params(c->cfg);
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  for (int l = t->maxlevel; l >= 1; l--)
    HIPCHK(launch_reflux_export(c->cfg.n, mixed, c->buf[c->cur], t->items_dev + t->item_start[l],
                                t->item_count[l], t->slot_dev, t->ovr_dev, P, c->stream));
  if (int rc = substep_core(c, dt, beta, dt_dev, true)) return rc;
  HIPCHK(launch_ledger(c->cfg.n, c->cfg.nf, t->ledger_dev, t->h[0], dt_dev, dt, beta, c->cfg.z_reflect, t->fo_dev,
                       c->stream));
  return inject(c);
}

double block_dt(double m, double h, double cfl) {  // grid/grid.cpp:474-481
  double dtinv = m / h;
  if (dtinv == 0.0) dtinv = 1.0e-10;
  double dt = 1.0 / dtinv;
  dt *= cfl;
  return dt;
}

double max_dt_rec(const TreeState* t, const std::vector<double>& eig, double cfl, int b) {
  double dtinv = 0.0;  // grid/grid_node.cpp:163-174
  for (int ci = 0; ci < 8; ci++)
    if (t->child[b][ci] >= 0) {
      const double cdtinv = 1.0 / max_dt_rec(t, eig, cfl, t->child[b][ci]);
      dtinv = cdtinv > dtinv ? cdtinv : dtinv;
    }
  const double own = 1.0 / block_dt(eig[b], t->h[b], cfl);
  dtinv = own > dtinv ? own : dtinv;
  return 1.0 / dtinv;
}

}  // namespace

// The rotating state's RK step on the octree with its dt deferred.  First: the ghost fill a
// substep makes, then the deferred substep level by level, finest first -- the reflux overrides,
// exports and ledger faces as the substep makes them (none depends on dt), the max face speed over
// every node into gmax_dev (the CFL scan's numbers: the face states are the same, and the scan's
// same-level ghosts are copies of the cells read here directly), D into dbuf.
int amrh::tree_deferred_first(amrh_ctx* c) {
  TreeState* t = c->tree;
  if (int rc = ctx_ready(c)) return rc;
  if (!c->dbuf) return fail(AMRH_ESTATE, "no D buffer");
  if (int rc = fill_ghosts(c, true)) return rc;
  RotArgs a = ctx_rot_args(c);
  a.Uin = c->buf[c->cur];
  a.dbuf = c->dbuf;
  a.gmax = c->gmax_dev;
  a.beta = 1.0;
  a.slot = t->slot_dev;
  a.ovr = t->ovr_dev;
  a.exp_slot = t->exp_dev;
  a.ledger = t->ledger_dev;
  a.ghost = t->ghost_dev;
  const bool mixed = c->cfg.precision == AMRH_MIXED;
  for (int l = t->maxlevel; l >= 0; l--) {
    a.nodes = t->level_nodes_dev + t->level_start[l];
    a.nblocks = t->level_count[l];
    HIPCHK(launch_rot_tree_deferred(c->cfg.g, c->cfg.recon, c->cfg.eos, mixed, a, level_grid(t, a.nblocks), c->stream));
  }
  return AMRH_OK;
}

// Then, once dt_dev holds dt and U^n is stored as the RK base: the first stage of every node
// (U^n + D dt with the floor and the pot reset, as the substep's epilogue) into the other buffer,
// and the flow-off ledger of that substep (beta = 1).  The restriction is the caller's.
int amrh::tree_first_stage(amrh_ctx* c, const double* dt_dev) {
  TreeState* t = c->tree;
  if (c->u0buf != c->cur) return fail(AMRH_ESTATE, "the first stage needs U^n stored as the RK base");
  const int out = free_buf(c);
  if (out < 0) return fail(AMRH_ESTATE, "no free state buffer");
  RotArgs a = ctx_rot_args(c);
  a.Uin = c->buf[c->cur];
  a.Uout = c->buf[out];
  a.dbuf = c->dbuf;
  a.dt_dev = dt_dev;
  a.beta = 1.0;
  HIPCHK(launch_rot_stage(c->cfg.g, c->cfg.eos, c->cfg.precision == AMRH_MIXED, a, c->stream));
  c->cur = out;
  HIPCHK(launch_ledger(c->cfg.n, c->cfg.nf, t->ledger_dev, t->h[0], dt_dev, 0.0, 1.0, c->cfg.z_reflect, t->fo_dev,
                       c->stream));
  return AMRH_OK;
}

int amrh::tree_substep_dev(amrh_ctx* c, const double* dt_dev, double beta, bool fill) {
  if (!c->tree || !c->have_u) return fail(AMRH_ESTATE, "tree or state not set");
  return tree_substep(c, 0.0, beta, dt_dev, fill);
}

// The low interior face of every node, Grid::xf(bw) (grid/grid.cpp:512-526):
//   (offset + bw) h - GRID_DIM - bw h0 - origin   in x and y;
//   z: the same without -origin_z, or, when the lower z boundary reflects, without -GRID_DIM
//   (the reference subtracts origin_z only under Z_REFLECT).
int amrh::tree_upload_x0(amrh_ctx* c) {
  const TreeState* t = c->tree;
  const int g = c->cfg.g;
  const double h0 = t->h[0];
  std::vector<double> x0(3 * (size_t)c->nb);
  for (int b = 0; b < c->nb; b++)
    for (int d = 0; d < 3; d++) {
      double x = double(t->off[b][d] + g) * t->h[b];
      if (d < 2) {
        x = x - c->cfg.grid_dim - g * h0 - c->origin[d];
      } else if (c->cfg.z_reflect) {
        x = x - g * h0 - c->origin[2];
      } else {
        x = x - c->cfg.grid_dim - g * h0;
      }
      x0[3 * b + d] = x;
    }
  if (!c->x0_dev) HIPCHK(hipMalloc(&c->x0_dev, sizeof(double) * 3 * c->nb));
  HIPCHK(hipMemcpy(c->x0_dev, x0.data(), sizeof(double) * x0.size(), hipMemcpyHostToDevice));
  c->have_x0 = true;
  return AMRH_OK;
}

extern "C" {

int amrh_tree_set(amrh_ctx* c, const int* parent, const int* child) {
  if (!c || !parent || !child) return fail(AMRH_EINVAL, "null argument");
  // the ideal gas: minmod (g = 2, n = 8 or 16) or PPM (g = 3, n = 8); the rotating-frame state: n = 8,
  // minmod g = 2 or PPM g = 3 (amrh_create checks the pairs)
  const bool euler_ok = c->cfg.physics == AMRH_PHYS_EULER &&
                        (c->cfg.recon == AMRH_RECON_MINMOD || (c->cfg.recon == AMRH_RECON_PPM && c->cfg.n == 8));
  const bool rot_ok = c->cfg.physics == AMRH_PHYS_ROTATING;
  if (!euler_ok && !rot_ok)
    return fail(AMRH_EINVAL, "the octree coupling runs the ideal gas (minmod, or PPM on n = 8) or the rotating state");
  // captured steps hold the old tree's tables, which build() frees
  drop_graphs(c);
  if (rot_ok && !c->buf[2]) {
    // the rotating octree's substep writes a buffer other than its input (rot_tree_substep)
    const size_t bytes = sizeof(double) * (size_t)c->nb * (size_t)c->L.block();
    HIPCHK(hipMalloc(&c->buf[2], bytes));
    HIPCHK(hipMemset(c->buf[2], 0, bytes));
  }
  if (int rc = build(c, parent, child)) return rc;
  HIPCHK(hipMemcpy(c->h_dev, c->tree->h.data(), sizeof(double) * c->nb, hipMemcpyHostToDevice));
  c->have_h = true;
  return tree_upload_x0(c);
}

int amrh_tree_set_origin(amrh_ctx* c, const double* origin) {
  if (!c || !origin) return fail(AMRH_EINVAL, "null argument");
  for (int d = 0; d < 3; d++)
    if (!std::isfinite(origin[d])) return fail(AMRH_EINVAL, "origin[%d] is not finite", d);
  for (int d = 0; d < 3; d++) c->origin[d] = origin[d];
  if (!c->tree) return AMRH_OK;  // applied by amrh_tree_set
  return tree_upload_x0(c);
}

int amrh_tree_get_origin(const amrh_ctx* c, double* origin) {
  if (!c || !origin) return fail(AMRH_EINVAL, "null argument");
  for (int d = 0; d < 3; d++) origin[d] = c->origin[d];
  return AMRH_OK;
}

int amrh_tree_fill_ghosts(amrh_ctx* c) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->tree || !c->have_u) return fail(AMRH_ESTATE, "tree or state not set");
  return fill_ghosts(c);
}

int amrh_tree_inject(amrh_ctx* c) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->tree || !c->have_u) return fail(AMRH_ESTATE, "tree or state not set");
  return inject(c);
}

int amrh_tree_substep(amrh_ctx* c, double dt, double beta) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (!c->tree || !c->have_u) return fail(AMRH_ESTATE, "tree or state not set");
  return tree_substep(c, dt, beta, nullptr, true);
}

int amrh_tree_substep_dev(amrh_ctx* c, const double* dt_dev, double beta) {
  if (!c || !dt_dev) return fail(AMRH_EINVAL, "null argument");
  if (!c->tree || !c->have_u) return fail(AMRH_ESTATE, "tree or state not set");
  return tree_substep(c, 0.0, beta, dt_dev, true);
}

int amrh_tree_max_dt(amrh_ctx* c, double* dt) {
  if (!c || !dt) return fail(AMRH_EINVAL, "null argument");
  if (!c->tree) return fail(AMRH_ESTATE, "tree not set");
  std::vector<double> eig(c->nb);
  if (int rc = amrh_block_max_eig(c, eig.data())) return rc;
  *dt = max_dt_rec(c->tree, eig, c->cfg.cfl, 0);
  return AMRH_OK;
}

int amrh_tree_flow_off(amrh_ctx* c, double* fo) {
  if (!c || !fo) return fail(AMRH_EINVAL, "null argument");
  if (!c->tree) return fail(AMRH_ESTATE, "tree not set");
  HIPCHK(hipMemcpyAsync(fo, c->tree->fo_dev, sizeof(double) * c->cfg.nf, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRH_OK;
}

}  // extern "C"
