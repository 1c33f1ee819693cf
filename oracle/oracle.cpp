// ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py CPU baseline).
// The product path never links this file; it exists to check the HIP kernels.
//
// FP64 restatement of the 3d-eulerian-amr hydro right-hand side for the ideal-gas physics:
//   reconstruct.cpp:22-102         minmod-theta and PPM face reconstruction
//   grid/grid.cpp:173-483          store / clear_difs / compute_*_flux / sum_*_difs /
//                                  sum_*_flow_off / add_difs / max_dt
//   grid/grid_node.cpp:59-174, 337-572, 1035-1274   the octree recursions around them
//   grid/grid_amr_bound.cpp:27-78  coarse->fine ghost interpolation
//   grid/grid_phys_bound.cpp:27-169 physical (outflow / z-reflect) ghosts
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).
#include "oracle.h"
#include "euler.hpp"
#include "recon.hpp"

#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

using namespace orc;

namespace {


struct Geo {
  int n, g, nf, G;  // G = GNX
  long G3;
  Phys P;
  int recon, zref;
  explicit Geo(const orc_cfg* c)
      : n(c->n), g(c->g), nf(c->nf), G(c->n + 2 * c->g), G3((long)G * G * G), recon(c->recon),
        zref(c->z_reflect) {
    P.gamma = c->gamma;
    P.rho_floor = c->rho_floor;
    P.ei_floor = c->ei_floor;
    P.theta = c->theta;
    P.cfl = c->cfl;
  }
  long at(int i, int j, int k) const { return ((long)k * G + j) * G + i; }
};

// Cell index of pencil position p for direction d through the pencil (a, b):
//   x: (p, a, b)   y: (a, p, b)   z: (a, b, p)   (grid/grid.cpp:222-227, 252-257, 282-287)
inline long pencil_at(const Geo& o, int d, int a, int b, int p) {
  if (d == 0) return o.at(p, a, b);
  if (d == 1) return o.at(a, p, b);
  return o.at(a, b, p);
}

void recon_minmod(const Geo& o, int len, double q0[][MAXF], double ql[][MAXF], double qr[][MAXF]) {
  orc::recon_minmod(o.nf, o.P.theta, len, q0, ql, qr);
}

void recon_ppm(const Geo& o, int len, double q0[][MAXF], double qlc[][MAXF], double qrc[][MAXF],
               const unsigned low_order_mask, const unsigned smooth_mask) {
  orc::recon_ppm(o.nf, len, q0, qlc, qrc, low_order_mask, smooth_mask);
}

void reconstruct(const Geo& o, int len, double q0[][MAXF], double ql[][MAXF], double qr[][MAXF]) {
  if (o.recon == 1)
    recon_ppm(o, len, q0, ql, qr, 0u, 0u);
  else
    recon_minmod(o, len, q0, ql, qr);
}

// The body of one pencil of Grid::compute_{x,y,z}_flux (grid/grid.cpp:225-241).
// q: conserved states along the pencil.  F[i] for faces g..len-g.  eig optional.
void pencil_kt(const Geo& o, int d, int len, double q[][MAXF], double F[][MAXF], double* eig) {
  double q0[MAXG][MAXF], ql[MAXG][MAXF], qr[MAXG][MAXF];
  for (int i = 0; i < len; i++) {
    for (int l = 0; l < o.nf; l++) q0[i][l] = q[i][l];
    to_prim(q0[i]);
  }
  reconstruct(o, len, q0, ql, qr);
  for (int i = o.g; i < len - o.g + 1; i++) {
    to_con(ql[i]);
    to_con(qr[i]);
    const double eL = max_abs_eigen(o.P, ql[i], d);
    const double eR = max_abs_eigen(o.P, qr[i], d);
    const double a = rmax(eL, eR);
    double fl[MAXF], fr[MAXF];
    flux(o.P, ql[i], d, fl, o.nf);
    flux(o.P, qr[i], d, fr, o.nf);
    for (int l = 0; l < o.nf; l++) F[i][l] = ((fl[l] + fr[l]) - (qr[i][l] - ql[i][l]) * a) * 0.5;
    if (eig) {
      eig[2 * i] = eL;
      eig[2 * i + 1] = eR;
    }
  }
}

// ---- per-block grid operations (padded SoA block: U[f*G3 + at(i,j,k)]) ----------------------

// Grid::compute_{x,y,z}_flux (grid/grid.cpp:215-302) for one block.
void block_flux(const Geo& o, const double* U, double* F, int d) {
  const int lo = o.g, hi = o.G - o.g - 1;
  for (int b = lo; b <= hi; b++)
    for (int a = lo; a <= hi; a++) {
      double q[MAXG][MAXF], Fp[MAXG][MAXF];
      for (int p = 0; p < o.G; p++) {
        const long c = pencil_at(o, d, a, b, p);
        for (int l = 0; l < o.nf; l++) q[p][l] = U[l * o.G3 + c];
      }
      pencil_kt(o, d, o.G, q, Fp, nullptr);
      for (int p = o.g; p < o.G - o.g + 1; p++) {
        const long c = pencil_at(o, d, a, b, p);
        for (int l = 0; l < o.nf; l++) F[l * o.G3 + c] = Fp[p][l];
      }
    }
}

// Grid::clear_difs (grid/grid.cpp:201-213)
void block_clear(const Geo& o, double* D) {
  for (int k = o.g; k < o.G - o.g; k++)
    for (int j = o.g; j < o.G - o.g; j++)
      for (int i = o.g; i < o.G - o.g; i++)
        for (int l = 0; l < o.nf; l++) D[l * o.G3 + o.at(i, j, k)] = 0.0;
}

// Grid::sum_{x,y,z}_difs (grid/grid.cpp:304-344): D -= (F[i+1]-F[i]) * (1/h)
void block_sum_difs(const Geo& o, double* D, const double* F, double h, int d) {
  const double hinv = 1.0 / h;
  const long step = d == 0 ? 1 : (d == 1 ? o.G : (long)o.G * o.G);
  for (int k = o.g; k < o.G - o.g; k++)
    for (int j = o.g; j < o.G - o.g; j++)
      for (int i = o.g; i < o.G - o.g; i++) {
        const long c = o.at(i, j, k);
        for (int l = 0; l < o.nf; l++) {
          const double* Fl = F + l * o.G3;
          D[l * o.G3 + c] -= (Fl[c + step] - Fl[c]) * hinv;
        }
      }
}

// Grid::add_difs (grid/grid.cpp:406-420): D += source (zero for the ideal gas, state.cpp:19-22);
// U = (U + D*dt)*beta + U0*(1-beta); floor.
void block_add_difs(const Geo& o, double* U, const double* U0, double* D, double dt, double beta) {
  const double omb = 1.0 - beta;
  for (int k = o.g; k < o.G - o.g; k++)
    for (int j = o.g; j < o.G - o.g; j++)
      for (int i = o.g; i < o.G - o.g; i++) {
        const long c = o.at(i, j, k);
        double u[MAXF];
        for (int l = 0; l < o.nf; l++) {
          double dl = D[l * o.G3 + c] + 0.0;
          D[l * o.G3 + c] = dl;
          u[l] = (U[l * o.G3 + c] + dl * dt) * beta + U0[l * o.G3 + c] * omb;
        }
        floor_state(o.P, u);
        for (int l = 0; l < o.nf; l++) U[l * o.G3 + c] = u[l];
      }
}

// Grid::max_dt's scan (grid/grid.cpp:430-473): max over the three sweeps of both face states.
double block_max_eig(const Geo& o, const double* U) {
  double m = 0.0;
  const int lo = o.g, hi = o.G - o.g - 1;
  for (int b = lo; b <= hi; b++)
    for (int a = lo; a <= hi; a++) {
      for (int d = 0; d < 3; d++) {
        double q[MAXG][MAXF], Fp[MAXG][MAXF], eig[2 * MAXG];
        for (int p = 0; p < o.G; p++) {
          const long c = pencil_at(o, d, a, b, p);
          for (int l = 0; l < o.nf; l++) q[p][l] = U[l * o.G3 + c];
        }
        pencil_kt(o, d, o.G, q, Fp, eig);
        for (int p = o.g; p < o.G - o.g + 1; p++) m = rmax(m, rmax(eig[2 * p], eig[2 * p + 1]));
      }
    }
  return m;
}

double block_dt(const Geo& o, double m, double h) {
  double dtinv = m / h;
  if (dtinv == 0.0) dtinv = 1.0e-10;
  double dt = 1.0 / dtinv;
  dt *= o.P.cfl;
  return dt;
}

// Physical boundary ghost of a lone block: ghost = copy of the adjacent interior cell with the
// outflow fix, or the z-mirror when z_reflect (grid/grid_phys_bound.cpp:27-54, 60-169).
// The boundary object on the node's face f has face invert(f) (grid/grid_node.cpp:1383-1391).
void block_phys_ghost_face(const Geo& o, double* U, int nodeface) {
  const int d = nodeface >> 1;
  const bool upper = nodeface & 1;
  const int bface = nodeface ^ 1;  // invert()
  const bool reflect = o.zref && d == 2 && !upper;
  for (int b = o.g; b < o.G - o.g; b++)
    for (int a = o.g; a < o.G - o.g; a++)
      for (int t = 0; t < o.g; t++) {
        const int gp = upper ? (o.G - o.g + t) : t;  // ghost position along d
        int src;
        if (reflect)
          src = 2 * o.g - 1 - gp;  // GNX-1-(GNX-2g+gp) mirrored about the face
        else
          src = upper ? (o.G - o.g - 1) : o.g;
        const long cg = pencil_at(o, d, a, b, gp), cs_ = pencil_at(o, d, a, b, src);
        double u[MAXF];
        for (int l = 0; l < o.nf; l++) u[l] = U[l * o.G3 + cs_];
        if (reflect)
          u[SZ] = -u[SZ];  // State::reflect_on_z (euler/state.cpp:94-96)
        else
          enforce_outflow(u, bface);
        for (int l = 0; l < o.nf; l++) U[l * o.G3 + cg] = u[l];
      }
}

// ---- octree -------------------------------------------------------------------------------

struct Node {
  int level = 0, loc[3] = {0, 0, 0}, off[3] = {0, 0, 0};
  double h = 0.0;
  int parent = -1, child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  std::vector<double> U, U0, D, F;
  double FO[MAXF] = {0}, FO0[MAXF] = {0}, DFO[MAXF] = {0};
};

}  // namespace

struct orc_tree {
  orc_cfg cfg;
  Geo o;
  double h0;
  std::vector<Node> nodes;
  std::map<std::tuple<int, int, int, int>, int> by_loc;
  explicit orc_tree(const orc_cfg* c) : cfg(*c), o(c) {
    h0 = 2.0 * c->grid_dim / double(c->n);  // grid/grid.cpp:70
  }
  Node& mk() {
    nodes.emplace_back();
    Node& nd = nodes.back();
    nd.U.assign(o.nf * o.G3, 0.0);
    nd.U0.assign(o.nf * o.G3, 0.0);
    nd.D.assign(o.nf * o.G3, 0.0);
    nd.F.assign(o.nf * o.G3, 0.0);
    return nd;
  }
  int find(int level, int x, int y, int z) const {
    auto it = by_loc.find(std::make_tuple(level, x, y, z));
    return it == by_loc.end() ? -1 : it->second;
  }
  bool phys(const Node& nd, int face) const {  // oct_node/oct_node.cpp:86-110
    const int d = face >> 1;
    return (face & 1) ? nd.loc[d] == (1 << nd.level) - 1 : nd.loc[d] == 0;
  }
  int sibling(const Node& nd, int face) const {
    const int d = face >> 1;
    int l[3] = {nd.loc[0], nd.loc[1], nd.loc[2]};
    l[d] += (face & 1) ? 1 : -1;
    if (l[d] < 0 || l[d] >= (1 << nd.level)) return -1;
    return find(nd.level, l[0], l[1], l[2]);
  }
  double& Uat(Node& nd, int f, int i, int j, int k) { return nd.U[f * o.G3 + o.at(i, j, k)]; }
  double Uat(const Node& nd, int f, int i, int j, int k) const { return nd.U[f * o.G3 + o.at(i, j, k)]; }
  std::vector<int> level_order() const {
    std::vector<int> ord;
    int maxl = 0;
    for (auto& nd : nodes) maxl = nd.level > maxl ? nd.level : maxl;
    for (int l = 0; l <= maxl; l++)
      for (int i = 0; i < (int)nodes.size(); i++)
        if (nodes[i].level == l) ord.push_back(i);
    return ord;
  }
  void postorder(int i, std::vector<int>& out) const {
    for (int c = 0; c < 8; c++)
      if (nodes[i].child[c] >= 0) postorder(nodes[i].child[c], out);
    out.push_back(i);
  }
  void preorder(int i, std::vector<int>& out) const {
    out.push_back(i);
    for (int c = 0; c < 8; c++)
      if (nodes[i].child[c] >= 0) preorder(nodes[i].child[c], out);
  }
};

namespace {

// GridAMRBound::operator() (grid/grid_amr_bound.cpp:27-78) evaluated for the fine node `nd`'s
// ghost cell at node-local (i,j,k) across node face `face`.
void amr_ghost(const orc_tree* t, const Node& nd, int face, int i, int j, int k, double* u) {
  const Geo& o = t->o;
  const Node& P = t->nodes[nd.parent];
  // fine global index -> parent-local: (nd.off + idx - 2*P.off); parity picks the half
  const int idx[3] = {i, j, k};
  int i1[3], par[3];
  for (int d = 0; d < 3; d++) {
    const int f = nd.off[d] + idx[d] - 2 * P.off[d];
    par[d] = f % 2;
    i1[d] = f / 2;
  }
  const int d = face >> 1;
  double up[MAXF], um[MAXF];
  for (int l = 0; l < o.nf; l++) u[l] = t->Uat(P, l, i1[0], i1[1], i1[2]);
  int p1[3] = {i1[0], i1[1], i1[2]}, m1[3] = {i1[0], i1[1], i1[2]};
  p1[d] += 1;
  m1[d] -= 1;
  for (int l = 0; l < o.nf; l++) {
    up[l] = t->Uat(P, l, p1[0], p1[1], p1[2]);
    um[l] = t->Uat(P, l, m1[0], m1[1], m1[2]);
  }
  for (int l = 0; l < o.nf; l++) {
    double m = minmod2(up[l] - u[l], u[l] - um[l]);
    if (par[d] == 1) m = 0.0 - m;
    u[l] -= m * 0.25;
  }
}

// GridNode::enforce_boundaries for one node (grid/grid_node.cpp:59-101): 6 face slabs of g layers
// over the interior (a, b) of the face.
void node_fill_ghosts(orc_tree* t, int ni) {
  const Geo& o = t->o;
  Node& nd = t->nodes[ni];
  for (int face = 0; face < 6; face++) {
    const int d = face >> 1;
    const bool upper = face & 1;
    const int s = t->sibling(nd, face);
    if (s < 0 && t->phys(nd, face)) {
      double* U = nd.U.data();
      block_phys_ghost_face(o, U, face);
      continue;
    }
    for (int b = o.g; b < o.G - o.g; b++)
      for (int a = o.g; a < o.G - o.g; a++)
        for (int tt = 0; tt < o.g; tt++) {
          const int gp = upper ? (o.G - o.g + tt) : tt;
          // the sibling-frame index the reference asks for: lower face GNX-2g+t, upper face t-...
          const int sp = upper ? (gp - o.G + 2 * o.g) : (o.G - 2 * o.g + gp);
          const long cg = pencil_at(o, d, a, b, gp);
          double u[MAXF];
          if (s >= 0) {
            const Node& S = t->nodes[s];
            const long cs_ = pencil_at(o, d, a, b, sp);
            for (int l = 0; l < o.nf; l++) u[l] = S.U[l * o.G3 + cs_];
          } else {
            // AMR bound: the bound object is asked for index sp in the frame of the missing
            // neighbour; rel_offset folds the neighbour's offset back onto the fine node
            // (grid/grid_amr_bound.cpp:104-131), i.e. the fine-frame index is gp.
            int ijk[3];
            if (d == 0) { ijk[0] = gp; ijk[1] = a; ijk[2] = b; }
            else if (d == 1) { ijk[0] = a; ijk[1] = gp; ijk[2] = b; }
            else { ijk[0] = a; ijk[1] = b; ijk[2] = gp; }
            amr_ghost(t, nd, face, ijk[0], ijk[1], ijk[2], u);
          }
          for (int l = 0; l < o.nf; l++) nd.U[l * o.G3 + cg] = u[l];
        }
  }
}

// GridNode::adjust_{x,y,z}_flux for one node (grid/grid_node.cpp:1076-1274)
void node_adjust_flux(orc_tree* t, int ni, int d) {
  const Geo& o = t->o;
  Node& nd = t->nodes[ni];
  auto child_of = [&](int node, int cx, int cy, int cz) -> int {
    if (node < 0) return -1;
    return t->nodes[node].child[cx | (cy << 1) | (cz << 2)];
  };
  const int sL = t->sibling(nd, 2 * d), sU = t->sibling(nd, 2 * d + 1);
  for (int l = 0; l < 3; l++)
    for (int ck = 0; ck < 2; ck++)
      for (int cj = 0; cj < 2; cj++) {
        // child index with the normal bit n and the two tangential bits (cj, ck) in the
        // reference's order: x: (n, cj, ck)  y: (cj, n, ck)  z: (cj, ck, n)
        auto ci = [&](int node, int nbit) {
          int c3[3];
          if (d == 0) { c3[0] = nbit; c3[1] = cj; c3[2] = ck; }
          else if (d == 1) { c3[0] = cj; c3[1] = nbit; c3[2] = ck; }
          else { c3[0] = cj; c3[1] = ck; c3[2] = nbit; }
          return child_of(node, c3[0], c3[1], c3[2]);
        };
        int cl, cr, i0;
        if (l == 0) { cr = ci(ni, 0); cl = ci(sL, 1); i0 = o.g; }
        else if (l == 1) { cl = ci(ni, 0); cr = ci(ni, 1); i0 = o.G / 2; }
        else { cr = ci(sU, 0); cl = ci(ni, 1); i0 = o.G - o.g; }
        int ch, i;
        if (cr < 0 && cl >= 0) { ch = cl; i = o.G - o.g; }
        else if (cl < 0 && cr >= 0) { ch = cr; i = o.g; }
        else continue;
        const Node& C = t->nodes[ch];
        for (int k = o.g; k <= o.G - o.g - 1; k += 2)
          for (int j = o.g; j <= o.G - o.g - 1; j += 2) {
            const int k0 = (o.g + k) / 2 + ck * (o.G / 2 - o.g);
            const int j0 = (o.g + j) / 2 + cj * (o.G / 2 - o.g);
            for (int f = 0; f < o.nf; f++) {
              const double* CF = C.F.data() + f * o.G3;
              double v = CF[pencil_at(o, d, j, k, i)];
              v += CF[pencil_at(o, d, j + 1, k, i)];
              v += CF[pencil_at(o, d, j, k + 1, i)];
              v += CF[pencil_at(o, d, j + 1, k + 1, i)];
              v *= 0.25;
              nd.F[f * o.G3 + pencil_at(o, d, j0, k0, i0)] = v;
            }
          }
      }
}

// Grid::sum_{x,y,z}_flow_off (grid/grid.cpp:350-404)
void node_flow_off(orc_tree* t, Node& nd, int d) {
  const Geo& o = t->o;
  for (int k = o.g; k < o.G - o.g; k++)
    for (int j = o.g; j < o.G - o.g; j++) {
      for (int f = 0; f < o.nf; f++) {
        const double* F = nd.F.data() + f * o.G3;
        double tmp;
        if (d == 2 && o.zref)
          tmp = F[pencil_at(o, d, j, k, o.G - o.g)] * nd.h * nd.h;
        else
          tmp = (F[pencil_at(o, d, j, k, o.G - o.g)] - F[pencil_at(o, d, j, k, o.g)]) * nd.h * nd.h;
        nd.DFO[f] += tmp;
      }
    }
}

// GridNode::inject_from_children for one (parent, child) pair (grid/grid_node.cpp:1043-1072)
void inject_child(orc_tree* t, int pi, int c) {
  const Geo& o = t->o;
  Node& P = t->nodes[pi];
  const Node& C = t->nodes[P.child[c]];
  const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  for (int k = o.g; k <= o.G - o.g - 1; k += 2)
    for (int j = o.g; j <= o.G - o.g - 1; j += 2)
      for (int i = o.g; i < o.G - o.g; i += 2) {
        const int k0 = (o.g + k) / 2 + cz * (o.G / 2 - o.g);
        const int j0 = (o.g + j) / 2 + cy * (o.G / 2 - o.g);
        const int i0 = (o.g + i) / 2 + cx * (o.G / 2 - o.g);
        double u[MAXF];
        for (int l = 0; l < o.nf; l++) {
          double s = 0.0;
          s += t->Uat(C, l, i + 0, j + 0, k + 0);
          s += t->Uat(C, l, i + 1, j + 0, k + 0);
          s += t->Uat(C, l, i + 0, j + 1, k + 0);
          s += t->Uat(C, l, i + 1, j + 1, k + 0);
          s += t->Uat(C, l, i + 0, j + 0, k + 1);
          s += t->Uat(C, l, i + 1, j + 0, k + 1);
          s += t->Uat(C, l, i + 0, j + 1, k + 1);
          s += t->Uat(C, l, i + 1, j + 1, k + 1);
          s *= 0.125;
          u[l] = s;
        }
        floor_state(o.P, u);
        for (int l = 0; l < o.nf; l++) t->Uat(P, l, i0, j0, k0) = u[l];
      }
}

// GridNode::inject_from_parent (grid/grid_node.cpp:193-233): a new child's interior from the
// parent's cells (ghosts filled) with minmod slopes per axis, then the floor
void prolong_child(orc_tree* t, int ni) {
  const Geo& o = t->o;
  Node& C = t->nodes[ni];
  const Node& P = t->nodes[C.parent];
  int c = 0;
  while (P.child[c] != ni) c++;
  const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  for (int k = o.g; k <= o.G - o.g - 1; k += 2)
    for (int j = o.g; j <= o.G - o.g - 1; j += 2) {
      const int k0 = (o.g + k) / 2 + cz * (o.G / 2 - o.g);
      const int j0 = (o.g + j) / 2 + cy * (o.G / 2 - o.g);
      int i0 = o.g + cx * (o.G / 2 - o.g);
      for (int i = o.g; i < o.G - o.g; i += 2, i0++) {
        double v[8][MAXF];
        for (int l = 0; l < o.nf; l++) {
          const double u = t->Uat(P, l, i0, j0, k0);
          const double s1 = minmod2(t->Uat(P, l, i0 + 1, j0, k0) - u, u - t->Uat(P, l, i0 - 1, j0, k0));
          const double s2 = minmod2(t->Uat(P, l, i0, j0 + 1, k0) - u, u - t->Uat(P, l, i0, j0 - 1, k0));
          const double s3 = minmod2(t->Uat(P, l, i0, j0, k0 + 1) - u, u - t->Uat(P, l, i0, j0, k0 - 1));
          v[0][l] = u - (s1 + s2 + s3) * 0.25;
          v[1][l] = u + (s1 - s2 - s3) * 0.25;
          v[2][l] = u - (s1 - s2 + s3) * 0.25;
          v[3][l] = u + (s1 + s2 - s3) * 0.25;
          v[4][l] = u - (s1 + s2 - s3) * 0.25;
          v[5][l] = u + (s1 - s2 + s3) * 0.25;
          v[6][l] = u - (s1 - s2 - s3) * 0.25;
          v[7][l] = u + (s1 + s2 + s3) * 0.25;
        }
        for (int q = 0; q < 8; q++) {
          floor_state(o.P, v[q]);
          for (int l = 0; l < o.nf; l++) t->Uat(C, l, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2)) = v[q][l];
        }
      }
    }
}

}  // namespace

// =============================================================================================
extern "C" {

int orc_pencil_flux(const orc_cfg* c, int dir, int len, const double* q, double* F, double* eig) {
  if (len > MAXG || c->nf > MAXF || dir < 0 || dir > 2) return 1;
  Geo o(c);
  double qq[MAXG][MAXF], FF[MAXG][MAXF];
  for (int i = 0; i < len; i++)
    for (int l = 0; l < o.nf; l++) qq[i][l] = q[i * o.nf + l];
  pencil_kt(o, dir, len, qq, FF, eig);
  for (int i = o.g; i < len - o.g + 1; i++)
    for (int l = 0; l < o.nf; l++) F[i * o.nf + l] = FF[i][l];
  return 0;
}

int orc_pencil_faces(const orc_cfg* c, int len, const double* q, double* ql, double* qr) {
  if (len > MAXG || c->nf > MAXF) return 1;
  Geo o(c);
  double q0[MAXG][MAXF], L[MAXG][MAXF], R[MAXG][MAXF];
  for (int i = 0; i < len; i++) {
    for (int l = 0; l < o.nf; l++) q0[i][l] = q[i * o.nf + l];
    to_prim(q0[i]);
  }
  reconstruct(o, len, q0, L, R);
  for (int i = o.g; i < len - o.g + 1; i++) {
    to_con(L[i]);
    to_con(R[i]);
    for (int l = 0; l < o.nf; l++) {
      ql[i * o.nf + l] = L[i][l];
      qr[i * o.nf + l] = R[i][l];
    }
  }
  return 0;
}

void orc_floor(const orc_cfg* c, double* u) {
  Geo o(c);
  floor_state(o.P, u);
}

void orc_outflow(const orc_cfg* c, int face, double* u) {
  (void)c;
  enforce_outflow(u, face);
}

int orc_batch_substep(const orc_cfg* c, int nb, double* U, const double* U0, const double* h,
                      double dt, double beta, double* D_out) {
  Geo o(c);
  if (o.G > MAXG || o.nf > MAXF) return 1;
#pragma omp parallel
  {
    std::vector<double> D(o.nf * o.G3), F(o.nf * o.G3);
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < nb; b++) {
      double* Ub = U + (long)b * o.nf * o.G3;
      const double* U0b = U0 + (long)b * o.nf * o.G3;
      block_clear(o, D.data());
      for (int d = 0; d < 3; d++) {
        block_flux(o, Ub, F.data(), d);
        block_sum_difs(o, D.data(), F.data(), h[b], d);
      }
      block_add_difs(o, Ub, U0b, D.data(), dt, beta);
      if (D_out) std::memcpy(D_out + (long)b * o.nf * o.G3, D.data(), sizeof(double) * o.nf * o.G3);
    }
  }
  return 0;
}

int orc_batch_max_eig(const orc_cfg* c, int nb, const double* U, double* eig) {
  Geo o(c);
  if (o.G > MAXG || o.nf > MAXF) return 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < nb; b++) eig[b] = block_max_eig(o, U + (long)b * o.nf * o.G3);
  return 0;
}

double orc_block_dt(const orc_cfg* c, double max_eig, double h) {
  Geo o(c);
  return block_dt(o, max_eig, h);
}

int orc_batch_outflow_ghosts(const orc_cfg* c, int nb, double* U) {
  Geo o(c);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; b++)
    for (int f = 0; f < 6; f++) block_phys_ghost_face(o, U + (long)b * o.nf * o.G3, f);
  return 0;
}

int orc_batch_rk_steps(const orc_cfg* c, int nb, double* U, double* U0, const double* h, double dt,
                       int rk, int nsteps) {
  Geo o(c);
  const long tot = (long)nb * o.nf * o.G3;
  static const double betas[3][3] = {{1.0, 0, 0}, {1.0, 0.5, 0}, {1.0, 0.25, 2.0 / 3.0}};
  for (int s = 0; s < nsteps; s++) {
    std::memcpy(U0, U, sizeof(double) * tot);
    for (int st = 0; st < rk; st++) {
      orc_batch_outflow_ghosts(c, nb, U);
      int rc = orc_batch_substep(c, nb, U, U0, h, dt, betas[rk - 1][st], nullptr);
      if (rc) return rc;
    }
  }
  orc_batch_outflow_ghosts(c, nb, U);
  return 0;
}

// ---- reference-shaped restatement (a CPU baseline that keeps the reference's structure) ------
// The same arithmetic as above, arranged as the reference arranges it: every node holds
// Array3d<State> U, U0 and Array3d<Vector> D, F as arrays of structures whose elements carry a
// virtual-table pointer before their fields (vector.h:40-45: sizeof(State) = 56 B for EULER), and
// a substep is a sequence of passes over the nodes, each pass one node at a time with an OpenMP
// parallel-for over that node's pencils or planes (grid/grid.cpp:177-420, the recursions of
// grid/grid_node.cpp): enforce_boundaries, clear_difs, compute_{x,y,z}_flux + sum_{x,y,z}_difs,
// add_difs; max_dt scans each node with a parallel-for and a critical max (grid/grid.cpp:422-483).
namespace {

struct AosVec {
  const void* vptr;  // the reference's Vector has a virtual method table
  double v[MAXF];
};
static_assert(MAXF >= 6, "");

struct RefShape {
  Geo o;
  int nb;
  std::vector<std::vector<AosVec>> U, U0, D, F;
  std::vector<double> h;
  explicit RefShape(const orc_cfg* c) : o(c), nb(0) {}
};

// Grid::compute_{x,y,z}_flux (grid/grid.cpp:215-302): parallel over the planes of pencils
void rs_flux(RefShape& R, int b, int d) {
  const Geo& o = R.o;
  const std::vector<AosVec>& U = R.U[b];
  std::vector<AosVec>& F = R.F[b];
  const int lo = o.g, hi = o.G - o.g - 1;
#pragma omp parallel for schedule(static)
  for (int bb = lo; bb <= hi; bb++)
    for (int a = lo; a <= hi; a++) {
      double q[MAXG][MAXF], Fp[MAXG][MAXF];
      for (int p = 0; p < o.G; p++) {
        const AosVec& u = U[pencil_at(o, d, a, bb, p)];
        for (int l = 0; l < o.nf; l++) q[p][l] = u.v[l];
      }
      pencil_kt(o, d, o.G, q, Fp, nullptr);
      for (int p = o.g; p < o.G - o.g + 1; p++) {
        AosVec& f = F[pencil_at(o, d, a, bb, p)];
        for (int l = 0; l < o.nf; l++) f.v[l] = Fp[p][l];
      }
    }
}

// Grid::sum_{x,y,z}_difs (grid/grid.cpp:304-344)
void rs_sum_difs(RefShape& R, int b, int d) {
  const Geo& o = R.o;
  std::vector<AosVec>& D = R.D[b];
  const std::vector<AosVec>& F = R.F[b];
  const double hinv = 1.0 / R.h[b];
  const long step = d == 0 ? 1 : (d == 1 ? o.G : (long)o.G * o.G);
#pragma omp parallel for schedule(static)
  for (int k = o.g; k < o.G - o.g; k++)
    for (int j = o.g; j < o.G - o.g; j++)
      for (int i = o.g; i < o.G - o.g; i++) {
        const long c = o.at(i, j, k);
        for (int l = 0; l < o.nf; l++) D[c].v[l] -= (F[c + step].v[l] - F[c].v[l]) * hinv;
      }
}

// Grid::clear_difs (grid/grid.cpp:201-213)
void rs_clear(RefShape& R, int b) {
  const Geo& o = R.o;
  std::vector<AosVec>& D = R.D[b];
#pragma omp parallel for schedule(static)
  for (int k = o.g; k < o.G - o.g; k++)
    for (int j = o.g; j < o.G - o.g; j++)
      for (int i = o.g; i < o.G - o.g; i++)
        for (int l = 0; l < o.nf; l++) D[o.at(i, j, k)].v[l] = 0.0;
}

// Grid::add_difs (grid/grid.cpp:406-420)
void rs_add_difs(RefShape& R, int b, double dt, double beta) {
  const Geo& o = R.o;
  std::vector<AosVec>& U = R.U[b];
  const std::vector<AosVec>& U0 = R.U0[b];
  std::vector<AosVec>& D = R.D[b];
  const double omb = 1.0 - beta;
#pragma omp parallel for schedule(static)
  for (int k = o.g; k < o.G - o.g; k++)
    for (int j = o.g; j < o.G - o.g; j++)
      for (int i = o.g; i < o.G - o.g; i++) {
        const long c = o.at(i, j, k);
        double u[MAXF];
        for (int l = 0; l < o.nf; l++) {
          const double dl = D[c].v[l] + 0.0;
          D[c].v[l] = dl;
          u[l] = (U[c].v[l] + dl * dt) * beta + U0[c].v[l] * omb;
        }
        floor_state(o.P, u);
        for (int l = 0; l < o.nf; l++) U[c].v[l] = u[l];
      }
}

// enforce_boundaries of a lone node (grid/grid_phys_bound.cpp:27-54): per face, parallel over the
// face's cells; the arithmetic of block_phys_ghost_face on one AoS node
void rs_ghosts(RefShape& R, int b) {
  const Geo& o = R.o;
  std::vector<AosVec>& U = R.U[b];
  for (int f = 0; f < 6; f++) {
    const int d = f >> 1;
    const bool upper = f & 1;
    const int bface = f ^ 1;
    const bool reflect = o.zref && d == 2 && !upper;
#pragma omp parallel for schedule(static)
    for (int bb = o.g; bb < o.G - o.g; bb++)
      for (int a = o.g; a < o.G - o.g; a++)
        for (int t = 0; t < o.g; t++) {
          const int gp = upper ? (o.G - o.g + t) : t;
          const int src = reflect ? 2 * o.g - 1 - gp : (upper ? (o.G - o.g - 1) : o.g);
          double u[MAXF];
          const AosVec& s = U[pencil_at(o, d, a, bb, src)];
          for (int l = 0; l < o.nf; l++) u[l] = s.v[l];
          if (reflect)
            u[SZ] = -u[SZ];
          else
            enforce_outflow(u, bface);
          AosVec& g = U[pencil_at(o, d, a, bb, gp)];
          for (int l = 0; l < o.nf; l++) g.v[l] = u[l];
        }
  }
}

}  // namespace

void* orc_refshape_create(const orc_cfg* c, int nb, const double* U, const double* h) {
  RefShape* R = new RefShape(c);
  const Geo& o = R->o;
  if (o.G > MAXG || o.nf > MAXF) {
    delete R;
    return nullptr;
  }
  R->nb = nb;
  R->h.assign(h, h + nb);
  static const char vtable_stand_in = 0;
  const AosVec zero{&vtable_stand_in, {}};
  for (auto* A : {&R->U, &R->U0, &R->D, &R->F}) A->assign(nb, std::vector<AosVec>(o.G3, zero));
  for (int b = 0; b < nb; b++)
    for (long e = 0; e < o.G3; e++)
      for (int l = 0; l < o.nf; l++) R->U[b][e].v[l] = U[((long)b * o.nf + l) * o.G3 + e];
  return R;
}

void orc_refshape_destroy(void* r) { delete static_cast<RefShape*>(r); }

int orc_refshape_download(const void* r, double* U) {
  const RefShape* R = static_cast<const RefShape*>(r);
  const Geo& o = R->o;
  for (int b = 0; b < R->nb; b++)
    for (long e = 0; e < o.G3; e++)
      for (int l = 0; l < o.nf; l++) U[((long)b * o.nf + l) * o.G3 + e] = R->U[b][e].v[l];
  return 0;
}

double orc_refshape_max_eig(void* r) {
  // Grid::max_dt (grid/grid.cpp:422-483), node by node; returns max over nodes of max_eig / h
  RefShape* R = static_cast<RefShape*>(r);
  const Geo& o = R->o;
  double m = 0.0;
  for (int b = 0; b < R->nb; b++) {
    double mb = 0.0;
    const std::vector<AosVec>& U = R->U[b];
    const int lo = o.g, hi = o.G - o.g - 1;
#pragma omp parallel for schedule(static)
    for (int bb = lo; bb <= hi; bb++)
      for (int a = lo; a <= hi; a++) {
        double mine = 0.0;
        for (int d = 0; d < 3; d++) {
          double q[MAXG][MAXF], Fp[MAXG][MAXF], eig[2 * MAXG];
          for (int p = 0; p < o.G; p++) {
            const AosVec& u = U[pencil_at(o, d, a, bb, p)];
            for (int l = 0; l < o.nf; l++) q[p][l] = u.v[l];
          }
          pencil_kt(o, d, o.G, q, Fp, eig);
          for (int p = o.g; p < o.G - o.g + 1; p++) mine = rmax(mine, rmax(eig[2 * p], eig[2 * p + 1]));
        }
#pragma omp critical
        mb = rmax(mb, mine);
      }
    m = rmax(m, mb / R->h[b]);
  }
  return m;
}

int orc_refshape_rk_step(void* r, double dt, int rk) {
  // HydroDriver::step (euler/hydro_driver.cpp:171-194) on lone outflow nodes: store, then per
  // substep enforce_boundaries and the sub_step passes (euler/hydro_driver.cpp:200-214)
  RefShape* R = static_cast<RefShape*>(r);
  const Geo& o = R->o;
  static const double betas[3][3] = {{1.0, 0, 0}, {1.0, 0.5, 0}, {1.0, 0.25, 2.0 / 3.0}};
  for (int b = 0; b < R->nb; b++) {
    std::vector<AosVec>& U = R->U[b];
    std::vector<AosVec>& U0 = R->U0[b];
#pragma omp parallel for schedule(static)
    for (long e = 0; e < o.G3; e++) U0[e] = U[e];
  }
  for (int st = 0; st < rk; st++) {
    const double beta = betas[rk - 1][st];
    for (int b = 0; b < R->nb; b++) rs_ghosts(*R, b);
    for (int b = 0; b < R->nb; b++) rs_clear(*R, b);
    for (int d = 0; d < 3; d++) {
      for (int b = 0; b < R->nb; b++) rs_flux(*R, b, d);
      for (int b = 0; b < R->nb; b++) rs_sum_difs(*R, b, d);
    }
    for (int b = 0; b < R->nb; b++) rs_add_difs(*R, b, dt, beta);
  }
  for (int b = 0; b < R->nb; b++) rs_ghosts(*R, b);
  return 0;
}

int orc_tree_prolong(orc_tree* t, int node) {
  if (node <= 0 || node >= (int)t->nodes.size()) return 1;
  prolong_child(t, node);
  return 0;
}

orc_tree* orc_tree_create(const orc_cfg* c) {
  Geo o(c);
  if (o.G > MAXG || o.nf > MAXF || (c->n % 2)) return nullptr;
  orc_tree* t = new orc_tree(c);
  Node& r = t->mk();
  r.h = t->h0;
  t->by_loc[std::make_tuple(0, 0, 0, 0)] = 0;
  return t;
}

void orc_tree_destroy(orc_tree* t) { delete t; }

int orc_tree_add_child(orc_tree* t, int parent, int child) {
  if (parent < 0 || parent >= (int)t->nodes.size() || child < 0 || child > 7) return -1;
  if (t->nodes[parent].child[child] >= 0) return -1;
  const int idx = (int)t->nodes.size();
  Node& nd = t->mk();
  Node& P = t->nodes[parent];
  const int cv[3] = {child & 1, (child >> 1) & 1, (child >> 2) & 1};
  nd.level = P.level + 1;
  for (int d = 0; d < 3; d++) {
    nd.loc[d] = 2 * P.loc[d] + cv[d];                            // oct_node/oct_node.cpp:177
    nd.off[d] = 2 * P.off[d] + t->o.g + cv[d] * (t->o.G - 2 * t->o.g);  // grid/grid_node.cpp:276
  }
  nd.h = P.h * 0.5;  // grid/grid_node.cpp:279
  nd.parent = parent;
  P.child[child] = idx;
  t->by_loc[std::make_tuple(nd.level, nd.loc[0], nd.loc[1], nd.loc[2])] = idx;
  return idx;
}

int orc_tree_nnodes(const orc_tree* t) { return (int)t->nodes.size(); }

int orc_tree_node_info(const orc_tree* t, int node, int* level, int* loc, int* off, double* h) {
  if (node < 0 || node >= (int)t->nodes.size()) return 1;
  const Node& nd = t->nodes[node];
  if (level) *level = nd.level;
  for (int d = 0; d < 3; d++) {
    if (loc) loc[d] = nd.loc[d];
    if (off) off[d] = nd.off[d];
  }
  if (h) *h = nd.h;
  return 0;
}

double* orc_tree_node_U(orc_tree* t, int node) { return t->nodes[node].U.data(); }
const double* orc_tree_node_F(const orc_tree* t, int node) { return t->nodes[node].F.data(); }

int orc_tree_enforce_boundaries(orc_tree* t) {
  // Level order: a node's fill reads same-level interiors, its own interior and its parent's
  // cells including the parent's ghosts, which the reference's pre-order recursion has filled.
  for (int ni : t->level_order()) node_fill_ghosts(t, ni);
  return 0;
}

int orc_tree_inject_from_children(orc_tree* t) {
  std::vector<int> po;
  t->postorder(0, po);
  for (int ni : po)
    for (int c = 0; c < 8; c++)
      if (t->nodes[ni].child[c] >= 0) inject_child(t, ni, c);
  return 0;
}

int orc_tree_store(orc_tree* t) {
  for (auto& nd : t->nodes) {
    for (int f = 0; f < t->o.nf; f++) nd.FO0[f] = nd.FO[f];
    nd.U0 = nd.U;
  }
  return 0;
}

static double tree_max_dt_rec(const orc_tree* t, int ni) {
  const Node& nd = t->nodes[ni];
  double dtinv = 0.0;
  for (int c = 0; c < 8; c++)
    if (nd.child[c] >= 0) {
      const double cdtinv = 1.0 / tree_max_dt_rec(t, nd.child[c]);
      dtinv = rmax(cdtinv, dtinv);
    }
  const double own = block_dt(t->o, block_max_eig(t->o, nd.U.data()), nd.h);
  dtinv = rmax(1.0 / own, dtinv);
  return 1.0 / dtinv;
}

double orc_tree_max_dt(const orc_tree* t) { return tree_max_dt_rec(t, 0); }

int orc_tree_substep(orc_tree* t, double dt, double beta) {
  const Geo& o = t->o;
  orc_tree_enforce_boundaries(t);
  for (auto& nd : t->nodes) {
    block_clear(o, nd.D.data());
    for (int f = 0; f < o.nf; f++) nd.DFO[f] = 0.0;
  }
  std::vector<int> po;
  t->postorder(0, po);
  for (int d = 0; d < 3; d++) {
    const int nn = (int)t->nodes.size();
#pragma omp parallel for schedule(dynamic, 1)
    for (int ni = 0; ni < nn; ni++) block_flux(o, t->nodes[ni].U.data(), t->nodes[ni].F.data(), d);
    for (int ni : po) node_adjust_flux(t, ni, d);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ni = 0; ni < nn; ni++)
      block_sum_difs(o, t->nodes[ni].D.data(), t->nodes[ni].F.data(), t->nodes[ni].h, d);
    node_flow_off(t, t->nodes[0], d);
  }
  const int nn = (int)t->nodes.size();
#pragma omp parallel for schedule(dynamic, 1)
  for (int ni = 0; ni < nn; ni++) {
    Node& nd = t->nodes[ni];
    block_add_difs(o, nd.U.data(), nd.U0.data(), nd.D.data(), dt, beta);
  }
  {
    Node& r = t->nodes[0];
    for (int f = 0; f < o.nf; f++) r.FO[f] = (r.FO[f] + r.DFO[f] * dt) * beta + r.FO0[f] * (1.0 - beta);
  }
  orc_tree_inject_from_children(t);
  return 0;
}

int orc_tree_state_sum(const orc_tree* t, double* sum) {
  const Geo& o = t->o;
  for (int f = 0; f < o.nf; f++) sum[f] = 0.0;
  std::vector<int> pre;
  t->preorder(0, pre);
  for (int ni : pre) {
    const Node& nd = t->nodes[ni];
    const double h3 = std::pow(nd.h, 3);
    for (int f = 0; f < o.nf; f++) {
      double a0 = 0.0;
      for (int k = o.g; k < o.G - o.g; k++)
        for (int j = o.g; j < o.G - o.g; j++)
          for (int i = o.g; i < o.G - o.g; i++) {
            // zone_is_refined: the child octant holding (i,j,k) exists
            const int c = (2 * i / o.G) | ((2 * j / o.G) << 1) | ((2 * k / o.G) << 2);
            if (nd.child[c] >= 0) continue;
            a0 += t->Uat(nd, f, i, j, k) * h3;
          }
      sum[f] += a0;
    }
  }
  return 0;
}

int orc_tree_flow_off(const orc_tree* t, double* fo) {
  for (int f = 0; f < t->o.nf; f++) fo[f] = t->nodes[0].FO[f];
  return 0;
}

}  // extern "C"
