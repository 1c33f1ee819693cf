// ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py CPU baseline).
//
// FP64 restatement of the reference's self-gravity solver on the octree, operation by operation:
//   poisson/poisson.cpp:14-55      vcycle_coarse_correction
//   poisson/poisson.cpp:57-84      vcycle0: phi boundaries, down sweep, up sweep
//   poisson/poisson.cpp:88-130     vcycle_relax: per-node Jacobi until the residual halves (<= 16)
//   poisson/poisson.cpp:223-247    compute_forces
//   poisson/poisson.cpp:249-472    adjust_{x,y,z}_dphi: coarse-fine flux matching in the residual
//   poisson/poisson.cpp:474-635    adjust_{x,y,z}_force: coarse face force = mean of 4 fine ones
//   poisson/poisson.cpp:637-668    solution_error
//   poisson/poisson.cpp:707-743    phi_from_children (64-point restriction)
//   poisson/poisson.cpp:745-902    enforce_phi_boundaries (faces, edges, vertices)
//   poisson/poisson.cpp:904-942    enforce_dphi_boundaries
//   poisson/poisson.cpp:944-1074   init_vcycle, poisson_zone_is_refined, vcycle_down, retire, vcycle
//   poisson/poisson_amr_bound.cpp  27-point coarse->fine phi ghosts, dphi ghosts from the parent
//   poisson/poisson_phys_bound.cpp boundary phi tables, reflection on z
//   binary/binary.cpp:938-1019     physical-boundary phi from per-octant monopoles
//   binary/binary_driver.cpp:294-344  solve(): the outer vcycle loop
// The source S is given per node (the reference sets S = 4 pi G rho, binary/state_ztwd.cpp:219-222).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <tuple>
#include <vector>

#include "oracle.h"

namespace {

constexpr double RELAX_RESID = 0.5, VDOWN_RESID = 1.0, VUP_RESID = 0.05;
enum { XL = 0, XU, YL, YU, ZL, ZU };
enum { K_SIB = 0, K_PHYS, K_AMR };

struct PNode {
  int level = 0, loc[3] = {0, 0, 0}, off[3] = {0, 0, 0};
  double h = 0.0;
  int parent = -1, cidx = -1, child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  std::vector<double> phi, phi0, dphi, dphi1, S, fx, fy, fz;
  double s0[8] = {0}, sx[8] = {0}, sy[8] = {0}, sz[8] = {0};
  bool all_refined = true;
  std::vector<double> tab[6];  // boundary phi per physical face, [a * P + b] over the other two axes
};

struct Face {
  int kind, node;
  bool reflect;
};

}  // namespace

struct orc_poisson {
  int n, P, bw, z_reflect, maxlevel = 0;
  double grid_dim, h0;
  std::vector<PNode> nd;
  std::map<std::tuple<int, int, int, int>, int> by_loc;

  long at(int i, int j, int k) const { return ((long)k * P + j) * P + i; }
  double& F(std::vector<double>& a, int i, int j, int k) { return a[at(i, j, k)]; }
  double F(const std::vector<double>& a, int i, int j, int k) const { return a[at(i, j, k)]; }

  bool phys(int b, int f) const {
    const int d = f >> 1;
    return (f & 1) ? nd[b].loc[d] == (1 << nd[b].level) - 1 : nd[b].loc[d] == 0;
  }
  int sibling(int b, int f) const {
    const int d = f >> 1;
    int l[3] = {nd[b].loc[0], nd[b].loc[1], nd[b].loc[2]};
    l[d] += (f & 1) ? 1 : -1;
    if (l[d] < 0 || l[d] >= (1 << nd[b].level)) return -1;
    auto it = by_loc.find(std::make_tuple(nd[b].level, l[0], l[1], l[2]));
    return it == by_loc.end() ? -1 : it->second;
  }
  // what sits across face f of node b: a node, a physical boundary, or an AMR boundary
  Face face(int b, int f) const {
    const int s = sibling(b, f);
    if (s >= 0) return {K_SIB, s, false};
    // PoissonPhysBound::set_sibling receives the inverted face: the node's ZL face reflects
    if (phys(b, f)) return {K_PHYS, b, z_reflect && f == ZL};
    return {K_AMR, b, false};
  }
  bool refined(int b, int i, int j, int k) const {  // poisson_zone_is_refined
    const int c = ((2 * i) / P) | (((2 * j) / P) << 1) | (((2 * k) / P) << 2);
    return nd[b].child[c] >= 0;
  }

  // ---- cell positions (Grid::xf / xc, grid/grid.cpp:498-526; Poisson::pxc = xc(i + BW - 1)) --
  double pc(int b, int d, int i) const {
    const int ii = i + bw - 1;
    const PNode& q = nd[b];
    double f;
    if (d == 2 && z_reflect)
      f = double(q.off[2] + ii) * q.h - double(bw) * h0 - 0.0;
    else if (d == 2)
      f = double(q.off[2] + ii) * q.h - grid_dim - double(bw) * h0;
    else
      f = double(q.off[d] + ii) * q.h - grid_dim - double(bw) * h0 - 0.0;
    return f + 0.5 * q.h;
  }

  // ---- PoissonAMRBound (poisson/poisson_amr_bound.cpp:33-127) ------------------------------
  // (i, j, k) are the indices the node passes to its neighbour object: the sibling frame.
  void amr_q(int b, int f, int i, int j, int k, int q[3]) const {
    const int d = f >> 1;
    const int v = (f & 1) ? 1 : -1;  // set_sibling sees the inverted face
    const PNode& c = nd[b];
    const PNode& p = nd[c.parent];
    const int idx[3] = {i, j, k};
    for (int a = 0; a < 3; a++) q[a] = c.off[a] + (a == d ? v * (P - 2) : 0) - 2 * p.off[a] - (bw - 1) + idx[a];
  }
  double amr_get_phi(int b, int f, int i, int j, int k) const {
    static const double w[3] = {-0.09375, 0.9375, 0.15625};  // quadratic Lagrange weights at 1/4
    int q[3];
    amr_q(b, f, i, j, k, q);
    int i1[3], i0[3];
    for (int a = 0; a < 3; a++) {
      i1[a] = q[a] % 2;
      i0[a] = q[a] / 2;
    }
    int par = nd[b].parent;
    for (int a = 0; a < 3; a++) {
      if (i0[a] > P - 2) {
        const Face s = face(par, 2 * a + 1);
        if (s.kind != K_SIB) return 0.0;
        par = s.node;
        i0[a] -= P - 2;
      } else if (i0[a] < 1) {
        const Face s = face(par, 2 * a);
        if (s.kind != K_SIB) return 0.0;
        par = s.node;
        i0[a] += P - 2;
      }
    }
    double sum = 0.0;
    for (int l = 0; l < 3; l++)
      for (int m = 0; m < 3; m++)
        for (int o = 0; o < 3; o++) {
          const int i2 = i0[0] + (l - 1) * (2 * i1[0] - 1);
          const int j2 = i0[1] + (m - 1) * (2 * i1[1] - 1);
          const int k2 = i0[2] + (o - 1) * (2 * i1[2] - 1);
          sum += (w[l] * w[m] * w[o]) * F(nd[par].phi, i2, j2, k2);
        }
    return sum;
  }
  double amr_get_dphi(int b, int f, int i, int j, int k) const {
    int q[3];
    amr_q(b, f, i, j, k, q);
    return F(nd[nd[b].parent].dphi, q[0] / 2, q[1] / 2, q[2] / 2);
  }
  // PoissonPhysBound::get_phi / get_dphi with reflection: the mirror of the node's own cells
  static int mirror(int P, int t) {
    if (t == 1) return P - 2;
    if (t == P - 2) return 1;
    fprintf(stderr, "oracle poisson: reflected index %d\n", t);  // the reference aborts here too
    std::abort();
  }
  double get(int b, int f, int i, int j, int k, bool dphi_field) const {
    const Face s = face(b, f);
    if (s.kind == K_SIB) return F(dphi_field ? nd[s.node].dphi : nd[s.node].phi, i, j, k);
    if (s.kind == K_AMR) return dphi_field ? amr_get_dphi(b, f, i, j, k) : amr_get_phi(b, f, i, j, k);
    const int d = f >> 1;
    if (!s.reflect) {
      if (dphi_field) return 0.0;
      const int a = d == 0 ? j : i, c = d == 2 ? j : k;
      return nd[b].tab[f][a * P + c];
    }
    int idx[3] = {i, j, k};
    idx[d] = mirror(P, idx[d]);
    return F(dphi_field ? nd[b].dphi : nd[b].phi, idx[0], idx[1], idx[2]);
  }
  double gp(int b, int f, int i, int j, int k) const { return get(b, f, i, j, k, false); }

  // ---- enforce_phi_boundaries for the nodes of one level --------------------------------
  void phi_faces(int b) {
    std::vector<double>& p = nd[b].phi;
    for (int k = 1; k < P - 1; k++)
      for (int j = 1; j < P - 1; j++) {
        F(p, 0, j, k) = gp(b, XL, P - 2, j, k);
        F(p, P - 1, j, k) = gp(b, XU, 1, j, k);
        F(p, j, 0, k) = gp(b, YL, j, P - 2, k);
        F(p, j, P - 1, k) = gp(b, YU, j, 1, k);
        F(p, j, k, 0) = gp(b, ZL, j, k, P - 2);
        F(p, j, k, P - 1) = gp(b, ZU, j, k, 1);
      }
  }
  void phi_edges(int b) {
    std::vector<double>& p = nd[b].phi;
    const int Q = P - 1, T = P - 2;
    for (int k = 1; k < P - 1; k++) {
      if (phys(b, XL) || phys(b, XU)) {
        F(p, 0, 0, k) = gp(b, XL, T, 0, k);
        F(p, 0, Q, k) = gp(b, XL, T, Q, k);
        F(p, 0, k, 0) = gp(b, XL, T, k, 0);
        F(p, 0, k, Q) = gp(b, XL, T, k, Q);
        F(p, Q, 0, k) = gp(b, XU, 1, 0, k);
        F(p, Q, Q, k) = gp(b, XU, 1, Q, k);
        F(p, Q, k, 0) = gp(b, XU, 1, k, 0);
        F(p, Q, k, Q) = gp(b, XU, 1, k, Q);
      } else {
        F(p, 0, 0, k) = gp(b, YL, 0, T, k);
        F(p, Q, 0, k) = gp(b, YL, Q, T, k);
        F(p, 0, Q, k) = gp(b, YU, 0, 1, k);
        F(p, Q, Q, k) = gp(b, YU, Q, 1, k);
        F(p, 0, k, 0) = gp(b, ZL, 0, k, T);
        F(p, Q, k, 0) = gp(b, ZL, Q, k, T);
        F(p, 0, k, Q) = gp(b, ZU, 0, k, 1);
        F(p, Q, k, Q) = gp(b, ZU, Q, k, 1);
      }
      if (phys(b, YL) || phys(b, YU)) {
        F(p, k, 0, 0) = gp(b, YL, k, T, 0);
        F(p, k, 0, Q) = gp(b, YL, k, T, Q);
        F(p, k, Q, 0) = gp(b, YU, k, 1, 0);
        F(p, k, Q, Q) = gp(b, YU, k, 1, Q);
      } else {
        F(p, k, 0, 0) = gp(b, ZL, k, 0, T);
        F(p, k, Q, 0) = gp(b, ZL, k, Q, T);
        F(p, k, 0, Q) = gp(b, ZU, k, 0, 1);
        F(p, k, Q, Q) = gp(b, ZU, k, Q, 1);
      }
    }
  }
  void phi_vertices(int b) {
    std::vector<double>& p = nd[b].phi;
    const int Q = P - 1, T = P - 2;
    if (phys(b, XL) || phys(b, XU)) {
      F(p, 0, 0, 0) = gp(b, XL, T, 0, 0);
      F(p, 0, Q, 0) = gp(b, XL, T, Q, 0);
      F(p, 0, 0, Q) = gp(b, XL, T, 0, Q);
      F(p, 0, Q, Q) = gp(b, XL, T, Q, Q);
      F(p, Q, 0, 0) = gp(b, XU, 1, 0, 0);
      F(p, Q, Q, 0) = gp(b, XU, 1, Q, 0);
      F(p, Q, 0, Q) = gp(b, XU, 1, 0, Q);
      F(p, Q, Q, Q) = gp(b, XU, 1, Q, Q);
    } else if (phys(b, YL) || phys(b, YU)) {
      F(p, 0, 0, 0) = gp(b, YL, 0, T, 0);
      F(p, Q, 0, 0) = gp(b, YL, Q, T, 0);
      F(p, 0, 0, Q) = gp(b, YL, 0, T, Q);
      F(p, Q, 0, Q) = gp(b, YL, Q, T, Q);
      F(p, 0, Q, 0) = gp(b, YU, 0, 1, 0);
      F(p, Q, Q, 0) = gp(b, YU, Q, 1, 0);
      F(p, 0, Q, Q) = gp(b, YU, 0, 1, Q);
      F(p, Q, Q, Q) = gp(b, YU, Q, 1, Q);
    } else {
      F(p, 0, 0, 0) = gp(b, ZL, 0, 0, T);
      F(p, 0, Q, 0) = gp(b, ZL, 0, Q, T);
      F(p, Q, 0, 0) = gp(b, ZL, Q, 0, T);
      F(p, Q, Q, 0) = gp(b, ZL, Q, Q, T);
      F(p, 0, 0, Q) = gp(b, ZU, 0, 0, 1);
      F(p, 0, Q, Q) = gp(b, ZU, 0, Q, 1);
      F(p, Q, 0, Q) = gp(b, ZU, Q, 0, 1);
      F(p, Q, Q, Q) = gp(b, ZU, Q, Q, 1);
    }
  }
  std::vector<int> at_level(int l) const {  // pre-order: the reference's recursion order
    std::vector<int> out;
    pre(0, l, out);
    return out;
  }
  void pre(int b, int l, std::vector<int>& out) const {
    if (nd[b].level == l) {
      out.push_back(b);
      return;
    }
    for (int c = 0; c < 8; c++)
      if (nd[b].child[c] >= 0) pre(nd[b].child[c], l, out);
  }
  void enforce_phi(int l) {
    const std::vector<int> lv = at_level(l);
    for (int b : lv) phi_faces(b);
    for (int b : lv) phi_edges(b);
    for (int b : lv) phi_vertices(b);
  }
  void enforce_dphi(int l) {
    for (int b : at_level(l)) {
      std::vector<double>& p = nd[b].dphi;
      for (int k = 1; k < P - 1; k++)
        for (int j = 1; j < P - 1; j++) {
          F(p, 0, j, k) = get(b, XL, P - 2, j, k, true);
          F(p, P - 1, j, k) = get(b, XU, 1, j, k, true);
          F(p, j, 0, k) = get(b, YL, j, P - 2, k, true);
          F(p, j, P - 1, k) = get(b, YU, j, 1, k, true);
          F(p, j, k, 0) = get(b, ZL, j, k, P - 2, true);
          F(p, j, k, P - 1) = get(b, ZU, j, k, 1, true);
        }
    }
  }

  // ---- array helpers (array3d.h:139-161) --------------------------------------------------
  double div(const std::vector<double>& a, int i, int j, int k) const {
    return (F(a, i + 1, j, k) + F(a, i - 1, j, k) + F(a, i, j + 1, k) + F(a, i, j - 1, k) + F(a, i, j, k + 1) +
            F(a, i, j, k - 1)) -
           (F(a, i, j, k) * 6.0);
  }
  double oct_avg(const std::vector<double>& a, int i, int j, int k) const {
    return (F(a, i, j, k) + F(a, i + 1, j, k) + F(a, i, j + 1, k) + F(a, i + 1, j + 1, k) + F(a, i, j, k + 1) +
            F(a, i + 1, j, k + 1) + F(a, i, j + 1, k + 1) + F(a, i + 1, j + 1, k + 1)) *
           0.125;
  }

  // ---- coarse-fine matching: the two children across the face at position l of axis d ------
  // returns the refined side's child and its face index, or -1 (grid_node-style quadrant walk)
  int cf_child(int b, int d, int l, int cj, int ck, int* fi, int* i0) const {
    auto kid = [&](int node, int x) {
      if (node < 0) return -1;
      int c3[3];
      c3[d] = x;
      c3[d == 0 ? 1 : 0] = cj;
      c3[d == 2 ? 1 : 2] = ck;
      return nd[node].child[c3[0] | (c3[1] << 1) | (c3[2] << 2)];
    };
    auto sib = [&](int f) {
      const Face s = face(b, f);
      return s.kind == K_SIB ? s.node : -1;
    };
    int cl, cr;
    if (l == 0) {
      cr = kid(b, 0);
      cl = kid(sib(2 * d), 1);
      *i0 = 1;
    } else if (l == 1) {
      cl = kid(b, 0);
      cr = kid(b, 1);
      *i0 = P / 2;
    } else {
      cr = kid(sib(2 * d + 1), 0);
      cl = kid(b, 1);
      *i0 = P - 1;
    }
    if (cr < 0 && cl >= 0) {
      *fi = P - 1;
      return cl;
    }
    if (cl < 0 && cr >= 0) {
      *fi = 1;
      return cr;
    }
    return -1;
  }
  // index (along d = i, across = j, k) -> (x, y, z) as adjust_{x,y,z} index their arrays
  void place(int d, int a, int j, int k, int* x, int* y, int* z) const {
    if (d == 0) {
      *x = a, *y = j, *z = k;
    } else if (d == 1) {
      *x = j, *y = a, *z = k;
    } else {
      *x = j, *y = k, *z = a;
    }
  }
  void adjust_dphi(int b, int d) {
    PNode& me = nd[b];
    for (int l = 0; l < 3; l++)
      for (int ck = 0; ck < 2; ck++)
        for (int cj = 0; cj < 2; cj++) {
          int i, i0;
          const int c = cf_child(b, d, l, cj, ck, &i, &i0);
          if (c < 0) continue;
          const PNode& ch = nd[c];
          for (int k = 1; k < P - 2; k += 2)
            for (int j = 1; j < P - 2; j += 2) {
              const int k0 = (1 + k) / 2 + ck * (P / 2 - 1);
              const int j0 = (1 + j) / 2 + cj * (P / 2 - 1);
              int x, y, z, xm, ym, zm;
              double v;
              place(d, i, j, k, &x, &y, &z);
              place(d, i - 1, j, k, &xm, &ym, &zm);
              v = +F(ch.phi, x, y, z) - F(ch.phi, xm, ym, zm);
              place(d, i, j + 1, k, &x, &y, &z);
              place(d, i - 1, j + 1, k, &xm, &ym, &zm);
              v += F(ch.phi, x, y, z) - F(ch.phi, xm, ym, zm);
              place(d, i, j, k + 1, &x, &y, &z);
              place(d, i - 1, j, k + 1, &xm, &ym, &zm);
              v += F(ch.phi, x, y, z) - F(ch.phi, xm, ym, zm);
              place(d, i, j + 1, k + 1, &x, &y, &z);
              place(d, i - 1, j + 1, k + 1, &xm, &ym, &zm);
              v += F(ch.phi, x, y, z) - F(ch.phi, xm, ym, zm);
              v *= 0.5;
              place(d, i0, j0, k0, &x, &y, &z);
              place(d, i0 - 1, j0, k0, &xm, &ym, &zm);
              const double dd = (F(me.phi, x, y, z) - F(me.phi, xm, ym, zm) - v) * (1.0 / 6.0);
              if (!refined(b, x, y, z)) F(me.dphi, x, y, z) += dd;
              if (!refined(b, xm, ym, zm)) F(me.dphi, xm, ym, zm) -= dd;
            }
        }
  }
  void adjust_force(int b, int d) {
    for (int c = 0; c < 8; c++)
      if (nd[b].child[c] >= 0) adjust_force(nd[b].child[c], d);
    PNode& me = nd[b];
    std::vector<double>& f = d == 0 ? me.fx : d == 1 ? me.fy : me.fz;
    for (int l = 0; l < 3; l++)
      for (int ck = 0; ck < 2; ck++)
        for (int cj = 0; cj < 2; cj++) {
          int i, i0;
          const int c = cf_child(b, d, l, cj, ck, &i, &i0);
          if (c < 0) continue;
          const PNode& ch = nd[c];
          const std::vector<double>& cf = d == 0 ? ch.fx : d == 1 ? ch.fy : ch.fz;
          for (int k = 1; k < P - 2; k += 2)
            for (int j = 1; j < P - 2; j += 2) {
              const int k0 = (1 + k) / 2 + ck * (P / 2 - 1);
              const int j0 = (1 + j) / 2 + cj * (P / 2 - 1);
              int x, y, z;
              double v;
              place(d, i, j, k, &x, &y, &z);
              v = +F(cf, x, y, z);
              place(d, i, j + 1, k, &x, &y, &z);
              v += F(cf, x, y, z);
              place(d, i, j, k + 1, &x, &y, &z);
              v += F(cf, x, y, z);
              place(d, i, j + 1, k + 1, &x, &y, &z);
              v += F(cf, x, y, z);
              place(d, i0, j0, k0, &x, &y, &z);
              F(f, x, y, z) = v * 0.25;
            }
        }
  }

  // ---- the V-cycle steps -------------------------------------------------------------------
  void vcycle_down(int b) {
    PNode& me = nd[b];
    const double h2 = me.h * me.h;
    for (int c = 0; c < 8; c++) {
      const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
      const int ch = me.child[c];
      for (int k = 1; k < P - 2; k += 2)
        for (int j = 1; j < P - 2; j += 2) {
          const int k0 = (1 + k) / 2 + cz * (P / 2 - 1);
          const int j0 = (1 + j) / 2 + cy * (P / 2 - 1);
          if (ch < 0) {
            const int i0min = 1 + cx * (P / 2 - 1), i0max = P / 2 + cx * (P / 2 - 1) - 1;
            for (int i0 = i0min; i0 <= i0max; i0++)
              F(me.dphi, i0, j0, k0) = (div(me.phi, i0, j0, k0) - h2 * F(me.S, i0, j0, k0)) * (1.0 / 6.0);
          } else {
            for (int i = 1; i < P - 2; i += 2) {
              const int i0 = (1 + i) / 2 + cx * (P / 2 - 1);
              F(me.dphi1, i0, j0, k0) = oct_avg(nd[ch].dphi1, i, j, k);
              F(me.dphi, i0, j0, k0) = oct_avg(nd[ch].dphi, i, j, k);
            }
          }
        }
    }
    adjust_dphi(b, 0);
    adjust_dphi(b, 1);
    adjust_dphi(b, 2);
    for (int k = 1; k < P - 1; k++)
      for (int j = 1; j < P - 1; j++)
        for (int i = 1; i < P - 1; i++)
          if (!refined(b, i, j, k)) F(me.dphi1, i, j, k) = F(me.dphi, i, j, k);
  }
  double relax_node(int b) {
    PNode& me = nd[b];
    std::vector<double> tmp(me.dphi.size(), 0.0);
    double resid = 0.0, resid0 = 0.0;
    bool first = true;
    int iters = 0;
    do {
      resid = 0.0;
      for (int k = 1; k < P - 1; k++)
        for (int j = 1; j < P - 1; j++)
          for (int i = 1; i < P - 1; i++) F(tmp, i, j, k) = F(me.dphi1, i, j, k) + div(me.dphi, i, j, k) * (1.0 / 6.0);
      for (int k = 1; k < P - 1; k++)
        for (int j = 1; j < P - 1; j++)
          for (int i = 1; i < P - 1; i++) {
            F(me.dphi, i, j, k) += F(tmp, i, j, k);
            resid += F(tmp, i, j, k) * F(tmp, i, j, k);
          }
      if (first) {
        resid0 = resid;
        first = false;
      }
      iters++;
    } while (resid / resid0 > RELAX_RESID && iters < 16);
    return resid;
  }
  double relax(int b, int l) {  // the reference's recursion and its order of summation
    if (nd[b].level == l) return relax_node(b);
    double r = 0.0;
    for (int c = 0; c < 8; c++)
      if (nd[b].child[c] >= 0) r += relax(nd[b].child[c], l);
    return r;
  }
  void coarse_correction(int b) {
    PNode& me = nd[b];
    const PNode& p = nd[me.parent];
    const int c = me.cidx, cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    for (int k = 1; k < P - 2; k += 2)
      for (int j = 1; j < P - 2; j += 2) {
        const int k0 = (1 + k) / 2 + cz * (P / 2 - 1);
        const int j0 = (1 + j) / 2 + cy * (P / 2 - 1);
        int i0 = 1 + cx * (P / 2 - 1);
        for (int i = 1; i < P - 1; i += 2, i0++) {
          const double a = F(p.dphi, i0, j0, k0) - oct_avg(me.dphi, i, j, k);
          for (int dz = 0; dz < 2; dz++)
            for (int dy = 0; dy < 2; dy++)
              for (int dx = 0; dx < 2; dx++) F(me.dphi, i + dx, j + dy, k + dz) += a;
        }
      }
  }
  void retire(int b) {
    PNode& me = nd[b];
    for (int k = 1; k < P - 1; k++)
      for (int j = 1; j < P - 1; j++)
        for (int i = 1; i < P - 1; i++) F(me.phi, i, j, k) += F(me.dphi, i, j, k);
  }
  void phi_from_children(int b) {
    for (int c = 0; c < 8; c++)
      if (nd[b].child[c] >= 0) phi_from_children(nd[b].child[c]);
    PNode& me = nd[b];
    for (int c = 0; c < 8; c++) {
      const int ch = me.child[c];
      if (ch < 0) continue;
      const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
      for (int k = 1; k < P - 2; k += 2)
        for (int j = 1; j < P - 2; j += 2) {
          const int k0 = (1 + k) / 2 + cz * (P / 2 - 1);
          const int j0 = (1 + j) / 2 + cy * (P / 2 - 1);
          for (int i = 1; i < P - 1; i += 2) {
            const int i0 = (1 + i) / 2 + cx * (P / 2 - 1);
            double sum = 0.0;
            // weights: 11 where all three offsets are inner (0 or 1), -1 where exactly two are
            for (int o3 = -1; o3 < 3; o3++)
              for (int o2 = -1; o2 < 3; o2++)
                for (int o1 = -1; o1 < 3; o1++) {
                  const int inner = (o1 == 0 || o1 == 1) + (o2 == 0 || o2 == 1) + (o3 == 0 || o3 == 1);
                  const int wgt = inner == 3 ? 11 : inner == 2 ? -1 : 0;
                  sum += wgt * F(nd[ch].phi, i + o1, j + o2, k + o3);
                }
            F(me.phi, i0, j0, k0) = sum / 64.0;
          }
        }
    }
  }
  void compute_forces_node(int b) {
    PNode& me = nd[b];
    const double h = me.h;
    for (int k = 1; k < P; k++)
      for (int j = 1; j < P; j++)
        for (int i = 1; i < P; i++) {
          F(me.fx, i, j, k) = -(F(me.phi, i, j, k) - F(me.phi, i - 1, j, k)) / h;
          F(me.fy, i, j, k) = -(F(me.phi, i, j, k) - F(me.phi, i, j - 1, k)) / h;
          F(me.fz, i, j, k) = -(F(me.phi, i, j, k) - F(me.phi, i, j, k - 1)) / h;
        }
  }
  void compute_forces() {
    for (int b = 0; b < (int)nd.size(); b++) compute_forces_node(b);
    adjust_force(0, 0);
    adjust_force(0, 1);
    adjust_force(0, 2);
  }
  double solution_error(int b) {
    PNode& me = nd[b];
    const double h1 = me.h;
    const double h3 = std::pow(h1, 3);
    const double h1inv = 1.0 / h1;
    double sum = 0.0;
    for (int k = 1; k < P - 1; k++)
      for (int j = 1; j < P - 1; j++)
        for (int i = 1; i < P - 1; i++) {
          if (!refined(b, i, j, k)) {
            double df = -(F(me.fx, i + 1, j, k) - F(me.fx, i, j, k)) * h1inv;
            df -= (F(me.fy, i, j + 1, k) - F(me.fy, i, j, k)) * h1inv;
            df -= (F(me.fz, i, j, k + 1) - F(me.fz, i, j, k)) * h1inv;
            F(me.dphi, i, j, k) = (df - F(me.S, i, j, k));
            sum += std::fabs(F(me.dphi, i, j, k)) * h3;
          } else {
            F(me.dphi, i, j, k) = 0.0;
          }
        }
    for (int c = 0; c < 8; c++)
      if (me.child[c] >= 0) sum += solution_error(me.child[c]);
    return sum;
  }
  void vcycle0() {
    phi_from_children(0);
    for (int l = 0; l <= maxlevel; l++) enforce_phi(l);
    for (auto& q : nd) {  // init_vcycle
      std::fill(q.dphi.begin(), q.dphi.end(), 0.0);
      std::fill(q.dphi1.begin(), q.dphi1.end(), 0.0);
    }
    for (int l = maxlevel; l >= 0; l--) {
      enforce_dphi(l);
      for (int b : at_level(l)) vcycle_down(b);
      enforce_dphi(l);
      const double resid0 = relax(0, l);
      double resid;
      do {
        enforce_dphi(l);
        resid = relax(0, l);
      } while (resid / resid0 > VDOWN_RESID);
    }
    for (int l = 0; l <= maxlevel; l++) {
      enforce_dphi(l);
      if (l > 0)
        for (int b : at_level(l)) coarse_correction(b);
      enforce_dphi(l);
      const double resid0 = relax(0, l);
      double resid;
      do {
        enforce_dphi(l);
        resid = relax(0, l);
      } while (resid / resid0 > VUP_RESID);
      for (int b : at_level(l)) retire(b);
      enforce_phi(l);
    }
  }

  // ---- physical boundary phi (binary/binary.cpp:938-1019) ----------------------------------
  void source_sums(int b) {
    PNode& me = nd[b];
    const double h3 = me.h * me.h * me.h;
    for (int c = 0; c < 8; c++) me.s0[c] = me.sx[c] = me.sy[c] = me.sz[c] = 0.0;
    me.all_refined = true;
    for (int k = 1; k < P - 1; k++)
      for (int j = 1; j < P - 1; j++)
        for (int i = 1; i < P - 1; i++) {
          if (refined(b, i, j, k)) continue;
          me.all_refined = false;
          const int ci = i / (P / 2) + ((j / (P / 2)) << 1) + ((k / (P / 2)) << 2);
          const double s = F(me.S, i, j, k);
          me.s0[ci] += s * h3;
          me.sx[ci] += s * pc(b, 0, i) * h3;
          me.sy[ci] += s * pc(b, 1, j) * h3;
          me.sz[ci] += s * pc(b, 2, k) * h3;
        }
    for (int c = 0; c < 8; c++) {
      if (!me.all_refined && me.s0[c] != 0.0) {
        me.sx[c] = me.sx[c] / me.s0[c];
        me.sy[c] = me.sy[c] / me.s0[c];
        me.sz[c] = me.sz[c] / me.s0[c];
        me.s0[c] = me.s0[c] / (4.0 * M_PI);
      } else {
        me.s0[c] = 0.0;
        me.sx[c] = me.sy[c] = me.sz[c] = 1.0;
      }
    }
    for (int c = 0; c < 8; c++)
      if (me.child[c] >= 0) source_sums(me.child[c]);
  }
  double compute_phi(int b, double x, double y, double z) const {
    const PNode& me = nd[b];
    double alpha = 1.0;
    if (z_reflect) alpha *= 2.0;
    double sum = 0.0;
    for (int c = 0; c < 8; c++) {
      const double dx = x - me.sx[c], dy = y - me.sy[c], dz = z - me.sz[c];
      if (!me.all_refined)
        sum -= alpha * me.s0[c] / std::sqrt(dx * dx + dy * dy + dz * dz);
      else
        sum = 0.0;
    }
    for (int c = 0; c < 8; c++)
      if (me.child[c] >= 0) sum += compute_phi(me.child[c], x, y, z);
    return sum;
  }
  void physical_boundaries() {
    for (int b = 0; b < (int)nd.size(); b++)
      for (int f = 0; f < 6; f++) {
        const Face s = face(b, f);
        if (s.kind != K_PHYS || s.reflect) continue;
        const int d = f >> 1, g = (f & 1) ? P - 1 : 0;
        std::vector<double>& t = nd[b].tab[f];
        t.assign(P * P, 0.0);
        for (int a = 0; a < P; a++)
          for (int c = 0; c < P; c++) {
            double x, y, z;
            if (d == 0) {
              x = pc(b, 0, g), y = pc(b, 1, a), z = pc(b, 2, c);
            } else if (d == 1) {
              x = pc(b, 0, a), y = pc(b, 1, g), z = pc(b, 2, c);
            } else {
              x = pc(b, 0, a), y = pc(b, 1, c), z = pc(b, 2, g);
            }
            t[a * P + c] = compute_phi(0, x, y, z);
          }
      }
  }
};

extern "C" {

orc_poisson* orc_poisson_create(int n, int bw, double grid_dim, int z_reflect, int nn, const int* parent,
                                const int* cidx) {
  if (n < 4 || (n % 2) || nn < 1 || parent[0] != -1) return nullptr;
  orc_poisson* p = new orc_poisson;
  p->n = n;
  p->P = n + 2;
  p->bw = bw;
  p->z_reflect = z_reflect;
  p->grid_dim = grid_dim;
  p->h0 = 2.0 * grid_dim / double(n);
  p->nd.resize(nn);
  const int GNX = n + 2 * bw;
  const long P3 = (long)p->P * p->P * p->P;
  for (int b = 0; b < nn; b++) {
    PNode& q = p->nd[b];
    for (auto* a : {&q.phi, &q.phi0, &q.dphi, &q.dphi1, &q.S, &q.fx, &q.fy, &q.fz}) a->assign(P3, 0.0);
    if (b == 0) {
      q.h = p->h0;
    } else {
      const int pa = parent[b], c = cidx[b];
      if (pa < 0 || pa >= b || c < 0 || c > 7 || p->nd[pa].child[c] >= 0) {
        delete p;
        return nullptr;
      }
      PNode& P_ = p->nd[pa];
      P_.child[c] = b;
      q.parent = pa;
      q.cidx = c;
      q.level = P_.level + 1;
      q.h = P_.h * 0.5;
      const int cv[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
      for (int d = 0; d < 3; d++) {
        q.loc[d] = 2 * P_.loc[d] + cv[d];
        q.off[d] = 2 * P_.off[d] + bw + cv[d] * (GNX - 2 * bw);  // grid/grid_node.cpp:275
      }
      if (q.level > p->maxlevel) p->maxlevel = q.level;
    }
    p->by_loc[std::make_tuple(q.level, q.loc[0], q.loc[1], q.loc[2])] = b;
  }
  return p;
}

void orc_poisson_destroy(orc_poisson* p) { delete p; }

double* orc_poisson_field(orc_poisson* p, int node, int which) {
  if (!p || node < 0 || node >= (int)p->nd.size()) return nullptr;
  PNode& q = p->nd[node];
  std::vector<double>* f[8] = {&q.phi, &q.S, &q.fx, &q.fy, &q.fz, &q.dphi, &q.dphi1, &q.phi0};
  if (which < 0 || which > 7) return nullptr;
  return f[which]->data();
}

int orc_poisson_prepare(orc_poisson* p) {
  p->source_sums(0);
  p->physical_boundaries();
  return 0;
}

double orc_poisson_boundary_phi(const orc_poisson* p, double x, double y, double z) {
  return p->compute_phi(0, x, y, z);
}

double orc_poisson_cell_centre(const orc_poisson* p, int node, int d, int i) { return p->pc(node, d, i); }

int orc_poisson_phi_from_children(orc_poisson* p) {
  p->phi_from_children(0);
  return 0;
}

int orc_poisson_enforce_phi(orc_poisson* p) {
  for (int l = 0; l <= p->maxlevel; l++) p->enforce_phi(l);
  return 0;
}

int orc_poisson_forces(orc_poisson* p) {
  p->compute_forces();
  return 0;
}

double orc_poisson_error(orc_poisson* p) { return p->solution_error(0); }

double orc_poisson_vcycle(orc_poisson* p) {
  p->vcycle0();
  p->compute_forces();
  return p->solution_error(0);
}

int orc_poisson_enforce_phi_level(orc_poisson* p, int l) {
  if (l < 0 || l > p->maxlevel) return 1;
  p->enforce_phi(l);
  return 0;
}

double* orc_poisson_boundary_table(orc_poisson* p, int node, int face) {
  if (!p || node < 0 || node >= (int)p->nd.size() || face < 0 || face > 5) return nullptr;
  std::vector<double>& t = p->nd[node].tab[face];
  return t.empty() ? nullptr : t.data();
}

int orc_poisson_solve(orc_poisson* p, double toler, int max_iter, int prepare, int* iters, double* error) {
  if (prepare) orc_poisson_prepare(p);
  p->compute_forces();
  int it = 0;
  double err;
  do {
    if (it > max_iter) return 1;
    err = orc_poisson_vcycle(p);
    it++;
  } while (err > toler);
  if (iters) *iters = it;
  if (error) *error = err;
  return 0;
}

// Poisson::store_phi / restore_phi / restore_half_phi (poisson/poisson.cpp:144-221)
int orc_poisson_store_phi(orc_poisson* p, int mode) {
  const int P = p->P;
  for (auto& q : p->nd)
    for (int k = 1; k < P - 1; k++)
      for (int j = 1; j < P - 1; j++)
        for (int i = 1; i < P - 1; i++) {
          const long e = p->at(i, j, k);
          if (mode == 0) {
            q.phi0[e] = q.phi[e];
          } else if (mode == 1) {
            q.phi[e] = q.phi0[e];
          } else {
            const double r = q.phi[e];
            q.phi[e] = (q.phi[e] + q.phi0[e]) * 0.5;
            q.phi0[e] = r;
          }
        }
  return 0;
}

}  // extern "C"
