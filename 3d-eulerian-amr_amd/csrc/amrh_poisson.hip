// Self-gravity on the octree: the reference's multigrid Poisson solver (poisson/poisson.cpp,
// poisson/poisson_amr_bound.cpp, poisson/poisson_phys_bound.cpp) with its physical-boundary phi
// from per-octant monopoles (binary/binary.cpp:938-1019), driven as BinaryDriver::solve does
// (binary/binary_driver.cpp:294-344).
//
// Every node keeps phi, phi0, dphi, dphi1, S, fx, fy, fz as (n+2)^3 cubes with one ghost layer,
// x fastest (Poisson's PNX = GNX - 2 BW + 2, poisson/poisson.h:4), node-major in one allocation.
// Each recursion of the reference over the tree becomes one launch over the nodes of a level (or
// of all levels where the recursion has no order).  vcycle_relax never exchanges ghosts between
// its Jacobi sweeps (poisson/poisson.cpp:88-130), so all sweeps of a node run in one workgroup
// with the node's dphi in LDS.  The host keeps the residual-ratio loops of vcycle0 and sums the
// per-node residuals and errors in the reference's recursion order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "amrh_ctx.h"
#include "amrh_internal.h"
#include "amrhydro.h"

using namespace amrh;

namespace {

constexpr double RELAX_RESID = 0.5, VDOWN_RESID = 1.0, VUP_RESID = 0.05;
constexpr int MAXDEPTH = 24;
constexpr int NTHREADS = 256;
enum { F_SIB = 0, F_TAB = 1, F_REFL = 2, F_AMR = 3 };
enum { PHI = 0, PHI0, DPHI, DPHI1, SRC, FX, FY, FZ, NFIELD };

struct PNodeDev {
  int off[3];
  int parent, cidx, level;
  int child[8];
  int fkind[6], fnode[6];  // per face: neighbour node | boundary table | - | parent
  double h;
};

struct PArgs {
  int P, bw, z_reflect, nn;
  long P3;
  double grid_dim, h0;
  const PNodeDev* nd;
  const int* list;  // the nodes of this launch
  double* f[NFIELD];
  double* tab;
  double* part;           // per-node partial results
  const double* src;      // per node 8 octants x (s0, x, y, z)
  const int* allref;      // per node: every octant refined
  const int* pre;         // nodes in pre-order
  const int2* slots;      // boundary table -> (node, face)
  double* exps;           // per node the six interior layers next to its faces, two copies
  struct LoopState* st;   // the residual-ratio loop of vcycle0 on the device
};

// One "resid0 = vcycle_relax(l); do { enforce_dphi_boundaries(l); resid = vcycle_relax(l); }
// while (resid / resid0 > limit)" loop of vcycle0 (poisson/poisson.cpp:66-84), kept on the device.
struct LoopState {
  double resid0, resid;
  int iters, done;
};

__device__ __forceinline__ long cell(int P, int i, int j, int k) { return ((long)k * P + j) * P + i; }

__device__ __forceinline__ bool phys_axis(const PNodeDev& nd, int d) {
  const int a = nd.fkind[2 * d], b = nd.fkind[2 * d + 1];
  return a == F_TAB || a == F_REFL || b == F_TAB || b == F_REFL;
}

// index along d and the two across it -> (x, y, z), as adjust_{x,y,z}_* index their arrays
__device__ __forceinline__ long placed(int P, int d, int a, int j, int k) {
  return d == 0 ? cell(P, a, j, k) : d == 1 ? cell(P, j, a, k) : cell(P, j, k, a);
}

// Grid::xc of Poisson cell i (Poisson::pxc = xc(i + BW - 1), grid/grid.cpp:498-526)
__device__ double centre(const PArgs& A, int b, int d, int i) {
  const PNodeDev& q = A.nd[b];
  const int ii = i + A.bw - 1;
  double f;
  if (d == 2 && A.z_reflect)
    f = double(q.off[2] + ii) * q.h - double(A.bw) * A.h0 - 0.0;
  else if (d == 2)
    f = double(q.off[2] + ii) * q.h - A.grid_dim - double(A.bw) * A.h0;
  else
    f = double(q.off[d] + ii) * q.h - A.grid_dim - double(A.bw) * A.h0 - 0.0;
  return f + 0.5 * q.h;
}

// index of a ghost's source in the parent frame (PoissonAMRBound::set_sibling's rel_offset,
// poisson/poisson_amr_bound.cpp:129-150, plus the sibling-frame index)
__device__ __forceinline__ void amr_q(const PArgs& A, int b, int f, const int idx[3], int q[3]) {
  const PNodeDev& me = A.nd[b];
  const PNodeDev& pa = A.nd[me.parent];
  const int d = f >> 1, v = (f & 1) ? 1 : -1;
  for (int a = 0; a < 3; a++) q[a] = me.off[a] + (a == d ? v * (A.P - 2) : 0) - 2 * pa.off[a] - (A.bw - 1) + idx[a];
}

// PoissonAMRBound::get_phi (poisson/poisson_amr_bound.cpp:33-111): quadratic Lagrange weights at
// a quarter cell, tensor product over 27 parent cells, through the parent's neighbour where the
// stencil centre leaves the parent
__device__ double amr_phi(const PArgs& A, int b, int f, const int idx[3]) {
  const double w[3] = {-0.09375, 0.9375, 0.15625};
  int q[3], i0[3], i1[3];
  amr_q(A, b, f, idx, q);
  for (int a = 0; a < 3; a++) {
    i1[a] = q[a] % 2;
    i0[a] = q[a] / 2;
  }
  int par = A.nd[b].parent;
  for (int a = 0; a < 3; a++) {
    if (i0[a] > A.P - 2) {
      const PNodeDev& p = A.nd[par];
      if (p.fkind[2 * a + 1] != F_SIB) return 0.0;
      par = p.fnode[2 * a + 1];
      i0[a] -= A.P - 2;
    } else if (i0[a] < 1) {
      const PNodeDev& p = A.nd[par];
      if (p.fkind[2 * a] != F_SIB) return 0.0;
      par = p.fnode[2 * a];
      i0[a] += A.P - 2;
    }
  }
  const double* ph = A.f[PHI] + (long)par * A.P3;
  const int sx = 2 * i1[0] - 1, sy = 2 * i1[1] - 1, sz = 2 * i1[2] - 1;
  double sum = 0.0;
  for (int l = 0; l < 3; l++)
    for (int m = 0; m < 3; m++)
      for (int o = 0; o < 3; o++)
        sum += (w[l] * w[m] * w[o]) *
               ph[cell(A.P, i0[0] + (l - 1) * sx, i0[1] + (m - 1) * sy, i0[2] + (o - 1) * sz)];
  return sum;
}

// what the neighbour object across face f returns for sibling-frame index (i, j, k)
template <bool ISD>
__device__ double across(const PArgs& A, int b, int f, int i, int j, int k) {
  const PNodeDev& nd = A.nd[b];
  const int kind = nd.fkind[f], d = f >> 1;
  int idx[3] = {i, j, k};
  const double* fld = A.f[ISD ? DPHI : PHI];
  if (kind == F_SIB) return fld[(long)nd.fnode[f] * A.P3 + cell(A.P, i, j, k)];
  if (kind == F_TAB) {  // PoissonPhysBound::get_phi: the boundary table; dphi is zero
    if (ISD) return 0.0;
    const int a = d == 0 ? j : i, c = d == 2 ? j : k;
    return A.tab[(long)nd.fnode[f] * A.P * A.P + a * A.P + c];
  }
  if (kind == F_REFL) {  // reflection: the node's own cell mirrored across the face
    idx[d] = idx[d] == 1 ? A.P - 2 : 1;
    return fld[(long)b * A.P3 + cell(A.P, idx[0], idx[1], idx[2])];
  }
  if (ISD) {  // PoissonAMRBound::get_dphi: the parent's dphi, piecewise constant
    int q[3];
    amr_q(A, b, f, idx, q);
    return fld[(long)nd.parent * A.P3 + cell(A.P, q[0] / 2, q[1] / 2, q[2] / 2)];
  }
  return amr_phi(A, b, f, idx);
}

// enforce_phi_boundaries (faces, then edges, then vertices: poisson/poisson.cpp:745-902) and
// enforce_dphi_boundaries (faces, :904-942) for the nodes of one level.  An edge or corner ghost
// is read from the neighbour across its first ghost axis (x, y, z order) that ends at the
// physical boundary, else across its last ghost axis -- the choice the reference spells out case
// by case.
template <bool ISD>
__global__ __launch_bounds__(NTHREADS) void ghost_kernel(PArgs A, int pass) {
  const int b = A.list[blockIdx.x];
  const PNodeDev& nd = A.nd[b];
  const int P = A.P, M = P - 2, Q = P - 1, T = P - 2;
  double* dst = A.f[ISD ? DPHI : PHI] + (long)b * A.P3;
  if (pass == 0) {
    for (int t = threadIdx.x; t < 6 * M * M; t += blockDim.x) {
      const int f = t / (M * M), r = t % (M * M), a = r % M + 1, c = r / M + 1;
      const int d = f >> 1, up = f & 1;
      int g[3];
      if (d == 0) {
        g[0] = up ? Q : 0, g[1] = a, g[2] = c;
      } else if (d == 1) {
        g[0] = a, g[1] = up ? Q : 0, g[2] = c;
      } else {
        g[0] = a, g[1] = c, g[2] = up ? Q : 0;
      }
      int s[3] = {g[0], g[1], g[2]};
      s[d] = up ? 1 : T;
      dst[cell(P, g[0], g[1], g[2])] = across<ISD>(A, b, f, s[0], s[1], s[2]);
    }
    return;
  }
  const int count = pass == 1 ? 12 * M : 8;
  for (int t = threadIdx.x; t < count; t += blockDim.x) {
    int g[3], ds;
    if (pass == 1) {
      const int e = t / M, along = t % M + 1;
      const int pair = e >> 2, s1 = e & 1, s2 = (e >> 1) & 1;  // ghost axes (x,y), (x,z), (y,z)
      const int d1 = pair == 2 ? 1 : 0, d2 = pair == 0 ? 1 : 2, d3 = 3 - d1 - d2;
      g[d1] = s1 ? Q : 0;
      g[d2] = s2 ? Q : 0;
      g[d3] = along;
      ds = phys_axis(nd, d1) ? d1 : d2;
    } else {
      g[0] = (t & 1) ? Q : 0, g[1] = (t & 2) ? Q : 0, g[2] = (t & 4) ? Q : 0;
      ds = phys_axis(nd, 0) ? 0 : phys_axis(nd, 1) ? 1 : 2;
    }
    int s[3] = {g[0], g[1], g[2]};
    s[ds] = g[ds] == 0 ? T : 1;
    const int f = 2 * ds + (g[ds] == 0 ? 0 : 1);
    dst[cell(P, g[0], g[1], g[2])] = across<false>(A, b, f, s[0], s[1], s[2]);
  }
}

__device__ __forceinline__ double divergence(const double* a, int P, int i, int j, int k) {
  const long e = cell(P, i, j, k);
  return (a[e + 1] + a[e - 1] + a[e + P] + a[e - P] + a[e + (long)P * P] + a[e - (long)P * P]) - (a[e] * 6.0);
}

__device__ __forceinline__ double oct_avg(const double* a, int P, int i, int j, int k) {
  const long e = cell(P, i, j, k), y = P, z = (long)P * P;
  return (a[e] + a[e + 1] + a[e + y] + a[e + y + 1] + a[e + z] + a[e + z + 1] + a[e + z + y] + a[e + z + y + 1]) * 0.125;
}

// The child across the face at position l (0: low face, 1: middle, 2: high face) of axis d in
// quadrant (cj, ck), when exactly one side of that face is refined: its index, the child's face
// index fi and the face's index i0 in the node (poisson/poisson.cpp:249-301, :474-526)
__device__ int cf_child(const PArgs& A, int b, int d, int l, int cj, int ck, int* fi, int* i0) {
  const PNodeDev& nd = A.nd[b];
  auto kid = [&](int node, int x) -> int {
    if (node < 0) return -1;
    int c3[3];
    c3[d] = x;
    c3[d == 0 ? 1 : 0] = cj;
    c3[d == 2 ? 1 : 2] = ck;
    return A.nd[node].child[c3[0] | (c3[1] << 1) | (c3[2] << 2)];
  };
  auto sib = [&](int f) { return nd.fkind[f] == F_SIB ? nd.fnode[f] : -1; };
  int cl, cr;
  if (l == 0) {
    cr = kid(b, 0);
    cl = kid(sib(2 * d), 1);
    *i0 = 1;
  } else if (l == 1) {
    cl = kid(b, 0);
    cr = kid(b, 1);
    *i0 = A.P / 2;
  } else {
    cr = kid(sib(2 * d + 1), 0);
    cl = kid(b, 1);
    *i0 = A.P - 1;
  }
  if (cr < 0 && cl >= 0) {
    *fi = A.P - 1;
    return cl;
  }
  if (cl < 0 && cr >= 0) {
    *fi = 1;
    return cr;
  }
  return -1;
}

// vcycle_down for the nodes of one level (poisson/poisson.cpp:989-1050): the residual of
// unrefined cells, the children's averages on refined ones, the coarse-fine flux matching of
// adjust_{x,y,z}_dphi (each cell meets at most one matched face per axis; x, y, z in turn), and
// dphi1 = dphi on unrefined cells.  Ghost cells the reference also touches are refilled before
// anything reads them, so they are left alone.
__global__ __launch_bounds__(NTHREADS) void down_kernel(PArgs A) {
  const int b = A.list[blockIdx.x];
  const PNodeDev& nd = A.nd[b];
  const int P = A.P, M = P - 2, H = P / 2;
  const double h2 = nd.h * nd.h;
  const long base = (long)b * A.P3;
  const double* phi = A.f[PHI] + base;
  const double* S = A.f[SRC] + base;
  double* dphi = A.f[DPHI] + base;
  double* dphi1 = A.f[DPHI1] + base;
  for (int t = threadIdx.x; t < M * M * M; t += blockDim.x) {
    const int idx[3] = {t % M + 1, (t / M) % M + 1, t / (M * M) + 1};
    const int cx = idx[0] >= H, cy = idx[1] >= H, cz = idx[2] >= H;
    const int ch = nd.child[cx | (cy << 1) | (cz << 2)];
    const long e = cell(P, idx[0], idx[1], idx[2]);
    double v;
    if (ch >= 0) {
      const int i = 2 * (idx[0] - cx * (H - 1)) - 1, j = 2 * (idx[1] - cy * (H - 1)) - 1,
                k = 2 * (idx[2] - cz * (H - 1)) - 1;
      dphi1[e] = oct_avg(A.f[DPHI1] + (long)ch * A.P3, P, i, j, k);
      dphi[e] = oct_avg(A.f[DPHI] + (long)ch * A.P3, P, i, j, k);
      continue;
    }
    v = (divergence(phi, P, idx[0], idx[1], idx[2]) - h2 * S[e]) * (1.0 / 6.0);
    for (int d = 0; d < 3; d++) {
      const int a = idx[d];
      int l, right;
      if (a == 1) {
        l = 0, right = 1;
      } else if (a == H) {
        l = 1, right = 1;
      } else if (a == H - 1) {
        l = 1, right = 0;
      } else if (a == P - 2) {
        l = 2, right = 0;
      } else {
        continue;
      }
      const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      const int cj = idx[t1] >= H, ck = idx[t2] >= H;
      int fi, f0;
      const int c = cf_child(A, b, d, l, cj, ck, &fi, &f0);
      if (c < 0) continue;
      const double* cp = A.f[PHI] + (long)c * A.P3;
      const int j = 2 * (idx[t1] - cj * (H - 1)) - 1, k = 2 * (idx[t2] - ck * (H - 1)) - 1;
      double w;
      w = +cp[placed(P, d, fi, j, k)] - cp[placed(P, d, fi - 1, j, k)];
      w += cp[placed(P, d, fi, j + 1, k)] - cp[placed(P, d, fi - 1, j + 1, k)];
      w += cp[placed(P, d, fi, j, k + 1)] - cp[placed(P, d, fi - 1, j, k + 1)];
      w += cp[placed(P, d, fi, j + 1, k + 1)] - cp[placed(P, d, fi - 1, j + 1, k + 1)];
      w *= 0.5;
      const double dd =
          (phi[placed(P, d, f0, idx[t1], idx[t2])] - phi[placed(P, d, f0 - 1, idx[t1], idx[t2])] - w) * (1.0 / 6.0);
      v = right ? v + dd : v - dd;
    }
    dphi[e] = v;
    dphi1[e] = v;
  }
}

// v of the lane DPP control CTRL names (0 where it names none), for the two halves of a double
template <int CTRL>
__device__ __forceinline__ double dpp64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Sum over the wavefront, the same on every lane: pairs, quads, half rows and rows of 16 lanes
// through DPP (VALU moves, no LDS traffic), then the four row sums
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp64<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp64<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp64<0x141>(v);  // row_half_mirror
  v += dpp64<0x140>(v);  // row_mirror
  return ((lane64(v, 0) + lane64(v, 16)) + lane64(v, 32)) + lane64(v, 48);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < (int)(blockDim.x >> 6); q++) s += red[q];
  return s;
}

// interior layer next to face f of a node: [copy][node][face][(c - 1) M + (a - 1)], (a, c) the
// face's other two axes in x, y, z order
__device__ __forceinline__ long exp_at(const PArgs& A, int copy, int b, int f, int a, int c) {
  const int M = A.P - 2;
  return (((long)copy * A.nn + b) * 6 + f) * M * M + (long)(c - 1) * M + (a - 1);
}

__device__ __forceinline__ long layer_cell(int P, int f, int a, int c, int t) {
  const int d = f >> 1;
  return d == 0 ? cell(P, t, a, c) : d == 1 ? cell(P, a, t, c) : cell(P, a, c, t);
}

// copy 1 of the layers from dphi, before a loop's first step
__global__ __launch_bounds__(NTHREADS) void export_kernel(PArgs A) {
  const int b = A.list[blockIdx.x];
  const int P = A.P, M = P - 2;
  const double* d = A.f[DPHI] + (long)b * A.P3;
  for (int t = threadIdx.x; t < 6 * M * M; t += blockDim.x) {
    const int f = t / (M * M), r = t % (M * M), a = r % M + 1, c = r / M + 1;
    A.exps[exp_at(A, 1, b, f, a, c)] = d[layer_cell(P, f, a, c, (f & 1) ? M : 1)];
  }
}

// Step k of the loop for the nodes of one level: enforce_dphi_boundaries (neighbours' layers from
// step k-1's copy, the parent's dphi, the mirror, or 0) and vcycle_relax (poisson/poisson.cpp:
// 88-130): Jacobi sweeps of dphi += dphi1 + div(dphi)/6 over the interior until the sum of squared
// updates falls to half of the first sweep's, at most 16.  Each node's residual goes to step k's
// copy of the per-node residuals (two copies, by step parity); the launch of step k+1 sums them
// and decides whether the loop goes on, so no workgroup waits for or fences against the others;
// once the loop has stopped, later launches return at once.
template <int N>
__global__ __launch_bounds__(NTHREADS) void relax_step_kernel(PArgs A, int k, double limit, int count) {
  constexpr int P = N + 2, M = N, CELLS = M * M * M, PER = (CELLS + NTHREADS - 1) / NTHREADS;
  constexpr long P3 = (long)P * P * P;
  // two copies of the cube where they fit: a sweep reads one and writes the other, so one barrier
  // per sweep suffices; otherwise the update waits for every thread's stencil
  constexpr bool TWO = 2 * P3 * 8 <= 65536;
  __shared__ double L[TWO ? 2 * P3 : P3];
  __shared__ double red[2][NTHREADS / 64];
  if (A.st->done) return;
  if (k > 0) {
    // the decision on step k-1, taken by every workgroup from the level's residuals that step
    // wrote (visible here without fences: they come from the previous launch), summed in the same
    // order everywhere
    const double* prev = A.part + (long)((k - 1) & 1) * A.nn;
    double s = 0.0;
    for (int q = threadIdx.x; q < count; q += NTHREADS) s += prev[A.list[q]];
    const double total = block_sum(s, red[0]);
    LoopState* st = A.st;
    if (k == 1) {
      if (blockIdx.x == 0 && threadIdx.x == 0) st->resid0 = total;
    } else if (!(total / st->resid0 > limit)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->resid = total;
        st->iters = k;
        st->done = 1;
      }
      return;
    }
  }
  const int b = A.list[blockIdx.x];
  const PNodeDev& nd = A.nd[b];
  const long base = (long)b * A.P3;
  double* g = A.f[DPHI] + base;
  double d1[PER];
  long at[PER];
#pragma unroll
  for (int c = 0; c < PER; c++) {
    const int t = threadIdx.x + c * NTHREADS;
    const int tt = t < CELLS ? t : 0;
    at[c] = cell(P, tt % M + 1, (tt / M) % M + 1, tt / (M * M) + 1);
    d1[c] = t < CELLS ? A.f[DPHI1][base + at[c]] : 0.0;
    if (t < CELLS) L[at[c]] = g[at[c]];
  }
  const int src = (k + 1) & 1;
  for (int t = threadIdx.x; t < 6 * M * M; t += NTHREADS) {
    const int f = t / (M * M), r = t % (M * M), a = r % M + 1, c = r / M + 1;
    const int kind = nd.fkind[f], up = f & 1;
    const long e = layer_cell(P, f, a, c, up ? P - 1 : 0);
    double v;
    if (kind == F_SIB) {
      v = A.exps[exp_at(A, src, nd.fnode[f], f ^ 1, a, c)];
    } else if (kind == F_TAB) {
      v = 0.0;
    } else if (kind == F_REFL) {
      v = g[layer_cell(P, f, a, c, up ? P - 2 : 1)];
    } else {
      const int d = f >> 1;
      int idx[3];
      idx[d] = up ? 1 : P - 2;  // the sibling-frame index the node asks for
      idx[d == 0 ? 1 : 0] = a;
      idx[d == 2 ? 1 : 2] = c;
      int q[3];
      amr_q(A, b, f, idx, q);
      v = A.f[DPHI][(long)nd.parent * A.P3 + cell(P, q[0] / 2, q[1] / 2, q[2] / 2)];
    }
    L[e] = v;
    if (TWO) L[P3 + e] = v;  // ghosts stay fixed through the sweeps
    g[e] = v;
  }
  __syncthreads();
  double resid = 0.0, resid0 = 0.0;
  int iters = 0, cur = 0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (;;) {
    const double* X = L + (TWO ? cur * P3 : 0);
    double* Y = L + (TWO ? (cur ^ 1) * P3 : 0);
    double tmp[PER];
    double part = 0.0;
#pragma unroll
    for (int c = 0; c < PER; c++) {
      const long e = at[c];
      tmp[c] = d1[c] + ((X[e + 1] + X[e - 1] + X[e + P] + X[e - P] + X[e + P * P] + X[e - P * P]) - (X[e] * 6.0)) *
                           (1.0 / 6.0);
    }
    if (!TWO) __syncthreads();
#pragma unroll
    for (int c = 0; c < PER; c++) {
      if (threadIdx.x + c * NTHREADS < CELLS) {
        Y[at[c]] = X[at[c]] + tmp[c];
        part += tmp[c] * tmp[c];
      }
    }
    part = wave_sum(part);
    if (lane == 0) red[iters & 1][wv] = part;
    __syncthreads();
    resid = 0.0;
    for (int q = 0; q < NTHREADS / 64; q++) resid += red[iters & 1][q];
    cur ^= 1;
    if (iters == 0) resid0 = resid;
    iters++;
    if (!(resid / resid0 > RELAX_RESID && iters < 16)) break;
  }
  const double* X = L + (TWO ? cur * P3 : 0);
#pragma unroll
  for (int c = 0; c < PER; c++)
    if (threadIdx.x + c * NTHREADS < CELLS) g[at[c]] = X[at[c]];
  for (int t = threadIdx.x; t < 6 * M * M; t += NTHREADS) {
    const int f = t / (M * M), r = t % (M * M), a = r % M + 1, c = r / M + 1;
    A.exps[exp_at(A, k & 1, b, f, a, c)] = X[layer_cell(P, f, a, c, (f & 1) ? M : 1)];
  }
  if (threadIdx.x == 0) A.part[(long)(k & 1) * A.nn + b] = resid;
}

// vcycle_coarse_correction for the nodes of one level (poisson/poisson.cpp:14-55)
__global__ __launch_bounds__(NTHREADS) void coarse_kernel(PArgs A) {
  const int b = A.list[blockIdx.x];
  const PNodeDev& nd = A.nd[b];
  const int P = A.P, H = P / 2, G = (P - 2) / 2;
  const int c = nd.cidx, cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  const double* pd = A.f[DPHI] + (long)nd.parent * A.P3;
  double* d = A.f[DPHI] + (long)b * A.P3;
  for (int t = threadIdx.x; t < G * G * G; t += blockDim.x) {
    const int i = 2 * (t % G) + 1, j = 2 * ((t / G) % G) + 1, k = 2 * (t / (G * G)) + 1;
    const int i0 = (1 + i) / 2 + cx * (H - 1), j0 = (1 + j) / 2 + cy * (H - 1), k0 = (1 + k) / 2 + cz * (H - 1);
    const double a = pd[cell(P, i0, j0, k0)] - oct_avg(d, P, i, j, k);
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) d[cell(P, i + x, j + y, k + z)] += a;
  }
}

// vcycle_retire_dphi (poisson/poisson.cpp:1052-1068) and the phi stores of
// poisson/poisson.cpp:144-221 (mode 0 store, 1 restore, 2 restore_half), interior only
__global__ __launch_bounds__(NTHREADS) void interior_kernel(PArgs A, int op) {
  const int b = A.list[blockIdx.x];
  const int P = A.P, M = P - 2;
  const long base = (long)b * A.P3;
  double* phi = A.f[PHI] + base;
  double* phi0 = A.f[PHI0] + base;
  const double* dphi = A.f[DPHI] + base;
  for (int t = threadIdx.x; t < M * M * M; t += blockDim.x) {
    const long e = cell(P, t % M + 1, (t / M) % M + 1, t / (M * M) + 1);
    if (op == 3) {
      phi[e] += dphi[e];
    } else if (op == 0) {
      phi0[e] = phi[e];
    } else if (op == 1) {
      phi[e] = phi0[e];
    } else {
      const double r = phi[e];
      phi[e] = (phi[e] + phi0[e]) * 0.5;
      phi0[e] = r;
    }
  }
}

// phi_from_children for the nodes of one level (poisson/poisson.cpp:707-743): weight 11 where
// all three offsets are inner (0 or 1), -1 where two are, 0 elsewhere, over 64.
__global__ __launch_bounds__(NTHREADS) void restrict_kernel(PArgs A) {
  const int b = A.list[blockIdx.x];
  const PNodeDev& nd = A.nd[b];
  const int P = A.P, M = P - 2, H = P / 2;
  double* phi = A.f[PHI] + (long)b * A.P3;
  for (int t = threadIdx.x; t < M * M * M; t += blockDim.x) {
    const int i0 = t % M + 1, j0 = (t / M) % M + 1, k0 = t / (M * M) + 1;
    const int cx = i0 >= H, cy = j0 >= H, cz = k0 >= H;
    const int ch = nd.child[cx | (cy << 1) | (cz << 2)];
    if (ch < 0) continue;
    const int i = 2 * (i0 - cx * (H - 1)) - 1, j = 2 * (j0 - cy * (H - 1)) - 1, k = 2 * (k0 - cz * (H - 1)) - 1;
    const double* cp = A.f[PHI] + (long)ch * A.P3;
    double sum = 0.0;
    for (int o3 = -1; o3 < 3; o3++)
      for (int o2 = -1; o2 < 3; o2++)
        for (int o1 = -1; o1 < 3; o1++) {
          const int inner = (o1 == 0 || o1 == 1) + (o2 == 0 || o2 == 1) + (o3 == 0 || o3 == 1);
          const int wgt = inner == 3 ? 11 : inner == 2 ? -1 : 0;
          sum += wgt * cp[cell(P, i + o1, j + o2, k + o3)];
        }
    phi[cell(P, i0, j0, k0)] = sum / 64.0;
  }
}

// compute_forces over every node (poisson/poisson.cpp:223-247)
__global__ __launch_bounds__(NTHREADS) void forces_kernel(PArgs A) {
  const int b = A.list[blockIdx.x];
  const double h = A.nd[b].h;
  const int P = A.P, M = P - 1;
  const long base = (long)b * A.P3;
  const double* phi = A.f[PHI] + base;
  for (int t = threadIdx.x; t < M * M * M; t += blockDim.x) {
    const int i = t % M + 1, j = (t / M) % M + 1, k = t / (M * M) + 1;
    const long e = cell(P, i, j, k);
    A.f[FX][base + e] = -(phi[e] - phi[e - 1]) / h;
    A.f[FY][base + e] = -(phi[e] - phi[e - P]) / h;
    A.f[FZ][base + e] = -(phi[e] - phi[e - (long)P * P]) / h;
  }
}

// adjust_{x,y,z}_force for the nodes of one level, deepest first (poisson/poisson.cpp:474-635):
// a coarse face next to a refined region takes the mean of the four fine face forces
__global__ __launch_bounds__(NTHREADS) void adjust_force_kernel(PArgs A) {
  const int b = A.list[blockIdx.x];
  const int P = A.P, H = P / 2, G = (P - 2) / 2;
  for (int t = threadIdx.x; t < 36 * G * G; t += blockDim.x) {
    const int d = t / (12 * G * G), r = t % (12 * G * G);
    const int l = r / (4 * G * G), r2 = r % (4 * G * G);
    const int q = r2 / (G * G), r3 = r2 % (G * G);
    const int cj = q & 1, ck = q >> 1;
    const int j = 2 * (r3 % G) + 1, k = 2 * (r3 / G) + 1;
    int fi, f0;
    const int c = cf_child(A, b, d, l, cj, ck, &fi, &f0);
    if (c < 0) continue;
    const int j0 = (1 + j) / 2 + cj * (H - 1), k0 = (1 + k) / 2 + ck * (H - 1);
    const double* cf = A.f[FX + d] + (long)c * A.P3;
    double v;
    v = +cf[placed(P, d, fi, j, k)];
    v += cf[placed(P, d, fi, j + 1, k)];
    v += cf[placed(P, d, fi, j, k + 1)];
    v += cf[placed(P, d, fi, j + 1, k + 1)];
    A.f[FX + d][(long)b * A.P3 + placed(P, d, f0, j0, k0)] = v * 0.25;
  }
}

// solution_error per node (poisson/poisson.cpp:637-668): dphi = div(force) - S on unrefined
// cells (0 on refined ones); part[b] = sum |dphi| h^3
__global__ __launch_bounds__(NTHREADS) void error_kernel(PArgs A) {
  __shared__ double red[NTHREADS / 64];
  const int b = A.list[blockIdx.x];
  const PNodeDev& nd = A.nd[b];
  const int P = A.P, M = P - 2;
  const long base = (long)b * A.P3;
  const double h1 = nd.h, h3 = pow(h1, 3.0), h1inv = 1.0 / h1;
  const double *fx = A.f[FX] + base, *fy = A.f[FY] + base, *fz = A.f[FZ] + base, *S = A.f[SRC] + base;
  double* dphi = A.f[DPHI] + base;
  double part = 0.0;
  for (int t = threadIdx.x; t < M * M * M; t += blockDim.x) {
    const int i = t % M + 1, j = (t / M) % M + 1, k = t / (M * M) + 1;
    const long e = cell(P, i, j, k);
    const bool ref = nd.child[((2 * i) / P) | (((2 * j) / P) << 1) | (((2 * k) / P) << 2)] >= 0;
    if (ref) {
      dphi[e] = 0.0;
      continue;
    }
    double df = -(fx[e + 1] - fx[e]) * h1inv;
    df -= (fy[e + P] - fy[e]) * h1inv;
    df -= (fz[e + (long)P * P] - fz[e]) * h1inv;
    dphi[e] = (df - S[e]);
    part += fabs(dphi[e]) * h3;
  }
  const double s = block_sum(part, red);
  if (threadIdx.x == 0) A.part[b] = s;
}

// Binary::compute_source_sums per node (binary/binary.cpp:985-1019): per octant the source mass
// and its centroid over unrefined cells
__global__ __launch_bounds__(NTHREADS) void source_sums_kernel(PArgs A, double* src, int* allref) {
  __shared__ double red[NTHREADS / 64];
  const int b = A.list[blockIdx.x];
  const PNodeDev& nd = A.nd[b];
  const int P = A.P, H = P / 2, G = (P - 2) / 2;
  const double h3 = nd.h * nd.h * nd.h;
  const double* S = A.f[SRC] + (long)b * A.P3;
  bool all_refined = true;
  for (int c = 0; c < 8; c++) {
    double s0 = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    if (nd.child[c] < 0) {
      all_refined = false;
      for (int t = threadIdx.x; t < G * G * G; t += blockDim.x) {
        const int i = t % G + 1 + cx * (H - 1), j = (t / G) % G + 1 + cy * (H - 1), k = t / (G * G) + 1 + cz * (H - 1);
        const double s = S[cell(P, i, j, k)];
        s0 += s * h3;
        sx += s * centre(A, b, 0, i) * h3;
        sy += s * centre(A, b, 1, j) * h3;
        sz += s * centre(A, b, 2, k) * h3;
      }
    }
    s0 = block_sum(s0, red);
    __syncthreads();
    sx = block_sum(sx, red);
    __syncthreads();
    sy = block_sum(sy, red);
    __syncthreads();
    sz = block_sum(sz, red);
    __syncthreads();
    if (threadIdx.x == 0) {
      double* o = src + (long)b * 32 + c * 4;
      o[0] = s0, o[1] = sx, o[2] = sy, o[3] = sz;
    }
  }
  if (threadIdx.x == 0) {
    for (int c = 0; c < 8; c++) {
      double* o = src + (long)b * 32 + c * 4;
      if (!all_refined && o[0] != 0.0) {
        o[1] = o[1] / o[0];
        o[2] = o[2] / o[0];
        o[3] = o[3] / o[0];
        o[0] = o[0] / (4.0 * M_PI);
      } else {
        o[0] = 0.0;
        o[1] = o[2] = o[3] = 1.0;
      }
    }
    allref[b] = all_refined ? 1 : 0;
  }
}

// PoissonPhysBound::compute_boundaries with Binary::compute_phi (binary/binary.cpp:958-983): the
// sum over every node's octant monopoles of -alpha m / r at each table entry.  32 lanes share an
// entry, each summing every 32nd node, and the partial sums are combined by a fixed butterfly, so
// the result is deterministic; it differs from the reference's recursion order only in rounding
// (all terms have one sign).
constexpr int TAB_LANES = 32;
__global__ __launch_bounds__(NTHREADS) void tables_kernel(PArgs A, int nslots) {
  const int P = A.P;
  const long t = (long)blockIdx.x * (NTHREADS / TAB_LANES) + threadIdx.x / TAB_LANES;
  const int lane = threadIdx.x % TAB_LANES;
  const bool valid = t < (long)nslots * P * P;
  double part = 0.0;
  if (valid) {
    const int slot = (int)(t / (P * P)), r = (int)(t % (P * P)), a = r / P, c = r % P;
    const int b = A.slots[slot].x, f = A.slots[slot].y, d = f >> 1, g = (f & 1) ? P - 1 : 0;
    double x, y, z;
    if (d == 0) {
      x = centre(A, b, 0, g), y = centre(A, b, 1, a), z = centre(A, b, 2, c);
    } else if (d == 1) {
      x = centre(A, b, 0, a), y = centre(A, b, 1, g), z = centre(A, b, 2, c);
    } else {
      x = centre(A, b, 0, a), y = centre(A, b, 1, c), z = centre(A, b, 2, g);
    }
    const double alpha = A.z_reflect ? 1.0 * 2.0 : 1.0;
    for (int nb = lane; nb < A.nn; nb += TAB_LANES) {
      if (A.allref[nb]) continue;  // every octant refined: the children hold the mass
      const double* s = A.src + (long)nb * 32;
#pragma unroll
      for (int o = 0; o < 8; o++) {
        const double dx = x - s[4 * o + 1], dy = y - s[4 * o + 2], dz = z - s[4 * o + 3];
        part -= alpha * s[4 * o] * rsqrt(dx * dx + dy * dy + dz * dz);
      }
    }
  }
  for (int off = TAB_LANES / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, TAB_LANES);
  if (valid && lane == 0) A.tab[t] = part;
}

// Poisson::set_source from a hydro batch: S = 4 pi G rho on the interior (binary/state_ztwd.cpp:219-222)
__global__ __launch_bounds__(NTHREADS) void source_from_state_kernel(PArgs A, const double* U, long rec, int nf,
                                                                    double coef) {
  const int b = blockIdx.x;
  const int P = A.P, M = P - 2;
  const double* rho = U + (long)b * nf * rec;
  double* S = A.f[SRC] + (long)b * A.P3;
  for (int t = threadIdx.x; t < M * M * M; t += blockDim.x) {
    const int i = t % M, j = (t / M) % M, k = t / (M * M);
    S[cell(P, i + 1, j + 1, k + 1)] = coef * rho[t];
  }
}

// Binary::pot_from_grid (binary/binary.cpp:857-872): phi = pot / rho + R^2 Omega0^2 / 2 on the
// interior, from a rotating-state batch of the same nodes
__global__ __launch_bounds__(NTHREADS) void phi_from_state_kernel(PArgs A, const double* U, long rec, int nf,
                                                                  const double* x0s, const double* hs,
                                                                  double omega0) {
  const int b = blockIdx.x;
  const int P = A.P, M = P - 2;
  const double* rho = U + (long)b * nf * rec;
  const double* pot = rho + 6 * rec;  // the pot field (binary/state_ztwd.cpp:15-28)
  double* phi = A.f[PHI] + (long)b * A.P3;
  const double h = hs[b];
  for (int t = threadIdx.x; t < M * M * M; t += blockDim.x) {
    const int i = t % M, j = (t / M) % M;
    const double x = x0s[3L * b] + (double)i * h + 0.5 * h;
    const double y = x0s[3L * b + 1] + (double)j * h + 0.5 * h;
    const double R2 = x * x + y * y;
    phi[cell(P, i + 1, j + 1, t / (M * M) + 1)] = pot[t] / rho[t] + 0.5 * R2 * omega0 * omega0;
  }
}

}  // namespace

struct amrh_poisson {
  int n = 0, P = 0, bw = 0, z_reflect = 0, nn = 0, maxlevel = 0;
  long P3 = 0;
  double grid_dim = 1.0, h0 = 0.0;
  hipStream_t stream = nullptr, own_stream = nullptr;
  std::vector<int> parent, cidx, level, pre;
  std::vector<std::array<int, 8>> child;
  std::vector<int> list_host, lv_start, lv_count, lv_parents_start, lv_parents_count;
  int all_start = 0;
  std::vector<int2> slots;
  PNodeDev* nd_dev = nullptr;
  int* list_dev = nullptr;
  int* pre_dev = nullptr;
  int2* slot_dev = nullptr;
  double* fields = nullptr;
  double* tab_dev = nullptr;
  double* src_dev = nullptr;
  int* allref_dev = nullptr;
  double* part_dev = nullptr;
  double* exps_dev = nullptr;
  LoopState* st_dev = nullptr;
  LoopState* st_host = nullptr;  // pinned
  std::vector<int> steps_last;   // steps the last loop of each (sweep, level) took
  std::vector<double> part;
  bool have_tables = false;

  PArgs args(int list_offset) const {
    PArgs a;
    a.P = P, a.bw = bw, a.z_reflect = z_reflect, a.nn = nn;
    a.P3 = P3;
    a.grid_dim = grid_dim, a.h0 = h0;
    a.nd = nd_dev;
    a.list = list_dev + list_offset;
    for (int q = 0; q < NFIELD; q++) a.f[q] = fields + (long)q * nn * P3;
    a.tab = tab_dev;
    a.part = part_dev;
    a.src = src_dev;
    a.allref = allref_dev;
    a.pre = pre_dev;
    a.slots = slot_dev;
    a.exps = exps_dev;
    a.st = st_dev;
    return a;
  }
  ~amrh_poisson() {
    for (void* p : {(void*)nd_dev, (void*)list_dev, (void*)pre_dev, (void*)slot_dev, (void*)fields, (void*)tab_dev,
                    (void*)src_dev, (void*)allref_dev, (void*)part_dev, (void*)exps_dev, (void*)st_dev})
      if (p) (void)hipFree(p);
    if (st_host) (void)hipHostFree(st_host);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }
};

namespace {

hipError_t launch_ghosts(amrh_poisson* p, int l, bool dphi) {
  if (!p->lv_count[l]) return hipSuccess;
  const PArgs A = p->args(p->lv_start[l]);
  const int passes = dphi ? 1 : 3;
  for (int pass = 0; pass < passes; pass++) {
    if (dphi)
      ghost_kernel<true><<<p->lv_count[l], NTHREADS, 0, p->stream>>>(A, pass);
    else
      ghost_kernel<false><<<p->lv_count[l], NTHREADS, 0, p->stream>>>(A, pass);
  }
  return hipGetLastError();
}

hipError_t launch_step(amrh_poisson* p, int l, int k, double limit) {
  const PArgs A = p->args(p->lv_start[l]);
  const int nb = p->lv_count[l];
  if (p->n == 8)
    relax_step_kernel<8><<<nb, NTHREADS, 0, p->stream>>>(A, k, limit, nb);
  else if (p->n == 12)
    relax_step_kernel<12><<<nb, NTHREADS, 0, p->stream>>>(A, k, limit, nb);
  else
    relax_step_kernel<16><<<nb, NTHREADS, 0, p->stream>>>(A, k, limit, nb);
  return hipGetLastError();
}

double error_sum(const amrh_poisson* p, int b) {  // solution_error's recursion
  double s = p->part[b];
  for (int c = 0; c < 8; c++)
    if (p->child[b][c] >= 0) s += error_sum(p, p->child[b][c]);
  return s;
}

int fetch_part(amrh_poisson* p) {
  hipError_t e = hipMemcpyAsync(p->part.data(), p->part_dev, sizeof(double) * p->nn, hipMemcpyDeviceToHost, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return fail(AMRH_EHIP, "poisson: %s", hipGetErrorString(e));
  return AMRH_OK;
}

#define PCHK(expr)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(AMRH_EHIP, "poisson: %s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// resid0 = relax; do { ghosts; resid = relax } while (resid / resid0 > limit): steps are enqueued
// in chunks sized by the last loop of the same (sweep, level); steps after the decision return
// at once, and the host reads the loop state once per chunk.
int relax_loop(amrh_poisson* p, int l, int sweep, double limit) {
  PCHK(hipMemsetAsync(p->st_dev, 0, sizeof(LoopState), p->stream));
  export_kernel<<<p->lv_count[l], NTHREADS, 0, p->stream>>>(p->args(p->lv_start[l]));
  PCHK(hipGetLastError());
  int& last = p->steps_last[sweep * (p->maxlevel + 1) + l];
  // the previous count (a loop's length varies little), plus the launch that sees the loop end
  int k = 0, chunk = std::max(2, last + 1);
  for (;;) {
    for (int q = 0; q < chunk; q++) PCHK(launch_step(p, l, k++, limit));
    PCHK(hipMemcpyAsync(p->st_host, p->st_dev, sizeof(LoopState), hipMemcpyDeviceToHost, p->stream));
    PCHK(hipStreamSynchronize(p->stream));
    if (p->st_host->done) break;
    if (k > 200000) return fail(AMRH_ESTATE, "poisson: relaxation at level %d does not settle", l);
    chunk = std::max(2, k / 4);  // a few more, growing if the loop keeps going
  }
  last = p->st_host->iters;
  return AMRH_OK;
}

int phi_from_children(amrh_poisson* p) {
  for (int l = p->maxlevel - 1; l >= 0; l--) {
    if (!p->lv_parents_count[l]) continue;
    restrict_kernel<<<p->lv_parents_count[l], NTHREADS, 0, p->stream>>>(p->args(p->lv_parents_start[l]));
    PCHK(hipGetLastError());
  }
  return AMRH_OK;
}

int forces_and_error(amrh_poisson* p, double* error) {
  forces_kernel<<<p->nn, NTHREADS, 0, p->stream>>>(p->args(p->all_start));
  PCHK(hipGetLastError());
  // every node of a level: a leaf next to a refined neighbour has a matched face too
  for (int l = p->maxlevel - 1; l >= 0; l--) {
    adjust_force_kernel<<<p->lv_count[l], NTHREADS, 0, p->stream>>>(p->args(p->lv_start[l]));
    PCHK(hipGetLastError());
  }
  if (!error) return AMRH_OK;
  error_kernel<<<p->nn, NTHREADS, 0, p->stream>>>(p->args(p->all_start));
  PCHK(hipGetLastError());
  if (int rc = fetch_part(p)) return rc;
  *error = error_sum(p, 0);
  return AMRH_OK;
}

// Poisson::vcycle0 (poisson/poisson.cpp:57-84)
int vcycle0(amrh_poisson* p) {
  if (int rc = phi_from_children(p)) return rc;
  for (int l = 0; l <= p->maxlevel; l++) PCHK(launch_ghosts(p, l, false));
  PCHK(hipMemsetAsync(p->fields + (long)DPHI * p->nn * p->P3, 0, sizeof(double) * 2 * p->nn * p->P3, p->stream));
  for (int l = p->maxlevel; l >= 0; l--) {
    PCHK(launch_ghosts(p, l, true));
    down_kernel<<<p->lv_count[l], NTHREADS, 0, p->stream>>>(p->args(p->lv_start[l]));
    PCHK(hipGetLastError());
    if (int rc = relax_loop(p, l, 0, VDOWN_RESID)) return rc;
  }
  for (int l = 0; l <= p->maxlevel; l++) {
    PCHK(launch_ghosts(p, l, true));
    if (l > 0) {
      coarse_kernel<<<p->lv_count[l], NTHREADS, 0, p->stream>>>(p->args(p->lv_start[l]));
      PCHK(hipGetLastError());
    }
    if (int rc = relax_loop(p, l, 1, VUP_RESID)) return rc;
    interior_kernel<<<p->lv_count[l], NTHREADS, 0, p->stream>>>(p->args(p->lv_start[l]), 3);
    PCHK(hipGetLastError());
    PCHK(launch_ghosts(p, l, false));
  }
  return AMRH_OK;
}

}  // namespace

extern "C" {

int amrh_poisson_create(int n, int bw, double grid_dim, int z_reflect, int nnodes, const int* parent,
                        const int* child, void* stream, amrh_poisson** out) {
  if (!parent || !child || !out) return fail(AMRH_EINVAL, "null argument");
  *out = nullptr;
  if (n != 8 && n != 12 && n != 16) return fail(AMRH_EINVAL, "poisson: n = 8, 12 or 16");
  if (bw < 1 || bw > 8) return fail(AMRH_EINVAL, "poisson: ghost width %d", bw);
  if (nnodes < 1 || parent[0] != -1) return fail(AMRH_EINVAL, "poisson: node 0 must be the root");
  if (!(grid_dim > 0.0)) return fail(AMRH_EINVAL, "poisson: grid_dim must be positive");
  amrh_poisson* p = new amrh_poisson;
  p->n = n, p->P = n + 2, p->bw = bw, p->z_reflect = z_reflect ? 1 : 0, p->nn = nnodes;
  p->P3 = (long)p->P * p->P * p->P;
  p->grid_dim = grid_dim;
  p->h0 = 2.0 * grid_dim / double(n);  // grid/grid.cpp:70
  p->parent.assign(parent, parent + nnodes);
  p->cidx.assign(nnodes, 0);
  p->level.assign(nnodes, 0);
  p->child.assign(nnodes, {-1, -1, -1, -1, -1, -1, -1, -1});
  std::vector<PNodeDev> nd(nnodes);
  std::vector<std::array<int, 3>> loc(nnodes, {0, 0, 0});
  std::map<std::tuple<int, int, int, int>, int> by_loc;
  const int GNX = n + 2 * bw;
  for (int b = 0; b < nnodes; b++) {
    PNodeDev& q = nd[b];
    memset(&q, 0, sizeof(q));
    for (int c = 0; c < 8; c++) q.child[c] = -1;
    q.parent = -1;
    if (b == 0) {
      q.h = p->h0;
    } else {
      const int pa = parent[b], c = child[b];
      if (pa < 0 || pa >= b || c < 0 || c > 7 || p->child[pa][c] >= 0) {
        delete p;
        return fail(AMRH_EINVAL, "poisson: node %d: bad parent %d / child index %d", b, pa, c);
      }
      p->child[pa][c] = b;
      nd[pa].child[c] = b;
      q.parent = pa;
      q.cidx = c;
      p->cidx[b] = c;
      q.level = p->level[b] = p->level[pa] + 1;
      if (q.level >= MAXDEPTH) {
        delete p;
        return fail(AMRH_EINVAL, "poisson: more than %d levels", MAXDEPTH - 1);
      }
      q.h = nd[pa].h * 0.5;
      const int cv[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
      for (int d = 0; d < 3; d++) {
        loc[b][d] = 2 * loc[pa][d] + cv[d];
        q.off[d] = 2 * nd[pa].off[d] + bw + cv[d] * (GNX - 2 * bw);  // grid/grid_node.cpp:275
      }
      p->maxlevel = std::max(p->maxlevel, q.level);
    }
    q.level = p->level[b];
    by_loc[std::make_tuple(q.level, loc[b][0], loc[b][1], loc[b][2])] = b;
  }
  // what sits across each face: a node, a boundary table (reflecting on z low), or the parent
  for (int b = 0; b < nnodes; b++)
    for (int f = 0; f < 6; f++) {
      const int d = f >> 1, lv = p->level[b];
      std::array<int, 3> l = loc[b];
      l[d] += (f & 1) ? 1 : -1;
      PNodeDev& q = nd[b];
      if (l[d] < 0 || l[d] >= (1 << lv)) {
        if (p->z_reflect && f == 4) {
          q.fkind[f] = F_REFL, q.fnode[f] = -1;
        } else {
          q.fkind[f] = F_TAB, q.fnode[f] = (int)p->slots.size();
          p->slots.push_back(make_int2(b, f));
        }
        continue;
      }
      auto it = by_loc.find(std::make_tuple(lv, l[0], l[1], l[2]));
      if (it != by_loc.end()) {
        q.fkind[f] = F_SIB, q.fnode[f] = it->second;
      } else {
        q.fkind[f] = F_AMR, q.fnode[f] = q.parent;
      }
    }
  // launch lists: per level in pre-order; per level the nodes with children; all nodes
  std::vector<int> pre;
  std::vector<int> stack = {0};
  while (!stack.empty()) {
    const int b = stack.back();
    stack.pop_back();
    pre.push_back(b);
    for (int c = 7; c >= 0; c--)
      if (p->child[b][c] >= 0) stack.push_back(p->child[b][c]);
  }
  p->pre = pre;
  const int L = p->maxlevel + 1;
  p->lv_start.assign(L, 0), p->lv_count.assign(L, 0);
  p->lv_parents_start.assign(L, 0), p->lv_parents_count.assign(L, 0);
  for (int l = 0; l < L; l++) {
    p->lv_start[l] = (int)p->list_host.size();
    for (int b : pre)
      if (p->level[b] == l) p->list_host.push_back(b);
    p->lv_count[l] = (int)p->list_host.size() - p->lv_start[l];
    p->lv_parents_start[l] = (int)p->list_host.size();
    for (int b : pre)
      if (p->level[b] == l && std::any_of(p->child[b].begin(), p->child[b].end(), [](int c) { return c >= 0; }))
        p->list_host.push_back(b);
    p->lv_parents_count[l] = (int)p->list_host.size() - p->lv_parents_start[l];
  }
  p->all_start = (int)p->list_host.size();
  for (int b = 0; b < nnodes; b++) p->list_host.push_back(b);
  p->part.assign(nnodes, 0.0);
  hipError_t e = hipSuccess;
  auto alloc = [&](void** ptr, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(ptr, std::max<size_t>(bytes, 8));
  };
  alloc((void**)&p->nd_dev, sizeof(PNodeDev) * nnodes);
  alloc((void**)&p->list_dev, sizeof(int) * p->list_host.size());
  alloc((void**)&p->pre_dev, sizeof(int) * nnodes);
  alloc((void**)&p->slot_dev, sizeof(int2) * p->slots.size());
  alloc((void**)&p->fields, sizeof(double) * NFIELD * nnodes * p->P3);
  alloc((void**)&p->tab_dev, sizeof(double) * p->slots.size() * p->P * p->P);
  alloc((void**)&p->src_dev, sizeof(double) * 32 * nnodes);
  alloc((void**)&p->allref_dev, sizeof(int) * nnodes);
  alloc((void**)&p->part_dev, sizeof(double) * 2 * nnodes);  // two copies for the relaxation loops
  alloc((void**)&p->exps_dev, sizeof(double) * 2 * nnodes * 6 * n * n);
  alloc((void**)&p->st_dev, sizeof(LoopState));
  if (e == hipSuccess) e = hipHostMalloc((void**)&p->st_host, sizeof(LoopState), hipHostMallocDefault);
  p->steps_last.assign(2 * (p->maxlevel + 1), 2);
  if (e == hipSuccess) e = hipMemcpy(p->nd_dev, nd.data(), sizeof(PNodeDev) * nnodes, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->list_dev, p->list_host.data(), sizeof(int) * p->list_host.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->pre_dev, pre.data(), sizeof(int) * nnodes, hipMemcpyHostToDevice);
  if (e == hipSuccess && !p->slots.empty())
    e = hipMemcpy(p->slot_dev, p->slots.data(), sizeof(int2) * p->slots.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p->fields, 0, sizeof(double) * NFIELD * nnodes * p->P3);
  if (e == hipSuccess && !p->slots.empty())
    e = hipMemset(p->tab_dev, 0, sizeof(double) * p->slots.size() * p->P * p->P);
  if (e == hipSuccess) {
    if (stream) {
      p->stream = (hipStream_t)stream;
    } else {
      e = hipStreamCreateWithFlags(&p->own_stream, hipStreamNonBlocking);
      p->stream = p->own_stream;
    }
  }
  if (e != hipSuccess) {
    delete p;
    return fail(AMRH_EHIP, "poisson create: %s", hipGetErrorString(e));
  }
  *out = p;
  return AMRH_OK;
}

void amrh_poisson_destroy(amrh_poisson* p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  delete p;
}

static int field_index(int which) { return which >= 0 && which < NFIELD ? which : -1; }

int amrh_poisson_upload(amrh_poisson* p, int field, const double* host) {
  if (!p || !host) return fail(AMRH_EINVAL, "null argument");
  const int q = field_index(field);
  if (q < 0) return fail(AMRH_EINVAL, "poisson: field %d", field);
  PCHK(hipMemcpyAsync(p->fields + (long)q * p->nn * p->P3, host, sizeof(double) * p->nn * p->P3,
                      hipMemcpyHostToDevice, p->stream));
  PCHK(hipStreamSynchronize(p->stream));
  return AMRH_OK;
}

int amrh_poisson_download(amrh_poisson* p, int field, double* host) {
  if (!p || !host) return fail(AMRH_EINVAL, "null argument");
  const int q = field_index(field);
  if (q < 0) return fail(AMRH_EINVAL, "poisson: field %d", field);
  PCHK(hipMemcpyAsync(host, p->fields + (long)q * p->nn * p->P3, sizeof(double) * p->nn * p->P3,
                      hipMemcpyDeviceToHost, p->stream));
  PCHK(hipStreamSynchronize(p->stream));
  return AMRH_OK;
}

const void* amrh_poisson_phi_device(const amrh_poisson* p) { return p ? p->fields + (long)PHI * p->nn * p->P3 : nullptr; }
const void* amrh_poisson_forces_device(const amrh_poisson* p) { return p ? p->fields + (long)FX * p->nn * p->P3 : nullptr; }

int amrh_poisson_source_from_state(amrh_poisson* p, const amrh_ctx* c, double G) {
  if (!p || !c) return fail(AMRH_EINVAL, "null argument");
  if (!c->have_u) return fail(AMRH_ESTATE, "the hydro context holds no state");
  if (c->nb != p->nn || c->cfg.n != p->n) return fail(AMRH_EINVAL, "poisson: the hydro batch must hold the same nodes");
  const double coef = 4.0 * M_PI * G;
  PCHK(hipStreamSynchronize(c->stream));  // the hydro stream may still be writing the state
  source_from_state_kernel<<<p->nn, NTHREADS, 0, p->stream>>>(p->args(p->all_start), c->buf[c->cur], c->L.rec(),
                                                              c->cfg.nf, coef);
  PCHK(hipGetLastError());
  PCHK(hipStreamSynchronize(p->stream));
  return AMRH_OK;
}

int amrh_poisson_phi_from_state(amrh_poisson* p, const amrh_ctx* c) {
  if (!p || !c) return fail(AMRH_EINVAL, "null argument");
  if (c->cfg.physics != AMRH_PHYS_ROTATING) return fail(AMRH_EINVAL, "phi from pot needs the rotating-frame state");
  if (!c->have_u || !c->have_x0) return fail(AMRH_ESTATE, "the hydro context holds no state or positions");
  if (c->nb != p->nn || c->cfg.n != p->n) return fail(AMRH_EINVAL, "poisson: the hydro batch must hold the same nodes");
  PCHK(hipStreamSynchronize(c->stream));
  phi_from_state_kernel<<<p->nn, NTHREADS, 0, p->stream>>>(p->args(p->all_start), c->buf[c->cur], c->L.rec(),
                                                            c->cfg.nf, c->x0_dev, c->h_dev, c->cfg.omega0);
  PCHK(hipGetLastError());
  PCHK(hipStreamSynchronize(p->stream));
  // the root's enforce_phi_boundaries for every level (binary/binary.cpp:878-883), once the
  // physical faces have their tables; before that the solve's own fill follows
  return p->have_tables ? amrh_poisson_phi_boundaries(p) : AMRH_OK;
}

int amrh_poisson_prepare(amrh_poisson* p) {
  if (!p) return fail(AMRH_EINVAL, "null argument");
  source_sums_kernel<<<p->nn, NTHREADS, 0, p->stream>>>(p->args(p->all_start), p->src_dev, p->allref_dev);
  PCHK(hipGetLastError());
  const long entries = (long)p->slots.size() * p->P * p->P;
  if (entries) {
    const long per_block = NTHREADS / TAB_LANES;
    tables_kernel<<<(unsigned)((entries + per_block - 1) / per_block), NTHREADS, 0, p->stream>>>(
        p->args(p->all_start), (int)p->slots.size());
    PCHK(hipGetLastError());
  }
  PCHK(hipStreamSynchronize(p->stream));
  p->have_tables = true;
  return AMRH_OK;
}

int amrh_poisson_boundary_tables(amrh_poisson* p, double* tab, int write) {
  if (!p || !tab) return fail(AMRH_EINVAL, "null argument");
  const long PP = (long)p->P * p->P;
  std::vector<double> flat(p->slots.size() * PP);
  if (write) {
    for (size_t s = 0; s < p->slots.size(); s++)
      memcpy(&flat[s * PP], tab + ((long)p->slots[s].x * 6 + p->slots[s].y) * PP, sizeof(double) * PP);
    if (!flat.empty())
      PCHK(hipMemcpyAsync(p->tab_dev, flat.data(), sizeof(double) * flat.size(), hipMemcpyHostToDevice, p->stream));
    PCHK(hipStreamSynchronize(p->stream));
    p->have_tables = true;
    return AMRH_OK;
  }
  if (!flat.empty())
    PCHK(hipMemcpyAsync(flat.data(), p->tab_dev, sizeof(double) * flat.size(), hipMemcpyDeviceToHost, p->stream));
  PCHK(hipStreamSynchronize(p->stream));
  memset(tab, 0, sizeof(double) * p->nn * 6 * PP);
  for (size_t s = 0; s < p->slots.size(); s++)
    memcpy(tab + ((long)p->slots[s].x * 6 + p->slots[s].y) * PP, &flat[s * PP], sizeof(double) * PP);
  return AMRH_OK;
}

int amrh_poisson_phi_from_children(amrh_poisson* p) {
  if (!p) return fail(AMRH_EINVAL, "null argument");
  if (int rc = phi_from_children(p)) return rc;
  PCHK(hipStreamSynchronize(p->stream));
  return AMRH_OK;
}

int amrh_poisson_phi_boundaries(amrh_poisson* p) {
  if (!p) return fail(AMRH_EINVAL, "null argument");
  if (!p->have_tables) return fail(AMRH_ESTATE, "poisson: boundary tables not set (amrh_poisson_prepare)");
  for (int l = 0; l <= p->maxlevel; l++) PCHK(launch_ghosts(p, l, false));
  PCHK(hipStreamSynchronize(p->stream));
  return AMRH_OK;
}

int amrh_poisson_forces(amrh_poisson* p, double* error) {
  if (!p) return fail(AMRH_EINVAL, "null argument");
  double err = 0.0;
  if (int rc = forces_and_error(p, &err)) return rc;
  if (error) *error = err;
  return AMRH_OK;
}

int amrh_poisson_vcycle(amrh_poisson* p, double* error) {
  if (!p) return fail(AMRH_EINVAL, "null argument");
  if (!p->have_tables) return fail(AMRH_ESTATE, "poisson: boundary tables not set (amrh_poisson_prepare)");
  if (int rc = vcycle0(p)) return rc;
  double err = 0.0;
  if (int rc = forces_and_error(p, &err)) return rc;
  if (error) *error = err;
  return AMRH_OK;
}

int amrh_poisson_solve(amrh_poisson* p, double toler, int max_iter, int prepare, int* iters, double* error) {
  if (!p) return fail(AMRH_EINVAL, "null argument");
  if (prepare) {
    if (int rc = amrh_poisson_prepare(p)) return rc;
  } else if (!p->have_tables) {
    return fail(AMRH_ESTATE, "poisson: boundary tables not set (amrh_poisson_prepare)");
  }
  if (int rc = forces_and_error(p, nullptr)) return rc;
  int it = 0;
  double err = 0.0;
  do {
    if (it > max_iter) return fail(AMRH_ESTATE, "poisson: no convergence in %d vcycles (error %g)", max_iter, err);
    if (int rc = amrh_poisson_vcycle(p, &err)) return rc;
    it++;
  } while (err > toler);
  if (iters) *iters = it;
  if (error) *error = err;
  return AMRH_OK;
}

int amrh_poisson_store_phi(amrh_poisson* p, int mode) {
  if (!p) return fail(AMRH_EINVAL, "null argument");
  if (mode < 0 || mode > 2) return fail(AMRH_EINVAL, "poisson: store mode %d", mode);
  interior_kernel<<<p->nn, NTHREADS, 0, p->stream>>>(p->args(p->all_start), mode);
  PCHK(hipGetLastError());
  PCHK(hipStreamSynchronize(p->stream));
  return AMRH_OK;
}

}  // extern "C"
