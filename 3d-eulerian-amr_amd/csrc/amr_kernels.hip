// Octree coupling on the device: ghost fill, reflux export, restriction, flow-off ledger.
//
//   tree_ghost_kernel    GridNode::enforce_boundaries (grid/grid_node.cpp:59-112) for one level:
//                        same-level sibling copy, coarse->fine interpolation from the parent
//                        (grid/grid_amr_bound.cpp:27-78) or physical outflow / z-mirror
//                        (grid/grid_phys_bound.cpp:27-54)
//   reflux_export_kernel the fine side of GridNode::adjust_{x,y,z}_flux (grid/grid_node.cpp:
//                        1076-1274): boundary face fluxes of a fine block, 2x2 means written to
//                        the coarse block's override slot
//   restrict_kernel      GridNode::inject_from_children (grid/grid_node.cpp:1035-1074)
//   ledger_kernel        Grid::sum_{x,y,z}_flow_off + the FO update of Grid::add_difs
//                        (grid/grid.cpp:350-404, 419)
#include <hip/hip_runtime.h>

#include "amr_kernels.h"
#include "amrh_layout.h"
#include "euler_phys.h"

namespace amrh {

namespace {

constexpr int GW = 2;

template <bool MIXED>
struct PT {
  using T = double;
};
template <>
struct PT<true> {
  using T = float;
};

template <bool MIXED>
__device__ __forceinline__ void prim_of(const double* u, typename PT<MIXED>::T* p) {
  if constexpr (!MIXED) {
#pragma unroll
    for (int l = 0; l < NFE; l++) p[l] = u[l];
    x64::to_prim(p);
  } else {
    const double inv = 1.0 / u[RHO];
    const double vx = u[SX] * inv, vy = u[SY] * inv, vz = u[SZ] * inv;
    const double ekin = 0.5 * (u[SX] * vx + u[SY] * vy + u[SZ] * vz);
    p[RHO] = (float)u[RHO];
    p[SX] = (float)vx;
    p[SY] = (float)vy;
    p[SZ] = (float)vz;
    p[ET] = (float)(u[ET] - ekin);
    p[TAU] = (float)u[TAU];
  }
}

// decode a slab element r of face `face` into the ghost's padded coordinates
template <int N, int G>
__device__ __forceinline__ void slab_coords(int face, int r, int* p, int* t) {
  if (face < 2) {
    *t = r % G;
    p[0] = (face & 1) ? N + G + *t : *t;
    p[1] = (r / G) % N + G;
    p[2] = r / (G * N) + G;
  } else if (face < 4) {
    *t = (r / N) % G;
    p[0] = r % N + G;
    p[1] = (face & 1) ? N + G + *t : *t;
    p[2] = r / (G * N) + G;
  } else {
    *t = r / (N * N);
    p[0] = r % N + G;
    p[1] = (r / N) % N + G;
    p[2] = (face & 1) ? N + G + *t : *t;
  }
}

template <int N, int G>
__global__ __launch_bounds__(256) void tree_ghost_kernel(double* U, const int* nodes, int count,
                                                         const int* ghost, const int* rel, int z_reflect,
                                                         int* err) {
  constexpr Layout L{N, G, NFE};
  constexpr long R = L.rec();
  constexpr int SC = G * N * N;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)count * 6 * SC) return;
  const int q = (int)(gid / (6 * SC));
  const int rem = (int)(gid % (6 * SC));
  const int face = rem / SC, r = rem % SC;
  const int node = nodes[q];
  const int type = ghost[(node * 6 + face) * 2];
  const int src = ghost[(node * 6 + face) * 2 + 1];
  const int d = face >> 1;
  const bool upper = face & 1;
  int p[3], t;
  slab_coords<N, G>(face, r, p, &t);
  double u[NFE];
  double* Un = U + (long)node * NFE * R;
  if (type == GHOST_SIB) {
    // sibling's interior: lower face reads its cells N-G+t, upper face its cells t
    int c[3] = {p[0] - G, p[1] - G, p[2] - G};
    c[d] = upper ? t : N - G + t;
    const double* Us = U + (long)src * NFE * R;
    const long e = L.interior(c[0], c[1], c[2]);
#pragma unroll
    for (int l = 0; l < NFE; l++) u[l] = Us[l * R + e];
  } else if (type == GHOST_PHYS) {
    int c[3] = {p[0] - G, p[1] - G, p[2] - G};
    const bool reflect = z_reflect && face == ZL;
    c[d] = reflect ? G - 1 - t : (upper ? N - 1 : 0);
    const long e = L.interior(c[0], c[1], c[2]);
#pragma unroll
    for (int l = 0; l < NFE; l++) u[l] = Un[l * R + e];
    if (reflect) {
      u[SZ] = -u[SZ];
    } else {
      const double v = (d == 2) ? u[SZ] : u[SX + d] / u[RHO];
      const bool inflow = upper ? (v < 0.0) : (v > 0.0);
      if (inflow) {
        const double s = u[SX + d];
        u[ET] = u[ET] - 0.5 * s * s / u[RHO];
        u[SX + d] = 0.0;
      }
    }
  } else {
    // coarse->fine: parent cell i1 = (rel + p)/2; normal slope minmod, sign by parity
    const double* Up = U + (long)src * NFE * R;
    int i1[3], par = 0;
#pragma unroll
    for (int e = 0; e < 3; e++) {
      const int f = rel[node * 3 + e] + p[e];
      i1[e] = f >> 1;
      if (e == d) par = f & 1;
    }
    int ip[3] = {i1[0], i1[1], i1[2]}, im[3] = {i1[0], i1[1], i1[2]};
    ip[d] += 1;
    im[d] -= 1;
    const long e0 = L.from_padded(i1[0], i1[1], i1[2]);
    const long ep = L.from_padded(ip[0], ip[1], ip[2]);
    const long em = L.from_padded(im[0], im[1], im[2]);
    if (e0 < 0 || ep < 0 || em < 0) {
      atomicOr(err, 2);
      return;
    }
#pragma unroll
    for (int l = 0; l < NFE; l++) {
      const double uc = Up[l * R + e0];
      double m = x64::minmod2(Up[l * R + ep] - uc, uc - Up[l * R + em]);
      if (par) m = 0.0 - m;
      u[l] = uc - m * 0.25;
    }
  }
  const long dst = L.slab(face) + r;
#pragma unroll
  for (int l = 0; l < NFE; l++) Un[l * R + dst] = u[l];
}

template <bool MIXED, int D>
__device__ __forceinline__ void face_from4(const PhysParams& P, const typename PT<MIXED>::T (*q)[NFE],
                                           typename PT<MIXED>::T* F) {
  using T = typename PT<MIXED>::T;
  T sl[NFE], sr[NFE];
  if constexpr (!MIXED) {
    x64::slope(q[0], q[1], q[2], sl, P.theta);
    x64::slope(q[1], q[2], q[3], sr, P.theta);
    double e;
    x64::kt_face<D>(P, q[1], sl, q[2], sr, F, e);
  } else {
    const f32::PhysF PF{(float)P.gamma, (float)(P.gamma - 1.0), (float)P.theta};
    f32::slope(q[0], q[1], q[2], sl, PF.theta);
    f32::slope(q[1], q[2], q[3], sr, PF.theta);
    float e;
    f32::kt_face<D>(PF, q[1], sl, q[2], sr, F, e);
  }
}

// One workgroup per export item: fine block C's boundary face plane (direction d, side), the n^2
// face fluxes, C's own overrides applied (finer levels were exported first), then 2x2 means.
template <int N, bool MIXED>
__global__ __launch_bounds__(N* N) void reflux_export_kernel(const double* U, const ExportItem* items,
                                                             const int* slot, double* ovr, PhysParams P) {
  using T = typename PT<MIXED>::T;
  constexpr Layout L{N, GW, NFE};
  constexpr long R = L.rec();
  __shared__ double Fs[N][N][NFE];
  const ExportItem it = items[blockIdx.x];
  const int ta = threadIdx.x % N, tb = threadIdx.x / N;
  const int d = it.dir;
  const double* Uc = U + (long)it.fine * NFE * R;
  T q[4][NFE];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int pn = it.side ? N + m : m;  // padded position along d
    int pc[3];
    if (d == 0) { pc[0] = pn; pc[1] = ta + GW; pc[2] = tb + GW; }
    else if (d == 1) { pc[0] = ta + GW; pc[1] = pn; pc[2] = tb + GW; }
    else { pc[0] = ta + GW; pc[1] = tb + GW; pc[2] = pn; }
    const long e = L.from_padded(pc[0], pc[1], pc[2]);
    double u[NFE];
#pragma unroll
    for (int l = 0; l < NFE; l++) u[l] = Uc[l * R + e];
    prim_of<MIXED>(u, q[m]);
  }
  T F[NFE];
  if (d == 0)
    face_from4<MIXED, 0>(P, q, F);
  else if (d == 1)
    face_from4<MIXED, 1>(P, q, F);
  else
    face_from4<MIXED, 2>(P, q, F);
  // the fine block's own reflux at this face (plane 0 or N), from an even finer level
  {
    const int pl = it.side ? 2 : 0;
    const int qa = ta >= N / 2, qb = tb >= N / 2;
    const int s = slot[((it.fine * 3 + d) * 3 + pl) * 4 + qa + 2 * qb];
    if (s >= 0) {
      const double* o = ovr + (((long)s * (N / 2) + (tb - qb * (N / 2))) * (N / 2) + (ta - qa * (N / 2))) * NFE;
#pragma unroll
      for (int l = 0; l < NFE; l++) F[l] = (T)o[l];
    }
  }
#pragma unroll
  for (int l = 0; l < NFE; l++) Fs[tb][ta][l] = (double)F[l];
  __syncthreads();
  if (threadIdx.x < (N / 2) * (N / 2)) {
    const int al = threadIdx.x % (N / 2), bl = threadIdx.x / (N / 2);
    const int a0 = 2 * al, b0 = 2 * bl;
    double* o = ovr + (((long)it.slot * (N / 2) + bl) * (N / 2) + al) * NFE;
#pragma unroll
    for (int l = 0; l < NFE; l++) {
      double v = Fs[b0][a0][l];
      v += Fs[b0][a0 + 1][l];
      v += Fs[b0 + 1][a0][l];
      v += Fs[b0 + 1][a0 + 1][l];
      v *= 0.25;
      o[l] = v;
    }
  }
}

template <int N, int G>
__global__ __launch_bounds__(256) void restrict_kernel(double* U, const int2* pairs, int count, PhysParams P) {
  constexpr Layout L{N, G, NFE};
  constexpr long R = L.rec();
  constexpr int H = N / 2;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)count * H * H * H) return;
  const int q = (int)(gid / (H * H * H));
  const int r = (int)(gid % (H * H * H));
  const int2 pc = pairs[q];  // (parent, child<<24 | child-node)
  const int parent = pc.x, c = (pc.y >> 24) & 7, child = pc.y & 0xffffff;
  const int ii = r % H, jj = (r / H) % H, kk = r / (H * H);
  const int i = 2 * ii, j = 2 * jj, k = 2 * kk;
  const double* Uc = U + (long)child * NFE * R;
  double u[NFE];
#pragma unroll
  for (int l = 0; l < NFE; l++) {
    const double* f = Uc + l * R;
    double s = 0.0;
    s += f[L.interior(i + 0, j + 0, k + 0)];
    s += f[L.interior(i + 1, j + 0, k + 0)];
    s += f[L.interior(i + 0, j + 1, k + 0)];
    s += f[L.interior(i + 1, j + 1, k + 0)];
    s += f[L.interior(i + 0, j + 0, k + 1)];
    s += f[L.interior(i + 1, j + 0, k + 1)];
    s += f[L.interior(i + 0, j + 1, k + 1)];
    s += f[L.interior(i + 1, j + 1, k + 1)];
    s *= 0.125;
    u[l] = s;
  }
  x64::floor_state(P, u);
  const int ci = (c & 1) * H + ii, cj = ((c >> 1) & 1) * H + jj, ck = ((c >> 2) & 1) * H + kk;
  double* Up = U + (long)parent * NFE * R;
  const long e = L.interior(ci, cj, ck);
#pragma unroll
  for (int l = 0; l < NFE; l++) Up[l * R + e] = u[l];
}

// The boundary terms tmp = (F_up - F_lo) h^2 (the upper face only for z with z_reflect) are
// formed by all threads into LDS; one thread per field then adds them in the reference's order
// (direction, then pencil row b, then pencil a), so the ledger is bit-identical to the serial sum.
template <int N, int NFL>
__global__ __launch_bounds__(256) void ledger_kernel(const double* ledger, double h, const double* dt_dev, double dt,
                                                     double beta, int z_reflect, double* fo /* [FO | FO0] */) {
  constexpr int TERMS = 3 * N * N;
  __shared__ double tmp[TERMS * NFL];
  for (int t = threadIdx.x; t < TERMS * NFL; t += blockDim.x) {
    const int q = t / NFL, l = t % NFL;  // q = (d * N + b) * N + a
    const int d = q / (N * N), ba = q % (N * N);
    const double lo = ledger[((d * 2 + 0) * N * N + ba) * NFL + l];
    const double up = ledger[((d * 2 + 1) * N * N + ba) * NFL + l];
    tmp[t] = (d == 2 && z_reflect) ? up * h * h : (up - lo) * h * h;
  }
  __syncthreads();
  if (threadIdx.x >= NFL) return;
  if (dt_dev) dt = *dt_dev;
  const int l = threadIdx.x;
  double dfo = 0.0;
  for (int q = 0; q < TERMS; q++) dfo += tmp[q * NFL + l];
  fo[l] = (fo[l] + dfo * dt) * beta + fo[NFL + l] * (1.0 - beta);
}

__global__ void ledger_store_kernel(double* fo, int nf) {
  if (threadIdx.x < nf) fo[nf + threadIdx.x] = fo[threadIdx.x];
}

__global__ void ledger_restore_kernel(double* fo, int nf) {
  if (threadIdx.x < nf) fo[threadIdx.x] = fo[nf + threadIdx.x];
}

}  // namespace

hipError_t launch_tree_ghosts(int n, int g, double* U, const int* nodes, int count, const int* ghost, const int* rel,
                              int z_reflect, int* err, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const long total = (long)count * 6 * g * n * n;
  const unsigned grid = (unsigned)((total + 255) / 256);
  if (n == 8 && g == 2)
    tree_ghost_kernel<8, 2><<<grid, 256, 0, s>>>(U, nodes, count, ghost, rel, z_reflect, err);
  else if (n == 8 && g == 3)
    tree_ghost_kernel<8, 3><<<grid, 256, 0, s>>>(U, nodes, count, ghost, rel, z_reflect, err);
  else if (n == 16 && g == 2)
    tree_ghost_kernel<16, 2><<<grid, 256, 0, s>>>(U, nodes, count, ghost, rel, z_reflect, err);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_reflux_export(int n, bool mixed, const double* U, const ExportItem* items, int count,
                                const int* slot, double* ovr, const PhysParams& P, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (n == 8) {
    if (mixed)
      reflux_export_kernel<8, true><<<count, 64, 0, s>>>(U, items, slot, ovr, P);
    else
      reflux_export_kernel<8, false><<<count, 64, 0, s>>>(U, items, slot, ovr, P);
  } else if (n == 16) {
    if (mixed)
      reflux_export_kernel<16, true><<<count, 256, 0, s>>>(U, items, slot, ovr, P);
    else
      reflux_export_kernel<16, false><<<count, 256, 0, s>>>(U, items, slot, ovr, P);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_restrict(int n, int g, double* U, const int2* pairs, int count, const PhysParams& P,
                           hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const long total = (long)count * (n / 2) * (n / 2) * (n / 2);
  const unsigned grid = (unsigned)((total + 255) / 256);
  if (n == 8 && g == 2)
    restrict_kernel<8, 2><<<grid, 256, 0, s>>>(U, pairs, count, P);
  else if (n == 8 && g == 3)
    restrict_kernel<8, 3><<<grid, 256, 0, s>>>(U, pairs, count, P);
  else if (n == 16 && g == 2)
    restrict_kernel<16, 2><<<grid, 256, 0, s>>>(U, pairs, count, P);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_ledger(int n, int nf, const double* ledger, double h, const double* dt_dev, double dt, double beta,
                         int z_reflect, double* fo, hipStream_t s) {
  if (n == 8 && nf == NFE)
    ledger_kernel<8, NFE><<<1, 256, 0, s>>>(ledger, h, dt_dev, dt, beta, z_reflect, fo);
  else if (n == 16 && nf == NFE)
    ledger_kernel<16, NFE><<<1, 256, 0, s>>>(ledger, h, dt_dev, dt, beta, z_reflect, fo);
  else if (n == 8 && nf == 13)
    ledger_kernel<8, 13><<<1, 256, 0, s>>>(ledger, h, dt_dev, dt, beta, z_reflect, fo);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_ledger_store(double* fo, int nf, hipStream_t s) {
  ledger_store_kernel<<<1, 64, 0, s>>>(fo, nf);
  return hipGetLastError();
}

hipError_t launch_ledger_restore(double* fo, int nf, hipStream_t s) {
  ledger_restore_kernel<<<1, 64, 0, s>>>(fo, nf);
  return hipGetLastError();
}

}  // namespace amrh
