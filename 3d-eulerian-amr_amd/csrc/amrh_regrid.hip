// Re-packing a batch after the tree changes (GridNode::check_for_refine, grid/grid_node.cpp:957-1033):
// surviving nodes' records are gathered from the old batch, and every node created by refinement
// is filled from its parent by GridNode::inject_from_parent (grid/grid_node.cpp:193-233), level by
// level; plus the density test of GridNode::set_refine_flags (grid/grid_node.cpp:647-667) as a
// device reduction.
#include <hip/hip_runtime.h>

#include <vector>

#include "amrh_ctx.h"
#include "amrh_internal.h"
#include "amrh_layout.h"
#include "amrhydro.h"
#include "euler_phys.h"

using namespace amrh;

namespace {

// copy whole block records: dst block b <- src block idx[b] (idx < 0: left alone)
__global__ __launch_bounds__(256) void gather_kernel(double* dst, const double* src, const int* idx, int nb,
                                                     long block_elems) {
  for (int b = blockIdx.y; b < nb; b += gridDim.y) {
    const int s = idx[b];
    if (s < 0) continue;
    const double* from = src + (long)s * block_elems;
    double* to = dst + (long)b * block_elems;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < block_elems; e += (long)gridDim.x * blockDim.x)
      to[e] = from[e];
  }
}

// inject_from_parent for the nodes in `nodes` (node, parent, child index packed as int3): one thread
// per 2x2x2 group of child cells, all fields
template <int N, int GW>
__global__ __launch_bounds__(64) void prolong_kernel(double* U, const int3* nodes, int count, PhysParams P) {
  constexpr Layout L{N, GW, NFE};
  constexpr long R = L.rec();
  constexpr int H = N / 2, G = N + 2 * GW;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)count * H * H * H) return;
  const int q = (int)(gid / (H * H * H)), r = (int)(gid % (H * H * H));
  const int3 t = nodes[q];
  const int node = t.x, parent = t.y, c = t.z;
  const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  const int ii = r % H, jj = (r / H) % H, kk = r / (H * H);
  // child padded cell (i, j, k) = (GW + 2 ii, ...); parent padded cell of grid/grid_node.cpp:204-206
  const int i = GW + 2 * ii, j = GW + 2 * jj, k = GW + 2 * kk;
  const int i0 = (GW + i) / 2 + cx * (G / 2 - GW);
  const int j0 = (GW + j) / 2 + cy * (G / 2 - GW);
  const int k0 = (GW + k) / 2 + cz * (G / 2 - GW);
  const double* Up = U + (long)parent * NFE * R;
  double* Uc = U + (long)node * NFE * R;
  const long e0 = L.from_padded(i0, j0, k0);
  const long exm = L.from_padded(i0 - 1, j0, k0), exp_ = L.from_padded(i0 + 1, j0, k0);
  const long eym = L.from_padded(i0, j0 - 1, k0), eyp = L.from_padded(i0, j0 + 1, k0);
  const long ezm = L.from_padded(i0, j0, k0 - 1), ezp = L.from_padded(i0, j0, k0 + 1);
  double v[8][NFE];
#pragma unroll
  for (int l = 0; l < NFE; l++) {
    const double* f = Up + l * R;
    const double u = f[e0];
    const double s1 = x64::minmod2(f[exp_] - u, u - f[exm]);
    const double s2 = x64::minmod2(f[eyp] - u, u - f[eym]);
    const double s3 = x64::minmod2(f[ezp] - u, u - f[ezm]);
    v[0][l] = u - (s1 + s2 + s3) * 0.25;
    v[1][l] = u + (s1 - s2 - s3) * 0.25;
    v[2][l] = u - (s1 - s2 + s3) * 0.25;
    v[3][l] = u + (s1 + s2 - s3) * 0.25;
    v[4][l] = u - (s1 + s2 - s3) * 0.25;
    v[5][l] = u + (s1 - s2 + s3) * 0.25;
    v[6][l] = u - (s1 - s2 - s3) * 0.25;
    v[7][l] = u + (s1 + s2 + s3) * 0.25;
  }
#pragma unroll
  for (int qq = 0; qq < 8; qq++) {
    x64::floor_state(P, v[qq]);
    const long e = L.interior(2 * ii + (qq & 1), 2 * jj + ((qq >> 1) & 1), 2 * kk + (qq >> 2));
#pragma unroll
    for (int l = 0; l < NFE; l++) Uc[l * R + e] = v[qq][l];
  }
}

// per block and octant: some interior cell has rho > threshold
template <int N, int GW>
__global__ __launch_bounds__(256) void density_flags_kernel(const double* U, int nf, double threshold,
                                                           unsigned* flags) {
  constexpr long R = (long)N * N * N + 6L * GW * N * N;
  const int b = blockIdx.x;
  const double* rho = U + (long)b * nf * R;
  unsigned f = 0;
  for (int e = threadIdx.x; e < N * N * N; e += blockDim.x) {
    const int i = e % N, j = (e / N) % N, k = e / (N * N);
    if (rho[e] > threshold) f |= 1u << ((i >= N / 2 ? 1 : 0) | (j >= N / 2 ? 2 : 0) | (k >= N / 2 ? 4 : 0));
  }
  if (f) atomicOr(&flags[b], f);
}

}  // namespace

extern "C" {

int amrh_tree_repack(amrh_ctx* dst, amrh_ctx* src, const int* src_index) {
  if (!dst || !src || !src_index) return fail(AMRH_EINVAL, "null argument");
  if (!dst->tree || !src->tree) return fail(AMRH_ESTATE, "both contexts need amrh_tree_set");
  if (!src->have_u) return fail(AMRH_ESTATE, "the source holds no state");
  if (dst->cfg.n != src->cfg.n || dst->cfg.g != src->cfg.g || dst->cfg.nf != src->cfg.nf)
    return fail(AMRH_EINVAL, "block shapes differ");
  if (dst->cfg.physics != src->cfg.physics) return fail(AMRH_EINVAL, "the two contexts run different physics");
  const bool rot = dst->cfg.physics == AMRH_PHYS_ROTATING;
  if (rot && dst->cfg.eos == AMRH_EOS_HELMHOLTZ && !dst->eos)
    return fail(AMRH_ESTATE, "eos = HELMHOLTZ needs amrh_attach_eos on the new context");
  // new nodes take their parent's origin (grid/grid_node.cpp:281): the new tree keeps the old one's
  if (int rc = amrh_tree_set_origin(dst, src->origin)) return rc;
  const TreeState* t = dst->tree;
  const int nb = dst->nb, n = dst->cfg.n, nf = dst->cfg.nf;
  std::vector<int> idx(src_index, src_index + nb);
  // A regrid adds at most one level below a surviving node (create_child clears the new node's
  // refine flags, grid/grid_node.cpp:295-297), and the new nodes read their parents' ghost cells as
  // they were before the regrid: all of them are filled in one pass from the gathered parents.
  std::vector<int3> fresh;
  for (int b = 0; b < nb; b++) {
    if (idx[b] >= src->nb) return fail(AMRH_EINVAL, "src_index[%d] = %d out of range", b, idx[b]);
    if (idx[b] < 0) {
      if (b == 0) return fail(AMRH_EINVAL, "the root must survive");
      if (idx[t->parent[b]] < 0)
        return fail(AMRH_EINVAL, "node %d: a new node's parent must survive the regrid", b);
      fresh.push_back(make_int3(b, t->parent[b], t->cidx[b]));
    }
  }
  // the parents' ghost cells feed the prolongation
  if (int rc = amrh_tree_fill_ghosts(src)) return rc;
  HIPCHK(hipStreamSynchronize(src->stream));
  int* idx_dev = nullptr;
  HIPCHK(hipMalloc(&idx_dev, sizeof(int) * nb));
  hipError_t e = hipMemcpyAsync(idx_dev, idx.data(), sizeof(int) * nb, hipMemcpyHostToDevice, dst->stream);
  const long be = dst->L.block();
  if (e == hipSuccess) {
    gather_kernel<<<dim3(8, nb < 65535 ? nb : 65535), 256, 0, dst->stream>>>(dst->buf[dst->cur], src->buf[src->cur], idx_dev, nb, be);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(t->fo_dev, src->tree->fo_dev, sizeof(double) * 2 * nf, hipMemcpyDeviceToDevice,
                                          dst->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(dst->stream);
  (void)hipFree(idx_dev);
  if (e != hipSuccess) return fail(AMRH_EHIP, "repack gather: %s", hipGetErrorString(e));
  dst->have_u = true;
  dst->u0buf = -1;
  if (!fresh.empty()) {
    int3* nodes_dev = nullptr;
    HIPCHK(hipMalloc(&nodes_dev, sizeof(int3) * fresh.size()));
    e = hipMemcpyAsync(nodes_dev, fresh.data(), sizeof(int3) * fresh.size(), hipMemcpyHostToDevice, dst->stream);
    const long total = (long)fresh.size() * (n / 2) * (n / 2) * (n / 2);
    if (e == hipSuccess) {
      const unsigned grid = (unsigned)((total + 63) / 64);
      if (rot) {
        const RotArgs ra = ctx_rot_args(dst);
        e = launch_rot_prolong(dst->cfg.g, dst->cfg.eos, dst->buf[dst->cur], nodes_dev, (int)fresh.size(), dst->x0_dev,
                               dst->h_dev, ra.c, ra.H, dst->stream);
      } else if (n == 8 && dst->cfg.g == 3)
        prolong_kernel<8, 3><<<grid, 64, 0, dst->stream>>>(dst->buf[dst->cur], nodes_dev, (int)fresh.size(),
                                                            params(dst->cfg));
      else if (n == 8)
        prolong_kernel<8, 2><<<grid, 64, 0, dst->stream>>>(dst->buf[dst->cur], nodes_dev, (int)fresh.size(),
                                                            params(dst->cfg));
      else
        prolong_kernel<16, 2><<<grid, 64, 0, dst->stream>>>(dst->buf[dst->cur], nodes_dev, (int)fresh.size(),
                                                             params(dst->cfg));
      if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(dst->stream);
    (void)hipFree(nodes_dev);
    if (e != hipSuccess) return fail(AMRH_EHIP, "prolongation: %s", hipGetErrorString(e));
  }
  // GridNode::check_for_refine ends with enforce_boundaries (grid/grid_node.cpp:1029-1031)
  return amrh_tree_fill_ghosts(dst);
}

int amrh_tree_density_flags(amrh_ctx* c, double min_density, unsigned char* flags) {
  if (!c || !flags) return fail(AMRH_EINVAL, "null argument");
  if (!c->have_u) return fail(AMRH_ESTATE, "no state uploaded");
  const int n = c->cfg.n, g = c->cfg.g;
  if (!((g == 2 && (n == 8 || n == 16)) || (g == 3 && n == 8)))
    return fail(AMRH_EINVAL, "n = 8 or 16 with g = 2, or n = 8 with g = 3");
  unsigned* f_dev = nullptr;
  HIPCHK(hipMalloc(&f_dev, sizeof(unsigned) * c->nb));
  std::vector<unsigned> f(c->nb, 0u);
  hipError_t e = hipMemsetAsync(f_dev, 0, sizeof(unsigned) * c->nb, c->stream);
  if (e == hipSuccess) {
    if (n == 8 && g == 3)
      density_flags_kernel<8, 3><<<c->nb, 256, 0, c->stream>>>(c->buf[c->cur], c->cfg.nf, min_density, f_dev);
    else if (n == 8)
      density_flags_kernel<8, 2><<<c->nb, 256, 0, c->stream>>>(c->buf[c->cur], c->cfg.nf, min_density, f_dev);
    else
      density_flags_kernel<16, 2><<<c->nb, 256, 0, c->stream>>>(c->buf[c->cur], c->cfg.nf, min_density, f_dev);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(f.data(), f_dev, sizeof(unsigned) * c->nb, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(f_dev);
  if (e != hipSuccess) return fail(AMRH_EHIP, "density flags: %s", hipGetErrorString(e));
  for (int b = 0; b < c->nb; b++) flags[b] = (unsigned char)f[b];
  return AMRH_OK;
}

}  // extern "C"
