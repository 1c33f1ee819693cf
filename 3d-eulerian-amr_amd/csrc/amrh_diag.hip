// Diagnostics as device reductions: GridNode::state_sum / state_max / state_min
// (grid/grid_node.cpp:1489-1589) over the uncovered interior cells of every block.
//
// Pass 1: one workgroup per block reduces, per field, the four moments sum(U h^3), sum(U h^3 x),
// sum(U h^3 y), sum(U h^3 z) and the extrema over the block's cells whose octant has no child
// (GridNode::zone_is_refined, grid/grid_node.cpp:1355-1362), in a fixed tree order.  Pass 2: one
// thread per (field, quantity) combines the blocks in index order.  The result does not depend on
// the launch shape; it differs from the reference's OpenMP reductions only by summation order.
#include <hip/hip_runtime.h>

#include <vector>

#include "amrh_ctx.h"
#include "amrh_internal.h"
#include "amrhydro.h"

using namespace amrh;

namespace {

constexpr int NQ = 6;  // a0, a1, a2, a3, max, min
constexpr int DT = 256;

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }  // real.h:20
__device__ __forceinline__ double dmin(double a, double b) { return a > b ? b : a; }  // real.h:38

template <int N, int G>
__global__ __launch_bounds__(DT) void diag_block_kernel(const double* __restrict__ U, int nf,
                                                       const double* __restrict__ x0, const double* __restrict__ h,
                                                       const unsigned char* __restrict__ covered,
                                                       double* __restrict__ part) {
  constexpr int N3 = N * N * N;
  constexpr long R = (long)N3 + 6L * G * N * N;
  __shared__ double red[NQ][DT];
  const int b = blockIdx.x, tid = threadIdx.x;
  const double hb = h[b];
  const double h3 = hb * hb * hb;
  const unsigned cov = covered ? covered[b] : 0u;
  const double ox = x0 ? x0[3 * b] : 0.0, oy = x0 ? x0[3 * b + 1] : 0.0, oz = x0 ? x0[3 * b + 2] : 0.0;
  for (int l = 0; l < nf; l++) {
    const double* Ub = U + ((long)b * nf + l) * R;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, mx = -1.0e+99, mn = 1.0e+99;
    for (int e = tid; e < N3; e += DT) {
      const int i = e % N, j = (e / N) % N, k = e / (N * N);
      const int oct = (i >= N / 2 ? 1 : 0) | (j >= N / 2 ? 2 : 0) | (k >= N / 2 ? 4 : 0);
      if ((cov >> oct) & 1u) continue;
      const double v = Ub[e];
      const double ds = v * h3;
      a0 += ds;
      a1 += ds * (ox + (double)i * hb + 0.5 * hb);
      a2 += ds * (oy + (double)j * hb + 0.5 * hb);
      a3 += ds * (oz + (double)k * hb + 0.5 * hb);
      mx = dmax(v, mx);
      mn = dmin(v, mn);
    }
    red[0][tid] = a0;
    red[1][tid] = a1;
    red[2][tid] = a2;
    red[3][tid] = a3;
    red[4][tid] = mx;
    red[5][tid] = mn;
    __syncthreads();
    for (int s = DT / 2; s > 0; s >>= 1) {
      if (tid < s) {
#pragma unroll
        for (int q = 0; q < 4; q++) red[q][tid] += red[q][tid + s];
        red[4][tid] = dmax(red[4][tid + s], red[4][tid]);
        red[5][tid] = dmin(red[5][tid + s], red[5][tid]);
      }
      __syncthreads();
    }
    if (tid < NQ) part[((long)b * nf + l) * NQ + tid] = red[tid][0];
    __syncthreads();
  }
}

// sums_layout = 0: out[l*NQ + q] for every quantity; 1: only the moments, as amrh_state_sum
// returns them, out[q*nf + l] for q < 4
__global__ void diag_combine_kernel(const double* __restrict__ part, int nb, int nf, double* __restrict__ out,
                                    int sums_layout) {
  const int t = threadIdx.x;
  if (t >= nf * NQ) return;
  const int l = t / NQ, q = t % NQ;
  if (sums_layout && q >= 4) return;
  double v = q < 4 ? 0.0 : (q == 4 ? -1.0e+99 : 1.0e+99);
  for (int b = 0; b < nb; b++) {
    const double p = part[((long)b * nf + l) * NQ + q];
    if (q < 4)
      v += p;
    else if (q == 4)
      v = dmax(p, v);
    else
      v = dmin(p, v);
  }
  out[sums_layout ? q * nf + l : t] = v;
}

}  // namespace

int amrh::diag_check(const amrh_ctx* c) {
  if (!c->have_u || !c->have_h) return fail(AMRH_ESTATE, "state or cell widths not set");
  if (c->cfg.nf * NQ > 1024) return fail(AMRH_EINVAL, "too many fields");
  const int n = c->cfg.n, g = c->cfg.g;
  if (!((n == 8 && (g == 2 || g == 3)) || (n == 16 && g == 2)))
    return fail(AMRH_EINVAL, "no diagnostics for n = %d, g = %d", n, g);
  return AMRH_OK;
}

long amrh::diag_part_elems(const amrh_ctx* c) { return (long)c->nb * c->cfg.nf * NQ; }

// pass 1 into part (diag_part_elems doubles), pass 2 into out: the state_sum layout (4*nf) when
// sums_layout, else [nf][NQ]; asynchronous on the context's stream, no allocation
int amrh::diag_launch(const amrh_ctx* c, double* part, double* out, bool sums_layout) {
  const int nf = c->cfg.nf, nb = c->nb;
  const unsigned char* cov = c->tree ? c->tree->covered_dev : nullptr;
  const double* x0 = c->have_x0 ? c->x0_dev : nullptr;
  const int n = c->cfg.n, g = c->cfg.g;
  if (n == 8 && g == 2)
    diag_block_kernel<8, 2><<<nb, DT, 0, c->stream>>>(c->buf[c->cur], nf, x0, c->h_dev, cov, part);
  else if (n == 8 && g == 3)
    diag_block_kernel<8, 3><<<nb, DT, 0, c->stream>>>(c->buf[c->cur], nf, x0, c->h_dev, cov, part);
  else
    diag_block_kernel<16, 2><<<nb, DT, 0, c->stream>>>(c->buf[c->cur], nf, x0, c->h_dev, cov, part);
  HIPCHK(hipGetLastError());
  diag_combine_kernel<<<1, 1024, 0, c->stream>>>(part, nb, nf, out, sums_layout ? 1 : 0);
  HIPCHK(hipGetLastError());
  return AMRH_OK;
}

namespace {

int run_diag(amrh_ctx* c, std::vector<double>& res) {
  if (!c) return fail(AMRH_EINVAL, "null ctx");
  if (int rc = diag_check(c)) return rc;
  const int nf = c->cfg.nf, nb = c->nb;
  if (!c->diag_part) {
    HIPCHK(hipMalloc(&c->diag_part, sizeof(double) * ((size_t)nb * nf * NQ + (size_t)nf * NQ)));
  }
  double* out = c->diag_part + (size_t)nb * nf * NQ;
  if (int rc = diag_launch(c, c->diag_part, out, false)) return rc;
  res.assign((size_t)nf * NQ, 0.0);
  HIPCHK(hipMemcpyAsync(res.data(), out, sizeof(double) * res.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return AMRH_OK;
}

}  // namespace

extern "C" {

int amrh_state_sum(amrh_ctx* c, double* sums) {
  if (!sums) return fail(AMRH_EINVAL, "null out");
  if (c && !c->tree && !c->have_x0)
    return fail(AMRH_ESTATE, "the moments need block positions: call amrh_set_block_origin");
  std::vector<double> r;
  if (int rc = run_diag(c, r)) return rc;
  const int nf = c->cfg.nf;
  for (int q = 0; q < 4; q++)
    for (int l = 0; l < nf; l++) sums[q * nf + l] = r[(size_t)l * NQ + q];
  return AMRH_OK;
}

int amrh_state_max(amrh_ctx* c, double* out) {
  if (!out) return fail(AMRH_EINVAL, "null out");
  std::vector<double> r;
  if (int rc = run_diag(c, r)) return rc;
  for (int l = 0; l < c->cfg.nf; l++) out[l] = r[(size_t)l * NQ + 4];
  return AMRH_OK;
}

int amrh_state_min(amrh_ctx* c, double* out) {
  if (!out) return fail(AMRH_EINVAL, "null out");
  std::vector<double> r;
  if (int rc = run_diag(c, r)) return rc;
  for (int l = 0; l < c->cfg.nf; l++) out[l] = r[(size_t)l * NQ + 5];
  return AMRH_OK;
}

}  // extern "C"
