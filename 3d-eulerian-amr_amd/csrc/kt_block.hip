// Block kernel: the fused KT substep and its CFL twin for schemes with a wide stencil (PPM, g = 3)
// on the reference's 8^3 blocks.
//
// One workgroup per block and one thread per interior cell (N^3 = 512 threads).  The block's
// primitives (interior + the six face ghost slabs, the record layout) are staged in LDS once; then
// for each direction in x, y, z order every face of the block is reconstructed and fluxed (a
// thread per face, the (N+1)*N^2 faces spread over the workgroup) into an LDS face buffer, and
// every cell subtracts its flux difference from its D held in registers, so D accumulates in the
// reference's order (grid/grid.cpp:304-344).  PPM reconstructs whole pencils first (a thread per
// field of a pencil, ppm_pencil: each slope, interface value and monotonisation formed once for
// the faces that share it) into the face states of every face, which the face threads then read.  The cell then forms U = (U + D*dt)*beta +
// U0*(1-beta), floors it and stores it (grid/grid.cpp:406-420).  HBM traffic per block is the
// algorithmic count: interior + face ghost slabs read once, interior written once, U0 read when
// beta != 1.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "amrh_internal.h"
#include "amrhydro.h"
#include "cell_ops.h"
#include "face_order.h"
#include "kt_kernels.h"
#include "rec_buf.h"
#include "recon.h"

namespace amrh {

namespace {

// record element of pencil position p in [-G, N+G) along direction d, at tangential cell
// indices (ta, tb): x pencils (j, k), y pencils (i, k), z pencils (i, j)
template <int N, int G>
__device__ __forceinline__ int pencil_elem(int d, int p, int ta, int tb) {
  constexpr int N3 = N * N * N, S = G * N * N;
  if (p >= 0 && p < N) {
    if (d == 0) return (tb * N + ta) * N + p;
    if (d == 1) return (tb * N + p) * N + ta;
    return (p * N + tb) * N + ta;
  }
  const int up = p >= N ? 1 : 0;
  const int t = up ? p - N : p + G;
  const int base = N3 + (2 * d + up) * S;
  if (d == 0) return base + (tb * N + ta) * G + t;
  if (d == 1) return base + (tb * G + t) * N + ta;
  return base + (t * N + tb) * N + ta;
}

template <int N, int G, int RECON, bool MIXED>
__device__ __forceinline__ void reconstruct(const typename Prec<MIXED>::T (*P)[N * N * N + 6 * G * N * N], int d,
                                            int f, int ta, int tb, typename Prec<MIXED>::T* ql,
                                            typename Prec<MIXED>::T* qr, double theta) {
  using T = typename Prec<MIXED>::T;
  if constexpr (RECON == AMRH_RECON_PPM) {
    int e[6];
#pragma unroll
    for (int m = 0; m < 6; m++) e[m] = pencil_elem<N, G>(d, f - 3 + m, ta, tb);
#pragma unroll
    for (int l = 0; l < NFE; l++) {
      T q[6];
#pragma unroll
      for (int m = 0; m < 6; m++) q[m] = P[l][e[m]];
      // every ideal-gas field is neither low-order nor smooth (euler/state.h:37-42)
      if constexpr (MIXED)
        f32::ppm_face(q, ql[l], qr[l], false, false);
      else
        x64::ppm_face(q, ql[l], qr[l], false, false);
    }
  } else {
    int e[4];
#pragma unroll
    for (int m = 0; m < 4; m++) e[m] = pencil_elem<N, G>(d, f - 2 + m, ta, tb);
#pragma unroll
    for (int l = 0; l < NFE; l++) {
      T q[4];
#pragma unroll
      for (int m = 0; m < 4; m++) q[m] = P[l][e[m]];
      if constexpr (MIXED)
        f32::minmod_face(q, ql[l], qr[l], (float)theta);
      else
        x64::minmod_face(q, ql[l], qr[l], theta);
    }
  }
}

// the KT flux of a face and its wavespeed max(|v_l| + c_l, |v_r| + c_r) -- the number the CFL
// scan forms for the same face states (eig_of_q)
template <bool MIXED, int D>
__device__ __forceinline__ double flux_q(const PhysParams& P, typename Prec<MIXED>::T* ql,
                                         typename Prec<MIXED>::T* qr, typename Prec<MIXED>::T* F) {
  if constexpr (!MIXED) {
    double e;
    x64::kt_flux<D>(P, ql, qr, F, e);
    return e;
  } else {
    const f32::PhysF PF{(float)P.gamma, (float)(P.gamma - 1.0), (float)P.theta};
    float e;
    f32::kt_flux<D>(PF, ql, qr, F, e);
    return (double)e;
  }
}

template <bool MIXED, int D>
__device__ __forceinline__ double eig_of_q(const PhysParams& P, typename Prec<MIXED>::T* ql,
                                           typename Prec<MIXED>::T* qr) {
  if constexpr (!MIXED) {
    return x64::eig_q<D>(P, ql, qr);
  } else {
    const f32::PhysF PF{(float)P.gamma, (float)(P.gamma - 1.0), (float)P.theta};
    return (double)f32::eig_q<D>(PF, ql, qr);
  }
}

// MODE 0: substep.  MODE 1: per-block max wavespeed (Grid::max_dt's scan, grid/grid.cpp:422-473).
// MODE 2: the first substep of an RK step with dt deferred (flat outflow batches): the substep's
// fluxes, the block's max face speed into a.gmax and each cell's D into a.dbuf instead of the
// update; block_first_stage_kernel forms the first stage once dt is known.
// TREE: one level of the octree (a.nodes): reflux overrides applied, the block's own boundary
// faces exported to the coarser nodes' slots, the root's boundary faces to the ledger.
// The mixed substep is held to 128 registers, four waves per SIMD: two workgroups share a CU
// (53 KB of LDS each), which hides the staging and barrier latency -- a small spill, and 32%
// faster than one workgroup at 155 registers on the 8192-block PPM batch.
template <int N, int G, int RECON, bool MIXED, int MODE, bool TREE>
__global__ __launch_bounds__(N* N* N) __attribute__((amdgpu_waves_per_eu(MIXED && MODE != 1 ? 4 : 1))) void
kt_block_kernel(SubstepArgs a, MaxEigArgs m) {
  constexpr bool SUB = MODE != 1;  // fluxes and the cells' D
  static_assert(!(TREE && MODE != 0), "the octree instantiation is the substep's");
  using T = typename Prec<MIXED>::T;
  constexpr int NT = N * N * N;
  constexpr int R = N * N * N + 6 * G * N * N;
  constexpr int NFACE = (N + 1) * N * N;
  // PPM reconstructs whole pencils (one field of one pencil per thread) into the face states of
  // every face, Fq[0] (low side) and Fq[1] (high side); a face's flux then overwrites its low side.
  // Minmod reconstructs face by face into its flux buffer Fq[0].
  constexpr bool PENCIL = RECON == AMRH_RECON_PPM;
  constexpr int NQ = PENCIL ? 2 : 1;
  constexpr bool QBUF = PENCIL || SUB;
  __shared__ T P[NFE][R];
  __shared__ T Fq[NQ][QBUF ? NFE : 1][QBUF ? NFACE : 1];
  auto& Fb = Fq[0];
  __shared__ double red[NT / 64];

  const int tid = threadIdx.x;
  const PhysParams PP = SUB ? a.P : m.P;
  const int nbl = SUB ? a.nblocks : m.nblocks;
  double dt = a.dt;
  if (MODE == 0 && a.dt_dev) dt = *a.dt_dev;
  const double beta = a.beta;
  const bool has_u0 = MODE == 0 && beta != 1.0;
  const int ci = tid % N, cj = (tid / N) % N, ck = tid / (N * N);

  for (int bb = blockIdx.x; bb < nbl; bb += gridDim.x) {
    const int b = SUB ? (TREE ? a.nodes[bb] : a.block0 + bb) : bb;
    const double* Ub = (SUB ? a.Uin : m.U) + (long)b * NFE * R;
    const RecBuf<R, NFE> ub(Ub);

    // stage the primitives; each thread keeps its own interior cell's conserved state.  An
    // isolated outflow block (not TREE) forms each ghost from the boundary cell it copies
    // (grid/grid_phys_bound.cpp:27-54); its ghost slabs are not read.  The slabs are
    // wavefront-aligned (G N^2 a multiple of 64), so the choice is uniform per wavefront.
    const bool self = !TREE && (SUB ? a.outflow : m.outflow) != 0;
    const bool zr = (SUB ? a.z_reflect : m.z_reflect) != 0;
    double ucur[NFE];
#pragma unroll 1
    for (int e = tid; e < R; e += NT) {
      double u[NFE];
      int post = -1;
      const unsigned src = (self && e >= NT) ? (unsigned)outflow_source<N, G>(e, zr, post) : (unsigned)e;
#pragma unroll
      for (int l = 0; l < NFE; l++) u[l] = ub.ld(l, src);
      ghost_post(u, post);
      if (e == tid) {  // the first pass: e = tid < N^3 is this thread's interior cell
#pragma unroll
        for (int l = 0; l < NFE; l++) ucur[l] = u[l];
      }
      T p[NFE];
      make_prim<MIXED>(u, p);
#pragma unroll
      for (int l = 0; l < NFE; l++) P[l][e] = p[l];
    }
    __syncthreads();

    double D[NFE];
#pragma unroll
    for (int l = 0; l < NFE; l++) D[l] = 0.0;
    double emax = 0.0;
    const double hinv = SUB ? 1.0 / a.h[b] : 0.0;

    auto sweep = [&](auto dc) {
      constexpr int d = decltype(dc)::value;
      if constexpr (PENCIL) {
        // every ideal-gas field is neither low-order nor smooth (euler/state.h:37-42)
#pragma unroll 1
        for (int t = tid; t < NFE * N * N; t += NT) {
          const int l = t / (N * N), ta = t % N, tb = (t / N) % N;
          T q[N + 6], ql[N + 1], qr[N + 1];
#pragma unroll
          for (int m = 0; m < N + 6; m++) q[m] = P[l][pencil_elem<N, G>(d, m - 3, ta, tb)];
          if constexpr (MIXED)
            f32::ppm_pencil<N>(q, ql, qr, false, false);
          else
            x64::ppm_pencil<N>(q, ql, qr, false, false);
#pragma unroll
          for (int f = 0; f <= N; f++) {
            const int q2 = face_index<N, d>(f, ta, tb);
            Fq[0][l][q2] = ql[f];
            Fq[NQ - 1][l][q2] = qr[f];
          }
        }
        __syncthreads();
      }
#pragma unroll 1
      for (int q = tid; q < NFACE; q += NT) {
        T ql[NFE], qr[NFE];
        if constexpr (PENCIL) {
#pragma unroll
          for (int l = 0; l < NFE; l++) {
            ql[l] = Fq[0][l][q];
            qr[l] = Fq[NQ - 1][l][q];
          }
        } else {
          int f, ta, tb;
          face_decode<N, d>(q, f, ta, tb);
          reconstruct<N, G, RECON, MIXED>(P, d, f, ta, tb, ql, qr, PP.theta);
        }
        if constexpr (SUB) {
          T Fo[NFE];
          const double e = flux_q<MIXED, d>(PP, ql, qr, Fo);
          if constexpr (MODE == 2) emax = x64::rmax(emax, e);
#pragma unroll
          for (int l = 0; l < NFE; l++) Fb[l][q] = Fo[l];
        } else {
          emax = x64::rmax(emax, eig_of_q<MIXED, d>(PP, ql, qr));
        }
      }
      if constexpr (SUB) {
        __syncthreads();
        if constexpr (TREE) {
          // adjust_*_flux (grid/grid_node.cpp:1076-1274): faces of the low, middle and high planes
          // next to a finer region take the means the finer level exported
#pragma unroll 1
          for (int q2 = tid; q2 < 3 * N * N; q2 += NT) {
            const int pl = q2 / (N * N), ta = q2 % N, tb = (q2 / N) % N;
            const int qa = ta >= N / 2, qb = tb >= N / 2;
            const int sl = a.slot[((b * 3 + d) * 3 + pl) * 4 + qa + 2 * qb];
            if (sl < 0) continue;
            const int f = pl == 0 ? 0 : (pl == 1 ? N / 2 : N);
            const double* o =
                a.ovr + (((long)sl * (N / 2) + (tb - qb * (N / 2))) * (N / 2) + (ta - qa * (N / 2))) * NFE;
            const int q = face_index<N, d>(f, ta, tb);
#pragma unroll
            for (int l = 0; l < NFE; l++) Fb[l][q] = (T)o[l];
          }
          __syncthreads();
          if (tid < 4 * (N / 2) * (N / 2)) {
            const int w = tid / ((N / 2) * (N / 2)), r = tid % ((N / 2) * (N / 2));
            const int side = w >> 1;
            const int sl = a.exp_slot[(b * 6 + 2 * d + side) * 2 + (w & 1)];
            if (sl >= 0) {
              const int al = r % (N / 2), bl = r / (N / 2);
              const int fp = side ? N : 0;
              const int q00 = face_index<N, d>(fp, 2 * al, 2 * bl), q10 = face_index<N, d>(fp, 2 * al + 1, 2 * bl);
              const int q01 = face_index<N, d>(fp, 2 * al, 2 * bl + 1);
              const int q11 = face_index<N, d>(fp, 2 * al + 1, 2 * bl + 1);
              double* o = a.ovr_out + (((long)sl * (N / 2) + bl) * (N / 2) + al) * NFE;
#pragma unroll
              for (int l = 0; l < NFE; l++) {
                double v = (double)Fb[l][q00];
                v += (double)Fb[l][q10];
                v += (double)Fb[l][q01];
                v += (double)Fb[l][q11];
                v *= 0.25;
                o[l] = v;
              }
            }
          }
          if (b == a.ledger_block && a.ledger) {
#pragma unroll 1
            for (int q2 = tid; q2 < 2 * N * N; q2 += NT) {
              const int side = q2 / (N * N), ta = q2 % N, tb = (q2 / N) % N;
              const int q = face_index<N, d>(side ? N : 0, ta, tb);
              double* o = a.ledger + ((((long)d * 2 + side) * N + tb) * N + ta) * NFE;
#pragma unroll
              for (int l = 0; l < NFE; l++) o[l] = (double)Fb[l][q];
            }
          }
        }
        const int p = d == 0 ? ci : (d == 1 ? cj : ck);
        const int ta = d == 0 ? cj : ci;
        const int tb = d == 2 ? cj : ck;
        const int q0 = face_index<N, d>(p, ta, tb), q1 = face_index<N, d>(p + 1, ta, tb);
#pragma unroll
        for (int l = 0; l < NFE; l++) {
          if constexpr (!MIXED)
            D[l] -= (Fb[l][q1] - Fb[l][q0]) * hinv;
          else  // FP32 fluxes differenced in FP64 (exact), so the divergence telescopes
            D[l] -= ((double)Fb[l][q1] - (double)Fb[l][q0]) * hinv;
        }
        __syncthreads();  // the next direction rewrites the face buffer
      } else if constexpr (PENCIL) {
        __syncthreads();  // the next direction rewrites the face states
      }
    };
    sweep(std::integral_constant<int, 0>{});
    sweep(std::integral_constant<int, 1>{});
    sweep(std::integral_constant<int, 2>{});

    if constexpr (MODE == 2) {
      // deferred: the cell's D (State::source is zero for the ideal gas) for block_first_stage_kernel
      const RecBuf<NT, NFE> dbo(a.dbuf + (long)b * NFE * NT);
#pragma unroll
      for (int l = 0; l < NFE; l++) dbo.st_nt(l, (unsigned)tid, D[l] + 0.0);
    }
    if constexpr (MODE == 0) {
      const long e = (long)tid;  // interior element of cell (ci, cj, ck)
      double u0v[NFE];
      if (has_u0) {
        const RecBuf<R, NFE> u0(a.U0 + (long)b * NFE * R);
#pragma unroll
        for (int l = 0; l < NFE; l++) u0v[l] = u0.ld(l, (unsigned)e);
      }
      double u[NFE];
#pragma unroll
      for (int l = 0; l < NFE; l++) {
        const double Dl = D[l] + 0.0;  // State::source is zero for the ideal gas (euler/state.cpp:19-22)
        if (has_u0)
          u[l] = (ucur[l] + Dl * dt) * beta + u0v[l] * (1.0 - beta);
        else
          u[l] = (ucur[l] + Dl * dt) * beta;
      }
      if constexpr (!MIXED)
        x64::floor_state(PP, u);
      else
        floor_mixed(PP, u);
      if (!(u[RHO] > 0.0) && a.err) atomicOr(a.err, 1);
      const RecBuf<R, NFE> uo(a.Uout + (long)b * NFE * R);
#pragma unroll
      for (int l = 0; l < NFE; l++) uo.st(l, (unsigned)e, u[l]);
    } else {
      for (int off = 32; off > 0; off >>= 1) emax = x64::rmax(emax, __shfl_xor(emax, off, 64));
      if ((tid & 63) == 0) red[tid >> 6] = emax;
      __syncthreads();
      if (tid == 0) {
        double v = red[0];
        for (int w = 1; w < NT / 64; w++) v = x64::rmax(v, red[w]);
        if constexpr (MODE == 1) {
          if (m.block_eig) m.block_eig[b] = v;
          if (m.gmax) {
            const double r = v / m.h[b];
            atomicMax(m.gmax, (unsigned long long)__double_as_longlong(r));
          }
        } else {
          const double r = v / a.h[b];
          atomicMax(a.gmax, (unsigned long long)__double_as_longlong(r));
        }
      }
    }
    __syncthreads();  // P (and red) are restaged by the next block of a persistent grid
  }
}

// The first stage of a deferred RK step, one thread per interior cell: the substep's update with
// beta = 1 (kt_block_kernel MODE 0) from the D the deferred substep stored, U1 = (U^n + D dt) 1,
// and the floor (euler/state.cpp:86-92); the same operations, so U1 has the same bits.
template <int N, int G, bool MIXED>
__global__ __launch_bounds__(256) void block_first_stage_kernel(SubstepArgs a) {
  constexpr int NT = N * N * N;
  constexpr int R = NT + 6 * G * N * N;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)a.nblocks * NT) return;
  const int b = (int)(gid / NT), e = (int)(gid % NT);
  const double dt = *a.dt_dev;
  const double beta = 1.0;
  const RecBuf<R, NFE> ub(a.Uin + (long)b * NFE * R), uo(a.Uout + (long)b * NFE * R);
  const RecBuf<NT, NFE> db(a.dbuf + (long)b * NFE * NT);
  double u[NFE];
#pragma unroll
  for (int l = 0; l < NFE; l++) u[l] = (ub.ld(l, (unsigned)e) + db.ld_nt(l, (unsigned)e) * dt) * beta;
  if constexpr (!MIXED)
    x64::floor_state(a.P, u);
  else
    floor_mixed(a.P, u);
  if (!(u[RHO] > 0.0) && a.err) atomicOr(a.err, 1);
#pragma unroll
  for (int l = 0; l < NFE; l++) uo.st(l, (unsigned)e, u[l]);
}

template <int N, int G, int RECON, bool MIXED, int MODE>
hipError_t launch_one(const SubstepArgs& a, const MaxEigArgs& m, int nblocks, int grid, hipStream_t s) {
  const int g = grid > 0 && grid < nblocks ? grid : nblocks;
  if (nblocks <= 0) return hipSuccess;
  if (MODE == 0 && a.nodes)
    kt_block_kernel<N, G, RECON, MIXED, MODE, MODE == 0><<<g, N * N * N, 0, s>>>(a, m);
  else
    kt_block_kernel<N, G, RECON, MIXED, MODE, false><<<g, N * N * N, 0, s>>>(a, m);
  return hipGetLastError();
}

}  // namespace

bool block_scheme_supported(int n, int g, int recon) {
  return n == 8 && ((recon == AMRH_RECON_PPM && g == 3) || (recon == AMRH_RECON_MINMOD && g == 2));
}

hipError_t launch_block_substep(int n, int g, int recon, const SubstepArgs& a, bool mixed, int grid,
                                hipStream_t s) {
  const MaxEigArgs m{};
  if (!block_scheme_supported(n, g, recon)) return hipErrorInvalidValue;
  if (recon == AMRH_RECON_PPM)
    return mixed ? launch_one<8, 3, AMRH_RECON_PPM, true, 0>(a, m, a.nblocks, grid, s)
                 : launch_one<8, 3, AMRH_RECON_PPM, false, 0>(a, m, a.nblocks, grid, s);
  return mixed ? launch_one<8, 2, AMRH_RECON_MINMOD, true, 0>(a, m, a.nblocks, grid, s)
               : launch_one<8, 2, AMRH_RECON_MINMOD, false, 0>(a, m, a.nblocks, grid, s);
}

hipError_t launch_block_deferred(int n, int g, int recon, const SubstepArgs& a, bool mixed, int grid,
                                 hipStream_t s) {
  const MaxEigArgs m{};
  if (!block_scheme_supported(n, g, recon)) return hipErrorInvalidValue;
  if (a.nodes || a.slot || !a.outflow || !a.dbuf || !a.gmax) return hipErrorInvalidValue;
  if (recon == AMRH_RECON_PPM)
    return mixed ? launch_one<8, 3, AMRH_RECON_PPM, true, 2>(a, m, a.nblocks, grid, s)
                 : launch_one<8, 3, AMRH_RECON_PPM, false, 2>(a, m, a.nblocks, grid, s);
  return mixed ? launch_one<8, 2, AMRH_RECON_MINMOD, true, 2>(a, m, a.nblocks, grid, s)
               : launch_one<8, 2, AMRH_RECON_MINMOD, false, 2>(a, m, a.nblocks, grid, s);
}

hipError_t launch_block_first_stage(int n, int g, const SubstepArgs& a, bool mixed, hipStream_t s) {
  if (n != 8 || (g != 2 && g != 3) || !a.dbuf || !a.dt_dev) return hipErrorInvalidValue;
  if (a.nblocks <= 0) return hipSuccess;
  const unsigned grid = (unsigned)(((long)a.nblocks * 512 + 255) / 256);
  if (g == 3)
    mixed ? block_first_stage_kernel<8, 3, true><<<grid, 256, 0, s>>>(a)
          : block_first_stage_kernel<8, 3, false><<<grid, 256, 0, s>>>(a);
  else
    mixed ? block_first_stage_kernel<8, 2, true><<<grid, 256, 0, s>>>(a)
          : block_first_stage_kernel<8, 2, false><<<grid, 256, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_block_maxeig(int n, int g, int recon, const MaxEigArgs& m, bool mixed, int grid, hipStream_t s) {
  const SubstepArgs a{};
  if (!block_scheme_supported(n, g, recon)) return hipErrorInvalidValue;
  if (recon == AMRH_RECON_PPM)
    return mixed ? launch_one<8, 3, AMRH_RECON_PPM, true, 1>(a, m, m.nblocks, grid, s)
                 : launch_one<8, 3, AMRH_RECON_PPM, false, 1>(a, m, m.nblocks, grid, s);
  return mixed ? launch_one<8, 2, AMRH_RECON_MINMOD, true, 1>(a, m, m.nblocks, grid, s)
               : launch_one<8, 2, AMRH_RECON_MINMOD, false, 1>(a, m, m.nblocks, grid, s);
}

}  // namespace amrh
