// Rotating-frame binary-star substep (the reference's BinaryDriver::sub_step with a given
// gravitational potential), its CFL scan and its outflow ghost fill, on 8^3 blocks; FP64, or the
// mixed form with FP32 faces.
//
// One workgroup of N^3 = 512 threads per block, a thread per interior cell.  For each direction
// in x, y, z order the primitives of the direction's pencils (interior + its two face ghost slabs)
// are staged in LDS, every face of the block is reconstructed (PPM with the potential field
// low-order and smooth, or minmod), converted back at the face position, given to the EOS, and
// its KT flux and the normal velocity of the mean face state (SAVE_RECONSTRUCTIONS,
// grid/grid.cpp:235-237) are written to LDS; each cell then subtracts its flux difference from D
// (grid/grid.cpp:304-344).  The cell finally adds the gravity source from the face-averaged force
// of the potential and the -p div u source (binary/binary.cpp:205-254, poisson/poisson.cpp:223-247),
// zeroes D[pot] and D[T], adds the geometric source and forms U = (U + D dt) beta + U0 (1 - beta)
// (grid/grid.cpp:406-420), floors it with the Newton solve for T (binary/state_ztwd.cpp:482-554) and
// resets pot = rho (phi - R^2 Omega0^2 / 2) (binary/binary.cpp:825-849).
// LDS (PPM, G = 3): FP64 stages one direction's 13 x 896 primitives beside 13 x 576 face fluxes
// and 576 face velocities (154 KB); the mixed form stages the whole record, 13 x 1664 FP32
// primitives, once per block beside two such face buffers in FP32 (148 KB).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "amr_kernels.h"
#include "amrh_internal.h"
#include "amrh_layout.h"
#include "amrhydro.h"
#include "face_order.h"
#include "kt_rot.h"
#include "rec_buf.h"
#include "recon.h"
#include "rot_phys.h"

namespace amrh {

namespace {

using namespace rot;

constexpr int N = 8;
constexpr int N2 = N * N, N3 = N * N * N;
constexpr int NFACE = (N + 1) * N * N;

// libm calls of the frame source, out of line: inlined, their FP64 constants are hoisted out of the
// block loop into registers that stay allocated for the whole kernel
__device__ __attribute__((noinline)) double atan2_ool(double y, double x) { return atan2(y, x); }
__device__ __attribute__((noinline)) double cos_ool(double x) { return cos(x); }
__device__ __attribute__((noinline)) double sin_ool(double x) { return sin(x); }

// record element -> interior-relative coordinates (i, j, k), any of them in [-G, N+G)
template <int G>
__device__ __forceinline__ void elem_ijk(int e, int& i, int& j, int& k) {
  constexpr int S = G * N2;
  if (e < N3) {
    i = e % N;
    j = (e / N) % N;
    k = e / N2;
    return;
  }
  const int s = (e - N3) / S, r = (e - N3) % S;
  const int up = s & 1;
  if (s < 2) {
    const int t = r % G;
    j = (r / G) % N;
    k = r / (G * N);
    i = up ? N + t : t - G;
  } else if (s < 4) {
    i = r % N;
    const int t = (r / N) % G;
    k = r / (G * N);
    j = up ? N + t : t - G;
  } else {
    i = r % N;
    j = (r / N) % N;
    const int t = r / N2;
    k = up ? N + t : t - G;
  }
}

// the region staged for direction d: interior, then the low and high slabs of d
template <int G>
__device__ __forceinline__ int region_to_elem(int d, int r) {
  constexpr int S = G * N2;
  return r < N3 ? r : r + 2 * d * S;
}
template <int G>
__device__ __forceinline__ int pencil_region(int d, int p, int ta, int tb) {
  constexpr int S = G * N2;
  if (p >= 0 && p < N) {
    if (d == 0) return (tb * N + ta) * N + p;
    if (d == 1) return (tb * N + p) * N + ta;
    return (p * N + tb) * N + ta;
  }
  const int up = p >= N ? 1 : 0;
  const int t = up ? p - N : p + G;
  const int base = N3 + up * S;
  if (d == 0) return base + (tb * N + ta) * G + t;
  if (d == 1) return base + (tb * G + t) * N + ta;
  return base + (t * N + tb) * N + ta;
}

__device__ __forceinline__ void cell_pos(const double* x0, double h, int i, int j, int k, double* X) {
  const int q[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double xf = x0[d] + (double)q[d] * h;
    X[d] = xf + 0.5 * h;
  }
}

// MIXED: the staged primitives, the reconstruction, the face states' EOS and fluxes and the KT
// combination in FP32 (rot::sp); the flux differences, sources, RK update and floor in FP64.
// TREE: the octree's reflux overrides, exports and the root's ledger (a.slot, a.exp_slot, a.ledger)
// The mixed form stages the whole record (interior + 6 slabs) once per block and computes the faces
// of all three directions as one list (with the ideal gas, as pairs of adjacent faces that share
// their reconstruction); the FP64 form stages one direction's pencils at a time.
// MODE 0: substep.  MODE 1: the CFL scan.  MODE 2: the first substep of an RK step with dt deferred
// (flat batches, and the levels of the octree): the substep's fluxes and sources, the block's max
// face speed (the numbers the scan forms from the same faces, before any reflux override) into
// a.gmax, and each cell's D (sources included) into a.dbuf instead of the update;
// rot_stage_kernel forms the first stage once dt is known; with a.gmax null, only D (the split
// form of a later substep).  The octree's reflux overrides,
// exports and ledger faces do not depend on dt, so MODE 2 applies and writes them as MODE 0 does.
template <int G, int RECON, int EOS, int MODE, bool MIXED, bool TREE>
__global__ __launch_bounds__(N3) void kt_rot_kernel(RotArgs a) {
  using T = typename std::conditional<MIXED, float, double>::type;
  constexpr bool SUB = MODE == 0 || MODE == 2;  // fluxes and the cells' D
  constexpr bool EIGM = MODE != 0;              // the block's max face speed
  static_assert(!(TREE && MODE == 1), "the octree instantiations are the substep's");
  constexpr bool ALL = MIXED;          // stage the whole record once
  constexpr int R = N3 + 6 * G * N2;   // record elements per field
  constexpr int RS = ALL ? R : N3 + 2 * G * N2;  // staged elements per field
  __shared__ T P[NF][RS];
  constexpr int NFB = SUB ? (ALL ? 2 : 1) : 1;  // face buffers (double-buffered in the pooled form)
  __shared__ T Fb[NFB][SUB ? NF : 1][SUB ? NFACE : 1];
  __shared__ T Vf[NFB][SUB ? NFACE : 1];
  __shared__ double red[N3 / 64];
  const sp::RotF cf = sp::make(a.c);

  const int tid = threadIdx.x;
  const RotParams& c = a.c;
  const helm::HelmDev& H = a.H;
  double dt = a.dt;
  if (MODE == 0 && a.dt_dev) dt = *a.dt_dev;
  const double beta = a.beta;
  const bool has_u0 = MODE == 0 && beta != 1.0;
  const int ci = tid % N, cj = (tid / N) % N, ck = tid / N2;

  for (int bi = blockIdx.x; bi < a.nblocks; bi += gridDim.x) {
    const int b = a.nodes ? a.nodes[bi] : bi;
    const double* Ub = a.Uin + (long)b * NF * R;
    const RecBuf<R, NF> ub(Ub);
    const double h = a.h[b];
    const double x0[3] = {a.x0[3 * b], a.x0[3 * b + 1], a.x0[3 * b + 2]};
    double D[NF];
#pragma unroll
    for (int l = 0; l < NF; l++) D[l] = 0.0;
    double vlo[3], vhi[3];
    double emax = 0.0;
    const double hinv = 1.0 / h;

    // the staged primitives of the cells of pencil position p of direction d
    auto staged = [&](int d, int p, int ta, int tb) {
      const int r = pencil_region<G>(d, p, ta, tb);
      return (ALL && (p < 0 || p >= N)) ? r + 2 * d * G * N2 : r;
    };
    auto stage = [&](int d) {
#pragma unroll 1
      for (int r = tid; r < RS; r += N3) {
        const int e = ALL ? r : region_to_elem<G>(d, r);
        int i, j, k;
        elem_ijk<G>(e, i, j, k);
        double u[NF];
        // octree: a ghost cell across a face with a same-level neighbour is that neighbour's cell
        // (what the ghost exchange would copy there), read from it directly; the slabs are
        // wavefront-aligned (G N^2 is a multiple of 64), so the branch is uniform
        const int* gt = (TREE && a.ghost && e >= N3) ? a.ghost + (b * 6 + (e - N3) / (G * N2)) * 2 : nullptr;
        if (gt && gt[0] == GHOST_SIB) {
          const int nf = (e - N3) / (G * N2);
          int s3[3] = {i, j, k};
          s3[nf >> 1] += (nf & 1) ? -N : N;
          const double* Us = a.Uin + (long)gt[1] * NF * R + (s3[2] * N + s3[1]) * N + s3[0];
#pragma unroll
          for (int l = 0; l < NF; l++) u[l] = Us[(long)l * R];
        } else if (!TREE && a.outflow && e >= N3) {
          // isolated outflow block: the ghost copies the boundary cell next to it, mirrored below
          // with z_reflect or given enforce_outflow at the boundary face (as rot_ghost_kernel)
          const int nf = (e - N3) / (G * N2), dd = nf >> 1, up = nf & 1;
          int s3[3] = {i, j, k};
          const bool mirror = a.z_reflect && nf == 4;
          s3[dd] = mirror ? -1 - s3[dd] : (up ? N - 1 : 0);
          const unsigned es = (unsigned)((s3[2] * N + s3[1]) * N + s3[0]);
#pragma unroll
          for (int l = 0; l < NF; l++) u[l] = ub.ld(l, es);
          if (mirror) {
            u[rot::SZ] = -u[rot::SZ];
          } else {
            double Xb[3];
            cell_pos(x0, h, s3[0], s3[1], s3[2], Xb);
            Xb[dd] = x0[dd] + (double)(up ? N : 0) * h;  // the boundary face
            enforce_outflow(c, u, Xb, nf);
          }
        } else {
#pragma unroll
          for (int l = 0; l < NF; l++) u[l] = ub.ld(l, (unsigned)e);
        }
        double X[3];
        cell_pos(x0, h, i, j, k, X);
        to_prim(c, u, X);
#pragma unroll
        for (int l = 0; l < NF; l++) P[l][r] = (T)u[l];
      }
      __syncthreads();
    };
    if constexpr (ALL) stage(0);

    // face q (face_order.h) of direction d: both sides reconstructed, converted back at the face
    // position and given to the EOS; the KT flux and the mean state's face velocity go to Fb[fb] and
    // Vf[fb] (MODE 0), the larger wavespeed to emax (MODE 1)
    auto face = [&](auto dc, int q, int fb) {
      constexpr int d = decltype(dc)::value;
      int f, ta, tb;
      face_decode<N, d>(q, f, ta, tb);
      if constexpr (MIXED) {
        float ql[NF], qr[NF];
        if constexpr (RECON == AMRH_RECON_PPM) {
          int e6[6];
#pragma unroll
          for (int m = 0; m < 6; m++) e6[m] = staged(d, f - 3 + m, ta, tb);
#pragma unroll
          for (int l = 0; l < NF; l++) {
            float v[6];
#pragma unroll
            for (int m = 0; m < 6; m++) v[m] = (float)P[l][e6[m]];
            f32::ppm_face(v, ql[l], qr[l], l == POT, l == POT);
          }
        } else {
          int e4[4];
#pragma unroll
          for (int m = 0; m < 4; m++) e4[m] = staged(d, f - 2 + m, ta, tb);
#pragma unroll
          for (int l = 0; l < NF; l++) {
            float v[4];
#pragma unroll
            for (int m = 0; m < 4; m++) v[m] = (float)P[l][e4[m]];
            f32::minmod_face(v, ql[l], qr[l], (float)c.theta);
          }
        }
        float X[3];
        {
          const int ii = d == 0 ? f : ta;
          const int jj = d == 0 ? ta : (d == 1 ? f : tb);
          const int kk = d == 2 ? f : tb;
          double Xd[3];
          cell_pos(x0, h, ii, jj, kk, Xd);
          Xd[d] = x0[d] + (double)f * h;
          X[0] = (float)Xd[0], X[1] = (float)Xd[1], X[2] = (float)Xd[2];
        }
        sp::to_con(cf, ql, X);
        sp::to_con(cf, qr, X);
        float fl[NF], fr[NF];
        const float el = sp::flux<EOS, d>(H, cf, ql, X, fl);
        const float er = sp::flux<EOS, d>(H, cf, qr, X, fr);
        const float amax = fmaxf(el, er);
        if constexpr (SUB) {
          float uf[NF];
#pragma unroll
          for (int l = 0; l < NF; l++) uf[l] = (ql[l] + qr[l]) * 0.5f;
          Vf[fb][q] = (T)sp::vel<d>(cf, uf, X, sp::Rc(X));
#pragma unroll
          for (int l = 0; l < NF; l++) Fb[fb][l][q] = (T)(((fl[l] + fr[l]) - (qr[l] - ql[l]) * amax) * 0.5f);
        }
        if constexpr (EIGM) emax = rmax(emax, (double)amax);
      } else {
        // the face's two reconstructed states, converted back at the face position X (cell
        // centres across, the face itself along d)
        auto face_states = [&](double* ql, double* qr, double* X) {
          if constexpr (RECON == AMRH_RECON_PPM) {
            int e6[6];
#pragma unroll
            for (int m = 0; m < 6; m++) e6[m] = staged(d, f - 3 + m, ta, tb);
#pragma unroll
            for (int l = 0; l < NF; l++) {
              double v[6];
#pragma unroll
              for (int m = 0; m < 6; m++) v[m] = (double)P[l][e6[m]];
              // the potential is low-order and smooth (binary/state_ztwd.cpp:50-56)
              x64::ppm_face(v, ql[l], qr[l], l == POT, l == POT);
            }
          } else {
            int e4[4];
#pragma unroll
            for (int m = 0; m < 4; m++) e4[m] = staged(d, f - 2 + m, ta, tb);
#pragma unroll
            for (int l = 0; l < NF; l++) {
              double v[4];
#pragma unroll
              for (int m = 0; m < 4; m++) v[m] = (double)P[l][e4[m]];
              x64::minmod_face(v, ql[l], qr[l], c.theta);
            }
          }
          const int ii = d == 0 ? f : ta;
          const int jj = d == 0 ? ta : (d == 1 ? f : tb);
          const int kk = d == 2 ? f : tb;
          cell_pos(x0, h, ii, jj, kk, X);
          X[d] = x0[d] + (double)f * h;
          to_con(c, ql, X);
          to_con(c, qr, X);
        };
        if constexpr (SUB) {
          double pl, csl, pr, csr;
          if constexpr (EOS == EOS_HELMHOLTZ) {
            // The Helmholtz evaluation gathers 36 table coefficients: with both face states live
            // through it the kernel spills.  So the states give the face velocity and the EOS
            // inputs, the two EOS calls run with nothing else live, and the states are formed
            // again from the staged primitives for the fluxes (the same values).
            double rl, Tl, Al, rr, Tr, Ar;
            {
              double ql[NF], qr[NF], X[3];
              face_states(ql, qr, X);
              double uf[NF];
#pragma unroll
              for (int l = 0; l < NF; l++) uf[l] = (ql[l] + qr[l]) * 0.5;
              Vf[fb][q] = (T)(d == 0 ? vx(c, uf, X) : (d == 1 ? vy(c, uf, X) : vz(uf)));
              rl = ql[D0], Tl = ql[TT], Al = Abar(c, ql);
              rr = qr[D0], Tr = qr[TT], Ar = Abar(c, qr);
            }
            eos_p_cs_in<EOS>(H, c, rl, Tl, Al, pl, csl);
            eos_p_cs_in<EOS>(H, c, rr, Tr, Ar, pr, csr);
            // the staged primitives are read again rather than kept through the EOS calls
            asm volatile("" ::: "memory");
          }
          double ql[NF], qr[NF], X[3];
          face_states(ql, qr, X);
          if constexpr (EOS != EOS_HELMHOLTZ) {
            double uf[NF];
#pragma unroll
            for (int l = 0; l < NF; l++) uf[l] = (ql[l] + qr[l]) * 0.5;
            Vf[fb][q] = (T)(d == 0 ? vx(c, uf, X) : (d == 1 ? vy(c, uf, X) : vz(uf)));
            eos_p_cs<EOS>(H, c, ql, pl, csl);
            eos_p_cs<EOS>(H, c, qr, pr, csr);
          }
          double fl[NF], fr[NF];
          const double el = flux_pc<d>(c, ql, X, pl, csl, fl);
          const double er = flux_pc<d>(c, qr, X, pr, csr, fr);
          const double amax = rmax(el, er);
#pragma unroll
          for (int l = 0; l < NF; l++) Fb[fb][l][q] = (T)(((fl[l] + fr[l]) - (qr[l] - ql[l]) * amax) * 0.5);
          if constexpr (EIGM) emax = rmax(emax, amax);
        } else {
          double ql[NF], qr[NF], X[3];
          face_states(ql, qr, X);
          double F[NF];
          const double el = flux<EOS, d>(H, c, ql, X, F);
          const double er = flux<EOS, d>(H, c, qr, X, F);
          emax = rmax(emax, rmax(el, er));
        }
      }
    };

    // MODE 0, once every face of direction d is in Fb[fb]: the octree's reflux overrides, exports
    // and ledger, then each cell subtracts its flux difference (D takes x, then y, then z:
    // grid/grid.cpp:304-344) and keeps its two face velocities
    auto finish = [&](auto dc, int fb) {
      constexpr int d = decltype(dc)::value;
      if constexpr (TREE) {
        // adjust_*_flux (grid/grid_node.cpp:1076-1274): faces of the node's low, middle and high
        // planes next to a finer region take the means its fine faces exported
#pragma unroll 1
        for (int q2 = tid; q2 < 3 * N2; q2 += N3) {
          const int pl = q2 / N2, ta = q2 % N, tb = (q2 / N) % N;
          const int qa = ta >= N / 2, qb = tb >= N / 2;
          const int sl = a.slot[((b * 3 + d) * 3 + pl) * 4 + qa + 2 * qb];
          if (sl < 0) continue;
          const int f = pl == 0 ? 0 : (pl == 1 ? N / 2 : N);
          const double* o = a.ovr + (((long)sl * (N / 2) + (tb - qb * (N / 2))) * (N / 2) + (ta - qa * (N / 2))) * NF;
          const int q = face_index<N, d>(f, ta, tb);
#pragma unroll
          for (int l = 0; l < NF; l++) Fb[fb][l][q] = (T)o[l];
        }
        __syncthreads();
        // this node's boundary faces, 2x2 means, to the coarse node's slot
        if (tid < 4 * (N / 2) * (N / 2)) {
          const int w = tid / ((N / 2) * (N / 2)), r = tid % ((N / 2) * (N / 2));
          const int side = w >> 1;
          const int sl = a.exp_slot[(b * 6 + 2 * d + side) * 2 + (w & 1)];
          if (sl >= 0) {
            const int al = r % (N / 2), bl = r / (N / 2);
            const int fp = side ? N : 0;
            const int q00 = face_index<N, d>(fp, 2 * al, 2 * bl), q10 = face_index<N, d>(fp, 2 * al + 1, 2 * bl);
            const int q01 = face_index<N, d>(fp, 2 * al, 2 * bl + 1), q11 = face_index<N, d>(fp, 2 * al + 1, 2 * bl + 1);
            double* o = a.ovr + (((long)sl * (N / 2) + bl) * (N / 2) + al) * NF;
#pragma unroll
            for (int l = 0; l < NF; l++) {
              double v = (double)Fb[fb][l][q00];
              v += (double)Fb[fb][l][q10];
              v += (double)Fb[fb][l][q01];
              v += (double)Fb[fb][l][q11];
              v *= 0.25;
              o[l] = v;
            }
          }
        }
        // the root's boundary faces for the flow-off ledger (grid/grid.cpp:350-404)
        if (b == 0 && a.ledger) {
#pragma unroll 1
          for (int q2 = tid; q2 < 2 * N2; q2 += N3) {
            const int side = q2 / N2, ta = q2 % N, tb = (q2 / N) % N;
            const int q = face_index<N, d>(side ? N : 0, ta, tb);
            double* o = a.ledger + ((((long)d * 2 + side) * N + tb) * N + ta) * NF;
#pragma unroll
            for (int l = 0; l < NF; l++) o[l] = (double)Fb[fb][l][q];
          }
        }
      }
      const int p = d == 0 ? ci : (d == 1 ? cj : ck);
      const int ta = d == 0 ? cj : ci;
      const int tb = d == 2 ? cj : ck;
      const int q0 = face_index<N, d>(p, ta, tb), q1 = face_index<N, d>(p + 1, ta, tb);
      // (with FP32 fluxes the FP64 difference is exact: the divergence still telescopes)
#pragma unroll
      for (int l = 0; l < NF; l++) D[l] -= ((double)Fb[fb][l][q1] - (double)Fb[fb][l][q0]) * hinv;
      vlo[d] = (double)Vf[fb][q0];
      vhi[d] = (double)Vf[fb][q1];
    };

    using X0 = std::integral_constant<int, 0>;
    using X1 = std::integral_constant<int, 1>;
    using X2 = std::integral_constant<int, 2>;
    if constexpr (!ALL) {
      // one direction's pencils staged at a time
      auto sweep = [&](auto dc) {
        constexpr int d = decltype(dc)::value;
        stage(d);
#pragma unroll 1
        for (int q = tid; q < NFACE; q += N3) face(dc, q, 0);
        __syncthreads();
        if constexpr (SUB) {
          finish(dc, 0);
          __syncthreads();  // P, Fb and Vf are restaged for the next direction
        }
      };
      sweep(X0{});
      sweep(X1{});
      sweep(X2{});
    } else if constexpr (EOS == EOS_HELMHOLTZ) {
      // Helmholtz: single faces (a face pair keeps the second face's sides live through the first
      // face's two EOS evaluations, which spills; measured 5.40 vs 5.33 ms)
      // The whole record is staged, so the faces of the three directions form one list of
      // 3 NFACE, handed out in rounds of N3 threads: four rounds (the last one 3/8 full) instead of
      // two per direction with the second one 1/8 full.  NFACE is a multiple of the wavefront, so
      // every wavefront's faces share a direction.  Fb and Vf are double-buffered (x -> 0, y -> 1,
      // z -> 0): the z faces of round 2 are written after x is finished, those of round 3 while y is.
      static_assert(NFACE % 64 == 0 && 2 * N3 >= NFACE && 2 * N3 <= 2 * NFACE && 3 * N3 >= 2 * NFACE &&
                        3 * N3 <= 3 * NFACE && 4 * N3 >= 3 * NFACE,
                    "the round schedule below assumes N = 8");
      auto face_any = [&](int g) {
        if (g < NFACE)
          face(X0{}, g, 0);
        else if (g < 2 * NFACE)
          face(X1{}, g - NFACE, 1);
        else
          face(X2{}, g - 2 * NFACE, 0);
      };
      if constexpr (SUB) {
        // round 0: x; 1: x, then y; 2: y, then z; 3: the rest of z (one call site keeps the face
        // code inlined once per direction)
#pragma unroll 1
        for (int r = 0; r < 4; r++) {
          const int g = tid + r * N3;
          if (g < 3 * NFACE) face_any(g);
          if (r == 1) {
            __syncthreads();
            finish(X0{}, 0);
            __syncthreads();  // Fb[0] takes z faces next
          } else if (r == 2) {
            __syncthreads();
            finish(X1{}, 1);
          }
        }
        __syncthreads();
        finish(X2{}, 0);
      } else {
#pragma unroll 1
        for (int g = tid; g < 3 * NFACE; g += N3) face_any(g);
      }
    } else {
      // The whole record is staged, so the faces of all three directions are handed out together,
      // as pairs of adjacent faces of one pencil (faces 2j and 2j+1 for j < N/2, then face N
      // alone): N/2 + 1 tasks per pencil, 960 per block, two rounds of N3 threads.  A pair shares
      // the reconstruction of its middle cells (ppm_face_pair: 5 slopes, 4 interfaces and 3
      // monotonisations for two faces instead of 8, 6 and 4).  x tasks run along each pencil
      // across the lanes (its cells are contiguous in LDS), y and z tasks across the pencils.  The
      // tasks of one direction fill whole wavefronts, so each wavefront has one direction.  Fb and
      // Vf are double-buffered (x -> 0, y -> 1, z -> 0): round 0 holds x and the first y tasks,
      // round 1 the rest of y and z, whose faces go to buffer 0 once x is finished.
      constexpr int TPD = (N / 2 + 1) * N2;  // tasks per direction
      static_assert(N == 8 && TPD % 64 == 0 && TPD <= N3 && 3 * TPD <= 2 * N3 && N3 < 2 * TPD,
                    "the round schedule below assumes N = 8");
      constexpr int W = RECON == AMRH_RECON_PPM ? 7 : 5;  // cells f0 - W/2 .. f0 + W/2 of a pair
      static_assert(W / 2 <= G, "a pair's cells lie in the staged pencil");
      // face f of direction d at (ta, tb) from its reconstructed primitive sides: the face position,
      // to_con, the EOS and the KT flux into Fb[fb], Vf[fb] (MODE 0) or the wavespeed into emax
      auto emit = [&](auto dc, int f, int ta, int tb, int fb, float* ql, float* qr) {
        constexpr int d = decltype(dc)::value;
        float X[3];
        {
          const int ii = d == 0 ? f : ta;
          const int jj = d == 0 ? ta : (d == 1 ? f : tb);
          const int kk = d == 2 ? f : tb;
          double Xd[3];
          cell_pos(x0, h, ii, jj, kk, Xd);
          Xd[d] = x0[d] + (double)f * h;
          X[0] = (float)Xd[0], X[1] = (float)Xd[1], X[2] = (float)Xd[2];
        }
        sp::to_con(cf, ql, X);
        sp::to_con(cf, qr, X);
        float fl[NF], fr[NF];
        const float el = sp::flux<EOS, d>(H, cf, ql, X, fl);
        const float er = sp::flux<EOS, d>(H, cf, qr, X, fr);
        const float amax = fmaxf(el, er);
        if constexpr (SUB) {
          const int q = face_index<N, d>(f, ta, tb);
          float uf[NF];
#pragma unroll
          for (int l = 0; l < NF; l++) uf[l] = (ql[l] + qr[l]) * 0.5f;
          Vf[fb][q] = (T)sp::vel<d>(cf, uf, X, sp::Rc(X));
#pragma unroll
          for (int l = 0; l < NF; l++) Fb[fb][l][q] = (T)(((fl[l] + fr[l]) - (qr[l] - ql[l]) * amax) * 0.5f);
        }
        if constexpr (EIGM) emax = rmax(emax, (double)amax);
      };
      auto pair_task = [&](auto dc, int r, int fb) {
        constexpr int d = decltype(dc)::value;
        const int j = d == 0 ? r % (N / 2 + 1) : r / N2;
        const int pen = d == 0 ? r / (N / 2 + 1) : r % N2;
        const int ta = pen % N, tb = pen / N;
        const int f0 = 2 * j;
        int e[W];
#pragma unroll
        for (int m = 0; m < W; m++) {
          int p = f0 - W / 2 + m;
          if (p > N + G - 1) p = N + G - 1;  // face N alone: its unused partner reads a staged cell
          e[m] = staged(d, p, ta, tb);
        }
        float ql[NF], qr[NF], ql1[NF], qr1[NF];
#pragma unroll
        for (int l = 0; l < NF; l++) {
          float v[W];
#pragma unroll
          for (int m = 0; m < W; m++) v[m] = (float)P[l][e[m]];
          if constexpr (RECON == AMRH_RECON_PPM)
            f32::ppm_face_pair(v, ql[l], qr[l], ql1[l], qr1[l], l == POT, l == POT);
          else
            f32::minmod_face_pair(v, ql[l], qr[l], ql1[l], qr1[l], (float)c.theta);
        }
        // one emit site for both faces: the second face's sides move into place after the first
#pragma unroll 1
        for (int s = 0; s < 2; s++) {
          if (s == 1) {
            if (f0 + 1 > N) break;
#pragma unroll
            for (int l = 0; l < NF; l++) {
              ql[l] = ql1[l];
              qr[l] = qr1[l];
            }
          }
          emit(dc, f0 + s, ta, tb, fb, ql, qr);
        }
      };
      auto task_any = [&](int g) {
        if (g < TPD)
          pair_task(X0{}, g, 0);
        else if (g < 2 * TPD)
          pair_task(X1{}, g - TPD, 1);
        else
          pair_task(X2{}, g - 2 * TPD, 0);
      };
      if constexpr (SUB) {
        // one call site keeps the task code inlined once per direction
#pragma unroll 1
        for (int rd = 0; rd < 2; rd++) {
          const int g = tid + rd * N3;
          if (g < 3 * TPD) task_any(g);
          __syncthreads();
          if (rd == 0) {
            finish(X0{}, 0);
            __syncthreads();  // Fb[0] takes z faces next
          }
        }
        finish(X1{}, 1);
        finish(X2{}, 0);
      } else {
#pragma unroll 1
        for (int g = tid; g < 3 * TPD; g += N3) task_any(g);
      }
    }

    if constexpr (SUB) {
      const long e = (long)tid;
      double s[NF];
#pragma unroll
      for (int l = 0; l < NF; l++) s[l] = ub.ld(l, (unsigned)e);
      double X[3];
      cell_pos(x0, h, ci, cj, ck, X);
      // face-averaged force of the potential (poisson/poisson.cpp:223-247, binary/binary.cpp:225-227)
      constexpr int PP = N + 2;
      const double* ph = a.phi + (long)b * PP * PP * PP;
      auto phi_at = [&](int i, int j, int k) { return ph[((k + 1) * PP + (j + 1)) * PP + (i + 1)]; };
      const double pc = phi_at(ci, cj, ck);
      double gf[3];
      if (a.force) {
        // Binary::add_difs (binary/binary.cpp:225-227) on the Poisson face forces
        const long P3 = (long)PP * PP * PP;
        const double* fx = a.force + (long)b * P3;
        const double* fy = fx + a.nforce * P3;
        const double* fz = fy + a.nforce * P3;
        auto at = [&](int i, int j, int k) { return ((long)k * PP + j) * PP + i; };
        gf[0] = 0.5 * (fx[at(ci + 1, cj + 1, ck + 1)] + fx[at(ci + 2, cj + 1, ck + 1)]);
        gf[1] = 0.5 * (fy[at(ci + 1, cj + 1, ck + 1)] + fy[at(ci + 1, cj + 2, ck + 1)]);
        gf[2] = 0.5 * (fz[at(ci + 1, cj + 1, ck + 1)] + fz[at(ci + 1, cj + 1, ck + 2)]);
      } else {
        const double fxl = -(pc - phi_at(ci - 1, cj, ck)) / h;
        const double fxh = -(phi_at(ci + 1, cj, ck) - pc) / h;
        const double fyl = -(pc - phi_at(ci, cj - 1, ck)) / h;
        const double fyh = -(phi_at(ci, cj + 1, ck) - pc) / h;
        const double fzl = -(pc - phi_at(ci, cj, ck - 1)) / h;
        const double fzh = -(phi_at(ci, cj, ck + 1) - pc) / h;
        gf[0] = 0.5 * (fxl + fxh);
        gf[1] = 0.5 * (fyl + fyh);
        gf[2] = 0.5 * (fzl + fzh);
      }
      // gravity_source (binary/state_ztwd.cpp:277-288); the zero components add exactly as
      // Vector += of a zero-initialised source
      double sv[NF];
#pragma unroll
      for (int l = 0; l < NF; l++) sv[l] = 0.0;
      const double dd = s[D0];
      sv[SR] += dd * (X[0] * gf[0] + X[1] * gf[1]) / Rc(X);
      sv[LZ] += dd * (X[0] * gf[1] - X[1] * gf[0]);
      sv[rot::SZ] += dd * gf[2];
      sv[rot::ET] += D[POT];
      sv[rot::ET] -= D[D0] * s[POT] / dd;
      sv[rot::ET] += s[D0] * (dvx(c, X) * gf[0] + dvy(c, X) * gf[1]);
#pragma unroll
      for (int l = 0; l < NF; l++) D[l] += sv[l];
      // internal_energy_source with the divergence of the saved face velocities
      double du = 0.0;
      du += vhi[0];
      du += vhi[1];
      du += vhi[2];
      du -= vlo[0];
      du -= vlo[1];
      du -= vlo[2];
      du /= h;
      double p, cs;
      if constexpr (MIXED && EOS == EOS_HELMHOLTZ) {
        // the mixed form's cell pressure for -p div u and the geometric source: FP32 EOS
        const double A = Abar(c, s);
        float pf, csf;
        helm::p_cs32(H, (float)s[D0], (float)s[TT], (float)A, (float)(A / 2.0), pf, csf);
        p = pf;
        cs = csf;
      } else {
        eos_p_cs<EOS>(H, c, s, p, cs);
      }
#pragma unroll
      for (int l = 0; l < NF; l++) sv[l] = 0.0;
      sv[rot::TAU] = -p * du;
#pragma unroll
      for (int l = 0; l < NF; l++) D[l] += sv[l];
      D[POT] = 0.0;
      D[TT] = 0.0;
      // State::source (binary/state_ztwd.cpp:347-358)
#pragma unroll
      for (int l = 0; l < NF; l++) sv[l] = 0.0;
      {
        const double Rr = Rc(X);
        sv[SR] += (p + f32::pow_out_of_line(s[LZ] / Rr, 2.0) / s[D0]) / Rr;
        const double th = atan2_ool(X[1], X[0]);
        const double cth = cos_ool(th - c.ftheta), sth = sin_ool(th - c.ftheta);
        sv[SR] += s[D0] * c.fR * cth;
        sv[LZ] -= s[D0] * c.fR * sth * Rr;
        sv[rot::ET] += s[SR] * c.fR * cth;
        sv[rot::ET] -= st(c, s, X) * c.fR * sth * Rr;
      }
#pragma unroll
      for (int l = 0; l < NF; l++) D[l] += sv[l];
      if constexpr (MODE == 2) {
        // deferred: the cell's D (its sources included) for rot_stage_kernel
        const RecBuf<N3, NF> dbo(a.dbuf + (long)b * NF * N3);
#pragma unroll
        for (int l = 0; l < NF; l++) dbo.st_nt(l, (unsigned)e, D[l]);
      } else {
      double u[NF];
      if (has_u0) {
        const RecBuf<R, NF> u0(a.U0 + (long)b * NF * R);
#pragma unroll
        for (int l = 0; l < NF; l++) u[l] = (s[l] + D[l] * dt) * beta + u0.ld(l, (unsigned)e) * (1.0 - beta);
      } else {
#pragma unroll
        for (int l = 0; l < NF; l++) u[l] = (s[l] + D[l] * dt) * beta + s[l] * (1.0 - beta);
      }
      floor_state<EOS, MIXED>(H, c, u, X);
      if (!(u[D0] > 0.0)) atomicOr(a.err, 1);
      const double R2 = X[0] * X[0] + X[1] * X[1];
      u[POT] = u[D0] * (pc - 0.5 * R2 * c.Omega0 * c.Omega0);
      const RecBuf<R, NF> uo(a.Uout + (long)b * NF * R);
#pragma unroll
      for (int l = 0; l < NF; l++) uo.st(l, (unsigned)e, u[l]);
      }
    }
    if constexpr (EIGM) {
      for (int off = 32; off > 0; off >>= 1) emax = rmax(emax, __shfl_xor(emax, off, 64));
      if ((tid & 63) == 0) red[tid >> 6] = emax;
      __syncthreads();
      if (tid == 0) {
        double v = red[0];
        for (int w = 1; w < N3 / 64; w++) v = rmax(v, red[w]);
        if (a.block_eig) a.block_eig[b] = v;
        if (a.gmax) atomicMax(a.gmax, (unsigned long long)__double_as_longlong(v / h));
      }
    }
    __syncthreads();
  }
}

// An RK stage from the D a MODE 2 launch stored, one thread per interior cell: the epilogue of
// kt_rot_kernel MODE 0, U = (U_in + D dt) beta + U0 (1 - beta) (U_in for U0 when beta = 1), the
// floor with the Newton solve for T at the cell (binary/state_ztwd.cpp:482-554), and
// pot = rho (phi - R^2 Omega0^2 / 2) (binary/binary.cpp:825-849).  Same operations in the same
// order, so the stage has the bits the substep's own epilogue would have stored.  Used for the
// first stage of a deferred step, and for the later substeps of the Helmholtz EOS, whose Newton
// solve in the substep's epilogue (every register of the wave taken by the fluxes' state) costs
// more than this separate pass and the D round trip.
template <int G, int EOS, bool MIXED>
__global__ __launch_bounds__(256) void rot_stage_kernel(RotArgs a) {
  constexpr int R = N3 + 6 * G * N2;
  constexpr int PP = N + 2;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)a.nblocks * N3) return;
  const int b = (int)(gid / N3), e = (int)(gid % N3);
  const int ci = e % N, cj = (e / N) % N, ck = e / N2;
  const double dt = *a.dt_dev;
  const double beta = a.beta;
  const RecBuf<R, NF> ub(a.Uin + (long)b * NF * R), uo(a.Uout + (long)b * NF * R);
  const RecBuf<N3, NF> db(a.dbuf + (long)b * NF * N3);
  double s[NF], D[NF], u[NF];
#pragma unroll
  for (int l = 0; l < NF; l++) {
    s[l] = ub.ld(l, (unsigned)e);
    D[l] = db.ld_nt(l, (unsigned)e);
  }
  if (beta != 1.0) {
    const RecBuf<R, NF> u0(a.U0 + (long)b * NF * R);
#pragma unroll
    for (int l = 0; l < NF; l++) u[l] = (s[l] + D[l] * dt) * beta + u0.ld(l, (unsigned)e) * (1.0 - beta);
  } else {
#pragma unroll
    for (int l = 0; l < NF; l++) u[l] = (s[l] + D[l] * dt) * beta + s[l] * (1.0 - beta);
  }
  const double x0[3] = {a.x0[3L * b], a.x0[3L * b + 1], a.x0[3L * b + 2]};
  double X[3];
  cell_pos(x0, a.h[b], ci, cj, ck, X);
  floor_state<EOS, MIXED>(a.H, a.c, u, X);
  if (!(u[D0] > 0.0)) atomicOr(a.err, 1);
  const double pc = a.phi[(long)b * PP * PP * PP + ((ck + 1) * PP + (cj + 1)) * PP + (ci + 1)];
  const double R2 = X[0] * X[0] + X[1] * X[1];
  u[POT] = u[D0] * (pc - 0.5 * R2 * a.c.Omega0 * a.c.Omega0);
#pragma unroll
  for (int l = 0; l < NF; l++) uo.st(l, (unsigned)e, u[l]);
}

// GridPhysBound for isolated blocks (grid/grid_phys_bound.cpp:27-54): every face ghost copies the
// boundary cell next to it and applies enforce_outflow at the boundary face's position; the lower
// z face mirrors (s_z -> -s_z) with z_reflect
template <int G>
__global__ __launch_bounds__(256) void rot_ghost_kernel(double* U, const double* x0s, const double* hs, int nblocks,
                                                        int z_reflect, RotParams c) {
  constexpr int S = G * N2, R = N3 + 6 * S;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)nblocks * 6 * S) return;
  const int b = (int)(gid / (6 * S));
  const int e = N3 + (int)(gid % (6 * S));
  int i, j, k;
  elem_ijk<G>(e, i, j, k);
  const int nf = (e - N3) / S;
  const int d = nf >> 1, up = nf & 1;
  int si = i, sj = j, sk = k;  // source cell
  const bool mirror = z_reflect && nf == 4;
  int* pd = d == 0 ? &si : (d == 1 ? &sj : &sk);
  const int pg = *pd;  // ghost coordinate along d
  if (mirror)
    *pd = -1 - pg;  // interior cell G-1-t for ghost t (padded 2G-1-t)
  else
    *pd = up ? N - 1 : 0;
  double* Ub = U + (long)b * NF * R;
  const long src = ((long)sk * N + sj) * N + si;
  double u[NF];
#pragma unroll
  for (int l = 0; l < NF; l++) u[l] = Ub[(long)l * R + src];
  if (mirror) {
    u[rot::SZ] = -u[rot::SZ];
  } else {
    const double h = hs[b];
    double X[3];
    cell_pos(x0s + 3L * b, h, si, sj, sk, X);
    X[d] = x0s[3L * b + d] + (double)(up ? N : 0) * h;  // the boundary face
    enforce_outflow(c, u, X, nf);
  }
#pragma unroll
  for (int l = 0; l < NF; l++) Ub[(long)l * R + e] = u[l];
}

// GridNode::enforce_boundaries (grid/grid_node.cpp:59-112) for one level of the octree: a face
// slab is copied from the same-level neighbour's interior, interpolated from the parent
// (GridAMRBound, grid/grid_amr_bound.cpp:27-78: the parent cell minus a quarter of the minmod slope
// along the face normal, the sign by the fine cell's parity) or filled by the physical boundary
// as for isolated blocks
template <int G>
__global__ __launch_bounds__(256) void rot_tree_ghost_kernel(double* U, const int* nodes, int count, const int* ghost,
                                                             const int* rel, const double* x0s, const double* hs,
                                                             int z_reflect, RotParams c, int* err,
                                                             const unsigned char* covered) {
  constexpr int S = G * N2, R = N3 + 6 * S;
  constexpr Layout L{N, G, NF};
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)count * 6 * S) return;
  const int b = nodes[gid / (6 * S)];
  const int e = N3 + (int)(gid % (6 * S));
  int q[3];
  elem_ijk<G>(e, q[0], q[1], q[2]);
  const int nf = (e - N3) / S;
  const int d = nf >> 1, up = nf & 1;
  const int type = ghost[(b * 6 + nf) * 2], src = ghost[(b * 6 + nf) * 2 + 1];
  // covered != nullptr: the sibling faces of leaves are left to the substep, which reads them
  // from the sibling; a refined node's are filled (its children's coarse ghosts read them)
  if (covered && type == GHOST_SIB && covered[b] == 0) return;
  double* Ub = U + (long)b * NF * R;
  double u[NF];
  if (type == GHOST_SIB) {
    int s3[3] = {q[0], q[1], q[2]};
    s3[d] = up ? q[d] - N : q[d] + N;  // the neighbour's cell at the same distance from the face
    const double* Us = U + (long)src * NF * R;
    const long es = ((long)s3[2] * N + s3[1]) * N + s3[0];
#pragma unroll
    for (int l = 0; l < NF; l++) u[l] = Us[(long)l * R + es];
  } else if (type == GHOST_PHYS) {
    int s3[3] = {q[0], q[1], q[2]};
    const bool mirror = z_reflect && nf == 4;
    s3[d] = mirror ? -1 - q[d] : (up ? N - 1 : 0);
    const long es = ((long)s3[2] * N + s3[1]) * N + s3[0];
#pragma unroll
    for (int l = 0; l < NF; l++) u[l] = Ub[(long)l * R + es];
    if (mirror) {
      u[rot::SZ] = -u[rot::SZ];
    } else {
      const double h = hs[b];
      double X[3];
      cell_pos(x0s + 3L * b, h, s3[0], s3[1], s3[2], X);
      X[d] = x0s[3L * b + d] + (double)(up ? N : 0) * h;
      enforce_outflow(c, u, X, nf);
    }
  } else {
    const double* Up = U + (long)src * NF * R;
    int i1[3], par = 0;
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const int f = rel[b * 3 + t] + q[t] + G;  // parent-padded index of the fine cell, doubled
      i1[t] = f >> 1;
      if (t == d) par = f & 1;
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
    for (int l = 0; l < NF; l++) {
      const double uc = Up[(long)l * R + e0];
      double m = x64::minmod2(Up[(long)l * R + ep] - uc, uc - Up[(long)l * R + em]);
      if (par) m = 0.0 - m;
      u[l] = uc - m * 0.25;
    }
  }
#pragma unroll
  for (int l = 0; l < NF; l++) Ub[(long)l * R + e] = u[l];
}

// GridNode::inject_from_children (grid/grid_node.cpp:1035-1074): a parent cell is the mean of its
// 8 children's cells, floored at the parent cell's position (binary/state_ztwd.cpp:482-554)
template <int G, int EOS>
__global__ __launch_bounds__(256) void rot_restrict_kernel(double* U, const int2* pairs, int count, const double* x0s,
                                                           const double* hs, RotParams c, helm::HelmDev H) {
  constexpr int R = N3 + 6 * G * N2;
  constexpr int HN = N / 2;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)count * HN * HN * HN) return;
  const int2 pc = pairs[gid / (HN * HN * HN)];
  const int r = (int)(gid % (HN * HN * HN));
  const int parent = pc.x, ch = (pc.y >> 24) & 7, child = pc.y & 0xffffff;
  const int ii = r % HN, jj = (r / HN) % HN, kk = r / (HN * HN);
  const int i = 2 * ii, j = 2 * jj, k = 2 * kk;
  const double* Uc = U + (long)child * NF * R;
  auto at = [](int x, int y, int z) { return ((long)z * N + y) * N + x; };
  double u[NF];
#pragma unroll
  for (int l = 0; l < NF; l++) {
    const double* f = Uc + (long)l * R;
    double s = 0.0;
    s += f[at(i + 0, j + 0, k + 0)];
    s += f[at(i + 1, j + 0, k + 0)];
    s += f[at(i + 0, j + 1, k + 0)];
    s += f[at(i + 1, j + 1, k + 0)];
    s += f[at(i + 0, j + 0, k + 1)];
    s += f[at(i + 1, j + 0, k + 1)];
    s += f[at(i + 0, j + 1, k + 1)];
    s += f[at(i + 1, j + 1, k + 1)];
    s *= 0.125;
    u[l] = s;
  }
  const int pi = (ch & 1) * HN + ii, pj = ((ch >> 1) & 1) * HN + jj, pk = ((ch >> 2) & 1) * HN + kk;
  double X[3];
  cell_pos(x0s + 3L * parent, hs[parent], pi, pj, pk, X);
  floor_state<EOS>(H, c, u, X);
  double* Up = U + (long)parent * NF * R;
#pragma unroll
  for (int l = 0; l < NF; l++) Up[(long)l * R + at(pi, pj, pk)] = u[l];
}

// GridNode::inject_from_parent (grid/grid_node.cpp:193-233) of the rotating state: one thread per
// 2x2x2 group of a new child's cells, minmod slopes of the parent's cells per axis, each child cell
// floored at its own position
template <int G, int EOS>
__global__ __launch_bounds__(64) void rot_prolong_kernel(double* U, const int3* nodes, int count, const double* x0s,
                                                         const double* hs, RotParams c, helm::HelmDev H) {
  constexpr int R = N3 + 6 * G * N2;
  constexpr int HN = N / 2, GP = N + 2 * G;
  constexpr Layout L{N, G, NF};
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)count * HN * HN * HN) return;
  const int3 t = nodes[gid / (HN * HN * HN)];
  const int r = (int)(gid % (HN * HN * HN));
  const int node = t.x, parent = t.y, ch = t.z;
  const int cx = ch & 1, cy = (ch >> 1) & 1, cz = (ch >> 2) & 1;
  const int ii = r % HN, jj = (r / HN) % HN, kk = r / (HN * HN);
  // the parent's padded cell of the group (grid/grid_node.cpp:204-206)
  const int i0 = (G + G + 2 * ii) / 2 + cx * (GP / 2 - G);
  const int j0 = (G + G + 2 * jj) / 2 + cy * (GP / 2 - G);
  const int k0 = (G + G + 2 * kk) / 2 + cz * (GP / 2 - G);
  const double* Up = U + (long)parent * NF * R;
  double* Uc = U + (long)node * NF * R;
  const long e0 = L.from_padded(i0, j0, k0);
  const long exm = L.from_padded(i0 - 1, j0, k0), exp_ = L.from_padded(i0 + 1, j0, k0);
  const long eym = L.from_padded(i0, j0 - 1, k0), eyp = L.from_padded(i0, j0 + 1, k0);
  const long ezm = L.from_padded(i0, j0, k0 - 1), ezp = L.from_padded(i0, j0, k0 + 1);
  double s1[NF], s2[NF], s3[NF], u0[NF];
#pragma unroll
  for (int l = 0; l < NF; l++) {
    const double* f = Up + (long)l * R;
    const double u = f[e0];
    u0[l] = u;
    s1[l] = x64::minmod2(f[exp_] - u, u - f[exm]);
    s2[l] = x64::minmod2(f[eyp] - u, u - f[eym]);
    s3[l] = x64::minmod2(f[ezp] - u, u - f[ezm]);
  }
#pragma unroll 1
  for (int q = 0; q < 8; q++) {
    const int qx = q & 1, qy = (q >> 1) & 1, qz = q >> 2;
    double v[NF];
#pragma unroll
    for (int l = 0; l < NF; l++) {
      const double u = u0[l];
      switch (q) {
        case 0: v[l] = u - (s1[l] + s2[l] + s3[l]) * 0.25; break;
        case 1: v[l] = u + (s1[l] - s2[l] - s3[l]) * 0.25; break;
        case 2: v[l] = u - (s1[l] - s2[l] + s3[l]) * 0.25; break;
        case 3: v[l] = u + (s1[l] + s2[l] - s3[l]) * 0.25; break;
        case 4: v[l] = u - (s1[l] + s2[l] - s3[l]) * 0.25; break;
        case 5: v[l] = u + (s1[l] - s2[l] + s3[l]) * 0.25; break;
        case 6: v[l] = u - (s1[l] - s2[l] - s3[l]) * 0.25; break;
        default: v[l] = u + (s1[l] + s2[l] + s3[l]) * 0.25; break;
      }
    }
    const int ci = 2 * ii + qx, cj = 2 * jj + qy, ck = 2 * kk + qz;
    double X[3];
    cell_pos(x0s + 3L * node, hs[node], ci, cj, ck, X);
    floor_state<EOS>(H, c, v, X);
    const long e = ((long)ck * N + cj) * N + ci;
#pragma unroll
    for (int l = 0; l < NF; l++) Uc[(long)l * R + e] = v[l];
  }
}

// Binary::pot_to_grid (binary/binary.cpp:825-849)
template <int G>
__global__ __launch_bounds__(256) void rot_pot_kernel(double* U, const double* phi, const double* x0s,
                                                      const double* hs, int nblocks, RotParams c) {
  constexpr int R = N3 + 6 * G * N2;
  constexpr int PP = N + 2;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)nblocks * N3) return;
  const int b = (int)(gid / N3), e = (int)(gid % N3);
  const int i = e % N, j = (e / N) % N, k = e / N2;
  double X[3];
  cell_pos(x0s + 3L * b, hs[b], i, j, k, X);
  double* Ub = U + (long)b * NF * R;
  const double d = Ub[(long)D0 * R + e];
  const double R2 = X[0] * X[0] + X[1] * X[1];
  const double ph = phi[(long)b * PP * PP * PP + ((k + 1) * PP + (j + 1)) * PP + (i + 1)];
  Ub[(long)POT * R + e] = d * (ph - 0.5 * R2 * c.Omega0 * c.Omega0);
}

template <int G, int RECON, int EOS, int MODE, bool MIXED>
hipError_t launch_one(const RotArgs& a, int grid, hipStream_t s) {
  if (a.nblocks <= 0) return hipSuccess;
  const int g = grid > 0 && grid < a.nblocks ? grid : a.nblocks;
  if (MODE != 1 && a.slot)
    kt_rot_kernel<G, RECON, EOS, MODE, MIXED, MODE != 1><<<g, N3, 0, s>>>(a);
  else
    kt_rot_kernel<G, RECON, EOS, MODE, MIXED, false><<<g, N3, 0, s>>>(a);
  return hipGetLastError();
}

template <int MODE, bool MIXED>
hipError_t dispatch(int g, int recon, int eos, const RotArgs& a, int grid, hipStream_t s) {
  if (g == 3 && recon == AMRH_RECON_PPM)
    return eos == EOS_HELMHOLTZ ? launch_one<3, AMRH_RECON_PPM, EOS_HELMHOLTZ, MODE, MIXED>(a, grid, s)
                                : launch_one<3, AMRH_RECON_PPM, EOS_IDEAL, MODE, MIXED>(a, grid, s);
  if (g == 2 && recon == AMRH_RECON_MINMOD)
    return eos == EOS_HELMHOLTZ ? launch_one<2, AMRH_RECON_MINMOD, EOS_HELMHOLTZ, MODE, MIXED>(a, grid, s)
                                : launch_one<2, AMRH_RECON_MINMOD, EOS_IDEAL, MODE, MIXED>(a, grid, s);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_rot_substep(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s) {
  return mixed ? dispatch<0, true>(g, recon, eos, a, grid, s) : dispatch<0, false>(g, recon, eos, a, grid, s);
}

hipError_t launch_rot_maxeig(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s) {
  return mixed ? dispatch<1, true>(g, recon, eos, a, grid, s) : dispatch<1, false>(g, recon, eos, a, grid, s);
}
hipError_t launch_rot_deferred(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s) {
  if (a.slot || a.nodes || !a.outflow || !a.dbuf) return hipErrorInvalidValue;
  return mixed ? dispatch<2, true>(g, recon, eos, a, grid, s) : dispatch<2, false>(g, recon, eos, a, grid, s);
}
hipError_t launch_rot_tree_deferred(int g, int recon, int eos, bool mixed, const RotArgs& a, int grid, hipStream_t s) {
  if (!a.slot || !a.nodes || a.outflow || !a.dbuf || !a.gmax) return hipErrorInvalidValue;
  return mixed ? dispatch<2, true>(g, recon, eos, a, grid, s) : dispatch<2, false>(g, recon, eos, a, grid, s);
}
hipError_t launch_rot_stage(int g, int eos, bool mixed, const RotArgs& a, hipStream_t s) {
  if (!a.dbuf || !a.dt_dev || (a.beta != 1.0 && !a.U0)) return hipErrorInvalidValue;
  if (a.nblocks <= 0) return hipSuccess;
  const unsigned grid = (unsigned)(((long)a.nblocks * N3 + 255) / 256);
  const bool helm = eos == EOS_HELMHOLTZ;
  if (g == 3) {
    if (mixed)
      helm ? rot_stage_kernel<3, EOS_HELMHOLTZ, true><<<grid, 256, 0, s>>>(a)
           : rot_stage_kernel<3, EOS_IDEAL, true><<<grid, 256, 0, s>>>(a);
    else
      helm ? rot_stage_kernel<3, EOS_HELMHOLTZ, false><<<grid, 256, 0, s>>>(a)
           : rot_stage_kernel<3, EOS_IDEAL, false><<<grid, 256, 0, s>>>(a);
  } else if (g == 2) {
    if (mixed)
      helm ? rot_stage_kernel<2, EOS_HELMHOLTZ, true><<<grid, 256, 0, s>>>(a)
           : rot_stage_kernel<2, EOS_IDEAL, true><<<grid, 256, 0, s>>>(a);
    else
      helm ? rot_stage_kernel<2, EOS_HELMHOLTZ, false><<<grid, 256, 0, s>>>(a)
           : rot_stage_kernel<2, EOS_IDEAL, false><<<grid, 256, 0, s>>>(a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rot_tree_ghosts(int g, double* U, const int* nodes, int count, const int* ghost, const int* rel,
                                  const double* x0, const double* h, int z_reflect, const RotParams& c, int* err,
                                  hipStream_t s, const unsigned char* covered) {
  const long total = (long)count * 6 * g * N2;
  if (total == 0) return hipSuccess;
  const unsigned grid = (unsigned)((total + 255) / 256);
  if (g == 3)
    rot_tree_ghost_kernel<3><<<grid, 256, 0, s>>>(U, nodes, count, ghost, rel, x0, h, z_reflect, c, err, covered);
  else if (g == 2)
    rot_tree_ghost_kernel<2><<<grid, 256, 0, s>>>(U, nodes, count, ghost, rel, x0, h, z_reflect, c, err, covered);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rot_restrict(int g, int eos, double* U, const int2* pairs, int count, const double* x0,
                               const double* h, const RotParams& c, const helm::HelmDev& H, hipStream_t s) {
  const long total = (long)count * (N / 2) * (N / 2) * (N / 2);
  if (total == 0) return hipSuccess;
  const unsigned grid = (unsigned)((total + 255) / 256);
  const bool helm = eos == EOS_HELMHOLTZ;
  if (g == 3 && helm)
    rot_restrict_kernel<3, EOS_HELMHOLTZ><<<grid, 256, 0, s>>>(U, pairs, count, x0, h, c, H);
  else if (g == 3)
    rot_restrict_kernel<3, EOS_IDEAL><<<grid, 256, 0, s>>>(U, pairs, count, x0, h, c, H);
  else if (g == 2 && helm)
    rot_restrict_kernel<2, EOS_HELMHOLTZ><<<grid, 256, 0, s>>>(U, pairs, count, x0, h, c, H);
  else if (g == 2)
    rot_restrict_kernel<2, EOS_IDEAL><<<grid, 256, 0, s>>>(U, pairs, count, x0, h, c, H);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rot_prolong(int g, int eos, double* U, const int3* nodes, int count, const double* x0,
                              const double* h, const RotParams& c, const helm::HelmDev& H, hipStream_t s) {
  const long total = (long)count * (N / 2) * (N / 2) * (N / 2);
  if (total == 0) return hipSuccess;
  const unsigned grid = (unsigned)((total + 63) / 64);
  const bool helm = eos == EOS_HELMHOLTZ;
  if (g == 3 && helm)
    rot_prolong_kernel<3, EOS_HELMHOLTZ><<<grid, 64, 0, s>>>(U, nodes, count, x0, h, c, H);
  else if (g == 3)
    rot_prolong_kernel<3, EOS_IDEAL><<<grid, 64, 0, s>>>(U, nodes, count, x0, h, c, H);
  else if (g == 2 && helm)
    rot_prolong_kernel<2, EOS_HELMHOLTZ><<<grid, 64, 0, s>>>(U, nodes, count, x0, h, c, H);
  else if (g == 2)
    rot_prolong_kernel<2, EOS_IDEAL><<<grid, 64, 0, s>>>(U, nodes, count, x0, h, c, H);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rot_pot_to_grid(int g, double* U, const double* phi, const double* x0, const double* h,
                                  int nblocks, const RotParams& c, hipStream_t s) {
  const long total = (long)nblocks * N3;
  if (total == 0) return hipSuccess;
  const unsigned grid = (unsigned)((total + 255) / 256);
  if (g == 3)
    rot_pot_kernel<3><<<grid, 256, 0, s>>>(U, phi, x0, h, nblocks, c);
  else if (g == 2)
    rot_pot_kernel<2><<<grid, 256, 0, s>>>(U, phi, x0, h, nblocks, c);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rot_ghosts(int g, double* U, const double* x0, const double* h, int nblocks, int z_reflect,
                             const RotParams& c, hipStream_t s) {
  const long total = (long)nblocks * 6 * g * N2;
  if (total == 0) return hipSuccess;
  const unsigned grid = (unsigned)((total + 255) / 256);
  if (g == 3)
    rot_ghost_kernel<3><<<grid, 256, 0, s>>>(U, x0, h, nblocks, z_reflect, c);
  else if (g == 2)
    rot_ghost_kernel<2><<<grid, 256, 0, s>>>(U, x0, h, nblocks, z_reflect, c);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace amrh
