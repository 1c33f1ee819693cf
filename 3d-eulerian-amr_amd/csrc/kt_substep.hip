// Fused KT substep, CFL wavespeed scan, ghost fill and layout conversion kernels (gfx950).
// See kt_kernels.h for the sweep scheme and amrh_layout.h for the record layout.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "amrh_internal.h"
#include "cell_ops.h"
#include "kt_kernels.h"
#include "rec_buf.h"

namespace amrh {

namespace {

constexpr int GW = 2;  // ghost width of the minmod scheme (euler/euler_defs.h:25)

template <bool MIXED, int D>
__device__ __forceinline__ void face_flux(const PhysParams& P, const typename Prec<MIXED>::T* pm2,
                                          const typename Prec<MIXED>::T* pm1,
                                          const typename Prec<MIXED>::T* p0,
                                          const typename Prec<MIXED>::T* pp1,
                                          typename Prec<MIXED>::T* F) {
  using T = typename Prec<MIXED>::T;
  T sl[NFE], sr[NFE];
  if constexpr (!MIXED) {
    x64::slope(pm2, pm1, p0, sl, P.theta);
    x64::slope(pm1, p0, pp1, sr, P.theta);
    double e;
    x64::kt_face<D>(P, pm1, sl, p0, sr, F, e);
  } else {
    const f32::PhysF PF{(float)P.gamma, (float)(P.gamma - 1.0), (float)P.theta};
    f32::slope(pm2, pm1, p0, sl, PF.theta);
    f32::slope(pm1, p0, pp1, sr, PF.theta);
    float e;
    f32::kt_face<D>(PF, pm1, sl, p0, sr, F, e);
  }
}

template <bool MIXED, int D>
__device__ __forceinline__ double face_eig(const PhysParams& P, const typename Prec<MIXED>::T* pm2,
                                           const typename Prec<MIXED>::T* pm1,
                                           const typename Prec<MIXED>::T* p0,
                                           const typename Prec<MIXED>::T* pp1) {
  using T = typename Prec<MIXED>::T;
  T sl[NFE], sr[NFE];
  if constexpr (!MIXED) {
    x64::slope(pm2, pm1, p0, sl, P.theta);
    x64::slope(pm1, p0, pp1, sr, P.theta);
    return x64::face_eig<D>(P, pm1, sl, p0, sr);
  } else {
    const f32::PhysF PF{(float)P.gamma, (float)(P.gamma - 1.0), (float)P.theta};
    f32::slope(pm2, pm1, p0, sl, PF.theta);
    f32::slope(pm1, p0, pp1, sr, PF.theta);
    return (double)f32::face_eig<D>(PF, pm1, sl, p0, sr);
  }
}

// Reflux: a face on plane 0, N/2 or N of direction d whose quadrant has a fine neighbour on exactly
// one side takes the mean of the four fine faces over it (grid/grid_node.cpp:1076-1140); the means
// were written to a.ovr by the export kernel.  (ta, tb): the face's tangential cell indices in
// (x, y, z) order with d left out.
template <int N, typename T>
__device__ __forceinline__ void apply_override(const SubstepArgs& a, int b, int d, int fi, int ta, int tb,
                                               T* F) {
  const int p = fi == 0 ? 0 : (fi == N / 2 ? 1 : (fi == N ? 2 : -1));
  if (p < 0) return;
  const int qa = ta >= N / 2, qb = tb >= N / 2;
  const int s = a.slot[((b * 3 + d) * 3 + p) * 4 + qa + 2 * qb];
  if (s < 0) return;
  const double* o = a.ovr + (((long)s * (N / 2) + (tb - qb * (N / 2))) * (N / 2) + (ta - qa * (N / 2))) * NFE;
#pragma unroll
  for (int l = 0; l < NFE; l++) F[l] = (T)o[l];
}

// element offset of the column cell (i, j, kk), kk in [-G, N+G)
template <int N>
__device__ __forceinline__ long col_elem(const Layout& L, int i, int j, int kk) {
  if (kk < 0) return L.slab(ZL) + ((long)(kk + GW) * N + j) * N + i;
  if (kk >= N) return L.slab(ZU) + ((long)(kk - N) * N + j) * N + i;
  return L.interior(i, j, kk);
}

template <bool MIXED>
__device__ __forceinline__ void slope_of(const PhysParams& P, const typename Prec<MIXED>::T* a,
                                         const typename Prec<MIXED>::T* b, const typename Prec<MIXED>::T* c,
                                         typename Prec<MIXED>::T* s) {
  if constexpr (!MIXED)
    x64::slope(a, b, c, s, P.theta);
  else
    f32::slope(a, b, c, s, (float)P.theta);
}

template <bool MIXED, int D>
__device__ __forceinline__ void face_of(const PhysParams& P, const typename Prec<MIXED>::T* pl,
                                        const typename Prec<MIXED>::T* sl, const typename Prec<MIXED>::T* pr,
                                        const typename Prec<MIXED>::T* sr, typename Prec<MIXED>::T* F) {
  if constexpr (!MIXED) {
    double e;
    x64::kt_face<D>(P, pl, sl, pr, sr, F, e);
  } else {
    const f32::PhysF PF{(float)P.gamma, (float)(P.gamma - 1.0), (float)P.theta};
    float e;
    f32::kt_face<D>(PF, pl, sl, pr, sr, F, e);
  }
}

template <bool MIXED, int D>
__device__ __forceinline__ double eig_of(const PhysParams& P, const typename Prec<MIXED>::T* pl,
                                         const typename Prec<MIXED>::T* sl, const typename Prec<MIXED>::T* pr,
                                         const typename Prec<MIXED>::T* sr) {
  if constexpr (!MIXED) {
    return x64::face_eig<D>(P, pl, sl, pr, sr);
  } else {
    const f32::PhysF PF{(float)P.gamma, (float)(P.gamma - 1.0), (float)P.theta};
    return (double)f32::face_eig<D>(PF, pl, sl, pr, sr);
  }
}

// position of boundary column (i, j) of an N x N plane along its perimeter, 0 .. 4N-5 (each corner
// once): the row j = 0, then the column i = N-1, the row j = N-1 backwards, the column i = 0 down
// The mixed form's outflow ghost from the boundary cell's FP32 primitives: the conserved-state
// fix (outflow_fix) zeroes the inflowing momentum and takes its kinetic energy out of E, which
// leaves the internal energy; so the ghost's primitives are the cell's with v_d zeroed on inflow
// (the sign of v_d is that of s_d: rho > 0), or v_z negated for the z mirror.  Exact operations.
__device__ __forceinline__ void ghost_post_prim(float* p, int post) {
  if (post == kMirror) {
    p[SZ] = -p[SZ];
  } else if (post >= 0) {
    const int d = post >> 1;
    const float v = p[SX + d];
    if ((post & 1) ? (v < 0.0f) : (v > 0.0f)) p[SX + d] = 0.0f;
  }
}

template <int N>
__device__ __forceinline__ int perim(int i, int j) {
  if (j == 0) return i;
  if (i == N - 1) return N - 1 + j;
  if (j == N - 1) return 3 * N - 3 - i;
  return 4 * N - 4 - j;
}  // ghost post-processing: the lower z face's mirror (sz -> -sz)


// Record element of column cell kk in [-G, N+G), and how the fetched cell becomes the value used:
// a stored cell (interior or ghost slab), or for a self-ghosting outflow block the boundary cell
// the ghost copies -- the nearest plane, or plane -1-kk mirrored below with z_reflect (the ghost
// at padded z = t holds padded 2G-1-t, grid/grid_phys_bound.cpp:139-166).
template <int N, bool SELF>
__device__ __forceinline__ unsigned col_src(const Layout& L, int i, int j, int kk, bool zr, int& post) {
  post = -1;
  if (!SELF || (kk >= 0 && kk < N)) return (unsigned)col_elem<N>(L, i, j, kk);
  if (kk < 0) {
    post = zr ? kMirror : (int)ZL;
    return (unsigned)L.interior(i, j, zr ? -1 - kk : 0);
  }
  post = ZU;
  return (unsigned)L.interior(i, j, N - 1);
}

// The first RK stage of a step whose dt was not known when its fluxes were formed (MODE 2 stored
// the divergence instead): U1 = floor((U^n + D dt) * 1), the same operations the substep's update
// makes with beta = 1 (grid/grid.cpp:406-420, euler/state.cpp:86-92), so U1 has the bits a direct
// first substep would have stored.  FP64: d = -(dFx + dFy + dFz)/h in the reference's order; mixed:
// d = the FP64 sum of the FP32 flux differences (h applied here, as the mixed update does).
template <bool MIXED>
__device__ __forceinline__ void first_stage(const PhysParams& P, double* u, const double* d, double dt,
                                            double hinv) {
#pragma unroll
  for (int l = 0; l < NFE; l++) {
    if constexpr (!MIXED)
      u[l] = (u[l] + d[l] * dt) * 1.0;
    else
      u[l] = fma(-dt * hinv, d[l], u[l]) * 1.0;
  }
  if constexpr (!MIXED)
    x64::floor_state(P, u);
  else
    floor_mixed(P, u);
}

template <bool MIXED, int D>
__device__ __forceinline__ void face_of_e(const PhysParams& P, const typename Prec<MIXED>::T* pl,
                                          const typename Prec<MIXED>::T* sl, const typename Prec<MIXED>::T* pr,
                                          const typename Prec<MIXED>::T* sr, typename Prec<MIXED>::T* F,
                                          double& eig) {
  if constexpr (!MIXED) {
    x64::kt_face<D>(P, pl, sl, pr, sr, F, eig);
  } else {
    const f32::PhysF PF{(float)P.gamma, (float)(P.gamma - 1.0), (float)P.theta};
    float e;
    f32::kt_face<D>(PF, pl, sl, pr, sr, F, e);
    eig = (double)e;
  }
}

// MODE 0: substep (flux + divergence + update).  MODE 1: max wavespeed scan only.
// MODE 2: the first substep of an RK step with dt deferred: the fluxes, the block's max wavespeed
//   (the KT face speeds are exactly Grid::max_dt's, grid/grid.cpp:422-483, since both reconstruct
//   the same state) and the divergence D, stored instead of the update.
// MODE 3: a substep (beta != 1) whose input is that first stage, formed on the fly from the step's
//   base U^n and D once dt is known; U^n is also its RK base.
// MODEs 2 and 3 run on self-ghosting outflow blocks only; together they replace the CFL scan,
// the dt, the store and the first two substeps of an RK step, and the scan's separate pass over
// the state disappears.
// Per plane k: stage the plane's primitives (with x/y ghosts) in LDS; x- and y-slopes of every
// cell once; the x/y faces from two cells and their slopes; the z-face from the column window;
// then each thread updates its cell.  Three barriers per plane.
// SELF: the blocks are isolated outflow domains (a.outflow / m.outflow), self-ghosting: each ghost
// is formed from the boundary cell it copies (grid/grid_phys_bound.cpp:27-54) where it is used, so
// their ghost slabs are neither read nor written.  A column's z ghosts come from its own end cells;
// the x/y ghosts of plane k from the boundary columns' input state of that plane, which their
// owners leave in an LDS ring (three planes deep) as they fetch it, three planes ahead.
// Occupancy, measured on the flat 4096 x 16^3 batch: the FP64 substep runs two waves per SIMD and
// reads its cell's conserved state back from memory (L2) at each plane instead of carrying three
// planes of FP64 column state in registers, which removes its spills (1.88 -> 1.54 ms; on the
// 8^3 multilevel batch 4.17 -> 3.39 ms); the mixed form keeps the column state, and runs three
// waves per SIMD on 16^3 blocks (168 registers) but two on 8^3 blocks (three are no faster there).
// Every record access goes through RecBuf (one 32-bit offset per cell, not a 64-bit pointer per
// field) and libm pow stays out of line; together they took the mixed 16^3 substep from 50 spilled
// registers to 4 (0.763 -> 0.711 ms).
// TREE: the octree's reflux overrides and the root's flow-off ledger (a.slot, a.ledger); the flat
// instantiation carries neither, which keeps its register count down.
// WPE: the waves per SIMD the register allocation is held to (the register budget: 2 -> 256 VGPRs,
// 3 -> 168, 4 -> 128); 0 takes the measured default above.  The flat substep is instantiated for
// 2, 3 and 4 as well, and the autotuner picks among them with the grid.
template <int N, bool MIXED, int MODE, bool TREE, bool SELF, int WPE = 0>
__global__ __launch_bounds__(N* N) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : (MIXED ? (N == 16 ? 3 : 1) : 2)))) void kt_sweep_kernel(
    SubstepArgs a, MaxEigArgs m) {
  using T = typename Prec<MIXED>::T;
  constexpr bool FLUX = MODE != 1;                 // face fluxes (MODE 1 forms wavespeeds only)
  constexpr bool EIG = MODE == 1 || MODE == 2;     // the block's max wavespeed
  constexpr bool UPDATE = MODE == 0 || MODE == 3;  // RK combination, floor, store
  constexpr bool FROMD = MODE == 3;                // input: the first stage from U^n and D
  static_assert(!TREE || MODE <= 1, "the octree instantiations are the substep and the scan");
  static_assert(!(TREE && SELF) && (MODE < 2 || SELF), "deferred steps run on self-ghosting flat batches");
  constexpr int PW = N + 2 * GW;  // padded plane edge
  constexpr int NT = N * N;
  constexpr int NFX = N * (N + 1);  // x faces of a plane (= y faces)
  constexpr int NSX = N * (N + 2);  // x slopes of a plane (= y slopes)
  constexpr int HN = GW * N;        // halo cells per face per plane
  constexpr Layout L{N, GW, NFE};
  constexpr long R = L.rec();
  constexpr long N3 = (long)N * N * N;

  __shared__ T plane[NFE][PW][PW];
  __shared__ T Sx[NFE][N][N + 2];  // x-slope of plane cell (padded x = idx+1, y = row+G)
  __shared__ T Sy[NFE][N + 2][N];  // y-slope of plane cell (padded y = idx+1, x = col+G)
  __shared__ T Fx[FLUX ? NFE : 1][N][N + 1];
  __shared__ T Fy[FLUX ? NFE : 1][N + 1][N];
  __shared__ double red[NT / 64 > 0 ? NT / 64 : 1];
  // PD: how many planes ahead a column cell is loaded.  The mixed form loads cell k+3 at the top of
  // plane k and converts it after the update (the most latency hidden; measured faster); the FP64
  // form loads cell k+2 and converts it after the x/y faces, which carries one plane of FP64 state
  // fewer (its kernels were register-bound) and shortens its L2 re-reads.  The mixed MODE 3 kernel
  // also looks two planes ahead: it reads each cell's U^n again at its plane as the RK base, and the
  // shorter distance keeps more of those re-reads in the L2 (0.60 -> 0.58 ms on 4096 x 16^3); its
  // loads of D, which it reads once, take the non-temporal policy so that they leave the L2 to U^n
  // (the FP64 form reads D again at the cell's plane and keeps the default policy).
  constexpr int PD = MIXED && MODE != 3 ? 3 : 2;
  // SELF: the boundary cells of planes k .. k+PD-1 (ring slot = plane % PD; plane k+PD's is written
  // over plane k's once the halo of k is staged) by position along the perimeter: their FP32
  // primitives in the mixed form, their FP64 input state in the FP64 form (whose ghosts are formed
  // as the reference forms them, from the conserved state)
  constexpr bool PRIMGHOST = SELF && MIXED;
  __shared__ typename std::conditional<MIXED, float, double>::type gst[SELF ? PD : 1][SELF ? 4 * N - 4 : 1][NFE];

  // persistent: a grid smaller than the batch walks it with a stride (launch shape chosen by the
  // autotuner); with one workgroup per block the loop runs once
  const int nbl = MODE == 1 ? m.nblocks : a.nblocks;
  for (int bb = blockIdx.x; bb < nbl; bb += gridDim.x) {
  const int b = (MODE == 1 ? 0 : a.block0) + bb;
  const int tid = threadIdx.x;
  const int i = tid % N, j = tid / N;
  const PhysParams P = MODE == 1 ? m.P : a.P;
  const double* Ub = (MODE == 1 ? m.U : a.Uin) + (long)b * NFE * R;
  double dt = a.dt;
  const double beta = a.beta;
  if (FROMD || (UPDATE && a.dt_dev)) dt = *a.dt_dev;
  const bool has_u0 = UPDATE && beta != 1.0;
  const bool zr = (MODE == 1 ? m.z_reflect : a.z_reflect) != 0;
  double* const Uob = UPDATE ? a.Uout + (long)b * NFE * R : nullptr;
  const RecBuf<R, NFE> ub(Ub), uo(Uob), u0(UPDATE && a.U0 ? a.U0 + (long)b * NFE * R : nullptr);
  // the deferred divergence, interior cells only (MODE 2 writes it, MODE 3 reads it)
  const RecBuf<N3, NFE> db(MODE >= 2 ? a.dbuf + (long)b * NFE * N3 : nullptr);
  const double hinv = MODE == 1 ? 0.0 : 1.0 / a.h[b];

  // ---- prologue: column cells k = -2 .. PD-1 ----------------------------------------------
  // the conserved FP64 input state of cells k .. k+2 (the mixed update carries it; FP64 reloads
  // cell k)
  constexpr bool CARRY = MIXED && UPDATE;
  double ucur[UPDATE ? NFE : 1], uc1[CARRY ? NFE : 1], uc2[CARRY ? NFE : 1];
  T c0[NFE], cp1[NFE], cp2[NFE], s0[NFE], Fzlo[NFE];
  double emax = 0.0;
  // an interior cell kk (input state u, primitives p) into the ring, where its column is a
  // boundary column
  auto stash = [&](int kk, const double* u, const T* p) {
    if constexpr (SELF) {
      if (i == 0 || i == N - 1 || j == 0 || j == N - 1) {
        const int q = perim<N>(i, j);
#pragma unroll
        for (int l = 0; l < NFE; l++) {
          if constexpr (MIXED)
            gst[kk % PD][q][l] = p[l];
          else
            gst[kk % PD][q][l] = u[l];
        }
      }
    }
  };
  // the mixed form's z ghost kk (< 0 or >= N) from the primitives of the column's end cell it
  // copies (zr: ghost -1-m mirrors cell m)
  auto zghost = [&](int kk, const T* src, T* out) {
#pragma unroll
    for (int l = 0; l < NFE; l++) out[l] = src[l];
    if constexpr (MIXED) ghost_post_prim(out, kk < 0 ? (zr ? kMirror : (int)ZL) : (int)ZU);
  };
  auto fetch_col = [&](int kk, double* u) {
    int post;
    const unsigned e = col_src<N, SELF>(L, i, j, kk, zr, post);
#pragma unroll
    for (int l = 0; l < NFE; l++) u[l] = ub.ld(l, e);
    if constexpr (FROMD) {
      double d[NFE];
#pragma unroll
      for (int l = 0; l < NFE; l++) d[l] = CARRY ? db.ld_nt(l, e) : db.ld(l, e);  // e: an interior cell
      first_stage<MIXED>(P, u, d, dt, hinv);
    }
    ghost_post(u, post);
  };
  {
    double u[NFE];
    T cm2[NFE], cm1[NFE], sm1[NFE];
    if constexpr (!PRIMGHOST) {
      fetch_col(-2, u);
      make_prim<MIXED>(u, cm2);
      fetch_col(-1, u);
      make_prim<MIXED>(u, cm1);
    }
    fetch_col(0, u);
    if constexpr (UPDATE)
#pragma unroll
      for (int l = 0; l < NFE; l++) ucur[l] = u[l];
    make_prim<MIXED>(u, c0);
    stash(0, u, c0);
    fetch_col(1, u);
    if constexpr (CARRY)
#pragma unroll
      for (int l = 0; l < NFE; l++) uc1[l] = u[l];
    make_prim<MIXED>(u, cp1);
    stash(1, u, cp1);
    if constexpr (PD == 3) {
      fetch_col(2, u);
      if constexpr (CARRY)
#pragma unroll
        for (int l = 0; l < NFE; l++) uc2[l] = u[l];
      make_prim<MIXED>(u, cp2);
      stash(2, u, cp2);
    }
    if constexpr (PRIMGHOST) {
      zghost(-1, c0, cm1);
      zghost(-2, zr ? cp1 : c0, cm2);
    }
    slope_of<MIXED>(P, cm2, cm1, c0, sm1);
    slope_of<MIXED>(P, cm1, c0, cp1, s0);
    if constexpr (FLUX) {
      double e;
      face_of_e<MIXED, 2>(P, cm1, sm1, c0, s0, Fzlo, e);  // face k = 0
      if constexpr (EIG) emax = x64::rmax(emax, e);
      if (TREE && a.slot) apply_override<N>(a, b, 2, 0, i, j, Fzlo);
      if (TREE && b == a.ledger_block) {
#pragma unroll
        for (int l = 0; l < NFE; l++) a.ledger[(((2 * 2 + 0) * N + j) * N + i) * NFE + l] = (double)Fzlo[l];
      }
    } else {
      emax = x64::rmax(emax, eig_of<MIXED, 2>(P, cm1, sm1, c0, s0));
    }
  }

  // the x/y halo: thread tid < 4*G*N stages ghost (face, r) of each plane -- from the stored
  // ghost slab (its raw value prefetched a plane ahead) or, SELF, from the ring
  double hraw[SELF ? 1 : NFE];
  const bool is_halo = tid < 4 * HN;
  int hpx = 0, hpy = 0, hface = 0, hpos = 0;
  unsigned hbase = 0;
  if (is_halo) {
    const int face = tid / HN, r = tid % HN;
    hface = face;
    if (face < 2) {
      const int jj = r / GW, t = r % GW;
      hpx = face == XL ? t : N + GW + t;
      hpy = jj + GW;
      hpos = perim<N>(face == XL ? 0 : N - 1, jj);
    } else {
      const int t = r / N, ii = r % N;
      hpx = ii + GW;
      hpy = face == YL ? t : N + GW + t;
      hpos = perim<N>(ii, face == YL ? 0 : N - 1);
    }
    if constexpr (!SELF) {
      hbase = (unsigned)(L.slab(face) + r);  // plane k adds k*HN (both slab kinds are plane-major)
#pragma unroll
      for (int l = 0; l < NFE; l++) hraw[l] = ub.ld_nt(l, hbase);
    }
  }
  if constexpr (SELF) __syncthreads();  // the ring's first three planes

  for (int k = 0; k < N; k++) {
    // A. stage plane k (interior from registers, halo from the prefetched raw values or the ring)
#pragma unroll
    for (int l = 0; l < NFE; l++) plane[l][j + GW][i + GW] = c0[l];
    if (is_halo) {
      T hp[NFE];
      if constexpr (PRIMGHOST) {
#pragma unroll
        for (int l = 0; l < NFE; l++) hp[l] = gst[k % PD][hpos][l];
        ghost_post_prim(hp, hface);
      } else if constexpr (SELF) {
        double g[NFE];
#pragma unroll
        for (int l = 0; l < NFE; l++) g[l] = gst[k % PD][hpos][l];
        ghost_post(g, hface);
        make_prim<MIXED>(g, hp);
      } else {
        make_prim<MIXED>(hraw, hp);
      }
#pragma unroll
      for (int l = 0; l < NFE; l++) plane[l][hpy][hpx] = hp[l];
    }
    // prefetch: halo of plane k+1 (stored slabs), column cell k+PD, U0 at k
    if constexpr (!SELF) {
      if (is_halo && k + 1 < N) {
#pragma unroll
        for (int l = 0; l < NFE; l++) hraw[l] = ub.ld_nt(l, hbase + (unsigned)((k + 1) * HN));
      }
    }
    double craw[NFE], cd[FROMD ? NFE : 1];
    int cpost = -1;
    // the mixed self-ghosting form forms the z ghosts N, N+1 from cell N-1's primitives
    const int kn = k + PD;  // the column cell loaded
    const bool have_next = PRIMGHOST ? kn < N : kn < N + GW;
    if (have_next) {
      const unsigned e = col_src<N, SELF>(L, i, j, kn, zr, cpost);
#pragma unroll
      for (int l = 0; l < NFE; l++) craw[l] = ub.ld(l, e);
      if constexpr (FROMD)
#pragma unroll
        for (int l = 0; l < NFE; l++) cd[l] = CARRY ? db.ld_nt(l, e) : db.ld(l, e);
    }
    // cell k's RK base (U0; for MODE 3 the step's base U^n, the input it was formed from) and,
    // where the column state is not carried, cell k's input state again (read PD planes ago: L2)
    double u0v[NFE], dk[(FROMD && !CARRY) ? NFE : 1];
    if constexpr (UPDATE) {
      const unsigned e = (unsigned)L.interior(i, j, k);
      if constexpr (!CARRY && !FROMD) {
        if (k > 0)
#pragma unroll
          for (int l = 0; l < NFE; l++) ucur[l] = ub.ld(l, e);
      }
      if constexpr (!CARRY && FROMD) {
        if (k > 0)
#pragma unroll
          for (int l = 0; l < NFE; l++) dk[l] = db.ld_nt(l, e);
      }
      if (has_u0)
#pragma unroll
        for (int l = 0; l < NFE; l++) u0v[l] = u0.ld_nt(l, e);
    }
    __syncthreads();

    // B. x- and y-slopes of every plane cell that a face uses (reconstruct.cpp:81-85)
#pragma unroll 1
    for (int q = tid; q < 2 * NSX; q += NT) {
      T pa[NFE], pb[NFE], pc[NFE], sl[NFE];
      if (q < NSX) {
        const int jj = q / (N + 2), sx = q % (N + 2);
        const int py = jj + GW;
#pragma unroll
        for (int l = 0; l < NFE; l++) {
          pa[l] = plane[l][py][sx];
          pb[l] = plane[l][py][sx + 1];
          pc[l] = plane[l][py][sx + 2];
        }
        slope_of<MIXED>(P, pa, pb, pc, sl);
#pragma unroll
        for (int l = 0; l < NFE; l++) Sx[l][jj][sx] = sl[l];
      } else {
        const int q2 = q - NSX;
        const int ii = q2 % N, sy = q2 / N;
        const int px = ii + GW;
#pragma unroll
        for (int l = 0; l < NFE; l++) {
          pa[l] = plane[l][sy][px];
          pb[l] = plane[l][sy + 1][px];
          pc[l] = plane[l][sy + 2][px];
        }
        slope_of<MIXED>(P, pa, pb, pc, sl);
#pragma unroll
        for (int l = 0; l < NFE; l++) Sy[l][sy][ii] = sl[l];
      }
    }
    __syncthreads();

    // C. x- and y-faces of plane k, then the z-face k+1 of the column
#pragma unroll 1
    for (int it = tid; it < 2 * NFX; it += NT) {
      T pl[NFE], pr[NFE], sl[NFE], sr[NFE];
      if (it < NFX) {
        const int fi = it % (N + 1), jj = it / (N + 1);
        const int py = jj + GW;
#pragma unroll
        for (int l = 0; l < NFE; l++) {
          pl[l] = plane[l][py][fi + 1];
          pr[l] = plane[l][py][fi + 2];
          sl[l] = Sx[l][jj][fi];
          sr[l] = Sx[l][jj][fi + 1];
        }
        if constexpr (FLUX) {
          T F[NFE];
          double e;
          face_of_e<MIXED, 0>(P, pl, sl, pr, sr, F, e);
          if constexpr (EIG) emax = x64::rmax(emax, e);
          if (TREE && a.slot) apply_override<N>(a, b, 0, fi, jj, k, F);
#pragma unroll
          for (int l = 0; l < NFE; l++) Fx[l][jj][fi] = F[l];
        } else {
          emax = x64::rmax(emax, eig_of<MIXED, 0>(P, pl, sl, pr, sr));
        }
      } else {
        const int it2 = it - NFX;
        const int ii = it2 % N, fj = it2 / N;
        const int px = ii + GW;
#pragma unroll
        for (int l = 0; l < NFE; l++) {
          pl[l] = plane[l][fj + 1][px];
          pr[l] = plane[l][fj + 2][px];
          sl[l] = Sy[l][fj][ii];
          sr[l] = Sy[l][fj + 1][ii];
        }
        if constexpr (FLUX) {
          T F[NFE];
          double e;
          face_of_e<MIXED, 1>(P, pl, sl, pr, sr, F, e);
          if constexpr (EIG) emax = x64::rmax(emax, e);
          if (TREE && a.slot) apply_override<N>(a, b, 1, fj, ii, k, F);
#pragma unroll
          for (int l = 0; l < NFE; l++) Fy[l][fj][ii] = F[l];
        } else {
          emax = x64::rmax(emax, eig_of<MIXED, 1>(P, pl, sl, pr, sr));
        }
      }
    }
    // the column cell loaded at the top of the plane becomes cell k+2 here (PD = 2) or k+3 after
    // the update (PD = 3); its state goes into the ring slot of plane k, whose halo every thread
    // staged before the first barrier
    auto take_next = [&]() {
      if (have_next) {
        if constexpr (FROMD) first_stage<MIXED>(P, craw, cd, dt, hinv);
        ghost_post(craw, cpost);
        if constexpr (CARRY)
#pragma unroll
          for (int l = 0; l < NFE; l++) uc2[l] = craw[l];
        make_prim<MIXED>(craw, cp2);
        if (kn < N) stash(kn, craw, cp2);
      } else if constexpr (PRIMGHOST) {
        if (kn < N + GW) zghost(kn, kn == N ? cp1 : c0, cp2);  // from cell N-1
      }
    };
    if constexpr (PD == 2) take_next();
    T sp1[NFE], Fzhi[NFE];
    slope_of<MIXED>(P, c0, cp1, cp2, sp1);
    if constexpr (FLUX) {
      double e;
      face_of_e<MIXED, 2>(P, c0, s0, cp1, sp1, Fzhi, e);
      if constexpr (EIG) emax = x64::rmax(emax, e);
      if (TREE && a.slot) apply_override<N>(a, b, 2, k + 1, i, j, Fzhi);
    } else {
      emax = x64::rmax(emax, eig_of<MIXED, 2>(P, c0, s0, cp1, sp1));
    }
    __syncthreads();

    // D. divergence (x, then y, then z), then the RK update of cell (i, j, k) -- or, deferred
    // (MODE 2), the divergence itself
    if constexpr (MODE == 2) {
      const unsigned e = (unsigned)L.interior(i, j, k);
#pragma unroll
      for (int l = 0; l < NFE; l++) {
        double D;
        if constexpr (!MIXED) {
          D = 0.0 - (Fx[l][j][i + 1] - Fx[l][j][i]) * hinv;
          D -= (Fy[l][j + 1][i] - Fy[l][j][i]) * hinv;
          D -= (Fzhi[l] - Fzlo[l]) * hinv;
          D += 0.0;  // State::source is zero for the ideal gas (euler/state.cpp:19-22)
        } else {
          D = (((double)Fx[l][j][i + 1] - (double)Fx[l][j][i]) + ((double)Fy[l][j + 1][i] - (double)Fy[l][j][i])) +
              ((double)Fzhi[l] - (double)Fzlo[l]);
        }
        db.st_nt(l, e, D);
      }
    }
    if constexpr (UPDATE) {
      if constexpr (FROMD && !CARRY) {
        // cell k's first stage again, from U^n (= U0) and D
        if (k > 0) {
#pragma unroll
          for (int l = 0; l < NFE; l++) ucur[l] = u0v[l];
          first_stage<MIXED>(P, ucur, dk, dt, hinv);
        }
      }
      double u[NFE];
#pragma unroll
      for (int l = 0; l < NFE; l++) {
        double D;
        if constexpr (!MIXED) {
          D = 0.0 - (Fx[l][j][i + 1] - Fx[l][j][i]) * hinv;
          D -= (Fy[l][j + 1][i] - Fy[l][j][i]) * hinv;
          D -= (Fzhi[l] - Fzlo[l]) * hinv;
          D += 0.0;  // State::source is zero for the ideal gas (euler/state.cpp:19-22)
          if (has_u0)
            u[l] = (ucur[l] + D * dt) * beta + u0v[l] * (1.0 - beta);
          else
            u[l] = (ucur[l] + D * dt) * beta;
        } else {
          // FP32 fluxes, differenced in FP64: the difference of two FP32 numbers is exact in FP64,
          // so the divergence still telescopes and mass and energy are conserved to roundoff
          const double dsum = (((double)Fx[l][j][i + 1] - (double)Fx[l][j][i]) +
                               ((double)Fy[l][j + 1][i] - (double)Fy[l][j][i])) +
                              ((double)Fzhi[l] - (double)Fzlo[l]);
          const double un = fma(-dt * hinv, dsum, ucur[l]);
          u[l] = has_u0 ? fma(beta, un, u0v[l] * (1.0 - beta)) : un * beta;
        }
      }
      if constexpr (!MIXED)
        x64::floor_state(P, u);
      else
        floor_mixed(P, u);
      if (!(u[RHO] > 0.0) && a.err) atomicOr(a.err, 1);
      if (TREE && b == a.ledger_block) {
        // boundary faces of the ledger block (grid/grid.cpp:350-404 sums them per pencil)
        double* lg = a.ledger;
        if (i == 0) {
#pragma unroll
          for (int l = 0; l < NFE; l++) {
            lg[(((0 * 2 + 0) * N + k) * N + j) * NFE + l] = (double)Fx[l][j][0];
            lg[(((0 * 2 + 1) * N + k) * N + j) * NFE + l] = (double)Fx[l][j][N];
          }
        }
        if (j == 0) {
#pragma unroll
          for (int l = 0; l < NFE; l++) {
            lg[(((1 * 2 + 0) * N + k) * N + i) * NFE + l] = (double)Fy[l][0][i];
            lg[(((1 * 2 + 1) * N + k) * N + i) * NFE + l] = (double)Fy[l][N][i];
          }
        }
        if (k == N - 1) {
#pragma unroll
          for (int l = 0; l < NFE; l++) lg[(((2 * 2 + 1) * N + j) * N + i) * NFE + l] = (double)Fzhi[l];
        }
      }
      const long e = L.interior(i, j, k);
#pragma unroll
      for (int l = 0; l < NFE; l++) uo.st_nt(l, (unsigned)e, u[l]);
    }

    // advance the column window
#pragma unroll
    for (int l = 0; l < NFE; l++) {
      c0[l] = cp1[l];
      cp1[l] = cp2[l];
      s0[l] = sp1[l];
      Fzlo[l] = Fzhi[l];
    }
    if constexpr (CARRY) {
#pragma unroll
      for (int l = 0; l < NFE; l++) {
        ucur[l] = uc1[l];
        uc1[l] = uc2[l];
      }
    }
    if constexpr (PD == 3) take_next();
  }

  if constexpr (EIG) {
    // block max: wavefront reduction, then across waves through LDS; the global max of eig/h as
    // the ordered bits of a non-negative double (order independent, so deterministic)
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
  }  // block loop
}

// Ghost fill of isolated blocks: every face slab cell copies the adjacent interior cell and
// removes inflow momentum (grid/grid_phys_bound.cpp:27-54, euler/state.cpp:45-84); the lower z face
// mirrors with sz -> -sz when z_reflect (grid/grid_phys_bound.cpp:298-313, euler/state.cpp:94-96).
template <int N, int GW>
__global__ __launch_bounds__(256) void outflow_ghost_kernel(double* U, int nblocks, int z_reflect) {
  constexpr Layout L{N, GW, NFE};
  constexpr long R = L.rec();
  constexpr int SC = GW * N * N;  // cells per slab
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)nblocks * 6 * SC;
  if (gid >= total) return;
  const int b = (int)(gid / (6 * SC));
  const int rem = (int)(gid % (6 * SC));
  const int face = rem / SC, r = rem % SC;
  int ii, jj, kk, t;  // interior coordinates of the face cell and the ghost depth
  if (face < 2) {
    t = r % GW;
    jj = (r / GW) % N;
    kk = r / (GW * N);
    ii = face == XL ? 0 : N - 1;
  } else if (face < 4) {
    ii = r % N;
    t = (r / N) % GW;
    kk = r / (GW * N);
    jj = face == YL ? 0 : N - 1;
  } else {
    ii = r % N;
    jj = (r / N) % N;
    t = r / (N * N);
    kk = face == ZL ? 0 : N - 1;
  }
  const bool reflect = z_reflect && face == ZL;
  if (reflect) kk = GW - 1 - t;  // padded source 2G-1-t
  double* Ub = U + (long)b * NFE * R;
  const long src = L.interior(ii, jj, kk);
  double u[NFE];
#pragma unroll
  for (int l = 0; l < NFE; l++) u[l] = Ub[l * R + src];
  if (reflect) {
    u[SZ] = -u[SZ];
  } else {
    // the boundary object's face is the node face inverted: lower node faces test v > 0
    const int d = face >> 1;
    const bool node_upper = face & 1;
    const double v = (d == 2) ? u[SZ] : u[SX + d] / u[RHO];
    const bool inflow = node_upper ? (v < 0.0) : (v > 0.0);
    if (inflow) {
      const double s = u[SX + d];
      u[ET] = u[ET] - 0.5 * s * s / u[RHO];
      u[SX + d] = 0.0;
    }
  }
  const long dst = L.slab(face) + r;
#pragma unroll
  for (int l = 0; l < NFE; l++) Ub[l * R + dst] = u[l];
}

// padded SoA <-> records (one thread per record element)
template <int N, int GW>
__global__ __launch_bounds__(256) void pack_kernel(const double* __restrict__ pad, double* __restrict__ rec,
                                                  long nrec_elems, int to_rec) {
  constexpr Layout L{N, GW, NFE};
  constexpr long R = L.rec();
  constexpr int PW = N + 2 * GW;
  constexpr long P3 = (long)PW * PW * PW;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nrec_elems) return;
  const long bf = gid / R;  // block*NF + field
  const int e = (int)(gid % R);
  int pi, pj, pk;
  if (e < N * N * N) {
    pi = e % N + GW;
    pj = (e / N) % N + GW;
    pk = e / (N * N) + GW;
  } else {
    const int s = (e - N * N * N) / (GW * N * N);
    const int r = (e - N * N * N) % (GW * N * N);
    const int lo = (s & 1) ? N + GW : 0;
    if (s < 2) {
      pi = lo + r % GW;
      pj = (r / GW) % N + GW;
      pk = r / (GW * N) + GW;
    } else if (s < 4) {
      pi = r % N + GW;
      pj = lo + (r / N) % GW;
      pk = r / (GW * N) + GW;
    } else {
      pi = r % N + GW;
      pj = (r / N) % N + GW;
      pk = lo + r / (N * N);
    }
  }
  const long pidx = bf * P3 + ((long)pk * PW + pj) * PW + pi;
  if (to_rec)
    rec[gid] = pad[pidx];
  else
    const_cast<double*>(pad)[pidx] = rec[gid];
}

// dt = scale * cfl / max_b(eig_b/h_b), optionally capped at growth * (previous dt)
// (euler/hydro_driver.cpp:107-113); resets the reduction word so the next scan starts at zero
// without a memset (keeps the step capturable in a graph).
__global__ void dt_from_gmax_kernel(unsigned long long* gmax, double cfl, double scale, double growth,
                                    double* dt) {
  const double m = __longlong_as_double((long long)*gmax);
  double dtinv = m;
  if (dtinv == 0.0) dtinv = 1.0e-10;
  double d = (1.0 / dtinv) * cfl * scale;
  if (growth > 0.0) {
    const double cap = growth * dt[0];
    d = cap < d ? cap : d;
  }
  dt[0] = d;
  *gmax = 0ull;
}

__global__ void reset_u64_kernel(unsigned long long* p) { *p = 0ull; }

}  // namespace

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
template <int N, bool MIXED, int MODE, bool SELF>
static void launch_flat_variant(const SubstepArgs& a, int g, int variant, hipStream_t s) {
  MaxEigArgs m{};
  switch (variant) {
    case 1: kt_sweep_kernel<N, MIXED, MODE, false, SELF, 2><<<g, N * N, 0, s>>>(a, m); break;
    case 2: kt_sweep_kernel<N, MIXED, MODE, false, SELF, 3><<<g, N * N, 0, s>>>(a, m); break;
    case 3: kt_sweep_kernel<N, MIXED, MODE, false, SELF, 4><<<g, N * N, 0, s>>>(a, m); break;
    default: kt_sweep_kernel<N, MIXED, MODE, false, SELF><<<g, N * N, 0, s>>>(a, m); break;
  }
}

template <int N, bool MIXED>
static void launch_flat_mode(const SubstepArgs& a, int g, int variant, int mode, hipStream_t s) {
  if (mode == 2)
    launch_flat_variant<N, MIXED, 2, true>(a, g, variant, s);
  else if (mode == 3)
    launch_flat_variant<N, MIXED, 3, true>(a, g, variant, s);
  else if (a.outflow)
    launch_flat_variant<N, MIXED, 0, true>(a, g, variant, s);
  else
    launch_flat_variant<N, MIXED, 0, false>(a, g, variant, s);
}

template <int N>
static hipError_t launch_substep_n(const SubstepArgs& a, bool mixed, int grid, int variant, int mode, hipStream_t s) {
  MaxEigArgs m{};
  if (a.nblocks <= 0) return hipSuccess;
  const int g = (grid > 0 && grid < a.nblocks) ? grid : a.nblocks;
  const bool tree = a.slot != nullptr || a.ledger != nullptr;
  if ((tree && a.outflow) ||
      (mode != 0 && (tree || !a.outflow || !a.dbuf || (mode == 2 && !a.gmax) ||
                     (mode == 3 && (!a.dt_dev || a.beta == 1.0 || a.U0 != a.Uin)))))
    return hipErrorInvalidValue;
  if (mixed && tree)
    kt_sweep_kernel<N, true, 0, true, false><<<g, N * N, 0, s>>>(a, m);
  else if (mixed)
    launch_flat_mode<N, true>(a, g, variant, mode, s);
  else if (tree)
    kt_sweep_kernel<N, false, 0, true, false><<<g, N * N, 0, s>>>(a, m);
  else
    launch_flat_mode<N, false>(a, g, variant, mode, s);
  return hipGetLastError();
}

hipError_t launch_substep(int n, const SubstepArgs& a, bool mixed, int grid, hipStream_t s, int variant, int mode) {
  if (variant < 0 || variant >= kSweepVariants) return hipErrorInvalidValue;
  if (mode != 0 && mode != 2 && mode != 3) return hipErrorInvalidValue;
  if (n == 8) return launch_substep_n<8>(a, mixed, grid, variant, mode, s);
  if (n == 16) return launch_substep_n<16>(a, mixed, grid, variant, mode, s);
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t launch_maxeig_n(const MaxEigArgs& m, bool mixed, int grid, hipStream_t s) {
  SubstepArgs a{};
  if (m.nblocks <= 0) return hipSuccess;
  const int g = (grid > 0 && grid < m.nblocks) ? grid : m.nblocks;
  if (mixed && m.outflow)
    kt_sweep_kernel<N, true, 1, false, true><<<g, N * N, 0, s>>>(a, m);
  else if (mixed)
    kt_sweep_kernel<N, true, 1, false, false><<<g, N * N, 0, s>>>(a, m);
  else if (m.outflow)
    kt_sweep_kernel<N, false, 1, false, true><<<g, N * N, 0, s>>>(a, m);
  else
    kt_sweep_kernel<N, false, 1, false, false><<<g, N * N, 0, s>>>(a, m);
  return hipGetLastError();
}

hipError_t launch_maxeig(int n, const MaxEigArgs& m, bool mixed, int grid, hipStream_t s) {
  if (n == 8) return launch_maxeig_n<8>(m, mixed, grid, s);
  if (n == 16) return launch_maxeig_n<16>(m, mixed, grid, s);
  return hipErrorInvalidValue;
}

hipError_t launch_outflow_ghosts(int n, int g, double* U, int nblocks, int z_reflect, hipStream_t s) {
  const long total = (long)nblocks * 6 * g * n * n;
  const unsigned grid = (unsigned)((total + 255) / 256);
  if (total == 0) return hipSuccess;
  if (n == 8 && g == 2)
    outflow_ghost_kernel<8, 2><<<grid, 256, 0, s>>>(U, nblocks, z_reflect);
  else if (n == 8 && g == 3)
    outflow_ghost_kernel<8, 3><<<grid, 256, 0, s>>>(U, nblocks, z_reflect);
  else if (n == 16 && g == 2)
    outflow_ghost_kernel<16, 2><<<grid, 256, 0, s>>>(U, nblocks, z_reflect);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_pack(int n, int g, const double* pad, double* rec, long nrec_elems, bool to_rec, hipStream_t s) {
  const unsigned grid = (unsigned)((nrec_elems + 255) / 256);
  if (nrec_elems == 0) return hipSuccess;
  if (n == 8 && g == 2)
    pack_kernel<8, 2><<<grid, 256, 0, s>>>(pad, rec, nrec_elems, to_rec ? 1 : 0);
  else if (n == 8 && g == 3)
    pack_kernel<8, 3><<<grid, 256, 0, s>>>(pad, rec, nrec_elems, to_rec ? 1 : 0);
  else if (n == 16 && g == 2)
    pack_kernel<16, 2><<<grid, 256, 0, s>>>(pad, rec, nrec_elems, to_rec ? 1 : 0);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_dt_from_gmax(unsigned long long* gmax, double cfl, double scale, double growth, double* dt,
                               hipStream_t s) {
  dt_from_gmax_kernel<<<1, 1, 0, s>>>(gmax, cfl, scale, growth, dt);
  return hipGetLastError();
}

hipError_t launch_reset_u64(unsigned long long* p, hipStream_t s) {
  reset_u64_kernel<<<1, 1, 0, s>>>(p);
  return hipGetLastError();
}

}  // namespace amrh
