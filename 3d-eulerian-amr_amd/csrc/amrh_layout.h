// Device-side layout of a batch of blocks ("halo-slab records").
//
// The reference keeps, per node, a padded GNX^3 array of AoS States (grid/grid.h:26-32).  Only
// the interior and the six face ghost slabs are ever read (grid/grid_node.cpp:59-112; edge and
// corner ghosts only under RECON3D, which is never defined), so a record here stores exactly
// those cells and nothing else.  Per block, per field, R = N^3 + 6*G*N^2 doubles:
//
//   [0, N^3)                 interior, (k*N + j)*N + i                    (i, j, k in [0, N))
//   slab s = XL,XU (0,1)     N^3 + s*G*N^2 + (k*N + j)*G + t             padded x = t | N+G+t
//   slab s = YL,YU (2,3)     N^3 + s*G*N^2 + (k*G + t)*N + i             padded y = t | N+G+t
//   slab s = ZL,ZU (4,5)     N^3 + s*G*N^2 + (t*N + j)*N + i             padded z = t | N+G+t
//
// Records are block-major: U[(b*NF + f)*R + e].  For N = 8, G = 2 a field record is 10,240 B
// (80 lines of 128 B); for N = 16, 57,344 B (448 lines): every field record starts 128-B aligned
// and a block reads exactly its algorithmic bytes.
#pragma once
#include <cstddef>

namespace amrh {

enum Face { XL = 0, XU = 1, YL = 2, YU = 3, ZL = 4, ZU = 5 };

struct Layout {
  int N, G, NF;
  __host__ __device__ constexpr long rec() const { return (long)N * N * N + 6L * G * N * N; }
  __host__ __device__ constexpr long block() const { return rec() * NF; }
  __host__ __device__ constexpr long slab(int s) const { return (long)N * N * N + (long)s * G * N * N; }
  __host__ __device__ constexpr long interior(int i, int j, int k) const { return ((long)k * N + j) * N + i; }
  // padded-cube index (p in [0, N+2G)^3) -> record element, or -1 for edge/corner cells
  __host__ __device__ long from_padded(int pi, int pj, int pk) const {
    const int lo = G, hi = N + G;
    const bool ix = pi >= lo && pi < hi, iy = pj >= lo && pj < hi, iz = pk >= lo && pk < hi;
    const int nin = (int)ix + (int)iy + (int)iz;
    if (nin == 3) return interior(pi - G, pj - G, pk - G);
    if (nin != 2) return -1;
    if (!ix) {
      const int s = pi < lo ? XL : XU, t = pi < lo ? pi : pi - hi;
      return slab(s) + ((long)(pk - G) * N + (pj - G)) * G + t;
    }
    if (!iy) {
      const int s = pj < lo ? YL : YU, t = pj < lo ? pj : pj - hi;
      return slab(s) + ((long)(pk - G) * G + t) * N + (pi - G);
    }
    const int s = pk < lo ? ZL : ZU, t = pk < lo ? pk : pk - hi;
    return slab(s) + ((long)t * N + (pj - G)) * N + (pi - G);
  }
};

}  // namespace amrh
