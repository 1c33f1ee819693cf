// Order of a block's faces in the LDS face buffers of the block and rotating kernels.
//
// Face q of direction d lies on plane f (0..N) along d at tangential cell indices (ta, tb): x faces
// at (j, k), y faces at (i, k), z faces at (i, j).  The tangential index that is contiguous in the
// staged pencils varies fastest, so the reads a wavefront makes of the staged primitives and of the
// face buffer are unit-stride in LDS for every direction.
#pragma once
#include <hip/hip_runtime.h>

namespace amrh {

template <int N, int d>
__device__ __forceinline__ int face_index(int f, int ta, int tb) {
  if constexpr (d == 0) return f + (N + 1) * (ta + N * tb);
  if constexpr (d == 1) return ta + N * (f + (N + 1) * tb);
  return ta + N * (tb + N * f);
}

template <int N, int d>
__device__ __forceinline__ void face_decode(int q, int& f, int& ta, int& tb) {
  if constexpr (d == 0) {
    f = q % (N + 1);
    ta = (q / (N + 1)) % N;
    tb = q / ((N + 1) * N);
  } else if constexpr (d == 1) {
    ta = q % N;
    f = (q / N) % (N + 1);
    tb = q / (N * (N + 1));
  } else {
    ta = q % N;
    tb = (q / N) % N;
    f = q / (N * N);
  }
}

}  // namespace amrh
