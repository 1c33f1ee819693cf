// Buffer-resource views of the per-block arrays the kernels stream: a wavefront-uniform base in
// SGPRs and a 32-bit per-thread byte offset, which buffer loads and stores take directly.  A thread
// then keeps one 32-bit offset per cell instead of a 64-bit pointer per field; the compiler hoists
// such pointers out of the plane and block loops, one register pair each, and spills them.
// Every access the kernels make is in range (the range check would return 0 / drop the store).
#pragma once
#include <hip/hip_runtime.h>

namespace amrh {

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// FP64 elements [0, n) at base (n * 8 bytes below 2 GiB)
struct F64Buf {
  __amdgpu_buffer_rsrc_t r;
  __device__ F64Buf(const double* base, long n)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, (int)(n * 8), 0x00020000)) {}
  // element e + off; off is wavefront-uniform (the field's offset)
  __device__ double ld(unsigned e, long off = 0) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, e * 8u, (int)(off * 8), 0));
  }
  __device__ void st(unsigned e, double v, long off = 0) const {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, e * 8u, (int)(off * 8), 0);
  }
  // streaming forms: the non-temporal cache policy (data read or written once), so that they do
  // not push data that is read again out of the L2
  __device__ double ld_nt(unsigned e, long off = 0) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, e * 8u, (int)(off * 8), kNT));
  }
  __device__ void st_nt(unsigned e, double v, long off = 0) const {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, e * 8u, (int)(off * 8), kNT);
  }
  static constexpr int kNT = 2;  // the cache-policy operand's NT bit on gfx950
};

// a block record: NF fields of R elements each
template <long R, int NF>
struct RecBuf : F64Buf {
  static_assert((long)NF * R * 8 < (1L << 31), "a block record must fit one buffer resource");
  __device__ explicit RecBuf(const double* blk) : F64Buf(blk, NF * R) {}
  __device__ double ld(int l, unsigned e) const { return F64Buf::ld(e, l * R); }
  __device__ void st(int l, unsigned e, double v) const { F64Buf::st(e, v, l * R); }
  __device__ double ld_nt(int l, unsigned e) const { return F64Buf::ld_nt(e, l * R); }
  __device__ void st_nt(int l, unsigned e, double v) const { F64Buf::st_nt(e, v, l * R); }
};

}  // namespace amrh
