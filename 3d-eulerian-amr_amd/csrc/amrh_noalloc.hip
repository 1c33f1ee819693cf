// amrh_state_sum_dev: amrh_state_sum without any allocation, for callers that cannot allocate
// device memory (work captured into a HIP graph, per-step loops that forbid hipMalloc).
//
// amrh_state_sum keeps its per-block partial sums in a buffer the context allocates on first use and
// copies the result to the host.  Here the partials live in `scratch`, a fixed device array of this
// library (one copy per GPU, held until the process ends), and the result goes to a device pointer
// the caller owns; the call only launches the same two kernels on the context's stream.
//
// scratch is reached through its device address, looked up once per device with
// hipGetSymbolAddress (in host code the array's name is not a device address).  Every block writes
// all of its partials on every call, so nothing needs clearing between calls.  All calls on one GPU
// share the array: calls that overlap in time on one GPU (two streams) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "amrh_ctx.h"
#include "amrhydro.h"

using namespace amrh;

namespace {

// 24 MiB: the partials of nblocks * nf <= 524,288 (6 per block and field), e.g. 87,381 blocks of
// the 6-field ideal gas or 40,329 blocks of the 13-field rotating state
constexpr long kScratchDoubles = AMRH_STATE_SUM_DEV_MAX_PARTIALS;
static_assert(kScratchDoubles == 3L << 20, "the header's limit");
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// the device address of this GPU's scratch, looked up on the first call on the device
int scratch_ptr(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(AMRH_EINVAL, "device %d beyond the library's table", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(AMRH_EHIP, "hipGetSymbolAddress gave a null address");
    g_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_ptr[dev];
  return AMRH_OK;
}

}  // namespace

extern "C" {

int amrh_state_sum_dev(amrh_ctx* c, double* sums_dev) {
  if (!c || !sums_dev) return fail(AMRH_EINVAL, "null argument");
  const long need = diag_part_elems(c);
  if (need > kScratchDoubles)
    return fail(AMRH_EINVAL,
                "%d blocks x %d fields need %ld partials; the no-allocation scratch holds %ld (use amrh_state_sum)",
                c->nb, c->cfg.nf, need, kScratchDoubles);
  if (!c->tree && !c->have_x0)
    return fail(AMRH_ESTATE, "the moments need block positions: call amrh_set_block_origin");
  if (int rc = diag_check(c)) return rc;
  double* part = nullptr;
  if (int rc = scratch_ptr(&part)) return rc;
  return diag_launch(c, part, sums_dev, true);
}

}  // extern "C"
