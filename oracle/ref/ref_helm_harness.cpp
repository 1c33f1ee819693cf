// ORACLE support -- test infrastructure only.  Linked with the reference's own
// helmholtz/helmeos.cpp (see ref/Makefile) in this container; never shipped, never run on the GPU
// box.  helmeos.cpp reads ./helmholtz.table when the library is loaded (helmholtz/helmeos.cpp:66-96),
// so the loader must make that file the current directory's first.  States are the reference's
// eos_t records (helmholtz/helmholtz.h:4-6), passed as 10 doubles each.
#include <cstring>

#include "helmholtz/helmholtz.h"

extern "C" {

void ref_helm_set_units(double cm, double g, double s, double K) { helmholtz_set_cgs_units(cm, g, s, K); }

void ref_helm_eos(int n, double* st) {
  for (int k = 0; k < n; k++) {
    eos_t q;
    std::memcpy(&q, st + 10 * (long)k, sizeof q);
    helmholtz_eos(&q);
    std::memcpy(st + 10 * (long)k, &q, sizeof q);
  }
}

void ref_helm_compute_T(int n, double* st) {
  for (int k = 0; k < n; k++) {
    eos_t q;
    std::memcpy(&q, st + 10 * (long)k, sizeof q);
    helmholtz_compute_T(&q);
    std::memcpy(st + 10 * (long)k, &q, sizeof q);
  }
}

}  // extern "C"
