// Device view of an EOS table object (amrh_eos_create), for kernels that evaluate the EOS.
#pragma once
#include "amrhydro.h"
#include "helmholtz.h"

namespace amrh {
const helm::HelmDev* eos_device(const amrh_eos* E);
}  // namespace amrh
