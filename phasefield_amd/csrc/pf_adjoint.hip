// pf_adjoint.hip — discrete adjoint of the fused step (gradient harness).
#include <hip/hip_runtime.h>

#include "../../include/phasefield_amd.h"

extern "C" int pf_adjoint_run(pf_handle *h, const pf_adjoint_spec *spec, pf_adjoint_result *out) {
  (void)h;
  (void)spec;
  (void)out;
  return PF_ERR_INVALID;
}
