// pf_sweep_mixed.hip — mixed-form (FP32 storage, FP64 decision) instances of the phase-field
// sweep.  Kernel definition: pf_sweep_impl.h.
#include "pf_sweep_impl.h"

namespace pfamd {
PF_INST_SWEEP_ALL(FormMixed, 4, 2, false)
PF_INST_SWEEP_ALL(FormMixed, 4, 3, false)
PF_INST_SWEEP_ALL(FormMixed, 6, 2, false)
PF_INST_SWEEP_ALL(FormMixed, 6, 3, false)
}  // namespace pfamd
