// pf_sweep_f64.hip — FP64-form instances of the phase-field sweep (and of the Iso_growth variant
// used by the seeded setup).  Kernel definition: pf_sweep_impl.h.
#include "pf_sweep_impl.h"

namespace pfamd {
PF_INST_SWEEP_ALL(FormF64, 4, 2, false)
PF_INST_SWEEP_ALL(FormF64, 4, 3, false)
PF_INST_SWEEP_ALL(FormF64, 6, 2, false)
PF_INST_SWEEP_ALL(FormF64, 6, 3, false)
PF_INST_SWEEP_ALL(FormF64, 4, 2, true)
PF_INST_SWEEP_ALL(FormF64, 4, 3, true)
PF_INST_SWEEP_ALL(FormF64, 6, 2, true)
PF_INST_SWEEP_ALL(FormF64, 6, 3, true)
}  // namespace pfamd
