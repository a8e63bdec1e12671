// pf_sweep_list.hip — listed-cell instances of the phase-field sweep (both forms) and the copy-cell
// classification that builds their list.  Kernel definitions: pf_sweep_impl.h.
#include "pf_sweep_impl.h"

namespace pfamd {
PF_INST_SWEEP_LIST(FormF64, 4, 2)
PF_INST_SWEEP_LIST(FormF64, 4, 3)
PF_INST_SWEEP_LIST(FormF64, 6, 2)
PF_INST_SWEEP_LIST(FormF64, 6, 3)
PF_INST_SWEEP_LIST(FormMixed, 4, 2)
PF_INST_SWEEP_LIST(FormMixed, 4, 3)
PF_INST_SWEEP_LIST(FormMixed, 6, 2)
PF_INST_SWEEP_LIST(FormMixed, 6, 3)
PF_INST_SWEEP_CLASSIFY(FormF64, 4)
PF_INST_SWEEP_CLASSIFY(FormF64, 6)
PF_INST_SWEEP_CLASSIFY(FormMixed, 4)
PF_INST_SWEEP_CLASSIFY(FormMixed, 6)
}  // namespace pfamd
