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
// FP64, K = 4 lists: the register path with its overflow list, then the general path over it
#define PF_INST_SWEEP_SPLIT(F, P)                            \
  PF_INST_SWEEP_M(F, 4, P, 4, false, false, true, 1)         \
  PF_INST_SWEEP_M(F, 4, P, 6, false, false, true, 1)         \
  PF_INST_SWEEP_M(F, 4, P, 6, true, false, true, 1)          \
  PF_INST_SWEEP_M(F, 4, P, 4, false, false, true, 2)         \
  PF_INST_SWEEP_M(F, 4, P, 6, false, false, true, 2)         \
  PF_INST_SWEEP_M(F, 4, P, 6, true, false, true, 2)
PF_INST_SWEEP_SPLIT(FormF64, 2)
PF_INST_SWEEP_SPLIT(FormF64, 3)
PF_INST_SWEEP_CLASSIFY(FormF64, 4)
PF_INST_SWEEP_CLASSIFY(FormF64, 6)
PF_INST_SWEEP_CLASSIFY(FormMixed, 4)
PF_INST_SWEEP_CLASSIFY(FormMixed, 6)
}  // namespace pfamd
