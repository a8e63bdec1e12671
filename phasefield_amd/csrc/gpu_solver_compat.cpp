// gpu_solver_compat.cpp — the reference's accelerator API with its own C++ linkage
// (C++-openmp-cuda/src/gpu_solver.h:18-23, included by phase_field.h:14).  The reference's
// phase_field.cpp calls these three names from a C++ translation unit, so the library exports them
// as C++ functions (mangled as the reference's prototypes mangle) over the pf_gpu_* C entry points.
#include <cstdio>

#include "../../include/phasefield_amd.h"

int GPU_Init(const int x, const int y, const int z) { return pf_gpu_init(x, y, z); }

void GPU_Diffusion_solver(double *hdiffusivity, double *hpartition_k, double *hpartition_c,
                          double *hvalue_phase, double *htotal_conc, int x, int y, int z,
                          double dt, double dx) {
  // the reference returns nothing; the C entry point reports a failure on stderr
  (void)pf_gpu_diffusion_solver(hdiffusivity, hpartition_k, hpartition_c, hvalue_phase,
                                htotal_conc, x, y, z, dt, dx);
}

void GPU_Cleanup() { pf_gpu_cleanup(); }
