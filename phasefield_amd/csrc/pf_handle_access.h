// pf_handle_access.h — internal view of a domain handle for the translation units that drive
// kernels outside pf_api.hip (the adjoint).  Defined in pf_api.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/phasefield_amd.h"
#include "pf_device.h"

struct HandleView {
  long N;
  int nx, ny, nz, K, KI, P, NF, nsolute, device;
  bool noflux, mixed, have_state;
  const pf_params *params;  // null until pf_set_params
  hipStream_t stream;
  uint16_t *id;  // current slot ids
  void *phi;     // current slot values
  void *pf;      // value_phase
  void *c0;      // current carbon concentration
  unsigned bulk_mask;  // phases (bit p) whose bulk cells take the exact reduced sub-step
};

HandleView handle_view(pf_handle *h);
// record an error that belongs to no handle (pf_last_error(NULL) returns it) and return `code`
int set_global_error(int code, const std::string &msg);
// record an error message on the handle and return `code`
int handle_fail(pf_handle *h, int code, const std::string &msg);
// the device state was replaced behind the step functions' back (drops cached bulk flags)
void handle_state_changed(pf_handle *h);
pfamd::PFConsts<double> handle_pf_consts(const pf_handle *h);
pfamd::DiffConsts<double> handle_diff_consts(const pf_handle *h, int solute, double factor);
// tile rows and z-chunk the plane-marching bulk kernel runs with on this handle (0, 0 until its
// first sub-step has picked them, or when that kernel is not in use)
void handle_march_shape(const pf_handle *h, int *rows, int *zchunk);
// one FP64 carbon sub-step from cin into cout with the step's own kernel, flags and factor
// (value_phase and bulk flags of the handle's current state)
int handle_diffuse_io(pf_handle *h, const double *cin, double *cout);
