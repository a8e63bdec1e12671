/*
 * phasefield_amd.h — C-ABI of the MI355X (gfx950) Steinbach multi-phase-field + solute-diffusion step.
 *
 * Plain C types only (pointers, sizes, PODs).  One handle owns all device memory of one simulation
 * domain on one GPU and one HIP stream.  Every entry point returns an int status (PF_OK == 0, <0 on
 * error) and never exits the process; pf_last_error() names the failure.
 *
 * Reference interfaces replaced (paths relative to the zhubq/PhaseField tree):
 *   GPU_Init / GPU_Diffusion_solver / GPU_Cleanup   C++-openmp-cuda/src/gpu_solver.h:18-23
 *   phase_field::Update(bool)                         C++-openmp-cuda/src/phase_field.cpp:243-305
 *   phase_field::Update_parameter()                   C++-openmp-cuda/src/phase_field.cpp:307-315
 *   phase_field::Read_data(const char*)               C++-openmp-cuda/src/phase_field.cpp:31-231
 *   phase_field::Construct(bool,char[])               C++-openmp-cuda/src/MicrostructureConstruction.cpp:4-65
 *   nucleation::Nucleate1in2(phase_field&)            C++-openmp-cuda/src/nucleation.cpp:3-43
 *   phase_field::Write_data(bool)                     C++-openmp-cuda/src/write_data.cpp:3-40
 */
#ifndef PHASEFIELD_AMD_H
#define PHASEFIELD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_MAX_PHASE 3   /* max_num_phase, constant.h:6 */
#define PF_MAX_SOLUTE 2  /* carbon (the reference's only solute) + manganese */
#define PF_MAX_DEPTH 6   /* max_num_coexist, constant.h:5 */

/* status codes */
#define PF_OK 0
#define PF_ERR_INVALID (-1)   /* bad argument / shape / parameter */
#define PF_ERR_NODEVICE (-2)  /* no HIP device (GPU_Init returning 0, gpu_solver.cpp:33-36) */
#define PF_ERR_OOM (-3)       /* device allocation failed (gpu_solver.cpp:50-53) */
#define PF_ERR_HIP (-4)       /* any other HIP runtime error */
#define PF_ERR_NONFINITE (-5) /* a reduction found NaN/Inf in the state */
#define PF_ERR_STATE (-6)     /* state not uploaded / inconsistent call order */
#define PF_ERR_IO (-7)        /* input/output file error */
#define PF_ERR_RANGE (-8)     /* too many grains per phase for the 16-bit slot index */

/* precision forms */
#define PF_FP64 0   /* phi, phase fractions and concentrations stored and computed in FP64 */
#define PF_MIXED 1  /* FP32 storage/compute, Kahan-compensated concentration increments,
                       FP64 renormalisation / pruning decision and FP64 reductions */

/* boundary conditions along z (x and y are always periodic, phase_field.h:58-68) */
#define PF_BC_PERIODIC 0
#define PF_BC_NOFLUX 1

/* Shape and storage form of one domain. */
typedef struct pf_desc {
  int nx, ny, nz;  /* cells; the reference's ix, iy, iz (phase_field.cpp:55) */
  int depth;       /* active-set depth K: 4 or 6 (max_num_coexist) */
  int nphase;      /* phases carried per cell: 2 or 3 (max_num_phase = 3; 2 when phase 3 is empty) */
  int nsolute;     /* 1 (carbon) or 2 (carbon, manganese) */
  int precision;   /* PF_FP64 or PF_MIXED */
  int bc_z;        /* PF_BC_PERIODIC or PF_BC_NOFLUX */
  int device;      /* HIP device ordinal */
} pf_desc;

/* Scalar parameters of one time step, as phase_field::Update_parameter leaves them
 * (phase_field.cpp:307-315, thermo_data.cpp:24-40).  Phase indices are 0-based here:
 * [i][j] holds the reference's (i+1, j+1). */
typedef struct pf_params {
  double dx;      /* grid spacing [um] */
  double eta;     /* interface width dx*i_intf [um] (phase_field.cpp:228) */
  double cutoff;  /* phase minimum (phase_field.cpp:67) */
  int dims;       /* the reference's `dimensions`: 3 - [iz==1] - [iy==1]; 2*dims stencil faces */
  int reserved0;
  double mobility[3][3];     /* M(i,j) at the current temperature */
  double sigma[3][3];        /* interface energy */
  double entropy[3][3];      /* thermodynamic.para_entropy */
  double eq_conc[3][3];      /* thermodynamic.para_equilibrium_conc */
  double eq_slope[3][3];     /* thermodynamic.para_equilibrium_dc_dT */
  double part_k[PF_MAX_SOLUTE][3][3]; /* partition slopes k(i,j) per solute (para_partition_k) */
  double part_c[PF_MAX_SOLUTE][3][3]; /* partition offsets c(i,j) per solute (para_partition_c) */
  double diffusivity[PF_MAX_SOLUTE][3]; /* D per solute per phase (C_diffusivity) */
  double pfm_dt;                        /* phase-field time step */
  double dt_d[PF_MAX_SOLUTE];           /* diffusion sub-step per solute */
  int n_sub[PF_MAX_SOLUTE];             /* diffusion sub-steps per PF step, per solute */
} pf_params;

/* Host-side state in the reference's layout (Blitz fortranArray, first index fastest):
 * cell n = i + nx*(j + ny*k), slots and phases fastest within a cell.  Grain ids are the
 * reference's (200000+g for Voronoi grains, 100000+n for nuclei, 0 = empty slot). */
typedef struct pf_host_state {
  int32_t *index;     /* [N][K]  index_grain1 */
  double *value;      /* [N][K]  value_grain1 */
  double *phase;      /* [N][3]  value_phase */
  double *conc;       /* [nsolute][N] total concentration (total_conc; solute 1 = Mn) */
  double *conc_comp;  /* [nsolute][N] Kahan compensation of the mixed form (effective value
                       * conc - conc_comp; stored as a multiple of 2^-7 ulp of conc, an upload
                       * rounds down to that grid), or NULL */
  double *phase_conc; /* [nsolute][N][3] partitioned concentrations (value_conc1); download only, or NULL */
} pf_host_state;

/* FP64 diagnostics, reduced on the device in a fixed order (deterministic). */
typedef struct pf_stats {
  double mass[PF_MAX_SOLUTE];       /* sum of total concentration over cells (compensated in mixed form) */
  double phase_frac[3];             /* sum of value_phase / N  (write_data.cpp:10-18) */
  double max_conc[PF_MAX_SOLUTE];
  double min_conc[PF_MAX_SOLUTE];
  double max_sum_err;               /* max over cells of |sum_s phi_s - 1| */
  int64_t n_interface;              /* cells with a second active slot */
  int64_t n_nonfinite;              /* cells holding NaN/Inf in phi or c */
} pf_stats;

typedef struct pf_handle pf_handle;

/* ---- device solver -------------------------------------------------------------------------- */

/* Allocate a domain on desc->device (replaces GPU_Init, gpu_solver.cpp:26-99). */
int pf_create(const pf_desc *desc, pf_handle **out);
/* Free everything the handle owns (replaces GPU_Cleanup, gpu_solver.cpp:223-239). */
int pf_destroy(pf_handle *h);
/* Text of the last error on this handle (or of the last failed pf_create when h == NULL). */
const char *pf_last_error(const pf_handle *h);
/* Set the step parameters (Update_parameter's outputs). */
int pf_set_params(pf_handle *h, const pf_params *p);
/* Copy a reference-layout host state onto the device (grain ids remapped to 16-bit codes).
 * The partitioned concentrations are derived on the device from conc and phase. */
int pf_upload_state(pf_handle *h, const pf_host_state *s);
/* Copy the device state back in the reference layout (synchronises the stream). */
int pf_download_state(pf_handle *h, pf_host_state *s);
/* n full steps of phase_field::Update(false) with the current parameters:
 * one phase-field sweep, then per solute n_sub diffusion sub-steps.  Asynchronous. */
int pf_step(pf_handle *h, int n_steps);
/* The phase-field sweep alone (phase_field::Update(true), phase_field.cpp:252-262). */
int pf_phase_field_sweep(pf_handle *h);
/* n diffusion sub-steps of one solute with sub-step dt_d (phase_field.cpp:273-283). */
int pf_diffuse_substeps(pf_handle *h, int solute, int n, double dt_d);
/* FP64 reductions over the current state (synchronises). */
int pf_reduce(pf_handle *h, pf_stats *out);
/* Wait for all queued work on the handle's stream. */
int pf_synchronize(pf_handle *h);
/* Device time in ms of the dominant kernel launches since the last reset, and how many launches
 * it covers (HIP event pairs on the handle's stream); kind 0 = diffusion sub-step (one pair around
 * each batch of back-to-back sub-steps, leaving out a batch's first sub-step, which also classifies
 * bulk cells), 1 = phase-field sweep, 2 = the plane-marching bulk kernel alone (one pair around
 * every eighth of those sub-steps' bulk launches). */
int pf_kernel_timing(pf_handle *h, int kind, double *total_ms, int64_t *launches);
int pf_kernel_timing_reset(pf_handle *h, int enable);
/* Lengths of the per-sweep cell lists the step works through (synchronises): the cells the last
 * phase-field sweep updated with the full formula (the rest were copied), and the cells the
 * diffusion sub-steps update with the full formula (the rest take the bulk update). */
int pf_list_lengths(pf_handle *h, int64_t *sweep_listed, int64_t *diffusion_listed);
/* Tile rows and z-chunk of the plane-marching diffusion kernel on this handle (picked by timing the
 * candidate shapes at the first sub-step); 0, 0 before that or when that kernel is not in use. */
int pf_march_shape(const pf_handle *h, int *tile_rows, int *z_chunk);
/* Size in bytes of one checkpoint of this handle.  save/restore copy the whole state, the mixed
 * form's sub-step counters and the grain-code tables to / from host memory of that size; restore
 * rejects a blob whose header names another domain shape, form or size. */
int pf_checkpoint_size(const pf_handle *h, size_t *bytes);
int pf_checkpoint_save(pf_handle *h, void *host_dst);
int pf_checkpoint_restore(pf_handle *h, const void *host_src);

/* ---- seeded setup on the device (MicrostructureConstruction.cpp, nucleation.cpp) ------------ */

/* Voronoi tessellation + diffuse-interface growth + initial partition (Construct(true,..)).
 * seeds: [3*G] seed coordinates (x1,y1,z1,x2,..) in 1-based cell units, as Tessellation draws them.
 * init_conc: [nsolute][3] per-phase initial concentration.  iso_factor = 0.5/dims*dx*dx.
 * Leaves value_phase, total_conc, slots set; time bookkeeping is the host's. */
int pf_setup_construct(pf_handle *h, const double *seeds, int n_grains, double reduction,
                       int n_iso_iterations, const double *init_conc);
/* Place n nuclei of phase 1 (ids 100001..) at first-max cells of the nucleation flag
 * (nucleation.cpp:3-43).  radius in cells; nucleus_conc[s] < 0 keeps the matrix concentration
 * of solute s; seed drives the per-cell noise. */
int pf_setup_nucleate(pf_handle *h, int n_nuclei, float radius, const double *nucleus_conc,
                      uint64_t seed);
/* Set one solute's total concentration to mean + stddev * g per cell, g an Irwin-Hall(12)
 * approximation of a standard Gaussian from the counter-based generator keyed by seed. */
int pf_setup_solute_noise(pf_handle *h, int solute, double mean, double stddev, uint64_t seed);
/* Copy the state of an FP64 domain into a domain of the same shape on the same device,
 * rounding to FP32 (compensation zero) when dst is a mixed domain.  Setup runs in FP64 only. */
int pf_copy_state(pf_handle *dst, const pf_handle *src);

/* ---- discrete adjoint (gradient harness) ---------------------------------------------------- */

/* Loss: sum over sample steps t_k of (f_gamma(t_k) - target_k)^2 + w_c * ||c_final - c_target||^2,
 * with f_gamma = mean value_phase of phase 1.  Gradients w.r.t. the mobility scale, the
 * interface-energy scale, the phase-1 carbon diffusivity and the initial carbon field, with
 * forward checkpoints every `checkpoint_every` steps recomputed in reverse.  dt and n_sub are
 * held fixed.  FP64 form only. */
typedef struct pf_adjoint_spec {
  int n_steps;
  int checkpoint_every;
  int n_samples;               /* number of sample steps (<= 64) */
  int sample_steps[64];        /* 1-based step numbers at which f_gamma enters the loss */
  double target_fraction[64];
  double weight_conc;          /* w_c */
  const double *target_conc;   /* host [N] carbon target, or NULL for no concentration term */
} pf_adjoint_spec;

typedef struct pf_adjoint_result {
  double loss;
  double d_mobility;   /* dL/d(scale of every mobility entry), at scale 1 */
  double d_sigma;      /* dL/d(scale of every interface-energy entry), at scale 1 */
  double d_diff_gamma; /* dL/dD(phase 1, carbon) */
  double *d_conc0;     /* host [N] dL/dc0 (initial carbon total concentration), or NULL */
  double fraction[64]; /* f_gamma at the sample steps (forward) */
  double forward_ms, reverse_ms;
} pf_adjoint_result;

/* Run forward with checkpoints from the current state, then the reverse sweep.  The device
 * state after the call is the final forward state. */
int pf_adjoint_run(pf_handle *h, const pf_adjoint_spec *spec, pf_adjoint_result *out);

/* ---- host model: input files, temperature path, parameter update ---------------------------- */

typedef struct pf_model pf_model;

/* Parse the positional main input file and the thermodynamics file it names
 * (phase_field.cpp:31-231, thermo_functor.cpp:11-53).  Optional `key=value` lines after the
 * GPU flag set extensions (steps, depth, precision, bc_z, nucleus_radius, nucleus_conc, seed,
 * nucleation_seed, solutes, mn_*).  The thermo path is resolved relative to the main file's
 * directory when it is not absolute. */
int pf_model_read(const char *path, pf_model **out);
int pf_model_free(pf_model *m);
const char *pf_model_last_error(void);
/* Integer/real/string fields by name (see pf_model_get_* keys in the driver docs). */
int pf_model_get_int(const pf_model *m, const char *key, int64_t *out);
int pf_model_get_real(const pf_model *m, const char *key, double *out);
int pf_model_get_string(const pf_model *m, const char *key, char *buf, size_t len);
int pf_model_set_real(pf_model *m, const char *key, double value);
int pf_model_set_int(pf_model *m, const char *key, int64_t value);
/* Fill a pf_desc from the model (shape, depth, phases, solutes, precision, bc). */
int pf_model_desc(const pf_model *m, pf_desc *out);
/* phase_field::Update_parameter: temperature at the current time, Arrhenius mobility and
 * diffusivity, pfm_dt / diff_dt, thermo tables; then time += pfm_dt. */
int pf_model_update_parameter(pf_model *m, pf_params *out);
/* Temperature_function(time) (phase_field.cpp:232-242, clamped to the last segment). */
double pf_model_temperature(const pf_model *m, double time);
/* Voronoi seed coordinates as Tessellation draws them (MicrostructureConstruction.cpp:68-85):
 * [3*G] doubles from the seeded generator (std::mt19937 seeded with G). */
int pf_model_voronoi_seeds(const pf_model *m, double *seeds, int n_grains);
/* Whether Write_data would write now (write_data.cpp:7-8): forced, at the output cadence, or
 * after a temperature change of more than 10 K. */
int pf_model_output_due(const pf_model *m, int forced, int *due);
/* Kinetics row (write_data.cpp:9-18) appended to <dir>/fraction_time.txt if due or forced. */
int pf_model_write_kinetics(pf_model *m, const pf_stats *st, double f_rex, int forced, int *written);
/* The kinetics file's last column: sum of slot phi over grain ids below 100000, divided by N
 * (write_data.cpp:18).  Exactly 0 without a table entry below 100000; else downloads the slots. */
int pf_rex_fraction(pf_handle *h, double *out);
/* ASCII VTK ImageData dump (write_data.cpp:42-82) of a reference-layout host state. */
int pf_write_vti(const char *path, int nx, int ny, int nz, int depth, double dx,
                 const int32_t *index, const double *conc);
/* The same fields (korn as Int32, the concentration as Float32) as a binary VTK ImageData file
 * with raw appended data.  No reference counterpart: the reference writes only the ASCII form. */
int pf_write_vti_raw(const char *path, int nx, int ny, int nz, double dx, const int32_t *korn,
                     const float *conc);
/* Asynchronous field dump of the device state (the fields of pf_write_vti_raw, carbon): the
 * fields are extracted on the device in stream order, copied to pinned host memory on a second
 * stream and written by a host thread while the caller keeps stepping.  A second call, and
 * pf_destroy, first wait for the previous dump.  No reference counterpart (write_data.cpp:42-82
 * downloads and formats synchronously); it is the dump path for domains where that stalls. */
int pf_dump_vti_async(pf_handle *h, const char *path, double dx);
/* Wait for the last pf_dump_vti_async; returns its status (PF_OK when none is pending). */
int pf_dump_wait(pf_handle *h);
/* Full-precision binary dump / reload of a reference-layout host state. */
int pf_write_state(const char *path, const pf_desc *d, double time, int64_t step,
                   const pf_host_state *s);
int pf_read_state(const char *path, const pf_desc *d, double *time, int64_t *step,
                  pf_host_state *s);

/* ---- compatibility shim: the reference's accelerator API (gpu_solver.h:18-23) --------------- */

/* C entry points of the shim (C callers, FFI).  C++ callers use the reference's own names below. */
/* Returns 1 if a device is usable for an x*y*z domain, else 0 (CPU fallback in the caller):
 * GPU_Init, gpu_solver.cpp:26-99. */
int pf_gpu_init(int x, int y, int z);
/* Diffusion-only offload with host pointers, as GPU_Diffusion_solver (gpu_solver.cpp:204-218):
 * copies value_phase[3N] and total_conc[N] in, runs ceil(dt/del_t) partition+diffusion sub-steps,
 * copies total_conc out.  Returns PF_OK or an error status (the reference returns nothing). */
int pf_gpu_diffusion_solver(const double *hdiffusivity, const double *hpartition_k,
                            const double *hpartition_c, const double *hvalue_phase,
                            double *htotal_conc, int x, int y, int z, double dt, double dx);
/* Frees the shim's domain: GPU_Cleanup, gpu_solver.cpp:223-239. */
void pf_gpu_cleanup(void);
/* No counterpart in the reference: the allocation-free form of pf_gpu_diffusion_solver, for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same sub-step count and arithmetic (bitwise the same total_conc), but
 * value_phase [N][3] and total_conc [N] are device pointers, the sub-steps are queued on `stream`
 * (a hipStream_t; NULL for the null stream) and the call returns without waiting.  It allocates
 * nothing: its second concentration buffer is a device array of the library with room for 2^21
 * cells (16 MiB on every GPU the library is loaded on, kept until the process ends); a larger
 * domain returns PF_ERR_RANGE.  Calls whose work overlaps on one GPU share that buffer and race;
 * calls ordered on one stream do not.  Errors: pf_last_error(NULL). */
int pf_gpu_diffusion_solver_noalloc(const double *hdiffusivity, const double *hpartition_k,
                                    const double *hpartition_c, const double *dvalue_phase,
                                    double *dtotal_conc, int x, int y, int z, double dt,
                                    double dx, void *stream);

#ifdef __cplusplus
}

/* The reference's accelerator API with the reference's C++ linkage, exactly as
 * C++-openmp-cuda/src/gpu_solver.h:18-23 declares it (real = double, real.h), so that the
 * reference's phase_field.cpp, which includes gpu_solver.h (phase_field.h:14), links against this
 * library unchanged.  Each forwards to the pf_gpu_* entry point above. */
int GPU_Init(const int x, const int y, const int z);
void GPU_Diffusion_solver(double *hdiffusivity, double *hpartition_k, double *hpartition_c,
                          double *hvalue_phase, double *htotal_conc, int x, int y, int z,
                          double dt, double dx);
void GPU_Cleanup();
#endif

#endif /* PHASEFIELD_AMD_H */
