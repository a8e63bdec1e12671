/*
 * pf_oracle.h — CPU oracle for the Steinbach multi-phase-field + solute-diffusion step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++/OpenMP restatement of the reference algorithm
 * (zhubq/PhaseField, C++-openmp-cuda/src), used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker and the timed CPU path.  Nothing in the product
 * (phasefield_amd/) links, loads or calls it.
 *
 * orc_params has exactly the field layout of pf_params in include/phasefield_amd.h, so one
 * parameter block can be handed to both.
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  double dx, eta, cutoff;
  int dims, reserved0;
  double mobility[3][3], sigma[3][3], entropy[3][3], eq_conc[3][3], eq_slope[3][3];
  double part_k[2][3][3], part_c[2][3][3], diffusivity[2][3];
  double pfm_dt;
  double dt_d[2];
  int n_sub[2];
} orc_params;

/* forms */
#define ORC_FP64 0
#define ORC_MIXED 1
#define ORC_DUAL 2 /* FP64 with forward-mode derivatives (reference gradient) */

void *orc_create(int nx, int ny, int nz, int depth, int nsolute, int form, int bc_z);
void orc_destroy(void *h);
void orc_set_params(void *h, const orc_params *p);
/* Reference layout: index/value [N][K], phase [N][3], conc [S][N].  value_conc is derived
 * (Partitioning of conc with phase), as after Construct (MicrostructureConstruction.cpp:54-62). */
void orc_set_state(void *h, const int32_t *index, const double *value, const double *phase,
                   const double *conc);
/* Mixed form: the stored compensations comp [S][N] after orc_set_state (values on the 4-bit code
 * grid, as a mixed device download gives them); ignored by the other forms. */
void orc_set_comp(void *h, const double *comp);
/* conc: stored total concentration (the mixed form's effective value is conc - comp); comp may
 * be NULL; phase_conc [S][N][3] may be NULL. */
void orc_get_state(void *h, int32_t *index, double *value, double *phase, double *conc,
                   double *comp, double *phase_conc);
void orc_pf_sweep(void *h);
void orc_diffusion_substeps(void *h, int solute, int n, double dt_d);
void orc_step(void *h, int n_steps);
/* Diffusion sub-steps with a precomputed factor dt_d/(dx*dx) (the CUDA offload's dt_dxdx). */
void orc_diffusion_substeps_factor(void *h, int solute, int n, double factor);
/* FP64 diagnostics: mass[S], phase_frac[3], max_conc[S], min_conc[S], max_sum_err. */
void orc_reduce(void *h, double *mass, double *phase_frac, double *max_conc, double *min_conc,
                double *max_sum_err, int64_t *n_interface);
int orc_num_threads(void);

/* ---- dual-number reference gradient (form ORC_DUAL) ---------------------------------------
 * Direction seeds: derivative component 0 = mobility scale, 1 = sigma scale, 2 = D(phase1,C),
 * 3 = initial carbon field along dc0_dir.  After orc_dual_seed the step functions propagate
 * derivatives; orc_dual_phase_frac returns d f_gamma / d theta_k; orc_dual_conc returns c and
 * dc/dtheta_k fields. */
void orc_dual_seed(void *h, const double *dc0_dir);
void orc_dual_phase_fraction(void *h, double *value, double *deriv4);
void orc_dual_conc(void *h, double *c, double *dc /* [4][N] */);

/* ---- setup restatements ---------------------------------------------------------------------- */
/* Parity unpinned at the Blitz++ boundary: the reference draws its seed coordinates and its
 * nucleation noise from Blitz ranlib streams and picks nuclei with Blitz's 3D maxIndex, none of
 * which is available here or fixed by a reference test.  These restatements use std::mt19937 seeds
 * (pinned by the C++ standard's 10000th output), a counter-based noise and the first maximum in
 * storage order, each pinned by brute force; the device setup is compared against them, and the
 * step's parity does not depend on them (GPU and oracle start from the same state). */
/* Tessellation (MicrostructureConstruction.cpp:68-109) over 1-based cells; grain_index [N]. */
void orc_tessellation(int nx, int ny, int nz, const double *seeds, int n_grains, double reduction,
                      int32_t *grain_index);
/* Construct(true,..) (MicrostructureConstruction.cpp:4-65) without the Update_parameter call:
 * slot 1 = 200000+grain, Iso_growth iterations, total_conc = dot(phase, init_conc), partition. */
void orc_construct(void *h, const int32_t *grain_index, int n_iso, const double *init_conc);
/* Nucleate1in2 (nucleation.cpp:3-43) with the counter-based noise, radius and nucleus
 * concentration options. */
void orc_nucleate(void *h, int n_nuclei, float radius, const double *nucleus_conc, uint64_t seed);
void orc_solute_noise(void *h, int solute, double mean, double stddev, uint64_t seed);

/* ---- host-parameter restatements ------------------------------------------------------------ */
/* Update_parameter (phase_field.cpp:307-315) + thermo_data::Get_data (thermo_data.cpp:24-40).
 * thermo: entropy lines [3][4] (i j a b), para lines [6][5] (i j a b c).  Outputs pf_params-
 * compatible tables; returns pfm_dt, diff_dt via pointers. */
void orc_update_parameter(double temperature, double dx, double dims, const double *mob0,
                          const double *mobQ, const double *sigma, const double *diff0,
                          const double *diffQ, const double *entropy_lines,
                          const double *para_lines, orc_params *out, double *pfm_dt,
                          double *diff_dt);
uint32_t orc_mt19937_nth(uint32_t seed, int n);
uint64_t orc_splitmix64(uint64_t x);
// the mixed form's stored compensation of comp for a cell holding value, with rounding key key
// (stochastic rounding to a multiple of 2^-7 ulp(value)), and the integer hash the keys come from
float orc_quantize_comp(float comp, float value, uint32_t key);
uint32_t orc_mix32(uint32_t x);

#ifdef __cplusplus
}
#endif
#endif
