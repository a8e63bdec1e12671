// pfm_main.cpp — the driver executable (C++-openmp-cuda/src/main.cpp:10-43):
//   pfm <input file> [ignored]
// Read the positional input file, construct the seeded microstructure, write the first kinetics
// row, nucleate, then advance with phase_field::Update until end_time (or `steps=N`), writing the
// kinetics file and VTI dumps at the output cadence.  The step runs on the GPU through the C-ABI;
// there is no CPU fallback (a missing device is an error).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <string>
#include <vector>

#include "phasefield_amd.h"

namespace {

int die(const char *what, const char *msg) {
  std::fprintf(stderr, "pfm: %s: %s\n", what, msg);
  return 1;
}

struct HostState {
  std::vector<int32_t> index;
  std::vector<double> value, phase, conc, comp;
  pf_host_state view{};
  HostState(const pf_desc &d) {
    const size_t N = size_t(d.nx) * d.ny * d.nz;
    index.resize(N * d.depth);
    value.resize(N * d.depth);
    phase.resize(N * 3);
    conc.resize(N * d.nsolute);
    comp.resize(N * d.nsolute);
    view.index = index.data();
    view.value = value.data();
    view.phase = phase.data();
    view.conc = conc.data();
    view.conc_comp = comp.data();
  }
};

// Write_data (write_data.cpp:3-40): kinetics row when due, VTI dump when enabled
int write_data(pf_model *m, pf_handle *h, const pf_desc &d, bool forced, int *vti_counter) {
  int due = 0;
  if (pf_model_output_due(m, forced ? 1 : 0, &due) != PF_OK) return PF_ERR_INVALID;
  if (!due) return PF_OK;
  pf_stats st{};
  int rc = pf_reduce(h, &st);
  if (rc != PF_OK) return rc;
  double f_rex = 0.0;
  if ((rc = pf_rex_fraction(h, &f_rex)) != PF_OK) return rc;
  int written = 0;
  if ((rc = pf_model_write_kinetics(m, &st, f_rex, 1, &written)) != PF_OK) return rc;
  int64_t micro = 0;
  pf_model_get_int(m, "is_micro_out", &micro);
  char fmt[64] = "ascii";  // extension key vti_format=binary: raw VTI, written asynchronously
  pf_model_get_string(m, "vti_format", fmt, sizeof fmt);
  if (micro && std::string(fmt) == "binary") {
    char dir[4096], name[64];
    pf_model_get_string(m, "directory", dir, sizeof dir);
    double t = 0, dx = 0;
    pf_model_get_real(m, "time", &t);
    pf_model_get_real(m, "dx", &dx);
    std::snprintf(name, sizeof name, "/field%07d_%d.vti", int(t * 100), (*vti_counter)++);
    rc = pf_dump_vti_async(h, (std::string(dir) + name).c_str(), dx);
  } else if (micro) {
    HostState hs(d);
    if ((rc = pf_download_state(h, &hs.view)) != PF_OK) return rc;
    char dir[4096], name[64];
    pf_model_get_string(m, "directory", dir, sizeof dir);
    double t = 0, dx = 0;
    pf_model_get_real(m, "time", &t);
    pf_model_get_real(m, "dx", &dx);
    std::snprintf(name, sizeof name, "/field%07d_%d.vti", int(t * 100), (*vti_counter)++);
    std::vector<double> eff(hs.conc.begin(), hs.conc.begin() + long(d.nx) * d.ny * d.nz);
    for (size_t n = 0; n < eff.size(); ++n) eff[n] -= hs.comp[n];
    rc = pf_write_vti((std::string(dir) + name).c_str(), d.nx, d.ny, d.nz, d.depth, dx,
                      hs.index.data(), eff.data());
  }
  return rc;
}

}  // namespace

int main(int argc, char **argv) {
  std::printf("phase-field step on gfx950: multi-phase field + solute diffusion (Steinbach model)\n");
  if (argc < 2) {
    std::printf("PROGRAM ARGUMENT IS NOT ENOUGH.\n");
    return 1;
  }
  pf_model *m = nullptr;
  if (pf_model_read(argv[1], &m) != PF_OK) return die("input", pf_model_last_error());
  pf_desc d{};
  pf_model_desc(m, &d);
  int64_t gpu_flag = 1, G = 0, n_iso = 0, steps = -1, seed_nuc = 0, seed_mn = 0, nsol = 1;
  pf_model_get_int(m, "gpu_flag", &gpu_flag);
  if (!gpu_flag) std::printf("GPU_Flag is 0 in the input; pfm always runs on the GPU\n");
  pf_model_get_int(m, "num_grain1", &G);
  pf_model_get_int(m, "num_init_iterations", &n_iso);
  pf_model_get_int(m, "steps", &steps);
  pf_model_get_int(m, "nucleation_seed", &seed_nuc);
  pf_model_get_int(m, "mn_seed", &seed_mn);
  pf_model_get_int(m, "solutes", &nsol);

  // setup always runs on an FP64 domain; a mixed run converts it afterwards
  pf_desc d64 = d;
  d64.precision = PF_FP64;
  pf_handle *h = nullptr;
  if (pf_create(&d64, &h) != PF_OK) return die("device", pf_last_error(nullptr));
  pf_params p{};
  pf_model_update_parameter(m, &p);  // Construct's Update_parameter (advances time by pfm_dt)
  if (pf_set_params(h, &p) != PF_OK) return die("params", pf_last_error(h));
  std::vector<double> seeds(3 * size_t(G));
  pf_model_voronoi_seeds(m, seeds.data(), int(G));
  double reduction = 0, ic[6] = {0, 0, 0, 0, 0, 0}, mn = 1.5;
  pf_model_get_real(m, "reduction", &reduction);
  for (int i = 0; i < 3; ++i) {
    char key[16];
    std::snprintf(key, sizeof key, "init_conc%d", i);
    pf_model_get_real(m, key, &ic[i]);
  }
  pf_model_get_real(m, "mn_conc", &mn);
  ic[3] = ic[4] = ic[5] = mn;
  if (pf_setup_construct(h, seeds.data(), int(G), reduction, int(n_iso), ic) != PF_OK)
    return die("construct", pf_last_error(h));
  double mn_noise = 0;
  pf_model_get_real(m, "mn_noise", &mn_noise);
  if (nsol > 1 && mn_noise > 0 &&
      pf_setup_solute_noise(h, 1, mn, mn_noise, uint64_t(seed_mn)) != PF_OK)
    return die("solute noise", pf_last_error(h));
  int vti = 0;
  if (write_data(m, h, d64, true, &vti) != PF_OK) return die("output", pf_last_error(h));
  double nn = 0, radius = 0, nconc[2] = {-1, -1};
  pf_model_get_real(m, "num_nuclei", &nn);
  pf_model_get_real(m, "nucleus_radius", &radius);
  pf_model_get_real(m, "nucleus_conc", &nconc[0]);
  pf_model_get_real(m, "mn_nucleus_conc", &nconc[1]);
  if (pf_setup_nucleate(h, int(nn), float(radius), nconc, uint64_t(seed_nuc)) != PF_OK)
    return die("nucleate", pf_last_error(h));
  if (d.precision == PF_MIXED) {
    pf_handle *hm = nullptr;
    if (pf_create(&d, &hm) != PF_OK) return die("device", pf_last_error(nullptr));
    if (pf_set_params(hm, &p) != PF_OK || pf_copy_state(hm, h) != PF_OK)
      return die("convert", pf_last_error(hm));
    pf_destroy(h);
    h = hm;
  }

  const clock_t c0 = std::clock();
  const auto w0 = std::chrono::steady_clock::now();
  int64_t done = 0;
  double t = 0, end_time = 0;
  pf_model_get_real(m, "end_time", &end_time);
  for (;;) {
    pf_model_get_real(m, "time", &t);
    if (steps >= 0 ? done >= steps : t >= end_time) break;
    pf_model_update_parameter(m, &p);
    if (pf_set_params(h, &p) != PF_OK || pf_step(h, 1) != PF_OK)
      return die("step", pf_last_error(h));
    ++done;
    pf_model_set_int(m, "step", done);
    if (write_data(m, h, d, false, &vti) != PF_OK) return die("output", pf_last_error(h));
  }
  if (pf_synchronize(h) != PF_OK) return die("step", pf_last_error(h));
  if (pf_dump_wait(h) != PF_OK) return die("output", pf_last_error(h));
  const double wall =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
  std::printf("SIMULATION IS FINISHED.\n\n");
  std::printf("Total CPU Running time(s)[physical_time*Num_threads]: %g\n",
              double(std::clock() - c0) / CLOCKS_PER_SEC);
  std::printf("Total Physical Time (s): %g\n", wall);
  const double cells = double(d.nx) * d.ny * d.nz;
  std::printf("steps %lld  cell-updates/s %.4g\n", (long long)done, done * cells / wall);
  char dump[4096];
  if (pf_model_get_string(m, "dump_state", dump, sizeof dump) == PF_OK && dump[0]) {
    HostState hs(d);
    if (pf_download_state(h, &hs.view) != PF_OK ||
        pf_write_state(dump, &d, t, done, &hs.view) != PF_OK)
      return die("dump", pf_last_error(h));
  }
  pf_destroy(h);
  pf_model_free(m);
  return 0;
}
