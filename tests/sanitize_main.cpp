// sanitize_main.cpp — a CPU driver that exercises the host model (phasefield_amd/csrc/host_model.cpp)
// and the oracle (oracle/pf_oracle.cpp) for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitizers.py compiles it with g++ and runs it).
//
//   sanitize_main <reference example input file> <scratch directory>
//
// Host model: parse the example input and its thermodynamics file, read and set fields, update the
// parameters along the time axis, draw Voronoi seeds, write and read back every file format, and
// take the error paths (missing and truncated input).  Oracle: seeded setup, full steps, sweeps and
// sub-steps in the FP64, mixed and dual-number forms, K = 4 and 6, periodic and no-flux z, 2D and
// 3D, reductions and state round trips.  Exit status 0 when every call behaved; the sanitizers abort
// the run on the first finding.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../include/phasefield_amd.h"
#include "../oracle/pf_oracle.h"

static int g_fail = 0;
#define CHECK(cond)                                                     \
  do {                                                                  \
    if (!(cond)) {                                                      \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
      ++g_fail;                                                         \
    }                                                                   \
  } while (0)

static void host_model(const char *input, const std::string &dir) {
  pf_model *m = nullptr;
  CHECK(pf_model_read(input, &m) == PF_OK);
  if (!m) return;
  int64_t ix = 0;
  double dx = 0;
  char buf[256];
  CHECK(pf_model_get_int(m, "ix", &ix) == PF_OK && ix > 0);
  CHECK(pf_model_get_real(m, "dx", &dx) == PF_OK && dx > 0);
  CHECK(pf_model_get_string(m, "directory", buf, sizeof(buf)) == PF_OK);
  CHECK(pf_model_get_string(m, "directory", buf, 2) == PF_OK && std::strlen(buf) < 2);
  CHECK(pf_model_get_real(m, "no_such_key", &dx) != PF_OK);
  CHECK(pf_model_set_real(m, "time", 0.0) == PF_OK);
  CHECK(pf_model_set_int(m, "steps", 3) == PF_OK);
  pf_desc d;
  CHECK(pf_model_desc(m, &d) == PF_OK);
  pf_params p;
  for (int k = 0; k < 5; ++k) CHECK(pf_model_update_parameter(m, &p) == PF_OK && p.pfm_dt > 0);
  for (double t : {-1.0, 0.0, 1.0, 5.0, 1e9}) CHECK(std::isfinite(pf_model_temperature(m, t)));
  std::vector<double> seeds(3 * 7);
  CHECK(pf_model_voronoi_seeds(m, seeds.data(), 7) == PF_OK);
  int due = 0;
  CHECK(pf_model_output_due(m, 1, &due) == PF_OK && due == 1);
  pf_model_free(m);

  // writers and readers on a small state
  const int nx = 5, ny = 4, nz = 3, K = 4, N = nx * ny * nz;
  std::vector<int32_t> idx(size_t(N) * K, 0), korn(N);
  std::vector<double> val(size_t(N) * K, 0.0), phase(size_t(N) * 3, 0.0), conc(N);
  std::vector<float> cf(N);
  for (int n = 0; n < N; ++n) {
    idx[size_t(n) * K] = n % 3 ? 200001 + n % 2 : 100001;
    val[size_t(n) * K] = 1.0;
    phase[size_t(n) * 3 + (n % 3 ? 1 : 0)] = 1.0;
    conc[n] = 0.02 + 1e-3 * n;
    korn[n] = n % 3 ? 2 : 1;
    cf[n] = float(conc[n]);
  }
  CHECK(pf_write_vti((dir + "/a.vti").c_str(), nx, ny, nz, K, 0.1, idx.data(), conc.data()) == PF_OK);
  CHECK(pf_write_vti_raw((dir + "/b.vti").c_str(), nx, ny, nz, 0.1, korn.data(), cf.data()) == PF_OK);
  pf_desc sd{nx, ny, nz, K, 2, 1, PF_FP64, PF_BC_PERIODIC, 0};
  pf_host_state hs{idx.data(), val.data(), phase.data(), conc.data(), nullptr, nullptr};
  CHECK(pf_write_state((dir + "/s.bin").c_str(), &sd, 1.5, 7, &hs) == PF_OK);
  std::vector<int32_t> idx2(idx.size());
  std::vector<double> val2(val.size()), phase2(phase.size()), conc2(conc.size());
  pf_host_state hs2{idx2.data(), val2.data(), phase2.data(), conc2.data(), nullptr, nullptr};
  double t = 0;
  int64_t step = 0;
  CHECK(pf_read_state((dir + "/s.bin").c_str(), &sd, &t, &step, &hs2) == PF_OK);
  CHECK(t == 1.5 && step == 7 && idx2 == idx && val2 == val && conc2 == conc);
  pf_desc other = sd;
  other.nx = 6;
  CHECK(pf_read_state((dir + "/s.bin").c_str(), &other, &t, &step, &hs2) != PF_OK);

  // error paths: missing file, an input cut short, an empty file
  pf_model *bad = nullptr;
  CHECK(pf_model_read((dir + "/missing.txt").c_str(), &bad) != PF_OK && !bad);
  {
    std::ifstream in(input);
    std::string txt((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    for (size_t cut : {size_t(0), txt.size() / 3, txt.size() / 2}) {
      std::ofstream out(dir + "/cut.txt");
      out << txt.substr(0, cut);
      out.close();
      CHECK(pf_model_read((dir + "/cut.txt").c_str(), &bad) != PF_OK && !bad);
      CHECK(std::strlen(pf_model_last_error()) > 0);
    }
  }
}

static orc_params params_from(const pf_params &p) {
  orc_params o;
  static_assert(sizeof(o) == sizeof(p), "layouts");
  std::memcpy(&o, &p, sizeof(o));
  return o;
}

static void oracle_runs(const char *input) {
  pf_model *m = nullptr;
  if (pf_model_read(input, &m) != PF_OK) {
    ++g_fail;
    return;
  }
  pf_params pp;
  pf_model_update_parameter(m, &pp);
  pf_model_free(m);
  struct Case {
    int nx, ny, nz, K, S, form, bc;
  };
  const Case cases[] = {{14, 11, 9, 4, 1, ORC_FP64, 0}, {13, 10, 1, 6, 1, ORC_FP64, 0},
                        {12, 10, 8, 4, 2, ORC_MIXED, 1}, {10, 9, 7, 4, 1, ORC_DUAL, 0},
                        {9, 8, 6, 6, 2, ORC_FP64, 1}};
  for (const Case &c : cases) {
    orc_params p = params_from(pp);
    p.dims = c.nz > 1 ? 3 : 2;
    p.n_sub[0] = 7;
    p.n_sub[1] = 1;
    p.dt_d[1] = p.pfm_dt;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        p.part_k[1][i][j] = p.part_k[0][i][j];
        p.part_c[1][i][j] = p.part_c[0][i][j];
      }
    p.diffusivity[1][0] = 1e-3;
    p.diffusivity[1][1] = 1e-2;
    const long N = long(c.nx) * c.ny * c.nz;
    void *h = orc_create(c.nx, c.ny, c.nz, c.K, c.S, c.form, c.bc);
    CHECK(h != nullptr);
    if (!h) continue;
    orc_set_params(h, &p);
    std::vector<double> seeds;
    for (int g = 0; g < 5; ++g)
      for (int d = 0; d < 3; ++d) seeds.push_back(1.0 + std::fmod(7.3 * (g + 1) * (d + 2), 8.0));
    std::vector<int32_t> grains(N);
    orc_tessellation(c.nx, c.ny, c.nz, seeds.data(), 5, 0.0, grains.data());
    const double ic[6] = {0.8, 0.02, 0.0, 1.5, 1.5, 1.5};
    orc_construct(h, grains.data(), 3, ic);
    if (c.S > 1) orc_solute_noise(h, 1, 1.5, 0.05, 11);
    const double nc[2] = {0.8, -1.0};
    orc_nucleate(h, 3, 2.0f, nc, 5);
    if (c.form == ORC_DUAL) {
      std::vector<double> dir(N, 1e-3);
      orc_dual_seed(h, dir.data());
    }
    orc_step(h, 2);
    orc_pf_sweep(h);
    orc_diffusion_substeps(h, 0, 3, p.dt_d[0]);
    orc_diffusion_substeps_factor(h, 0, 2, p.dt_d[0] / (p.dx * p.dx));
    double mass[2], pf[3], mx[2], mn[2], serr;
    int64_t nif;
    orc_reduce(h, mass, pf, mx, mn, &serr, &nif);
    CHECK(std::isfinite(mass[0]) && serr < 1e-6);
    std::vector<int32_t> idx(size_t(N) * c.K);
    std::vector<double> val(size_t(N) * c.K), ph(size_t(N) * 3), cc(size_t(N) * c.S),
        comp(size_t(N) * c.S), pc(size_t(N) * c.S * 3);
    orc_get_state(h, idx.data(), val.data(), ph.data(), cc.data(), comp.data(), pc.data());
    orc_set_state(h, idx.data(), val.data(), ph.data(), cc.data());
    orc_set_comp(h, comp.data());  // every form: the mixed one stores them, the others ignore them
    if (c.form == ORC_DUAL) {
      double v, d4[4];
      orc_dual_phase_fraction(h, &v, d4);
      std::vector<double> cv(N), dc(size_t(4) * N);
      orc_dual_conc(h, cv.data(), dc.data());
    }
    orc_destroy(h);
  }
  (void)orc_mt19937_nth(5489, 10000);
  for (uint32_t k = 0; k < 1000; ++k)
    (void)orc_quantize_comp(float(k) * 1e-12f - 5e-10f, 0.02f + float(k) * 1e-4f, orc_mix32(k));
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: sanitize_main <input file> <scratch dir>\n");
    return 2;
  }
  host_model(argv[1], argv[2]);
  oracle_runs(argv[1]);
  if (g_fail) std::fprintf(stderr, "%d checks failed\n", g_fail);
  std::printf("sanitize_main: %s\n", g_fail ? "FAILED" : "ok");
  return g_fail ? 1 : 0;
}
