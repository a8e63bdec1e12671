// host_model.cpp — driver surface on the host: the positional input file, the thermodynamics
// file, the temperature path, the per-step parameter update and the kinetics / field writers.
//
// Reference behaviour followed (zhubq/PhaseField, C++-openmp-cuda/src):
//   phase_field::Read_data          phase_field.cpp:31-231   (positional, prompt lines skipped)
//   thermo_functor::Read_data       thermo_functor.cpp:11-53 (3 entropy, 6 NPLE1, 6 NPLE2, 6 PARA lines)
//   thermo_data::Get_data           thermo_data.cpp:24-40
//   phase_field::Temperature_function  phase_field.cpp:232-242 (clamped outside the path)
//   phase_field::Update_parameter   phase_field.cpp:307-315
//   phase_field::Write_data / Vti_write  write_data.cpp:3-82
// Extensions (not in the reference's input dialect) are optional `key=value` lines after the
// GPU flag; without them the file is read exactly as the reference reads it.

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/phasefield_amd.h"

namespace {

constexpr double kGasR = 8.314;  // constant.h:8

thread_local std::string g_model_error;

int model_fail(int code, const std::string &msg) {
  g_model_error = msg;
  return code;
}

// a*T + b  (linear_functor.h) and a*T*T + b*T + c (quadratic_functor.h)
struct Line {
  double a = 0.0, b = 0.0;
  double at(double T) const { return a * T + b; }
};
struct Quad {
  double a = 0.0, b = 0.0, c = 0.0;
  double at(double x) const { return a * x * x + b * x + c; }
};

}  // namespace

struct pf_model {
  // positional main-file fields
  std::string directory, thermo_file, main_path;
  int ix = 1, iy = 1, iz = 1, i_intf = 1;
  double dx = 1.0, cutoff = 1e-5;
  int num_grain[3] = {0, 0, 0};
  double phase_fraction[3] = {0, 0, 0};
  double init_conc[3] = {0, 0, 0};
  std::vector<double> t_time, t_temp;
  double reduction = 0.0;
  int num_init_iterations = 0;
  double timestep_output = 1.0;
  bool is_micro_out = false;
  double num_nuclei = 0.0;
  double mobility0[3][3] = {{0}}, mobility_q[3][3] = {{0}};
  double intf_energy[3][3] = {{0}};
  double diff0[3] = {0, 0, 0}, diff_q[3] = {0, 0, 0};
  int gpu_flag = 0;
  // thermodynamics file
  Line entropy[3][3];
  Quad nple1[3][3], nple2[3][3], para_eq[3][3];
  Line para_slope[3][3];
  // extensions
  std::map<std::string, double> ext_real;
  std::map<std::string, std::string> ext_str;
  // time bookkeeping (phase_field members)
  double time = 0.0, temperature = 0.0, end_time = 0.0, pfm_dt = 0.0, diff_dt = 0.0;
  double last_temperature = 0.0;  // Write_data's static
  int micro_count = 0;            // Write_data's static
  int64_t step = 0;

  int dims() const { return 3 - (iz == 1 ? 1 : 0) - (iy == 1 ? 1 : 0); }
  double ext(const std::string &k, double dflt) const {
    auto it = ext_real.find(k);
    return it == ext_real.end() ? dflt : it->second;
  }
  std::string exts(const std::string &k, const std::string &dflt) const {
    auto it = ext_str.find(k);
    return it == ext_str.end() ? dflt : it->second;
  }
};

namespace {

// istream reader with the reference's skip-a-line primitive (ignore(1000,'\n'))
struct Reader {
  std::istream &in;
  explicit Reader(std::istream &s) : in(s) {}
  void skip(int n = 1) {
    for (int i = 0; i < n; ++i) in.ignore(1000, '\n');
  }
  template <class T> T get() {
    T v{};
    in >> v;
    return v;
  }
  std::string line() {
    std::string s;
    std::getline(in, s, '\n');
    return s;
  }
};

std::string dir_of(const std::string &p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

int read_thermo(pf_model *m, const std::string &path) {
  std::ifstream f(path.c_str());
  if (!f) return model_fail(PF_ERR_IO, "cannot open thermodynamics file " + path);
  Reader r(f);
  r.skip(3);
  for (int k = 0; k < 3; ++k) {  // (max_num_phase-1)*max_num_phase/2 entropy pairs
    int i = r.get<int>(), j = r.get<int>();
    double a = r.get<double>(), b = r.get<double>();
    if (i < 1 || i > 3 || j < 1 || j > 3) return model_fail(PF_ERR_IO, "bad entropy line");
    m->entropy[i - 1][j - 1] = {a, b};
    m->entropy[j - 1][i - 1] = {-a, -b};
  }
  r.skip(4);
  Quad(*tables[3])[3] = {m->nple1, m->nple2, m->para_eq};
  for (int t = 0; t < 3; ++t) {
    if (t > 0) r.skip(2);
    for (int k = 0; k < 6; ++k) {
      int i = r.get<int>(), j = r.get<int>();
      double a = r.get<double>(), b = r.get<double>(), c = r.get<double>();
      if (!f || i < 1 || i > 3 || j < 1 || j > 3)
        return model_fail(PF_ERR_IO, "bad phase-diagram line in " + path);
      tables[t][i - 1][j - 1] = {a, b, c};
      if (t == 2) m->para_slope[i - 1][j - 1] = {2 * a, b};
    }
  }
  return PF_OK;
}

void parse_extensions(pf_model *m, std::istream &in) {
  std::string ln;
  while (std::getline(in, ln)) {
    size_t hash = ln.find('#');
    if (hash != std::string::npos) ln = ln.substr(0, hash);
    std::istringstream ls(ln);
    std::string tok;
    while (ls >> tok) {
      size_t eq = tok.find('=');
      if (eq == std::string::npos || eq == 0) continue;
      std::string key = tok.substr(0, eq), v = tok.substr(eq + 1);
      m->ext_str[key] = v;
      char *end = nullptr;
      double d = std::strtod(v.c_str(), &end);
      if (end && *end == '\0' && !v.empty()) m->ext_real[key] = d;
    }
  }
}

}  // namespace

extern "C" {

const char *pf_model_last_error(void) { return g_model_error.c_str(); }

// phase_field::Read_data, phase_field.cpp:31-231
int pf_model_read(const char *path, pf_model **out) {
  if (!path || !out) return model_fail(PF_ERR_INVALID, "null argument");
  std::ifstream f(path);
  if (!f) return model_fail(PF_ERR_IO, std::string("cannot open input file ") + path);
  pf_model *m = new pf_model();
  m->main_path = path;
  Reader r(f);
  r.skip();
  m->directory = r.line();
  r.skip(3);
  m->ix = r.get<int>();
  m->iy = r.get<int>();
  m->iz = r.get<int>();
  r.skip(2);
  m->dx = r.get<double>();
  r.skip(2);
  m->i_intf = r.get<int>();  // "5.00" reads as 5, the rest of the line is skipped
  r.skip(2);
  m->cutoff = r.get<double>();
  r.skip(2);
  for (int i = 0; i < 3; ++i) m->num_grain[i] = r.get<int>();
  r.skip(2);
  for (int i = 0; i < 3; ++i) m->phase_fraction[i] = r.get<double>();
  r.skip(2);
  for (int i = 0; i < 3; ++i) m->init_conc[i] = r.get<double>();
  r.skip(2);
  int nT = r.get<int>();
  if (!f || nT < 1 || nT > 100000) {
    delete m;
    return model_fail(PF_ERR_IO, "bad temperature path in input file");
  }
  for (int i = 0; i < nT; ++i) {
    m->t_time.push_back(r.get<double>());
    m->t_temp.push_back(r.get<double>());
  }
  m->time = m->t_time.front();
  m->temperature = m->t_temp.front();
  m->end_time = m->t_time.back();
  r.skip(2);
  m->reduction = r.get<double>();
  r.skip(2);
  m->num_init_iterations = r.get<int>();
  r.skip(2);
  m->timestep_output = r.get<double>();
  r.skip(2);
  m->is_micro_out = r.get<bool>();
  r.skip();
  r.skip();
  m->num_nuclei = r.get<double>();
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      r.skip(2);
      m->mobility0[i][j] = r.get<double>();
      m->mobility_q[i][j] = r.get<double>();
      m->mobility0[j][i] = m->mobility0[i][j];
      m->mobility_q[j][i] = m->mobility_q[i][j];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      r.skip(2);
      m->intf_energy[i][j] = r.get<double>();
      m->intf_energy[j][i] = m->intf_energy[i][j];
    }
  for (int i = 0; i < 3; ++i) {
    r.skip(2);
    m->diff0[i] = r.get<double>();
    m->diff_q[i] = r.get<double>();
  }
  r.skip(2);
  m->thermo_file = trim(r.line());
  r.skip();
  m->gpu_flag = r.get<int>();
  if (!f) {
    delete m;
    return model_fail(PF_ERR_IO, std::string("input file ended early: ") + path);
  }
  parse_extensions(m, f);
  if (m->ix < 1 || m->iy < 1 || m->iz < 1 || m->dx <= 0.0 || m->i_intf < 1) {
    delete m;
    return model_fail(PF_ERR_IO, "bad domain size, cell size or interface width");
  }
  std::string tp = m->thermo_file;
  if (!tp.empty() && tp[0] != '/') {
    std::ifstream probe(tp.c_str());
    if (!probe) tp = dir_of(m->main_path) + "/" + tp;
  }
  int rc = read_thermo(m, tp);
  if (rc != PF_OK) {
    delete m;
    return rc;
  }
  *out = m;
  return PF_OK;
}

int pf_model_free(pf_model *m) {
  delete m;
  return PF_OK;
}

int pf_model_get_int(const pf_model *m, const char *key, int64_t *out) {
  if (!m || !key || !out) return PF_ERR_INVALID;
  std::string k(key);
  if (k == "ix") *out = m->ix;
  else if (k == "iy") *out = m->iy;
  else if (k == "iz") *out = m->iz;
  else if (k == "i_intf") *out = m->i_intf;
  else if (k == "dims") *out = m->dims();
  else if (k == "num_grain0") *out = m->num_grain[0];
  else if (k == "num_grain1") *out = m->num_grain[1];
  else if (k == "num_grain2") *out = m->num_grain[2];
  else if (k == "num_init_iterations") *out = m->num_init_iterations;
  else if (k == "is_micro_out") *out = m->is_micro_out ? 1 : 0;
  else if (k == "num_nuclei") *out = int64_t(m->num_nuclei);
  else if (k == "gpu_flag") *out = m->gpu_flag;
  else if (k == "n_temperature_points") *out = int64_t(m->t_time.size());
  else if (k == "step") *out = m->step;
  else if (k == "steps") *out = int64_t(m->ext("steps", -1));
  else if (k == "seed") *out = int64_t(m->ext("seed", m->num_grain[1]));
  else if (k == "nucleation_seed") *out = int64_t(m->ext("nucleation_seed", 12345));
  else if (k == "mn_seed") *out = int64_t(m->ext("mn_seed", 54321));
  else if (k == "solutes") *out = int64_t(m->ext("solutes", 1));
  else return model_fail(PF_ERR_INVALID, "unknown integer key " + k);
  return PF_OK;
}

int pf_model_get_real(const pf_model *m, const char *key, double *out) {
  if (!m || !key || !out) return PF_ERR_INVALID;
  std::string k(key);
  if (k == "dx") *out = m->dx;
  else if (k == "cutoff") *out = m->cutoff;
  else if (k == "eta") *out = m->dx * m->i_intf;
  else if (k == "reduction") *out = m->reduction;
  else if (k == "timestep_output") *out = m->timestep_output;
  else if (k == "time") *out = m->time;
  else if (k == "temperature") *out = m->temperature;
  else if (k == "end_time") *out = m->end_time;
  else if (k == "pfm_dt") *out = m->pfm_dt;
  else if (k == "diff_dt") *out = m->diff_dt;
  else if (k == "num_nuclei") *out = m->num_nuclei;
  else if (k.rfind("init_conc", 0) == 0 && k.size() == 10) *out = m->init_conc[k[9] - '0'];
  else if (k.rfind("phase_fraction", 0) == 0 && k.size() == 15)
    *out = m->phase_fraction[k[14] - '0'];
  else if (k.rfind("mobility0_", 0) == 0 && k.size() == 12)
    *out = m->mobility0[k[10] - '1'][k[11] - '1'];
  else if (k.rfind("mobilityQ_", 0) == 0 && k.size() == 12)
    *out = m->mobility_q[k[10] - '1'][k[11] - '1'];
  else if (k.rfind("sigma_", 0) == 0 && k.size() == 8)
    *out = m->intf_energy[k[6] - '1'][k[7] - '1'];
  else if (k.rfind("diff0_", 0) == 0 && k.size() == 7) *out = m->diff0[k[6] - '1'];
  else if (k.rfind("diffQ_", 0) == 0 && k.size() == 7) *out = m->diff_q[k[6] - '1'];
  else if (k == "temperature_t0") *out = m->t_temp.front();
  else if (k == "nucleus_radius") *out = m->ext("nucleus_radius", double(m->i_intf));
  else if (k == "nucleus_conc") *out = m->ext("nucleus_conc", -1.0);
  else if (k == "mn_conc") *out = m->ext("mn_conc", 1.5);
  else if (k == "mn_noise") *out = m->ext("mn_noise", 0.0);
  else if (k == "mn_nucleus_conc") *out = m->ext("mn_nucleus_conc", -1.0);
  else if (m->ext_real.count(k)) *out = m->ext_real.at(k);
  else return model_fail(PF_ERR_INVALID, "unknown real key " + k);
  return PF_OK;
}

int pf_model_get_string(const pf_model *m, const char *key, char *buf, size_t len) {
  if (!m || !key || !buf || len == 0) return PF_ERR_INVALID;
  std::string k(key), v;
  if (k == "directory") v = m->directory;
  else if (k == "thermo_file") v = m->thermo_file;
  else if (m->ext_str.count(k)) v = m->ext_str.at(k);
  else return model_fail(PF_ERR_INVALID, "unknown string key " + k);
  std::snprintf(buf, len, "%s", v.c_str());
  return PF_OK;
}

int pf_model_set_real(pf_model *m, const char *key, double value) {
  if (!m || !key) return PF_ERR_INVALID;
  std::string k(key);
  if (k == "time") m->time = value;
  else if (k == "temperature") m->temperature = value;
  else if (k == "end_time") m->end_time = value;
  else m->ext_real[k] = value;
  return PF_OK;
}

int pf_model_set_int(pf_model *m, const char *key, int64_t value) {
  if (!m || !key) return PF_ERR_INVALID;
  std::string k(key);
  if (k == "step") m->step = value;
  else m->ext_real[k] = double(value);
  return PF_OK;
}

int pf_model_desc(const pf_model *m, pf_desc *d) {
  if (!m || !d) return PF_ERR_INVALID;
  d->nx = m->ix;
  d->ny = m->iy;
  d->nz = m->iz;
  d->depth = int(m->ext("depth", 6));  // max_num_coexist = 6 in the reference
  d->nphase = int(m->ext("nphase", 2));
  d->nsolute = int(m->ext("solutes", 1));
  std::string prec = m->exts("precision", "fp64");
  d->precision = (prec == "mixed" || prec == "fp32") ? PF_MIXED : PF_FP64;
  std::string bc = m->exts("bc_z", "periodic");
  d->bc_z = bc == "noflux" ? PF_BC_NOFLUX : PF_BC_PERIODIC;
  d->device = int(m->ext("device", 0));
  return PF_OK;
}

// Temperature_function, phase_field.cpp:232-242: piecewise-linear in time; outside the path the
// temperature is held at the nearest end point (the reference leaves it uninitialised there)
double pf_model_temperature(const pf_model *m, double t) {
  const size_t n = m->t_time.size();
  if (n == 1 || t < m->t_time.front()) return m->t_temp.front();
  for (size_t i = 0; i + 1 < n; ++i)
    if (t >= m->t_time[i] && t < m->t_time[i + 1])
      return (m->t_temp[i + 1] - m->t_temp[i]) / (m->t_time[i + 1] - m->t_time[i]) *
                 (t - m->t_time[i]) +
             m->t_temp[i];
  return m->t_temp.back();
}

// Update_parameter, phase_field.cpp:307-315, thermo_data::Get_data, thermo_data.cpp:24-40
int pf_model_update_parameter(pf_model *m, pf_params *p) {
  if (!m || !p) return PF_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  const double T = pf_model_temperature(m, m->time);
  m->temperature = T;
  const double dims = double(m->dims());
  double mmax = -1e308, dmax = -1e308;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      p->mobility[i][j] = m->mobility0[i][j] * std::exp(-m->mobility_q[i][j] / (kGasR * T));
      p->sigma[i][j] = m->intf_energy[i][j];
      mmax = std::max(mmax, p->mobility[i][j] * p->sigma[i][j]);
    }
  for (int i = 0; i < 3; ++i) {
    p->diffusivity[0][i] = m->diff0[i] * std::exp(-m->diff_q[i] / (kGasR * T));
    dmax = std::max(dmax, p->diffusivity[0][i]);
  }
  m->pfm_dt = m->dx * m->dx / dims * 0.2 / mmax;
  m->diff_dt = m->dx * m->dx / dmax / dims * 0.2;
  // thermodynamic tables
  double slope[3][3];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      p->entropy[i][j] = m->entropy[i][j].at(T);
      p->eq_conc[i][j] = m->para_eq[i][j].at(T);
      slope[i][j] = m->para_slope[i][j].at(T);
    }
  for (int i = 0; i < 3; ++i) slope[i][i] = 1.;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      p->eq_slope[i][j] = slope[i][j];
      p->part_k[0][i][j] = slope[j][i] / slope[i][j];
      p->part_c[0][i][j] = p->eq_conc[j][i] - p->part_k[0][i][j] * p->eq_conc[i][j];
    }
  // manganese (extension): fixed partition ratio k = c_gamma/c_alpha, Arrhenius diffusivity
  const double kmn = m->ext("mn_k", 2.0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double k = 1.0;
      if (i == 0 && j == 1) k = 1.0 / kmn;  // c_alpha = c_gamma / k
      if (i == 1 && j == 0) k = kmn;        // c_gamma = k * c_alpha
      p->part_k[1][i][j] = k;
      p->part_c[1][i][j] = 0.0;
    }
  for (int i = 0; i < 3; ++i) {
    char k0[16], kq[16];
    std::snprintf(k0, sizeof k0, "mn_D0_%d", i + 1);
    std::snprintf(kq, sizeof kq, "mn_Q_%d", i + 1);
    p->diffusivity[1][i] = m->ext(k0, 0.0) * std::exp(-m->ext(kq, 0.0) / (kGasR * T));
  }
  p->dx = m->dx;
  p->eta = m->dx * m->i_intf;
  p->cutoff = m->cutoff;
  p->dims = m->dims();
  p->pfm_dt = m->pfm_dt;
  // phase_field.cpp:247-248
  p->n_sub[0] = std::max(int(std::ceil(m->pfm_dt / m->diff_dt)), 1);
  p->dt_d[0] = m->pfm_dt / p->n_sub[0];
  const int mn_sub = int(m->ext("mn_substeps", 1));
  p->n_sub[1] = m->ext("solutes", 1) > 1 ? std::max(mn_sub, 1) : 0;
  p->dt_d[1] = p->n_sub[1] > 0 ? m->pfm_dt / p->n_sub[1] : 0.0;
  m->time += m->pfm_dt;
  return PF_OK;
}

// Tessellation's seed draw (MicrostructureConstruction.cpp:69-85): 3*G closed-uniform numbers
// in column-major (x1,y1,z1,x2,...) order, scaled to [1, n_d].  The generator is std::mt19937
// seeded with the `seed` extension (default: the grain count, as the reference seeds with it).
int pf_model_voronoi_seeds(const pf_model *m, double *seeds, int G) {
  if (!m || !seeds || G < 1) return PF_ERR_INVALID;
  std::mt19937 gen(uint32_t(m->ext("seed", m->num_grain[1])));
  for (int n = 0; n < 3 * G; ++n) seeds[n] = double(gen()) / 4294967295.0;
  const int ext[3] = {m->ix, m->iy, m->iz};
  for (int n = 0; n < G; ++n)
    for (int d = 0; d < 3; ++d) {
      seeds[3 * n + d] *= double(ext[d] - 1);
      seeds[3 * n + d] += 1.0;
    }
  return PF_OK;
}

// write_data.cpp:7-8
int pf_model_output_due(const pf_model *m, int forced, int *due) {
  if (!m || !due) return PF_ERR_INVALID;
  const double t = m->time, dto = m->timestep_output;
  *due = (forced || std::fabs(t - dto * std::floor(t / dto + 0.5)) < 0.5 * m->pfm_dt ||
          std::fabs(m->temperature - m->last_temperature) > 10.)
             ? 1
             : 0;
  return PF_OK;
}

// Write_data, write_data.cpp:3-40: a kinetics row when forced, at the output cadence or after a
// 10 K temperature change; mkdir of the output directory on first use
int pf_model_write_kinetics(pf_model *m, const pf_stats *st, double f_rex, int forced,
                            int *written) {
  if (!m || !st) return PF_ERR_INVALID;
  if (written) *written = 0;
  const double t = m->time;
  int due = 0;
  pf_model_output_due(m, forced, &due);
  if (!due) return PF_OK;
  ::mkdir(m->directory.c_str(), 0755);
  std::ofstream k((m->directory + "/fraction_time.txt").c_str(), std::ios::app);
  if (!k) return model_fail(PF_ERR_IO, "cannot write " + m->directory + "/fraction_time.txt");
  k << std::setw(8) << std::setprecision(4) << t << " " << m->temperature << " ";
  for (int i = 0; i < 3; ++i) k << std::setprecision(4) << float(st->phase_frac[i]) << " ";
  k << f_rex << std::endl;
  std::cout << "time:" << t << " dt: " << m->pfm_dt << " " << m->diff_dt << std::endl;
  m->last_temperature = m->temperature;
  if (written) *written = 1;
  return PF_OK;
}

// Vti_write, write_data.cpp:42-82: `korn` = 0 where slot 2 is occupied, else the phase of
// slot 1; `conc` = total concentration with 3 significant digits
int pf_write_vti(const char *path, int nx, int ny, int nz, int K, double dx, const int32_t *index,
                 const double *conc) {
  std::ofstream f(path);
  if (!f) return model_fail(PF_ERR_IO, std::string("cannot write ") + path);
  f << "<?xml version=\"1.0\"?>" << '\n'
    << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">"
    << "\n"
    << "<ImageData WholeExtent=\"" << 1 << ' ' << nx << ' ' << 1 << ' ' << ny << ' ' << 1 << ' '
    << nz << '"' << " Origin=" << "\"0 0 0\" Spacing=\"" << dx << ' ' << dx << ' ' << dx << "\">"
    << std::endl;
  f << "<Piece Extent=\"" << 1 << ' ' << nx << ' ' << 1 << ' ' << ny << ' ' << 1 << ' ' << nz
    << "\">" << std::endl;
  f << "<PointData Scalars=\"energy\">" << std::endl;
  f << "<DataArray type=\"Int32\" Name=\"korn\" format=\"ascii\">" << std::endl;
  const long N = long(nx) * ny * nz;
  for (long n = 0; n < N; ++n) {
    const int32_t g1 = index[n * K], g2 = K > 1 ? index[n * K + 1] : 0;
    int ph = g1 == 0 ? 0 : (g1 / 100000 > 0 ? g1 / 100000 : 1);
    f << (g2 > 0 ? 0 : ph) << " ";
  }
  f << "\n</DataArray>" << std::endl;
  f << "<DataArray type=\"Float32\" Name=\"conc\" format=\"ascii\">" << std::endl;
  for (long n = 0; n < N; ++n) f << std::setprecision(3) << conc[n] << " ";
  f << "</DataArray>" << std::endl
    << "</PointData>" << std::endl
    << "</Piece>" << std::endl
    << "</ImageData>" << std::endl
    << "</VTKFile>" << std::endl;
  return f ? PF_OK : model_fail(PF_ERR_IO, std::string("write failed: ") + path);
}

// The same two fields as a binary VTK ImageData file (raw appended data, UInt64 block headers):
// korn as Int32 and the concentration as Float32, the form ParaView reads without parsing text.
// No counterpart in the reference; at 1024^3 its ASCII file is ~10 GB of text.
int pf_write_vti_raw(const char *path, int nx, int ny, int nz, double dx, const int32_t *korn,
                     const float *conc) {
  if (!path || !korn || !conc || nx <= 0 || ny <= 0 || nz <= 0)
    return model_fail(PF_ERR_INVALID, "bad VTI arguments");
  FILE *f = std::fopen(path, "wb");
  if (!f) return model_fail(PF_ERR_IO, std::string("cannot write ") + path);
  const uint64_t N = uint64_t(nx) * ny * nz, bk = N * 4, bc = N * 4;
  char head[1024];
  const int n = std::snprintf(
      head, sizeof head,
      "<?xml version=\"1.0\"?>\n"
      "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" "
      "header_type=\"UInt64\">\n"
      "<ImageData WholeExtent=\"1 %d 1 %d 1 %d\" Origin=\"0 0 0\" Spacing=\"%.17g %.17g %.17g\">\n"
      "<Piece Extent=\"1 %d 1 %d 1 %d\">\n"
      "<PointData Scalars=\"energy\">\n"
      "<DataArray type=\"Int32\" Name=\"korn\" format=\"appended\" offset=\"0\"/>\n"
      "<DataArray type=\"Float32\" Name=\"conc\" format=\"appended\" offset=\"%llu\"/>\n"
      "</PointData>\n</Piece>\n</ImageData>\n<AppendedData encoding=\"raw\">\n_",
      nx, ny, nz, dx, dx, dx, nx, ny, nz, (unsigned long long)(8 + bk));
  bool ok = n > 0 && size_t(n) < sizeof head && std::fwrite(head, 1, size_t(n), f) == size_t(n);
  ok = ok && std::fwrite(&bk, 8, 1, f) == 1 && std::fwrite(korn, 4, N, f) == N;
  ok = ok && std::fwrite(&bc, 8, 1, f) == 1 && std::fwrite(conc, 4, N, f) == N;
  static const char tail[] = "\n</AppendedData>\n</VTKFile>\n";
  ok = ok && std::fwrite(tail, 1, sizeof tail - 1, f) == sizeof tail - 1;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? PF_OK : model_fail(PF_ERR_IO, std::string("write failed: ") + path);
}

// Full-precision binary state dump (no counterpart in the reference, whose ASCII VTI keeps
// 3 significant digits).  Header: magic, version, shape, K, S, time, step.
static const char kMagic[8] = {'P', 'F', 'A', 'M', 'D', 'S', 'T', '1'};

int pf_write_state(const char *path, const pf_desc *d, double time, int64_t step,
                   const pf_host_state *s) {
  if (!path || !d || !s || !s->index || !s->value || !s->phase || !s->conc)
    return model_fail(PF_ERR_INVALID, "null argument");
  FILE *f = std::fopen(path, "wb");
  if (!f) return model_fail(PF_ERR_IO, std::string("cannot write ") + path);
  const int64_t N = int64_t(d->nx) * d->ny * d->nz;
  int32_t hdr[6] = {d->nx, d->ny, d->nz, d->depth, d->nsolute, s->conc_comp ? 1 : 0};
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 6, f) == 6 &&
            std::fwrite(&time, 8, 1, f) == 1 && std::fwrite(&step, 8, 1, f) == 1;
  ok = ok && std::fwrite(s->index, 4, size_t(N * d->depth), f) == size_t(N * d->depth);
  ok = ok && std::fwrite(s->value, 8, size_t(N * d->depth), f) == size_t(N * d->depth);
  ok = ok && std::fwrite(s->phase, 8, size_t(N * 3), f) == size_t(N * 3);
  ok = ok && std::fwrite(s->conc, 8, size_t(N * d->nsolute), f) == size_t(N * d->nsolute);
  if (s->conc_comp)
    ok = ok &&
         std::fwrite(s->conc_comp, 8, size_t(N * d->nsolute), f) == size_t(N * d->nsolute);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? PF_OK : model_fail(PF_ERR_IO, std::string("write failed: ") + path);
}

int pf_read_state(const char *path, const pf_desc *d, double *time, int64_t *step,
                  pf_host_state *s) {
  if (!path || !d || !s || !s->index || !s->value || !s->phase || !s->conc)
    return model_fail(PF_ERR_INVALID, "null argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) return model_fail(PF_ERR_IO, std::string("cannot read ") + path);
  char mg[8];
  int32_t hdr[6];
  double t = 0;
  int64_t st = 0;
  bool ok = std::fread(mg, 1, 8, f) == 8 && std::memcmp(mg, kMagic, 8) == 0 &&
            std::fread(hdr, 4, 6, f) == 6 && std::fread(&t, 8, 1, f) == 1 &&
            std::fread(&st, 8, 1, f) == 1;
  ok = ok && hdr[0] == d->nx && hdr[1] == d->ny && hdr[2] == d->nz && hdr[3] == d->depth &&
       hdr[4] == d->nsolute;
  if (!ok) {
    std::fclose(f);
    return model_fail(PF_ERR_IO, std::string("state file header mismatch: ") + path);
  }
  const int64_t N = int64_t(d->nx) * d->ny * d->nz;
  ok = std::fread(s->index, 4, size_t(N * d->depth), f) == size_t(N * d->depth);
  ok = ok && std::fread(s->value, 8, size_t(N * d->depth), f) == size_t(N * d->depth);
  ok = ok && std::fread(s->phase, 8, size_t(N * 3), f) == size_t(N * 3);
  ok = ok && std::fread(s->conc, 8, size_t(N * d->nsolute), f) == size_t(N * d->nsolute);
  if (hdr[5] && s->conc_comp)
    ok = ok &&
         std::fread(s->conc_comp, 8, size_t(N * d->nsolute), f) == size_t(N * d->nsolute);
  else if (s->conc_comp)
    std::memset(s->conc_comp, 0, size_t(N * d->nsolute) * 8);
  std::fclose(f);
  if (!ok) return model_fail(PF_ERR_IO, std::string("state file truncated: ") + path);
  if (time) *time = t;
  if (step) *step = st;
  return PF_OK;
}

}  // extern "C"
