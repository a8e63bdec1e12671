// pf_oracle.cpp — CPU oracle (TEST INFRASTRUCTURE ONLY; see pf_oracle.h).
//
// A plain C++/OpenMP restatement of the reference's step, written from the reference's
// behaviour (zhubq/PhaseField, paths relative to its root):
//   Phase_field_solver      C++-openmp-cuda/src/solver.cpp:147-304
//   Diffusion_solver        C++-openmp-cuda/src/solver.cpp:343-372
//   Driving_force           C++-openmp-cuda/src/solver.cpp:374-387
//   Partitioning            C++-openmp-cuda/src/solver.cpp:389-405
//   Sort_xyz                C++-openmp-cuda/src/solver.cpp:426-432
//   Get_neighbor / Phase_of_grain  C++-openmp-cuda/src/phase_field.h:25-29,58-68
//   Update (loop schedules) C++-openmp-cuda/src/phase_field.cpp:243-305
//   Update_parameter        C++-openmp-cuda/src/phase_field.cpp:307-315, thermo_data.cpp:24-40
//   Construct/Tessellation/Iso_growth  C++-openmp-cuda/src/MicrostructureConstruction.cpp:4-248
//   Nucleate1in2/distance   C++-openmp-cuda/src/nucleation.cpp:3-54
// Every expression keeps the reference's evaluation order (left to right, no FMA: build with
// -ffp-contract=off).  Blitz++ containers are replaced by flat vectors in the same column-major
// layout.  The reference itself cannot be compiled here (Blitz++ is not installed).
//
// Three number forms share the code: FP64 (the reference), MIXED (FP32 storage and arithmetic,
// FP64 for the clamp/prune decision and renormalisation, Kahan-compensated concentration update
// with the compensation held as a 4-bit fixed-point fraction of an ulp, rounded stochastically)
// and DUAL (FP64 with forward-mode derivatives, the reference gradient for the adjoint).

#include "pf_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <type_traits>
#include <utility>
#include <vector>

namespace {

constexpr double kPi = 3.1415926;  // constant.h:7 (truncated on purpose)
constexpr double kR = 8.314;       // constant.h:8
constexpr int kStencil = 10;       // max_num_nonzero_stencil, solver.cpp:148

// ------------------------------------------------------------------------------------------
// forward-mode dual number with 4 derivative components
struct Dual {
  double v;
  double d[4];
  Dual() : v(0.0) { d[0] = d[1] = d[2] = d[3] = 0.0; }
  Dual(double x) : v(x) { d[0] = d[1] = d[2] = d[3] = 0.0; }  // NOLINT
};
inline Dual operator+(const Dual &a, const Dual &b) {
  Dual r(a.v + b.v);
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
inline Dual operator-(const Dual &a, const Dual &b) {
  Dual r(a.v - b.v);
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
inline Dual operator*(const Dual &a, const Dual &b) {
  Dual r(a.v * b.v);
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
inline Dual operator/(const Dual &a, const Dual &b) {
  Dual r(a.v / b.v);
  for (int k = 0; k < 4; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
  return r;
}
inline Dual &operator+=(Dual &a, const Dual &b) { return a = a + b; }
inline Dual &operator-=(Dual &a, const Dual &b) { return a = a - b; }
inline Dual &operator*=(Dual &a, const Dual &b) { return a = a * b; }
inline bool operator>(const Dual &a, double b) { return a.v > b; }
inline bool operator<(const Dual &a, double b) { return a.v < b; }
inline bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
inline Dual sqrt(const Dual &a) {
  Dual r(std::sqrt(a.v));
  // derivative convention: sqrt'(0) := 0 (structural zeros of the phase products)
  for (int k = 0; k < 4; ++k) r.d[k] = a.v > 0.0 ? a.d[k] / (2.0 * r.v) : 0.0;
  return r;
}
inline double val(const Dual &a) { return a.v; }
inline double val(double a) { return a; }
inline double val(float a) { return a; }

template <class T> inline T cvt(double x) { return T(x); }
template <> inline float cvt<float>(double x) { return static_cast<float>(x); }

// conversions between the storage, arithmetic and decision types of a form
template <class To, class From> inline To as(const From &x) { return To(x); }
template <> inline float as<float, double>(const double &x) { return static_cast<float>(x); }
template <> inline double as<double, Dual>(const Dual &x) { return x.v; }

using std::sqrt;

// Mixed form: the Kahan compensation is kept as a 4-bit signed fixed-point multiple of 2^-3 ulp of
// the stored value (comp = q * 2^(e - 27), value = m * 2^e with 0.5 <= |m| < 1, q saturating at
// [-8, 7]), rounded stochastically without bias: X = the compensation in units of 2^-24 of that
// quantum (nearest integer, saturated to [-8, 7] quanta), U = the high 24 bits of a hash of the
// cell index and the sub-step (key), q = floor((X + U) / 2^24) in integers.  This returns the
// compensation that representation holds.
inline uint32_t mix32(uint32_t x) {  // xorshift-multiply integer hash ("lowbias32" form)
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
constexpr int kQuantumShift = 27;  // quantum 2^(e - 27)
inline float quantize_comp(float comp, float value, uint32_t key) {
  int e = 0;
  (void)std::frexp(value, &e);
  float s = std::ldexp(comp, kQuantumShift + 24 - e);
  s = std::fmin(std::fmax(s, -8.f * 16777216.f), 7.f * 16777216.f);
  const int64_t X = int64_t(std::nearbyint(s));
  const int64_t q = (X + int64_t(key >> 8)) >> 24;  // floor: arithmetic shift of a signed value
  return std::ldexp(float(q), e - kQuantumShift);
}
// the mean of quantize_comp over all 2^24 thresholds (exact in int64)
inline double quantize_comp_mean(float comp, float value) {
  int64_t sum = 0;
  int e = 0;
  (void)std::frexp(value, &e);
  for (uint32_t u = 0; u < (1u << 24); ++u)
    sum += int64_t(std::ldexp(quantize_comp(comp, value, u << 8), kQuantumShift - e));
  return std::ldexp(double(sum), e - kQuantumShift - 24);
}
template <class T> inline T quantize_comp(T comp, T, uint32_t) { return comp; }

template <class S, class A, class D, bool K> struct Form {
  using Store = S;  // phi, value_phase, concentrations
  using Acc = A;    // pair terms, partition, diffusion flux
  using Dec = D;    // clamp / prune / sort / renormalisation
  static constexpr bool kahan = K;
};
using FormF64 = Form<double, double, double, false>;
using FormMixed = Form<float, float, double, true>;
using FormDual = Form<Dual, Dual, Dual, false>;

inline int phase_of_grain(int g) {  // phase_field.h:25-29
  if (g == 0) return 0;
  return g / 100000 > 0 ? g / 100000 : 1;
}

inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
// per-cell stream key for the counter-based setup noise
inline uint64_t noise_bits(uint64_t seed, uint64_t counter) {
  return splitmix64(splitmix64(seed) + counter);
}

struct DomainBase {
  virtual ~DomainBase() {}
  int nsolute = 1;  // solutes the domain carries
  virtual void set_params(const orc_params *p) = 0;
  virtual void set_state(const int32_t *, const double *, const double *, const double *) = 0;
  virtual void set_comp(const double *) = 0;
  virtual void get_state(int32_t *, double *, double *, double *, double *, double *) = 0;
  virtual void pf_sweep() = 0;
  virtual void diffuse(int s, int n, double factor) = 0;
  virtual void reduce(double *, double *, double *, double *, double *, int64_t *) = 0;
  virtual void construct(const int32_t *grain, int n_iso, const double *init_conc) = 0;
  virtual void nucleate(int n, float radius, const double *nuc_conc, uint64_t seed) = 0;
  virtual void solute_noise(int s, double mean, double sd, uint64_t seed) = 0;
  virtual void dual_seed(const double *dir) = 0;
  virtual void dual_phase_fraction(double *v, double *d) = 0;
  virtual void dual_conc(double *c, double *dc) = 0;
  orc_params p{};
};

template <class F> struct Domain : DomainBase {
  using S = typename F::Store;
  using A = typename F::Acc;
  using D = typename F::Dec;

  int nx, ny, nz, K, NS, bcz;
  long N;
  std::vector<int32_t> idx1, idx2;
  std::vector<S> val1, val2, vph;
  std::vector<S> conc[2], comp[2], vc1[2], vc2[2];
  uint32_t substeps[2] = {0, 0};  // sub-steps per solute since the state was set (rounding keys)
  // 1-based phase tables; row/column 0 belong to "no phase" (empty slot) and are kept finite
  A M[4][4], sig[4][4], ent[4][4], eqc[4][4], slope[4][4];
  S pk[2][4][4], pc[2][4][4];
  A dif[2][4];

  Domain(int x, int y, int z, int k, int ns, int bc)
      : nx(x), ny(y), nz(z), K(k), NS(ns), bcz(bc), N(long(x) * y * z) {
    nsolute = ns;
    idx1.assign(N * K, 0);
    idx2.assign(N * K, 0);
    val1.assign(N * K, S(0.0));
    val2.assign(N * K, S(0.0));
    vph.assign(N * 3, S(0.0));
    for (int s = 0; s < NS; ++s) {
      conc[s].assign(N, S(0.0));
      comp[s].assign(N, S(0.0));
      vc1[s].assign(N * 3, S(0.0));
      vc2[s].assign(N * 3, S(0.0));
    }
    std::memset(&p, 0, sizeof(p));
  }

  void set_params(const orc_params *q) override {
    p = *q;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        bool in = i > 0 && j > 0;
        M[i][j] = cvt<A>(in ? q->mobility[i - 1][j - 1] : 0.0);
        sig[i][j] = cvt<A>(in ? q->sigma[i - 1][j - 1] : 0.0);
        ent[i][j] = cvt<A>(in ? q->entropy[i - 1][j - 1] : 0.0);
        eqc[i][j] = cvt<A>(in ? q->eq_conc[i - 1][j - 1] : 0.0);
        slope[i][j] = cvt<A>(in ? q->eq_slope[i - 1][j - 1] : 1.0);
        for (int s = 0; s < 2; ++s) {
          pk[s][i][j] = cvt<S>(in ? q->part_k[s][i - 1][j - 1] : 0.0);
          pc[s][i][j] = cvt<S>(in ? q->part_c[s][i - 1][j - 1] : 0.0);
        }
      }
    for (int s = 0; s < 2; ++s) {
      dif[s][0] = cvt<A>(0.0);
      for (int i = 1; i < 4; ++i) dif[s][i] = cvt<A>(q->diffusivity[s][i - 1]);
    }
  }

  inline long cell(int i, int j, int k) const { return i + long(nx) * (j + long(ny) * k); }

  // neighbour cells in the order (-x,+x,-y,+y,-z,+z), Get_neighbor wrap (phase_field.h:58-68);
  // no-flux z mirrors the boundary cell onto itself
  inline void neighbours(int i, int j, int k, long nb[6]) const {
    int im = i - 1 < 0 ? i - 1 + nx : i - 1, ip = i + 1 >= nx ? i + 1 - nx : i + 1;
    int jm = j - 1 < 0 ? j - 1 + ny : j - 1, jp = j + 1 >= ny ? j + 1 - ny : j + 1;
    int km, kp;
    if (bcz == 1) {
      km = k - 1 < 0 ? k : k - 1;
      kp = k + 1 >= nz ? k : k + 1;
    } else {
      km = k - 1 < 0 ? k - 1 + nz : k - 1;
      kp = k + 1 >= nz ? k + 1 - nz : k + 1;
    }
    nb[0] = cell(im, j, k);
    nb[1] = cell(ip, j, k);
    nb[2] = cell(i, jm, k);
    nb[3] = cell(i, jp, k);
    nb[4] = cell(i, j, km);
    nb[5] = cell(i, j, kp);
  }

  // Partitioning, solver.cpp:389-405
  inline void partition(int s, S c, const S *phase, S *out) const {
    for (int i = 1; i <= 3; ++i) {
      S t1 = S(0.0), t2 = S(0.0);
      for (int j = 1; j <= 3; ++j) {
        t1 += phase[j - 1] * pk[s][i][j];
        t2 += phase[j - 1] * pc[s][i][j];
      }
      out[i - 1] = (c - t2) / t1;
    }
  }

  // Driving_force, solver.cpp:374-387
  inline A driving_force(int pa, int pb, A ca) const {
    A df = A(0.0);
    if (pa != pb) df += ent[pa][pb] * (ca - eqc[pa][pb]) / slope[pa][pb];
    return df;
  }

  void set_state(const int32_t *index, const double *value, const double *phase,
                 const double *c) override {
    for (long n = 0; n < N * K; ++n) {
      idx1[n] = index[n];
      val1[n] = cvt<S>(value[n]);
      idx2[n] = 0;
      val2[n] = S(0.0);
    }
    for (long n = 0; n < N * 3; ++n) vph[n] = cvt<S>(phase[n]);
    for (int s = 0; s < NS; ++s) {
      substeps[s] = 0;
      for (long n = 0; n < N; ++n) {
        conc[s][n] = cvt<S>(c[s * N + n]);
        comp[s][n] = S(0.0);
      }
    }
    repartition();
  }

  // the mixed form's stored compensations [S][N] (values on its code grid, as a download of a
  // mixed device state gives them); the FP64 and dual forms keep none
  void set_comp(const double *c) override {
    if (!std::is_same<S, float>::value) return;
    for (int s = 0; s < NS; ++s)
      for (long n = 0; n < N; ++n) comp[s][n] = cvt<S>(c[s * N + n]);
  }

  void repartition() {
    for (int s = 0; s < NS; ++s)
      for (long n = 0; n < N; ++n) partition(s, conc[s][n], &vph[n * 3], &vc1[s][n * 3]);
  }

  void get_state(int32_t *index, double *value, double *phase, double *c, double *cc,
                 double *pcv) override {
    if (index)
      for (long n = 0; n < N * K; ++n) index[n] = idx1[n];
    if (value)
      for (long n = 0; n < N * K; ++n) value[n] = val(val1[n]);
    if (phase)
      for (long n = 0; n < N * 3; ++n) phase[n] = val(vph[n]);
    for (int s = 0; s < NS; ++s)
      for (long n = 0; n < N; ++n) {
        if (c) c[s * N + n] = val(conc[s][n]);
        if (cc) cc[s * N + n] = val(comp[s][n]);
        if (pcv)
          for (int q = 0; q < 3; ++q) pcv[(s * N + n) * 3 + q] = val(vc1[s][n * 3 + q]);
      }
  }

  // ---------------------------------------------------------------------------------------
  // Phase_field_solver, solver.cpp:147-304
  void pf_cell(int i, int j, int k, const A f1, const A f2, const A f3, const double dt) {
    const long c = cell(i, j, k);
    const double cutoff = p.cutoff;
    int Lg[kStencil] = {0}, Lp[kStencil] = {0};
    S V1[kStencil], V2[kStencil];
    for (int e = 0; e < kStencil; ++e) V1[e] = V2[e] = S(0.0);
    int n = 0;
    for (int s = 0; s < K; ++s) {
      if (val1[c * K + s] > cutoff) {
        ++n;
        Lg[s] = idx1[c * K + s];
        Lp[s] = phase_of_grain(Lg[s]);
        V1[s] = val1[c * K + s];
      }
    }
    long nb[6];
    neighbours(i, j, k, nb);
    for (int f = 0; f < 2 * p.dims; ++f) {
      for (int s = 0; s < K; ++s) {
        const int g = idx1[nb[f] * K + s];
        if (g == 0) break;
        const S v = val1[nb[f] * K + s];
        int kk;
        for (kk = 0; kk < kStencil; ++kk) {
          if (g == Lg[kk]) {
            V2[kk] += v;
            break;
          }
        }
        if (kk == kStencil && n < kStencil) {
          Lg[n] = g;
          Lp[n] = phase_of_grain(g);
          V2[n] += v;
          ++n;
        }
      }
    }
    if (n == 1) {
      for (int s = 0; s < K; ++s) {
        idx2[c * K + s] = idx1[c * K + s];
        val2[c * K + s] = val1[c * K + s];
      }
      return;
    }
    A lc[4];
    lc[0] = A(0.0);
    for (int q = 1; q <= 3; ++q) lc[q] = A(vc1[0][c * 3 + q - 1]);
    D del[kStencil];
    for (int e = 0; e < kStencil; ++e) del[e] = D(0.0);
    for (int a = 0; a < n; ++a) {
      const int pa = Lp[a];
      for (int b = 0; b < n; ++b) {
        if (a == b) continue;
        const int pb = Lp[b];
        if (std::abs(pa - pb) == 2) continue;
        A t = (A(V2[a]) * A(V1[b]) - A(V2[b]) * A(V1[a])) * f1;
        t += (A(V1[a]) - A(V1[b])) * f2;
        t *= sig[pa][pb];
        t += f3 * sqrt(A(V1[a]) * A(V1[b])) * driving_force(pa, pb, lc[pa]);
        t *= M[pa][pb];
        del[a] += as<D>(t);
      }
      del[a] *= D(dt);
      del[a] += as<D>(V1[a]);
      if (del[a] > 1.0)
        del[a] = D(1.0);
      else if (del[a] < cutoff)
        del[a] = D(0.0);
    }
    // Sort_xyz: selection sort, descending, first maximum wins ties
    for (int a = 0; a < kStencil; ++a) {
      int m = a;
      for (int b = a + 1; b < kStencil; ++b)
        if (del[b] > del[m]) m = b;
      std::swap(del[a], del[m]);
      std::swap(Lg[a], Lg[m]);
    }
    D sum = D(0.0);
    D vp[4] = {D(0.0), D(0.0), D(0.0), D(0.0)};
    D out[PF_DEPTH_MAX];
    for (int s = 0; s < K; ++s) {
      if (del[s] > cutoff) {
        idx2[c * K + s] = Lg[s];
        out[s] = del[s];
        vp[phase_of_grain(Lg[s])] += del[s];
        sum += del[s];
      } else {
        idx2[c * K + s] = 0;
        out[s] = D(0.0);
      }
    }
    const D inv = D(1.0) / sum;
    for (int s = 0; s < K; ++s) val2[c * K + s] = as<S>(out[s] * inv);
    for (int q = 1; q <= 3; ++q) vph[c * 3 + q - 1] = as<S>(vp[q] * inv);
    for (int sidx = 0; sidx < NS; ++sidx)
      partition(sidx, conc[sidx][c], &vph[c * 3], &vc1[sidx][c * 3]);
  }
  static constexpr int PF_DEPTH_MAX = 6;

  void pf_sweep() override {
    const A f1 = cvt<A>(1. / (p.dx * p.dx));
    const A f2 = cvt<A>(kPi * kPi / (2. * p.eta * p.eta));
    const A f3 = cvt<A>(kPi / p.eta);
    const double dt = p.pfm_dt;
#pragma omp parallel for schedule(guided, 1) collapse(2)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) pf_cell(i, j, k, f1, f2, f3, dt);
    std::swap(idx1, idx2);
    std::swap(val1, val2);
  }

  // ---------------------------------------------------------------------------------------
  // Diffusion_solver, solver.cpp:343-372 (factor = diff_dt/(dx*dx))
  void diffuse(int s, int nsteps, double factor) override {
    const A r = cvt<A>(factor);
    for (int it = 0; it < nsteps; ++it) {
      // the sub-step's rounding key: mix32(k * 2 + s) for the k-th sub-step of solute s
      const uint32_t key = mix32(substeps[s]++ * 2u + uint32_t(s));
#pragma omp parallel for schedule(static) collapse(2)
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            const long c = cell(i, j, k);
            long nb[6];
            neighbours(i, j, k, nb);
            A del = A(0.0);
            if (F::kahan) {
              // Mixed form (no reference order to keep): per phase the face sum of
              // w * (c_nb - c), then the sum over phases of that times D_q * r, rounded once.
              // A bulk cell gives (sum_f (c_f - c)) * (D_p * r), the bulk kernels' form.
              A sq[4] = {A(0.0), A(0.0), A(0.0), A(0.0)};
              for (int f = 0; f < 2 * p.dims; ++f)
                for (int q = 1; q <= 3; ++q)
                  sq[q] += sqrt(A(vph[nb[f] * 3 + q - 1]) * A(vph[c * 3 + q - 1])) *
                           (A(vc1[s][nb[f] * 3 + q - 1]) - A(vc1[s][c * 3 + q - 1]));
              for (int q = 1; q <= 3; ++q) del += sq[q] * (dif[s][q] * r);
            } else {
              for (int f = 0; f < 2 * p.dims; ++f)
                for (int q = 1; q <= 3; ++q)
                  del += sqrt(A(vph[nb[f] * 3 + q - 1]) * A(vph[c * 3 + q - 1])) *
                         (A(vc1[s][nb[f] * 3 + q - 1]) - A(vc1[s][c * 3 + q - 1])) * dif[s][q];
              del *= r;
            }
            if (F::kahan) {
              const S y = S(del) - comp[s][c];
              const S t = conc[s][c] + y;
              // the cell's rounding hash: key ^ (c * 0x9E3779B9)
              comp[s][c] = quantize_comp((t - conc[s][c]) - y, t, key ^ (uint32_t(c) * 0x9E3779B9u));
              conc[s][c] = t;
            } else {
              conc[s][c] += S(del);
            }
            partition(s, conc[s][c], &vph[c * 3], &vc2[s][c * 3]);
          }
      std::swap(vc1[s], vc2[s]);
    }
  }

  void reduce(double *mass, double *pfrac, double *cmax, double *cmin, double *serr,
              int64_t *nif) override {
    for (int s = 0; s < NS; ++s) {
      double m = 0.0, mx = -1e300, mn = 1e300;
      for (long n = 0; n < N; ++n) {
        double v = val(conc[s][n]) - val(comp[s][n]);
        m += v;
        mx = std::max(mx, v);
        mn = std::min(mn, v);
      }
      if (mass) mass[s] = m;
      if (cmax) cmax[s] = mx;
      if (cmin) cmin[s] = mn;
    }
    if (pfrac)
      for (int q = 0; q < 3; ++q) {
        double a = 0.0;
        for (long n = 0; n < N; ++n) a += val(vph[n * 3 + q]);
        pfrac[q] = a / double(N);
      }
    double e = 0.0;
    int64_t ni = 0;
    for (long n = 0; n < N; ++n) {
      double sphi = 0.0;
      for (int s = 0; s < K; ++s) sphi += val(val1[n * K + s]);
      e = std::max(e, std::fabs(sphi - 1.0));
      if (K > 1 && idx1[n * K + 1] != 0) ++ni;
    }
    if (serr) *serr = e;
    if (nif) *nif = ni;
  }

  // ---------------------------------------------------------------------------------------
  // Iso_growth, MicrostructureConstruction.cpp:130-248.  Undefined behaviour of the reference
  // (uninitialised kk for neighbour entries <= cutoff) is defined as "skip the entry"; a
  // single-grain cell copies all K slots.
  void iso_cell(int i, int j, int k, double f1, double f2, double fac) {
    const long c = cell(i, j, k);
    const double cutoff = p.cutoff;
    int Lg[kStencil] = {0};
    double V1[kStencil] = {0}, V2[kStencil] = {0}, del[kStencil] = {0};
    int n = 0;
    for (int s = 0; s < K; ++s)
      if (val(val1[c * K + s]) > cutoff) {
        ++n;
        Lg[s] = idx1[c * K + s];
        V1[s] = val(val1[c * K + s]);
      }
    long nb[6];
    neighbours(i, j, k, nb);
    for (int f = 0; f < 2 * p.dims; ++f)
      for (int s = 0; s < K; ++s) {
        const int g = idx1[nb[f] * K + s];
        if (g == 0) break;
        const double v = val(val1[nb[f] * K + s]);
        if (!(v > cutoff)) continue;
        int kk;
        for (kk = 0; kk < kStencil; ++kk)
          if (g == Lg[kk]) {
            V2[kk] += v;
            break;
          }
        if (kk == kStencil && n < kStencil) {
          Lg[n] = g;
          V2[n] += v;
          ++n;
        }
      }
    if (n == 1) {
      for (int s = 0; s < K; ++s) {
        idx2[c * K + s] = idx1[c * K + s];
        val2[c * K + s] = val1[c * K + s];
      }
      return;
    }
    for (int a = 0; a < n; ++a) {
      for (int b = 0; b < n; ++b) {
        if (a == b) continue;
        double t = (V2[a] * V1[b] - V2[b] * V1[a]) * f1 + (V1[a] - V1[b]) * f2;
        del[a] += t;
      }
      del[a] *= fac;
      del[a] += V1[a];
      if (del[a] > 1.0)
        del[a] = 1.0;
      else if (del[a] < cutoff)
        del[a] = 0.;
    }
    for (int a = 0; a < kStencil; ++a) {
      int m = a;
      for (int b = a + 1; b < kStencil; ++b)
        if (del[b] > del[m]) m = b;
      std::swap(del[a], del[m]);
      std::swap(Lg[a], Lg[m]);
    }
    double sum = 0.0, vp[4] = {0, 0, 0, 0}, out[PF_DEPTH_MAX];
    for (int s = 0; s < K; ++s) {
      if (del[s] > cutoff) {
        idx2[c * K + s] = Lg[s];
        out[s] = del[s];
        vp[phase_of_grain(Lg[s])] += del[s];
        sum += del[s];
      } else {
        idx2[c * K + s] = 0;
        out[s] = 0.0;
      }
    }
    const double inv = 1. / sum;
    for (int s = 0; s < K; ++s) val2[c * K + s] = cvt<S>(out[s] * inv);
    for (int q = 1; q <= 3; ++q) vph[c * 3 + q - 1] = cvt<S>(vp[q] * inv);
  }

  void construct(const int32_t *grain, int n_iso, const double *init_conc) override {
    std::fill(idx1.begin(), idx1.end(), 0);
    std::fill(idx2.begin(), idx2.end(), 0);
    std::fill(val1.begin(), val1.end(), S(0.0));
    std::fill(val2.begin(), val2.end(), S(0.0));
    std::fill(vph.begin(), vph.end(), S(0.0));
    for (long n = 0; n < N; ++n) {
      idx1[n * K] = grain[n] + 200000;
      val1[n * K] = S(1.0);
      vph[n * 3 + phase_of_grain(idx1[n * K]) - 1] = S(1.0);
    }
    for (int s = 0; s < NS; ++s)
      for (long n = 0; n < N; ++n)
        for (int q = 0; q < 3; ++q) vc1[s][n * 3 + q] = cvt<S>(init_conc[s * 3 + q]);
    const double f1 = 1. / (p.dx * p.dx);
    const double f2 = kPi * kPi / (2. * p.eta * p.eta);
    const double fac = 0.5 / double(p.dims) * p.dx * p.dx;
    for (int it = 0; it < n_iso; ++it) {
#pragma omp parallel for collapse(2)
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) iso_cell(i, j, k, f1, f2, fac);
      std::swap(idx1, idx2);
      std::swap(val1, val2);
    }
    // total_conc = dot(value_phase, value_conc1); Partitioning (MicrostructureConstruction.cpp:54-62)
    for (int s = 0; s < NS; ++s) {
      substeps[s] = 0;
      for (long n = 0; n < N; ++n) {
        S t = S(0.0);
        for (int q = 0; q < 3; ++q) t += vph[n * 3 + q] * vc1[s][n * 3 + q];
        conc[s][n] = t;
        comp[s][n] = S(0.0);
      }
    }
    repartition();
  }

  // Nucleate1in2, nucleation.cpp:3-43, with u(cell) from the counter-based generator
  void nucleate(int nn, float radius, const double *nuc_conc, uint64_t seed) override {
    std::vector<float> flag(N);
    for (long n = 0; n < N; ++n) {
      int cnt = 0;
      for (int s = 0; s < K; ++s) cnt += idx1[n * K + s] > 0 ? 1 : 0;
      const float u = float(double(noise_bits(seed, uint64_t(n)) >> 40) * (1.0 / 16777216.0));
      const float c2 = float(val(vc1[0][n * 3 + 1]));
      flag[n] = float(double(float(cnt)) - double(c2) * (1. - 0.001 * double(u)));
    }
    std::vector<char> is_nuc(N, 0);
    const int ext[3] = {nx, ny, nz};
    for (int q = 1; q <= nn; ++q) {
      long m = 0;
      for (long n = 1; n < N; ++n)
        if (flag[n] > flag[m]) m = n;
      const int ci = int(m % nx), cj = int((m / nx) % ny), ck = int(m / (long(nx) * ny));
      const int ctr[3] = {ci, cj, ck};
      const int R = int(std::ceil(2.0f * radius));
      int lo[3], cnt[3];
      for (int d = 0; d < 3; ++d) {
        if (2 * R + 1 >= ext[d]) {
          lo[d] = 0;
          cnt[d] = ext[d];
        } else {
          lo[d] = ctr[d] - R;
          cnt[d] = 2 * R + 1;
        }
      }
      for (int dk = 0; dk < cnt[2]; ++dk)
        for (int dj = 0; dj < cnt[1]; ++dj)
          for (int di = 0; di < cnt[0]; ++di) {
            int pos[3] = {lo[0] + di, lo[1] + dj, lo[2] + dk};
            for (int d = 0; d < 3; ++d) pos[d] = ((pos[d] % ext[d]) + ext[d]) % ext[d];
            float dist = 0.f;
            for (int d = 0; d < 3; ++d) {
              const int dd = ctr[d] - pos[d];
              const int ad = dd < 0 ? -dd : dd;
              const float s1 = float(dd * dd), s2 = float((ext[d] - ad) * (ext[d] - ad));
              dist += std::min(s1, s2);
            }
            const float Dist = std::sqrt(dist);
            const long n = cell(pos[0], pos[1], pos[2]);
            if (Dist < radius) {
              for (int s = 0; s < K; ++s) {
                idx1[n * K + s] = 0;
                val1[n * K + s] = S(0.0);
              }
              idx1[n * K] = q + 100000;
              val1[n * K] = S(1.0);
              vph[n * 3 + 0] = S(1.0);
              vph[n * 3 + 1] = S(0.0);
              vph[n * 3 + 2] = S(0.0);
              flag[n] = 0.f;
              is_nuc[n] = 1;
            } else if (Dist < 2 * radius) {
              flag[n] = 0.f;
            }
          }
    }
    for (int s = 0; s < NS; ++s)
      if (nuc_conc && nuc_conc[s] >= 0.0)
        for (long n = 0; n < N; ++n)
          if (is_nuc[n]) {
            conc[s][n] = cvt<S>(nuc_conc[s]);
            comp[s][n] = S(0.0);
          }
    repartition();
  }

  void solute_noise(int s, double mean, double sd, uint64_t seed) override {
    for (long n = 0; n < N; ++n) {
      double acc = 0.0;
      for (int q = 0; q < 12; ++q)
        acc += double(noise_bits(seed, uint64_t(n) * 12 + q) >> 11) * (1.0 / 9007199254740992.0);
      conc[s][n] = cvt<S>(mean + sd * (acc - 6.0));
      comp[s][n] = S(0.0);
    }
    repartition();
  }

  void dual_seed(const double *) override {}
  void dual_phase_fraction(double *, double *) override {}
  void dual_conc(double *, double *) override {}
};

// dual-number specialisations of the derivative seeding / extraction
template <> void Domain<FormDual>::dual_seed(const double *dir) {
  for (int i = 1; i < 4; ++i)
    for (int j = 1; j < 4; ++j) {
      M[i][j].d[0] = M[i][j].v;
      sig[i][j].d[1] = sig[i][j].v;
    }
  dif[0][1].d[2] = 1.0;
  if (dir)
    for (long n = 0; n < N; ++n) conc[0][n].d[3] = dir[n];
  repartition();
}
template <> void Domain<FormDual>::dual_phase_fraction(double *v, double *d) {
  Dual a(0.0);
  for (long n = 0; n < N; ++n) a += vph[n * 3 + 0];
  *v = a.v / double(N);
  for (int k = 0; k < 4; ++k) d[k] = a.d[k] / double(N);
}
template <> void Domain<FormDual>::dual_conc(double *c, double *dc) {
  for (long n = 0; n < N; ++n) {
    if (c) c[n] = conc[0][n].v;
    if (dc)
      for (int k = 0; k < 4; ++k) dc[k * N + n] = conc[0][n].d[k];
  }
}

DomainBase *dom(void *h) { return static_cast<DomainBase *>(h); }

}  // namespace

extern "C" {

void *orc_create(int nx, int ny, int nz, int depth, int nsolute, int form, int bc_z) {
  if (nx < 1 || ny < 1 || nz < 1 || depth < 1 || depth > 6 || nsolute < 1 || nsolute > 2)
    return nullptr;
  switch (form) {
    case ORC_FP64: return new Domain<FormF64>(nx, ny, nz, depth, nsolute, bc_z);
    case ORC_MIXED: return new Domain<FormMixed>(nx, ny, nz, depth, nsolute, bc_z);
    case ORC_DUAL: return new Domain<FormDual>(nx, ny, nz, depth, nsolute, bc_z);
  }
  return nullptr;
}
void orc_destroy(void *h) { delete dom(h); }
void orc_set_params(void *h, const orc_params *p) { dom(h)->set_params(p); }
void orc_set_state(void *h, const int32_t *index, const double *value, const double *phase,
                   const double *conc) {
  dom(h)->set_state(index, value, phase, conc);
}
void orc_set_comp(void *h, const double *comp) { dom(h)->set_comp(comp); }
void orc_get_state(void *h, int32_t *index, double *value, double *phase, double *conc,
                   double *comp, double *phase_conc) {
  dom(h)->get_state(index, value, phase, conc, comp, phase_conc);
}
void orc_pf_sweep(void *h) { dom(h)->pf_sweep(); }
void orc_diffusion_substeps(void *h, int solute, int n, double dt_d) {
  const orc_params &p = dom(h)->p;
  dom(h)->diffuse(solute, n, dt_d / (p.dx * p.dx));
}
void orc_diffusion_substeps_factor(void *h, int solute, int n, double factor) {
  dom(h)->diffuse(solute, n, factor);
}
// Update(false), phase_field.cpp:243-283, with the parameters already set
void orc_step(void *h, int n_steps) {
  DomainBase *d = dom(h);
  for (int it = 0; it < n_steps; ++it) {
    d->pf_sweep();
    for (int s = 0; s < d->nsolute; ++s)
      if (d->p.n_sub[s] > 0) d->diffuse(s, d->p.n_sub[s], d->p.dt_d[s] / (d->p.dx * d->p.dx));
  }
}
void orc_reduce(void *h, double *mass, double *phase_frac, double *max_conc, double *min_conc,
                double *max_sum_err, int64_t *n_interface) {
  dom(h)->reduce(mass, phase_frac, max_conc, min_conc, max_sum_err, n_interface);
}
int orc_num_threads(void) { return omp_get_max_threads(); }

void orc_dual_seed(void *h, const double *dir) { dom(h)->dual_seed(dir); }
void orc_dual_phase_fraction(void *h, double *v, double *d) { dom(h)->dual_phase_fraction(v, d); }
void orc_dual_conc(void *h, double *c, double *dc) { dom(h)->dual_conc(c, dc); }

// Tessellation, MicrostructureConstruction.cpp:68-109 (1-based cells, strict '>' keeps the
// lowest seed index on ties, per-axis periodic distance, reduction stretches x and squeezes y)
void orc_tessellation(int nx, int ny, int nz, const double *seeds, int G, double reduction,
                      int32_t *grain_index) {
  const int shape[3] = {nx, ny, nz};
#pragma omp parallel for
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        int index = 0;
        const double a0 = shape[0] / (1. - reduction), a1 = shape[1] / (1. - reduction),
                     a2 = shape[2] / (1. - reduction);
        double temp1 = a0 * a0 + a1 * a1 + a2 * a2;
        for (int n = 1; n <= G; ++n) {
          const double px = seeds[3 * (n - 1)], py = seeds[3 * (n - 1) + 1],
                       pz = seeds[3 * (n - 1) + 2];
          const double dx = std::fabs(px - i) < 0.5 * shape[0] ? std::fabs(px - i)
                                                                : shape[0] - std::fabs(px - i);
          const double dy = std::fabs(py - j) < 0.5 * shape[1] ? std::fabs(py - j)
                                                                : shape[1] - std::fabs(py - j);
          const double dz = std::fabs(pz - k) < 0.5 * shape[2] ? std::fabs(pz - k)
                                                                : shape[2] - std::fabs(pz - k);
          const double ex = dx / (1. - reduction), ey = dy * (1. - reduction);
          const double temp2 = ex * ex + ey * ey + dz * dz;
          if (temp1 > temp2) {
            temp1 = temp2;
            index = n;
          }
        }
        grain_index[(i - 1) + long(nx) * ((j - 1) + long(ny) * (k - 1))] = index;
      }
}

void orc_construct(void *h, const int32_t *grain_index, int n_iso, const double *init_conc) {
  dom(h)->construct(grain_index, n_iso, init_conc);
}
void orc_nucleate(void *h, int n_nuclei, float radius, const double *nucleus_conc, uint64_t seed) {
  dom(h)->nucleate(n_nuclei, radius, nucleus_conc, seed);
}
void orc_solute_noise(void *h, int solute, double mean, double stddev, uint64_t seed) {
  dom(h)->solute_noise(solute, mean, stddev, seed);
}

// Update_parameter, phase_field.cpp:307-315 and thermo_data::Get_data, thermo_data.cpp:24-40,
// with thermo_functor::Read_data's table filling (thermo_functor.cpp:11-53)
void orc_update_parameter(double T, double dx, double dims, const double *mob0,
                          const double *mobQ, const double *sigma, const double *diff0,
                          const double *diffQ, const double *entropy_lines,
                          const double *para_lines, orc_params *out, double *pfm_dt,
                          double *diff_dt) {
  double mob[3][3], mx = -1e308;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      mob[i][j] = mob0[i * 3 + j] * std::exp(-mobQ[i * 3 + j] / (kR * T));
      out->mobility[i][j] = mob[i][j];
      out->sigma[i][j] = sigma[i * 3 + j];
      mx = std::max(mx, mob[i][j] * sigma[i * 3 + j]);
    }
  double dmax = -1e308;
  for (int i = 0; i < 3; ++i) {
    out->diffusivity[0][i] = diff0[i] * std::exp(-diffQ[i] / (kR * T));
    dmax = std::max(dmax, out->diffusivity[0][i]);
  }
  *pfm_dt = dx * dx / dims * 0.2 / mx;
  *diff_dt = dx * dx / dmax / dims * 0.2;
  // functor tables
  double ea[3][3] = {{0}}, eb[3][3] = {{0}};
  double qa[3][3] = {{0}}, qb[3][3] = {{0}}, qc[3][3] = {{0}};
  for (int l = 0; l < 3; ++l) {
    const int i = int(entropy_lines[l * 4]) - 1, j = int(entropy_lines[l * 4 + 1]) - 1;
    ea[i][j] = entropy_lines[l * 4 + 2];
    eb[i][j] = entropy_lines[l * 4 + 3];
    ea[j][i] = -entropy_lines[l * 4 + 2];
    eb[j][i] = -entropy_lines[l * 4 + 3];
  }
  for (int l = 0; l < 6; ++l) {
    const int i = int(para_lines[l * 5]) - 1, j = int(para_lines[l * 5 + 1]) - 1;
    qa[i][j] = para_lines[l * 5 + 2];
    qb[i][j] = para_lines[l * 5 + 3];
    qc[i][j] = para_lines[l * 5 + 4];
  }
  double slope[3][3];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      out->entropy[i][j] = ea[i][j] * T + eb[i][j];
      out->eq_conc[i][j] = qa[i][j] * T * T + qb[i][j] * T + qc[i][j];
      slope[i][j] = 2 * qa[i][j] * T + qb[i][j];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (i == j) slope[i][j] = 1.;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      out->eq_slope[i][j] = slope[i][j];
      out->part_k[0][i][j] = slope[j][i] / slope[i][j];
      out->part_c[0][i][j] = out->eq_conc[j][i] - out->part_k[0][i][j] * out->eq_conc[i][j];
    }
}

uint32_t orc_mt19937_nth(uint32_t seed, int n) {
  std::mt19937 g(seed);
  uint32_t v = 0;
  for (int i = 0; i < n; ++i) v = g();
  return v;
}
uint64_t orc_splitmix64(uint64_t x) { return splitmix64(x); }
double orc_quantize_comp_mean(float comp, float value) { return quantize_comp_mean(comp, value); }
float orc_quantize_comp(float comp, float value, uint32_t key) {
  return quantize_comp(comp, value, key);
}
uint32_t orc_mix32(uint32_t x) { return mix32(x); }

}  // extern "C"
