// pf_misc.hip — deterministic FP64 reductions, seeded-setup kernels and form conversion.
// All floating-point expressions keep the reference's evaluation order; the library is built
// with -ffp-contract=off so no multiply-add is fused, and HIP's division and square root are
// IEEE correctly rounded on gfx950.
#include "pf_kernels.h"

namespace pfamd {
// ----------------------------------------------------------------------------------------------
// Deterministic FP64 reductions.  Stage 1: each block reduces a contiguous cell range with a
// fixed tree; stage 2: one block reduces the partials in index order.
template <class F, int K, int P>
__global__ __launch_bounds__(256) void reduce_stage1(const uint16_t *__restrict__ id,
                                                     const typename F::T *__restrict__ phi,
                                                     const typename F::T *__restrict__ pfrac,
                                                     const typename F::C *__restrict__ c0,
                                                     const typename F::C *__restrict__ c1, int ns,
                                                     long ncell, long per_block,
                                                     RedPartial *__restrict__ out) {
  __shared__ RedPartial sh[256];
  RedPartial r;
  r.mass[0] = r.mass[1] = 0.0;
  r.pf[0] = r.pf[1] = r.pf[2] = 0.0;
  r.cmax[0] = r.cmax[1] = -1e300;
  r.cmin[0] = r.cmin[1] = 1e300;
  r.serr = 0.0;
  r.nif = r.nnf = 0;
  const long beg = long(blockIdx.x) * per_block;
  const long end = min(beg + per_block, ncell);
  for (long c = beg + threadIdx.x; c < end; c += 256) {
    bool bad = false;
    for (int s = 0; s < ns; ++s) {
      const double v = conc_effective<F>(s == 0 ? c0 : c1, c, ncell);
      r.mass[s] += v;
      r.cmax[s] = fmax(r.cmax[s], v);
      r.cmin[s] = fmin(r.cmin[s], v);
      bad = bad || !isfinite(v);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const double v = double(pfrac[c * P + p]);
      r.pf[p] += v;
      bad = bad || !isfinite(v);
    }
    double sp = 0.0;
#pragma unroll
    for (int s = 0; s < K; ++s) sp += double(phi[c * K + s]);
    bad = bad || !isfinite(sp);
    r.serr = fmax(r.serr, fabs(sp - 1.0));
    uint32_t g[K];
    load_ids<K>(id, c, g);
    r.nif += (K > 1 && g[1] != 0) ? 1 : 0;
    r.nnf += bad ? 1 : 0;
  }
  sh[threadIdx.x] = r;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      RedPartial &a = sh[threadIdx.x];
      const RedPartial &o = sh[threadIdx.x + w];
      for (int s = 0; s < 2; ++s) {
        a.mass[s] += o.mass[s];
        a.cmax[s] = fmax(a.cmax[s], o.cmax[s]);
        a.cmin[s] = fmin(a.cmin[s], o.cmin[s]);
      }
      for (int p = 0; p < 3; ++p) a.pf[p] += o.pf[p];
      a.serr = fmax(a.serr, o.serr);
      a.nif += o.nif;
      a.nnf += o.nnf;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void reduce_stage2(RedPartial *__restrict__ part, int nparts) {
  __shared__ RedPartial sh[256];
  RedPartial r;
  r.mass[0] = r.mass[1] = 0.0;
  r.pf[0] = r.pf[1] = r.pf[2] = 0.0;
  r.cmax[0] = r.cmax[1] = -1e300;
  r.cmin[0] = r.cmin[1] = 1e300;
  r.serr = 0.0;
  r.nif = r.nnf = 0;
  // thread t reduces partials t*chunk .. (t+1)*chunk-1 in order
  const int chunk = (nparts + 255) / 256;
  for (int i = threadIdx.x * chunk; i < min(nparts, (threadIdx.x + 1) * chunk); ++i) {
    const RedPartial &o = part[i];
    for (int s = 0; s < 2; ++s) {
      r.mass[s] += o.mass[s];
      r.cmax[s] = fmax(r.cmax[s], o.cmax[s]);
      r.cmin[s] = fmin(r.cmin[s], o.cmin[s]);
    }
    for (int p = 0; p < 3; ++p) r.pf[p] += o.pf[p];
    r.serr = fmax(r.serr, o.serr);
    r.nif += o.nif;
    r.nnf += o.nnf;
  }
  sh[threadIdx.x] = r;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      RedPartial &a = sh[threadIdx.x];
      const RedPartial &o = sh[threadIdx.x + w];
      for (int s = 0; s < 2; ++s) {
        a.mass[s] += o.mass[s];
        a.cmax[s] = fmax(a.cmax[s], o.cmax[s]);
        a.cmin[s] = fmin(a.cmin[s], o.cmin[s]);
      }
      for (int p = 0; p < 3; ++p) a.pf[p] += o.pf[p];
      a.serr = fmax(a.serr, o.serr);
      a.nif += o.nif;
      a.nnf += o.nnf;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[0] = sh[0];
}

// ----------------------------------------------------------------------------------------------
// Seeded setup kernels (FP64 domains).

// Tessellation (MicrostructureConstruction.cpp:68-109): nearest seed under the per-axis periodic
// distance with the reduction stretch; strict '>' keeps the lowest seed index on ties.  Writes
// slot 0 = grain code (phase 2), phi = 1, value_phase = (0,1[,0]).
template <int K, int P>
__global__ __launch_bounds__(256) void voronoi_kernel(const double *__restrict__ seeds, int G,
                                                      double reduction, int nx, int ny, int nz,
                                                      long ncell, uint16_t *__restrict__ id,
                                                      double *__restrict__ phi,
                                                      double *__restrict__ pfrac) {
  __shared__ double sx[256], sy[256], sz[256];
  const long cell = long(blockIdx.x) * 256 + threadIdx.x;
  const bool valid = cell < ncell;
  const int i = valid ? int(cell % nx) + 1 : 1;
  const int j = valid ? int((cell / nx) % ny) + 1 : 1;
  const int k = valid ? int(cell / (long(nx) * ny)) + 1 : 1;
  const double a0 = nx / (1. - reduction), a1 = ny / (1. - reduction), a2 = nz / (1. - reduction);
  double temp1 = a0 * a0 + a1 * a1 + a2 * a2;
  int index = 0;
  for (int base = 0; base < G; base += 256) {
    __syncthreads();
    if (base + int(threadIdx.x) < G) {
      sx[threadIdx.x] = seeds[3 * (base + threadIdx.x)];
      sy[threadIdx.x] = seeds[3 * (base + threadIdx.x) + 1];
      sz[threadIdx.x] = seeds[3 * (base + threadIdx.x) + 2];
    }
    __syncthreads();
    const int cnt = min(256, G - base);
    for (int q = 0; q < cnt; ++q) {
      const double ax = fabs(sx[q] - i), ay = fabs(sy[q] - j), az = fabs(sz[q] - k);
      const double dx = ax < 0.5 * nx ? ax : nx - ax;
      const double dy = ay < 0.5 * ny ? ay : ny - ay;
      const double dz = az < 0.5 * nz ? az : nz - az;
      const double ex = dx / (1. - reduction), ey = dy * (1. - reduction);
      const double temp2 = ex * ex + ey * ey + dz * dz;
      if (temp1 > temp2) {
        temp1 = temp2;
        index = base + q + 1;
      }
    }
  }
  if (!valid) return;
  uint32_t g[K];
#pragma unroll
  for (int s = 0; s < K; ++s) g[s] = 0;
  g[0] = (2u << 14) | uint32_t(index);
  store_ids<K>(id, cell, g);
#pragma unroll
  for (int s = 0; s < K; ++s) phi[cell * K + s] = s == 0 ? 1.0 : 0.0;
#pragma unroll
  for (int p = 0; p < P; ++p) pfrac[cell * P + p] = p == 1 ? 1.0 : 0.0;
}

// total_conc = dot(value_phase, per-phase initial concentration) (MicrostructureConstruction.cpp:58)
template <int P>
__global__ __launch_bounds__(256) void construct_conc_kernel(const double *__restrict__ pfrac,
                                                             double c1, double c2, double c3,
                                                             long ncell, double *__restrict__ c) {
  const long cell = long(blockIdx.x) * 256 + threadIdx.x;
  if (cell >= ncell) return;
  const double ic[3] = {c1, c2, c3};
  double t = 0.0;
#pragma unroll
  for (int p = 0; p < P; ++p) t += pfrac[cell * P + p] * ic[p];
  c[cell] = t;
}

__device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Nucleation flag (nucleation.cpp:12): slots in use minus the matrix-phase carbon of the cell
// times (1 - 0.001 u); u from the counter-based generator keyed by (seed, cell).
template <int K, int P>
__global__ __launch_bounds__(256) void nucleation_flag_kernel(
    const uint16_t *__restrict__ id, const double *__restrict__ pfrac,
    const double *__restrict__ c0, const DiffConsts<double> dc, uint64_t seed_key, long ncell,
    float *__restrict__ flag, uint8_t *__restrict__ mark) {
  const long cell = long(blockIdx.x) * 256 + threadIdx.x;
  if (cell >= ncell) return;
  uint32_t g[K];
  load_ids<K>(id, cell, g);
  int cnt = 0;
#pragma unroll
  for (int s = 0; s < K; ++s) cnt += g[s] > 0 ? 1 : 0;
  double ph[P];
#pragma unroll
  for (int p = 0; p < P; ++p) ph[p] = pfrac[cell * P + p];
  // value_conc1(2,cell): the full partition of phase 2
  double t1 = 0.0, t2 = 0.0;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    t1 += ph[j] * dc.pk[1][j];
    t2 += ph[j] * dc.pc[1][j];
  }
  const float c2 = float((c0[cell] - t2) / t1);
  const float u = float(double(splitmix64(seed_key + uint64_t(cell)) >> 40) * (1.0 / 16777216.0));
  flag[cell] = float(double(float(cnt)) - double(c2) * (1. - 0.001 * double(u)));
  mark[cell] = 0;
}

// first maximum in storage order: (value, index) pairs, larger value wins, then smaller index
__global__ __launch_bounds__(256) void argmax_stage1(const float *__restrict__ flag, long ncell,
                                                     long per_block, float *__restrict__ pv,
                                                     long long *__restrict__ pi) {
  __shared__ float sv[256];
  __shared__ long long si[256];
  const long beg = long(blockIdx.x) * per_block;
  const long end = min(beg + per_block, ncell);
  float bv = -INFINITY;
  long long bi = -1;
  for (long c = beg + threadIdx.x; c < end; c += 256) {
    const float v = flag[c];
    if (bi < 0 || v > bv) {
      bv = v;
      bi = c;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const float ov = sv[threadIdx.x + w];
      const long long oi = si[threadIdx.x + w];
      const long long mi = si[threadIdx.x];
      if (oi >= 0 && (mi < 0 || ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < mi))) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

__global__ __launch_bounds__(256) void argmax_stage2(float *__restrict__ pv,
                                                     long long *__restrict__ pi, int nparts) {
  __shared__ float sv[256];
  __shared__ long long si[256];
  float bv = -INFINITY;
  long long bi = -1;
  for (int q = threadIdx.x; q < nparts; q += 256) {
    const float v = pv[q];
    const long long i = pi[q];
    if (i >= 0 && (bi < 0 || v > bv || (v == bv && i < bi))) {
      bv = v;
      bi = i;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const float ov = sv[threadIdx.x + w];
      const long long oi = si[threadIdx.x + w];
      const long long mi = si[threadIdx.x];
      if (oi >= 0 && (mi < 0 || ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < mi))) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[0] = sv[0];
    pi[0] = si[0];
  }
}

// Paint nucleus q around the argmax cell (nucleation.cpp:15-37): cells closer than radius become
// nucleus q (slot 0 = phase-1 code, phi = 1, value_phase = (1,0,0)); cells closer than 2*radius
// leave the flag at zero.  The box of half-width R = ceil(2 radius) around the centre holds every
// affected cell; along an axis shorter than 2R+1 the whole axis is scanned.
template <int K, int P>
__global__ __launch_bounds__(256) void nucleus_paint_kernel(
    const long long *__restrict__ center, int code, float radius, int R, int nx, int ny, int nz,
    int cx_n, int cy_n, int cz_n, uint16_t *__restrict__ id, double *__restrict__ phi,
    double *__restrict__ pfrac, float *__restrict__ flag, uint8_t *__restrict__ mark) {
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  const long total = long(cx_n) * cy_n * cz_n;
  if (t >= total) return;
  const long m = center[0];
  const int ctr[3] = {int(m % nx), int((m / nx) % ny), int(m / (long(nx) * ny))};
  const int ext[3] = {nx, ny, nz};
  const int cntv[3] = {cx_n, cy_n, cz_n};
  const int off[3] = {int(t % cx_n), int((t / cx_n) % cy_n), int(t / (long(cx_n) * cy_n))};
  int pos[3];
  for (int d = 0; d < 3; ++d) {
    const int lo = (cntv[d] == ext[d]) ? 0 : ctr[d] - R;
    pos[d] = (((lo + off[d]) % ext[d]) + ext[d]) % ext[d];
  }
  float dist = 0.f;
  for (int d = 0; d < 3; ++d) {
    const int dd = ctr[d] - pos[d];
    const int ad = dd < 0 ? -dd : dd;
    const float s1 = float(dd * dd), s2 = float((ext[d] - ad) * (ext[d] - ad));
    dist += fminf(s1, s2);
  }
  const float Dist = sqrtf(dist);
  const long cell = pos[0] + long(nx) * (pos[1] + long(ny) * pos[2]);
  if (Dist < radius) {
    uint32_t g[K];
#pragma unroll
    for (int s = 0; s < K; ++s) g[s] = 0;
    g[0] = uint32_t(code) & 0xffffu;
    store_ids<K>(id, cell, g);
#pragma unroll
    for (int s = 0; s < K; ++s) phi[cell * K + s] = s == 0 ? 1.0 : 0.0;
#pragma unroll
    for (int p = 0; p < P; ++p) pfrac[cell * P + p] = p == 0 ? 1.0 : 0.0;
    flag[cell] = 0.f;
    mark[cell] = 1;
  } else if (Dist < 2 * radius) {
    flag[cell] = 0.f;
  }
}

__global__ __launch_bounds__(256) void nucleus_conc_kernel(const uint8_t *__restrict__ mark,
                                                           double value, long ncell,
                                                           double *__restrict__ c) {
  const long cell = long(blockIdx.x) * 256 + threadIdx.x;
  if (cell >= ncell) return;
  if (mark[cell]) c[cell] = value;
}

// mean + sd * (sum of 12 uniforms - 6): an Irwin-Hall approximation of a Gaussian whose value
// is reproduced exactly by any IEEE double evaluation in this order
__global__ __launch_bounds__(256) void solute_noise_kernel(double mean, double sd,
                                                           uint64_t seed_key, long ncell,
                                                           double *__restrict__ c) {
  const long cell = long(blockIdx.x) * 256 + threadIdx.x;
  if (cell >= ncell) return;
  double acc = 0.0;
  for (int q = 0; q < 12; ++q)
    acc += double(splitmix64(seed_key + uint64_t(cell) * 12 + q) >> 11) *
           (1.0 / 9007199254740992.0);
  c[cell] = mean + sd * (acc - 6.0);
}

// FP64 domain -> mixed domain (round each stored value; compensation starts at zero)
// grid-stride: n reaches N*K elements (2^32 at 1024^3, K = 4), beyond one launch's work-items
__global__ __launch_bounds__(256) void convert_f64_to_f32(const double *__restrict__ src,
                                                          float *__restrict__ dst, long n) {
  for (long i = long(blockIdx.x) * 256 + threadIdx.x; i < n; i += long(gridDim.x) * 256)
    dst[i] = float(src[i]);
}
__global__ __launch_bounds__(256) void convert_conc_f64_to_mixed(const double *__restrict__ src,
                                                                 float *__restrict__ dst, long n) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) {
    dst[i] = float(src[i]);
    if (i < (n + 1) / 2) comp_bytes(dst, n)[i] = 0;  // the code bytes of cells 2i, 2i + 1
  }
}

// explicit instantiations used by the host dispatch (pf_api.hip)
#define PF_INST_RED(F, K, P)                                                                     \
  template __global__ void reduce_stage1<F, K, P>(const uint16_t *, const F::T *, const F::T *,  \
                                                  const F::C *, const F::C *, int, long, long,   \
                                                  RedPartial *);
PF_INST_RED(FormF64, 4, 2)
PF_INST_RED(FormF64, 4, 3)
PF_INST_RED(FormF64, 6, 2)
PF_INST_RED(FormF64, 6, 3)
PF_INST_RED(FormMixed, 4, 2)
PF_INST_RED(FormMixed, 4, 3)
PF_INST_RED(FormMixed, 6, 2)
PF_INST_RED(FormMixed, 6, 3)

// Fields of the VTI dump (write_data.cpp:42-82): korn = 0 where slot 2 holds a grain, else the
// phase of slot 1 (the grain code's top bits; 0 for an empty cell); conc = the total carbon
// concentration (mixed form: the stored value minus its compensation, as pf_download_state
// reports it), as float.
template <class F>
__global__ __launch_bounds__(256) void vti_fields_kernel(const uint16_t *__restrict__ id,
                                                         const int KI,
                                                         const typename F::C *__restrict__ c,
                                                         const long ncell,
                                                         int32_t *__restrict__ korn,
                                                         float *__restrict__ conc) {
  const long cell = long(blockIdx.x) * 256 + threadIdx.x;
  if (cell >= ncell) return;
  const int g1 = id[cell * KI], g2 = id[cell * KI + 1];
  korn[cell] = g2 != 0 ? 0 : (g1 >> 14);
  conc[cell] = float(conc_effective<F>(c, cell, ncell));
}
template __global__ void vti_fields_kernel<FormF64>(const uint16_t *, int, const double *, long,
                                                    int32_t *, float *);
template __global__ void vti_fields_kernel<FormMixed>(const uint16_t *, int, const float *, long,
                                                      int32_t *, float *);

#define PF_INST_SETUP(K, P)                                                                      \
  template __global__ void voronoi_kernel<K, P>(const double *, int, double, int, int, int,      \
                                                long, uint16_t *, double *, double *);           \
  template __global__ void nucleation_flag_kernel<K, P>(const uint16_t *, const double *,        \
                                                        const double *, const DiffConsts<double>, \
                                                        uint64_t, long, float *, uint8_t *);     \
  template __global__ void nucleus_paint_kernel<K, P>(const long long *, int, float, int, int,  \
                                                      int, int, int, int, int, uint16_t *,       \
                                                      double *, double *, float *, uint8_t *);
PF_INST_SETUP(4, 2)
PF_INST_SETUP(4, 3)
PF_INST_SETUP(6, 2)
PF_INST_SETUP(6, 3)
template __global__ void construct_conc_kernel<2>(const double *, double, double, double, long,
                                                  double *);
template __global__ void construct_conc_kernel<3>(const double *, double, double, double, long,
                                                  double *);

}  // namespace pfamd
