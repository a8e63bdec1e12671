// pf_adjoint.hip — discrete adjoint of the FP64 step (gradient harness).
//
// Forward step k (phase_field::Update, C++-openmp-cuda/src/phase_field.cpp:243-283):
//   (phi_k, pf_k, c_k) --sweep--> (phi_{k+1}, pf_{k+1}, c_k) --n_sub diffusion sub-steps--> c_{k+1}
// The adjoint treats the discrete structure of each step (active-set union, clamp and prune
// decisions, the sort permutation, the top-K cut) as fixed and differentiates the arithmetic on
// it; sqrt'(0) := 0; dt, n_sub and the sub-step length are held fixed.
//
// Kernels:
//   adj_diffuse_kernel   reverse of one diffusion sub-step (solver.cpp:343-372): a gather stencil
//                        on the concentration adjoint with the forward weights times dc_p/dc, plus
//                        the value_phase adjoint (weights and partition) and dL/dD_p partials.
//   adj_sweep_kernel     reverse of the phase-field sweep of one cell (solver.cpp:147-304):
//                        recompute the union, pair terms, clamps and sort, then propagate through
//                        renormalisation -> sort -> clamp -> pair terms -> (own phi, neighbour sums,
//                        local concentration, mobility and interface-energy scales).
//   adj_gather_kernel    scatter of the neighbour-sum adjoints back to the slots that fed them, as
//                        a gather over the stencil (a neighbour slot feeds the entry with its id).
// Schedule: the forward run keeps a device checkpoint every `checkpoint_every` steps; the reverse
// sweep recomputes each segment, then per step recomputes and stores the sub-step concentrations
// and runs the reverse kernels backwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/phasefield_amd.h"
#include "pf_handle_access.h"
#include "pf_kernels.h"

namespace pfamd {

constexpr int kAdjBS = 64;

__device__ inline void nbrs(long cell, int nx, int ny, int nz, bool noflux, long (&nb)[6]) {
  int x, y, z;
  cell_coords(cell, nx, ny, x, y, z);
  const long plane = long(nx) * ny;
  nb[0] = wrap(x - 1, nx) + long(nx) * y + plane * z;
  nb[1] = wrap(x + 1, nx) + long(nx) * y + plane * z;
  nb[2] = x + long(nx) * wrap(y - 1, ny) + plane * z;
  nb[3] = x + long(nx) * wrap(y + 1, ny) + plane * z;
  const int zm = noflux ? (z == 0 ? z : z - 1) : wrap(z - 1, nz);
  const int zp = noflux ? (z == nz - 1 ? z : z + 1) : wrap(z + 1, nz);
  nb[4] = x + long(nx) * y + plane * zm;
  nb[5] = x + long(nx) * y + plane * zp;
}

// partition of all P phases with its derivatives: c_p, 1/t1_p and dc_p/dphi_q (0 where phi_p == 0,
// as the forward kernels leave absent phases at zero weight)
template <int P> struct PartD {
  double cp[3], g[3], h[3][3];
};
template <int P>
__device__ inline void part_deriv(double c, const double *ph, const DiffConsts<double> &dc,
                                  PartD<P> &o) {
#pragma unroll
  for (int i = 0; i < P; ++i) {
    o.cp[i] = 0.0;
    o.g[i] = 0.0;
#pragma unroll
    for (int q = 0; q < P; ++q) o.h[i][q] = 0.0;
    if (ph[i] > 0.0) {
      double t1 = 0.0, t2 = 0.0;
#pragma unroll
      for (int j = 0; j < P; ++j) {
        t1 += ph[j] * dc.pk[i][j];
        t2 += ph[j] * dc.pc[i][j];
      }
      o.cp[i] = (c - t2) / t1;
      o.g[i] = 1.0 / t1;
#pragma unroll
      for (int q = 0; q < P; ++q) o.h[i][q] = -(dc.pc[i][q] + o.cp[i] * dc.pk[i][q]) / t1;
    }
  }
}

// sum of one double over the workgroup (256 lanes = 4 wavefronts), added to acc[blockIdx.x]
__device__ inline void block_accumulate(double v, double *__restrict__ acc) {
  __shared__ double part[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) part[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) acc[blockIdx.x] += (part[0] + part[1]) + (part[2] + part[3]);
}

// Reverse of one sub-step: lam_out = dL/dc^s given lam_in = dL/dc^{s+1} and c^s; accumulates
// dL/dvalue_phase into lam_pf and a per-workgroup partial of dL/dD(phase 1) into accD.
// Bulk cells (flag p: the cell and its stencil hold exactly e_p, and the partition of phase p is
// the identity) have unit weights and dc_p/dc = 1, so lam_out = lam + r D_p sum_faces (lam_nb -
// lam).  The value_phase adjoint of a cell whose value_phase is exactly a unit vector is never
// read (the renormalisation of such an output has zero derivative, a copied cell passes it on to
// the same unit vector, and the initial value_phase is not a gradient output), so those cells
// skip it.
// Full reverse of one cell's sub-step (any cell; the formula reduces to the bulk form on bulk
// cells): writes lam_out[j], adds the value_phase adjoint of non-unit cells, returns the cell's
// dL/dD(phase 1) term.
template <int P, int NF, bool NOFLUX>
__device__ inline double adj_cell(long j, const double *__restrict__ c,
                                  const double *__restrict__ pf, const double *__restrict__ lam_in,
                                  double *__restrict__ lam_out, double *__restrict__ lam_pf,
                                  const DiffConsts<double> &dc, int nx, int ny, int nz) {
  long nb[6];
  nbrs(j, nx, ny, nz, NOFLUX, nb);
  const double lj = lam_in[j];
  double dD0 = 0.0;
  double phj[P];
  int ones = 0, zeros = 0;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    phj[p] = pf[j * P + p];
    ones += phj[p] == 1.0;
    zeros += phj[p] == 0.0;
  }
  const bool unit = ones == 1 && zeros == P - 1;
  PartD<P> me;
  part_deriv<P>(c[j], phj, dc, me);
  double lout = lj, lpf[P];
#pragma unroll
  for (int p = 0; p < P; ++p) lpf[p] = 0.0;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const long n = nb[f];
    double phn[P];
#pragma unroll
    for (int p = 0; p < P; ++p) phn[p] = pf[n * P + p];
    PartD<P> o;
    part_deriv<P>(c[n], phn, dc, o);
    const double ln = lam_in[n];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const double prod = phn[p] * phj[p];
      if (prod > 0.0) {
        const double w = sqrt(prod);
        const double dcp = o.cp[p] - me.cp[p];
        lout += dc.r * dc.D[p] * w * me.g[p] * (ln - lj);
        if (!unit) {
          const double dw = phn[p] / (2.0 * w);  // dw/dphi_p(j)
          lpf[p] += dc.r * dc.D[p] * dw * dcp * (lj - ln);
#pragma unroll
          for (int q = 0; q < P; ++q) lpf[q] += dc.r * dc.D[p] * w * me.h[p][q] * (ln - lj);
        }
        if (p == 0) dD0 += dc.r * lj * w * dcp;
      }
    }
  }
  lam_out[j] = lout;
  if (!unit)
#pragma unroll
    for (int p = 0; p < P; ++p) lam_pf[j * P + p] += lpf[p];
  return dD0;
}

// One thread per cell; bulk cells take the reduced form.
template <int P, int NF, bool NOFLUX>
__global__ __launch_bounds__(256) void adj_diffuse_kernel(
    const double *__restrict__ c, const double *__restrict__ pf, const uint8_t *__restrict__ flag,
    const double *__restrict__ lam_in, double *__restrict__ lam_out, double *__restrict__ lam_pf,
    double *__restrict__ accD, const DiffConsts<double> dc, int nx, int ny, int nz, long ncell) {
  const long j = long(blockIdx.x) * 256 + threadIdx.x;
  double dD0 = 0.0;
  if (j < ncell) {
    const int fp = flag[j];
    if (fp != 0) {
      long nb[6];
      nbrs(j, nx, ny, nz, NOFLUX, nb);
      const double lj = lam_in[j];
      double sl = 0.0;
#pragma unroll
      for (int f = 0; f < NF; ++f) sl += lam_in[nb[f]] - lj;
      lam_out[j] = lj + dc.r * dc.D[fp - 1] * sl;
      if (fp == 1) {
        const double cj = c[j];
        double sc = 0.0;
#pragma unroll
        for (int f = 0; f < NF; ++f) sc += c[nb[f]] - cj;
        dD0 = dc.r * lj * sc;
      }
    } else {
      dD0 = adj_cell<P, NF, NOFLUX>(j, c, pf, lam_in, lam_out, lam_pf, dc, nx, ny, nz);
    }
  }
  block_accumulate(dD0, accD);
}

// Split form: the listed cells (flag 0 after adj_list_flag_kernel: interface cells and bulk cells
// of phase 1, whose terms carry dL/dD(phase 1); with the plane-marching kernel also the bulk cells
// of phases other than its bulk phase) in a grid-stride pass, after the bulk cells were updated by
// the forward bulk kernel run on lambda (for a bulk cell of phase p the reverse sub-step
// lam + r D_p sum_faces (lam_nb - lam) is that kernel's update with lam for c).
template <int P, int NF, bool NOFLUX>
__global__ __launch_bounds__(256) void adj_general_kernel(
    const double *__restrict__ c, const double *__restrict__ pf, const int *__restrict__ list,
    const int *__restrict__ count, const double *__restrict__ lam_in, double *__restrict__ lam_out,
    double *__restrict__ lam_pf, double *__restrict__ accD, const DiffConsts<double> dc, int nx,
    int ny, int nz) {
  const int n = *count;
  double dD0 = 0.0;
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256)
    dD0 += adj_cell<P, NF, NOFLUX>(list[k], c, pf, lam_in, lam_out, lam_pf, dc, nx, ny, nz);
  block_accumulate(dD0, accD);
}

// bulk flags of phase 1 -> 0, so those cells join the listed pass
__global__ __launch_bounds__(256) void adj_list_flag_kernel(uint8_t *__restrict__ flag,
                                                            long ncell) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i < ncell && flag[i] == 1) flag[i] = 0;
}

// Reverse of the phase-field sweep of one cell.  Inputs: the sweep's input state (ids, phi, pf,
// c), lam_phi2/lam_pf2 = adjoints of its outputs.  Outputs: lam_phi1 (own-slot part), lam_pf1,
// lam_c (+=), the neighbour-sum adjoint list of the cell (entry ids and values, count 0 when the
// cell was copied) and per-block partials of dL/d(mobility scale), dL/d(sigma scale).
template <int K, int P, int NF, bool NOFLUX>
__global__ __launch_bounds__(kAdjBS) void adj_sweep_kernel(
    const uint16_t *__restrict__ id1, const double *__restrict__ phi1,
    const double *__restrict__ pf1, const double *__restrict__ c,
    const double *__restrict__ lam_phi2, const double *__restrict__ lam_pf2,
    double *__restrict__ lam_phi1, double *__restrict__ lam_pf1, double *__restrict__ lam_c,
    uint32_t *__restrict__ v2_id, double *__restrict__ v2_lam, uint8_t *__restrict__ v2_n,
    double *__restrict__ accMS, const PFConsts<double> cs, int nx, int ny, int nz, long ncell) {
  __shared__ uint32_t sLg[kStencilMax][kAdjBS];
  __shared__ double sV1[kStencilMax][kAdjBS], sV2[kStencilMax][kAdjBS];
  __shared__ double sD[kStencilMax][kAdjBS], sL1[kStencilMax][kAdjBS], sL2[kStencilMax][kAdjBS];
  __shared__ uint8_t sPerm[kStencilMax][kAdjBS], sLive[kStencilMax][kAdjBS];
  __shared__ double red[2][kAdjBS];
  const int t = threadIdx.x;
  const long i = long(blockIdx.x) * kAdjBS + t;
  double dMs = 0.0, dSs = 0.0;
  const double cutoff = cs.cutoff, dt = cs.dt;
  if (i < ncell) {
    for (int e = 0; e < kStencilMax; ++e) {
      sLg[e][t] = 0;
      sV1[e][t] = sV2[e][t] = sD[e][t] = sL1[e][t] = sL2[e][t] = 0.0;
      sPerm[e][t] = uint8_t(e);
      sLive[e][t] = 0;
    }
    uint32_t g0[K];
    load_ids<K>(id1, i, g0);
    int n = 0;
    for (int s = 0; s < K; ++s)
      if (phi1[i * K + s] > cutoff) {
        ++n;
        sLg[s][t] = g0[s];
        sV1[s][t] = phi1[i * K + s];
      }
    long nb[6];
    nbrs(i, nx, ny, nz, NOFLUX, nb);
    for (int f = 0; f < NF; ++f) {
      uint32_t gn[K];
      load_ids<K>(id1, nb[f], gn);
      for (int s = 0; s < K; ++s) {
        const uint32_t g = gn[s];
        if (g == 0) break;
        const double v = phi1[nb[f] * K + s];
        int kk = 0;
        for (; kk < kStencilMax; ++kk)
          if (sLg[kk][t] == g) break;
        if (kk < kStencilMax) {
          sV2[kk][t] += v;
        } else if (n < kStencilMax) {
          sLg[n][t] = g;
          sV2[n][t] += v;
          ++n;
        }
      }
    }
    if (n == 1) {
      // copied cell: slots and value_phase pass through unchanged
      for (int s = 0; s < K; ++s) lam_phi1[i * K + s] += lam_phi2[i * K + s];
      for (int p = 0; p < P; ++p) lam_pf1[i * P + p] += lam_pf2[i * P + p];
      v2_n[i] = 0;
    } else {
      // local concentrations and driving forces (as the forward sweep)
      double ph[P], lc[4] = {0, 0, 0, 0}, t1v[4] = {1, 1, 1, 1};
      for (int j = 0; j < P; ++j) ph[j] = pf1[i * P + j];
      for (int q = 1; q <= P; ++q) {
        double t1 = 0.0, t2 = 0.0;
        for (int j = 0; j < P; ++j) {
          t1 += ph[j] * cs.pk[q][j + 1];
          t2 += ph[j] * cs.pc[q][j + 1];
        }
        lc[q] = (c[i] - t2) / t1;
        t1v[q] = t1;
      }
      auto dG = [&](int a, int b) -> double {
        return (a != b && a >= 1 && b >= 1) ? cs.ent[a][b] * (lc[a] - cs.eqc[a][b]) / cs.slope[a][b]
                                            : 0.0;
      };
      auto tab = [&](const double (&tb)[4][4], int a, int b) -> double {
        return (a >= 1 && b >= 1 && a <= 3 && b <= 3) ? tb[a][b] : 0.0;
      };
      // pre-clamp updates; live[e]: the clamp and prune left the value differentiable
      for (int a = 0; a < n; ++a) {
        const int pa = code_phase(sLg[a][t]);
        double acc = 0.0;
        for (int b = 0; b < n; ++b) {
          if (b == a) continue;
          const int pb = code_phase(sLg[b][t]);
          if (pa - pb == 2 || pb - pa == 2) continue;
          double tt = (sV2[a][t] * sV1[b][t] - sV2[b][t] * sV1[a][t]) * cs.f1;
          tt += (sV1[a][t] - sV1[b][t]) * cs.f2;
          tt *= tab(cs.sig, pa, pb);
          tt += cs.f3 * sqrt(sV1[a][t] * sV1[b][t]) * dG(pa, pb);
          tt *= tab(cs.M, pa, pb);
          acc += tt;
        }
        acc *= dt;
        acc += sV1[a][t];
        uint8_t live = 1;
        if (acc > 1.0) {
          acc = 1.0;
          live = 0;
        } else if (acc < cutoff) {
          acc = 0.0;
          live = 0;
        }
        sD[a][t] = acc;
        sLive[a][t] = live;
      }
      // selection sort (first K passes) recording the permutation
      for (int a = 0; a < K; ++a) {
        int m = a;
        double best = sD[a][t];
        for (int b = a + 1; b < kStencilMax; ++b)
          if (sD[b][t] > best) {
            best = sD[b][t];
            m = b;
          }
        if (m != a) {
          sD[m][t] = sD[a][t];
          sD[a][t] = best;
          const uint8_t tp = sPerm[a][t];
          sPerm[a][t] = sPerm[m][t];
          sPerm[m][t] = tp;
        }
      }
      // renormalisation adjoint: out(s) = D_s/S, pf'(p) = sum_{s of phase p} D_s / S
      double S = 0.0, vp[4] = {0, 0, 0, 0};
      bool kept[6];
      int phs[6];
      for (int s = 0; s < K; ++s) {
        const uint32_t g = sLg[sPerm[s][t]][t];
        kept[s] = sD[s][t] > cutoff;
        phs[s] = code_phase(g);
        if (kept[s]) {
          S += sD[s][t];
          vp[phs[s]] += sD[s][t];
        }
      }
      double dot = 0.0, dotp = 0.0;
      for (int s = 0; s < K; ++s)
        if (kept[s]) dot += lam_phi2[i * K + s] * (sD[s][t] / S);
      for (int p = 1; p <= P; ++p) dotp += lam_pf2[i * P + p - 1] * (vp[p] / S);
      double lamD[kStencilMax];
      for (int e = 0; e < kStencilMax; ++e) lamD[e] = 0.0;
      for (int s = 0; s < K; ++s)
        if (kept[s]) {
          const int a = sPerm[s][t];
          const double lp = phs[s] >= 1 ? lam_pf2[i * P + phs[s] - 1] : 0.0;
          const double l = (lam_phi2[i * K + s] - dot + lp - dotp) / S;
          lamD[a] = sLive[a][t] ? l : 0.0;
        }
      // pair-term adjoint
      double lamLc[4] = {0, 0, 0, 0};
      for (int a = 0; a < n; ++a) {
        const double la = lamD[a];
        if (la == 0.0) continue;
        const int pa = code_phase(sLg[a][t]);
        const double v1a = sV1[a][t], v2a = sV2[a][t];
        sL1[a][t] += la;  // dD_a/dV1_a: the Euler identity term
        for (int b = 0; b < n; ++b) {
          if (b == a) continue;
          const int pb = code_phase(sLg[b][t]);
          if (pa - pb == 2 || pb - pa == 2) continue;
          const double v1b = sV1[b][t], v2b = sV2[b][t];
          const double sg = tab(cs.sig, pa, pb), mo = tab(cs.M, pa, pb), g = dG(pa, pb);
          const double X = (v2a * v1b - v2b * v1a) * cs.f1 + (v1a - v1b) * cs.f2;
          const double prod = v1a * v1b, sq = sqrt(prod);
          const double k = la * dt * mo;
          // t_ab = M [sigma X + f3 sqrt(v1a v1b) dG]
          sL1[a][t] += k * (sg * (-v2b * cs.f1 + cs.f2) + (prod > 0 ? cs.f3 * g * v1b / (2 * sq) : 0));
          sL1[b][t] += k * (sg * (v2a * cs.f1 - cs.f2) + (prod > 0 ? cs.f3 * g * v1a / (2 * sq) : 0));
          sL2[a][t] += k * sg * v1b * cs.f1;
          sL2[b][t] += -k * sg * v1a * cs.f1;
          if (pa != pb && pa >= 1 && pb >= 1)
            lamLc[pa] += k * cs.f3 * sq * cs.ent[pa][pb] / cs.slope[pa][pb];
          dMs += la * dt * mo * (sg * X + cs.f3 * sq * g);
          dSs += la * dt * mo * sg * X;
        }
      }
      // own slot values entered as V1 of the entry at their slot position
      for (int s = 0; s < K; ++s)
        if (phi1[i * K + s] > cutoff) lam_phi1[i * K + s] += sL1[s][t];
      // local concentration: lc_q = (c - t2_q)/t1_q of the input value_phase
      double lcv = 0.0;
      for (int q = 1; q <= P; ++q) {
        if (lamLc[q] == 0.0) continue;
        lcv += lamLc[q] / t1v[q];
        for (int r = 0; r < P; ++r)
          lam_pf1[i * P + r] +=
              lamLc[q] * (-(cs.pc[q][r + 1] + lc[q] * cs.pk[q][r + 1]) / t1v[q]);
      }
      lam_c[i] += lcv;
      for (int e = 0; e < kStencilMax; ++e) {
        v2_id[i * kStencilMax + e] = e < n ? sLg[e][t] : 0u;
        v2_lam[i * kStencilMax + e] = e < n ? sL2[e][t] : 0.0;
      }
      v2_n[i] = uint8_t(n);
    }
  }
  red[0][t] = dMs;
  red[1][t] = dSs;
  __syncthreads();
  for (int w = kAdjBS / 2; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] += red[0][t + w];
      red[1][t] += red[1][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    accMS[blockIdx.x * 2] += red[0][0];
    accMS[blockIdx.x * 2 + 1] += red[1][0];
  }
}

// Neighbour-sum adjoints back to the slots that fed them: slot s of cell j (before the first empty
// slot) was added to the entry with its id in every cell that has j as a stencil neighbour.
template <int K, int NF, bool NOFLUX>
__global__ __launch_bounds__(256) void adj_gather_kernel(
    const uint16_t *__restrict__ id1, const uint32_t *__restrict__ v2_id,
    const double *__restrict__ v2_lam, const uint8_t *__restrict__ v2_n,
    double *__restrict__ lam_phi1, int nx, int ny, int nz, long ncell) {
  const long j = long(blockIdx.x) * 256 + threadIdx.x;
  if (j >= ncell) return;
  uint32_t g[K];
  load_ids<K>(id1, j, g);
  long nb[6];
  nbrs(j, nx, ny, nz, NOFLUX, nb);
  for (int s = 0; s < K; ++s) {
    if (g[s] == 0) break;
    double acc = 0.0;
    for (int f = 0; f < NF; ++f) {
      const long i = nb[f];
      const int n = v2_n[i];
      for (int e = 0; e < n; ++e)
        if (v2_id[i * kStencilMax + e] == g[s]) {
          acc += v2_lam[i * kStencilMax + e];
          break;
        }
    }
    lam_phi1[j * K + s] += acc;
  }
}

__global__ __launch_bounds__(256) void adj_seed_fraction_kernel(double *__restrict__ lam_pf,
                                                                 int P, double v, long ncell) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i < ncell) lam_pf[i * P] += v;
}

__global__ __launch_bounds__(256) void adj_seed_conc_kernel(const double *__restrict__ c,
                                                            const double *__restrict__ target,
                                                            double w, double *__restrict__ lam,
                                                            double *__restrict__ part,
                                                            long ncell) {
  __shared__ double sh[256];
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  double l = 0.0;
  if (i < ncell) {
    const double d = c[i] - target[i];
    lam[i] += 2.0 * w * d;
    l = w * d * d;
  }
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int wdt = 128; wdt > 0; wdt >>= 1) {
    if (int(threadIdx.x) < wdt) sh[threadIdx.x] += sh[threadIdx.x + wdt];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

}  // namespace pfamd

// ================================================================================================
// host orchestration (the handle's buffers through the internal view defined in pf_api.hip)

using namespace pfamd;

namespace {

struct DevState {  // one FP64 state on the device
  uint16_t *id = nullptr;
  double *phi = nullptr, *pf = nullptr, *c = nullptr;
};

// the device allocations of one adjoint run, released together
struct Arena {
  std::vector<void *> ptrs;
  hipError_t err = hipSuccess;
  template <class T> T *get(size_t count) {
    void *q = nullptr;
    if (err == hipSuccess) err = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (err != hipSuccess) return nullptr;
    ptrs.push_back(q);
    return static_cast<T *>(q);
  }
  DevState state(long N, int KI, int K, int P) {
    DevState d;
    d.id = get<uint16_t>(size_t(N) * KI);
    d.phi = get<double>(size_t(N) * K);
    d.pf = get<double>(size_t(N) * P);
    d.c = get<double>(size_t(N));
    return d;
  }
  ~Arena() {
    for (void *q : ptrs) hipFree(q);
  }
};

// Workgroup layout of the bulk sub-step kernels (64 lanes x 4 rows, `cols` x cells per tile row,
// z split into chunks of at least 8 planes until ~32768 workgroups), as the forward step uses.
struct TileGeom {
  int tiles_x, tiles_y, zchunk, nblocks;
};
TileGeom tile_geom(int nx, int ny, int nz, int nf, int cols) {
  TileGeom g;
  g.tiles_x = (nx + cols - 1) / cols;
  g.tiles_y = (ny + 3) / 4;
  const long tiles = long(g.tiles_x) * g.tiles_y;
  int zc = nz;
  if (nf == 6) {
    const long nzc = std::max(1L, std::min<long>(nz, (32768 + tiles - 1) / tiles));
    zc = int((nz + nzc - 1) / nzc);
    zc = std::max(zc, std::min(nz, 8));
  }
  g.zchunk = zc;
  g.nblocks = int(tiles * ((nz + zc - 1) / zc));
  return g;
}

// One step recomputed in the reverse sweep: its input state, value_phase after its sweep with the
// bulk flags of that value_phase, and the carbon field before every sub-step (csub + N*s,
// s = 0..n_sub; s = n_sub is the step's output).
struct StepData {
  DevState in;
  double *pf_new = nullptr;
  uint8_t *flag = nullptr;
  unsigned *count = nullptr;  // [4] bulk cells per phase of that value_phase
  double *csub = nullptr;
};

template <int K, int P, int NF, bool NOF> struct Adj {
  static int run(pf_handle *h, const pf_adjoint_spec *sp, pf_adjoint_result *out) {
    // the run starts from the state as it is, partitioned with the current tables (which the run
    // keeps fixed)
    handle_state_changed(h);
    const HandleView v = handle_view(h);
    const long N = v.N;
    const int KI = v.KI;
    const hipStream_t st = v.stream;
    const pf_params &p = *v.params;
    const int nsteps = sp->n_steps, CK = std::min(sp->checkpoint_every, sp->n_steps);
    const int nsub = std::max(p.n_sub[0], 0);
    const int nseg = (nsteps + CK - 1) / CK;
    const DiffConsts<double> dc = handle_diff_consts(h, 0, p.dt_d[0] / (p.dx * p.dx));
    const PFConsts<double> cs = handle_pf_consts(h);
    const unsigned gb = unsigned((N + 255) / 256), gs = unsigned((N + kAdjBS - 1) / kAdjBS);
    const size_t n8 = size_t(N) * 8;

    Arena ar;
    std::vector<DevState> ckpt;
    for (int q = 0; q < nseg; ++q) ckpt.push_back(ar.state(N, KI, K, P));
    const DevState fin = ar.state(N, KI, K, P);
    double *lam_c = ar.get<double>(N), *lam_c2 = ar.get<double>(N);
    double *lam_pf = ar.get<double>(size_t(N) * P), *lam_pf2 = ar.get<double>(size_t(N) * P);
    double *lam_phi = ar.get<double>(size_t(N) * K), *lam_phi2 = ar.get<double>(size_t(N) * K);
    double *v2_lam = ar.get<double>(size_t(N) * kStencilMax);
    uint32_t *v2_id = ar.get<uint32_t>(size_t(N) * kStencilMax);
    uint8_t *v2_n = ar.get<uint8_t>(N);
    double *accD = ar.get<double>(gb), *accMS = ar.get<double>(size_t(gs) * 2);
    double *part = ar.get<double>(gb);
    double *target = sp->target_conc ? ar.get<double>(N) : nullptr;
    // split reverse sub-step: bulk cells by the forward bulk kernel on lambda, listed cells apart
    const char *env_split = std::getenv("PFAMD_ADJ_SPLIT");
    const bool split = !(env_split && std::strcmp(env_split, "0") == 0);
    const bool vec = v.nx % 2 == 0;
    const unsigned nl = unsigned((N + 256L * kListChunk - 1) / (256L * kListChunk));
    const unsigned ngen = std::min<unsigned>(gb, 2048);
    uint8_t *unit_tmp = ar.get<uint8_t>(N);  // scratch of the bulk classification
    int *alist = split ? ar.get<int>(N) : nullptr;
    int *ablock = split ? ar.get<int>(nl + 1) : nullptr;
    const TileGeom gvec = tile_geom(v.nx, v.ny, v.nz, NF, 128);
    const TileGeom gsc = tile_geom(v.nx, v.ny, v.nz, NF, 64);
    std::vector<StepData> seg(CK);
    for (auto &sd : seg) sd.in = ar.state(N, KI, K, P);
    // Sub-step concentrations for a whole segment when they fit (every step is then recomputed
    // once in the reverse sweep), else for one step (steps are recomputed from the segment's
    // checkpoint to collect their inputs, then once more each).
    const size_t per_step = n8 * (nsub + 1) + n8 * P + size_t(N);
    size_t mem_free = 0, mem_total = 0;
    if (ar.err == hipSuccess) ar.err = hipMemGetInfo(&mem_free, &mem_total);
    const char *env = std::getenv("PFAMD_ADJ_RECOMPUTE");
    const bool force_step = env && std::strcmp(env, "step") == 0;
    const bool store_segment = !force_step && double(per_step) * CK < 0.8 * double(mem_free);
    const int nstore = store_segment ? CK : 1;
    for (int q = 0; q < nstore; ++q) {
      seg[q].pf_new = ar.get<double>(size_t(N) * P);
      seg[q].flag = ar.get<uint8_t>(N);
      seg[q].count = ar.get<unsigned>(4);
      seg[q].csub = ar.get<double>(size_t(N) * (nsub + 1));
    }
    if (ar.err != hipSuccess)
      return handle_fail(h, PF_ERR_OOM,
                         std::string("adjoint buffers: ") + hipGetErrorString(ar.err));

    // first error of the asynchronous copies and clears below (reported by launch_ok)
    hipError_t async_err = hipSuccess;
    auto async_ok = [&](hipError_t e) {
      if (async_err == hipSuccess) async_err = e;
    };
    auto copy_state = [&](const DevState &dst, const DevState &src) {
      async_ok(hipMemcpyAsync(dst.id, src.id, size_t(N) * KI * 2, hipMemcpyDeviceToDevice, st));
      async_ok(hipMemcpyAsync(dst.phi, src.phi, n8 * K, hipMemcpyDeviceToDevice, st));
      async_ok(hipMemcpyAsync(dst.pf, src.pf, n8 * P, hipMemcpyDeviceToDevice, st));
      async_ok(hipMemcpyAsync(dst.c, src.c, n8, hipMemcpyDeviceToDevice, st));
    };
    auto on_handle = [&]() {
      const HandleView w = handle_view(h);
      DevState d;
      d.id = w.id;
      d.phi = static_cast<double *>(w.phi);
      d.pf = static_cast<double *>(w.pf);
      d.c = static_cast<double *>(w.c0);
      return d;
    };
    auto save = [&](const DevState &dst) { copy_state(dst, on_handle()); };
    auto load = [&](const DevState &src) {
      copy_state(on_handle(), src);
      handle_state_changed(h);
    };
    auto sample_index = [&](int step) -> int {
      for (int q = 0; q < sp->n_samples; ++q)
        if (sp->sample_steps[q] == step) return q;
      return -1;
    };
    auto launch_ok = [&]() -> int {
      const hipError_t e = async_err != hipSuccess ? async_err : hipGetLastError();
      return e == hipSuccess ? PF_OK
                             : handle_fail(h, PF_ERR_HIP, std::string("adjoint: ") +
                                                              hipGetErrorString(e));
    };

    // ---- forward with checkpoints ----
    if (target) async_ok(hipMemcpyAsync(target, sp->target_conc, n8, hipMemcpyHostToDevice, st));
    const auto t0 = std::chrono::steady_clock::now();
    double loss = 0.0;
    std::vector<double> frac(sp->n_samples, 0.0);
    int rc = PF_OK;
    for (int k = 0; k < nsteps && rc == PF_OK; ++k) {
      if (k % CK == 0) save(ckpt[k / CK]);
      rc = pf_step(h, 1);
      const int q = sample_index(k + 1);
      if (rc == PF_OK && q >= 0) {
        pf_stats s;
        rc = pf_reduce(h, &s);
        frac[q] = s.phase_frac[0];
        const double d = frac[q] - sp->target_fraction[q];
        loss += d * d;
      }
    }
    if (rc != PF_OK) return rc;
    save(fin);
    async_ok(hipMemsetAsync(lam_c, 0, n8, st));
    async_ok(hipMemsetAsync(lam_phi, 0, n8 * K, st));
    async_ok(hipMemsetAsync(lam_pf, 0, n8 * P, st));
    async_ok(hipMemsetAsync(accD, 0, size_t(gb) * 8, st));
    async_ok(hipMemsetAsync(accMS, 0, size_t(gs) * 16, st));
    if (target && sp->weight_conc != 0.0) {
      hipLaunchKernelGGL(adj_seed_conc_kernel, dim3(gb), dim3(256), 0, st, fin.c, target,
                         sp->weight_conc, lam_c, part, N);
      std::vector<double> hp(gb);
      async_ok(hipMemcpyAsync(hp.data(), part, size_t(gb) * 8, hipMemcpyDeviceToHost, st));
      if (hipStreamSynchronize(st) != hipSuccess || async_err != hipSuccess) return launch_ok();
      for (double x : hp) loss += x;
    }
    if (hipStreamSynchronize(st) != hipSuccess || async_err != hipSuccess) return launch_ok();
    const auto t1 = std::chrono::steady_clock::now();

    // ---- reverse ----
    // recompute the step whose input the handle holds into slot q
    auto recompute = [&](int q) -> int {
      StepData &sd = seg[q];
      int r = pf_phase_field_sweep(h);
      if (r != PF_OK) return r;
      const DevState cur = on_handle();
      async_ok(hipMemcpyAsync(sd.pf_new, cur.pf, n8 * P, hipMemcpyDeviceToDevice, st));
      (void)launch_classify<double, P, NF, NOF>(st, sd.pf_new, unit_tmp, sd.flag, v.nx, v.ny,
                                                 v.nz, N, v.bulk_mask, sd.count);
      async_ok(hipMemcpyAsync(sd.csub, cur.c, n8, hipMemcpyDeviceToDevice, st));
      for (int s = 0; s < nsub && r == PF_OK; ++s)
        r = handle_diffuse_io(h, sd.csub + size_t(N) * s, sd.csub + size_t(N) * (s + 1));
      if (r != PF_OK) return r;
      // the step's output carbon, the next step's input
      async_ok(hipMemcpyAsync(cur.c, sd.csub + size_t(N) * nsub, n8, hipMemcpyDeviceToDevice, st));
      return launch_ok();
    };
    // reverse step k from slot q (sub-step data) and its input state
    auto reverse = [&](int q, int k, const DevState &in) -> int {
      const StepData &sd = seg[q];
      const int qs = sample_index(k + 1);
      if (qs >= 0) {
        const double g = 2.0 * (frac[qs] - sp->target_fraction[qs]) / double(N);
        hipLaunchKernelGGL(adj_seed_fraction_kernel, dim3(gb), dim3(256), 0, st, lam_pf, P, g, N);
      }
      // the forward step's plane-marching shape, when its first sub-step has picked one
      int mby = 0, mzc = 0;
      handle_march_shape(h, &mby, &mzc);
      if (split) {
        // the listed cells of this step: interface cells and bulk cells of phase 1, and, when the
        // plane-marching kernel (one bulk phase for every cell) runs, every cell not of its phase
        const unsigned *keep = vec && mby ? sd.count : nullptr;
        hipLaunchKernelGGL(adj_list_flag_kernel, dim3(gb), dim3(256), 0, st, sd.flag, N);
        hipLaunchKernelGGL(general_count_kernel, dim3(nl), dim3(256), 0, st, sd.flag, keep, N,
                           ablock, ListOrder{}, static_cast<int *>(nullptr));
        hipLaunchKernelGGL(general_scan_kernel, dim3(1), dim3(1024), 0, st, ablock, int(nl),
                           static_cast<int *>(nullptr));
        hipLaunchKernelGGL(general_scatter_kernel, dim3(nl), dim3(256), 0, st, sd.flag, keep, N,
                           ablock, alist, ListOrder{});
      }
      const int mtx = (v.nx + 127) / 128, mty = mby ? (v.ny + mby - 1) / mby : 0;
      const long mnb = mby ? long(mtx) * mty * ((v.nz + mzc - 1) / mzc) : 0;
      for (int s = nsub - 1; s >= 0; --s) {
        if (split) {
          if (vec && mby == 8)
            hipLaunchKernelGGL((diffuse_bulk_march_kernel<FormF64, P, NF, NOF, 8, 2, 1>),
                               dim3(mnb), dim3(64, 8), 0, st, lam_c, lam_c2, sd.count, dc, v.nx,
                               v.ny, v.nz, mzc, mtx, mty, int(mnb));
          else if (vec && mby == 16)
            hipLaunchKernelGGL((diffuse_bulk_march_kernel<FormF64, P, NF, NOF, 16, 2, 1>),
                               dim3(mnb), dim3(64, 16), 0, st, lam_c, lam_c2, sd.count, dc, v.nx,
                               v.ny, v.nz, mzc, mtx, mty, int(mnb));
          else if (vec && mby == 32)
            hipLaunchKernelGGL((diffuse_bulk_march_kernel<FormF64, P, NF, NOF, 32, 2, 2>),
                               dim3(mnb), dim3(64, 16), 0, st, lam_c, lam_c2, sd.count, dc, v.nx,
                               v.ny, v.nz, mzc, mtx, mty, int(mnb));
          else if (vec)
            hipLaunchKernelGGL((diffuse_bulk_vec_kernel<FormF64, P, NF, NOF, 4, 2>),
                               dim3(gvec.nblocks), dim3(64, 4), 0, st, lam_c, lam_c2, sd.flag, dc,
                               v.nx, v.ny, v.nz, gvec.zchunk, gvec.tiles_x, gvec.tiles_y,
                               gvec.nblocks);
          else
            hipLaunchKernelGGL((diffuse_bulk_kernel<FormF64, P, NF, NOF, 4, true>),
                               dim3(gsc.nblocks), dim3(64, 4), 0, st, lam_c, lam_c2, sd.pf_new,
                               sd.flag, dc, v.nx, v.ny, v.nz, gsc.zchunk, gsc.tiles_x, gsc.tiles_y,
                               gsc.nblocks);
          hipLaunchKernelGGL((adj_general_kernel<P, NF, NOF>), dim3(ngen), dim3(256), 0, st,
                             sd.csub + size_t(N) * s, sd.pf_new, alist, ablock + nl, lam_c, lam_c2,
                             lam_pf, accD, dc, v.nx, v.ny, v.nz);
        } else {
          hipLaunchKernelGGL((adj_diffuse_kernel<P, NF, NOF>), dim3(gb), dim3(256), 0, st,
                             sd.csub + size_t(N) * s, sd.pf_new, sd.flag, lam_c, lam_c2, lam_pf,
                             accD, dc, v.nx, v.ny, v.nz, N);
        }
        std::swap(lam_c, lam_c2);
      }
      async_ok(hipMemsetAsync(lam_phi2, 0, n8 * K, st));
      async_ok(hipMemsetAsync(lam_pf2, 0, n8 * P, st));
      hipLaunchKernelGGL((adj_sweep_kernel<K, P, NF, NOF>), dim3(gs), dim3(kAdjBS), 0, st, in.id,
                         in.phi, in.pf, in.c, lam_phi, lam_pf, lam_phi2, lam_pf2, lam_c, v2_id,
                         v2_lam, v2_n, accMS, cs, v.nx, v.ny, v.nz, N);
      hipLaunchKernelGGL((adj_gather_kernel<K, NF, NOF>), dim3(gb), dim3(256), 0, st, in.id,
                         v2_id, v2_lam, v2_n, lam_phi2, v.nx, v.ny, v.nz, N);
      std::swap(lam_phi, lam_phi2);
      std::swap(lam_pf, lam_pf2);
      return launch_ok();
    };
    for (int sg = nseg - 1; sg >= 0 && rc == PF_OK; --sg) {
      const int k0 = sg * CK, k1 = std::min(nsteps, k0 + CK);
      load(ckpt[sg]);
      if (store_segment) {
        for (int k = k0; k < k1 && rc == PF_OK; ++k) {
          save(seg[k - k0].in);
          rc = recompute(k - k0);
        }
        for (int k = k1 - 1; k >= k0 && rc == PF_OK; --k) rc = reverse(k - k0, k, seg[k - k0].in);
      } else {
        for (int k = k0; k < k1 && rc == PF_OK; ++k) {
          save(seg[k - k0].in);
          if (k + 1 < k1) rc = pf_step(h, 1);
        }
        for (int k = k1 - 1; k >= k0 && rc == PF_OK; --k) {
          load(seg[k - k0].in);
          rc = recompute(0);
          if (rc == PF_OK) rc = reverse(0, k, seg[k - k0].in);
        }
      }
    }
    if (rc == PF_OK) {
      std::vector<double> hD(gb), hMS(size_t(gs) * 2);
      async_ok(hipMemcpyAsync(hD.data(), accD, hD.size() * 8, hipMemcpyDeviceToHost, st));
      async_ok(hipMemcpyAsync(hMS.data(), accMS, hMS.size() * 8, hipMemcpyDeviceToHost, st));
      if (out->d_conc0)
        async_ok(hipMemcpyAsync(out->d_conc0, lam_c, n8, hipMemcpyDeviceToHost, st));
      hipError_t e = hipStreamSynchronize(st);
      if (e == hipSuccess) e = async_err;
      if (e != hipSuccess) rc = handle_fail(h, PF_ERR_HIP, hipGetErrorString(e));
      const auto t2 = std::chrono::steady_clock::now();
      double dD = 0.0, dM = 0.0, dS = 0.0;
      for (double x : hD) dD += x;
      for (unsigned b = 0; b < gs; ++b) {
        dM += hMS[size_t(b) * 2];
        dS += hMS[size_t(b) * 2 + 1];
      }
      out->loss = loss;
      out->d_mobility = dM;
      out->d_sigma = dS;
      out->d_diff_gamma = dD;
      for (int q = 0; q < sp->n_samples; ++q) out->fraction[q] = frac[q];
      out->forward_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
      out->reverse_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    // leave the final forward state on the handle
    load(fin);
    const hipError_t e = hipStreamSynchronize(st);
    if (rc == PF_OK && e != hipSuccess) rc = handle_fail(h, PF_ERR_HIP, hipGetErrorString(e));
    return rc;
  }
};

template <int K, int P> int run_nf(pf_handle *h, const pf_adjoint_spec *sp, pf_adjoint_result *o) {
  HandleView v = handle_view(h);
  if (v.NF == 4) return Adj<K, P, 4, false>::run(h, sp, o);
  if (v.noflux) return Adj<K, P, 6, true>::run(h, sp, o);
  return Adj<K, P, 6, false>::run(h, sp, o);
}

}  // namespace

extern "C" int pf_adjoint_run(pf_handle *h, const pf_adjoint_spec *spec, pf_adjoint_result *out) {
  if (!h || !spec || !out) return PF_ERR_INVALID;
  HandleView v = handle_view(h);
  if (v.mixed) return handle_fail(h, PF_ERR_INVALID, "the adjoint runs on FP64 domains");
  if (v.nsolute != 1) return handle_fail(h, PF_ERR_INVALID, "the adjoint covers the carbon-only step");
  if (!v.have_state || !v.params) return handle_fail(h, PF_ERR_STATE, "state or parameters missing");
  if (spec->n_steps < 1 || spec->checkpoint_every < 1 || spec->n_samples < 0 ||
      spec->n_samples > 64)
    return handle_fail(h, PF_ERR_INVALID, "bad adjoint spec");
  hipSetDevice(v.device);
  if (v.K != 4 && v.K != 6) return handle_fail(h, PF_ERR_INVALID, "depth must be 4 or 6");
  if (v.K == 4 && v.P == 2) return run_nf<4, 2>(h, spec, out);
  if (v.K == 4 && v.P == 3) return run_nf<4, 3>(h, spec, out);
  if (v.K == 6 && v.P == 2) return run_nf<6, 2>(h, spec, out);
  return run_nf<6, 3>(h, spec, out);
}
