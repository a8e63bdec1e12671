// pf_diffuse.hip — diffusion sub-step kernel (one solute, one sub-step per launch).
// All floating-point expressions keep the reference's evaluation order; the library is built
// with -ffp-contract=off so no multiply-add is fused, and HIP's division and square root are
// IEEE correctly rounded on gfx950.
#include "pf_kernels.h"

namespace pfamd {
// ----------------------------------------------------------------------------------------------
// One diffusion sub-step of one solute (Diffusion_solver, solver.cpp:343-372; the partition of
// the previous sub-step is recomputed from the stored total concentration, which is the value the
// reference keeps in value_conc1):
//   del = sum_faces sum_p sqrt(phi_p(nb)*phi_p) * (c_p(nb) - c_p) * D_p ;  c += del * dt/dx^2
// Workgroup = 64 x TY threads owning a 64 x TY column tile; it marches through a z-chunk.  Per
// plane the partitioned concentrations and phase fractions of the tile plus a one-cell x/y halo
// are staged in LDS (double-buffered, one barrier per plane); the z neighbours of a column live
// in registers.  A phase absent from a cell (phi_p == 0) gets zero weight on all its faces, so
// its partitioned value is not needed and is not computed; the weights sqrt(0) and sqrt(1) are
// taken exactly without the square-root sequence.
template <class A> struct CellPF {
  A c[3];
  A ph[3];
};

template <class F, int P>
__device__ inline void partition_cell(const typename F::C cv, const typename F::T *phrow,
                                      const DiffConsts<typename F::A> &dc,
                                      CellPF<typename F::A> &o) {
  using A = typename F::A;
  const A c = A(conc_hi(cv));
#pragma unroll
  for (int j = 0; j < P; ++j) o.ph[j] = A(phrow[j]);
#pragma unroll
  for (int i = 0; i < P; ++i) {
    A v = A(0);
    if (o.ph[i] > A(0)) {
      A t1 = A(0), t2 = A(0);
#pragma unroll
      for (int j = 0; j < P; ++j) {
        t1 += o.ph[j] * dc.pk[i][j];
        t2 += o.ph[j] * dc.pc[i][j];
      }
      v = c - t2;
      if (t1 != A(1)) v = v / t1;
    }
    o.c[i] = v;
  }
}

template <class A> __device__ inline A face_weight(A prod) {
  return (prod == A(0) || prod == A(1)) ? prod : sqrt(prod);
}

template <class F> __device__ inline typename F::C conc_update(typename F::C c, typename F::A del);
template <> __device__ inline double conc_update<FormF64>(double c, double del) { return c + del; }
template <> __device__ inline float2 conc_update<FormMixed>(float2 c, float del) {
  // Kahan-compensated increment: y = del - comp; t = c + y; comp = (t - c) - y; c = t
  const float y = del - c.y;
  const float t = c.x + y;
  return make_float2(t, (t - c.x) - y);
}

template <class F, int P, int NF, bool NOFLUX>
__global__ __launch_bounds__(64 * kDiffTY) void diffuse_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const typename F::T *__restrict__ pfrac, const DiffConsts<typename F::A> dc, const int nx,
    const int ny, const int nz, const int zchunk, const int tiles_x, const int tiles_y,
    const int nblocks) {
  using T = typename F::T;
  using A = typename F::A;
  using C = typename F::C;
  constexpr int TY = kDiffTY;
  constexpr int LW = 66, LH = TY + 2;
  __shared__ A lc[2][P][LH][LW];
  __shared__ A lp[2][P][LH][LW];

  // XCD-aware mapping: workgroup b runs on XCD b%8; give each XCD a contiguous range of tiles
  // (x fastest, then y, then z-chunks) so y-adjacent tiles share that XCD's L2.
  const int b = blockIdx.x;
  const int per = nblocks >> 3, rem = nblocks & 7, xcd = b & 7;
  const int logical = xcd * per + (xcd < rem ? xcd : rem) + (b >> 3);
  const int tx_t = logical % tiles_x;
  const int t2 = logical / tiles_x;
  const int ty_t = t2 % tiles_y;
  const int tz = t2 / tiles_y;
  const int x0 = tx_t * 64, y0 = ty_t * TY, z0 = tz * zchunk;
  const int z1 = min(z0 + zchunk, nz);

  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const bool own_valid = (x0 + tx < nx) && (y0 + ty < ny);
  const int ox = (x0 + tx) % nx, oy = (y0 + ty) % ny;
  const long plane = long(nx) * ny;
  const long ocol = ox + long(nx) * oy;

  // halo slot of this thread: top row, bottom row, left column, right column
  int hlx = -1, hly = -1;
  long hcol = 0;
  if (NF >= 4) {
    if (tid < 64) {
      hlx = tid + 1;
      hly = 0;
    } else if (tid < 128) {
      hlx = tid - 64 + 1;
      hly = TY + 1;
    } else if (tid < 128 + TY) {
      hlx = 0;
      hly = tid - 128 + 1;
    } else if (tid < 128 + 2 * TY) {
      hlx = LW - 1;
      hly = tid - 128 - TY + 1;
    }
  } else {
    if (tid < TY) {
      hlx = 0;
      hly = tid + 1;
    } else if (tid < 2 * TY) {
      hlx = LW - 1;
      hly = tid - TY + 1;
    }
  }
  if (hlx >= 0) {
    const int hx = ((x0 - 1 + hlx) % nx + nx) % nx;
    const int hy = ((y0 - 1 + hly) % ny + ny) % ny;
    hcol = hx + long(nx) * hy;
  }

  auto zidx = [&](int z) -> int {
    if (NOFLUX) return z < 0 ? 0 : (z >= nz ? nz - 1 : z);
    return z < 0 ? z + nz : (z >= nz ? z - nz : z);
  };

  CellPF<A> prv, cur, nxt;
  C ccur;
  {
    const long cz = ocol + plane * z0;
    ccur = cin[cz];
    partition_cell<F, P>(ccur, pfrac + cz * P, dc, cur);
    if (NF == 6) {
      const long pz = ocol + plane * zidx(z0 - 1);
      partition_cell<F, P>(cin[pz], pfrac + pz * P, dc, prv);
    }
  }

  for (int z = z0; z < z1; ++z) {
    const int buf = (z - z0) & 1;
    C cnext{};
    if (NF == 6) {
      const long nzc = ocol + plane * zidx(z + 1);
      cnext = cin[nzc];
      partition_cell<F, P>(cnext, pfrac + nzc * P, dc, nxt);
    }
    if (hlx >= 0) {
      const long hc = hcol + plane * z;
      CellPF<A> h;
      partition_cell<F, P>(cin[hc], pfrac + hc * P, dc, h);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        lc[buf][p][hly][hlx] = h.c[p];
        lp[buf][p][hly][hlx] = h.ph[p];
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      lc[buf][p][ty + 1][tx + 1] = cur.c[p];
      lp[buf][p][ty + 1][tx + 1] = cur.ph[p];
    }
    __syncthreads();
    if (own_valid) {
      A del = A(0);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
          A nc, np;
          if (f < 4) {
            const int ly = ty + 1 + (f == 2 ? -1 : (f == 3 ? 1 : 0));
            const int lx = tx + 1 + (f == 0 ? -1 : (f == 1 ? 1 : 0));
            nc = lc[buf][p][ly][lx];
            np = lp[buf][p][ly][lx];
          } else if (f == 4) {
            nc = prv.c[p];
            np = prv.ph[p];
          } else {
            nc = nxt.c[p];
            np = nxt.ph[p];
          }
          del += face_weight<A>(np * cur.ph[p]) * (nc - cur.c[p]) * dc.D[p];
        }
      }
      del *= dc.r;
      cout[ocol + plane * z] = conc_update<F>(ccur, del);
    }
    if (NF == 6) {
      prv = cur;
      cur = nxt;
      ccur = cnext;
    } else if (z + 1 < z1) {
      const long cz = ocol + plane * (z + 1);
      ccur = cin[cz];
      partition_cell<F, P>(ccur, pfrac + cz * P, dc, cur);
    }
  }
}

// explicit instantiations used by the host dispatch (pf_api.hip)
#define PF_INST_DIFF(F, P)                                                                       \
  template __global__ void diffuse_kernel<F, P, 4, false>(const F::C *, F::C *, const F::T *,    \
                                                          const DiffConsts<F::A>, int, int, int, \
                                                          int, int, int, int);                   \
  template __global__ void diffuse_kernel<F, P, 6, false>(const F::C *, F::C *, const F::T *,    \
                                                          const DiffConsts<F::A>, int, int, int, \
                                                          int, int, int, int);                   \
  template __global__ void diffuse_kernel<F, P, 6, true>(const F::C *, F::C *, const F::T *,     \
                                                         const DiffConsts<F::A>, int, int, int,  \
                                                         int, int, int, int);
PF_INST_DIFF(FormF64, 2)
PF_INST_DIFF(FormF64, 3)
PF_INST_DIFF(FormMixed, 2)
PF_INST_DIFF(FormMixed, 3)


}  // namespace pfamd
