// pf_diffuse.hip — one diffusion sub-step of one solute (Diffusion_solver,
// C++-openmp-cuda/src/solver.cpp:343-372):
//   del = sum_faces sum_p sqrt(phi_p(nb)*phi_p) * (c_p(nb) - c_p) * D_p ;  c += del * dt/dx^2
// where c_p is the quasi-equilibrium partition (solver.cpp:389-405) of the cell's total
// concentration.  The reference stores the partition of the previous sub-step (value_conc1); here
// it is recomputed from the stored total concentration, which gives the same bits.
// All floating-point expressions keep the reference's evaluation order; the library is built with
// -ffp-contract=off, and HIP's division and square root are IEEE correctly rounded on gfx950.
//
// Work layout: a workgroup owns a 64 x TY column tile and marches through a z-chunk, one plane per
// iteration.  It has TY + 3 wavefronts: TY interior rows, the rows above and below the tile, and
// one wavefront for the 2*TY cells left and right of the tile.  Every lane therefore partitions
// exactly one cell per plane (no wavefront waits on a busier one at the barrier), writes its
// partitioned concentrations and phase fractions to LDS (double-buffered, one barrier per plane),
// and the interior lanes then apply the 7-point flux.  z neighbours of a column stay in
// registers; the global loads of a column run two planes ahead of their use.
// A phase absent from a cell (phi_p == 0) gives zero weight on all its faces, so its partition is
// not computed; weights sqrt(0) and sqrt(1) are exact and skip the square-root sequence; the weight
// of a z face is shared by the two planes it separates (the product is commutative).
#include "pf_kernels.h"

namespace pfamd {

template <class A> struct CellPF {
  A c[3];
  A ph[3];
};

template <class F, int P> struct RawCell {
  typename F::C c;
  typename F::T ph[P];
};

template <class F, int P>
__device__ inline void partition_cell(const RawCell<F, P> &r, const DiffConsts<typename F::A> &dc,
                                      CellPF<typename F::A> &o) {
  using A = typename F::A;
  const A c = A(conc_hi(r.c));
#pragma unroll
  for (int j = 0; j < P; ++j) o.ph[j] = A(r.ph[j]);
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
  A w = prod;
  if (prod != A(0) && prod != A(1)) w = sqrt(prod);
  return w;
}

template <class F> __device__ inline typename F::C conc_update(typename F::C c, typename F::A del);
template <> __device__ inline double conc_update<FormF64>(double c, double del) { return c + del; }
template <> __device__ inline float2 conc_update<FormMixed>(float2 c, float del) {
  // Kahan-compensated increment: y = del - comp; t = c + y; comp = (t - c) - y; c = t
  const float y = del - c.y;
  const float t = c.x + y;
  return make_float2(t, (t - c.x) - y);
}

template <class F, int P, int NF, bool NOFLUX, int TY>
__global__ __launch_bounds__(64 * (TY + 3)) void diffuse_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const typename F::T *__restrict__ pfrac, const DiffConsts<typename F::A> dc, const int nx,
    const int ny, const int nz, const int zchunk, const int tiles_x, const int tiles_y,
    const int nblocks) {
  using A = typename F::A;
  using C = typename F::C;
  using R = RawCell<F, P>;
  constexpr int LW = 66, LH = TY + 2;
  __shared__ A lc[2][P][LH][LW];
  __shared__ A lp[2][P][LH][LW];

  // XCD-aware mapping: workgroup b runs on XCD b%8; each XCD gets a contiguous range of tiles
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

  // role of this lane: the LDS slot (ly, lx) it fills each plane.  The spare lanes of the x-halo
  // wavefront fill the corner slot (0, 0), which no stencil reads; they run the same code so every
  // load in the plane loop is unconditional and the load counters stay exact across iterations.
  const int tx = threadIdx.x, w = threadIdx.y;
  int lx, ly;
  bool interior = false;
  if (w < TY) {
    lx = tx + 1;
    ly = w + 1;
    interior = (x0 + tx < nx) && (y0 + w < ny);
  } else if (w == TY) {
    lx = tx + 1;
    ly = 0;
  } else if (w == TY + 1) {
    lx = tx + 1;
    ly = TY + 1;
  } else if (tx < TY) {
    lx = 0;
    ly = tx + 1;
  } else if (tx < 2 * TY) {
    lx = LW - 1;
    ly = tx - TY + 1;
  } else {
    lx = ly = 0;
  }
  const int gx = ((x0 - 1 + lx) % nx + nx) % nx;
  const int gy = ((y0 - 1 + ly) % ny + ny) % ny;
  const long plane = long(nx) * ny;
  const long col = gx + long(nx) * gy;

  auto zidx = [&](int z) -> int {
    if (NOFLUX) return z < 0 ? 0 : (z >= nz ? nz - 1 : z);
    return z < 0 ? z + nz : (z >= nz ? z - nz : z);
  };
  auto load = [&](int z) -> R {
    R r;
    const long i = col + plane * zidx(z);
    r.c = cin[i];
#pragma unroll
    for (int p = 0; p < P; ++p) r.ph[p] = pfrac[i * P + p];
    return r;
  };
  // last plane whose data a column needs: the +z neighbour of the chunk's last plane
  const int last = NF == 6 ? z1 : z1 - 1;

  CellPF<A> prv, cur, nxt;
  C ccur{}, cnext{};
  R r1{}, r2{};
  A wmz[P];
  {
    const R r0 = load(z0);
    ccur = r0.c;
    partition_cell<F, P>(r0, dc, cur);
    if (NF == 6 && interior) {
      partition_cell<F, P>(load(z0 - 1), dc, prv);
#pragma unroll
      for (int p = 0; p < P; ++p) wmz[p] = face_weight<A>(prv.ph[p] * cur.ph[p]);
    }
    r1 = load(min(z0 + 1, last));
    r2 = load(min(z0 + 2, last));
  }

  // One plane.  `ra` holds plane z+1 (loaded two planes earlier) and is refilled with plane z+3
  // (clamped to the chunk's last needed plane); the loop below alternates two such buffers so no
  // register still being loaded is ever copied (a copy would force the load to complete and cut
  // the prefetch distance to one plane).
  auto plane_step = [&](const int z, R &ra) {
    const int buf = (z - z0) & 1;
    {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        lc[buf][p][ly][lx] = cur.c[p];
        lp[buf][p][ly][lx] = cur.ph[p];
      }
      if (z + 1 <= last) {
        cnext = ra.c;
        partition_cell<F, P>(ra, dc, nxt);
      }
      ra = load(min(z + 3, last));
    }
    __syncthreads();
    if (interior) {
      A del = A(0);
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        if (f < NF) {
          const int yy = ly + (f == 2 ? -1 : (f == 3 ? 1 : 0));
          const int xx = lx + (f == 0 ? -1 : (f == 1 ? 1 : 0));
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const A np = lp[buf][p][yy][xx];
            const A nc = lc[buf][p][yy][xx];
            del += face_weight<A>(np * cur.ph[p]) * (nc - cur.c[p]) * dc.D[p];
          }
        }
      }
      if (NF == 6) {
        A wpz[P];
#pragma unroll
        for (int p = 0; p < P; ++p) del += wmz[p] * (prv.c[p] - cur.c[p]) * dc.D[p];
#pragma unroll
        for (int p = 0; p < P; ++p) {
          wpz[p] = face_weight<A>(nxt.ph[p] * cur.ph[p]);
          del += wpz[p] * (nxt.c[p] - cur.c[p]) * dc.D[p];
        }
#pragma unroll
        for (int p = 0; p < P; ++p) wmz[p] = wpz[p];
      }
      del *= dc.r;
      cout[col + plane * z] = conc_update<F>(ccur, del);
    }
    prv = cur;
    cur = nxt;
    ccur = cnext;
  };
  for (int z = z0; z < z1; z += 2) {
    plane_step(z, r1);
    if (z + 1 < z1) plane_step(z + 1, r2);
  }
}

// Column-marching variant without LDS or barriers: a lane owns one (x, y) column of a 64 x BY tile
// and walks a z-chunk.  Its own column's partitions at z-1, z, z+1 stay in registers (one new
// plane is loaded, two planes ahead); the four x/y neighbours of the current plane are loaded
// (cache hits: they are other lanes' own cells) and partitioned again by this lane.  Latency is
// hidden by occupancy instead of software pipelining across a barrier.
template <class F, int P, int NF, bool NOFLUX, int BY>
__global__ __launch_bounds__(64 * BY) void diffuse_march_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const typename F::T *__restrict__ pfrac, const DiffConsts<typename F::A> dc, const int nx,
    const int ny, const int nz, const int zchunk, const int tiles_x, const int tiles_y,
    const int nblocks) {
  using A = typename F::A;
  using C = typename F::C;
  using R = RawCell<F, P>;
  const int b = blockIdx.x;
  const int per = nblocks >> 3, rem = nblocks & 7, xcd = b & 7;
  const int logical = xcd * per + (xcd < rem ? xcd : rem) + (b >> 3);
  const int tx_t = logical % tiles_x;
  const int t2 = logical / tiles_x;
  const int ty_t = t2 % tiles_y;
  const int tz = t2 / tiles_y;
  const int x = tx_t * 64 + threadIdx.x, y = ty_t * BY + threadIdx.y, z0 = tz * zchunk;
  if (x >= nx || y >= ny) return;  // no barrier in this kernel
  const int z1 = min(z0 + zchunk, nz);
  const long plane = long(nx) * ny;
  const long c0 = x + long(nx) * y;
  long cn[4];
  cn[0] = wrap(x - 1, nx) + long(nx) * y;
  cn[1] = wrap(x + 1, nx) + long(nx) * y;
  cn[2] = x + long(nx) * wrap(y - 1, ny);
  cn[3] = x + long(nx) * wrap(y + 1, ny);
  auto zidx = [&](int z) -> int {
    if (NOFLUX) return z < 0 ? 0 : (z >= nz ? nz - 1 : z);
    return z < 0 ? z + nz : (z >= nz ? z - nz : z);
  };
  auto load = [&](long colm, int z) -> R {
    R r;
    const long i = colm + plane * zidx(z);
    r.c = cin[i];
#pragma unroll
    for (int p = 0; p < P; ++p) r.ph[p] = pfrac[i * P + p];
    return r;
  };
  CellPF<A> prv, cur, nxt;
  C ccur;
  A wmz[P];
  {
    const R r0 = load(c0, z0);
    ccur = r0.c;
    partition_cell<F, P>(r0, dc, cur);
    if (NF == 6) {
      partition_cell<F, P>(load(c0, z0 - 1), dc, prv);
#pragma unroll
      for (int p = 0; p < P; ++p) wmz[p] = face_weight<A>(prv.ph[p] * cur.ph[p]);
    }
  }
  for (int z = z0; z < z1; ++z) {
    R rn[4];
#pragma unroll
    for (int f = 0; f < 4; ++f)
      if (f < NF) rn[f] = load(cn[f], z);
    R rz;
    if (NF == 6) rz = load(c0, z + 1);
    A del = A(0);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (f < NF) {
        CellPF<A> nb;
        partition_cell<F, P>(rn[f], dc, nb);
#pragma unroll
        for (int p = 0; p < P; ++p)
          del += face_weight<A>(nb.ph[p] * cur.ph[p]) * (nb.c[p] - cur.c[p]) * dc.D[p];
      }
    }
    C cnext{};
    if (NF == 6) {
      cnext = rz.c;
      partition_cell<F, P>(rz, dc, nxt);
      A wpz[P];
#pragma unroll
      for (int p = 0; p < P; ++p) del += wmz[p] * (prv.c[p] - cur.c[p]) * dc.D[p];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        wpz[p] = face_weight<A>(nxt.ph[p] * cur.ph[p]);
        del += wpz[p] * (nxt.c[p] - cur.c[p]) * dc.D[p];
      }
#pragma unroll
      for (int p = 0; p < P; ++p) wmz[p] = wpz[p];
    }
    del *= dc.r;
    cout[c0 + plane * z] = conc_update<F>(ccur, del);
    if (NF == 6) {
      prv = cur;
      cur = nxt;
      ccur = cnext;
    } else if (z + 1 < z1) {
      const R r = load(c0, z + 1);
      ccur = r.c;
      partition_cell<F, P>(r, dc, cur);
    }
  }
}

// explicit instantiations used by the host dispatch (pf_api.hip)
#define PF_INST_DIFF_ONE(F, P, NF, NOF)                                                           \
  template __global__ void diffuse_kernel<F, P, NF, NOF, kDiffTY>(                                \
      const F::C *, F::C *, const F::T *, const DiffConsts<F::A>, int, int, int, int, int, int, \
      int);                                                                                     \
  template __global__ void diffuse_march_kernel<F, P, NF, NOF, kMarchBY>(                         \
      const F::C *, F::C *, const F::T *, const DiffConsts<F::A>, int, int, int, int, int, int, \
      int);
#define PF_INST_DIFF(F, P)               \
  PF_INST_DIFF_ONE(F, P, 4, false)       \
  PF_INST_DIFF_ONE(F, P, 6, false)       \
  PF_INST_DIFF_ONE(F, P, 6, true)
PF_INST_DIFF(FormF64, 2)
PF_INST_DIFF(FormF64, 3)
PF_INST_DIFF(FormMixed, 2)
PF_INST_DIFF(FormMixed, 3)

}  // namespace pfamd
