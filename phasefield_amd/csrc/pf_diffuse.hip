// pf_diffuse.hip — one diffusion sub-step of one solute (Diffusion_solver,
// C++-openmp-cuda/src/solver.cpp:343-372):
//   del = sum_faces sum_p sqrt(phi_p(nb)*phi_p) * (c_p(nb) - c_p) * D_p ;  c += del * dt/dx^2
// where c_p is the quasi-equilibrium partition (solver.cpp:389-405) of the cell's total
// concentration.  The reference stores the partition of the previous sub-step (value_conc1); here
// it is recomputed from the stored total concentration, which gives the same bits.
// All floating-point expressions keep the reference's evaluation order; the library is built with
// -ffp-contract=off, and HIP's division and square root are IEEE correctly rounded on gfx950.
//
// Kernels in this file:
//   unit_phase(4)_kernel, classify(4)_kernel  once per phase-field sweep: per-cell flag of "bulk"
//                          cells whose whole stencil sits in one phase (exact reduction of the
//                          formula, see below) and the bulk cells per phase
//   diffuse_bulk_kernel    column-marching, U planes in flight per lane, bulk cells skip the
//                          partition and weights (used when nx is not a multiple of the run length)
//   diffuse_linear_kernel  the same arithmetic, one thread per cell in storage order (small domains)
//   diffuse_bulk_vec_kernel    bulk cells only, 16-byte runs per lane, U planes per batch
//   diffuse_bulk_march_kernel  the bulk update of the most common bulk phase for every cell, no
//                          per-cell flag read, 16-byte runs per lane, plane-marching with the
//                          tile's rows in LDS (the default sub-step when nx is a multiple of the
//                          run length; its tile height and z-chunk are picked by timing)
//   diffuse_general_kernel the full formula over the per-sweep list of the cells the bulk kernels
//                          leave (non-bulk cells; with the plane-marching kernel, every cell
//                          whose flag is not its bulk phase)
//   diffuse_kernel         full formula for every cell, LDS-tiled (selected with
//                          PFAMD_DIFF_KERNEL=lds; kept as an independent path for the parity tests)
//   general_count/scan/scatter_kernel  the storage-order list of the cells a bulk kernel leaves
//
// diffuse_kernel layout: a workgroup owns a 64 x TY column tile and marches through a z-chunk, one
// plane per iteration.  It has TY + 3 wavefronts: TY interior rows, the rows above and below the tile, and
// one wavefront for the 2*TY cells left and right of the tile.  Every lane therefore partitions
// exactly one cell per plane (no wavefront waits on a busier one at the barrier), writes its
// partitioned concentrations and phase fractions to LDS (double-buffered, one barrier per plane),
// and the interior lanes then apply the 7-point flux.  z neighbours of a column stay in
// registers; the global loads of a column run two planes ahead of their use.
// A phase absent from a cell (phi_p == 0) gives zero weight on all its faces, so its partition is
// not computed; weights sqrt(0) and sqrt(1) are exact and skip the square-root sequence; the weight
// of a z face is shared by the two planes it separates (the product is commutative).
#include "pf_diffuse_common.h"

namespace pfamd {

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
      DelAcc<F, P> acc;
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        if (f < NF) {
          const int yy = ly + (f == 2 ? -1 : (f == 3 ? 1 : 0));
          const int xx = lx + (f == 0 ? -1 : (f == 1 ? 1 : 0));
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const A np = lp[buf][p][yy][xx];
            const A nc = lc[buf][p][yy][xx];
            acc.add(p, face_weight<A>(np * cur.ph[p]) * (nc - cur.c[p]), dc);
          }
        }
      }
      if (NF == 6) {
        A wpz[P];
#pragma unroll
        for (int p = 0; p < P; ++p) acc.add(p, wmz[p] * (prv.c[p] - cur.c[p]), dc);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          wpz[p] = face_weight<A>(nxt.ph[p] * cur.ph[p]);
          acc.add(p, wpz[p] * (nxt.c[p] - cur.c[p]), dc);
        }
#pragma unroll
        for (int p = 0; p < P; ++p) wmz[p] = wpz[p];
      }
      const A del = acc.result(dc);
      conc_update<F>(cin, cout, col + plane * z, plane * nz, ccur, del, dc.key);
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

// ------------------------------------------------------------------------------------------------
// Bulk classification, run once per phase-field sweep (value_phase is constant over the diffusion
// sub-cycle): flag = p (1-based) when the cell and all 2*dims stencil neighbours hold exactly the
// unit vector e_p in value_phase and bit p of `mask` is set, else 0.  The host sets bit p when
// every solute's partition table has k(p,p) == 1 and c(p,p) == 0.  For such a cell the
// reference's sub-step reduces exactly to del = sum_faces (c_nb - c) * D_p:
//   weights are sqrt(1*1) = 1 for phase p and sqrt(0*0) = 0 for the others, whose terms are
//   0 * finite = +-0 and leave the running sum unchanged (it starts at +0 and can never become
//   -0 by round-to-nearest addition); the partition of phase p is (c - t2)/t1 with
//   t1 = 1*k(p,p) + sum 0*k(p,j) = 1 and t2 = 1*c(p,p) + sum 0*c(p,j) = +-0, i.e. exactly c.
//
// Two passes: unit_phase_kernel streams value_phase once and writes one byte per cell (p when the
// cell's value_phase is exactly e_p and bit p of `mask` is set, else 0); classify_kernel then
// compares each cell's byte with its stencil's.  The bytes (1/16 of FP64 value_phase) are what the
// stencil re-reads, from cache.  classify_kernel also counts the bulk cells of each phase into
// count[1..3] (zeroed by unit_phase_kernel, which runs first on the same stream): the phase with
// the most bulk cells is the one the plane-marching kernel applies to every cell (bulk_phase).
template <class T, int P>
__global__ __launch_bounds__(256) void unit_phase_kernel(const T *__restrict__ pfrac,
                                                         const long ncell, const unsigned mask,
                                                         uint8_t *__restrict__ unit,
                                                         unsigned *__restrict__ count) {
  const long cell = long(blockIdx.x) * 256 + threadIdx.x;
  if (cell < 4) count[cell] = 0u;
  if (cell >= ncell) return;
  int p = 0, ones = 0, zeros = 0;
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const T v = pfrac[cell * P + q];
    if (v == T(1)) {
      ++ones;
      p = q + 1;
    } else if (v == T(0)) {
      ++zeros;
    }
  }
  if (!(ones == 1 && zeros == P - 1) || !((mask >> p) & 1u)) p = 0;
  unit[cell] = uint8_t(p);
}

// the same for four consecutive cells per thread (ncell % 4 == 0): P*4 values read as vectors, one
// 32-bit word of unit bytes written
template <class T, int P>
__global__ __launch_bounds__(256) void unit_phase4_kernel(const T *__restrict__ pfrac,
                                                          const long ncell, const unsigned mask,
                                                          uint8_t *__restrict__ unit,
                                                          unsigned *__restrict__ count) {
  const long q = long(blockIdx.x) * 256 + threadIdx.x;
  if (q < 4) count[q] = 0u;
  if (q >= ncell / 4) return;
  struct alignas(16) Quad {
    T v[4 * P];
  };
  const Quad ph = reinterpret_cast<const Quad *>(pfrac)[q];
  uint32_t out = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    int p = 0, ones = 0, zeros = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const T v = ph.v[c * P + k];
      if (v == T(1)) {
        ++ones;
        p = k + 1;
      } else if (v == T(0)) {
        ++zeros;
      }
    }
    if (!(ones == 1 && zeros == P - 1) || !((mask >> p) & 1u)) p = 0;
    out |= uint32_t(p) << (8 * c);
  }
  reinterpret_cast<uint32_t *>(unit)[q] = out;
}

constexpr int kClassifyBlocks = 4096;  // grid-stride, so the per-phase counts take few atomics

__global__ __launch_bounds__(64) void zero_counts_kernel(unsigned *__restrict__ count) {
  if (threadIdx.x < 4) count[threadIdx.x] = 0u;
}

// A workgroup's per-phase bulk counts (wave-uniform n1..n3 in each of its four wavefronts) added to
// count[1..3] with one atomic per phase: atomics on one address serialise, and one per wavefront
// (3 x 16384 at 512^3) cost as much as the pass itself.  Every thread of the workgroup calls it.
__device__ inline void add_phase_counts(unsigned n1, unsigned n2, unsigned n3,
                                        unsigned *__restrict__ count) {
  __shared__ unsigned part[4][3];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[wave][0] = n1;
    part[wave][1] = n2;
    part[wave][2] = n3;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const unsigned n = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                       part[3][threadIdx.x];
    if (n) atomicAdd(count + 1 + threadIdx.x, n);
  }
}

template <int NF, bool NOFLUX>
__global__ __launch_bounds__(256) void classify_kernel(const uint8_t *__restrict__ unit,
                                                       uint8_t *__restrict__ flag, const int nx,
                                                       const int ny, const int nz,
                                                       const long ncell,
                                                       unsigned *__restrict__ count) {
  unsigned n1 = 0, n2 = 0, n3 = 0;  // wave-uniform
  const long stride = long(gridDim.x) * 256;
  for (long base = long(blockIdx.x) * 256; base < ncell; base += stride) {
    const long cell = base + threadIdx.x;
    int p = 0;
    if (cell < ncell) {
      int x, y, z;
      cell_coords(cell, nx, ny, x, y, z);
      p = unit[cell];
      if (p != 0) {
        const long plane = long(nx) * ny;
        long nb[6];
        nb[0] = wrap(x - 1, nx) + long(nx) * y + plane * z;
        nb[1] = wrap(x + 1, nx) + long(nx) * y + plane * z;
        nb[2] = x + long(nx) * wrap(y - 1, ny) + plane * z;
        nb[3] = x + long(nx) * wrap(y + 1, ny) + plane * z;
        const int zm = NOFLUX ? (z == 0 ? z : z - 1) : wrap(z - 1, nz);
        const int zp = NOFLUX ? (z == nz - 1 ? z : z + 1) : wrap(z + 1, nz);
        nb[4] = x + long(nx) * y + plane * zm;
        nb[5] = x + long(nx) * y + plane * zp;
#pragma unroll
        for (int f = 0; f < NF; ++f)
          if (unit[nb[f]] != p) p = 0;
      }
      flag[cell] = uint8_t(p);
    }
    n1 += unsigned(__popcll(__ballot(p == 1)));
    n2 += unsigned(__popcll(__ballot(p == 2)));
    n3 += unsigned(__popcll(__ballot(p == 3)));
  }
  add_phase_counts(n1, n2, n3, count);
}

// The same for nx % 4 == 0, four x-consecutive cells per thread: their own bytes, their -y/+y and
// -z/+z neighbours are one aligned 32-bit word each, and only the two x ends are single bytes.
template <int NF, bool NOFLUX>
__global__ __launch_bounds__(256) void classify4_kernel(const uint8_t *__restrict__ unit,
                                                        uint8_t *__restrict__ flag, const int nx,
                                                        const int ny, const int nz,
                                                        const long ncell,
                                                        unsigned *__restrict__ count) {
  const uint32_t *__restrict__ u32 = reinterpret_cast<const uint32_t *>(unit);
  unsigned n1 = 0, n2 = 0, n3 = 0;  // wave-uniform
  const long nq = ncell / 4, stride = long(gridDim.x) * 256;
  // word offsets of the neighbours (32-bit: a domain holds fewer than 2^31 cells)
  const uint32_t qx = uint32_t(nx) / 4, qplane = qx * uint32_t(ny), qvol = qplane * uint32_t(nz);
  for (long base = long(blockIdx.x) * 256; base < nq; base += stride) {
    const long q = base + threadIdx.x;
    uint32_t out = 0;
    if (q < nq) {
      const uint32_t q32 = uint32_t(q), c = 4 * q32;
      int x, y, z;
      cell_coords(c, nx, ny, x, y, z);
      // every load issued together, without a branch on the quad's own bytes (a quad of zeros
      // gives zeros below)
      const uint32_t me = u32[q32];
      uint32_t ym = me, yp = me, zm = me, zp = me;
      if (NF >= 4) {
        ym = u32[y == 0 ? q32 + qplane - qx : q32 - qx];
        yp = u32[y == ny - 1 ? q32 + qx - qplane : q32 + qx];
      }
      if (NF == 6) {
        zm = u32[z == 0 ? (NOFLUX ? q32 : q32 + qvol - qplane) : q32 - qplane];
        zp = u32[z == nz - 1 ? (NOFLUX ? q32 : q32 + qplane - qvol) : q32 + qplane];
      }
      const uint32_t xl = unit[x == 0 ? c + uint32_t(nx) - 1u : c - 1u];
      const uint32_t xr = unit[x + 4 == nx ? c - uint32_t(x) : c + 4u];
      // byte v of `left` / `right` is the -x / +x neighbour of cell v; a cell keeps its byte when
      // all its neighbours' bytes equal it (byte-wise, no carry between bytes)
      const uint32_t left = (me << 8) | xl, right = (me >> 8) | (xr << 24);
      const uint32_t diff = (me ^ left) | (me ^ right) | (me ^ ym) | (me ^ yp) | (me ^ zm) |
                            (me ^ zp);
      const uint32_t same = zero_bytes(diff);  // 0x80 in the bytes of the cells that keep theirs
      out = me & (same | (same - (same >> 7)));
      reinterpret_cast<uint32_t *>(flag)[q32] = out;
    }
    int p[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) p[v] = int((out >> (8 * v)) & 0xffu);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      n1 += unsigned(__popcll(__ballot(p[v] == 1)));
      n2 += unsigned(__popcll(__ballot(p[v] == 2)));
      n3 += unsigned(__popcll(__ballot(p[v] == 3)));
    }
  }
  add_phase_counts(n1, n2, n3, count);
}

template <class T, int P, int NF, bool NOFLUX>
hipError_t launch_classify(hipStream_t st, const T *pfrac, uint8_t *unit, uint8_t *flag, int nx,
                           int ny, int nz, long ncell, unsigned mask, unsigned *count,
                           bool unit_ready) {
  const unsigned nb = unsigned((ncell + 255) / 256);
  if (unit_ready)  // the bytes are current: only the counters to clear
    hipLaunchKernelGGL(zero_counts_kernel, dim3(1), dim3(64), 0, st, count);
  else if (nx % 4 == 0)
    hipLaunchKernelGGL((unit_phase4_kernel<T, P>), dim3((nb + 3) / 4), dim3(256), 0, st, pfrac,
                       ncell, mask, unit, count);
  else
    hipLaunchKernelGGL((unit_phase_kernel<T, P>), dim3(nb), dim3(256), 0, st, pfrac, ncell, mask,
                       unit, count);
  if (nx % 4 == 0)
    hipLaunchKernelGGL((classify4_kernel<NF, NOFLUX>),
                       dim3(std::min<unsigned>((nb + 3) / 4, kClassifyBlocks)), dim3(256), 0, st,
                       unit, flag, nx, ny, nz, ncell, count);
  else
    hipLaunchKernelGGL((classify_kernel<NF, NOFLUX>),
                       dim3(std::min<unsigned>(nb, kClassifyBlocks)), dim3(256), 0, st, unit, flag,
                       nx, ny, nz, ncell, count);
  return hipGetLastError();
}

// Bulk-aware sub-step: a lane owns one (x, y) column of a 64 x BY tile and walks a z-chunk, U
// planes per iteration with all of their loads issued before any arithmetic (the first touch of a
// plane misses to HBM; U planes in flight per lane hide that latency).  Bulk cells (flag p) read
// only the stored concentrations of the stencil and apply del = sum_faces (c_nb - c) * D_p; other
// cells run the full partition + weighted flux on their seven cells.  No LDS and no barrier:
// neighbouring lanes and planes hit in cache.  Plane bases are wave-uniform (scalar) pointers and
// lanes index them with 32-bit column offsets.
template <class F> __device__ inline typename F::A load_hi(const typename F::C *__restrict__ p,
                                                          int i) {
  return typename F::A(p[i]);
}

template <class F, int P, int NF, bool NOFLUX, int BY, bool SPLIT>
__global__ __launch_bounds__(64 * BY) void diffuse_bulk_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const typename F::T *__restrict__ pfrac, const uint8_t *__restrict__ flag,
    const DiffConsts<typename F::A> dc, const int nx, const int ny, const int nz,
    const int zchunk, const int tiles_x, const int tiles_y, const int nblocks) {
  using A = typename F::A;
  using C = typename F::C;
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
  const int oc = x + nx * y;  // column offset within a plane (N <= 2^31 cells)
  int on[4];
  on[0] = wrap(x - 1, nx) + nx * y;
  on[1] = wrap(x + 1, nx) + nx * y;
  on[2] = x + nx * wrap(y - 1, ny);
  on[3] = x + nx * wrap(y + 1, ny);
  auto zidx = [&](int z) -> int {
    if (NOFLUX) return z < 0 ? 0 : (z >= nz ? nz - 1 : z);
    return z < 0 ? z + nz : (z >= nz ? z - nz : z);
  };
  constexpr int U = kDiffUnroll;
  for (int z = z0; z < z1; z += U) {
    const int nv = min(U, z1 - z);  // planes of this batch (wave-uniform)
    uint8_t fl[U];
    C cc[U];
    A xy[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long base = plane * min(z + u, z1 - 1);
      const C *__restrict__ pl = cin + base;
      fl[u] = flag[base + oc];
      cc[u] = pl[oc];
#pragma unroll
      for (int f = 0; f < 4; ++f)
        if (f < NF) xy[u][f] = load_hi<F>(pl, on[f]);
    }
    A zlo = A(0), zhi = A(0);
    if (NF == 6) {
      zlo = load_hi<F>(cin + plane * zidx(z - 1), oc);
      zhi = load_hi<F>(cin + plane * zidx(z + nv), oc);
    }
    unsigned general = 0;  // planes of this batch whose cell needs the full formula
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (u < nv) {
        if (fl[u] != 0) {
          A Dp = dc.D[0];
#pragma unroll
          for (int q = 1; q < P; ++q)
            if (fl[u] == q + 1) Dp = dc.D[q];
          const A c = A(conc_hi(cc[u]));
          BulkAcc<F> acc;
#pragma unroll
          for (int f = 0; f < 4; ++f)
            if (f < NF) acc.add(xy[u][f] - c, Dp);
          if (NF == 6) {
            const A m = u == 0 ? zlo : A(conc_hi(cc[u > 0 ? u - 1 : 0]));
            const A p = u + 1 < nv ? A(conc_hi(cc[u + 1 < U ? u + 1 : u])) : zhi;
            acc.add(m - c, Dp);
            acc.add(p - c, Dp);
          }
          const A del = acc.result(dc.r, Dp * dc.r);
          conc_update<F>(cin, cout, plane * (z + u) + oc, plane * nz, cc[u], del, dc.key);
        } else {
          general |= 1u << u;
        }
      }
    }
    // cells next to a phase boundary: the full partition + weighted flux, one plane at a time
    // (SPLIT: diffuse_general_kernel updates them from the per-sweep list instead)
    if (!SPLIT) {
      while (general) {
        const int u = __builtin_ctz(general);
        general &= general - 1;
        const long base = plane * (z + u);
        long nbi[6];
#pragma unroll
        for (int f = 0; f < 4; ++f) nbi[f] = base + on[f];
        nbi[4] = oc + plane * zidx(z + u - 1);
        nbi[5] = oc + plane * zidx(z + u + 1);
        const long i = base + oc;
        full_update<F, P, NF>(cin, cout, pfrac, i, plane * nz, nbi, dc);
      }
    }
  }
}

// Vector form of the bulk-only sub-step (the full-formula cells are left to the listed pass): a
// lane owns V consecutive x cells of one row (16-byte loads of its own and its y rows), takes the
// x neighbours at the ends of its run with two scalar loads, and walks a z-chunk U planes at a
// time.  No branch on the flag: cells with flag 0 get a value that diffuse_general_kernel
// overwrites afterwards (same stream), so every store is a full vector.  Needs nx % V == 0.
// Streaming stores of a run (one `global_store ... nt` each): the sub-step's output is not read
// again in the same launch, and non-temporal stores keep it from displacing the planes the
// workgroups are still reading (512^3 sub-step: mixed 301 -> 291 us, FP64 445 -> 434 us; the same
// hint on the loads cost 16-30 %, since it evicts the rows the neighbouring tiles re-read).
template <class T, int V> __device__ inline void store_nt(VecA<T, V> *p, const VecA<T, V> &v) {
#pragma unroll
  for (int k = 0; k < V; ++k) __builtin_nontemporal_store(v.v[k], &p->v[k]);
}
// codes of the four cells i..i+3 (i a multiple of 4): one 16-bit load, four bits per cell
__device__ inline VecCode<4> load_codes4(const float *__restrict__ buf, long n, long i) {
  const uint32_t w = *reinterpret_cast<const uint16_t *>(comp_bytes(buf, n) + (i >> 1));
  VecCode<4> q;
#pragma unroll
  for (int v = 0; v < 4; ++v) q.q[v] = comp_unpack(w >> (4 * v));
  return q;
}
__device__ inline void store_codes4_nt(float *__restrict__ buf, long n, long i,
                                       const VecCode<4> &q) {
  uint32_t w = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v) w |= (uint32_t(q.q[v]) & 15u) << (4 * v);
  __builtin_nontemporal_store(uint16_t(w), reinterpret_cast<uint16_t *>(comp_bytes(buf, n) + (i >> 1)));
}

template <class F, int V>
__device__ inline void vec_update(const typename F::C *__restrict__ in, typename F::C *__restrict__ out,
                                  long i, long n, const VecA<typename F::A, V> &c,
                                  const typename F::A (&del)[V], const VecCode<V> &codes,
                                  uint32_t key);
template <>
__device__ inline void vec_update<FormF64, 2>(const double *__restrict__, double *__restrict__ out,
                                              long i, long, const VecA<double, 2> &c,
                                              const double (&del)[2], const VecCode<2> &,
                                              uint32_t) {
  VecA<double, 2> o;
  o.v[0] = c.v[0] + del[0];
  o.v[1] = c.v[1] + del[1];
  store_nt(reinterpret_cast<VecA<double, 2> *>(out + i), o);
}
// The mixed form's update of a run of four cells i..i+3: values stored at vdst, their four 4-bit
// codes (one 16-bit word) at tdst; hbase = i * kCompCellMul gives the rounding hashes of comp_hash
// by additions.
__device__ inline void run_update_mixed(float *__restrict__ vdst, uint16_t *__restrict__ tdst,
                                        const VecA<float, 4> &c, const float (&del)[4],
                                        const VecCode<4> &qi, uint32_t key, uint32_t hbase) {
  VecA<float, 4> o;
  uint32_t w = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float comp = comp_decode(qi.q[v], c.v[v]);
    const float y = del[v] - comp;
    const float t = c.v[v] + y;
    o.v[v] = t;
    const CompCode q = comp_encode((t - c.v[v]) - y, t, key ^ (hbase + uint32_t(v) * kCompCellMul));
    w |= (uint32_t(q) & 15u) << (4 * v);
  }
  store_nt(reinterpret_cast<VecA<float, 4> *>(vdst), o);
  __builtin_nontemporal_store(uint16_t(w), tdst);
}
template <>
__device__ inline void vec_update<FormMixed, 4>(const float *__restrict__ in, float *__restrict__ out,
                                                long i, long n, const VecA<float, 4> &c,
                                                const float (&del)[4], const VecCode<4> &qi,
                                                uint32_t key) {
  run_update_mixed(out + i, reinterpret_cast<uint16_t *>(comp_bytes(out, n) + (i >> 1)), c, del,
                   qi, key, uint32_t(i) * kCompCellMul);
}

template <class F, int P, int NF, bool NOFLUX, int BY, int V>
__global__ __launch_bounds__(64 * BY) void diffuse_bulk_vec_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const uint8_t *__restrict__ flag, const DiffConsts<typename F::A> dc, const int nx,
    const int ny, const int nz, const int zchunk, const int tiles_x, const int tiles_y,
    const int nblocks) {
  using A = typename F::A;
  using VA = VecA<A, V>;
  const int b = blockIdx.x;
  const int per = nblocks >> 3, rem = nblocks & 7, xcd = b & 7;
  const int logical = xcd * per + (xcd < rem ? xcd : rem) + (b >> 3);
  const int tx_t = logical % tiles_x;
  const int t2 = logical / tiles_x;
  const int ty_t = t2 % tiles_y;
  const int tz = t2 / tiles_y;
  const int x0 = (tx_t * 64 + int(threadIdx.x)) * V, y = ty_t * BY + threadIdx.y, z0 = tz * zchunk;
  if (x0 >= nx || y >= ny) return;  // no barrier in this kernel
  const int z1 = min(z0 + zchunk, nz);
  const long plane = long(nx) * ny;
  const long ncell = plane * nz;
  const int oc = x0 + nx * y;  // column offset of the run within a plane
  const int oxm = wrap(x0 - 1, nx) + nx * y, oxp = wrap(x0 + V, nx) + nx * y;
  const int oym = x0 + nx * wrap(y - 1, ny), oyp = x0 + nx * wrap(y + 1, ny);
  auto zidx = [&](int z) -> int {
    if (NOFLUX) return z < 0 ? 0 : (z >= nz ? nz - 1 : z);
    return z < 0 ? z + nz : (z >= nz ? z - nz : z);
  };
  auto ldv = [&](const typename F::C *pl, int o) -> VA {
    return *reinterpret_cast<const VA *>(pl + o);
  };
  constexpr int U = 4;
  for (int z = z0; z < z1; z += U) {
    const int nv = min(U, z1 - z);  // planes of this batch (wave-uniform)
    VA cc[U], ym[U], yp[U];
    A xm[U], xp[U];
    VecFlag<V> fl[U];
    VecCode<V> qc[U];  // mixed: compensation codes
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long base = plane * min(z + u, z1 - 1);
      const typename F::C *__restrict__ pl = cin + base;
      if constexpr (F::kMixed) qc[u] = load_codes4(cin, ncell, base + oc);
      fl[u] = *reinterpret_cast<const VecFlag<V> *>(flag + base + oc);
      cc[u] = ldv(pl, oc);
      xm[u] = A(pl[oxm]);
      xp[u] = A(pl[oxp]);
      if (NF >= 4) {
        ym[u] = ldv(pl, oym);
        yp[u] = ldv(pl, oyp);
      }
    }
    VA zlo, zhi;
    if (NF == 6) {
      zlo = ldv(cin + plane * zidx(z - 1), oc);
      zhi = ldv(cin + plane * zidx(z + nv), oc);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (u < nv) {
        A del[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int fv = fl[u].f[v];
          A Dp = dc.D[0];
#pragma unroll
          for (int q = 1; q < P; ++q)
            if (fv == q + 1) Dp = dc.D[q];
          const A c = cc[u].v[v];
          const A l = v == 0 ? xm[u] : cc[u].v[v > 0 ? v - 1 : 0];
          const A r = v == V - 1 ? xp[u] : cc[u].v[v + 1 < V ? v + 1 : v];
          BulkAcc<F> d;
          d.add(l - c, Dp);
          d.add(r - c, Dp);
          if (NF >= 4) {
            d.add(ym[u].v[v] - c, Dp);
            d.add(yp[u].v[v] - c, Dp);
          }
          if (NF == 6) {
            const A m = u == 0 ? zlo.v[v] : cc[u > 0 ? u - 1 : 0].v[v];
            const A p = u + 1 < nv ? cc[u + 1 < U ? u + 1 : u].v[v] : zhi.v[v];
            d.add(m - c, Dp);
            d.add(p - c, Dp);
          }
          del[v] = d.result(dc.r, Dp * dc.r);
        }
        vec_update<F, V>(cin, cout, plane * (z + u) + oc, ncell, cc[u], del, qc[u], dc.key);
      }
    }
  }
}

// Workgroup barrier ordering LDS only.  __syncthreads() also orders global memory, which makes the
// compiler wait for every outstanding global load (the next plane's prefetch included) before each
// barrier; the plane-marching kernel exchanges data between wavefronts through LDS alone.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Plane-marching form of the bulk-only sub-step.  A workgroup owns a tile of BY rows x 64*V cells,
// R consecutive rows per wavefront, and walks its z-chunk one plane at a time.  The -z / own / +z
// runs of each lane stay in registers; the y neighbours come from the lane's own other rows, from
// the tile's rows staged in LDS (double-buffered, one barrier per plane) or, at the tile's edge
// rows, from global memory; the x neighbours at the ends of a lane's run come from the adjacent
// lanes' registers (lane shuffles), from global memory only at the tile's two ends.  Only the
// tile's edge rows and the z-chunk's two end planes are read from HBM a second time, which is what
// the batched kernel above spends ~30 % of its read traffic on (2/BY of the rows here).  Loads are
// issued in the order they are consumed: this plane's edge rows and tile-end x values first, then
// plane z+2 and its tags, so the wait before the update covers only loads from the previous
// iteration and the prefetch stays in flight across the barrier.  No per-cell flag is read: every
// cell gets the bulk update of the phase with the most bulk cells (bulk_phase), and the cells whose
// flag is another value are listed and overwritten by diffuse_general_kernel.  A cell moves only
// its value (and, mixed, its 4-bit compensation code): FP64 16 B, mixed 9 B per sub-step.  Same
// per-cell arithmetic and face order as diffuse_bulk_vec_kernel (-x, +x, -y, +y, -z, +z), so the
// two are bitwise interchangeable on the cells both keep.  Needs nx % V == 0; partial tiles (x or y) are masked but keep every barrier.
template <class F, int P, int NF, bool NOFLUX, int BY, int V, int R>
__global__ __launch_bounds__(64 * BY / R) void diffuse_bulk_march_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const unsigned *__restrict__ count, const DiffConsts<typename F::A> dc, const int nx,
    const int ny, const int nz, const int zchunk, const int tiles_x, const int tiles_y,
    const int nblocks) {
  using A = typename F::A;
  using VA = VecA<A, V>;
  static_assert(BY % R == 0, "rows per wavefront must divide the tile height");
  // per-cell tags of a plane: the mixed form's V 4-bit compensation codes as one raw word, kept
  // packed until the plane is updated (see the note in the plane step); none for FP64
  struct Tag {
    uint32_t w;
  };
  static_assert(F::kMixed ? V == 4 : V == 2, "tag words assume 16-byte runs");
  // the R rows of one lane in one plane
  struct Rows {
    VA c[R];
  };
  struct Tags {
    Tag t[R];
  };
  __shared__ VA s[2][BY][64];
  const int b = blockIdx.x;
  const int per = nblocks >> 3, rem = nblocks & 7, xcd = b & 7;
  const int logical = xcd * per + (xcd < rem ? xcd : rem) + (b >> 3);
  const int tx_t = logical % tiles_x;
  const int t2 = logical / tiles_x;
  const int ty_t = t2 % tiles_y;
  const int tz = t2 / tiles_y;
  const int lane = threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.y));  // blockDim.x == 64
  const int tx0 = tx_t * 64 * V;
  const int x0 = tx0 + lane * V, z0 = tz * zchunk;
  const int z1 = min(z0 + zchunk, nz);
  const long plane = long(nx) * ny;
  const long ncell = plane * nz;
  const int xc = x0 < nx ? x0 : tx0;
  // the diffusivity of the bulk phase applied to every cell (wave-uniform)
  const int pd = bulk_phase(count);
  A Dp = dc.D[0];
#pragma unroll
  for (int q = 1; q < P; ++q)
    if (pd == q + 1) Dp = dc.D[q];
  const A Dpr = Dp * dc.r;  // the mixed form's factored D_p * r
  // per row r of the lane: tile row tr = wv*R + r, domain row y (rows past the domain load a
  // valid row and store nothing)
  int oc[R], oym[R], oyp[R], oxm[R], oxp[R];
  bool act[R], ym_g[R], yp_g[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int tr = wv * R + r;
    const int y = ty_t * BY + tr;
    const int yc = min(y, ny - 1);
    act[r] = x0 < nx && y < ny;
    oc[r] = xc + nx * yc;
    oym[r] = xc + nx * wrap(yc - 1, ny);
    oyp[r] = xc + nx * wrap(yc + 1, ny);
    oxm[r] = wrap(xc - 1, nx) + nx * yc;
    oxp[r] = wrap(xc + V, nx) + nx * yc;
    // y neighbours from global memory (wave-uniform): the tile's first row, its last row or the
    // domain's last row; the others are in the lane's registers or in LDS
    ym_g[r] = tr == 0;
    yp_g[r] = tr == BY - 1 || y + 1 >= ny;
  }
  auto zidx = [&](int z) -> int {
    if (NOFLUX) return z < 0 ? 0 : (z >= nz ? nz - 1 : z);
    return z < 0 ? z + nz : (z >= nz ? z - nz : z);
  };
  auto ldv = [&](int z, int o) -> VA { return *reinterpret_cast<const VA *>(cin + plane * z + o); };
  // A lane's own run in plane z: the plane's base is wave-uniform and the run's offset a 32-bit
  // byte count (the launcher keeps a plane below 2^32 bytes), so the address is a scalar base plus
  // a vector offset with no 64-bit arithmetic per plane.  Mixed: plane is even (nx % 4 == 0), so the
  // code byte of cell plane * z + o is byte plane * z / 2 + o / 2 of the code region.
  uint32_t ob[R], ot[R], oh[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    ob[r] = uint32_t(oc[r]) * uint32_t(sizeof(typename F::C));
    ot[r] = uint32_t(oc[r]) >> 1;
    oh[r] = uint32_t(oc[r]) * kCompCellMul;
  }
  auto ldrows = [&](int z, Rows &d) {
    const char *pb = reinterpret_cast<const char *>(cin + plane * z);
#pragma unroll
    for (int r = 0; r < R; ++r) d.c[r] = *reinterpret_cast<const VA *>(pb + ob[r]);
  };
  auto ldtags = [&](int z, Tags &d) {
    if constexpr (F::kMixed) {
      const uint8_t *tb = comp_bytes(cin, ncell) + (plane * z >> 1);
#pragma unroll
      for (int r = 0; r < R; ++r) d.t[r].w = *reinterpret_cast<const uint16_t *>(tb + ot[r]);
    }
  };
  // Register rings, indexed by the plane's position in the chunk modulo 4: planes z-1, z, z+1 are
  // in use while z+2 is loading, tags of z and z+1 while z+2 is loading.  The loop is unrolled by
  // 4 so that no register still being loaded is ever copied (a copy would wait for the load).
  Rows r0{}, r1, r2{}, r3{};
  Tags tg0, tg1{}, tg2{}, tg3{};
  ldrows(z0, r1);
  if (NF == 6) ldrows(zidx(z0 - 1), r0);
  ldrows(NF == 6 ? zidx(z0 + 1) : min(z0 + 1, z1 - 1), r2);
  ldtags(z0, tg0);
  if (z0 + 1 < z1) ldtags(z0 + 1, tg1);
#pragma unroll
  for (int r = 0; r < R; ++r) s[0][wv * R + r][lane] = r1.c[r];
  lds_barrier();
  // one plane: pm, pc, pp = planes z-1, z, z+1; pn receives plane z+2; tc = tags of plane z, tn
  // receives the tags of plane z+2
  // dz carries c(z) - c(z-1) of the lane's cells into the plane step and leaves c(z+1) - c(z) for
  // the next: the -z term of plane z + 1 is the negated +z term of plane z (a - b == -(b - a)
  // exactly), one subtraction per cell fewer.  Likewise along x inside a run.
  Rows dz;
  if (NF == 6) {
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int v = 0; v < V; ++v) dz.c[r].v[v] = r1.c[r].v[v] - r0.c[r].v[v];
  }
  auto step = [&](const int z, const Rows &pm, const Rows &pc, const Rows &pp, Rows &pn,
                  const Tags &tc, Tags &tn) {
    const VA *sp = &s[(z - z0) & 1][0][0];
    // this plane's values from outside the lane's registers and LDS
    VA ym[R], yp[R];
    A xm[R], xp[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int tr = wv * R + r;
      if (r > 0) ym[r] = pc.c[r > 0 ? r - 1 : 0];
      else if (ym_g[r]) ym[r] = ldv(z, oym[r]);
      else ym[r] = sp[(tr - 1) * 64 + lane];
      if (yp_g[r]) yp[r] = ldv(z, oyp[r]);
      else if (r + 1 < R) yp[r] = pc.c[r + 1 < R ? r + 1 : r];
      else yp[r] = sp[(tr + 1) * 64 + lane];
      // x neighbours of the run's two ends: the adjacent lanes' registers, global memory only at
      // the tile's ends and at the domain's last run (its +x neighbour wraps to x = 0)
      xm[r] = __shfl_up(pc.c[r].v[V - 1], 1);
      xp[r] = __shfl_down(pc.c[r].v[0], 1);
      if (lane == 0) xm[r] = A(cin[plane * z + oxm[r]]);
      if (lane == 63 || x0 + V >= nx) xp[r] = A(cin[plane * z + oxp[r]]);
    }
    // keep the prefetch below behind these loads: a wait for them then leaves it in flight
    __builtin_amdgcn_sched_barrier(0);
    // then the prefetch: plane z+2 (the +z neighbour of the next plane) and its tags.  Issued on
    // every plane, clamped at the chunk's end to planes already read (cache hits): a load issued
    // on some planes only makes the compiler wait for it inside the branch.
    ldrows(NF == 6 ? zidx(min(z + 2, z1)) : min(z + 2, z1 - 1), pn);
    ldtags(min(z + 2, z1 - 1), tn);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      // consume the x loads here, outside the store's branch, so they are not sunk behind the
      // prefetch (a wait for a load issued after the prefetch would wait for the prefetch too)
      asm volatile("" ::"v"(xm[r]), "v"(xp[r]));
      // The empty asm pins the tag word as loaded: without it the compiler unpacks the ring's
      // tags at the loop's back edge, which waits there for the tags still being prefetched.
      VecCode<V> qc{};
      if constexpr (F::kMixed) {
        Tag tg = tc.t[r];
        asm volatile("" : "+v"(tg.w));
#pragma unroll
        for (int v = 0; v < V; ++v) qc.q[v] = comp_unpack(tg.w >> (4 * v));
      }
      // +x differences of the run (the last one against the next lane's first cell) and the +z
      // differences, for every lane (the next plane needs them whether or not this row stores)
      A dx[V], dzp[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const A c = pc.c[r].v[v];
        dx[v] = (v == V - 1 ? xp[r] : pc.c[r].v[v + 1 < V ? v + 1 : v]) - c;
        if (NF == 6) dzp[v] = pp.c[r].v[v] - c;
      }
      if (act[r]) {
        A del[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const A c = pc.c[r].v[v];
          BulkAcc<F> d;
          d.first(v == 0 ? xm[r] - c : -dx[v > 0 ? v - 1 : 0], Dp);  // l - c
          d.add(dx[v], Dp);                                        // r - c
          if (NF >= 4) {
            d.add(ym[r].v[v] - c, Dp);
            d.add(yp[r].v[v] - c, Dp);
          }
          if (NF == 6) {
            d.add(-dz.c[r].v[v], Dp);  // c(z-1) - c
            d.add(dzp[v], Dp);         // c(z+1) - c
          }
          del[v] = d.result(dc.r, Dpr);
        }
        char *vb = reinterpret_cast<char *>(cout + plane * z) + ob[r];
        if constexpr (F::kMixed) {
          uint8_t *tb = comp_bytes(cout, ncell) + (plane * z >> 1) + ot[r];
          run_update_mixed(reinterpret_cast<float *>(vb), reinterpret_cast<uint16_t *>(tb),
                           pc.c[r], del, qc, dc.key, uint32_t(plane * z) * kCompCellMul + oh[r]);
        } else {
          vec_update<F, V>(cin, reinterpret_cast<typename F::C *>(vb), 0, ncell, pc.c[r], del, qc,
                           dc.key);
        }
      }
      if (NF == 6) {
#pragma unroll
        for (int v = 0; v < V; ++v) dz.c[r].v[v] = dzp[v];
      }
    }
    // plane z+1 into the other buffer (unread after the chunk's last plane); its readers wait at
    // the barrier below, and the buffer read in this plane is only overwritten after the next one
#pragma unroll
    for (int r = 0; r < R; ++r) s[((z - z0) & 1) ^ 1][wv * R + r][lane] = pp.c[r];
    lds_barrier();
  };
  int z = z0;
  for (; z + 3 < z1; z += 4) {
    step(z, r0, r1, r2, r3, tg0, tg2);
    step(z + 1, r1, r2, r3, r0, tg1, tg3);
    step(z + 2, r2, r3, r0, r1, tg2, tg0);
    step(z + 3, r3, r0, r1, r2, tg3, tg1);
  }
  if (z < z1) step(z, r0, r1, r2, r3, tg0, tg2);
  if (z + 1 < z1) step(z + 1, r1, r2, r3, r0, tg1, tg3);
  if (z + 2 < z1) step(z + 2, r2, r3, r0, r1, tg2, tg0);
}

// Storage-order variant: one thread per cell, workgroups over consecutive cells (sequential HBM
// streams); the six neighbours come from the caches.  Same arithmetic as diffuse_bulk_kernel.
template <class F, int P, int NF, bool NOFLUX, bool SPLIT>
__global__ __launch_bounds__(256) void diffuse_linear_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const typename F::T *__restrict__ pfrac, const uint8_t *__restrict__ flag,
    const DiffConsts<typename F::A> dc, const int nx, const int ny, const int nz) {
  using A = typename F::A;
  const long plane = long(nx) * ny;
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i >= plane * nz) return;
  int x, y, z;
  cell_coords(i, nx, ny, x, y, z);
  long nbi[6];
  nbi[0] = i - x + wrap(x - 1, nx);
  nbi[1] = i - x + wrap(x + 1, nx);
  nbi[2] = i + long(nx) * (wrap(y - 1, ny) - y);
  nbi[3] = i + long(nx) * (wrap(y + 1, ny) - y);
  const int zm = NOFLUX ? (z == 0 ? z : z - 1) : wrap(z - 1, nz);
  const int zp = NOFLUX ? (z == nz - 1 ? z : z + 1) : wrap(z + 1, nz);
  nbi[4] = i + plane * (zm - z);
  nbi[5] = i + plane * (zp - z);
  const uint8_t fl = flag[i];
  if (fl != 0) {
    const typename F::C cv = cin[i];
    A Dp = dc.D[0];
#pragma unroll
    for (int q = 1; q < P; ++q)
      if (fl == q + 1) Dp = dc.D[q];
    const A c = A(conc_hi(cv));
    A nv[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) nv[f] = A(conc_hi(cin[nbi[f]]));
    BulkAcc<F> acc;
#pragma unroll
    for (int f = 0; f < NF; ++f) acc.add(nv[f] - c, Dp);
    conc_update<F>(cin, cout, i, plane * nz, cv, acc.result(dc.r, Dp * dc.r), dc.key);
  } else if (!SPLIT) {
    full_update<F, P, NF>(cin, cout, pfrac, i, plane * nz, nbi, dc);
  }
}

// The listed cells (next to a phase boundary, and with the plane-marching kernel also the bulk
// cells of the other phases), from the list built once per sweep: the full partition + weighted
// flux, grid-stride over the device-side count of list entries.
template <class F, int P, int NF, bool NOFLUX>
__global__ __launch_bounds__(256) void diffuse_general_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const typename F::T *__restrict__ pfrac, const int *__restrict__ list,
    const int *__restrict__ count, const DiffConsts<typename F::A> dc, const int nx,
    const int ny, const int nz) {
  using A = typename F::A;
  const int n = *count;
  const long plane = long(nx) * ny, ncell = plane * nz;
  auto stencil = [&](long i, long (&nbi)[6]) {
    int x, y, z;
    cell_coords(i, nx, ny, x, y, z);
    nbi[0] = i - x + wrap(x - 1, nx);
    nbi[1] = i - x + wrap(x + 1, nx);
    nbi[2] = i + long(nx) * (wrap(y - 1, ny) - y);
    nbi[3] = i + long(nx) * (wrap(y + 1, ny) - y);
    const int zm = NOFLUX ? (z == 0 ? z : z - 1) : wrap(z - 1, nz);
    const int zp = NOFLUX ? (z == nz - 1 ? z : z + 1) : wrap(z + 1, nz);
    nbi[4] = i + plane * (zm - z);
    nbi[5] = i + plane * (zp - z);
  };
  const int stride = gridDim.x * 256;
  if constexpr (F::kMixed) {
    // Mixed: the list's entries are byte pairs (ListOrder::pairs), j | mask << 30, and a thread
    // owns the code byte of cells 2j and 2j + 1: their values and the byte are written with plain
    // stores (a cell of the pair that is not listed keeps the code the bulk kernel wrote).  The
    // loads of both cells are issued before the arithmetic.
    if ((nx & 1) == 0) {
      // nx even: cells 2j and 2j + 1 are x neighbours in one row, so the pair's two stencils share
      // the pair's own cells: 12 distinct cells and partitions instead of 14, and the face
      // between the two (its weight sqrt(phi_a phi_b) is the same product both ways) is weighed
      // once.  Both cells are computed, which keeps the wavefront uniform; only the listed ones
      // are stored.
      for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += stride) {
        const uint32_t e = uint32_t(list[k]);
        const long j = long(e & 0x3fffffffu);
        const uint32_t m = e >> 30;
        const long a = 2 * j, b = a + 1;
        // a's stencil (x even, so a + 1 = b is its +x neighbour); b's: its -x neighbour is a, its
        // +x neighbour wraps to the row's start after the row's last cell, its y and z neighbours
        // are a's shifted by one
        int xa, ya, za;
        cell_coords(a, nx, ny, xa, ya, za);
        long na[6];
        na[0] = xa == 0 ? a + nx - 1 : a - 1;
        na[1] = b;
        na[2] = a + long(nx) * (wrap(ya - 1, ny) - ya);
        na[3] = a + long(nx) * (wrap(ya + 1, ny) - ya);
        {
          const int zm = NOFLUX ? (za == 0 ? za : za - 1) : wrap(za - 1, nz);
          const int zp = NOFLUX ? (za == nz - 1 ? za : za + 1) : wrap(za + 1, nz);
          na[4] = a + plane * (zm - za);
          na[5] = a + plane * (zp - za);
        }
        const long b_xp = xa + 2 < nx ? b + 1 : b + 1 - nx;
        const long idx[12] = {a,         b,         na[0],     b_xp,      na[2],     na[3],
                              na[4],     na[5],     na[2] + 1, na[3] + 1, na[4] + 1, na[5] + 1};
        RawCell<F, P> r[12];
#pragma unroll
        for (int q = 0; q < 12; ++q) {
          r[q].c = cin[idx[q]];
#pragma unroll
          for (int p = 0; p < P; ++p) r[q].ph[p] = pfrac[idx[q] * P + p];
        }
        uint8_t *cb = comp_bytes(cout, ncell) + j;
        const uint32_t cin_byte = comp_bytes(cin, ncell)[j];
        const uint32_t old = m == 3u ? 0u : uint32_t(*cb);
        CellPF<A> pc[12];
#pragma unroll
        for (int q = 0; q < 12; ++q) partition_cell<F, P>(r[q], dc, pc[q]);
        // the face weights: the shared a|b face, then a's other faces and b's other faces
        constexpr int NW = (2 * NF - 1) * P;
        A prod[NW], w[NW];
        const int fa[6] = {2, 1, 4, 5, 6, 7}, fb[6] = {0, 3, 8, 9, 10, 11};
        auto wi = [](int side, int f, int p) {  // index in prod/w of face f of cell a (0) / b (1)
          return side == 0 ? (f == 1 ? p : P + ((f == 0 ? 0 : f - 1) * P) + p)
                           : (f == 0 ? p : P + (NF - 1) * P + (f - 1) * P + p);
        };
#pragma unroll
        for (int p = 0; p < P; ++p) prod[p] = pc[1].ph[p] * pc[0].ph[p];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int p = 0; p < P; ++p) {
            if (f != 1) prod[wi(0, f, p)] = pc[fa[f]].ph[p] * pc[0].ph[p];
            if (f != 0) prod[wi(1, f, p)] = pc[fb[f]].ph[p] * pc[1].ph[p];
          }
        face_weights<A>(prod, w);
        // faces in the reference's order (-x, +x, -y, +y, -z, +z)
        DelAcc<F, P> da, db;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const A term = w[wi(0, f, p)] * (pc[fa[f]].c[p] - pc[0].c[p]);
            if (f == 0) da.first(p, term, dc);
            else da.add(p, term, dc);
          }
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const A term = w[wi(1, f, p)] * (pc[fb[f]].c[p] - pc[1].c[p]);
            if (f == 0) db.first(p, term, dc);
            else db.add(p, term, dc);
          }
        CompCode qa, qb;
        const float ta = mixed_step(r[0].c, da.result(dc), comp_unpack(cin_byte), comp_hash(uint32_t(a), dc.key), qa);
        const float tb = mixed_step(r[1].c, db.result(dc), comp_unpack(cin_byte >> 4), comp_hash(uint32_t(b), dc.key), qb);
        uint32_t byte = old;
        if (m & 1u) {
          cout[a] = ta;
          byte = (byte & 0xf0u) | (uint32_t(qa) & 15u);
        }
        if (m & 2u) {
          cout[b] = tb;
          byte = (byte & 0x0fu) | ((uint32_t(qb) & 15u) << 4);
        }
        *cb = uint8_t(byte);
      }
      return;
    }
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += stride) {
      const uint32_t e = uint32_t(list[k]);
      const long j = long(e & 0x3fffffffu);
      const uint32_t m = e >> 30;
      const long i0 = 2 * j + ((m & 1u) ? 0 : 1);  // the first listed cell of the pair
      const long i1 = m == 3u ? i0 + 1 : i0;
      long nb0[6], nb1[6];
      stencil(i0, nb0);
      stencil(i1, nb1);
      StencilLoad<F, P, NF> s0, s1;
      stencil_load<F, P, NF>(cin, pfrac, i0, ncell, nb0, s0);
      stencil_load<F, P, NF>(cin, pfrac, i1, ncell, nb1, s1);
      uint8_t *cb = comp_bytes(cout, ncell) + j;
      uint32_t byte = m == 3u ? 0u : uint32_t(*cb);
      CompCode q;
      cout[i0] = mixed_step(s0.me.c, stencil_del<F, P, NF>(s0, dc), s0.q,
                            comp_hash(uint32_t(i0), dc.key), q);
      const int sh0 = 4 * int(i0 & 1);
      byte = (byte & ~(15u << sh0)) | ((uint32_t(q) & 15u) << sh0);
      if (m == 3u) {
        cout[i1] = mixed_step(s1.me.c, stencil_del<F, P, NF>(s1, dc), s1.q,
                              comp_hash(uint32_t(i1), dc.key), q);
        byte = (byte & 15u) | ((uint32_t(q) & 15u) << 4);
      }
      *cb = uint8_t(byte);
    }
    return;
  }
  // FP64: one cell per entry; two entries per iteration, the second's loads issued before the
  // first's arithmetic (the pass is bound by the latency of its gathers, not by bandwidth)
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += 2 * stride) {
    const bool two = k + stride < n;
    const long i0 = list[k], i1 = two ? long(list[k + stride]) : i0;
    long nb0[6], nb1[6];
    stencil(i0, nb0);
    stencil(i1, nb1);
    StencilLoad<F, P, NF> s0, s1;
    stencil_load<F, P, NF>(cin, pfrac, i0, ncell, nb0, s0);
    stencil_load<F, P, NF>(cin, pfrac, i1, ncell, nb1, s1);
    stencil_update<F, P, NF>(cout, i0, ncell, s0, dc);
    if (two) stencil_update<F, P, NF>(cout, i1, ncell, s1, dc);
  }
}

// The reference keeps each cell's partitioned concentrations (value_conc1) from the last time it
// computed them: the last diffusion sub-step for the cells a sweep copies, the sweep itself for the
// cells it updates (solver.cpp:207-216, 300-302; phase_field.cpp:273-283).  When the partition tables
// change between steps (a temperature path), the first sub-step after the sweep therefore reads
// partitions of the previous step's tables in the copied cells and of the current tables in the
// others.  This kernel reproduces that sub-step: the full formula for every cell, one thread per
// cell, each of the seven cells partitioned with the tables its flag names.
template <class F, int P, int NF, bool NOFLUX>
__global__ __launch_bounds__(256) void diffuse_retable_kernel(
    const typename F::C *__restrict__ cin, typename F::C *__restrict__ cout,
    const typename F::T *__restrict__ pfrac, const uint8_t *__restrict__ copied,
    const DiffConsts<typename F::A> dc, const DiffConsts<typename F::A> dc_prev, const int nx,
    const int ny, const int nz) {
  using A = typename F::A;
  const long plane = long(nx) * ny, ncell = plane * nz;
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i >= ncell) return;
  int x, y, z;
  cell_coords(i, nx, ny, x, y, z);
  long nbi[6];
  nbi[0] = i - x + wrap(x - 1, nx);
  nbi[1] = i - x + wrap(x + 1, nx);
  nbi[2] = i + long(nx) * (wrap(y - 1, ny) - y);
  nbi[3] = i + long(nx) * (wrap(y + 1, ny) - y);
  const int zm = NOFLUX ? (z == 0 ? z : z - 1) : wrap(z - 1, nz);
  const int zp = NOFLUX ? (z == nz - 1 ? z : z + 1) : wrap(z + 1, nz);
  nbi[4] = i + plane * (zm - z);
  nbi[5] = i + plane * (zp - z);
  StencilLoad<F, P, NF> s;
  stencil_load<F, P, NF>(cin, pfrac, i, ncell, nbi, s);
  CellPF<A> me;
  partition_cell<F, P>(s.me, (copied[i] & 1u) ? dc_prev : dc, me);
  DelAcc<F, P> acc;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    CellPF<A> nb;
    partition_cell<F, P>(s.nb[f], (copied[nbi[f]] & 1u) ? dc_prev : dc, nb);
#pragma unroll
    for (int p = 0; p < P; ++p)
      acc.add(p, face_weight<A>(nb.ph[p] * me.ph[p]) * (nb.c[p] - me.c[p]), dc);
  }
  conc_store<F>(cout, i, ncell, s.me.c, acc.result(dc), s.q, dc.key);
}

// ---- list of the cells that need the full formula ---------------------------------------------
__device__ inline bool listed_cell(int f, const unsigned *count, int keep) {
  return count ? f != keep : f == 0;
}
// Blocks of 256 threads x kListChunk consecutive list-order positions (one 16-byte load of flags
// per thread when the chunk's cells are consecutive in storage: always in storage order, and in
// row-tile order when nx % kListChunk == 0, since every row then starts at a multiple of
// kListChunk).  count: listed cells per block; scan: exclusive offsets (one workgroup); scatter:
// each thread writes its listed cells in order at the block's offset plus the exclusive sum of the
// counts of the threads before it, so the list follows the list order.
static_assert(kListChunk == 16, "one uint4 of flags per thread");
__device__ inline bool list_chunk_contiguous(const ListOrder &o) {
  return o.rows == 0 || o.nx % kListChunk == 0;
}
// listed bits of the thread's kListChunk positions starting at `first` (positions at or past
// ncell: none)
__device__ inline unsigned listed_bits(const uint8_t *__restrict__ flag, long first, long ncell,
                                       const unsigned *count, int keep, const ListOrder &o) {
  unsigned bits = 0;
  if (first + kListChunk <= ncell && list_chunk_contiguous(o)) {
    const uint4 w = *reinterpret_cast<const uint4 *>(flag + list_order_cell(o, first));
    const uint32_t word[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int j = 0; j < kListChunk; ++j)
      if (listed_cell(int((word[j >> 2] >> (8 * (j & 3))) & 0xffu), count, keep)) bits |= 1u << j;
  } else {
    for (int j = 0; j < kListChunk && first + j < ncell; ++j)
      if (listed_cell(flag[list_order_cell(o, first + j)], count, keep)) bits |= 1u << j;
  }
  return bits;
}

// the pairs (2j, 2j + 1) of a chunk's listed bits with a listed cell: bit 2j' set for pair j'
__device__ inline unsigned pair_bits(unsigned bits) { return (bits | (bits >> 1)) & 0x5555u; }

__global__ __launch_bounds__(256) void general_count_kernel(const uint8_t *__restrict__ flag,
                                                            const unsigned *__restrict__ count,
                                                            long ncell, int *__restrict__ bcount,
                                                            const ListOrder order,
                                                            int *__restrict__ cells) {
  __shared__ int wsum[4];
  const long first = long(blockIdx.x) * (256L * kListChunk) + long(threadIdx.x) * kListChunk;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ int csum[4];
  const int keep = count ? bulk_phase(count) : 0;
  const unsigned bits = listed_bits(flag, first, ncell, count, keep, order);
  int n = __popc(order.pairs ? pair_bits(bits) : bits);
  int nc = __popc(bits);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    n += __shfl_xor(n, d);
    nc += __shfl_xor(nc, d);
  }
  if (lane == 0) {
    wsum[wave] = n;
    csum[wave] = nc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    bcount[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (cells) cells[blockIdx.x] = csum[0] + csum[1] + csum[2] + csum[3];
  }
}

__global__ __launch_bounds__(8) void zero_queue_kernel(int *__restrict__ queue) {
  queue[32 * threadIdx.x] = 0;
}

// exclusive prefix sum of bcount[0..nb) in place; bcount[nb] = total.  One workgroup of 1024, in
// passes of 16 consecutive counts per thread (four 16-byte loads issued together, a wave scan and
// the 16 wave totals from LDS), the running total carried from pass to pass.
__global__ __launch_bounds__(1024) void general_scan_kernel(int *__restrict__ bcount, int nb,
                                                             int *__restrict__ total_out) {
  __shared__ int wsum[16];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int carry = 0;
  for (int base = 0; base < nb; base += 1024 * 16) {
    const int i0 = base + t * 16;
    const bool full = i0 + 16 <= nb;
    int v[16];
    if (full) {
      const int4 *p = reinterpret_cast<const int4 *>(bcount + i0);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int4 x = p[q];
        v[4 * q] = x.x;
        v[4 * q + 1] = x.y;
        v[4 * q + 2] = x.z;
        v[4 * q + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = i0 + q < nb ? bcount[i0 + q] : 0;
    }
    int sum = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) sum += v[q];
    int incl = sum;  // inclusive scan over the wave's lanes
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(incl, d);
      if (lane >= d) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int woff = 0, pass_total = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      const int x = wsum[w];
      woff += w < wv ? x : 0;
      pass_total += x;
    }
    int run = carry + woff + incl - sum;  // exclusive start of this thread's counts
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = v[q];
      v[q] = run;
      run += c;
    }
    if (full) {
      int4 *p = reinterpret_cast<int4 *>(bcount + i0);
#pragma unroll
      for (int q = 0; q < 4; ++q) p[q] = make_int4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (i0 + q < nb) bcount[i0 + q] = v[q];
    }
    carry += pass_total;
    __syncthreads();  // wsum is rewritten by the next pass
  }
  if (t == 0) {
    bcount[nb] = carry;
    if (total_out) *total_out = carry;  // the list length, kept for diagnostics
  }
}

// Each lane writes its entries, in order, into its wave's LDS row at the exclusive sum of the
// counts of the lanes before it; the wave then copies the row to the list with consecutive lanes
// on consecutive entries (whole-line stores instead of up to 16 scattered stores per lane).
__global__ __launch_bounds__(256) void general_scatter_kernel(const uint8_t *__restrict__ flag,
                                                              const unsigned *__restrict__ count,
                                                              long ncell,
                                                              const int *__restrict__ boff,
                                                              int *__restrict__ list,
                                                              const ListOrder order) {
  __shared__ int wsum[4];
  __shared__ int stage[4][64 * kListChunk];
  const long first = long(blockIdx.x) * (256L * kListChunk) + long(threadIdx.x) * kListChunk;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int keep = count ? bulk_phase(count) : 0;
  unsigned bits = listed_bits(flag, first, ncell, count, keep, order);
  unsigned sel = order.pairs ? pair_bits(bits) : bits;
  const int c = __popc(sel);
  int v = c;  // inclusive scan over the wave's lanes
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(v, d);
    if (lane >= d) v += y;
  }
  const int total = __shfl(v, 63);
  int *row = stage[wave];
  int k = v - c;
  if (order.pairs) {
    while (sel) {
      const int j = __builtin_ctz(sel);  // even: the pair's first cell
      sel &= sel - 1;
      row[k++] = int(uint32_t((first + j) >> 1) | (((bits >> j) & 3u) << 30));
    }
  } else if (sel && list_chunk_contiguous(order)) {
    const long base = list_order_cell(order, first);
    while (sel) {
      const int j = __builtin_ctz(sel);
      sel &= sel - 1;
      row[k++] = int(base + j);
    }
  } else {
    while (sel) {
      const int j = __builtin_ctz(sel);
      sel &= sel - 1;
      row[k++] = int(list_order_cell(order, first + j));
    }
  }
  if (lane == 0) wsum[wave] = total;
  __syncthreads();  // the LDS rows and the wave totals
  int pos = boff[blockIdx.x];
  for (int w = 0; w < wave; ++w) pos += wsum[w];
  for (int q = lane; q < total; q += 64) list[pos + q] = row[q];
}

// explicit instantiations used by the host dispatch (pf_api.hip)
#define PF_INST_CLASSIFY(T, P)                                                                 \
  template hipError_t launch_classify<T, P, 4, false>(hipStream_t, const T *, uint8_t *, uint8_t *, \
                                                      int, int, int, long, unsigned, unsigned *, bool);  \
  template hipError_t launch_classify<T, P, 6, false>(hipStream_t, const T *, uint8_t *, uint8_t *, \
                                                      int, int, int, long, unsigned, unsigned *, bool);  \
  template hipError_t launch_classify<T, P, 6, true>(hipStream_t, const T *, uint8_t *, uint8_t *,  \
                                                     int, int, int, long, unsigned, unsigned *, bool);
PF_INST_CLASSIFY(double, 2)
PF_INST_CLASSIFY(double, 3)
PF_INST_CLASSIFY(float, 2)
PF_INST_CLASSIFY(float, 3)
#define PF_INST_BULK(F, P, NF, NOF, BY, SPLIT)                                                  \
  template __global__ void diffuse_bulk_kernel<F, P, NF, NOF, BY, SPLIT>(                       \
      const F::C *, F::C *, const F::T *, const uint8_t *, const DiffConsts<F::A>, int, int, int, \
      int, int, int, int);
#define PF_INST_DIFF_ONE(F, P, NF, NOF)                                                           \
  template __global__ void diffuse_linear_kernel<F, P, NF, NOF, false>(                           \
      const F::C *, F::C *, const F::T *, const uint8_t *, const DiffConsts<F::A>, int, int, int); \
  template __global__ void diffuse_linear_kernel<F, P, NF, NOF, true>(                            \
      const F::C *, F::C *, const F::T *, const uint8_t *, const DiffConsts<F::A>, int, int, int); \
  template __global__ void diffuse_general_kernel<F, P, NF, NOF>(                                 \
      const F::C *, F::C *, const F::T *, const int *, const int *, const DiffConsts<F::A>, int,  \
      int, int);                                                                                \
  template __global__ void diffuse_retable_kernel<F, P, NF, NOF>(                                 \
      const F::C *, F::C *, const F::T *, const uint8_t *, const DiffConsts<F::A>,              \
      const DiffConsts<F::A>, int, int, int);                                                   \
  PF_INST_BULK(F, P, NF, NOF, 4, false)                                                         \
  PF_INST_BULK(F, P, NF, NOF, 8, false)                                                         \
  PF_INST_BULK(F, P, NF, NOF, 16, false)                                                        \
  PF_INST_BULK(F, P, NF, NOF, 4, true)                                                          \
  PF_INST_BULK(F, P, NF, NOF, 8, true)                                                          \
  PF_INST_BULK(F, P, NF, NOF, 16, true)                                                         \
  template __global__ void diffuse_kernel<F, P, NF, NOF, kDiffTY>(                                \
      const F::C *, F::C *, const F::T *, const DiffConsts<F::A>, int, int, int, int, int, int, \
      int);
#define PF_INST_VEC(F, P, V)                                                                       \
  template __global__ void diffuse_bulk_vec_kernel<F, P, 4, false, 4, V>(                         \
      const F::C *, F::C *, const uint8_t *, const DiffConsts<F::A>, int, int, int, int, int, int, \
      int);                                                                                      \
  template __global__ void diffuse_bulk_vec_kernel<F, P, 6, false, 4, V>(                         \
      const F::C *, F::C *, const uint8_t *, const DiffConsts<F::A>, int, int, int, int, int, int, \
      int);                                                                                      \
  template __global__ void diffuse_bulk_vec_kernel<F, P, 6, true, 4, V>(                          \
      const F::C *, F::C *, const uint8_t *, const DiffConsts<F::A>, int, int, int, int, int, int, \
      int);
#define PF_INST_MARCH_ONE(F, P, NF, NOF, BY, V, R)                                                 \
  template __global__ void diffuse_bulk_march_kernel<F, P, NF, NOF, BY, V, R>(                     \
      const F::C *, F::C *, const unsigned *, const DiffConsts<F::A>, int, int, int, int, int,    \
      int, int);
#define PF_INST_MARCH_NF(F, P, BY, V, R)          \
  PF_INST_MARCH_ONE(F, P, 4, false, BY, V, R)     \
  PF_INST_MARCH_ONE(F, P, 6, false, BY, V, R)     \
  PF_INST_MARCH_ONE(F, P, 6, true, BY, V, R)
#define PF_INST_MARCH(F, P, V)          \
  PF_INST_MARCH_NF(F, P, 8, V, 1)       \
  PF_INST_MARCH_NF(F, P, 16, V, 1)      \
  PF_INST_MARCH_NF(F, P, 32, V, 2)
PF_INST_MARCH(FormF64, 2, 2)
PF_INST_MARCH(FormF64, 3, 2)
PF_INST_MARCH(FormMixed, 2, 4)
PF_INST_MARCH(FormMixed, 3, 4)
PF_INST_VEC(FormF64, 2, 2)
PF_INST_VEC(FormF64, 3, 2)
PF_INST_VEC(FormMixed, 2, 4)
PF_INST_VEC(FormMixed, 3, 4)
#define PF_INST_DIFF(F, P)               \
  PF_INST_DIFF_ONE(F, P, 4, false)       \
  PF_INST_DIFF_ONE(F, P, 6, false)       \
  PF_INST_DIFF_ONE(F, P, 6, true)
PF_INST_DIFF(FormF64, 2)
PF_INST_DIFF(FormF64, 3)
PF_INST_DIFF(FormMixed, 2)
PF_INST_DIFF(FormMixed, 3)

}  // namespace pfamd
