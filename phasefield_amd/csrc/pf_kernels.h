// pf_kernels.h — declarations of the gfx950 kernels (definitions and explicit instantiations in
// pf_kernels.hip) for the host dispatch in pf_api.hip.
#pragma once
#include "pf_device.h"

namespace pfamd {

constexpr int kDiffTY = 8;     // rows of the diffusion tile (64 x kDiffTY threads)
constexpr int kSweepBS = 128;  // threads per phase-field sweep block (one cell each)

// MODE (LIST, K = 4, FP64 form): 0 both paths of the union; 1 the register path only, a cell whose union
// outgrows it appended to olist (its count in olist[ncell], zeroed before); 2 the general path
// only, over list = olist, count = olist + ncell
template <class F, int K, int P, int NF, bool NOFLUX, bool ISO, bool LIST, int MODE = 0>
__global__ void pf_sweep_kernel(const uint16_t *__restrict__ id1,
                                const typename F::T *__restrict__ phi1,
                                uint16_t *__restrict__ id2, typename F::T *__restrict__ phi2,
                                typename F::T *__restrict__ pfrac,
                                const typename F::C *__restrict__ conc0,
                                const PFConsts<typename F::A> cs, const int nx, const int ny,
                                const int nz, const long ncell, const int *__restrict__ list,
                                const int *__restrict__ count, int *__restrict__ queue,
                                uint8_t *__restrict__ sflag, uint8_t *__restrict__ unit,
                                const unsigned umask, int *__restrict__ olist);
// count[0..3] set to zero (one workgroup of 64)
__global__ void zero_counts_kernel(unsigned *__restrict__ count);
// the LIST walk's per-XCD counters (queue[32 * xcd], xcd < 8) set to zero: one workgroup of 8
__global__ void zero_queue_kernel(int *__restrict__ queue);

// copy cells of the sweep written directly, the others flagged 0 for the listed sweep
template <class F, int K, int NF, bool NOFLUX>
__global__ void sweep_classify_kernel(const uint16_t *__restrict__ id1,
                                      const typename F::T *__restrict__ phi1,
                                      uint16_t *__restrict__ id2, typename F::T *__restrict__ phi2,
                                      uint8_t *__restrict__ sflag, const int history,
                                      const double cutoff, const float cutoff_f, const int nx,
                                      const int ny, const int nz, const long ncell);

template <class F, int P, int NF, bool NOFLUX, int TY>
__global__ void diffuse_kernel(const typename F::C *__restrict__ cin,
                               typename F::C *__restrict__ cout,
                               const typename F::T *__restrict__ pfrac,
                               const DiffConsts<typename F::A> dc, const int nx, const int ny,
                               const int nz, const int zchunk, const int tiles_x,
                               const int tiles_y, const int nblocks);

constexpr int kMarchBY = 8;     // default rows of the bulk diffusion tile (64 x kMarchBY lanes)
constexpr int kDiffUnroll = 4;  // planes per iteration of a column (loads issued together)

// SPLIT: cells with flag 0 are left to diffuse_general_kernel (list built per sweep)
template <class F, int P, int NF, bool NOFLUX, bool SPLIT>
__global__ void diffuse_linear_kernel(const typename F::C *__restrict__ cin,
                                      typename F::C *__restrict__ cout,
                                      const typename F::T *__restrict__ pfrac,
                                      const uint8_t *__restrict__ flag,
                                      const DiffConsts<typename F::A> dc, const int nx,
                                      const int ny, const int nz);

// bulk-only sub-step with V consecutive x cells per lane (nx % V == 0); full-formula cells are
// overwritten by diffuse_general_kernel afterwards
template <class F, int P, int NF, bool NOFLUX, int BY, int V>
__global__ void diffuse_bulk_vec_kernel(const typename F::C *__restrict__ cin,
                                        typename F::C *__restrict__ cout,
                                        const uint8_t *__restrict__ flag,
                                        const DiffConsts<typename F::A> dc, const int nx,
                                        const int ny, const int nz, const int zchunk,
                                        const int tiles_x, const int tiles_y, const int nblocks);

// Bulk cell counts per phase of the last classification (count[1..3]; count[0] unused): the phase
// with the most bulk cells, the first of equal counts, 1 when there are none.  The plane-marching
// kernel applies that phase's bulk update to every cell; every other cell is listed.
__host__ __device__ inline int bulk_phase(const unsigned *count) {
  int p = 1;
  unsigned best = count[1];
  if (count[2] > best) {
    p = 2;
    best = count[2];
  }
  if (count[3] > best) p = 3;
  return p;
}

// plane-marching kernel: the bulk update of phase bulk_phase(count) applied to every cell, without
// reading any per-cell flag; the cells it does not hold for (flag != that phase) are overwritten by
// diffuse_general_kernel afterwards.  Needs nx % V == 0.
template <class F, int P, int NF, bool NOFLUX, int BY, int V, int R>
__global__ void diffuse_bulk_march_kernel(const typename F::C *__restrict__ cin,
                                          typename F::C *__restrict__ cout,
                                          const unsigned *__restrict__ count,
                                          const DiffConsts<typename F::A> dc, const int nx,
                                          const int ny, const int nz, const int zchunk,
                                          const int tiles_x, const int tiles_y, const int nblocks);

template <class F, int P, int NF, bool NOFLUX>
__global__ void diffuse_general_kernel(const typename F::C *__restrict__ cin,
                                       typename F::C *__restrict__ cout,
                                       const typename F::T *__restrict__ pfrac,
                                       const int *__restrict__ list, const int *__restrict__ count,
                                       const DiffConsts<typename F::A> dc, const int nx,
                                       const int ny, const int nz);

// The first sub-step after a sweep whose partition tables differ from those the stored
// concentrations were last partitioned with: the full formula for every cell, each cell's partition
// taken with the previous tables (dc_prev) if the sweep copied it (bit 0 of copied[cell]), else with the
// current ones.  Diffusivities, factor and rounding key come from dc.
template <class F, int P, int NF, bool NOFLUX>
__global__ void diffuse_retable_kernel(const typename F::C *__restrict__ cin,
                                       typename F::C *__restrict__ cout,
                                       const typename F::T *__restrict__ pfrac,
                                       const uint8_t *__restrict__ copied,
                                       const DiffConsts<typename F::A> dc,
                                       const DiffConsts<typename F::A> dc_prev, const int nx,
                                       const int ny, const int nz);

// List of the cells a full-formula pass must cover: flag == 0, or, with a non-null `count`,
// flag != bulk_phase(count) (the cells the plane-marching kernel gets wrong).
constexpr int kListChunk = 16;  // 256-cell chunks per block of the full-formula list build
// Order of the list.  rows == 0: storage order.  rows > 0: row tiles — `rows` full-width rows
// (the last tile may hold fewer), a tile's planes one after another, the tiles in y order — so a
// walk over the list meets a cell's -z and +z neighbour rows one tile-plane (rows * nx cells)
// before and after the cell instead of one whole plane.  Each listed cell's update reads only the
// source buffers, so the order changes which rows share the cache, never a result.
// pairs (storage order only): one entry per byte pair of cells (2j, 2j + 1) holding a listed
// cell, j | mask << 30 with mask bit b set when cell 2j + b is listed (cell indices stay below
// 2^31).  The mixed form's 4-bit compensation codes of the two cells share a byte, so a pass that
// gives a pair to one thread writes that byte with a plain store.
struct ListOrder {
  int nx = 0, ny = 0, nz = 0, rows = 0;
  bool pairs = false;
};
// storage index of list-order position p (0 <= p < nx * ny * nz)
// (every quantity is below 2^31, the limit on a domain's cells, so the divisions are 32-bit: a
// 64-bit division is several times longer on the device)
__host__ __device__ inline long list_order_cell(const ListOrder &o, long p) {
  if (o.rows == 0) return p;
  const uint32_t up = uint32_t(p), nx = uint32_t(o.nx), rows = uint32_t(o.rows);
  const uint32_t column = rows * nx * uint32_t(o.nz);  // cells of one full tile
  const uint32_t ty = up / column;
  const uint32_t left = uint32_t(o.ny) - ty * rows;  // rows of the last tile may be fewer
  const uint32_t h = left < rows ? left : rows;
  const uint32_t q = up - ty * column;
  const uint32_t plane = h * nx;
  const uint32_t z = q / plane;
  const uint32_t r = q - z * plane;
  const uint32_t yy = r / nx;
  const uint32_t x = r - yy * nx;
  return (long(z) * o.ny + long(ty * rows + yy)) * o.nx + x;
}
// The split sweep's per-cell byte (sflag).  The classification writes 0 for a cell the listed sweep
// must update, nonzero for a copy cell; the listed sweep then rewrites its cells' bytes.  Besides
// bit 0 (copied by this sweep) the byte summarises the cell's new own slots for the next sweep's
// classification, which then reads no phi: bit 1 when at most one slot is above the cutoff, bit 2
// when exactly one is, bits 3-5 that slot's index.
__host__ __device__ inline uint8_t slot_summary(int above, int slot) {
  return above > 1 ? uint8_t(0) : uint8_t(2u | (above == 1 ? 4u | (unsigned(slot) << 3) : 0u));
}
// 0x80 in each byte of x that is 0, 0 in the others: the low seven bits of a byte plus 0x7f set its
// top bit unless they are all 0, and no sum carries into the next byte
__host__ __device__ inline uint32_t zero_bytes(uint32_t x) {
  const uint32_t t = (x & 0x7f7f7f7fu) + 0x7f7f7f7fu;
  return ~(t | x | 0x7f7f7f7fu);
}
// Whether a neighbour's row of four 16-bit ids (slot 0 in the low bits) breaks a single-entry
// active-set union of grain G: its ids are read in slot order up to the first 0 and each one read
// must be G, so slot s breaks it when it is neither 0 nor G and every slot before it is a live G
// (G != 0).  Branch-free: the wavefront's cells evaluate it together.
__host__ __device__ inline bool id_row_breaks(uint64_t w, uint32_t G) {
  const bool live_g = G != 0u;
  const uint32_t lo = uint32_t(w), hi = uint32_t(w >> 32);
  const uint32_t i0 = lo & 0xffffu, i1 = lo >> 16, i2 = hi & 0xffffu, i3 = hi >> 16;
  const bool e0 = live_g & (i0 == G), e1 = live_g & (i1 == G), e2 = live_g & (i2 == G);
  const bool b0 = (i0 != G) & (i0 != 0u), b1 = (i1 != G) & (i1 != 0u);
  const bool b2 = (i2 != G) & (i2 != 0u), b3 = (i3 != G) & (i3 != 0u);
  return b0 | (e0 & b1) | (e0 & e1 & b2) | (e0 & e1 & e2 & b3);
}
// `cells` (optional): the count kernel's per-block counts of listed cells, for lists whose
// entries are not cells (pairs); general_scan_kernel over them gives the total
__global__ void general_count_kernel(const uint8_t *__restrict__ flag,
                                     const unsigned *__restrict__ count, long ncell,
                                     int *__restrict__ bcount, const ListOrder order,
                                     int *__restrict__ cells);
__global__ void general_scan_kernel(int *__restrict__ bcount, int nb,
                                    int *__restrict__ total_out);
__global__ void general_scatter_kernel(const uint8_t *__restrict__ flag,
                                       const unsigned *__restrict__ count, long ncell,
                                       const int *__restrict__ boff, int *__restrict__ list,
                                       const ListOrder order);

// Bulk flags of a value_phase field into `flag` and the bulk cells per phase into count[1..3]
// (two kernels; `unit` is an N-byte array of unit-phase bytes, computed here unless `unit_ready`
// says it already describes this value_phase and mask)
template <class T, int P, int NF, bool NOFLUX>
hipError_t launch_classify(hipStream_t st, const T *pfrac, uint8_t *unit, uint8_t *flag, int nx,
                           int ny, int nz, long ncell, unsigned mask, unsigned *count,
                           bool unit_ready = false);

template <class F, int P, int NF, bool NOFLUX, int BY, bool SPLIT>
__global__ void diffuse_bulk_kernel(const typename F::C *__restrict__ cin,
                                    typename F::C *__restrict__ cout,
                                    const typename F::T *__restrict__ pfrac,
                                    const uint8_t *__restrict__ flag,
                                    const DiffConsts<typename F::A> dc, const int nx, const int ny,
                                    const int nz, const int zchunk, const int tiles_x,
                                    const int tiles_y, const int nblocks);

struct RedPartial {
  double mass[2], pf[3], cmax[2], cmin[2], serr;
  long long nif, nnf;
};

template <class F, int K, int P>
__global__ void reduce_stage1(const uint16_t *__restrict__ id,
                              const typename F::T *__restrict__ phi,
                              const typename F::T *__restrict__ pfrac,
                              const typename F::C *__restrict__ c0,
                              const typename F::C *__restrict__ c1, int ns, long ncell,
                              long per_block, RedPartial *__restrict__ out);
__global__ void reduce_stage2(RedPartial *__restrict__ part, int nparts);

template <int K, int P>
__global__ void voronoi_kernel(const double *__restrict__ seeds, int G, double reduction, int nx,
                               int ny, int nz, long ncell, uint16_t *__restrict__ id,
                               double *__restrict__ phi, double *__restrict__ pfrac);
template <int P>
__global__ void construct_conc_kernel(const double *__restrict__ pfrac, double c1, double c2,
                                      double c3, long ncell, double *__restrict__ c);
template <int K, int P>
__global__ void nucleation_flag_kernel(const uint16_t *__restrict__ id,
                                       const double *__restrict__ pfrac,
                                       const double *__restrict__ c0, const DiffConsts<double> dc,
                                       uint64_t seed_key, long ncell, float *__restrict__ flag,
                                       uint8_t *__restrict__ mark);
__global__ void argmax_stage1(const float *__restrict__ flag, long ncell, long per_block,
                              float *__restrict__ pv, long long *__restrict__ pi);
__global__ void argmax_stage2(float *__restrict__ pv, long long *__restrict__ pi, int nparts);
template <int K, int P>
__global__ void nucleus_paint_kernel(const long long *__restrict__ center, int code, float radius,
                                     int R, int nx, int ny, int nz, int cx_n, int cy_n, int cz_n,
                                     uint16_t *__restrict__ id, double *__restrict__ phi,
                                     double *__restrict__ pfrac, float *__restrict__ flag,
                                     uint8_t *__restrict__ mark);
template <class F>
__global__ void vti_fields_kernel(const uint16_t *__restrict__ id, int KI,
                                  const typename F::C *__restrict__ c, long ncell,
                                  int32_t *__restrict__ korn, float *__restrict__ conc);
__global__ void nucleus_conc_kernel(const uint8_t *__restrict__ mark, double value, long ncell,
                                    double *__restrict__ c);
__global__ void solute_noise_kernel(double mean, double sd, uint64_t seed_key, long ncell,
                                    double *__restrict__ c);
__global__ void convert_f64_to_f32(const double *__restrict__ src, float *__restrict__ dst,
                                   long n);
__global__ void convert_conc_f64_to_mixed(const double *__restrict__ src, float *__restrict__ dst,
                                          long n);

}  // namespace pfamd
