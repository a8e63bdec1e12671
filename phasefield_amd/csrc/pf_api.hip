// pf_api.hip — the C-ABI (include/phasefield_amd.h): domain handles, host<->device layout
// conversion with the 16-bit grain-code remap, step orchestration on one HIP stream, FP64
// reductions, checkpoints, device-side seeded setup, and the reference's three-function
// accelerator API (gpu_solver.h:18-23) as a thin shim over the same kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/phasefield_amd.h"
#include "pf_handle_access.h"
#include "pf_kernels.h"

using namespace pfamd;

namespace {

thread_local std::string g_create_error;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct pf_handle {
  pf_desc d{};
  long N = 0;
  int KI = 4, NF = 6, dims = 3;
  size_t tsz = 8;   // bytes of one phi/value_phase element
  size_t cbuf = 0;  // bytes of one concentration buffer (FP64: 8 N; mixed: mixed_conc_bytes(N))
  hipStream_t stream = nullptr;
  uint16_t *id[2] = {nullptr, nullptr};
  void *phi[2] = {nullptr, nullptr};
  void *pfrac = nullptr;
  void *conc[PF_MAX_SOLUTE][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int cur = 0;                       // active slot buffer
  int ccur[PF_MAX_SOLUTE] = {0, 0};  // active concentration buffer per solute
  // sub-steps applied per solute since the state was set: the mixed form's rounding keys
  uint32_t csub[PF_MAX_SOLUTE] = {0, 0};
  RedPartial *red = nullptr;
  int red_blocks = 0;
  pf_params p{};
  bool have_params = false, have_state = false;
  std::vector<int32_t> code_to_id[4];  // per phase: serial -> reference grain id
  std::string err;
  // timing of the dominant launches
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[3];
  std::vector<int> ev_count[3];  // launches covered by each event pair
  size_t ev_used[3] = {0, 0, 0};
  double t_ms[3] = {0, 0, 0};
  int64_t t_n[3] = {0, 0, 0};
  int diff_zchunk_override = 0;
  int diff_variant = 0;  // 0: bulk-aware column kernel, 1: LDS-tiled full-formula kernel
  uint8_t *flag = nullptr;   // per-cell bulk phase (0 = needs the full formula)
  uint8_t *unit = nullptr;   // scratch of the classification: per-cell unit-vector phase
  unsigned *bulk_count = nullptr;  // [4] bulk cells per phase of the last classification
  bool flag_valid = false;   // flags describe the current value_phase and stencil
  bool unit_valid = false;   // unit bytes describe the current value_phase and bulk_mask
  int flag_nf = 0;           // stencil faces the flags were computed for
  int diff_by = -1;          // rows of the bulk diffusion tile (0: storage order, -1: by size)
  bool tune_log = false;     // print the plane-marching shape timings (PFAMD_TUNE_LOG=1)
  bool diff_split = true;    // full-formula cells from a per-sweep list in their own kernel
  bool diff_vec = true;      // bulk-only sub-step with 16-byte runs per lane (when nx allows)
  bool diff_march = true;    // ... in its plane-marching form
  int march_by = 0;          // its tile rows and z-chunk, picked by timing at first use
  int march_zc = 0;
  int *list_len = nullptr;   // [0]: cells the last sweep listed, [1]: cells the diffusion lists
  int *ccount = nullptr;     // pair lists: listed cells per list block ([nblocks + 1])
  bool ccount_pending = false;  // ccount not yet summed into list_len[1]
  bool time_march = false;   // time the next plane-marching launch (timing kind 2)
  int *glist = nullptr;      // indices of the listed cells (sweep: row tiles, diffusion: storage)
  int *gblock = nullptr;     // per-list-block offsets into glist; [nblocks] = list length
  bool sweep_split = true;   // copy cells classified and written apart from the listed sweep
  uint8_t *sflag = nullptr;  // per cell: bit 0 copied by the last sweep, bits 1-5 slot_summary
  int sweep_rows = -1;       // row tiles of the sweep's list (0: storage order, -1: by size)
  bool sweep_queue = true;   // listed sweep: per-XCD work counters instead of a fixed stride
  int *squeue = nullptr;     // those counters, 128 B apart
  int *olist = nullptr;      // K = 4 split sweep: cells for the general path; [N]: their count
  int num_cu = 0;            // compute units of the handle's device
  std::map<const void *, int> occupancy;  // resident workgroups per CU, by kernel, on that device
  bool sweep_hist = false;   // sflag describes the last sweep of the current buffers
  unsigned bulk_mask = 0;    // bit p: phase p may take the bulk-cell reduction
  // The reference stores the partitioned concentrations (value_conc1) and reads them until they are
  // recomputed, so after a parameter change it still reads partitions of the previous tables: in
  // the sweep's driving force, and in the first sub-step for the cells the sweep copied.  The
  // tables the stored state was last partitioned with, per solute:
  double prev_pk[PF_MAX_SOLUTE][3][3] = {}, prev_pc[PF_MAX_SOLUTE][3][3] = {};
  bool have_prev = false;
  // next sub-step of solute s mixes two tables: 1 = the cells the last sweep copied keep the
  // previous ones, 2 = every cell does (no sweep since the change)
  int retable[PF_MAX_SOLUTE] = {0, 0};
  // asynchronous VTI dump: device fields -> pinned host copy (dump stream) -> writer thread
  void *dump_dev = nullptr;      // korn [N] int32, conc [N] float
  void *dump_host = nullptr;     // pinned copy of dump_dev
  hipStream_t dump_stream = nullptr;
  hipEvent_t dump_ev[2] = {nullptr, nullptr};
  std::thread dump_thread;
  bool dump_busy = false;
  int dump_status = PF_OK;
  std::string dump_error;
};

namespace {

int fail(pf_handle *h, int code, const std::string &msg) {
  if (h)
    h->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define HIPCHK(h, call)                                                                      \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return fail((h), e_ == hipErrorOutOfMemory ? PF_ERR_OOM : PF_ERR_HIP,                  \
                  std::string(#call) + ": " + hipGetErrorString(e_));                        \
  } while (0)

bool is_mixed(const pf_handle *h) { return h->d.precision == PF_MIXED; }

// ---- constants ----------------------------------------------------------------------------------
template <class A> PFConsts<A> make_pf_consts(const pf_handle *h) {
  const pf_params &p = h->p;
  PFConsts<A> c;
  std::memset(&c, 0, sizeof(c));
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const bool in = i > 0 && j > 0;
      c.M[i][j] = A(in ? p.mobility[i - 1][j - 1] : 0.0);
      c.sig[i][j] = A(in ? p.sigma[i - 1][j - 1] : 0.0);
      c.ent[i][j] = A(in ? p.entropy[i - 1][j - 1] : 0.0);
      c.eqc[i][j] = A(in ? p.eq_conc[i - 1][j - 1] : 0.0);
      c.slope[i][j] = A(in ? p.eq_slope[i - 1][j - 1] : 1.0);
      // local_conc of Driving_force: the stored partition, i.e. of the tables last used
      c.pk[i][j] = A(in ? (h->have_prev ? h->prev_pk[0] : p.part_k[0])[i - 1][j - 1] : 0.0);
      c.pc[i][j] = A(in ? (h->have_prev ? h->prev_pc[0] : p.part_c[0])[i - 1][j - 1] : 0.0);
      c.pkS[i][j] = A(in ? p.part_k[1][i - 1][j - 1] : 0.0);
      c.pcS[i][j] = A(in ? p.part_c[1][i - 1][j - 1] : 0.0);
    }
  // solver.cpp:226-228
  c.f1 = A(1. / (p.dx * p.dx));
  c.f2 = A(kPi * kPi / (2. * p.eta * p.eta));
  c.f3 = A(kPi / p.eta);
  c.dt = p.pfm_dt;
  c.cutoff = p.cutoff;
  return c;
}

template <class A> DiffConsts<A> make_diff_consts(const pf_params &p, int s, double factor) {
  DiffConsts<A> d;
  std::memset(&d, 0, sizeof(d));
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      d.pk[i][j] = A(p.part_k[s][i][j]);
      d.pc[i][j] = A(p.part_c[s][i][j]);
    }
    d.D[i] = A(p.diffusivity[s][i]);
  }
  d.r = A(factor);
  return d;
}

// Bit p (1-based phase) when every solute's partition table has k(p,p) == 1 and c(p,p) == 0 in
// the form's arithmetic type: the condition under which the bulk-cell reduction of the diffusion
// sub-step is exact.
void update_bulk_mask(pf_handle *h) {
  unsigned m = 0;
  for (int p = 0; p < h->d.nphase; ++p) {
    bool ok = true;
    for (int s = 0; s < h->d.nsolute; ++s) {
      const double k = h->p.part_k[s][p][p], c = h->p.part_c[s][p][p];
      if (is_mixed(h))
        ok = ok && float(k) == 1.0f && float(c) == 0.0f;
      else
        ok = ok && k == 1.0 && c == 0.0;
    }
    if (ok) m |= 1u << (p + 1);
  }
  if (m != h->bulk_mask) h->unit_valid = false;  // the unit bytes test the mask's phases
  h->bulk_mask = m;
  h->flag_valid = false;
}

// the stored state now counts as partitioned with the current tables of solute s (all when s < 0)
void tables_current(pf_handle *h, int s = -1) {
  for (int q = 0; q < PF_MAX_SOLUTE; ++q)
    if (s < 0 || s == q) {
      std::memcpy(h->prev_pk[q], h->p.part_k[q], sizeof(h->prev_pk[q]));
      std::memcpy(h->prev_pc[q], h->p.part_c[q], sizeof(h->prev_pc[q]));
      h->retable[q] = 0;
    }
  if (s < 0) h->have_prev = h->have_params;
}
// the state was replaced: partitioned with the current tables, or with the first ones set
void tables_state_set(pf_handle *h) {
  tables_current(h);
  h->have_prev = h->have_params;
}
// whether solute s's previous and current tables differ in the form's arithmetic type
bool tables_differ(const pf_handle *h, int s) {
  if (!h->have_prev) return false;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double a = h->prev_pk[s][i][j], b = h->p.part_k[s][i][j];
      const double c = h->prev_pc[s][i][j], d = h->p.part_c[s][i][j];
      if (is_mixed(h) ? (float(a) != float(b) || float(c) != float(d)) : (a != b || c != d))
        return true;
    }
  return false;
}

// ---- launch helpers -----------------------------------------------------------------------------
inline unsigned blocks_for(long n, int bs = 256) { return unsigned((n + bs - 1) / bs); }
// the plane-marching launch is timed on every kMarchTimingEvery-th sub-step of a batch (timing
// kind 2): a sample of its launches, with few events between the kernels of the step
constexpr int kMarchTimingEvery = 8;
// workgroups of the full-formula list build (kListChunk * 256 cells each)
inline unsigned list_blocks(long n) { return blocks_for(n, 256 * kListChunk); }

// One event pair on the handle's stream around `count` back-to-back launches of one kind (events
// between every launch would add their own cost to what they measure).
struct Timer {
  pf_handle *h;
  int kind;
  hipEvent_t b = nullptr, e = nullptr;
  Timer(pf_handle *hh, int k, int count = 1) : h(hh), kind(k) {
    if (!h->timing || count <= 0) return;
    auto &pool = h->ev[kind];
    if (h->ev_used[kind] == pool.size()) {
      hipEvent_t x, y;
      if (hipEventCreate(&x) != hipSuccess || hipEventCreate(&y) != hipSuccess) return;
      pool.emplace_back(x, y);
      h->ev_count[kind].push_back(0);
    }
    b = pool[h->ev_used[kind]].first;
    e = pool[h->ev_used[kind]].second;
    h->ev_count[kind][h->ev_used[kind]] = count;
    h->ev_used[kind]++;
    hipEventRecord(b, h->stream);
  }
  ~Timer() {
    if (e) hipEventRecord(e, h->stream);
  }
};

// Order of the sweep's list: row tiles of about 32768 cells per tile-plane (nx = 512: 64 rows), so
// that the -z/+z neighbour rows of the cells a walk is on were touched a tile-plane earlier and are
// still in L2, and a tile's -y/+y edge rows, read again by the next tile, are few; storage order for
// one plane (2D) or when PFAMD_SWEEP_ROWS=0.  Measured at 512^3 with the per-XCD walk (read bytes
// per listed cell, mixed / FP64): 8 rows 83 / 118, 32 rows 76 / 108, 64 rows 75 / 111,
// 128 rows 90 / 179, storage order 150 / 211.
inline ListOrder sweep_list_order(const pf_handle *h) {
  int rows = h->sweep_rows;
  if (rows < 0) rows = std::max(1, 32768 / std::max(h->d.nx, 1));
  if (h->d.nz == 1) rows = 0;
  return ListOrder{h->d.nx, h->d.ny, h->d.nz, std::min(rows, h->d.ny)};
}

// the largest float not above the cutoff: for a float v, v > it exactly when double(v) > cutoff
inline float cutoff_below(double cutoff) {
  float f = float(cutoff);
  if (double(f) > cutoff) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}

template <class F, int K, int P, int NF, bool NOF, bool ISO>
hipError_t launch_sweep(pf_handle *h, const PFConsts<typename F::A> &cs) {
  using T = typename F::T;
  using C = typename F::C;
  const int a = h->cur, b = 1 - h->cur;
  const T *phi1 = static_cast<const T *>(h->phi[a]);
  T *phi2 = static_cast<T *>(h->phi[b]);
  const C *c0 = static_cast<const C *>(h->conc[0][h->ccur[0]]);
  // the split form records which cells it copied, which a change of partition tables needs
  if (!ISO && (h->sweep_split || h->retable[0] || h->retable[1])) {
    // copy cells first (written, or skipped when the destination already holds them), then the
    // full update over the list of the others
    const unsigned nb = blocks_for(h->N), nl = list_blocks(h->N);
    hipLaunchKernelGGL((sweep_classify_kernel<F, K, NF, NOF>), dim3(nb), dim3(256), 0, h->stream,
                       h->id[a], phi1, h->id[b], phi2, h->sflag, h->sweep_hist ? 1 : 0,
                       h->p.cutoff, cutoff_below(h->p.cutoff), h->d.nx, h->d.ny, h->d.nz, h->N);
    const ListOrder order = sweep_list_order(h);
    hipLaunchKernelGGL(general_count_kernel, dim3(nl), dim3(256), 0, h->stream, h->sflag,
                       static_cast<const unsigned *>(nullptr), h->N, h->gblock, order,
                       static_cast<int *>(nullptr));
    hipLaunchKernelGGL(general_scan_kernel, dim3(1), dim3(1024), 0, h->stream, h->gblock, int(nl),
                       h->list_len);
    hipLaunchKernelGGL(general_scatter_kernel, dim3(nl), dim3(256), 0, h->stream, h->sflag,
                       static_cast<const unsigned *>(nullptr), h->N, h->gblock, h->glist, order);
    // Every workgroup resident at once (the device's CUs times the kernel's occupancy), each
    // XCD's wavefronts taking the next entries of its eighth of the list from a counter: the
    // entries in process stay one compact front (512^3: 190 -> 83 B read per listed cell, mixed).
    // With a fixed stride the workgroups drift apart as their cells' work differs, and the front
    // outgrows the XCD's L2.
    // FP64, K = 4: the register path of unions up to four entries is a kernel of its own (no LDS
    // arrays, 81 instead of 106 VGPRs, 6 instead of 486 spilled SGPRs), and the few cells whose
    // union outgrows it are appended to an overflow list that the general path then takes (512^3:
    // 2128 -> 1944 + 51 us).  The mixed form's listed sweep is VALU-bound at 93 % busy, so the
    // extra occupancy buys nothing there and the second launch costs 48 us: it keeps both paths in
    // one kernel.
    constexpr int MODE = (K == 4 && !F::kMixed) ? 1 : 0;
    const void *kfn =
        reinterpret_cast<const void *>(pf_sweep_kernel<F, K, P, NF, NOF, false, true, MODE>);
    int &per_cu = h->occupancy[kfn];
    if (per_cu == 0) {
      int n = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
              &n, pf_sweep_kernel<F, K, P, NF, NOF, false, true, MODE>, kSweepBS, 0) !=
              hipSuccess ||
          n < 1)
        n = 1;
      per_cu = n;
    }
    const unsigned ng = unsigned(std::min<long>(blocks_for(h->N, kSweepBS),
                                                long(std::max(h->num_cu, 1)) * per_cu));
    if (h->sweep_queue)
      hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(8), 0, h->stream, h->squeue);
    if (MODE == 1)  // the overflow count, olist[N]
      hipLaunchKernelGGL(zero_counts_kernel, dim3(1), dim3(64), 0, h->stream,
                         reinterpret_cast<unsigned *>(h->olist + h->N));
    hipLaunchKernelGGL((pf_sweep_kernel<F, K, P, NF, NOF, false, true, MODE>), dim3(ng),
                       dim3(kSweepBS), 0, h->stream, h->id[a], phi1, h->id[b], phi2,
                       static_cast<T *>(h->pfrac), c0, cs, h->d.nx, h->d.ny, h->d.nz, h->N, h->glist,
                       h->gblock + nl, h->sweep_queue ? h->squeue : static_cast<int *>(nullptr),
                       h->sflag, h->unit, h->bulk_mask, h->olist);
    if (MODE == 1)
      hipLaunchKernelGGL((pf_sweep_kernel<F, K, P, NF, NOF, false, true, 2>),
                         dim3(unsigned(std::max(h->num_cu, 1))), dim3(kSweepBS), 0, h->stream,
                         h->id[a], phi1, h->id[b], phi2, static_cast<T *>(h->pfrac), c0, cs,
                         h->d.nx, h->d.ny, h->d.nz, h->N, h->olist, h->olist + h->N,
                         static_cast<int *>(nullptr), h->sflag, h->unit, h->bulk_mask,
                         static_cast<int *>(nullptr));
    h->sweep_hist = true;
  } else {
    hipLaunchKernelGGL((pf_sweep_kernel<F, K, P, NF, NOF, ISO, false>),
                       dim3(blocks_for(h->N, kSweepBS)), dim3(kSweepBS), 0, h->stream, h->id[a],
                       phi1, h->id[b], phi2, static_cast<T *>(h->pfrac), c0, cs, h->d.nx, h->d.ny,
                       h->d.nz, h->N, nullptr, nullptr, nullptr, nullptr, nullptr, 0u, nullptr);
    h->sweep_hist = false;
    h->unit_valid = false;  // value_phase rewritten without its unit bytes
  }
  // (the split sweep keeps the unit bytes current: copy cells keep their value_phase, the listed
  // sweep writes the bytes of the cells it updates)
  h->cur = b;
  h->flag_valid = false;
  return hipGetLastError();
}

template <class F, int K, int P, bool ISO>
hipError_t dispatch_sweep_nf(pf_handle *h, const PFConsts<typename F::A> &cs) {
  if (h->NF == 4) return launch_sweep<F, K, P, 4, false, ISO>(h, cs);
  if (h->d.bc_z == PF_BC_NOFLUX) return launch_sweep<F, K, P, 6, true, ISO>(h, cs);
  return launch_sweep<F, K, P, 6, false, ISO>(h, cs);
}

template <class F, bool ISO>
hipError_t dispatch_sweep(pf_handle *h, const PFConsts<typename F::A> &cs) {
  const int K = h->d.depth, P = h->d.nphase;
  if (K == 4 && P == 2) return dispatch_sweep_nf<F, 4, 2, ISO>(h, cs);
  if (K == 4 && P == 3) return dispatch_sweep_nf<F, 4, 3, ISO>(h, cs);
  if (K == 6 && P == 2) return dispatch_sweep_nf<F, 6, 2, ISO>(h, cs);
  return dispatch_sweep_nf<F, 6, 3, ISO>(h, cs);
}

struct DiffGeom {
  int tiles_x, tiles_y, zchunk, nblocks;
};

// rows: tile height; target: workgroups to aim for (z is split into chunks until reached);
// cols: x cells per tile row
DiffGeom diff_geom(const pf_handle *h, int rows, long target, int min_chunk, int cols = 64) {
  DiffGeom g;
  g.tiles_x = (h->d.nx + cols - 1) / cols;
  g.tiles_y = (h->d.ny + rows - 1) / rows;
  const int nz = h->d.nz;
  int zc;
  if (h->NF != 6) {
    zc = nz;  // planes are independent; one chunk walks all of them
  } else if (h->diff_zchunk_override > 0) {
    zc = std::min(h->diff_zchunk_override, nz);
  } else {
    const long tiles = long(g.tiles_x) * g.tiles_y;
    long nzc = std::max(1L, std::min<long>(nz, (target + tiles - 1) / tiles));
    zc = int((nz + nzc - 1) / nzc);
    zc = std::max(zc, std::min(nz, min_chunk));
  }
  g.zchunk = zc;
  g.nblocks = g.tiles_x * g.tiles_y * ((nz + zc - 1) / zc);
  return g;
}

// The plane-marching bulk kernel.  Its speed depends on how many workgroups are resident per CU
// and how long each one marches (the best shape differs between the forms and sizes), which no
// closed rule captured.  So at the first sub-step of a handle each candidate tile height (8, 16
// rows, one per wavefront) and z-chunk (8..256 planes) is launched and timed into the output
// buffer.  Every candidate writes the same bits (the listed pass rewrites the cells the kernel
// gets wrong), so the choice changes speed only.  PFAMD_DIFF_BY (8, 16, or 32: two rows per
// wavefront, which halves the re-read edge rows but measured 6-7 % slower at 512^3 for its
// VGPRs) and PFAMD_DIFF_ZCHUNK pin either choice; PFAMD_TUNE_LOG=1 prints the timings.
template <class F, int P, int NF, bool NOF>
hipError_t launch_march(pf_handle *h, const DiffConsts<typename F::A> &dc,
                        const typename F::C *cin, typename F::C *cout) {
  constexpr int V = F::kMixed ? 4 : 2;
  const int nx = h->d.nx, ny = h->d.ny, nz = h->d.nz;
  auto launch = [&](int by, int zc) -> hipError_t {
    const int tiles_x = (nx + 64 * V - 1) / (64 * V), tiles_y = (ny + by - 1) / by;
    const long nblocks = long(tiles_x) * tiles_y * ((nz + zc - 1) / zc);
    if (nblocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    if (by == 8)
      hipLaunchKernelGGL((diffuse_bulk_march_kernel<F, P, NF, NOF, 8, V, 1>), dim3(nblocks),
                         dim3(64, 8), 0, h->stream, cin, cout, h->bulk_count, dc, nx, ny, nz, zc,
                         tiles_x, tiles_y, int(nblocks));
    else if (by == 16)
      hipLaunchKernelGGL((diffuse_bulk_march_kernel<F, P, NF, NOF, 16, V, 1>), dim3(nblocks),
                         dim3(64, 16), 0, h->stream, cin, cout, h->bulk_count, dc, nx, ny, nz, zc,
                         tiles_x, tiles_y, int(nblocks));
    else  // 32 rows: two per wavefront
      hipLaunchKernelGGL((diffuse_bulk_march_kernel<F, P, NF, NOF, 32, V, 2>), dim3(nblocks),
                         dim3(64, 16), 0, h->stream, cin, cout, h->bulk_count, dc, nx, ny, nz, zc,
                         tiles_x, tiles_y, int(nblocks));
    return hipGetLastError();
  };
  if (h->march_by == 0) {
    std::vector<int> bys, zcs;
    if (h->diff_by == 8 || h->diff_by == 16 || h->diff_by == 32) bys = {h->diff_by};
    else bys = {8, 16};
    if (h->diff_zchunk_override > 0) {
      zcs = {std::min(h->diff_zchunk_override, nz)};
    } else {
      for (int zc : {8, 16, 32, 64, 128, 256}) {
        const int c = std::min(zc, nz);
        if (std::find(zcs.begin(), zcs.end(), c) == zcs.end()) zcs.push_back(c);
      }
    }
    if (bys.size() * zcs.size() == 1) {
      h->march_by = bys[0];
      h->march_zc = zcs[0];
    } else {
      hipEvent_t e0, e1;
      hipError_t err = hipEventCreate(&e0);
      if (err != hipSuccess) return err;
      if ((err = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return err;
      }
      // each shape: one warm-up launch, then the best of three timed launches; the kept shape is
      // the longest z-chunk (then the taller tile) within 3 % of the fastest, since longer chunks
      // re-read fewer end planes and the timings of close shapes differ by noise
      struct Cand {
        int by, zc;
        float ms;
      };
      std::vector<Cand> cands;
      for (int by : bys)
        for (int zc : zcs) {
          if ((err = launch(by, zc)) != hipSuccess) break;  // warm-up
          float best_ms = 0.f;
          for (int r = 0; r < 3 && err == hipSuccess; ++r) {
            float ms = 0.f;
            if ((err = hipEventRecord(e0, h->stream)) != hipSuccess) break;
            if ((err = launch(by, zc)) != hipSuccess) break;
            if ((err = hipEventRecord(e1, h->stream)) != hipSuccess) break;
            if ((err = hipEventSynchronize(e1)) != hipSuccess) break;
            if ((err = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
            if (r == 0 || ms < best_ms) best_ms = ms;
          }
          if (err != hipSuccess) break;
          cands.push_back({by, zc, best_ms});
          if (h->tune_log)
            std::fprintf(stderr, "[pfamd] march %s rows %d z-chunk %d: %.1f us\n",
                         F::kMixed ? "mixed" : "fp64", by, zc, 1e3 * best_ms);
        }
      if (err == hipSuccess && !cands.empty()) {
        float fastest = cands[0].ms;
        for (const Cand &c : cands) fastest = std::min(fastest, c.ms);
        const Cand *pick = nullptr;
        for (const Cand &c : cands)
          if (c.ms <= 1.03f * fastest &&
              (!pick || c.zc > pick->zc || (c.zc == pick->zc && c.by > pick->by)))
            pick = &c;
        h->march_by = pick->by;
        h->march_zc = pick->zc;
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (err != hipSuccess) {
        h->march_by = h->march_zc = 0;
        return err;
      }
    }
  }
  return launch(h->march_by, h->march_zc);
}

// one sub-step from `cin` into `cout` with the handle's kernel choice and bulk flags
template <class F, int P, int NF, bool NOF>
hipError_t launch_diffuse_io(pf_handle *h, const DiffConsts<typename F::A> &dc,
                             const typename F::C *cin, typename F::C *cout) {
  using T = typename F::T;
  const T *pf = static_cast<const T *>(h->pfrac);
  if (h->diff_variant == 1) {
    const DiffGeom g = diff_geom(h, kDiffTY, 2048, 16);
    hipLaunchKernelGGL((diffuse_kernel<F, P, NF, NOF, kDiffTY>), dim3(g.nblocks),
                       dim3(64, kDiffTY + 3), 0, h->stream, cin, cout, pf, dc, h->d.nx, h->d.ny,
                       h->d.nz, g.zchunk, g.tiles_x, g.tiles_y, g.nblocks);
  } else {
    const unsigned nb = blocks_for(h->N);
    constexpr int V = F::kMixed ? 4 : 2;  // 16-byte runs per lane
    // (the marching kernel addresses a plane with 32-bit byte offsets)
    const bool march = h->diff_split && h->diff_vec && h->diff_march && h->d.nx % V == 0 &&
                       long(h->d.nx) * h->d.ny * long(sizeof(typename F::C)) < (1L << 32) &&
                       (h->diff_by < 0 || h->diff_by == 8 || h->diff_by == 16 || h->diff_by == 32);
    if (!h->flag_valid || h->flag_nf != NF) {
      const hipError_t ec =
          launch_classify<T, P, NF, NOF>(h->stream, pf, h->unit, h->flag, h->d.nx, h->d.ny,
                                         h->d.nz, h->N, h->bulk_mask, h->bulk_count,
                                         h->unit_valid);
      if (ec != hipSuccess) return ec;
      h->unit_valid = true;
      if (h->diff_split) {
        // the plane-marching kernel leaves every cell whose flag is not its bulk phase; the other
        // bulk kernels leave the flag-0 cells
        const unsigned *keep = march ? h->bulk_count : nullptr;
        const unsigned nl = list_blocks(h->N);
        // mixed: entries are byte pairs of cells (their codes share a byte); the count kernel
        // keeps per-block counts of the listed cells, summed into list_len[1] when they are asked
        // for (pf_list_lengths); FP64: list_len[1] is the list's length
        ListOrder order;
        order.pairs = F::kMixed;
        int *cells = F::kMixed ? h->ccount : nullptr;
        hipLaunchKernelGGL(general_count_kernel, dim3(nl), dim3(256), 0, h->stream, h->flag, keep,
                           h->N, h->gblock, order, cells);
        hipLaunchKernelGGL(general_scan_kernel, dim3(1), dim3(1024), 0, h->stream, h->gblock,
                           int(nl), F::kMixed ? static_cast<int *>(nullptr) : h->list_len + 1);
        hipLaunchKernelGGL(general_scatter_kernel, dim3(nl), dim3(256), 0, h->stream, h->flag,
                           keep, h->N, h->gblock, h->glist, order);
        h->ccount_pending = F::kMixed;
      }
      h->flag_valid = true;
      h->flag_nf = NF;
    }
    auto go = [&](auto by, auto split) {
      constexpr int BY = decltype(by)::value;
      constexpr bool SPLIT = decltype(split)::value;
      // z-chunks of ~16 planes: the measured optimum at 512^3 (4 tile rows, 8192+ workgroups)
      const DiffGeom g = diff_geom(h, BY, 131072 / BY, 8);
      hipLaunchKernelGGL((diffuse_bulk_kernel<F, P, NF, NOF, BY, SPLIT>), dim3(g.nblocks),
                         dim3(64, BY), 0, h->stream, cin, cout, pf, h->flag, dc, h->d.nx, h->d.ny,
                         h->d.nz, g.zchunk, g.tiles_x, g.tiles_y, g.nblocks);
    };
    auto go_split = [&](auto split) {
      constexpr bool SPLIT = decltype(split)::value;
      // below ~64M cells the storage-order kernel measured faster (caches hold more of the domain)
      const int by = h->diff_by >= 0 ? h->diff_by : (h->N < (1L << 26) ? 0 : 4);
      if (by == 0)
        hipLaunchKernelGGL((diffuse_linear_kernel<F, P, NF, NOF, SPLIT>), dim3(nb), dim3(256), 0,
                           h->stream, cin, cout, pf, h->flag, dc, h->d.nx, h->d.ny, h->d.nz);
      else if (by == 16)
        go(std::integral_constant<int, 16>(), split);
      else if (by == 8)
        go(std::integral_constant<int, 8>(), split);
      else
        go(std::integral_constant<int, 4>(), split);
    };
    if (march) {
      // kind 2: this launch alone, on sampled sub-steps once the shape is picked
      Timer t(h, 2, h->time_march && h->march_by != 0 ? 1 : 0);
      const hipError_t e = launch_march<F, P, NF, NOF>(h, dc, cin, cout);
      if (e != hipSuccess) return e;
    } else if (h->diff_split && h->diff_vec && h->d.nx % V == 0 &&
               (h->diff_by < 0 || h->diff_by == 4)) {
      const DiffGeom g = diff_geom(h, 4, 131072 / 4, 8, 64 * V);
      hipLaunchKernelGGL((diffuse_bulk_vec_kernel<F, P, NF, NOF, 4, V>), dim3(g.nblocks),
                         dim3(64, 4), 0, h->stream, cin, cout, h->flag, dc, h->d.nx, h->d.ny,
                         h->d.nz, g.zchunk, g.tiles_x, g.tiles_y, g.nblocks);
    } else if (h->diff_split) {
      go_split(std::true_type());
    }
    if (h->diff_split) {
      // the listed cells: a grid-stride pass sized for the whole chip, bounded by the list length
      const unsigned ng = std::min<unsigned>(nb, 2048);
      hipLaunchKernelGGL((diffuse_general_kernel<F, P, NF, NOF>), dim3(ng), dim3(256), 0,
                         h->stream, cin, cout, pf, h->glist, h->gblock + list_blocks(h->N), dc,
                         h->d.nx, h->d.ny, h->d.nz);
    } else {
      go_split(std::false_type());
    }
  }
  return hipGetLastError();
}

template <class F, int P, int NF, bool NOF>
hipError_t launch_diffuse(pf_handle *h, int s, const DiffConsts<typename F::A> &dc) {
  using C = typename F::C;
  const int a = h->ccur[s], b = 1 - a;
  const C *cin = static_cast<const C *>(h->conc[s][a]);
  C *cout = static_cast<C *>(h->conc[s][b]);
  hipError_t e;
  if (h->retable[s]) {
    // the first sub-step after a sweep that changed the partition tables: full formula for every
    // cell, copied cells partitioned with the previous tables (their stored partitions)
    DiffConsts<typename F::A> prev = dc;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        prev.pk[i][j] = typename F::A(h->prev_pk[s][i][j]);
        prev.pc[i][j] = typename F::A(h->prev_pc[s][i][j]);
      }
    hipLaunchKernelGGL((diffuse_retable_kernel<F, P, NF, NOF>), dim3(blocks_for(h->N)), dim3(256),
                       0, h->stream, cin, cout, static_cast<const typename F::T *>(h->pfrac),
                       h->retable[s] == 1 ? h->sflag : static_cast<const uint8_t *>(nullptr), dc,
                       prev, h->d.nx, h->d.ny, h->d.nz);
    e = hipGetLastError();
    tables_current(h, s);
  } else {
    e = launch_diffuse_io<F, P, NF, NOF>(h, dc, cin, cout);
  }
  h->ccur[s] = b;
  return e;
}

template <class F, int P>
hipError_t dispatch_diffuse_p(pf_handle *h, int s, const DiffConsts<typename F::A> &dc) {
  if (h->NF == 4) return launch_diffuse<F, P, 4, false>(h, s, dc);
  if (h->d.bc_z == PF_BC_NOFLUX) return launch_diffuse<F, P, 6, true>(h, s, dc);
  return launch_diffuse<F, P, 6, false>(h, s, dc);
}

template <class F>
int run_diffuse(pf_handle *h, int s, int n, double factor) {
  auto dc = make_diff_consts<typename F::A>(h->p, s, factor);
  if (n <= 0) return PF_OK;
  if (h->have_state && !h->have_prev) tables_state_set(h);
  // a change of tables with no sweep since: the first sub-step reads every cell's partition of
  // the previous tables; every cell is partitioned with the current tables after this batch
  if (!h->retable[s] && tables_differ(h, s)) h->retable[s] = 2;
  // the first launch classifies bulk cells when the flags are stale; keep it out of the timing
  dc.key = comp_step_key(h->csub[s]++, s);
  hipError_t e = h->d.nphase == 2 ? dispatch_diffuse_p<F, 2>(h, s, dc)
                                  : dispatch_diffuse_p<F, 3>(h, s, dc);
  if (e != hipSuccess) return fail(h, PF_ERR_HIP, std::string("diffuse: ") + hipGetErrorString(e));
  Timer t(h, 0, n - 1);
  for (int it = 1; it < n; ++it) {
    dc.key = comp_step_key(h->csub[s]++, s);
    h->time_march = h->timing && it % kMarchTimingEvery == 0;
    e = h->d.nphase == 2 ? dispatch_diffuse_p<F, 2>(h, s, dc) : dispatch_diffuse_p<F, 3>(h, s, dc);
    h->time_march = false;
    if (e != hipSuccess) return fail(h, PF_ERR_HIP, std::string("diffuse: ") + hipGetErrorString(e));
  }
  tables_current(h, s);
  return PF_OK;
}

template <class F> int run_sweep(pf_handle *h) {
  if (h->have_state && !h->have_prev) tables_state_set(h);
  // cells the sweep updates are partitioned with the current tables from here on, the copied ones
  // keep the tables they were last partitioned with: the next sub-step of a solute whose tables
  // changed mixes the two
  for (int s = 0; s < h->d.nsolute; ++s) h->retable[s] = tables_differ(h, s) ? 1 : 0;
  const auto cs = make_pf_consts<typename F::A>(h);
  Timer t(h, 1);
  hipError_t e = dispatch_sweep<F, false>(h, cs);
  if (e != hipSuccess) return fail(h, PF_ERR_HIP, std::string("sweep: ") + hipGetErrorString(e));
  return PF_OK;
}

int check_ready(pf_handle *h) {
  if (!h) return PF_ERR_INVALID;
  if (!h->have_params) return fail(h, PF_ERR_STATE, "parameters not set (pf_set_params)");
  if (!h->have_state) return fail(h, PF_ERR_STATE, "state not uploaded or constructed");
  return PF_OK;
}

int diffuse_any(pf_handle *h, int s, int n, double factor) {
  return is_mixed(h) ? run_diffuse<FormMixed>(h, s, n, factor)
                     : run_diffuse<FormF64>(h, s, n, factor);
}

void free_all(pf_handle *h) {
  for (int b = 0; b < 2; ++b) {
    if (h->id[b]) hipFree(h->id[b]);
    if (h->phi[b]) hipFree(h->phi[b]);
    for (int s = 0; s < PF_MAX_SOLUTE; ++s)
      if (h->conc[s][b]) hipFree(h->conc[s][b]);
  }
  if (h->pfrac) hipFree(h->pfrac);
  if (h->flag) hipFree(h->flag);
  if (h->unit) hipFree(h->unit);
  if (h->bulk_count) hipFree(h->bulk_count);
  if (h->glist) hipFree(h->glist);
  if (h->list_len) hipFree(h->list_len);
  if (h->ccount) hipFree(h->ccount);
  if (h->sflag) hipFree(h->sflag);
  if (h->squeue) hipFree(h->squeue);
  if (h->olist) hipFree(h->olist);
  if (h->gblock) hipFree(h->gblock);
  if (h->red) hipFree(h->red);
  for (int k = 0; k < 3; ++k)
    for (auto &pr : h->ev[k]) {
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
  if (h->dump_thread.joinable()) h->dump_thread.join();
  if (h->dump_dev) hipFree(h->dump_dev);
  if (h->dump_host) hipHostFree(h->dump_host);
  if (h->dump_stream) hipStreamDestroy(h->dump_stream);
  for (hipEvent_t &ev : h->dump_ev)
    if (ev) hipEventDestroy(ev);
  if (h->stream) hipStreamDestroy(h->stream);
}

int phase_of_id(int32_t g) {  // phase_field.h:25-29
  if (g == 0) return 0;
  return g / 100000 > 0 ? g / 100000 : 1;
}

}  // namespace

// =================================================================================================
extern "C" {

const char *pf_last_error(const pf_handle *h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

int pf_create(const pf_desc *desc, pf_handle **out) {
  if (!desc || !out) return fail(nullptr, PF_ERR_INVALID, "null argument");
  const pf_desc &d = *desc;
  if (d.nx < 1 || d.ny < 1 || d.nz < 1)
    return fail(nullptr, PF_ERR_INVALID, "domain extents must be >= 1");
  if (d.depth != 4 && d.depth != 6) return fail(nullptr, PF_ERR_INVALID, "depth must be 4 or 6");
  if (d.nphase != 2 && d.nphase != 3) return fail(nullptr, PF_ERR_INVALID, "nphase must be 2 or 3");
  if (d.nsolute < 1 || d.nsolute > PF_MAX_SOLUTE)
    return fail(nullptr, PF_ERR_INVALID, "nsolute must be 1 or 2");
  if (d.precision != PF_FP64 && d.precision != PF_MIXED)
    return fail(nullptr, PF_ERR_INVALID, "precision must be PF_FP64 or PF_MIXED");
  if (d.bc_z != PF_BC_PERIODIC && d.bc_z != PF_BC_NOFLUX)
    return fail(nullptr, PF_ERR_INVALID, "bc_z must be periodic or no-flux");
  const int dims = 3 - (d.nz == 1 ? 1 : 0) - (d.ny == 1 ? 1 : 0);
  if (dims < 2) return fail(nullptr, PF_ERR_INVALID, "one-dimensional domains are not supported");
  if (long(d.nx) * d.ny * d.nz > 2147483647L)
    return fail(nullptr, PF_ERR_INVALID, "more than 2^31 - 1 cells in one domain");
  if (d.bc_z == PF_BC_NOFLUX && dims != 3)
    return fail(nullptr, PF_ERR_INVALID, "no-flux z needs a 3D domain");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, PF_ERR_NODEVICE, "no HIP device available");
  if (d.device < 0 || d.device >= ndev) return fail(nullptr, PF_ERR_INVALID, "bad device ordinal");
  if (hipSetDevice(d.device) != hipSuccess) return fail(nullptr, PF_ERR_HIP, "hipSetDevice failed");

  auto h = std::make_unique<pf_handle>();
  h->d = d;
  h->N = long(d.nx) * d.ny * d.nz;
  if (hipDeviceGetAttribute(&h->num_cu, hipDeviceAttributeMultiprocessorCount, d.device) !=
          hipSuccess ||
      h->num_cu < 1)
    h->num_cu = 1;
  h->KI = d.depth == 4 ? 4 : 8;
  h->dims = dims;
  h->NF = 2 * dims;
  h->tsz = d.precision == PF_MIXED ? 4 : 8;
  // doubles, or float values followed by the 4-bit compensation codes
  h->cbuf = d.precision == PF_MIXED ? size_t(mixed_conc_bytes(h->N)) : size_t(h->N) * 8;
  if (const char *e = std::getenv("PFAMD_DIFF_ZCHUNK")) h->diff_zchunk_override = std::atoi(e);
  if (const char *e = std::getenv("PFAMD_DIFF_KERNEL"))
    h->diff_variant = std::strcmp(e, "lds") == 0 ? 1 : 0;
  if (const char *e = std::getenv("PFAMD_DIFF_BY")) h->diff_by = std::atoi(e);
  if (const char *e = std::getenv("PFAMD_TUNE_LOG")) h->tune_log = std::atoi(e) != 0;
  if (const char *e = std::getenv("PFAMD_DIFF_SPLIT")) h->diff_split = std::atoi(e) != 0;
  if (const char *e = std::getenv("PFAMD_DIFF_VEC")) h->diff_vec = std::atoi(e) != 0;
  if (const char *e = std::getenv("PFAMD_DIFF_MARCH")) h->diff_march = std::atoi(e) != 0;
  if (const char *e = std::getenv("PFAMD_SWEEP_SPLIT")) h->sweep_split = std::atoi(e) != 0;
  if (const char *e = std::getenv("PFAMD_SWEEP_ROWS")) h->sweep_rows = std::atoi(e);
  if (const char *e = std::getenv("PFAMD_SWEEP_QUEUE")) h->sweep_queue = std::atoi(e) != 0;
  pf_handle *hp = h.get();
  auto chk = [&](hipError_t e, const char *what) -> int {
    if (e == hipSuccess) return PF_OK;
    free_all(hp);
    return fail(nullptr, e == hipErrorOutOfMemory ? PF_ERR_OOM : PF_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
  };
  int rc;
  if ((rc = chk(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "stream")) != PF_OK)
    return rc;
  const size_t N = size_t(h->N);
  for (int b = 0; b < 2; ++b) {
    if ((rc = chk(hipMalloc(&h->id[b], N * h->KI * 2), "alloc ids")) != PF_OK) return rc;
    if ((rc = chk(hipMalloc(&h->phi[b], N * d.depth * h->tsz), "alloc phi")) != PF_OK) return rc;
    for (int s = 0; s < d.nsolute; ++s)
      if ((rc = chk(hipMalloc(&h->conc[s][b], h->cbuf), "alloc conc")) != PF_OK) return rc;
  }
  if ((rc = chk(hipMalloc(&h->pfrac, N * d.nphase * h->tsz), "alloc value_phase")) != PF_OK)
    return rc;
  if ((rc = chk(hipMalloc(&h->flag, N), "alloc bulk flags")) != PF_OK) return rc;
  if ((rc = chk(hipMalloc(&h->unit, N), "alloc bulk flags")) != PF_OK) return rc;
  if ((rc = chk(hipMalloc(&h->bulk_count, 4 * sizeof(unsigned)), "alloc bulk counts")) != PF_OK)
    return rc;
  if ((rc = chk(hipMalloc(&h->glist, N * 4), "alloc full-formula list")) != PF_OK) return rc;
  if ((rc = chk(hipMalloc(&h->list_len, 2 * sizeof(int)), "alloc list lengths")) != PF_OK) return rc;
  if ((rc = chk(hipMalloc(&h->ccount, (list_blocks(h->N) + 1) * sizeof(int)), "alloc list counts")) !=
      PF_OK)
    return rc;
  if ((rc = chk(hipMemsetAsync(h->list_len, 0, 2 * sizeof(int), h->stream), "clear")) != PF_OK)
    return rc;
  if ((rc = chk(hipMalloc(&h->sflag, N), "alloc sweep copy flags")) != PF_OK) return rc;
  if ((rc = chk(hipMalloc(&h->squeue, 8 * 32 * sizeof(int)), "alloc sweep counters")) != PF_OK)
    return rc;
  if (d.depth == 4 && !is_mixed(h.get()) &&
      (rc = chk(hipMalloc(&h->olist, (h->N + 4) * sizeof(int)), "alloc sweep overflow list")) !=
          PF_OK)
    return rc;
  if ((rc = chk(hipMalloc(&h->gblock, (list_blocks(h->N) + 1) * sizeof(int)),
                "alloc list offsets")) != PF_OK)
    return rc;
  h->red_blocks = int(std::min<long>(1024, std::max<long>(1, (h->N + 4095) / 4096)));
  if ((rc = chk(hipMalloc(&h->red, sizeof(RedPartial) * h->red_blocks), "alloc partials")) != PF_OK)
    return rc;
  for (int b = 0; b < 2; ++b) {
    if ((rc = chk(hipMemsetAsync(h->id[b], 0, N * h->KI * 2, h->stream), "clear")) != PF_OK)
      return rc;
    if ((rc = chk(hipMemsetAsync(h->phi[b], 0, N * d.depth * h->tsz, h->stream), "clear")) != PF_OK)
      return rc;
  }
  if ((rc = chk(hipStreamSynchronize(h->stream), "sync")) != PF_OK) return rc;
  *out = h.release();
  return PF_OK;
}

int pf_destroy(pf_handle *h) {
  if (!h) return PF_OK;
  hipSetDevice(h->d.device);
  if (h->dump_busy) (void)pf_dump_wait(h);
  if (h->stream) hipStreamSynchronize(h->stream);
  free_all(h);
  delete h;
  return PF_OK;
}

int pf_dump_wait(pf_handle *h) {
  if (!h) return PF_ERR_INVALID;
  if (!h->dump_busy) return PF_OK;
  h->dump_thread.join();
  h->dump_busy = false;
  if (h->dump_status != PF_OK) return fail(h, h->dump_status, h->dump_error);
  return PF_OK;
}

int pf_dump_vti_async(pf_handle *h, const char *path, double dx) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!path) return fail(h, PF_ERR_INVALID, "null path");
  if ((rc = pf_dump_wait(h)) != PF_OK) return rc;  // the buffers are the previous dump's
  hipSetDevice(h->d.device);
  const long N = h->N;
  if (!h->dump_dev) {
    HIPCHK(h, hipMalloc(&h->dump_dev, size_t(N) * 8));
    HIPCHK(h, hipHostMalloc(&h->dump_host, size_t(N) * 8, hipHostMallocDefault));
    HIPCHK(h, hipStreamCreateWithFlags(&h->dump_stream, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreateWithFlags(&h->dump_ev[0], hipEventDisableTiming));
    HIPCHK(h, hipEventCreateWithFlags(&h->dump_ev[1], hipEventDisableTiming));
  }
  int32_t *korn = static_cast<int32_t *>(h->dump_dev);
  float *conc = reinterpret_cast<float *>(korn + N);
  // fields of the current state, in stream order with the steps around the call
  if (is_mixed(h))
    hipLaunchKernelGGL(vti_fields_kernel<FormMixed>, dim3(blocks_for(N)), dim3(256), 0, h->stream,
                       h->id[h->cur], h->KI, static_cast<const float *>(h->conc[0][h->ccur[0]]), N,
                       korn, conc);
  else
    hipLaunchKernelGGL(vti_fields_kernel<FormF64>, dim3(blocks_for(N)), dim3(256), 0, h->stream,
                       h->id[h->cur], h->KI, static_cast<const double *>(h->conc[0][h->ccur[0]]),
                       N, korn, conc);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventRecord(h->dump_ev[0], h->stream));
  HIPCHK(h, hipStreamWaitEvent(h->dump_stream, h->dump_ev[0], 0));
  HIPCHK(h, hipMemcpyAsync(h->dump_host, h->dump_dev, size_t(N) * 8, hipMemcpyDeviceToHost,
                           h->dump_stream));
  HIPCHK(h, hipEventRecord(h->dump_ev[1], h->dump_stream));
  h->dump_busy = true;
  h->dump_status = PF_OK;
  h->dump_error.clear();
  const std::string file(path);
  const int nx = h->d.nx, ny = h->d.ny, nz = h->d.nz, dev = h->d.device;
  const hipEvent_t done = h->dump_ev[1];
  const int32_t *hk = static_cast<const int32_t *>(h->dump_host);
  const float *hc = reinterpret_cast<const float *>(hk + N);
  h->dump_thread = std::thread([h, file, nx, ny, nz, dx, dev, done, hk, hc]() {
    hipSetDevice(dev);
    const hipError_t e = hipEventSynchronize(done);
    if (e != hipSuccess) {
      h->dump_status = PF_ERR_HIP;
      h->dump_error = std::string("dump copy: ") + hipGetErrorString(e);
      return;
    }
    if (pf_write_vti_raw(file.c_str(), nx, ny, nz, dx, hk, hc) != PF_OK) {
      h->dump_status = PF_ERR_IO;
      h->dump_error = "cannot write " + file;
    }
  });
  return PF_OK;
}

int pf_set_params(pf_handle *h, const pf_params *p) {
  if (!h || !p) return PF_ERR_INVALID;
  if (!(p->dx > 0) || !(p->eta > 0) || !(p->pfm_dt > 0))
    return fail(h, PF_ERR_INVALID, "dx, eta and pfm_dt must be positive");
  if (p->dims != h->dims)
    return fail(h, PF_ERR_INVALID, "params.dims does not match the domain shape");
  for (int s = 0; s < h->d.nsolute; ++s)
    if (p->n_sub[s] < 0) return fail(h, PF_ERR_INVALID, "negative sub-step count");
  // the split sweep's per-cell summaries count own slots above the cutoff
  if (h->have_params && p->cutoff != h->p.cutoff) h->sweep_hist = false;
  h->p = *p;
  h->have_params = true;
  update_bulk_mask(h);
  return PF_OK;
}

int pf_upload_state(pf_handle *h, const pf_host_state *s) {
  if (!h || !s || !s->index || !s->value || !s->phase || !s->conc)
    return fail(h, PF_ERR_INVALID, "null state pointer");
  hipSetDevice(h->d.device);
  const long N = h->N;
  const int K = h->d.depth, KI = h->KI, P = h->d.nphase;
  // grain-code tables: per phase, sorted unique ids -> serials 1..
  std::vector<int32_t> ids[4];
  for (long n = 0; n < N * K; ++n) {
    const int32_t g = s->index[n];
    if (g == 0) continue;
    if (g < 0) return fail(h, PF_ERR_INVALID, "negative grain id");
    const int ph = phase_of_id(g);
    if (ph > 3 || ph > P)
      return fail(h, PF_ERR_INVALID, "grain id of a phase the domain does not carry");
    ids[ph].push_back(g);
  }
  std::map<int32_t, uint16_t> code;
  for (int ph = 1; ph <= 3; ++ph) {
    std::sort(ids[ph].begin(), ids[ph].end());
    ids[ph].erase(std::unique(ids[ph].begin(), ids[ph].end()), ids[ph].end());
    if (ids[ph].size() > 16383)
      return fail(h, PF_ERR_RANGE, "more than 16383 grains of one phase");
    h->code_to_id[ph].assign(1, 0);
    for (size_t q = 0; q < ids[ph].size(); ++q) {
      h->code_to_id[ph].push_back(ids[ph][q]);
      code[ids[ph][q]] = uint16_t((ph << 14) | int(q + 1));
    }
  }
  std::vector<uint16_t> hid(size_t(N) * KI, 0);
  for (long n = 0; n < N; ++n)
    for (int k = 0; k < K; ++k) {
      const int32_t g = s->index[n * K + k];
      hid[size_t(n) * KI + k] = g == 0 ? 0 : code[g];
    }
  for (long n = 0; n < N; ++n)
    if (P == 2 && s->phase[n * 3 + 2] != 0.0)
      return fail(h, PF_ERR_INVALID, "value_phase of phase 3 is nonzero but nphase == 2");
  const size_t NK = size_t(N) * K;
  HIPCHK(h, hipMemcpyAsync(h->id[0], hid.data(), hid.size() * 2, hipMemcpyHostToDevice, h->stream));
  if (is_mixed(h)) {
    std::vector<float> v(NK), ph(size_t(N) * P);
    for (size_t n = 0; n < NK; ++n) v[n] = float(s->value[n]);
    for (long n = 0; n < N; ++n)
      for (int p = 0; p < P; ++p) ph[size_t(n) * P + p] = float(s->phase[n * 3 + p]);
    HIPCHK(h, hipMemcpyAsync(h->phi[0], v.data(), NK * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->pfrac, ph.data(), ph.size() * 4, hipMemcpyHostToDevice, h->stream));
    for (int sol = 0; sol < h->d.nsolute; ++sol) {
      // [values][compensation codes] per buffer
      std::vector<float> c(h->cbuf / 4, 0.f);
      for (long n = 0; n < N; ++n) {
        c[n] = float(s->conc[sol * N + n]);
        if (s->conc_comp)
          comp_put_host(c.data(), N, n, comp_encode(float(s->conc_comp[sol * N + n]), c[n], 0u));
      }
      HIPCHK(h, hipMemcpyAsync(h->conc[sol][0], c.data(), h->cbuf, hipMemcpyHostToDevice,
                               h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
  } else {
    std::vector<double> ph(size_t(N) * P);
    for (long n = 0; n < N; ++n)
      for (int p = 0; p < P; ++p) ph[size_t(n) * P + p] = s->phase[n * 3 + p];
    HIPCHK(h, hipMemcpyAsync(h->phi[0], s->value, NK * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->pfrac, ph.data(), ph.size() * 8, hipMemcpyHostToDevice, h->stream));
    for (int sol = 0; sol < h->d.nsolute; ++sol)
      HIPCHK(h, hipMemcpyAsync(h->conc[sol][0], s->conc + sol * N, N * 8, hipMemcpyHostToDevice,
                               h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  h->cur = 0;
  h->ccur[0] = h->ccur[1] = 0;
  h->csub[0] = h->csub[1] = 0;
  h->have_state = true;
  h->flag_valid = false;
  h->unit_valid = false;
  h->sweep_hist = false;
  tables_state_set(h);
  return PF_OK;
}

int pf_download_state(pf_handle *h, pf_host_state *s) {
  if (!h || !s) return PF_ERR_INVALID;
  if (!h->have_state) return fail(h, PF_ERR_STATE, "no state on the device");
  hipSetDevice(h->d.device);
  const long N = h->N;
  const int K = h->d.depth, KI = h->KI, P = h->d.nphase;
  const size_t NK = size_t(N) * K;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (s->index) {
    std::vector<uint16_t> hid(size_t(N) * KI);
    HIPCHK(h, hipMemcpy(hid.data(), h->id[h->cur], hid.size() * 2, hipMemcpyDeviceToHost));
    for (long n = 0; n < N; ++n)
      for (int k = 0; k < K; ++k) {
        const uint16_t c = hid[size_t(n) * KI + k];
        const int ph = c >> 14, ser = c & 0x3fff;
        int32_t g = 0;
        if (c != 0) {
          if (ser >= int(h->code_to_id[ph].size()))
            return fail(h, PF_ERR_STATE, "grain code without a table entry");
          g = h->code_to_id[ph][ser];
        }
        s->index[n * K + k] = g;
      }
  }
  std::vector<double> phase(size_t(N) * 3, 0.0);
  std::vector<double> conc(size_t(N) * h->d.nsolute), comp(size_t(N) * h->d.nsolute, 0.0);
  if (is_mixed(h)) {
    std::vector<float> v(NK), ph(size_t(N) * P);
    HIPCHK(h, hipMemcpy(v.data(), h->phi[h->cur], NK * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ph.data(), h->pfrac, ph.size() * 4, hipMemcpyDeviceToHost));
    if (s->value)
      for (size_t n = 0; n < NK; ++n) s->value[n] = v[n];
    for (long n = 0; n < N; ++n)
      for (int p = 0; p < P; ++p) phase[size_t(n) * 3 + p] = ph[size_t(n) * P + p];
    for (int sol = 0; sol < h->d.nsolute; ++sol) {
      std::vector<float> c(h->cbuf / 4);
      HIPCHK(h, hipMemcpy(c.data(), h->conc[sol][h->ccur[sol]], h->cbuf, hipMemcpyDeviceToHost));
      for (long n = 0; n < N; ++n) {
        conc[sol * N + n] = c[n];
        comp[sol * N + n] = comp_decode(comp_load(c.data(), N, n), c[n]);
      }
    }
  } else {
    std::vector<double> ph(size_t(N) * P);
    if (s->value)
      HIPCHK(h, hipMemcpy(s->value, h->phi[h->cur], NK * 8, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ph.data(), h->pfrac, ph.size() * 8, hipMemcpyDeviceToHost));
    for (long n = 0; n < N; ++n)
      for (int p = 0; p < P; ++p) phase[size_t(n) * 3 + p] = ph[size_t(n) * P + p];
    for (int sol = 0; sol < h->d.nsolute; ++sol)
      HIPCHK(h, hipMemcpy(conc.data() + sol * N, h->conc[sol][h->ccur[sol]], N * 8,
                          hipMemcpyDeviceToHost));
  }
  if (s->phase) std::memcpy(s->phase, phase.data(), phase.size() * 8);
  if (s->conc) std::memcpy(s->conc, conc.data(), conc.size() * 8);
  if (s->conc_comp) std::memcpy(s->conc_comp, comp.data(), comp.size() * 8);
  if (s->phase_conc) {
    // value_conc1 = Partitioning(total_conc, value_phase), solver.cpp:389-405, in the form's
    // precision over all three phases
    for (int sol = 0; sol < h->d.nsolute; ++sol)
      for (long n = 0; n < N; ++n)
        for (int i = 0; i < 3; ++i) {
          if (is_mixed(h)) {
            float t1 = 0.f, t2 = 0.f;
            for (int j = 0; j < 3; ++j) {
              t1 += float(phase[n * 3 + j]) * float(h->p.part_k[sol][i][j]);
              t2 += float(phase[n * 3 + j]) * float(h->p.part_c[sol][i][j]);
            }
            s->phase_conc[(sol * N + n) * 3 + i] = (float(conc[sol * N + n]) - t2) / t1;
          } else {
            double t1 = 0.0, t2 = 0.0;
            for (int j = 0; j < 3; ++j) {
              t1 += phase[n * 3 + j] * h->p.part_k[sol][i][j];
              t2 += phase[n * 3 + j] * h->p.part_c[sol][i][j];
            }
            s->phase_conc[(sol * N + n) * 3 + i] = (conc[sol * N + n] - t2) / t1;
          }
        }
  }
  return PF_OK;
}

int pf_phase_field_sweep(pf_handle *h) {
  int rc = check_ready(h);
  if (rc) return rc;
  hipSetDevice(h->d.device);
  return is_mixed(h) ? run_sweep<FormMixed>(h) : run_sweep<FormF64>(h);
}

int pf_diffuse_substeps(pf_handle *h, int solute, int n, double dt_d) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (solute < 0 || solute >= h->d.nsolute || n < 0)
    return fail(h, PF_ERR_INVALID, "bad solute index or sub-step count");
  hipSetDevice(h->d.device);
  // phase_field.cpp:366: del_conc *= diff_dt/(dx*dx)
  return diffuse_any(h, solute, n, dt_d / (h->p.dx * h->p.dx));
}

int pf_step(pf_handle *h, int n_steps) {
  int rc = check_ready(h);
  if (rc) return rc;
  hipSetDevice(h->d.device);
  for (int it = 0; it < n_steps; ++it) {
    rc = is_mixed(h) ? run_sweep<FormMixed>(h) : run_sweep<FormF64>(h);
    if (rc) return rc;
    for (int s = 0; s < h->d.nsolute; ++s) {
      if (h->p.n_sub[s] <= 0) continue;
      rc = diffuse_any(h, s, h->p.n_sub[s], h->p.dt_d[s] / (h->p.dx * h->p.dx));
      if (rc) return rc;
    }
  }
  return PF_OK;
}

int pf_synchronize(pf_handle *h) {
  if (!h) return PF_ERR_INVALID;
  hipSetDevice(h->d.device);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return PF_OK;
}

int pf_reduce(pf_handle *h, pf_stats *out) {
  if (!h || !out) return PF_ERR_INVALID;
  if (!h->have_state) return fail(h, PF_ERR_STATE, "no state on the device");
  hipSetDevice(h->d.device);
  const long per = (h->N + h->red_blocks - 1) / h->red_blocks;
  const int nb = int((h->N + per - 1) / per);
  const int K = h->d.depth, P = h->d.nphase;
  const void *c1 = h->d.nsolute > 1 ? h->conc[1][h->ccur[1]] : h->conc[0][h->ccur[0]];
#define PF_RED(F, KK, PP)                                                                          \
  hipLaunchKernelGGL((reduce_stage1<F, KK, PP>), dim3(nb), dim3(256), 0, h->stream, h->id[h->cur], \
                     static_cast<const F::T *>(h->phi[h->cur]),                                    \
                     static_cast<const F::T *>(h->pfrac),                                          \
                     static_cast<const F::C *>(h->conc[0][h->ccur[0]]),                            \
                     static_cast<const F::C *>(c1), h->d.nsolute, h->N, per, h->red)
  if (is_mixed(h)) {
    if (K == 4 && P == 2) PF_RED(FormMixed, 4, 2);
    else if (K == 4) PF_RED(FormMixed, 4, 3);
    else if (P == 2) PF_RED(FormMixed, 6, 2);
    else PF_RED(FormMixed, 6, 3);
  } else {
    if (K == 4 && P == 2) PF_RED(FormF64, 4, 2);
    else if (K == 4) PF_RED(FormF64, 4, 3);
    else if (P == 2) PF_RED(FormF64, 6, 2);
    else PF_RED(FormF64, 6, 3);
  }
#undef PF_RED
  HIPCHK(h, hipGetLastError());
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(256), 0, h->stream, h->red, nb);
  HIPCHK(h, hipGetLastError());
  RedPartial r;
  HIPCHK(h, hipMemcpyAsync(&r, h->red, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::memset(out, 0, sizeof(*out));
  for (int s = 0; s < PF_MAX_SOLUTE; ++s) {
    const bool in = s < h->d.nsolute;
    out->mass[s] = in ? r.mass[s] : 0.0;
    out->max_conc[s] = in ? r.cmax[s] : 0.0;
    out->min_conc[s] = in ? r.cmin[s] : 0.0;
  }
  for (int p = 0; p < 3; ++p) out->phase_frac[p] = r.pf[p] / double(h->N);
  out->max_sum_err = r.serr;
  out->n_interface = r.nif;
  out->n_nonfinite = r.nnf;
  if (r.nnf > 0) return fail(h, PF_ERR_NONFINITE, "non-finite values in the state");
  return PF_OK;
}

int pf_rex_fraction(pf_handle *h, double *out) {
  if (!h || !out) return PF_ERR_INVALID;
  bool any = false;
  for (int ph = 0; ph < 4; ++ph)
    for (size_t q = 1; q < h->code_to_id[ph].size(); ++q) any = any || h->code_to_id[ph][q] < 100000;
  *out = 0.0;
  if (!any) return PF_OK;
  const long N = h->N;
  const int K = h->d.depth;
  std::vector<int32_t> idx(size_t(N) * K);
  std::vector<double> val(size_t(N) * K);
  std::vector<double> phase(size_t(N) * 3), conc(size_t(N) * h->d.nsolute);
  pf_host_state s{};
  s.index = idx.data();
  s.value = val.data();
  s.phase = phase.data();
  s.conc = conc.data();
  int rc = pf_download_state(h, &s);
  if (rc != PF_OK) return rc;
  double acc = 0.0;
  for (size_t n = 0; n < idx.size(); ++n) acc += idx[n] < 100000 ? val[n] : 0.0;
  *out = acc / double(N);
  return PF_OK;
}

int pf_kernel_timing_reset(pf_handle *h, int enable) {
  if (!h) return PF_ERR_INVALID;
  hipSetDevice(h->d.device);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->timing = enable != 0;
  for (int k = 0; k < 3; ++k) {
    h->ev_used[k] = 0;
    h->t_ms[k] = 0.0;
    h->t_n[k] = 0;
  }
  return PF_OK;
}

int pf_kernel_timing(pf_handle *h, int kind, double *total_ms, int64_t *launches) {
  if (!h || kind < 0 || kind > 2) return PF_ERR_INVALID;
  hipSetDevice(h->d.device);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t i = 0; i < h->ev_used[kind]; ++i) {
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev[kind][i].first, h->ev[kind][i].second));
    h->t_ms[kind] += ms;
    h->t_n[kind] += h->ev_count[kind][i];
  }
  h->ev_used[kind] = 0;
  if (total_ms) *total_ms = h->t_ms[kind];
  if (launches) *launches = h->t_n[kind];
  return PF_OK;
}

int pf_march_shape(const pf_handle *h, int *tile_rows, int *z_chunk) {
  if (!h) return PF_ERR_INVALID;
  handle_march_shape(h, tile_rows, z_chunk);
  return PF_OK;
}

int pf_list_lengths(pf_handle *h, int64_t *sweep_listed, int64_t *diffusion_listed) {
  if (!h) return PF_ERR_INVALID;
  hipSetDevice(h->d.device);
  int v[2] = {0, 0};
  if (h->ccount_pending) {  // the pair list's per-block cell counts, summed once
    hipLaunchKernelGGL(general_scan_kernel, dim3(1), dim3(1024), 0, h->stream, h->ccount,
                       int(list_blocks(h->N)), h->list_len + 1);
    HIPCHK(h, hipGetLastError());
    h->ccount_pending = false;
  }
  HIPCHK(h, hipMemcpyAsync(v, h->list_len, sizeof(v), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (sweep_listed) *sweep_listed = v[0];
  if (diffusion_listed) *diffusion_listed = v[1];
  return PF_OK;
}

// ---- checkpoints (host memory, raw device layout) -----------------------------------------------
// Blob: a 64-byte header (magic, pf_desc, the mixed form's sub-step counters, the blob's size), the
// grain-code tables (per phase 1..3: how many entries, then up to kCodeCap reference ids, so that a
// restore into any handle maps codes back to the same ids), then the device buffers of the current
// state in the device layout.
namespace {
constexpr char kCkptMagic[9] = "PFCKPT04";
constexpr size_t kCodeCap = 16384;  // serials 0..16383 of a 14-bit grain code
constexpr size_t kCkptHeader = 64;
constexpr size_t kCkptTables = 4 * sizeof(int32_t) + 3 * kCodeCap * sizeof(int32_t);
size_t checkpoint_bytes(const pf_handle *h) {
  const size_t N = size_t(h->N);
  return kCkptHeader + kCkptTables + N * h->KI * 2 + N * h->d.depth * h->tsz +
         N * h->d.nphase * h->tsz + h->cbuf * h->d.nsolute;
}
}  // namespace

int pf_checkpoint_size(const pf_handle *h, size_t *bytes) {
  if (!h || !bytes) return PF_ERR_INVALID;
  *bytes = checkpoint_bytes(h);
  return PF_OK;
}

int pf_checkpoint_save(pf_handle *h, void *dst) {
  if (!h || !dst) return PF_ERR_INVALID;
  if (!h->have_state) return fail(h, PF_ERR_STATE, "no state on the device");
  hipSetDevice(h->d.device);
  const size_t N = size_t(h->N);
  const uint64_t total = checkpoint_bytes(h);
  char *p = static_cast<char *>(dst);
  static_assert(8 + sizeof(pf_desc) + sizeof(pf_handle::csub) + sizeof(uint64_t) <= kCkptHeader,
                "checkpoint header");
  std::memset(p, 0, kCkptHeader + kCkptTables);
  std::memcpy(p, kCkptMagic, 8);
  std::memcpy(p + 8, &h->d, sizeof(pf_desc));
  std::memcpy(p + 8 + sizeof(pf_desc), h->csub, sizeof(h->csub));
  std::memcpy(p + 8 + sizeof(pf_desc) + sizeof(h->csub), &total, sizeof(total));
  p += kCkptHeader;
  int32_t count[4] = {0, 0, 0, 0};
  for (int ph = 1; ph <= 3; ++ph) {
    if (h->code_to_id[ph].size() > kCodeCap)
      return fail(h, PF_ERR_RANGE, "grain-code table larger than the checkpoint holds");
    count[ph] = int32_t(h->code_to_id[ph].size());
    if (count[ph])
      std::memcpy(p + sizeof(count) + (ph - 1) * kCodeCap * sizeof(int32_t),
                  h->code_to_id[ph].data(), count[ph] * sizeof(int32_t));
  }
  std::memcpy(p, count, sizeof(count));
  p += kCkptTables;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(p, h->id[h->cur], N * h->KI * 2, hipMemcpyDeviceToHost));
  p += N * h->KI * 2;
  HIPCHK(h, hipMemcpy(p, h->phi[h->cur], N * h->d.depth * h->tsz, hipMemcpyDeviceToHost));
  p += N * h->d.depth * h->tsz;
  HIPCHK(h, hipMemcpy(p, h->pfrac, N * h->d.nphase * h->tsz, hipMemcpyDeviceToHost));
  p += N * h->d.nphase * h->tsz;
  for (int s = 0; s < h->d.nsolute; ++s) {
    HIPCHK(h, hipMemcpy(p, h->conc[s][h->ccur[s]], h->cbuf, hipMemcpyDeviceToHost));
    p += h->cbuf;
  }
  return PF_OK;
}

int pf_checkpoint_restore(pf_handle *h, const void *src) {
  if (!h || !src) return PF_ERR_INVALID;
  hipSetDevice(h->d.device);
  const size_t N = size_t(h->N);
  const char *p = static_cast<const char *>(src);
  pf_desc d;
  uint64_t total = 0;
  std::memcpy(&d, p + 8, sizeof(d));
  std::memcpy(&total, p + 8 + sizeof(pf_desc) + sizeof(h->csub), sizeof(total));
  if (std::memcmp(p, kCkptMagic, 8) != 0 || std::memcmp(&d, &h->d, sizeof(d)) != 0 ||
      total != checkpoint_bytes(h))
    return fail(h, PF_ERR_INVALID, "checkpoint does not match this domain");
  uint32_t csub[PF_MAX_SOLUTE];
  std::memcpy(csub, p + 8 + sizeof(pf_desc), sizeof(csub));
  p += kCkptHeader;
  int32_t count[4];
  std::memcpy(count, p, sizeof(count));
  std::vector<int32_t> tables[4];
  tables[0].assign(1, 0);
  for (int ph = 1; ph <= 3; ++ph) {
    if (count[ph] < 0 || size_t(count[ph]) > kCodeCap)
      return fail(h, PF_ERR_INVALID, "checkpoint holds a bad grain-code table");
    tables[ph].resize(size_t(count[ph]));
    if (count[ph])
      std::memcpy(tables[ph].data(), p + sizeof(count) + (ph - 1) * kCodeCap * sizeof(int32_t),
                  count[ph] * sizeof(int32_t));
  }
  p += kCkptTables;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(h->id[h->cur], p, N * h->KI * 2, hipMemcpyHostToDevice));
  p += N * h->KI * 2;
  HIPCHK(h, hipMemcpy(h->phi[h->cur], p, N * h->d.depth * h->tsz, hipMemcpyHostToDevice));
  p += N * h->d.depth * h->tsz;
  HIPCHK(h, hipMemcpy(h->pfrac, p, N * h->d.nphase * h->tsz, hipMemcpyHostToDevice));
  p += N * h->d.nphase * h->tsz;
  for (int s = 0; s < h->d.nsolute; ++s) {
    HIPCHK(h, hipMemcpy(h->conc[s][h->ccur[s]], p, h->cbuf, hipMemcpyHostToDevice));
    p += h->cbuf;
  }
  std::memcpy(h->csub, csub, sizeof(csub));
  for (int ph = 0; ph < 4; ++ph) h->code_to_id[ph] = tables[ph];
  h->have_state = true;
  h->flag_valid = false;
  h->unit_valid = false;
  h->sweep_hist = false;
  tables_state_set(h);
  return PF_OK;
}

// ---- device-side seeded setup -----------------------------------------------------------------
int pf_setup_construct(pf_handle *h, const double *seeds, int G, double reduction, int n_iso,
                       const double *init_conc) {
  if (!h || !seeds || G < 1 || !init_conc) return fail(h, PF_ERR_INVALID, "bad setup arguments");
  if (is_mixed(h))
    return fail(h, PF_ERR_INVALID,
                "set up the microstructure on an FP64 domain and convert it with pf_copy_state");
  if (!h->have_params) return fail(h, PF_ERR_STATE, "parameters not set (pf_set_params)");
  if (G > 16383) return fail(h, PF_ERR_RANGE, "more than 16383 Voronoi grains");
  if (reduction >= 1.0) return fail(h, PF_ERR_INVALID, "reduction must be < 1");
  hipSetDevice(h->d.device);
  const int K = h->d.depth, P = h->d.nphase;
  double *dseeds = nullptr;
  HIPCHK(h, hipMalloc(&dseeds, sizeof(double) * 3 * G));
  hipError_t e = hipMemcpyAsync(dseeds, seeds, sizeof(double) * 3 * G, hipMemcpyHostToDevice,
                                h->stream);
  double *phi = static_cast<double *>(h->phi[0]), *pf = static_cast<double *>(h->pfrac);
  if (e == hipSuccess) {
    h->cur = 0;
#define PF_VOR(KK, PP)                                                                        \
  hipLaunchKernelGGL((voronoi_kernel<KK, PP>), dim3(blocks_for(h->N)), dim3(256), 0, h->stream, \
                     dseeds, G, reduction, h->d.nx, h->d.ny, h->d.nz, h->N, h->id[0], phi, pf)
    if (K == 4 && P == 2) PF_VOR(4, 2);
    else if (K == 4) PF_VOR(4, 3);
    else if (P == 2) PF_VOR(6, 2);
    else PF_VOR(6, 3);
#undef PF_VOR
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(dseeds);
  if (e != hipSuccess) return fail(h, PF_ERR_HIP, std::string("voronoi: ") + hipGetErrorString(e));
  // slot-1 buffer must hold zeros where Iso_growth leaves slots unwritten
  HIPCHK(h, hipMemsetAsync(h->id[1], 0, size_t(h->N) * h->KI * 2, h->stream));
  HIPCHK(h, hipMemsetAsync(h->phi[1], 0, size_t(h->N) * K * 8, h->stream));
  // Iso_growth iterations with factor 0.5/dims*dx*dx (MicrostructureConstruction.cpp:216)
  PFConsts<double> cs = make_pf_consts<double>(h);
  cs.dt = 0.5 / double(h->dims) * h->p.dx * h->p.dx;
  for (int it = 0; it < n_iso; ++it) {
    hipError_t le = dispatch_sweep<FormF64, true>(h, cs);
    if (le != hipSuccess) return fail(h, PF_ERR_HIP, std::string("iso: ") + hipGetErrorString(le));
  }
  for (int s = 0; s < h->d.nsolute; ++s) {
    h->ccur[s] = 0;
    h->csub[s] = 0;
    double *c = static_cast<double *>(h->conc[s][0]);
    if (P == 2)
      hipLaunchKernelGGL((construct_conc_kernel<2>), dim3(blocks_for(h->N)), dim3(256), 0,
                         h->stream, pf, init_conc[3 * s], init_conc[3 * s + 1],
                         init_conc[3 * s + 2], h->N, c);
    else
      hipLaunchKernelGGL((construct_conc_kernel<3>), dim3(blocks_for(h->N)), dim3(256), 0,
                         h->stream, pf, init_conc[3 * s],
                         init_conc[3 * s + 1], init_conc[3 * s + 2], h->N, c);
    HIPCHK(h, hipGetLastError());
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int ph = 0; ph < 4; ++ph) h->code_to_id[ph].assign(1, 0);
  for (int g = 1; g <= G; ++g) h->code_to_id[2].push_back(200000 + g);
  h->have_state = true;
  h->flag_valid = false;
  h->unit_valid = false;
  h->sweep_hist = false;
  tables_state_set(h);
  return PF_OK;
}

int pf_setup_nucleate(pf_handle *h, int n_nuclei, float radius, const double *nucleus_conc,
                      uint64_t seed) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (is_mixed(h)) return fail(h, PF_ERR_INVALID, "nucleate on an FP64 domain");
  if (n_nuclei < 0 || n_nuclei > 16383) return fail(h, PF_ERR_RANGE, "bad nucleus count");
  if (!(radius > 0.f)) return fail(h, PF_ERR_INVALID, "radius must be positive");
  // Nucleus q takes the reference's id 100000 + q (nucleation.cpp:29).  Grains are the same grain
  // exactly when their ids are equal, so a nucleus whose id a phase-1 grain of the state already
  // has takes that grain's code, and every other id gets the next free serial; the codes of the
  // grains already on the device are kept.
  std::vector<int32_t> table = h->code_to_id[1];
  if (table.empty()) table.assign(1, 0);
  std::vector<int> codes(size_t(n_nuclei) + 1, 0);
  {
    std::map<int32_t, int> serial;
    for (size_t s = 1; s < table.size(); ++s) serial[table[s]] = int(s);
    for (int q = 1; q <= n_nuclei; ++q) {
      const int32_t gid = 100000 + q;
      auto it = serial.find(gid);
      int ser;
      if (it != serial.end()) {
        ser = it->second;
      } else {
        ser = int(table.size());
        if (ser > 16383) return fail(h, PF_ERR_RANGE, "more than 16383 grains of phase 1");
        table.push_back(gid);
        serial[gid] = ser;
      }
      codes[q] = (1 << 14) | ser;
    }
  }
  hipSetDevice(h->d.device);
  const int K = h->d.depth, P = h->d.nphase;
  const long N = h->N;
  float *flag = nullptr, *pv = nullptr;
  uint8_t *mark = nullptr;
  long long *pi = nullptr;
  const int nparts = int(std::min<long>(1024, (N + 4095) / 4096));
  const long per = (N + nparts - 1) / nparts;
  const int nb = int((N + per - 1) / per);
  auto cleanup = [&]() {
    if (flag) hipFree(flag);
    if (mark) hipFree(mark);
    if (pv) hipFree(pv);
    if (pi) hipFree(pi);
  };
  hipError_t e = hipMalloc(&flag, N * 4);
  if (e == hipSuccess) e = hipMalloc(&mark, N);
  if (e == hipSuccess) e = hipMalloc(&pv, nparts * 4);
  if (e == hipSuccess) e = hipMalloc(&pi, nparts * 8);
  if (e != hipSuccess) {
    cleanup();
    return fail(h, PF_ERR_OOM, "nucleation scratch allocation failed");
  }
  const DiffConsts<double> dc = make_diff_consts<double>(h->p, 0, 0.0);
  const uint64_t key = [&] {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }();
  double *phi = static_cast<double *>(h->phi[h->cur]), *pf = static_cast<double *>(h->pfrac);
  const double *c0 = static_cast<const double *>(h->conc[0][h->ccur[0]]);
#define PF_FLAG(KK, PP)                                                                     \
  hipLaunchKernelGGL((nucleation_flag_kernel<KK, PP>), dim3(blocks_for(N)), dim3(256), 0,      \
                     h->stream, h->id[h->cur], pf, c0, dc, key, N, flag, mark)
  if (K == 4 && P == 2) PF_FLAG(4, 2);
  else if (K == 4) PF_FLAG(4, 3);
  else if (P == 2) PF_FLAG(6, 2);
  else PF_FLAG(6, 3);
#undef PF_FLAG
  e = hipGetLastError();
  const int R = int(std::ceil(2.0f * radius));
  const int ext[3] = {h->d.nx, h->d.ny, h->d.nz};
  int cntv[3];
  for (int d = 0; d < 3; ++d) cntv[d] = (2 * R + 1 >= ext[d]) ? ext[d] : 2 * R + 1;
  const long box = long(cntv[0]) * cntv[1] * cntv[2];
  for (int q = 1; q <= n_nuclei && e == hipSuccess; ++q) {
    hipLaunchKernelGGL(argmax_stage1, dim3(nb), dim3(256), 0, h->stream, flag, N, per, pv, pi);
    hipLaunchKernelGGL(argmax_stage2, dim3(1), dim3(256), 0, h->stream, pv, pi, nb);
#define PF_PAINT(KK, PP)                                                                      \
  hipLaunchKernelGGL((nucleus_paint_kernel<KK, PP>), dim3(blocks_for(box)), dim3(256), 0,      \
                     h->stream, pi, codes[q], radius, R, h->d.nx, h->d.ny, h->d.nz, cntv[0],  \
                     cntv[1],                                                                 \
                     cntv[2], h->id[h->cur], phi, pf, flag, mark)
    if (K == 4 && P == 2) PF_PAINT(4, 2);
    else if (K == 4) PF_PAINT(4, 3);
    else if (P == 2) PF_PAINT(6, 2);
    else PF_PAINT(6, 3);
#undef PF_PAINT
    e = hipGetLastError();
  }
  for (int s = 0; s < h->d.nsolute && e == hipSuccess; ++s)
    if (nucleus_conc && nucleus_conc[s] >= 0.0) {
      hipLaunchKernelGGL(nucleus_conc_kernel, dim3(blocks_for(N)), dim3(256), 0, h->stream, mark,
                         nucleus_conc[s], N, static_cast<double *>(h->conc[s][h->ccur[s]]));
      e = hipGetLastError();
    }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  cleanup();
  if (e != hipSuccess) return fail(h, PF_ERR_HIP, std::string("nucleate: ") + hipGetErrorString(e));
  h->code_to_id[1] = table;
  h->flag_valid = false;
  h->unit_valid = false;
  h->sweep_hist = false;
  tables_state_set(h);
  return PF_OK;
}

int pf_setup_solute_noise(pf_handle *h, int solute, double mean, double stddev, uint64_t seed) {
  if (!h) return PF_ERR_INVALID;
  if (!h->have_state) return fail(h, PF_ERR_STATE, "no state on the device");
  if (is_mixed(h)) return fail(h, PF_ERR_INVALID, "set solute noise on an FP64 domain");
  if (solute < 0 || solute >= h->d.nsolute) return fail(h, PF_ERR_INVALID, "bad solute index");
  hipSetDevice(h->d.device);
  uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  const uint64_t key = z ^ (z >> 31);
  hipLaunchKernelGGL(solute_noise_kernel, dim3(blocks_for(h->N)), dim3(256), 0, h->stream, mean,
                     stddev, key, h->N, static_cast<double *>(h->conc[solute][h->ccur[solute]]));
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  tables_state_set(h);
  return PF_OK;
}

int pf_copy_state(pf_handle *dst, const pf_handle *src) {
  if (!dst || !src) return PF_ERR_INVALID;
  if (!src->have_state) return fail(dst, PF_ERR_STATE, "source has no state");
  const pf_desc &a = dst->d, &b = src->d;
  if (a.nx != b.nx || a.ny != b.ny || a.nz != b.nz || a.depth != b.depth || a.nphase != b.nphase ||
      a.nsolute != b.nsolute || a.device != b.device)
    return fail(dst, PF_ERR_INVALID, "domains differ in shape, depth, phases, solutes or device");
  if (is_mixed(src)) return fail(dst, PF_ERR_INVALID, "source must be an FP64 domain");
  hipSetDevice(a.device);
  HIPCHK(dst, hipStreamSynchronize(src->stream));
  const size_t N = size_t(dst->N);
  dst->cur = 0;
  HIPCHK(dst, hipMemcpyAsync(dst->id[0], src->id[src->cur], N * dst->KI * 2,
                             hipMemcpyDeviceToDevice, dst->stream));
  if (is_mixed(dst)) {
    hipLaunchKernelGGL(convert_f64_to_f32, dim3(std::min(blocks_for(long(N) * a.depth), 65536u)),
                       dim3(256), 0,
                       dst->stream, static_cast<const double *>(src->phi[src->cur]),
                       static_cast<float *>(dst->phi[0]), long(N) * a.depth);
    hipLaunchKernelGGL(convert_f64_to_f32, dim3(std::min(blocks_for(long(N) * a.nphase), 65536u)),
                       dim3(256), 0,
                       dst->stream, static_cast<const double *>(src->pfrac),
                       static_cast<float *>(dst->pfrac), long(N) * a.nphase);
    for (int s = 0; s < a.nsolute; ++s) {
      dst->ccur[s] = 0;
      dst->csub[s] = 0;
      hipLaunchKernelGGL(convert_conc_f64_to_mixed, dim3(blocks_for(long(N))), dim3(256), 0,
                         dst->stream, static_cast<const double *>(src->conc[s][src->ccur[s]]),
                         static_cast<float *>(dst->conc[s][0]), long(N));
    }
    HIPCHK(dst, hipGetLastError());
  } else {
    HIPCHK(dst, hipMemcpyAsync(dst->phi[0], src->phi[src->cur], N * a.depth * 8,
                               hipMemcpyDeviceToDevice, dst->stream));
    HIPCHK(dst, hipMemcpyAsync(dst->pfrac, src->pfrac, N * a.nphase * 8, hipMemcpyDeviceToDevice,
                               dst->stream));
    for (int s = 0; s < a.nsolute; ++s) {
      dst->ccur[s] = 0;
      dst->csub[s] = 0;
      HIPCHK(dst, hipMemcpyAsync(dst->conc[s][0], src->conc[s][src->ccur[s]], N * 8,
                                 hipMemcpyDeviceToDevice, dst->stream));
    }
  }
  HIPCHK(dst, hipStreamSynchronize(dst->stream));
  for (int ph = 0; ph < 4; ++ph) dst->code_to_id[ph] = src->code_to_id[ph];
  dst->have_state = true;
  dst->flag_valid = false;
  dst->unit_valid = false;
  dst->sweep_hist = false;
  tables_state_set(dst);
  return PF_OK;
}

// ---- the reference's accelerator API (gpu_solver.h:18-23) ------------------------------------
// One process-wide domain, as the reference keeps static device buffers.  value_phase carries all
// three phases; the stencil always has six faces (the z faces of a one-cell-thick domain meet the
// cell itself and carry no flux), as in Diff3D (gpu_solver.cpp:160-188).  These are the C entry
// points; gpu_solver_compat.cpp exports the reference's C++-linkage names over them.
static pf_handle *g_compat = nullptr;
static int g_compat_dims = 3;

int pf_gpu_init(int x, int y, int z) {
  if (g_compat) {
    pf_destroy(g_compat);
    g_compat = nullptr;
  }
  pf_desc d{};
  d.nx = x;
  d.ny = y;
  d.nz = z;
  d.depth = 4;
  d.nphase = 3;
  d.nsolute = 1;
  d.precision = PF_FP64;
  d.bc_z = PF_BC_PERIODIC;
  d.device = 0;
  if (y == 1 && z == 1) return 0;  // one-dimensional domains fall back to the caller's CPU path
  if (pf_create(&d, &g_compat) != PF_OK) {
    std::fprintf(stderr, "GPU_Init: %s\n", pf_last_error(nullptr));
    g_compat = nullptr;
    return 0;
  }
  g_compat_dims = z > 1 ? 3 : 2;  // gpu_solver.cpp:93-94
  return 1;
}

int pf_gpu_diffusion_solver(const double *hdiffusivity, const double *hpartition_k,
                            const double *hpartition_c, const double *hvalue_phase,
                            double *htotal_conc, int x, int y, int z, double dt, double dx) {
  pf_handle *h = g_compat;
  if (!h || h->d.nx != x || h->d.ny != y || h->d.nz != z) {
    // The reference's Update (phase_field.cpp:289-290) calls this without a GPU_Init (only
    // cudatest.cpp:10 calls it), so the domain is set up here on first use or a new shape.
    if (pf_gpu_init(x, y, z) != 1) {
      std::fprintf(stderr, "GPU_Diffusion_solver: no usable device for a %dx%dx%d domain\n", x, y,
                   z);
      return PF_ERR_NODEVICE;
    }
    h = g_compat;
  }
  if (!hdiffusivity || !hpartition_k || !hpartition_c || !hvalue_phase || !htotal_conc) {
    std::fprintf(stderr, "GPU_Diffusion_solver: null pointer\n");
    return PF_ERR_INVALID;
  }
  hipSetDevice(h->d.device);
  pf_params p{};
  p.dx = dx;
  p.eta = dx;
  p.dims = h->dims;
  p.pfm_dt = dt;
  for (int i = 0; i < 3; ++i) {
    p.diffusivity[0][i] = hdiffusivity[i];
    for (int j = 0; j < 3; ++j) {
      p.part_k[0][i][j] = hpartition_k[i + 3 * j];  // column-major (i,j) at i+3j
      p.part_c[0][i][j] = hpartition_c[i + 3 * j];
    }
  }
  // gpu_solver.cpp:205-209, including the single-precision 0.2f
  const double max_diff = *std::max_element(hdiffusivity, hdiffusivity + 3);
  const double del_t = dx * dx / max_diff / g_compat_dims * 0.2f;
  const int num_loop = int(std::ceil(dt / del_t));
  const double dt_dxdx = dt / num_loop / (dx * dx);
  p.n_sub[0] = num_loop;
  p.dt_d[0] = dt / num_loop;
  h->p = p;
  h->have_params = true;
  update_bulk_mask(h);
  // only what the sub-steps read: value_phase [N][3] and total_conc [N], both already in the
  // device layout (gpu_solver.cpp:106-129 copies the same two arrays plus the tables)
  const size_t N = size_t(h->N);
  int rc = PF_OK;
  auto report = [&](int code, const char *what, hipError_t e) {
    std::fprintf(stderr, "GPU_Diffusion_solver: %s: %s\n", what, hipGetErrorString(e));
    return fail(h, code, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e = hipMemcpyAsync(h->pfrac, hvalue_phase, N * 3 * 8, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h->conc[0][0], htotal_conc, N * 8, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) return report(PF_ERR_HIP, "upload", e);
  h->ccur[0] = 0;
  h->csub[0] = 0;
  h->have_state = true;
  h->flag_valid = false;
  h->unit_valid = false;
  tables_state_set(h);  // the reference partitions the uploaded state first (gpu_solver.cpp:210)
  h->NF = 6;  // Diff3D always sums six faces
  rc = diffuse_any(h, 0, num_loop, dt_dxdx);
  h->NF = 2 * h->dims;
  if (rc != PF_OK) {
    std::fprintf(stderr, "GPU_Diffusion_solver: %s\n", h->err.c_str());
    return rc;
  }
  e = hipMemcpyAsync(htotal_conc, h->conc[0][h->ccur[0]], N * 8, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return report(PF_ERR_HIP, "copy back", e);
  return PF_OK;
}

void pf_gpu_cleanup(void) {
  if (g_compat) pf_destroy(g_compat);
  g_compat = nullptr;
}

}  // extern "C"

// ---- internal view for the adjoint translation unit ---------------------------------------------
HandleView handle_view(pf_handle *h) {
  HandleView v;
  v.N = h->N;
  v.nx = h->d.nx;
  v.ny = h->d.ny;
  v.nz = h->d.nz;
  v.K = h->d.depth;
  v.KI = h->KI;
  v.P = h->d.nphase;
  v.NF = h->NF;
  v.nsolute = h->d.nsolute;
  v.device = h->d.device;
  v.noflux = h->d.bc_z == PF_BC_NOFLUX;
  v.mixed = is_mixed(h);
  v.have_state = h->have_state;
  v.params = h->have_params ? &h->p : nullptr;
  v.stream = h->stream;
  v.id = h->id[h->cur];
  v.phi = h->phi[h->cur];
  v.pf = h->pfrac;
  v.c0 = h->conc[0][h->ccur[0]];
  v.bulk_mask = h->bulk_mask;
  return v;
}
int handle_fail(pf_handle *h, int code, const std::string &msg) { return fail(h, code, msg); }
int set_global_error(int code, const std::string &msg) { return fail(nullptr, code, msg); }
void handle_march_shape(const pf_handle *h, int *rows, int *zchunk) {
  const bool used = h->diff_split && h->diff_vec && h->diff_march && h->diff_variant == 0;
  *rows = used ? h->march_by : 0;
  *zchunk = used ? h->march_zc : 0;
}

namespace {
template <int P>
hipError_t diffuse_io_p(pf_handle *h, const DiffConsts<double> &dc, const double *cin,
                        double *cout) {
  if (h->NF == 4) return launch_diffuse_io<FormF64, P, 4, false>(h, dc, cin, cout);
  if (h->d.bc_z == PF_BC_NOFLUX) return launch_diffuse_io<FormF64, P, 6, true>(h, dc, cin, cout);
  return launch_diffuse_io<FormF64, P, 6, false>(h, dc, cin, cout);
}
}  // namespace

int handle_diffuse_io(pf_handle *h, const double *cin, double *cout) {
  // the expression pf_step uses for the carbon sub-step factor (bit-identical results)
  const double factor = h->p.dt_d[0] / (h->p.dx * h->p.dx);
  const DiffConsts<double> dc = make_diff_consts<double>(h->p, 0, factor);
  const hipError_t e = h->d.nphase == 2 ? diffuse_io_p<2>(h, dc, cin, cout)
                                        : diffuse_io_p<3>(h, dc, cin, cout);
  if (e != hipSuccess) return fail(h, PF_ERR_HIP, std::string("diffuse: ") + hipGetErrorString(e));
  return PF_OK;
}
void handle_state_changed(pf_handle *h) {
  h->flag_valid = false;
  h->unit_valid = false;
  h->sweep_hist = false;
  tables_state_set(h);
}
PFConsts<double> handle_pf_consts(const pf_handle *h) { return make_pf_consts<double>(h); }
DiffConsts<double> handle_diff_consts(const pf_handle *h, int solute, double factor) {
  return make_diff_consts<double>(h->p, solute, factor);
}
