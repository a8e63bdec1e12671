// pf_sweep_impl.h — the phase-field sweep kernel (definition; instantiated in pf_sweep_f64.hip and
// pf_sweep_mixed.hip).
//
// One thread per cell (Phase_field_solver, C++-openmp-cuda/src/solver.cpp:147-304):
//  1. active set: own slots with phi > cutoff at their slot position; then the slots of the
//     neighbours (-x,+x,-y,+y,-z,+z; first 2*dims faces; a zero id ends a neighbour's list) are
//     matched by id (first match) or appended while fewer than 10 entries exist.
//  2. a single entry copies the slots unchanged (value_phase untouched).
//  3. per entry a: sum over b != a (phases not 1<->3) of
//       M_ab * ( sigma_ab*[(S_a phi_b - S_b phi_a)/dx^2 + pi^2/(2 eta^2)(phi_a - phi_b)]
//                + pi/eta sqrt(phi_a phi_b) dG_ab ),
//     Euler step, clamp to [0,1], zero below cutoff.
//  4. selection sort (descending, first maximum wins), keep K, value_phase, renormalise by
//     multiplying with 1/sum.
// ISO variant (Iso_growth, MicrostructureConstruction.cpp:130-248): neighbour entries are used
// only when phi > cutoff, the pair term has no sigma/M/driving force, the Euler factor is
// 0.5/dims*dx*dx (passed as dt) and no phase pair is skipped.
// The mixed form accumulates the per-entry update, clamps, prunes, sorts and renormalises in FP64.
//
// The per-thread union lists (16-bit ids, neighbour sums, updated values) live in LDS as
// [entry][thread] arrays so the data-dependent match/append/sort indices stay cheap; an entry's
// own value is read from the thread's slot registers.
#pragma once
#include <type_traits>

#include "pf_kernels.h"

namespace pfamd {

// table value for the phase pair (pa, pb), 1-based, zero when either phase is 0 (empty slot)
template <class A, int P>
__device__ inline A pair_pick(const A (&tab)[4][4], int pa, int pb) {
  A r = A(0);
#pragma unroll
  for (int i = 1; i <= P; ++i)
#pragma unroll
    for (int j = 1; j <= P; ++j)
      if (pa == i && pb == j) r = tab[i][j];
  return r;
}

// LIST: the cells come from list[0 .. *count) (grid-stride); otherwise one cell per thread.
template <class F, int K, int P, int NF, bool NOFLUX, bool ISO, bool LIST, int MODE>
__global__ __launch_bounds__(kSweepBS) void pf_sweep_kernel(
    const uint16_t *__restrict__ id1, const typename F::T *__restrict__ phi1,
    uint16_t *__restrict__ id2, typename F::T *__restrict__ phi2,
    typename F::T *__restrict__ pfrac, const typename F::C *__restrict__ conc0,
    const PFConsts<typename F::A> cs, const int nx, const int ny, const int nz,
    const long ncell, const int *__restrict__ list, const int *__restrict__ count,
    int *__restrict__ queue, uint8_t *__restrict__ sflag, uint8_t *__restrict__ unit,
    const unsigned umask, int *__restrict__ olist) {
  static_assert(MODE == 0 || (LIST && K == 4 && !ISO), "the split walk is for K = 4 lists");
  using T = typename F::T;
  using A = typename F::A;
  __shared__ uint16_t sLg[kStencilMax][kSweepBS];
  __shared__ T sV2[kStencilMax][kSweepBS];
  __shared__ double sDel[kStencilMax][kSweepBS];
  const int t = threadIdx.x;
  // the cell's new row: ids, slots renormalised by *= 1./sum, value_phase; the listed sweep also
  // records its own-slot summary for the next classification (slot_summary), from the stored
  // values compared as the classification compares them
  auto store_cell = [&](const long cell, const uint32_t (&go)[K], const double (&out)[K],
                        const double (&vp)[4], const double sum) {
    const double inv = 1. / sum;
    store_ids<K>(id2, cell, go);
    int above = 0, slot = 0;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const T v = T(out[s] * inv);
      phi2[cell * K + s] = v;
      if (double(v) > cs.cutoff) {
        ++above;
        slot = s;
      }
    }
    int up = 0, ones = 0, zeros = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const T v = T(vp[p + 1] * inv);
      pfrac[cell * P + p] = v;
      if (v == T(1)) {
        ++ones;
        up = p + 1;
      } else if (v == T(0)) {
        ++zeros;
      }
    }
    if (LIST) {
      sflag[cell] = slot_summary(above, slot);
      // the diffusion flag pass's unit-phase byte of the new value_phase (unit_phase_kernel's
      // rule), so that pass need not read value_phase again after a split sweep
      unit[cell] = uint8_t(ones == 1 && zeros == P - 1 && ((umask >> up) & 1u) ? up : 0);
    }
  };
  // a cell whose union has one entry keeps its row (solver.cpp:207-216)
  auto copy_cell = [&](const long cell, const T (&v0)[K]) {
    using V = typename IdRowT<K>::V;
    reinterpret_cast<V *>(id2)[cell] = reinterpret_cast<const V *>(id1)[cell];
    int above = 0, slot = 0;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      phi2[cell * K + s] = v0[s];
      if (double(v0[s]) > cs.cutoff) {
        ++above;
        slot = s;
      }
    }
    if (LIST) sflag[cell] = slot_summary(above, slot);
  };
  // no block-wide barrier below: threads run their cells independently
  auto body = [&](const long cell) {
  int x, y, z;
  cell_coords(cell, nx, ny, x, y, z);
  const double cutoff = cs.cutoff;

  long nb[6];
  {
    const long plane = long(nx) * ny;
    const long row = long(nx) * y + plane * z;
    nb[0] = wrap(x - 1, nx) + row;
    nb[1] = wrap(x + 1, nx) + row;
    nb[2] = x + long(nx) * wrap(y - 1, ny) + plane * z;
    nb[3] = x + long(nx) * wrap(y + 1, ny) + plane * z;
    int zm, zp;
    if (NOFLUX) {
      zm = z == 0 ? z : z - 1;
      zp = z == nz - 1 ? z : z + 1;
    } else {
      zm = wrap(z - 1, nz);
      zp = wrap(z + 1, nz);
    }
    nb[4] = x + long(nx) * y + plane * zm;
    nb[5] = x + long(nx) * y + plane * zp;
  }
  // Every load of the cell is issued here, before any of them is used: its own and its
  // neighbours' whole id and phi rows (vectors, unconditionally; only the slots before a row's first
  // zero id are used), its value_phase and concentration.  One round of memory latency per cell
  // instead of one per face.
  uint32_t g0[K], gn[NF][K];
  T v0[K], vn[NF][K];
  load_ids<K>(id1, cell, g0);
  load_row<T, K>(phi1, cell, v0);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    load_ids<K>(id1, nb[f], gn[f]);
    load_row<T, K>(phi1, nb[f], vn[f]);
  }
  A ph[P], c = A(0);
  if (!ISO) {
#pragma unroll
    for (int j = 0; j < P; ++j) ph[j] = A(pfrac[cell * P + j]);
    c = A(conc_hi(conc0[cell]));
  }

  // local_conc (value_conc1 of this cell before the update) and the driving force per phase
  // pair (Driving_force, solver.cpp:374-387: only the concentration of the first phase enters)
  auto driving_forces = [&](A (&dG)[4][4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) dG[a][b] = A(0);
    if (!ISO) {
      A lc[4];
      lc[0] = A(0);
#pragma unroll
      for (int i = 1; i <= P; ++i) {
        A t1 = A(0), t2 = A(0);
#pragma unroll
        for (int j = 0; j < P; ++j) {
          t1 += ph[j] * cs.pk[i][j + 1];
          t2 += ph[j] * cs.pc[i][j + 1];
        }
        lc[i] = (c - t2) / t1;
      }
#pragma unroll
      for (int a = 1; a <= P; ++a)
#pragma unroll
        for (int b = 1; b <= P; ++b) {
          A df = A(0);
          if (a != b) df += cs.ent[a][b] * (lc[a] - cs.eqc[a][b]) / cs.slope[a][b];
          dG[a][b] = df;
        }
    }
  };
  // the entry values of the own slots: an entry's own value is the slot value at its position
  // when that slot is above the cutoff (positions >= K hold neighbour-only entries: 0)
  auto v1_of = [&](int e) -> T {
    T r = T(0);
#pragma unroll
    for (int s = 0; s < K; ++s)
      if (e == s && double(v0[s]) > cutoff) r = v0[s];
    return r;
  };
  // the pair sum of entry a (solver.cpp:232-272): sum over b != a, in order, then Euler step,
  // clamp and prune; Lb/V2b/V1b give entry b's id, neighbour sum and own value
  auto entry_update = [&](int pa, A v1a, A v2a, int a, int nb_, const A (&dG)[4][4], auto Lb,
                          auto V2b, auto V1b) -> double {
    double acc = 0.0;
    for (int b = 0; b < nb_; ++b) {
      if (b == a) continue;
      const A v1b = V1b(b), v2b = V2b(b);
      if (ISO) {
        const A tt = (v2a * v1b - v2b * v1a) * cs.f1 + (v1a - v1b) * cs.f2;
        acc += double(tt);
      } else {
        const int pb = code_phase(Lb(b));
        if (pa - pb == 2 || pb - pa == 2) continue;
        A tt = (v2a * v1b - v2b * v1a) * cs.f1;
        tt += (v1a - v1b) * cs.f2;
        tt *= pair_pick<A, P>(cs.sig, pa, pb);
        tt += cs.f3 * sqrt(v1a * v1b) * pair_pick<A, P>(dG, pa, pb);
        tt *= pair_pick<A, P>(cs.M, pa, pb);
        acc += double(tt);
      }
    }
    acc *= cs.dt;
    acc += double(v1a);
    if (acc > 1.0)
      acc = 1.0;
    else if (acc < cutoff)
      acc = 0.0;
    return acc;
  };

  if constexpr (K == 4 && !ISO && MODE != 2) {
    // Registers for a union of at most four entries (almost every listed cell: two grains at a
    // grain boundary, three at a triple line or an austenite front), the same algorithm as the
    // general path below: own slots at their slot positions, first-match search, appends at
    // position n, pair sums in entry order, first-maximum selection sort.  With K = 4 the own
    // entries lie in positions 0..3 and an append reaches position 4 only as the fifth entry, so
    // until then positions 4..9 are empty and searching them finds nothing; a cell whose union
    // would need a fifth entry takes the general path from the start.
    uint32_t L[4];
    T S2[4];
    int m = 0;
    bool over = false;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool own = double(v0[e]) > cutoff;
      L[e] = own ? g0[e] : 0u;
      S2[e] = T(0);
      m += own ? 1 : 0;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      bool live = true;
#pragma unroll
      for (int sl = 0; sl < K; ++sl) {
        const uint32_t g = gn[f][sl];
        live = live && (g != 0);
        if (live) {
          const T v = vn[f][sl];
          const int kk = L[0] == g ? 0 : L[1] == g ? 1 : L[2] == g ? 2 : L[3] == g ? 3 : -1;
          if (kk >= 0) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (e == kk) S2[e] += v;
          } else if (m < 4) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (e == m) {
                L[e] = g;
                S2[e] += v;
              }
            ++m;
          } else {
            over = true;
          }
        }
      }
    }
    if (!over) {
      if (m == 1) {
        copy_cell(cell, v0);
        return;
      }
      A dG[4][4];
      driving_forces(dG);
      // the pair sums with entry indices known at compile time (no selects into the arrays)
      A V1[4], V2[4];
      int PH[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        V1[e] = A(double(v0[e]) > cutoff ? v0[e] : T(0));
        V2[e] = A(S2[e]);
        PH[e] = code_phase(L[e]);
      }
      double D[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        D[a] = 0.0;
        if (a < m) {
          double acc = 0.0;
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            if (b == a || b >= m) continue;
            const int pa = PH[a], pb = PH[b];
            if (pa - pb == 2 || pb - pa == 2) continue;
            A tt = (V2[a] * V1[b] - V2[b] * V1[a]) * cs.f1;
            tt += (V1[a] - V1[b]) * cs.f2;
            tt *= pair_pick<A, P>(cs.sig, pa, pb);
            tt += cs.f3 * sqrt(V1[a] * V1[b]) * pair_pick<A, P>(dG, pa, pb);
            tt *= pair_pick<A, P>(cs.M, pa, pb);
            acc += double(tt);
          }
          acc *= cs.dt;
          acc += double(V1[a]);
          if (acc > 1.0)
            acc = 1.0;
          else if (acc < cutoff)
            acc = 0.0;
          D[a] = acc;
        }
      }
      // Sort_xyz's first K passes over positions 0..3; positions 4..9 hold 0, which never
      // exceeds a value that survived the clamp (0 or >= cutoff), so they are never chosen
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        int mi = a;
        double best = D[a];
#pragma unroll
        for (int b = a + 1; b < 4; ++b)
          if (D[b] > best) {
            best = D[b];
            mi = b;
          }
        const double da = D[a];
        const uint32_t la = L[a];
        uint32_t lm = la;
#pragma unroll
        for (int b = a + 1; b < 4; ++b)
          if (mi == b) lm = L[b];
#pragma unroll
        for (int b = a + 1; b < 4; ++b)
          if (mi == b) {
            D[b] = da;
            L[b] = la;
          }
        D[a] = best;
        L[a] = lm;
      }
      double sum = 0.0, vp[4] = {0.0, 0.0, 0.0, 0.0};
      uint32_t go[K];
      double out[K];
#pragma unroll
      for (int s = 0; s < K; ++s) {
        const double d = D[s];
        if (d > cutoff) {
          go[s] = L[s];
          out[s] = d;
          const int p = code_phase(go[s]);
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (q == p) vp[q] += d;
          sum += d;
        } else {
          go[s] = 0;
          out[s] = 0.0;
        }
      }
      store_cell(cell, go, out, vp, sum);
      return;
    }
    if constexpr (MODE == 1) {
      // The general path is a kernel of its own (MODE 2), so this one keeps no LDS arrays and
      // fewer registers: the cell goes to the overflow list, one atomic per wavefront.
      const unsigned long long act = __ballot(1);
      const int lane = t & 63, leader = __ffsll(static_cast<long long>(act)) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(olist + ncell, __popcll(act));
      base = __shfl(base, leader);
      olist[base + __popcll(act & ((1ull << lane) - 1ull))] = int(cell);
      return;
    }
  }

  for (int e = 0; e < kStencilMax; ++e) {
    sLg[e][t] = 0;
    sV2[e][t] = T(0);
  }
  int n = 0;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (double(v0[s]) > cutoff) {
      ++n;
      sLg[s][t] = uint16_t(g0[s]);
    }
  }

#pragma unroll
  for (int f = 0; f < NF; ++f) {
    bool live = true;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const uint32_t g = gn[f][s];
      live = live && (g != 0);
      if (live) {
        const T v = vn[f][s];
        if (!ISO || double(v) > cutoff) {
          int kk = 0;
          for (; kk < kStencilMax; ++kk)
            if (sLg[kk][t] == g) break;
          if (kk < kStencilMax) {
            sV2[kk][t] += v;
          } else if (n < kStencilMax) {
            sLg[n][t] = uint16_t(g);
            sV2[n][t] += v;
            ++n;
          }
        }
      }
    }
  }

  if (n == 1) {
    copy_cell(cell, v0);
    return;
  }

  A dG[4][4];
  driving_forces(dG);

  for (int a = 0; a < kStencilMax; ++a) sDel[a][t] = 0.0;
  for (int a = 0; a < n; ++a)
    sDel[a][t] = entry_update(code_phase(sLg[a][t]), A(v1_of(a)), A(sV2[a][t]), a, n, dG,
                              [&](int b) -> uint32_t { return sLg[b][t]; },
                              [&](int b) -> A { return A(sV2[b][t]); },
                              [&](int b) -> A { return A(v1_of(b)); });

  // Sort_xyz (solver.cpp:426-432): only the first K positions are read afterwards, and the
  // first K passes of the selection sort fix them.
  for (int a = 0; a < K; ++a) {
    int m = a;
    double best = sDel[a][t];
    for (int b = a + 1; b < kStencilMax; ++b) {
      const double v = sDel[b][t];
      if (v > best) {
        best = v;
        m = b;
      }
    }
    if (m != a) {
      sDel[m][t] = sDel[a][t];
      sDel[a][t] = best;
      const uint16_t tg = sLg[a][t];
      sLg[a][t] = sLg[m][t];
      sLg[m][t] = tg;
    }
  }

  double sum = 0.0, vp[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t go[K];
  double out[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const double d = sDel[s][t];
    if (d > cutoff) {
      go[s] = sLg[s][t];
      out[s] = d;
      const int p = code_phase(go[s]);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q == p) vp[q] += d;
      sum += d;
    } else {
      go[s] = 0;
      out[s] = 0.0;
    }
  }
  store_cell(cell, go, out, vp, sum);
  };
  if (MODE == 2) {  // the overflow list: grid-stride
    const int n = *count;
    for (int k = blockIdx.x * kSweepBS + t; k < n; k += gridDim.x * kSweepBS) body(list[k]);
  } else if (LIST) {
    // Workgroup b runs on XCD b % 8: each XCD walks a contiguous eighth of the list (row-tile
    // order, all its workgroups resident, one front of consecutive entries), so the neighbour
    // rows of a workgroup's cells are in its own XCD's L2.  (Fewer than 8 workgroups: one plain
    // grid-stride walk, so that every range has a walker.)
    const int n = *count;
    if (gridDim.x < 8) {
      for (int k = blockIdx.x * kSweepBS + t; k < n; k += gridDim.x * kSweepBS) body(list[k]);
    } else if (queue) {
      // dynamic: each wavefront takes the next 64 entries of its XCD's eighth from that XCD's
      // counter (queue[32 * xcd], zeroed before the launch), so the entries in process stay one
      // compact front however unevenly the cells' work is spread
      const int xcd = int(blockIdx.x & 7u), lane = t & 63;
      const int lo = int((long(n) * xcd) / 8), hi = int((long(n) * (xcd + 1)) / 8);
      for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&queue[32 * xcd], 64);
        base = __shfl(base, 0);
        if (lo + base >= hi) break;  // wave-uniform; the counter only grows, so every wave ends
        if (lo + base + lane < hi) body(list[lo + base + lane]);
      }
    } else {
      const int xcd = int(blockIdx.x & 7u), rank = int(blockIdx.x >> 3);
      const int nb_x = (int(gridDim.x) - xcd + 7) / 8;  // workgroups on this XCD (>= 1)
      const int lo = int((long(n) * xcd) / 8), hi = int((long(n) * (xcd + 1)) / 8);
      for (int k = lo + rank * kSweepBS + t; k < hi; k += nb_x * kSweepBS) body(list[k]);
    }
  } else {
    const long cell = long(blockIdx.x) * kSweepBS + t;
    if (cell < ncell) body(cell);
  }
}

// Copy cells of the sweep (solver.cpp:207-216: a cell whose active-set union holds a single entry
// keeps its slots, value_phase untouched).  The union has one entry exactly when at most one own
// slot is above the cutoff and every neighbour id before its list's first zero is that grain (or,
// with no own entry, all those ids are one grain).  Copy cells are written here; the others get
// flag 0 for the listed full sweep.  With `history` (the previous sweep was split like this one
// and nothing has changed the state since) sflag holds what that sweep left: bit 0 for a cell it
// copied, whose destination row then already equals its source row (the buffers swap every sweep
// and a copy makes them equal), so the write is skipped; bits 1-5 the own-slot summary
// (slot_summary), so the slots are not read at all.
// The FP32 form compares its slots with cutoff_f, the largest float not above the cutoff: for a
// float v, v > cutoff_f exactly when double(v) > cutoff.
template <class F, int K, int NF, bool NOFLUX>
__global__ __launch_bounds__(256) void sweep_classify_kernel(
    const uint16_t *__restrict__ id1, const typename F::T *__restrict__ phi1,
    uint16_t *__restrict__ id2, typename F::T *__restrict__ phi2, uint8_t *__restrict__ sflag,
    const int history, const double cutoff, const float cutoff_f, const int nx, const int ny,
    const int nz, const long ncell) {
  using T = typename F::T;
  // Workgroup b runs on XCD b % 8.  With the grid a multiple of 128 workgroups, XCD j takes the
  // groups of 16 consecutive workgroups (4096 cells: 8 rows at nx = 512) numbered j, j + 8, ...:
  // a cell's -y/+y neighbour rows are then in its own XCD's L2 (in plain order they belong to
  // other XCDs), and so are its -z/+z rows, one eighth of a plane away on that XCD.
  long blk = blockIdx.x;
  if ((gridDim.x & 127u) == 0) {
    const long k = blockIdx.x >> 3;
    blk = (((k >> 4) * 8 + (blockIdx.x & 7u)) << 4) | (k & 15);
  }
  const long cell = blk * 256 + threadIdx.x;
  if (cell >= ncell) return;
  // Every load is issued before any of them is used: the own ids and slots, the history byte and
  // (K = 4) the six neighbour id rows, so a thread waits on memory once.
  // With history the byte the previous sweep left (slot_summary) says how many own slots are above
  // the cutoff and which one, so the slots themselves are read only to write a copy.
  uint32_t g0[K];
  T v0[K];
  load_ids<K>(id1, cell, g0);
  if (!history) load_row<T, K>(phi1, cell, v0);
  const uint8_t prev = sflag[cell];
  // the six neighbours (wrap(x - 1, nx) + nx * y + plane * z and so on) as 32-bit offsets from the
  // cell: a domain holds fewer than 2^31 cells
  int x, y, z;
  cell_coords(cell, nx, ny, x, y, z);
  const uint32_t c = uint32_t(cell), unx = uint32_t(nx), plane = unx * uint32_t(ny);
  const uint32_t vol = plane * uint32_t(nz);
  uint32_t nb[6];
  nb[0] = x == 0 ? c + unx - 1u : c - 1u;
  nb[1] = x == nx - 1 ? c + 1u - unx : c + 1u;
  nb[2] = y == 0 ? c + plane - unx : c - unx;
  nb[3] = y == ny - 1 ? c + unx - plane : c + unx;
  nb[4] = z == 0 ? (NOFLUX ? c : c + vol - plane) : c - plane;
  nb[5] = z == nz - 1 ? (NOFLUX ? c : c + plane - vol) : c + plane;
  uint64_t w[NF];
  if constexpr (K == 4) {
    const uint64_t *rows = reinterpret_cast<const uint64_t *>(id1);
#pragma unroll
    for (int f = 0; f < NF; ++f) w[f] = rows[nb[f]];
  }
  // keep the arithmetic on the loaded values below every load (the scheduler would otherwise
  // move a compare of the own slots above the neighbour loads and wait for it there)
  __builtin_amdgcn_sched_barrier(0);
  int n0 = 0, slot = 0;
  uint32_t G = 0;
  if (history) {
    n0 = (prev & 2) ? ((prev & 4) ? 1 : 0) : 2;  // 2: two or more
    slot = (prev >> 3) & 7;
#pragma unroll
    for (int s = 0; s < K; ++s)
      if (n0 == 1 && s == slot) G = g0[s];
  } else {
#pragma unroll
    for (int s = 0; s < K; ++s) {
      bool above;
      if constexpr (std::is_same<T, float>::value) above = v0[s] > cutoff_f;
      else above = double(v0[s]) > cutoff;
      if (above) {
        ++n0;
        G = g0[s];
        slot = s;
      }
    }
  }
  bool copy = n0 <= 1;
  if constexpr (K == 4) {
    // A neighbour's ids are read in slot order up to its first 0, and each one read must be G
    // (with no own entry, G is the first id read: slot 0 of the first neighbour whose slot 0 is not
    // 0).  Four ids are one 64-bit row (id_row_breaks).  Evaluated for every cell, without a branch.
    if (n0 == 0) {
#pragma unroll
      for (int f = NF - 1; f >= 0; --f)
        if ((w[f] & 0xffffu) != 0u) G = uint32_t(w[f] & 0xffffu);
    }
    bool bad = !(n0 == 1 || G != 0u);  // no id at all: not a single-entry union
#pragma unroll
    for (int f = 0; f < NF; ++f) bad = bad | id_row_breaks(w[f], G);
    copy = copy && !bad;
  } else if (copy) {
    bool any = n0 == 1;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      uint32_t gn[K];
      load_ids<K>(id1, nb[f], gn);
      bool live = true;
#pragma unroll
      for (int s = 0; s < K; ++s) {
        live = live && gn[s] != 0;
        if (live) {
          if (!any) {
            G = gn[s];
            any = true;
          } else if (gn[s] != G) {
            copy = false;
          }
        }
      }
    }
    copy = copy && any;
  }
  if (copy) {
    if (!(history && (prev & 1))) {  // not copied by the previous sweep: the rows differ
      using V = typename IdRowT<K>::V;
      reinterpret_cast<V *>(id2)[cell] = reinterpret_cast<const V *>(id1)[cell];
      if (history) load_row<T, K>(phi1, cell, v0);
#pragma unroll
      for (int s = 0; s < K; ++s) phi2[cell * K + s] = v0[s];
    }
    sflag[cell] = uint8_t(1u | slot_summary(n0, slot));
  } else {
    sflag[cell] = 0;
  }
}

#define PF_INST_SWEEP_M(F, K, P, NF, NOF, ISO, LIST, MODE)                                       \
  template __global__ void pf_sweep_kernel<F, K, P, NF, NOF, ISO, LIST, MODE>(                   \
      const uint16_t *, const F::T *, uint16_t *, F::T *, F::T *, const F::C *,                 \
      const PFConsts<F::A>, int, int, int, long, const int *, const int *, int *, uint8_t *,    \
      uint8_t *, unsigned, int *);
#define PF_INST_SWEEP(F, K, P, NF, NOF, ISO, LIST) PF_INST_SWEEP_M(F, K, P, NF, NOF, ISO, LIST, 0)
#define PF_INST_SWEEP_ALL(F, K, P, ISO)          \
  PF_INST_SWEEP(F, K, P, 4, false, ISO, false)   \
  PF_INST_SWEEP(F, K, P, 6, false, ISO, false)   \
  PF_INST_SWEEP(F, K, P, 6, true, ISO, false)
#define PF_INST_SWEEP_LIST(F, K, P)             \
  PF_INST_SWEEP(F, K, P, 4, false, false, true)  \
  PF_INST_SWEEP(F, K, P, 6, false, false, true)  \
  PF_INST_SWEEP(F, K, P, 6, true, false, true)
#define PF_INST_SWEEP_CLASSIFY(F, K)                                                        \
  template __global__ void sweep_classify_kernel<F, K, 4, false>(                          \
      const uint16_t *, const F::T *, uint16_t *, F::T *, uint8_t *, int, double, float, int, \
      int, int, long);                                                                     \
  template __global__ void sweep_classify_kernel<F, K, 6, false>(                          \
      const uint16_t *, const F::T *, uint16_t *, F::T *, uint8_t *, int, double, float, int, \
      int, int, long);                                                                     \
  template __global__ void sweep_classify_kernel<F, K, 6, true>(                           \
      const uint16_t *, const F::T *, uint16_t *, F::T *, uint8_t *, int, double, float, int, \
      int, int, long);

}  // namespace pfamd
