// pf_diffuse_common.h — device helpers of the diffusion sub-step kernels (pf_diffuse.hip): the
// quasi-equilibrium partition of one cell, face weights, the concentration update of both forms,
// the full-formula update of one cell and the 16-byte run types.
#pragma once
#include <type_traits>

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
      // the sums start from their first term: 0 + x differs from x only for x = -0, and a sum
      // that differs by the sign of a zero gives the same v below (t1 holds the term
      // ph[i] * k(i,i) > 0; c - (+0) == c - (-0) for every c other than -0)
      A t1 = o.ph[0] * dc.pk[i][0], t2 = o.ph[0] * dc.pc[i][0];
#pragma unroll
      for (int j = 1; j < P; ++j) {
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
// The correctly rounded FP32 square root for x = +0 and x >= 2^-96: the hardware square root
// (within 1 ulp) and its two neighbours, the neighbour below taken when x <= s_dn * s, the one
// above when x > s_up * s (residuals by fma).  This is the compiler's sequence for sqrtf without
// the two parts that change nothing on that range: the rescaling of inputs below 2^-96 and the
// select that returns +0 and +inf unchanged (for +0 the residuals keep s = +0: s_dn is a NaN and
// the +1 ulp residual is +0).  Bitwise equal to sqrtf there.
__device__ inline float sqrt_rn_unscaled(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float dn = __uint_as_float(__float_as_uint(s) - 1u);
  const float up = __uint_as_float(__float_as_uint(s) + 1u);
  float r = s;
  if (fmaf(-dn, s, x) <= 0.f) r = dn;
  if (fmaf(-up, s, x) > 0.f) r = up;
  return r;
}
// The weights of N faces, face_weight of each product.  FP32: the unscaled sequence above, which
// gives face_weight's bits for products that are 0, 1 or at least 2^-96 (sqrt(0) = 0 and
// sqrt(1) = 1, so the shortcut is not needed), unless a product of the wavefront lies in
// (0, 2^-96): then the whole wavefront takes face_weight.  One branch for the N faces, so the N
// square-root sequences interleave.
template <class A, int N> __device__ inline void face_weights(const A (&prod)[N], A (&w)[N]) {
  if constexpr (std::is_same<A, float>::value) {
    // bits - 1 is below bits(2^-96) - 1 exactly for the products in (0, 2^-96)
    uint32_t m = 0xffffffffu;
#pragma unroll
    for (int i = 0; i < N; ++i) m = min(m, __float_as_uint(prod[i]) - 1u);
    if (__builtin_expect(__ballot(m < 0x0f800000u - 1u) == 0, 1)) {
#pragma unroll
      for (int i = 0; i < N; ++i) w[i] = sqrt_rn_unscaled(prod[i]);
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) w[i] = face_weight<A>(prod[i]);
}

// Store q of cell i while other threads may store the other cells of its 32-bit word: each thread
// changes only its own four bits, by an atomic XOR with the difference.  Its own bits were last
// written by an earlier kernel (or are this thread's alone), so a plain load reads them.
__device__ inline void comp_put_atomic(float *buf, long n, long i, CompCode q) {
  uint32_t *w = reinterpret_cast<uint32_t *>(comp_bytes(buf, n)) + (i >> 3);
  const int sh = 4 * int(i & 7);
  const uint32_t d = ((*w >> sh) ^ uint32_t(q)) & 15u;
  if (d) atomicXor(w, d << sh);
}

// total_conc += del for cell i (solver.cpp:366) of stored value c, written to `out` (n cells per
// buffer).  Mixed form: Kahan-compensated with the cell's compensation code q, y = del - comp;
// t = c + y; comp = (t - c) - y; c = t, the new code stored after the n values of the buffer and
// rounded with the sub-step's key.
template <class F>
__device__ inline void conc_store(typename F::C *__restrict__ out, long i, long n, typename F::C c,
                                  typename F::A del, CompCode q, uint32_t key);
template <>
__device__ inline void conc_store<FormF64>(double *__restrict__ out, long i, long, double c,
                                           double del, CompCode, uint32_t) {
  out[i] = c + del;
}
// the mixed form's Kahan step: new value t and its code from value c, code q, increment del
__device__ inline float mixed_step(float c, float del, CompCode q, uint32_t h, CompCode &qo) {
  const float comp = comp_decode(q, c);
  const float y = del - comp;
  const float t = c + y;
  qo = comp_encode((t - c) - y, t, h);
  return t;
}
template <>
__device__ inline void conc_store<FormMixed>(float *__restrict__ out, long i, long n, float c,
                                             float del, CompCode q, uint32_t key) {
  CompCode qo;
  out[i] = mixed_step(c, del, q, comp_hash(uint32_t(i), key), qo);
  comp_put_atomic(out, n, i, qo);
}
// the same with the code read from the input buffer `in`
template <class F>
__device__ inline void conc_update(const typename F::C *__restrict__ in,
                                   typename F::C *__restrict__ out, long i, long n,
                                   typename F::C c, typename F::A del, uint32_t key) {
  CompCode q = 0;
  if constexpr (F::kMixed) q = comp_load(in, n, i);
  conc_store<F>(out, i, n, c, del, q, key);
}

// The stored values the reference's sub-step of one cell reads: its own and its stencil's
// concentration and value_phase (and, mixed, its compensation code), loaded before any arithmetic
// so that a thread can have the loads of several cells in flight.
template <class F, int P, int NF> struct StencilLoad {
  RawCell<F, P> me, nb[NF];
  CompCode q;
};
template <class F, int P, int NF>
__device__ inline void stencil_load(const typename F::C *__restrict__ cin,
                                    const typename F::T *__restrict__ pfrac, long i, long n,
                                    const long (&nbi)[6], StencilLoad<F, P, NF> &s) {
  s.me.c = cin[i];
#pragma unroll
  for (int p = 0; p < P; ++p) s.me.ph[p] = pfrac[i * P + p];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    s.nb[f].c = cin[nbi[f]];
#pragma unroll
    for (int p = 0; p < P; ++p) s.nb[f].ph[p] = pfrac[nbi[f] * P + p];
  }
  s.q = 0;
  if constexpr (F::kMixed) s.q = comp_load(cin, n, i);
}
// A cell's diffusion increment, accumulated face by face (solver.cpp:352-366).  FP64 keeps the
// reference's order: del = sum over faces and phases of w * (c_nb - c) * D_p, then del *= r.  The
// mixed form sums w * (c_nb - c) per phase over the faces and adds those sums times D_p * r (one
// product, rounded once), a multiply per face and phase fewer.  For a bulk cell of phase p the
// other phases add 0 * x = +-0, so the mixed full formula gives (sum_f (c_f - c)) * (D_p * r)
// bitwise, the bulk kernels' form (BulkAcc), as the reference's two forms agree in FP64.
template <class F, int P> struct DelAcc {
  using A = typename F::A;
  static constexpr int NS = F::kMixed ? P : 1;
  A s[NS];
  __device__ DelAcc() {
#pragma unroll
    for (int p = 0; p < NS; ++p) s[p] = A(0);
  }
  // the term of phase p on one face: w * (c_nb - c)
  __device__ void add(int p, A term, const DiffConsts<A> &dc) {
    if constexpr (F::kMixed) s[p] += term;
    else s[0] += term * dc.D[p];
  }
  // the same for the first face: the mixed form's per-phase sums start from it (0 + x is x but
  // for x = -0, and a del that differs only by the sign of a zero gives the same value and code)
  __device__ void first(int p, A term, const DiffConsts<A> &dc) {
    if constexpr (F::kMixed) s[p] = term;
    else s[0] += term * dc.D[p];
  }
  __device__ A result(const DiffConsts<A> &dc) const {
    if constexpr (F::kMixed) {
      A del = s[0] * (dc.D[0] * dc.r);
#pragma unroll
      for (int p = 1; p < P; ++p) del += s[p] * (dc.D[p] * dc.r);
      return del;
    } else {
      return s[0] * dc.r;
    }
  }
};
// The bulk-cell form, one phase's D_p for every face: FP64 sum_f (c_f - c) * D_p, then * r; mixed
// (sum_f (c_f - c)) * Dpr with Dpr = D_p * r.
template <class F> struct BulkAcc {
  using A = typename F::A;
  A s = A(0);
  __device__ void add(A diff, A Dp) {
    if constexpr (F::kMixed) s += diff;
    else s += diff * Dp;
  }
  // the first face's term: the sum starts from it (0 + x is x but for x = -0, and a sum that
  // differs only by the sign of a zero gives the same update)
  __device__ void first(A diff, A Dp) {
    if constexpr (F::kMixed) s = diff;
    else s = diff * Dp;
  }
  __device__ A result(A r, A Dpr) const {
    if constexpr (F::kMixed) return s * Dpr;
    else return s * r;
  }
};

// partition (solver.cpp:389-405), weighted flux (solver.cpp:352-365): the cell's increment
template <class F, int P, int NF>
__device__ inline typename F::A stencil_del(const StencilLoad<F, P, NF> &s,
                                            const DiffConsts<typename F::A> &dc) {
  using A = typename F::A;
  CellPF<A> me;
  partition_cell<F, P>(s.me, dc, me);
  CellPF<A> nb[NF];
  A prod[NF * P], w[NF * P];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    partition_cell<F, P>(s.nb[f], dc, nb[f]);
#pragma unroll
    for (int p = 0; p < P; ++p) prod[f * P + p] = nb[f].ph[p] * me.ph[p];
  }
  face_weights<A>(prod, w);
  DelAcc<F, P> acc;
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const A term = w[f * P + p] * (nb[f].c[p] - me.c[p]);
      if (f == 0) acc.first(p, term, dc);
      else acc.add(p, term, dc);
    }
  return acc.result(dc);
}
// ... and the update
template <class F, int P, int NF>
__device__ inline void stencil_update(typename F::C *__restrict__ cout, long i, long n,
                                      const StencilLoad<F, P, NF> &s,
                                      const DiffConsts<typename F::A> &dc) {
  conc_store<F>(cout, i, n, s.me.c, stencil_del<F, P, NF>(s, dc), s.q, dc.key);
}

// The reference's sub-step for one cell from its own and its stencil's stored concentration and
// value_phase.
template <class F, int P, int NF>
__device__ inline void full_update(const typename F::C *__restrict__ cin,
                                   typename F::C *__restrict__ cout,
                                   const typename F::T *__restrict__ pfrac, long i, long n,
                                   const long (&nbi)[6], const DiffConsts<typename F::A> &dc) {
  StencilLoad<F, P, NF> s;
  stencil_load<F, P, NF>(cin, pfrac, i, n, nbi, s);
  stencil_update<F, P, NF>(cout, i, n, s, dc);
}

template <class A, int V> struct alignas(sizeof(A) * V) VecA {
  A v[V];
};
template <int V> struct alignas(V) VecFlag {
  uint8_t f[V];
};
template <int V> struct VecCode {  // decoded compensation codes of a run
  CompCode q[V];
};

}  // namespace pfamd
