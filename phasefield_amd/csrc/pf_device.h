// pf_device.h — device-side types shared by the gfx950 kernels.
//
// Native HBM layout (one domain, N = nx*ny*nz cells, cell n = x + nx*(y + ny*z), x contiguous):
//   slot ids    uint16 [N][KI]   KI = 4 for K = 4, 8 for K = 6 (padded, 16-byte rows)
//   slot phi    T      [N][K]    double (FP64 form) or float (mixed form), double-buffered
//   phase frac  T      [N][P]    value_phase, P = 2 or 3 phases carried
//   conc        C      [S][N]    per solute, double-buffered (Jacobi sub-steps):
//                                FP64: double [N];  mixed: float value [N] followed by the 4-bit
//                                Kahan compensation codes, two cells per byte, in the same
//                                buffer (4.5 B per cell)
// A slot id is a 16-bit grain code: bits 15..14 = phase (1..3), bits 13..0 = grain serial;
// 0 = empty slot.  The host maps the reference's int32 grain ids (100000+n, 200000+g) to codes
// and back; the step only ever compares codes for equality, so results do not depend on it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfamd {

constexpr double kPi = 3.1415926;  // the reference's truncated pi (constant.h:7)
constexpr int kStencilMax = 10;    // union capacity (solver.cpp:148)

__host__ __device__ inline int code_phase(uint32_t code) { return int(code >> 14); }

// Storage/arithmetic forms.
struct FormF64 {
  using T = double;    // phi, value_phase
  using A = double;    // pair terms, partition, diffusion flux
  using C = double;    // stored total concentration
  static constexpr bool kMixed = false;
};
struct FormMixed {
  using T = float;
  using A = float;
  using C = float;     // stored value; its Kahan compensation code follows the N values
  static constexpr bool kMixed = true;
};

template <class C> __host__ __device__ inline C conc_hi(C c) { return c; }

// Mixed-form Kahan compensation, stored in four bits per cell: comp = q * 2^(e - 27) where
// c = m * 2^e (frexp, 0.5 <= |m| < 1), i.e. a multiple of 2^-3 ulp(c) (ulp(c) = 2^(e - 24)).
// |comp| <= ulp/2 after a Kahan update, so |q| <= 4 with 2x headroom; q saturates at [-8, 7].
// The effective concentration is c - comp.  The codes of a buffer of n cells follow its n float
// values, two cells per byte: cell i in bits 4 * (i & 1) of byte i / 2.
//
// q is rounded stochastically and without bias: with X = comp in units of 2^-24 quanta (rounded to
// the nearest integer, saturated to [-8, 7] quanta) and a threshold U = the high 24 bits of a
// per-cell hash of the sub-step, q = floor((X + U) / 2^24), computed in integers.  Over uniformly
// distributed U the mean of q is X / 2^24 exactly.  (Thresholds with fewer fraction bits than the
// compensation round down more often than up: 16-bit thresholds leave the stored compensation low
// by up to 2^-16 quanta in every cell with a nonzero increment, a systematic mass gain that grows
// as the diffusion front reaches more cells.)  Rounding to nearest would drop every increment below
// half a quantum, and those losses do not cancel between neighbours of different exponents.  A
// compensation already on the grid is kept by every threshold; threshold 0 rounds down.
using CompCode = int8_t;  // q, sign-extended from its four stored bits
constexpr int kCompShift = 27;
constexpr int kCompMin = -8, kCompMax = 7;
__host__ __device__ inline int comp_exp(float c) {
  int e = 0;
  (void)frexpf(c, &e);
  return e;
}
__host__ __device__ inline float comp_decode(CompCode code, float c) {
  return ldexpf(float(code), comp_exp(c) - kCompShift);
}
// q of four stored bits (the low nibble of `bits`)
__host__ __device__ inline CompCode comp_unpack(uint32_t bits) {
  return CompCode(int32_t(bits << 28) >> 28);
}
// 32-bit integer mixer (xorshift-multiply rounds of the "lowbias32" form)
__host__ __device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
// key of one sub-step of solute s: k counts the sub-steps applied to that solute since the state
// was set; the rounding hash of cell i is key ^ (i * 0x9E3779B9) (comp_hash).  Over the sub-steps
// of one cell its thresholds are as uniform as the keys' high 24 bits, which is all the unbiased
// rounding needs.
__host__ __device__ inline uint32_t comp_step_key(uint32_t k, int s) {
  return mix32(k * 2u + uint32_t(s));
}
constexpr uint32_t kCompCellMul = 0x9E3779B9u;
__host__ __device__ inline uint32_t comp_hash(uint32_t i, uint32_t key) {
  return key ^ (i * kCompCellMul);
}
// the code of compensation comp of value c with rounding hash h (h = 0 rounds down: exact for a
// compensation already on the grid, as in uploads)
__host__ __device__ inline CompCode comp_encode(float comp, float c, uint32_t h) {
  float s = ldexpf(comp, kCompShift + 24 - comp_exp(c));
  s = fminf(fmaxf(s, float(kCompMin) * 16777216.f), float(kCompMax) * 16777216.f);
  const int X = int(rintf(s));
  return CompCode((X + int(h >> 8)) >> 24);  // floor((X + U) / 2^24), never beyond [-8, 7]
}
// bytes of a mixed concentration buffer of n cells: n floats, then the codes, padded to 16 B
__host__ __device__ inline long mixed_conc_bytes(long n) {
  return 4 * n + ((n + 1) / 2 + 15) / 16 * 16;
}
// the code bytes of a mixed concentration buffer of n cells
__host__ __device__ inline const uint8_t *comp_bytes(const float *buf, long n) {
  return reinterpret_cast<const uint8_t *>(buf + n);
}
__host__ __device__ inline uint8_t *comp_bytes(float *buf, long n) {
  return reinterpret_cast<uint8_t *>(buf + n);
}
// q of cell i
__host__ __device__ inline CompCode comp_load(const float *buf, long n, long i) {
  return comp_unpack(uint32_t(comp_bytes(buf, n)[i >> 1]) >> (4 * int(i & 1)));
}
// host side: store q of cell i (no other writer of its byte at the same time)
inline void comp_put_host(float *buf, long n, long i, CompCode q) {
  uint8_t &b = comp_bytes(buf, n)[i >> 1];
  const int sh = 4 * int(i & 1);
  b = uint8_t((b & ~(15u << sh)) | ((uint32_t(q) & 15u) << sh));
}
// effective stored concentration of cell i of a buffer of n cells (mixed: value - compensation)
template <class F> __device__ inline double conc_effective(const typename F::C *buf, long i, long n);
template <> __device__ inline double conc_effective<FormF64>(const double *buf, long i, long) {
  return buf[i];
}
template <> __device__ inline double conc_effective<FormMixed>(const float *buf, long i, long n) {
  return double(buf[i]) - double(comp_decode(comp_load(buf, n, i), buf[i]));
}

// (x, y, z) of a storage-order cell index.  Cell indices stay below 2^31 (pf_create rejects larger
// domains), so 32-bit unsigned division replaces the much longer 64-bit division sequence.
__device__ inline void cell_coords(long cell, int nx, int ny, int &x, int &y, int &z) {
  const unsigned c = unsigned(cell), ux = unsigned(nx), uy = unsigned(ny);
  const unsigned yz = c / ux;
  const unsigned zz = yz / uy;
  x = int(c - yz * ux);
  y = int(yz - zz * uy);
  z = int(zz);
}

// periodic wrap of i+d for |d| <= n (Get_neighbor, phase_field.h:58-68, 0-based)
__device__ inline int wrap(int i, int n) {
  i = i < 0 ? i + n : i;
  return i >= n ? i - n : i;
}

// Step constants for the phase-field sweep (1-based phase tables; row/column 0 = empty slot).
template <class A> struct PFConsts {
  A M[4][4], sig[4][4], ent[4][4], eqc[4][4], slope[4][4];
  A pk[4][4], pc[4][4];  // carbon partition (local_conc of Driving_force)
  A pkS[4][4], pcS[4][4];  // partition of the second solute (re-partition after the sweep)
  A f1, f2, f3;
  double dt, cutoff;
};

// Step constants for one solute's diffusion sub-step (0-based phase tables).
template <class A> struct DiffConsts {
  A pk[3][3], pc[3][3];
  A D[3];
  A r;           // dt_d / (dx*dx)
  uint32_t key;  // mixed form: the sub-step's compensation rounding key (comp_step_key)
};

// Slot-row storage of the ids: KI 16-bit codes per cell.
template <int K> struct IdRowT;
template <> struct IdRowT<4> {
  static constexpr int KI = 4;
  using V = uint2;  // 8 bytes
};
template <> struct IdRowT<6> {
  static constexpr int KI = 8;
  using V = uint4;  // 16 bytes (2 padding codes)
};

template <int K> __device__ inline void load_ids(const uint16_t *base, long cell, uint32_t (&g)[K]) {
  using V = typename IdRowT<K>::V;
  const V v = reinterpret_cast<const V *>(base)[cell];
  const uint32_t *w = reinterpret_cast<const uint32_t *>(&v);
#pragma unroll
  for (int s = 0; s < K; ++s) g[s] = (w[s >> 1] >> ((s & 1) * 16)) & 0xffffu;
}
// the K slot values of a cell's row with vector loads (rows are 8-byte aligned for every K and
// type here; 16-byte pieces where the row size allows)
template <class T, int K> __device__ inline void load_row(const T *base, long cell, T (&v)[K]) {
  constexpr int B = int(sizeof(T)) * K;  // row bytes: 16, 24, 32 or 48
  if constexpr (B % 16 == 0) {
    using V = uint4;
    const V *p = reinterpret_cast<const V *>(base + cell * K);
#pragma unroll
    for (int q = 0; q < B / 16; ++q) {
      const V w = p[q];
      __builtin_memcpy(reinterpret_cast<char *>(v) + 16 * q, &w, 16);
    }
  } else {
    using V = uint2;
    const V *p = reinterpret_cast<const V *>(base + cell * K);
#pragma unroll
    for (int q = 0; q < B / 8; ++q) {
      const V w = p[q];
      __builtin_memcpy(reinterpret_cast<char *>(v) + 8 * q, &w, 8);
    }
  }
}
template <int K> __device__ inline void store_ids(uint16_t *base, long cell, const uint32_t (&g)[K]) {
  using V = typename IdRowT<K>::V;
  V v;
  uint32_t *w = reinterpret_cast<uint32_t *>(&v);
#pragma unroll
  for (int q = 0; q < int(sizeof(V) / 4); ++q) w[q] = 0;
#pragma unroll
  for (int s = 0; s < K; ++s) w[s >> 1] |= (g[s] & 0xffffu) << ((s & 1) * 16);
  reinterpret_cast<V *>(base)[cell] = v;
}

}  // namespace pfamd
