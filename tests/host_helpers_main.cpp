// host_helpers_main.cpp — CPU checks of the device helpers that also compile for the host
// (phasefield_amd/csrc/pf_kernels.h, pf_device.h); tests/test_host_helpers.py builds it with hipcc
// (host code only) against the oracle's library and runs it.
//
//  1. list_order_cell (the sweep list's row-tile order): a bijection onto the cells, tiles in y
//     order, a tile's planes in z order, rows then x inside a plane; 16 consecutive positions
//     starting at a multiple of 16 are 16 consecutive cells whenever nx % 16 == 0.
//  2. The mixed form's 4-bit code storage: comp_put_host / comp_load round-trip every code in
//     [-8, 7] at even and odd cells and leave the other cell of the byte unchanged;
//     mixed_conc_bytes leaves room for every code and for the 32-bit words holding them.
//  3. comp_encode / comp_decode (the device's rounding of the Kahan compensation) against the
//     oracle's independent quantize_comp (orc_quantize_comp) on random and edge inputs.
//  4. id_row_breaks (the copy-cell rule on a neighbour's four ids as one 64-bit word) against the
//     slot-by-slot walk it replaces, on rows drawn from ids that include 0 and G.
//  5. zero_bytes (the diffusion flag pass's carry-free zero-byte test) against a byte loop.
// Exit status 0 when every check holds.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../phasefield_amd/csrc/pf_kernels.h"

extern "C" float orc_quantize_comp(float comp, float value, uint32_t key);

using namespace pfamd;

static int g_fail = 0;
#define CHECK(cond, ...)                        \
  do {                                          \
    if (!(cond)) {                              \
      std::fprintf(stderr, "check failed: " __VA_ARGS__); \
      std::fprintf(stderr, "\n");               \
      ++g_fail;                                 \
    }                                           \
  } while (0)

static void list_order() {
  const int shapes[][4] = {{512, 512, 4, 8}, {70, 13, 19, 5}, {64, 48, 40, 7}, {16, 16, 16, 16},
                           {32, 9, 3, 4},    {96, 17, 5, 1},  {128, 100, 6, 32}, {33, 5, 7, 2}};
  for (const auto &sh : shapes) {
    ListOrder o;
    o.nx = sh[0];
    o.ny = sh[1];
    o.nz = sh[2];
    o.rows = sh[3] < sh[1] ? sh[3] : sh[1];
    const long N = long(o.nx) * o.ny * o.nz;
    std::vector<char> seen(N, 0);
    std::vector<long> cell(N);
    for (long p = 0; p < N; ++p) {
      const long c = list_order_cell(o, p);
      cell[p] = c;
      CHECK(c >= 0 && c < N && !seen[c], "list order %dx%dx%d rows %d: position %ld -> %ld", o.nx,
            o.ny, o.nz, o.rows, p, c);
      if (c >= 0 && c < N) seen[c] = 1;
    }
    for (long p = 1; p < N; ++p) {
      const long a = cell[p - 1], b = cell[p];
      const long xa = a % o.nx, ya = (a / o.nx) % o.ny, za = a / (long(o.nx) * o.ny);
      const long xb = b % o.nx, yb = (b / o.nx) % o.ny, zb = b / (long(o.nx) * o.ny);
      const long ta = ya / o.rows, tb = yb / o.rows;
      const bool ok = ta < tb || (ta == tb && (za < zb || (za == zb && (ya < yb || (ya == yb && xa < xb)))));
      CHECK(ok, "list order %dx%dx%d rows %d: positions %ld, %ld out of order", o.nx, o.ny, o.nz,
            o.rows, p - 1, p);
    }
    if (o.nx % 16 == 0)
      for (long f = 0; f + 16 <= N; f += 16)
        for (int j = 1; j < 16; ++j)
          CHECK(cell[f + j] == cell[f] + j, "list order %dx%dx%d: chunk at %ld not contiguous",
                o.nx, o.ny, o.nz, f);
    ListOrder plain;  // rows == 0: storage order
    for (long p = 0; p < N; p += 97) CHECK(list_order_cell(plain, p) == p, "storage order");
  }
}

static void code_storage() {
  std::mt19937 rng(7);
  for (long n : {1L, 2L, 3L, 17L, 1000L, 4097L}) {
    // the code region after the n values: a multiple of 16 bytes, every code, and the 32-bit
    // words the one-cell-per-thread kernels change with atomics (word i / 8 for cell i)
    const long bytes = mixed_conc_bytes(n), code_bytes = bytes - 4 * n;
    CHECK(code_bytes % 16 == 0 && code_bytes >= (n + 1) / 2 && code_bytes >= 4 * ((n + 7) / 8),
          "mixed_conc_bytes(%ld) = %ld", n, bytes);
    std::vector<float> buf(bytes / 4, 0.f);
    std::vector<int> want(n);
    for (long i = 0; i < n; ++i) {
      want[i] = int(rng() % 16) - 8;
      comp_put_host(buf.data(), n, i, CompCode(want[i]));
    }
    for (long i = 0; i < n; ++i)
      CHECK(comp_load(buf.data(), n, i) == want[i], "code %ld of %ld: %d != %d", i, n,
            int(comp_load(buf.data(), n, i)), want[i]);
    // rewriting one cell keeps the other cell of its byte
    for (long i = 0; i < n; ++i) {
      const long j = i ^ 1;
      comp_put_host(buf.data(), n, i, CompCode(-8 + (want[i] + 9) % 16));
      if (j < n) CHECK(comp_load(buf.data(), n, j) == want[j], "neighbour code %ld changed", j);
      comp_put_host(buf.data(), n, i, CompCode(want[i]));
    }
    for (long i = 0; i < n; ++i) CHECK(buf[i] == 0.f, "value %ld of %ld written by a code store", i, n);
  }
}

static void code_rounding() {
  std::mt19937 rng(11);
  std::uniform_real_distribution<float> cval(1e-3f, 1.5f), frac(-9.f, 9.f);
  long checked = 0;
  auto one = [&](float comp, float c, uint32_t h) {
    const float dev = comp_decode(comp_encode(comp, c, h), c);
    const float orc = orc_quantize_comp(comp, c, h);
    CHECK(dev == orc, "comp %.9g value %.9g key %08x: device %.9g oracle %.9g", comp, c, h, dev, orc);
    ++checked;
  };
  for (int k = 0; k < 200000; ++k) {
    const float c = cval(rng);
    int e = 0;
    (void)std::frexp(c, &e);
    const float quantum = std::ldexp(1.f, e - 27);
    one(frac(rng) * quantum, c, uint32_t(rng()));
  }
  // edge cases: on the grid, saturation, zero compensation, thresholds 0 and 2^32 - 1, tiny values
  for (float c : {0.02f, 0.5f, 0.75f, 1.f, 1.7f, 3e-3f}) {
    int e = 0;
    (void)std::frexp(c, &e);
    const float quantum = std::ldexp(1.f, e - 27);
    for (float m : {0.f, -0.f, 1.f, -1.f, 7.f, -8.f, 7.5f, -8.5f, 20.f, -20.f, 1e-6f, -1e-6f, 0.5f,
                    -0.5f})
      for (uint32_t h : {0u, 0xffffffffu, 0x80000000u, 0x12345678u}) one(m * quantum, c, h);
  }
  std::printf("code rounding: %ld cases equal to the oracle\n", checked);
}

// the copy-cell rule of one neighbour row as a slot walk (the K = 6 path of the classification and
// the reference's loop): ids up to the first 0 are live; with a grain already known every live id
// must equal it, otherwise the first live id becomes the grain
static bool walk_breaks(const uint32_t (&id)[4], uint32_t &G, bool &any) {
  bool bad = false;
  for (int s = 0; s < 4 && id[s] != 0; ++s) {
    if (!any) {
      G = id[s];
      any = true;
    } else if (id[s] != G) {
      bad = true;
    }
  }
  return bad;
}

static void id_rows() {
  std::mt19937 rng(5);
  long checked = 0;
  const uint32_t pool[] = {0u, 1u, 2u, 7u, 0xffffu, 0x8000u};
  for (int k = 0; k < 400000; ++k) {
    uint32_t id[4];
    for (int s = 0; s < 4; ++s) id[s] = pool[rng() % 6];
    const uint64_t w = uint64_t(id[0]) | uint64_t(id[1]) << 16 | uint64_t(id[2]) << 32 |
                       uint64_t(id[3]) << 48;
    for (uint32_t G : {0u, 1u, 2u, 7u, 0xffffu}) {
      // a known grain G (own entry, or an earlier neighbour; G = 0 only as an own slot's id)
      uint32_t g = G;
      bool any = true;
      CHECK(walk_breaks(id, g, any) == id_row_breaks(w, G), "row %016llx G %u",
            (unsigned long long)w, G);
      ++checked;
    }
    // no grain known yet: the walk takes slot 0 (when not 0) as the grain, as the kernel does
    uint32_t g = 0;
    bool any = false;
    const bool walk = walk_breaks(id, g, any);
    const uint32_t G = id[0];
    CHECK(walk == (G != 0u && id_row_breaks(w, G)), "row %016llx first grain", (unsigned long long)w);
    CHECK(any == (G != 0u), "row %016llx live", (unsigned long long)w);
    ++checked;
  }
  std::printf("id rows: %ld cases equal to the slot walk\n", checked);
}

// zero_bytes (the flag pass's byte-wise equality test) against a byte loop
static void zero_byte_words() {
  std::mt19937 rng(9);
  long checked = 0;
  auto one = [&](uint32_t x) {
    uint32_t want = 0;
    for (int b = 0; b < 4; ++b)
      if (((x >> (8 * b)) & 0xffu) == 0u) want |= 0x80u << (8 * b);
    CHECK(zero_bytes(x) == want, "zero_bytes(%08x) = %08x, want %08x", x, zero_bytes(x), want);
    ++checked;
  };
  // every word whose bytes are drawn from {0, 1, 0x7f, 0x80, 0xff, random}
  const uint32_t pick[] = {0u, 1u, 0x7fu, 0x80u, 0xffu};
  for (int k = 0; k < 6 * 6 * 6 * 6; ++k) {
    uint32_t x = 0;
    int r = k;
    for (int b = 0; b < 4; ++b, r /= 6) x |= (r % 6 < 5 ? pick[r % 6] : (rng() & 0xffu)) << (8 * b);
    one(x);
  }
  for (int k = 0; k < 200000; ++k) one(uint32_t(rng()) & uint32_t(rng() | rng()));
  std::printf("zero bytes: %ld words equal to the byte loop\n", checked);
}

int main() {
  list_order();
  id_rows();
  zero_byte_words();
  code_storage();
  code_rounding();
  std::printf("host_helpers_main: %s\n", g_fail ? "FAILED" : "ok");
  return g_fail ? 1 : 0;
}
