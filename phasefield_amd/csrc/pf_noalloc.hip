// pf_noalloc.hip — pf_gpu_diffusion_solver_noalloc: the diffusion-only offload of
// pf_gpu_diffusion_solver (GPU_Diffusion_solver, gpu_solver.cpp:204-218) for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  The caller passes value_phase and total_conc in device memory and a stream; the
// sub-steps run asynchronously on that stream with the LDS-tiled full-formula kernel
// (diffuse_kernel, bitwise equal to the bulk + listed path of the handle API).
//
// Jacobi sub-steps need a second concentration buffer.  It is `scratch`, a file-scope device array
// of kNoAllocMaxCells doubles (16 MiB), which every GPU the library is loaded on holds until the
// process ends.  Host code never passes the array's name to a kernel: its device address is looked
// up once per device with hipGetSymbolAddress (with that device current) and kept per device.  No
// part of the buffer needs clearing: the first sub-step writes every cell of it before any read.
// Calls on one GPU share the buffer, so calls whose work overlaps on one GPU (two streams, two
// threads) race; calls ordered on one stream do not.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/phasefield_amd.h"
#include "pf_handle_access.h"
#include "pf_kernels.h"

using namespace pfamd;

namespace {

constexpr long kNoAllocMaxCells = 1L << 21;  // 2,097,152 cells (e.g. 128^3)
__device__ __align__(256) double scratch[kNoAllocMaxCells];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
std::array<double *, kMaxDevices> g_scratch_ptr{};

// device address of `scratch` on the current device (looked up on the first call there)
hipError_t scratch_on(int dev, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

__global__ __launch_bounds__(256) void copy_conc_kernel(const double *__restrict__ src,
                                                        double *__restrict__ dst, long n) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

extern "C" int pf_gpu_diffusion_solver_noalloc(const double *hdiffusivity,
                                               const double *hpartition_k,
                                               const double *hpartition_c,
                                               const double *dvalue_phase, double *dtotal_conc,
                                               int x, int y, int z, double dt, double dx,
                                               void *stream) {
  if (!hdiffusivity || !hpartition_k || !hpartition_c || !dvalue_phase || !dtotal_conc)
    return set_global_error(PF_ERR_INVALID, "pf_gpu_diffusion_solver_noalloc: null pointer");
  if (x < 1 || y < 1 || z < 1 || (y == 1 && z == 1))
    return set_global_error(PF_ERR_INVALID,
                            "pf_gpu_diffusion_solver_noalloc: extents must be >= 1 and not 1D");
  const long N = long(x) * y * z;
  if (N > kNoAllocMaxCells)
    return set_global_error(
        PF_ERR_RANGE, "pf_gpu_diffusion_solver_noalloc: " + std::to_string(N) +
                          " cells exceed its scratch buffer of " +
                          std::to_string(kNoAllocMaxCells) + " cells; use pf_gpu_diffusion_solver");
  if (!(dx > 0) || !(dt >= 0))
    return set_global_error(PF_ERR_INVALID, "pf_gpu_diffusion_solver_noalloc: bad dt or dx");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return set_global_error(PF_ERR_NODEVICE, "pf_gpu_diffusion_solver_noalloc: no current device");
  double *buf = nullptr;
  if ((e = scratch_on(dev, &buf)) != hipSuccess)
    return set_global_error(PF_ERR_HIP, std::string("pf_gpu_diffusion_solver_noalloc: "
                                                    "hipGetSymbolAddress: ") +
                                            hipGetErrorString(e));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  // sub-step count and factor exactly as pf_gpu_diffusion_solver (gpu_solver.cpp:205-209)
  const int dims = z > 1 ? 3 : 2;
  const double max_diff = *std::max_element(hdiffusivity, hdiffusivity + 3);
  const double del_t = dx * dx / max_diff / dims * 0.2f;
  const int num_loop = int(std::ceil(dt / del_t));
  if (num_loop <= 0) return PF_OK;
  DiffConsts<double> dc;
  std::memset(&dc, 0, sizeof(dc));
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      dc.pk[i][j] = hpartition_k[i + 3 * j];  // column-major (i,j) at i+3j
      dc.pc[i][j] = hpartition_c[i + 3 * j];
    }
    dc.D[i] = hdiffusivity[i];
  }
  dc.r = dt / num_loop / (dx * dx);
  // tile geometry of the LDS kernel: 64 x kDiffTY columns, z split into chunks of >= 16 planes
  // until about 2048 workgroups (six faces always, as Diff3D sums them, gpu_solver.cpp:160-188)
  const int tiles_x = (x + 63) / 64, tiles_y = (y + kDiffTY - 1) / kDiffTY;
  const long tiles = long(tiles_x) * tiles_y;
  const long nzc = std::max(1L, std::min<long>(z, (2048 + tiles - 1) / tiles));
  int zc = int((z + nzc - 1) / nzc);
  zc = std::max(zc, std::min(z, 16));
  const long nblocks = tiles * ((z + zc - 1) / zc);
  if (nblocks > 0x7fffffffL)
    return set_global_error(PF_ERR_INVALID, "pf_gpu_diffusion_solver_noalloc: grid too large");
  const double *pf = dvalue_phase;
  double *a = dtotal_conc, *b = buf;
  for (int it = 0; it < num_loop; ++it) {
    hipLaunchKernelGGL((diffuse_kernel<FormF64, 3, 6, false, kDiffTY>), dim3(unsigned(nblocks)),
                       dim3(64, kDiffTY + 3), 0, s, a, b, pf, dc, x, y, z, zc, tiles_x, tiles_y,
                       int(nblocks));
    std::swap(a, b);
  }
  if (a != dtotal_conc)  // an odd count leaves the result in the scratch buffer
    hipLaunchKernelGGL(copy_conc_kernel, dim3(unsigned((N + 255) / 256)), dim3(256), 0, s, a,
                       dtotal_conc, N);
  if ((e = hipGetLastError()) != hipSuccess)
    return set_global_error(PF_ERR_HIP, std::string("pf_gpu_diffusion_solver_noalloc: launch: ") +
                                            hipGetErrorString(e));
  return PF_OK;
}
