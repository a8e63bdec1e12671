"""Test helpers: small workloads, oracle-built initial states and random active-set states."""
from __future__ import annotations

import os
import tempfile

import numpy as np

import oracle
from phasefield_amd import Model, configs
from phasefield_amd._capi import PfParams


def small_model(tmpdir: str, workload: str = "fec_3d_256", **over) -> Model:
    """Parse a (shrunken) named workload with the library's own input parser."""
    path = configs.write_inputs(tmpdir, workload, **over)
    return Model(path)


def model_params(m: Model) -> PfParams:
    """Parameters after Construct's Update_parameter (as the reference computes them)."""
    return m.update_parameter()


def oracle_initial_state(m: Model, p: PfParams, form=oracle.ORC_FP64) -> oracle.OracleDomain:
    """Seeded setup on the CPU oracle: Tessellation + Iso_growth + partition + nucleation
    (+ manganese noise for two-solute workloads)."""
    d = m.desc()
    od = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, d.nsolute, form, d.bc_z)
    od.set_params(p)
    seeds = m.voronoi_seeds()
    grains = oracle.tessellation(d.nx, d.ny, d.nz, seeds, m.get_real("reduction"))
    ic = [m.get_real(f"init_conc{i}") for i in range(3)]
    if d.nsolute > 1:
        ic += [m.get_real("mn_conc")] * 3
    od.construct(grains, m.get_int("num_init_iterations"), ic)
    if d.nsolute > 1 and m.get_real("mn_noise") > 0:
        od.solute_noise(1, m.get_real("mn_conc"), m.get_real("mn_noise"), m.get_int("mn_seed"))
    od.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
                (m.get_real("nucleus_conc"), m.get_real("mn_nucleus_conc")),
                m.get_int("nucleation_seed"))
    return od


def random_state(rng, nx, ny, nz, K, n_grains=6, nphase=2, nsolute=1, max_slots=None):
    """Random valid active-set state: per cell 1..max_slots distinct grains with phi sorted
    descending and summing to one; value_phase consistent; random concentrations."""
    N = nx * ny * nz
    max_slots = max_slots or K
    ids = np.zeros((N, K), np.int32)
    val = np.zeros((N, K))
    pool = [200001 + g for g in range(n_grains)] + [100001 + g for g in range(max(2, n_grains // 2))]
    if nphase == 3:
        pool += [300001, 300002]
    ncell = rng.integers(1, max_slots + 1, size=N)
    for n in range(N):
        k = ncell[n]
        g = rng.choice(pool, size=k, replace=False)
        v = rng.random(k) + 0.05
        v /= v.sum()
        order = np.argsort(-v, kind="stable")
        ids[n, :k] = g[order]
        val[n, :k] = v[order]
    phase = np.zeros((N, 3))
    for s in range(K):
        ph = np.where(ids[:, s] >= 300000, 2, np.where(ids[:, s] >= 200000, 1, 0))
        live = ids[:, s] != 0
        np.add.at(phase, (np.nonzero(live)[0], ph[live]), val[live, s])
    conc = np.concatenate([0.02 + 0.8 * rng.random(N) for _ in range(nsolute)])
    return ids.reshape(-1), val.reshape(-1), phase.reshape(-1), conc


def tmpdir():
    return tempfile.mkdtemp(prefix="pfamd_")


# ---- a reference-shaped C++ caller of the accelerator API ----------------------------------------
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIB_DIR = os.path.join(ROOT, "phasefield_amd")
LIB_SO = os.path.join(LIB_DIR, "libpfamd.so")

# The three prototypes as C++-openmp-cuda/src/gpu_solver.h:18-23 declares them (C++ linkage,
# real = double from real.h): the interface phase_field.cpp is compiled against.
REFERENCE_GPU_SOLVER_PROTOTYPES = """typedef double real;
extern int GPU_Init(const int x,const int y, const int z);
extern void GPU_Diffusion_solver(real*hdiffusivity, real*hpartition_k,real*hpartition_c,real*hvalue_phase,real*htotal_conc,int x,int y, int z,real  dt,real dx);
extern void GPU_Cleanup();
"""

_SHIM_MAIN = r"""
#include <cstdio>
#include <vector>
// in.bin: int x, y, z; double dt, dx, D[3], k[9], c[9], value_phase[3N], total_conc[N]
// out.bin: total_conc[N] after GPU_Diffusion_solver
int main(int argc, char **argv) {
  if (argc < 3) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int n[3];
  double sc[2], D[3], k[9], c[9];
  if (std::fread(n, sizeof(int), 3, f) != 3 || std::fread(sc, 8, 2, f) != 2 ||
      std::fread(D, 8, 3, f) != 3 || std::fread(k, 8, 9, f) != 9 || std::fread(c, 8, 9, f) != 9)
    return 2;
  const long N = long(n[0]) * n[1] * n[2];
  std::vector<double> vp(3 * N), tc(N);
  if (std::fread(vp.data(), 8, 3 * N, f) != size_t(3 * N) ||
      std::fread(tc.data(), 8, N, f) != size_t(N))
    return 2;
  std::fclose(f);
  const int ok = GPU_Init(n[0], n[1], n[2]);
  std::printf("GPU_Init=%d\n", ok);
  if (!ok) return 3;
  GPU_Diffusion_solver(D, k, c, vp.data(), tc.data(), n[0], n[1], n[2], sc[0], sc[1]);
  GPU_Cleanup();
  FILE *g = std::fopen(argv[2], "wb");
  if (!g || std::fwrite(tc.data(), 8, N, g) != size_t(N)) return 2;
  std::fclose(g);
  return 0;
}
"""


def build_reference_shim_program(tmpdir: str) -> str:
    """Compile and link (g++) a caller that sees only the reference's prototypes."""
    import subprocess

    src = os.path.join(tmpdir, "ref_caller.cpp")
    exe = os.path.join(tmpdir, "ref_caller")
    with open(src, "w") as f:
        f.write(REFERENCE_GPU_SOLVER_PROTOTYPES + _SHIM_MAIN)
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", exe, src, "-L", LIB_DIR, "-lpfamd",
                    f"-Wl,-rpath,{LIB_DIR}", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return exe


def write_shim_input(path, x, y, z, dt, dx, D, k, c, value_phase, total_conc):
    with open(path, "wb") as f:
        f.write(np.array([x, y, z], np.int32).tobytes())
        f.write(np.array([dt, dx], np.float64).tobytes())
        for a in (D, k, c, value_phase, total_conc):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
