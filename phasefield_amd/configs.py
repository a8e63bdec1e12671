"""Named workloads written as input files in the reference's positional dialect.

The reference reads a positional main file and a thermodynamics file
(C++-openmp-cuda/src/phase_field.cpp:31-231, thermo_functor.cpp:11-53; example pair:
C++-openmp-cuda/build/log_file.txt, build/thermo_data.txt).  Each workload below is written as such
a pair and read back with the library's own parser, so every run exercises the drop-in input path.

Physics shared by all workloads (Fe-C at T = 1013 K; units um, s, J):
  phase 1 = austenite (nuclei, ids 100000+n), phase 2 = ferrite (Voronoi grains, ids 200000+g),
  phase 3 = unused placeholder with finite partition coefficients.
  dx = 0.1 um, interface width 6 cells, cutoff 1e-5.
  sigma = 0.5 J/m^2 = 0.5e-12 J/um^2 and M = 1e-12 m^4/(J s) = 1e12 um^4/(J s) for all pairs of
  phases 1 and 2; zero for pairs with phase 3.
  D_alpha = 1e-10 m^2/s = 100 um^2/s, D_gamma = 5e-13 m^2/s = 0.5 um^2/s.
  Linearised phase diagram: straight lines through (T0 = 1185 K, 0 wt%) with c_gamma_eq = 0.45 and
  c_alpha_eq = 0.018 wt% at 1013 K, so k(1,2) = 0.018/0.45 = 0.04 and c(1,2) ~ 0.
  Entropy S(1,2) = -0.2e-12 J/(K um^3): with the negative slope of the gamma line this makes the
  driving force for austenite positive when its carbon exceeds c_gamma_eq.
  Manganese (two-solute workloads): 1.5 wt% + seeded noise (sd 0.05 wt%), k_gamma/alpha = 2,
  D_Mn,alpha = 1e-16 m^2/s = 1e-4 um^2/s, D_Mn,gamma = 1e-18 m^2/s = 1e-6 um^2/s, one explicit
  update per phase-field step (D_Mn * pfm_dt / dx^2 ~ 1e-5, far inside the stability limit).
"""
from __future__ import annotations

import os

T_K = 1013.0
T0_K = 1185.0
C_GAMMA_EQ = 0.45
C_ALPHA_EQ = 0.018

WORKLOADS = {
    # 2D 256x256, 12 grains + 20 nuclei, 1000 steps: the reference's own CPU-runnable case
    "fec_2d_256": dict(nx=256, ny=256, nz=1, grains=12, nuclei=20, steps=1000, depth=6,
                       precision="fp64", solutes=1, bc_z="periodic"),
    # 3D 256^3 Fe-C, 128 grains + 200 nuclei, K=4, FP64 parity case
    "fec_3d_256": dict(nx=256, ny=256, nz=256, grains=128, nuclei=200, steps=2000, depth=4,
                       precision="fp64", solutes=1, bc_z="periodic"),
    # 3D 512^3 Fe-C-Mn, 512 grains + 1000 nuclei, K=4, mixed vs FP64
    "fecmn_3d_512": dict(nx=512, ny=512, nz=512, grains=512, nuclei=1000, steps=1000, depth=4,
                         precision="mixed", solutes=2, bc_z="periodic"),
    # 3D 128^3 Fe-C, 32 grains + 40 nuclei, 400 steps: discrete-adjoint case
    "fec_3d_128_adjoint": dict(nx=128, ny=128, nz=128, grains=32, nuclei=40, steps=400, depth=4,
                               precision="fp64", solutes=1, bc_z="periodic"),
    # 3D 1024^3 Fe-C-Mn, 4096 grains + 8000 nuclei, no-flux z, K=4 mixed
    "fecmn_3d_1024": dict(nx=1024, ny=1024, nz=1024, grains=4096, nuclei=8000, steps=200,
                          depth=4, precision="mixed", solutes=2, bc_z="noflux"),
}


def thermo_text() -> str:
    """Thermodynamics file: 3 entropy lines, 6 NPLE1, 6 NPLE2 and 6 PARA lines (a T^2 + b T + c)."""
    lines = {}
    for (i, j), ceq in (((1, 2), C_GAMMA_EQ), ((2, 1), C_ALPHA_EQ)):
        b = ceq / (T_K - T0_K)
        lines[(i, j)] = (0.0, b, -b * T0_K)
    # phase-3 placeholder: finite, nonzero slopes so every partition coefficient is finite
    lines[(1, 3)] = (0.0, -1e-3, 1.5)
    lines[(3, 1)] = (0.0, -1e-3, 1.2)
    lines[(2, 3)] = (0.0, -1e-3, 1.4)
    lines[(3, 2)] = (0.0, -1e-3, 1.1)
    order = [(1, 2), (2, 1), (1, 3), (3, 1), (2, 3), (3, 2)]
    para = "\n".join(f"{i} {j} {a!r} {b!r} {c!r}" for (i, j) in order for (a, b, c) in [lines[(i, j)]])
    return (
        "#three phases: austenite (1), ferrite (2), placeholder (3)\n"
        "#driving force G = k*(c - c_eq)/slope, k = a*T + b (entropy)\n"
        "#input of k: phase_i phase_j a b\n"
        "1 2 0. -0.2e-12\n"
        "1 3 0. 0.\n"
        "2 3 0. 0.\n"
        "#phase diagram [NPLE1, NPLE2, PARA]\n"
        "#for each line a quadratic fit wt% = a*T^2 + b*T + c: phase_i phase_j a b c\n"
        "#NPLE1 (not used by the step)\n"
        f"{para}\n"
        "#NPLE2 (not used by the step)\n"
        f"{para}\n"
        "#PARA wt%\n"
        f"{para}\n"
    )


def main_text(wl: dict, outdir: str, thermo_name: str) -> str:
    """Main input file in the reference's positional dialect plus `key=value` extensions."""
    g = wl["grains"]
    ext = {
        "steps": wl["steps"],
        "depth": wl["depth"],
        "nphase": 2,
        "precision": wl["precision"],
        "bc_z": wl["bc_z"],
        "solutes": wl["solutes"],
        "nucleus_radius": wl.get("nucleus_radius", 4),
        "nucleus_conc": wl.get("nucleus_conc", 0.8),
        "seed": wl.get("seed", g),
        "nucleation_seed": wl.get("nucleation_seed", 20251),
    }
    if wl["solutes"] > 1:
        ext.update({"mn_conc": 1.5, "mn_noise": wl.get("mn_noise", 0.05),
                    "mn_seed": wl.get("mn_seed", 777), "mn_k": 2.0,
                    "mn_D0_1": 1e-6, "mn_D0_2": 1e-4, "mn_D0_3": 0.0, "mn_substeps": 1})
    ext.update(wl.get("extra", {}))
    ext_lines = "\n".join(f"{k}={v}" for k, v in ext.items())
    # optional overrides: an Arrhenius mobility / diffusivities (prefactor, Q in J/mol) and a
    # piecewise-linear temperature path [(time, T), ...]
    def pair(v):
        return f"{float(v[0])!r} {float(v[1])!r}"

    M = pair(wl["mobility"]) if "mobility" in wl else "1e12 0"
    sig = "0.5e-12"
    d_gam = (pair(wl["diff_gamma"]) if "diff_gamma" in wl
             else f"{wl.get('D_gamma', 0.5)} 0")
    d_alp = (pair(wl["diff_alpha"]) if "diff_alpha" in wl
             else f"{wl.get('D_alpha', 100.0)} 0")
    t_lines = (f"2\n0 {T_K}\n1e6 {T_K}" if "temperature" not in wl else
               "\n".join([str(len(wl["temperature"]))] +
                         [pair(tT) for tT in wl["temperature"]]))
    return f"""OUTPUT NAME?
{outdir}
GRID SIZE?
 (FOR 2D CALCULATIONS:IZ=1;FOR 1D CALCULATIONS:IY=1,IZ=1)
 [IX IY IZ]?
        {wl['nx']}       {wl['ny']}       {wl['nz']}
 CELL DIMENSION (MICRON):
 {wl.get('dx', 0.1)}
 INTERFACE THICKNESS (IN CELLS):
{wl.get('i_intf', 6)}
 PHASE MINIMUM:
{wl.get('cutoff', 1e-5)}
 NUMBER OF GRAINS (phase 1, phase 2, phase 3):
     0 {g} 0
VOLUME FRACTION OF SECOND PHASE (<1):
0. 1. 0.
solute concentration in each phase (austenite, ferrite, placeholder)?
0.8 0.02 0.
 temperature [K]?
{t_lines}
 rolling reduction(<1):
{wl.get('reduction', 0.)}
 Number of iterations for initialisation:
         {wl.get('iso', 10)}
 the time step to write the result into file:
{wl.get('output_dt', 0.5)}
Microstructure output?
{wl.get('micro_out', 0)}
austenite nuclei number in ferrite?
 {wl['nuclei']}
 austenite-austenite interface mobility: prefactor mu0[um^4/J.s] and activation energy Q[J]?
 {M}
 austenite-ferrite interface mobility: prefactor mu0[um^4/J.s] and activation energy Q[J]?
 {M}
 austenite-placeholder interface mobility?
 0 0
 ferrite-ferrite interface mobility?
 {M}
 ferrite-placeholder interface mobility?
 0 0
 placeholder-placeholder interface mobility?
 0 0
 austenite-austenite interface energy[J/um^2]?
 {sig}
 austenite-ferrite interface energy[J/um^2]?
 {sig}
 austenite-placeholder interface energy[J/um^2]?
 0
 ferrite-ferrite interface energy[J/um^2]?
 {sig}
 ferrite-placeholder interface energy[J/um^2]?
 0
 placeholder-placeholder interface energy[J/um^2]?
 0
 austenite diffusivity D0 [um^2/s], Activation energy Q[J]?
 {d_gam}
 ferrite diffusivity D0,Activation energy Q?
 {d_alp}
 placeholder diffusivity D0,Activation energy Q?
 0 0
thermodynamic file?
{thermo_name}
Use GPU_Computing if possible?
1
{ext_lines}
"""


def write_inputs(directory: str, workload: str | dict, **overrides) -> str:
    """Write <directory>/input.txt and thermo.txt for a named workload; returns the input path.

    Overrides replace workload fields (nx, ny, nz, grains, nuclei, steps, depth, precision, ...).
    """
    wl = dict(WORKLOADS[workload]) if isinstance(workload, str) else dict(workload)
    wl.update(overrides)
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "thermo.txt"), "w") as f:
        f.write(thermo_text())
    path = os.path.join(directory, "input.txt")
    with open(path, "w") as f:
        f.write(main_text(wl, os.path.join(directory, "out"), "thermo.txt"))
    return path
