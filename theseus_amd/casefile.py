"""Running a reference case file end to end on the engine: the JSON
config's ``runTime`` keys (src/Simulation.cpp:105-313, through
``RunTime.from_json``), its mesh file (MFEM v1.0 or Gmsh 2.2, ``meshio``),
``ser_ref_levels`` uniform refinement (:343-353), ``mesh_ordering: Hilbert``
(:363-374) and the initial condition named in
``conditions.initial_conditions`` with positional parameters x1..xN
(include/StateInit.hpp; nodal interpolation at the GLL points, which is what
ProjectCoefficient does for a DG GLL space, src/Simulation.cpp:1076-1109).

Initial conditions restated: IsentropicVortexIC, TaylorGreenVortexIC,
TaylorGreenVortex2DIC, CompressibleTaylorGreenVortexIC,
KelvinHelmholtzInstabilityIC.  The Helmholtz LTE gas needs a table; the
reference builds it with PLATO (SimFactory_impl.hpp:140-216), which is not
available here, so a caller passes one (``lte_table``) or the synthetic
dissociating gas is tabulated on the config's grid (N_rho, N_T, rho_min,
rho_max, T_min, T_max, rho_dist, T_dist).
"""
from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import cases
from .config import RunTime
from .meshio import AffineMesh, hilbert_order, read_mesh, reorder, uniform_refine

# name -> (number of parameters, function(mesh, order, *params))
INITIAL_CONDITIONS = {
    "IsentropicVortexIC": (5, cases.reference_isentropic_vortex),
    "TaylorGreenVortexIC": (2, cases.taylor_green),
    "TaylorGreenVortex2DIC": (2, cases.taylor_green_2d),
    "CompressibleTaylorGreenVortexIC": (1, cases.compressible_taylor_green),
    "KelvinHelmholtzInstabilityIC": (1, cases.kelvin_helmholtz),
}


@dataclass
class CaseFile:
    runtime: RunTime
    mesh: AffineMesh
    U0: np.ndarray
    config: dict
    final_time: float
    lte_table: object = None


def resolve_mesh_path(config_path: Path, mesh_file: str) -> Path:
    """The reference's mesh paths are relative to its build directory
    ("../../TestCases/..."); try the path as given, relative to the config's
    directory, then the file name next to the config."""
    cfg_dir = Path(config_path).resolve().parent
    for cand in (Path(mesh_file), cfg_dir / mesh_file, cfg_dir / Path(mesh_file).name):
        if cand.exists():
            return cand
    raise FileNotFoundError(f"mesh_file {mesh_file!r} not found next to {config_path}")


def initial_condition(name: str, params: dict, mesh, order: int) -> np.ndarray:
    if name not in INITIAL_CONDITIONS:
        raise NotImplementedError(f"initial condition {name!r} is not restated (have: {sorted(INITIAL_CONDITIONS)})")
    n, fn = INITIAL_CONDITIONS[name]
    args = [float(params[f"x{i + 1}"]) for i in range(n)]
    return fn(mesh, order, *args)


def load_case(config_path, mesh_path=None, lte_table=None, bc_types=2) -> CaseFile:
    config_path = Path(config_path)
    cfg = json.loads(config_path.read_text())
    rtj = cfg["runTime"]
    rt = RunTime.from_json(rtj, cfg.get("compileTime"))
    mpath = Path(mesh_path) if mesh_path else resolve_mesh_path(config_path, rtj["mesh_file"])
    mesh = read_mesh(mpath, bc_types)
    if mesh.dim != rt.dim:
        raise ValueError(f"mesh dimension {mesh.dim} != runTime dim {rt.dim}")
    levels = int(rtj.get("ser_ref_levels", 0))
    if levels:
        mesh = uniform_refine(mesh, levels)
    ordering = rtj.get("mesh_ordering")
    if ordering == "Hilbert":
        mesh = reorder(mesh, hilbert_order(mesh))
    elif ordering not in (None, ""):
        raise NotImplementedError(f"mesh_ordering {ordering!r} (Hilbert only)")
    ic = cfg["runTime"]["conditions"]["initial_conditions"]
    U0 = initial_condition(ic["function"], ic.get("params", {}), mesh, rt.order)
    table = lte_table
    if table is None and rt.volume_flux == cases.A.VOLUME_LTEEC:
        from .lte import synthetic_table

        table = synthetic_table(int(rtj.get("N_rho", 101)), int(rtj.get("N_T", 101)),
                                (float(rtj.get("rho_min", 0.01)), float(rtj.get("rho_max", 3.0))),
                                (float(rtj.get("T_min", 100.0)), float(rtj.get("T_max", 2000.0))),
                                rho_dist=rtj.get("rho_dist", "log"), T_dist=rtj.get("T_dist", "linear"))
    return CaseFile(rt, mesh, U0, cfg, float(rtj.get("final_time", 0.0)), table)
