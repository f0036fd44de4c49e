"""Running a reference case file end to end on the engine: the JSON
config's ``runTime`` keys (src/Simulation.cpp:105-313, through
``RunTime.from_json``), its mesh file (MFEM v1.0 or Gmsh 2.2, ``meshio``),
``ser_ref_levels`` uniform refinement (:343-353), ``mesh_ordering: Hilbert``
(:363-374) and the initial condition named in
``conditions.initial_conditions`` with positional parameters x1..xN
(include/StateInit.hpp; nodal interpolation at the GLL points, which is what
ProjectCoefficient does for a DG GLL space, src/Simulation.cpp:1076-1109).

Initial conditions restated: IsentropicVortexIC, LTEVortexIC, TaylorGreenVortexIC,
TaylorGreenVortex2DIC, CompressibleTaylorGreenVortexIC,
KelvinHelmholtzInstabilityIC, LidDrivenCavityIC, ForwardFacingStepIC.  Boundaries:
``conditions.boundary_conditions`` by Gmsh physical name, with the
reference's boundary types and its named constant boundary data.  The Helmholtz LTE gas needs a table; the
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
from .meshio import AffineMesh, gmsh_physical_names, hilbert_order, read_mesh, reorder, uniform_refine

# name -> (number of parameters, function(mesh, order, *params))
INITIAL_CONDITIONS = {
    "LidDrivenCavityIC": (2, cases.lid_driven_cavity),
    "IsentropicVortexIC": (5, cases.reference_isentropic_vortex),
    "LTEVortexIC": (5, cases.lte_vortex),
    "TaylorGreenVortexIC": (2, cases.taylor_green),
    "TaylorGreenVortex2DIC": (2, cases.taylor_green_2d),
    "CompressibleTaylorGreenVortexIC": (1, cases.compressible_taylor_green),
    "KelvinHelmholtzInstabilityIC": (1, cases.kelvin_helmholtz),
    "ForwardFacingStepIC": (1, cases.forward_facing_step),
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


# The reference's named constant boundary data (StateInit.hpp, registered
# with Prandtl::ConditionFactory): conservative states for inflow, wall
# velocities and wall heat fluxes / temperatures.
NAMED_VECTORS = {
    "LidDrivenCavityWallVelBCVector": (0.0, 0.0),                                     # :296
    "LidDrivenCavityLidVelBCVector": (1.0, 0.0),                                      # :309
    "BackwardFacingStepLeftBCVector": (5.9970, 5.9970 * 98.5914, 0.0,
                                       11666.5 / 0.4 + 0.5 * 5.9970 * 98.5914 * 98.5914),  # :477-478
    "DoubleMachReflectionLeftBottom1BCVector": (8.0, 8.0 * 7.144709581221619, -8.0 * 4.125,
                                                116.5 / 0.4 + 0.5 * 8.0 * (7.144709581221619 ** 2 + 4.125 ** 2)),  # :544-545
    "RampInletBCVector": (1.0, 2.0, 0.0, (1.0 / 1.4) / 0.4 + 0.5 * 1.0 * 2.0 * 2.0),  # :733
    "ViscousSphereFreestreamBCVector": (1.0, 0.3, 0.0, (1.0 / 1.4) / 0.4 + 0.5 * 0.3 * 0.3),  # :788-789
    "ForwardFacingStepLeftBCVector": (1.4, 1.4 * 3.0, 0.0, 1.0 / 0.4 + 0.5 * 1.4 * 3.0 * 3.0),  # :819
}
NAMED_SCALARS = {
    "LidDrivenCavityAdiaBCScalar": 0.0,    # :280
    "LidDrivenCavityIsoBCScalar": 0.007,   # :283
}


def _named(table, key, what):
    if key not in table:
        raise NotImplementedError(f"named {what} {key!r} is not restated (have: {sorted(table)})")
    return table[key]


def boundary_spec(spec: dict, dim: int, gamma: float, R_gas: float):
    """One entry of conditions.boundary_conditions as the reference's
    configuration reads it (src/Simulation.cpp:555-900): returns the engine's
    boundary type, or (type, payload).  Forms the reference itself rejects
    (function-valued or time-dependent data, specified-state) raise
    NotImplementedError, as do the axisymmetric axis and radial profiles."""
    from . import _abi as A

    t = spec["type"]
    if t == "slip":
        return A.BC_SLIP_WALL
    if t == "symmetry":
        return A.BC_SYMMETRY
    if t == "supersonic-outflow":
        return A.BC_SUPERSONIC_OUTFLOW
    if t == "supersonic-inflow":
        if "vector" not in spec:
            raise NotImplementedError("supersonic-inflow with a function or time-dependent state")
        return A.BC_SUPERSONIC_INFLOW, tuple(_named(NAMED_VECTORS, spec["vector"], "state vector"))
    if t == "no-slip-adiabatic":
        if "vector" not in spec.get("velocity", {}):
            raise NotImplementedError("no-slip-adiabatic with function-valued wall data")
        vel = _named(NAMED_VECTORS, spec["velocity"]["vector"], "wall velocity")
        heat = _named(NAMED_SCALARS, spec["heat"]["scalar"], "wall heat flux")
        if len(vel) != dim:
            # the payload is the velocity followed by the heat flux at index dim
            raise ValueError("Wall velocity dimension mismatch")
        return A.BC_NO_SLIP_ADIABATIC, tuple(float(x) for x in vel) + (float(heat),)
    if t == "no-slip-isothermal":
        v, T = spec.get("velocity", {}), spec.get("temperature", {})
        vel = v["value"] if "value" in v else _named(NAMED_VECTORS, v["vector"], "wall velocity")
        temp = T["value"] if "value" in T else _named(NAMED_SCALARS, T["scalar"], "wall temperature")
        if len(vel) != dim:
            raise ValueError("Wall velocity dimension mismatch")
        if not (np.isfinite(temp) and temp > 0):
            raise ValueError("Wall temperature must be positive")
        return A.BC_NO_SLIP_ISOTHERMAL, tuple(float(x) for x in vel) + (float(temp),)
    if t == "exterior-state":
        return A.BC_PRESCRIBED_STATE, exterior_state(spec, dim, gamma, R_gas)
    if t == "cpg-exterior-state":
        # Simulation.cpp:777-801: 2-D calorically perfect gas only; CPGState
        # (RadialProfile.hpp:108-116) checks p, T > 0
        if dim != 2:
            raise ValueError("CPG exterior BC requires 2D CPG; LTE is not supported")
        p, T = float(spec["pressure"]), float(spec["temperature"])
        ux, ur = float(spec.get("ux", 0.0)), float(spec.get("ur", 0.0))
        if not (p > 0 and T > 0) or not np.isfinite(p + T + ux + ur):
            raise ValueError("Invalid CPG primitive state")
        rho = p / (R_gas * T)
        return A.BC_PRESCRIBED_STATE, (rho, rho * ux, rho * ur, p / (gamma - 1.0) + 0.5 * rho * (ux * ux + ur * ur))
    raise NotImplementedError(f"boundary type {t!r} is not supported")


def exterior_state(spec: dict, dim: int, gamma: float, R_gas: float):
    """The conservative state of an ``exterior-state`` boundary as the
    reference parses and checks it (BoundaryStateConfig.hpp:18-45,
    PhysicalStateConfig.hpp:13-50, PhysicalState.hpp:33-57): only the keys
    type and state; state has exactly two of density, pressure and
    temperature plus a velocity array of the case's dimension, all finite,
    the thermodynamic values positive; ideal-gas conversion p = rho R T."""
    extra = set(spec) - {"type", "state"}
    if extra:
        raise ValueError(f"Unsupported boundary field '{sorted(extra)[0]}'")
    st = spec.get("state")

    def bad(msg):
        return ValueError(f"state: {msg}")

    if not isinstance(st, dict):
        raise bad("must be an object")
    for key in st:
        if key not in ("density", "pressure", "temperature", "velocity"):
            raise bad(f"unsupported field '{key}'")
    have = [k for k in ("density", "pressure", "temperature") if k in st]
    if len(have) != 2:
        raise bad("requires exactly two of density, pressure, temperature")
    vals = {}
    for k in have:
        if not isinstance(st[k], (int, float)) or isinstance(st[k], bool):
            raise bad(f"{k} must be a number")
        vals[k] = float(st[k])
    if "velocity" not in st:
        raise bad("velocity must be an array")
    vel = st["velocity"]
    if not isinstance(vel, (list, tuple)):
        raise bad("velocity must be an array")
    if any(not isinstance(u, (int, float)) or isinstance(u, bool) for u in vel):
        raise bad("velocity entries must be numbers")
    vel = [float(u) for u in vel]
    if len(vel) != dim:
        raise bad("Physical state velocity dimension does not match gas layout")
    if not all(np.isfinite(vel)):
        raise bad("Velocity must be finite")
    for k in have:
        if not (np.isfinite(vals[k]) and vals[k] > 0):
            raise bad(f"{k.capitalize()} must be finite and positive")
    if "density" not in vals:
        rho = vals["pressure"] / (R_gas * vals["temperature"])
        p = vals["pressure"]
    elif "pressure" not in vals:
        rho = vals["density"]
        p = rho * R_gas * vals["temperature"]
    else:
        rho, p = vals["density"], vals["pressure"]
    E = p / (gamma - 1.0) + 0.5 * rho * sum(x * x for x in vel)
    return (rho, *[rho * x for x in vel], E)


def boundary_types(cfg: dict, mesh_path: Path, dim: int, gamma: float, R_gas: float):
    """attribute -> boundary spec for read_mesh, from the config's named
    boundaries and the mesh's Gmsh physical names."""
    bcs = cfg["runTime"].get("conditions", {}).get("boundary_conditions")
    if not bcs:
        return 2
    names = gmsh_physical_names(mesh_path) if Path(mesh_path).read_text()[:11] == "$MeshFormat" else {}
    if not names:
        raise NotImplementedError("named boundaries need a Gmsh mesh with $PhysicalNames")
    out = {}
    for name, spec in bcs.items():
        if name not in names:
            continue  # the reference skips names without faces (Simulation.cpp:529-533)
        out[names[name]] = boundary_spec(spec, dim, gamma, R_gas)
    return out


def initial_condition(name: str, params: dict, mesh, order: int) -> np.ndarray:
    if name not in INITIAL_CONDITIONS:
        raise NotImplementedError(f"initial condition {name!r} is not restated (have: {sorted(INITIAL_CONDITIONS)})")
    n, fn = INITIAL_CONDITIONS[name]
    args = [float(params[f"x{i + 1}"]) for i in range(n)]
    return fn(mesh, order, *args)


def load_case(config_path, mesh_path=None, lte_table=None, bc_types=None) -> CaseFile:
    """bc_types None: from the config's boundary_conditions (supersonic outflow
    where it has none); else as read_mesh takes it."""
    config_path = Path(config_path)
    cfg = json.loads(config_path.read_text())
    mpath = Path(mesh_path) if mesh_path else resolve_mesh_path(config_path, cfg["runTime"]["mesh_file"])
    return load_case_config(cfg, mpath, lte_table, bc_types)


def load_case_config(cfg: dict, mesh_path, lte_table=None, bc_types=None) -> CaseFile:
    """load_case for a configuration already parsed ({"compileTime": ...,
    "runTime": ...}) and an explicit mesh file."""
    rtj = cfg["runTime"]
    rt = RunTime.from_json(rtj, cfg.get("compileTime"))
    mpath = Path(mesh_path)
    if bc_types is None:
        bc_types = boundary_types(cfg, mpath, rt.dim, rt.gamma, rt.R_gas)
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
    if table is None and (rt.lte or rt.volume_flux == cases.A.VOLUME_LTEEC):
        from .lte import synthetic_table

        table = synthetic_table(int(rtj.get("N_rho", 101)), int(rtj.get("N_T", 101)),
                                (float(rtj.get("rho_min", 0.01)), float(rtj.get("rho_max", 3.0))),
                                (float(rtj.get("T_min", 100.0)), float(rtj.get("T_max", 2000.0))),
                                rho_dist=rtj.get("rho_dist", "log"), T_dist=rtj.get("T_dist", "linear"))
    return CaseFile(rt, mesh, U0, cfg, float(rtj.get("final_time", 0.0)), table)
