"""JSON ``runTime`` case configuration -> engine descriptor.

Mirrors the reference factory's string handling (SimFactory_impl.hpp:63-128):
lower-cased ``numerical_flux`` / ``flow_model`` / ``gas_model`` with the same
aliases and precedence, and the PhysicsConstants defaults
(SimFactory_impl.hpp:133-137, Physics.hpp:40: gamma 1.4, Pr 0.72,
R_gas 287.05, mu 0.02, mu_bulk 2/3, and mu0 1.716e-5, the Sutherland
reference viscosity of the tabulated gas's SUTHERLAND build) plus
``alpha_max`` (src/Simulation.cpp:482-488).

Two keys have no counterpart in the reference and exist for pairings the
reference cannot select (one ``numerical_flux`` string fixes both fluxes
there): ``volume_flux`` and ``surface_flux`` override either half of the
pair.  ``blending`` (default true) mirrors the reference's compile-time
SUBCELL_FV_BLENDING switch (CMakeLists.txt:24).
"""
from __future__ import annotations

from dataclasses import dataclass

from . import _abi as A

# numerical_flux -> (volume flux, surface flux) (SimFactory_impl.hpp:253-334)
PAIRINGS = {
    "chandrashekar": (A.VOLUME_CHANDRASHEKAR, A.SURFACE_CHANDRASHEKAR),
    "ranocha": (A.VOLUME_RANOCHA, A.SURFACE_RANOCHA),
    "pirozzoli": (A.VOLUME_PIROZZOLI, A.SURFACE_PIROZZOLI),
    "hll": (A.VOLUME_CENTRAL, A.SURFACE_HLL),
    "llf": (A.VOLUME_CENTRAL, A.SURFACE_LLF),
    "roe": (A.VOLUME_CENTRAL, A.SURFACE_ROE),
    "lteec": (A.VOLUME_LTEEC, A.SURFACE_LTEEC),
}
# flux families the tabulated (Helmholtz) LTE gas is built with
LTE_FAMILIES = ("lteec", "pirozzoli", "hll", "llf")
VOLUME_NAMES = {
    "central": A.VOLUME_CENTRAL,
    "chandrashekar": A.VOLUME_CHANDRASHEKAR,
    "ranocha": A.VOLUME_RANOCHA,
    "pirozzoli": A.VOLUME_PIROZZOLI,
    "lteec": A.VOLUME_LTEEC,
}
SURFACE_NAMES = {
    "chandrashekar": A.SURFACE_CHANDRASHEKAR,
    "ranocha": A.SURFACE_RANOCHA,
    "pirozzoli": A.SURFACE_PIROZZOLI,
    "roe": A.SURFACE_ROE,
    "hll": A.SURFACE_HLL,
    "llf": A.SURFACE_LLF,
    "lteec": A.SURFACE_LTEEC,
}


def numerical_flux_name(s: str | None) -> str:
    """Canonical flux family for a ``numerical_flux`` string, with the
    reference's aliases and test order (SimFactory_impl.hpp:78-118,253-334)."""
    f = (s or "").lower()
    if f in ("lteec", "lteecflux"):
        return "lteec"
    if f == "" or f == "chandrashekar" or f.startswith("chan"):
        return "chandrashekar"
    if f == "ranocha":
        return "ranocha"
    if f in ("pirozzoli", "piflux") or f.startswith("pi"):
        return "pirozzoli"
    if f == "hll":
        return "hll"
    if f in ("llf", "lfr", "laxfriedrichs", "lax_friedrichs") or f.startswith("lax"):
        return "llf"
    if f == "roe":
        return "roe"
    raise ValueError(f"Invalid Numerical Flux Type specified: {s!r}")


def is_viscous(flow_model: str | None) -> bool:
    f = (flow_model or "").lower()
    return f.startswith(("visc", "nav", "cns", "ns"))


@dataclass
class RunTime:
    dim: int
    order: int
    flow_model: int
    volume_flux: int
    surface_flux: int
    gamma: float = 1.4
    Pr: float = 0.72
    R_gas: float = 287.05
    mu: float = 0.02
    mu_bulk: float = 2.0 / 3.0
    mu0: float = 1.716e-5
    alpha_max: float = 0.5
    alpha_min: float = 0.001
    blending: bool = True
    precision: int = A.PRECISION_FP64
    cfl: float = 1.0
    dt: float | None = None
    variable_dt: bool = True
    max_dt_growth: float = 1.2
    lte: bool = False  # the tabulated (Helmholtz) LTE gas: the operator needs its table

    @classmethod
    def from_json(cls, runtime: dict, compile_time: dict | None = None) -> "RunTime":
        """`runtime` is a case file's "runTime" object; `compile_time` its
        "compileTime" object (SUBCELL_FV_BLENDING sets the blending default)."""
        gas = (runtime.get("gas_model") or "").lower()
        fam = numerical_flux_name(runtime.get("numerical_flux"))
        # SimFactory_impl.hpp:75-95: the Helmholtz LTE gas is gas_model
        # "lte_helmholtz", or "lte" together with the LTEEC flux
        helmholtz = gas == "lte_helmholtz" or (gas == "lte" and fam == "lteec")
        if gas not in ("", "cpg", "ideal", "ideal_gas") and not helmholtz:
            raise NotImplementedError(f"gas_model {gas!r} with numerical_flux {runtime.get('numerical_flux')!r} is "
                                      "not implemented (calorically perfect gas, or the Helmholtz LTE gas)")
        # SimFactory_impl.hpp:140-145: the Helmholtz gas takes LTEEC, Pirozzoli, HLL or LLF
        if helmholtz and fam not in LTE_FAMILIES:
            raise ValueError("Helmholtz LTE supports LTEECFlux, Pirozzoli, HLL and LLF")
        if fam == "lteec" and not helmholtz:
            raise ValueError("the LTEEC flux needs gas_model 'lte' or 'lte_helmholtz'")
        vol, surf = PAIRINGS[fam]
        if "volume_flux" in runtime:
            vol = VOLUME_NAMES[str(runtime["volume_flux"]).lower()]
        if "surface_flux" in runtime:
            surf = SURFACE_NAMES[str(runtime["surface_flux"]).lower()]
        prec = str(runtime.get("precision_mode", "fp64")).lower()
        dim = int(runtime["dim"])
        if "num_equations" in runtime and int(runtime["num_equations"]) != dim + 2:
            raise NotImplementedError("passive scalars are not implemented (num_equations must be dim+2)")
        if str(runtime.get("ode_solver", "RK3SSP")).upper() not in ("RK3SSP",):
            raise NotImplementedError("only the RK3SSP time integrator is implemented")
        return cls(
            dim=dim,
            order=int(runtime["order"]),
            flow_model=A.FLOW_NAVIER_STOKES if is_viscous(runtime.get("flow_model")) else A.FLOW_EULER,
            volume_flux=vol,
            surface_flux=surf,
            gamma=float(runtime.get("gamma", 1.4)),
            Pr=float(runtime.get("Pr", 0.72)),
            R_gas=float(runtime.get("R_gas", 287.05)),
            mu=float(runtime.get("mu", 0.02)),
            mu0=float(runtime.get("mu0", 1.716e-5)),
            alpha_max=float(runtime.get("alpha_max", 0.5)),
            blending=bool(runtime.get("blending", (compile_time or {}).get("SUBCELL_FV_BLENDING", True))),
            precision=A.PRECISION_MIXED if prec == "mixed" else A.PRECISION_FP64,
            cfl=float(runtime.get("cfl", 1.0)),
            dt=float(runtime["dt"]) if "dt" in runtime else None,
            variable_dt=bool(runtime.get("variable_dt", "dt" not in runtime)),
            max_dt_growth=float(runtime.get("max_dt_growth", 1.2)),
            lte=helmholtz,
        )
