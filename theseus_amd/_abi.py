"""ctypes view of the C-ABI in include/theseus_rhs.h.

Loading is explicit and loud: ``load()`` raises if libtheseus_rhs.so has not
been built; nothing in this package falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libtheseus_rhs.so"

# error codes
OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_UNSUPPORTED = -2
ERR_DEVICE = -3
ERR_NONFINITE = -4
ERR_CAPACITY = -5

FLOW_EULER = 0
FLOW_NAVIER_STOKES = 1

VOLUME_CENTRAL = 0
VOLUME_CHANDRASHEKAR = 1
VOLUME_RANOCHA = 2
VOLUME_PIROZZOLI = 3
VOLUME_LTEEC = 4

SURFACE_CHANDRASHEKAR = 0
SURFACE_RANOCHA = 1
SURFACE_PIROZZOLI = 2
SURFACE_ROE = 3
SURFACE_HLL = 4
SURFACE_LLF = 5
SURFACE_LTEEC = 6

BC_SLIP_WALL = 0
BC_SUPERSONIC_INFLOW = 1
BC_SUPERSONIC_OUTFLOW = 2
BC_PRESCRIBED_STATE = 3
BC_SYMMETRY = 4
BC_AXIS = 5
BC_NO_SLIP_ADIABATIC = 6
BC_NO_SLIP_ISOTHERMAL = 7
BC_DATA_STRIDE = 8
POOL_MAX_BC = 16

PRECISION_FP64 = 0
PRECISION_MIXED = 1

EXPORTED_SYMBOLS = (
    "theseus_rhs_create",
    "theseus_rhs_create_lte",
    "theseus_rhs_destroy",
    "theseus_rhs_sizes",
    "theseus_rhs_eval",
    "theseus_rhs_estimate_stability",
    "theseus_rhs_step_ssprk3",
    "theseus_rhs_get_alpha",
    "theseus_rhs_integral_measures",
    "theseus_rhs_integral_measures_baseline",
    "theseus_rhs_integral_measures_noalloc",
    "theseus_rhs_entropy_production",
    "theseus_orient_elements",
    "theseus_orient_node",
    "theseus_rhs_step_ssprk3_cfl",
    "theseus_rhs_enable_kernel_timing",
    "theseus_rhs_kernel_timing",
    "theseus_last_error",
    "theseus_pool_create",
    "theseus_pool_destroy",
    "theseus_pool_info",
    "theseus_pool_admit",
    "theseus_pool_step",
    "theseus_pool_read",
    "theseus_pool_integral_measures",
    "theseus_pool_release",
    "theseus_pool_group_create",
    "theseus_pool_group_step",
    "theseus_pool_group_destroy",
)


class CaseDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("order", C.c_int),
        ("num_elements", C.c_int),
        ("flow_model", C.c_int),
        ("volume_flux", C.c_int),
        ("surface_flux", C.c_int),
        ("precision", C.c_int),
        ("blending", C.c_int),
        ("gamma", C.c_double),
        ("Pr", C.c_double),
        ("R_gas", C.c_double),
        ("mu", C.c_double),
        ("mu_bulk", C.c_double),
        ("alpha_max", C.c_double),
        ("alpha_min", C.c_double),
        ("elem_jac", C.POINTER(C.c_double)),
        ("elem_adj", C.POINTER(C.c_double)),
        ("face_neighbor", C.POINTER(C.c_int)),
        ("num_bc", C.c_int),
        ("bc_type", C.POINTER(C.c_int)),
        ("bc_data", C.POINTER(C.c_double)),
        ("elem_orient", C.POINTER(C.c_ubyte)),
    ]


class LteTable(C.Structure):
    _fields_ = [
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("rho_grid", C.POINTER(C.c_double)),
        ("T_grid", C.POINTER(C.c_double)),
        ("e_grid", C.POINTER(C.c_double)),
        ("planes", C.POINTER(C.c_double)),
        ("inverse", C.POINTER(C.c_double)),
        ("viscosity_model", C.c_int),
        ("mu0", C.c_double),
        ("T0", C.c_double),
        ("Ts", C.c_double),
    ]


class PoolDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("order", C.c_int),
        ("flow_model", C.c_int),
        ("volume_flux", C.c_int),
        ("surface_flux", C.c_int),
        ("precision", C.c_int),
        ("blending", C.c_int),
        ("gamma", C.c_double),
        ("Pr", C.c_double),
        ("R_gas", C.c_double),
        ("mu", C.c_double),
        ("mu_bulk", C.c_double),
        ("alpha_max", C.c_double),
        ("alpha_min", C.c_double),
        ("hbm_budget_bytes", C.c_longlong),
        ("max_cases", C.c_int),
        ("tile", C.c_int),
        ("xcd_remap", C.c_int),
    ]


class Stability(C.Structure):
    _fields_ = [("advective_rate", C.c_double), ("surface_rate", C.c_double), ("diffusive_rate", C.c_double)]


class IntegralMeasures(C.Structure):
    """theseus_integral_measures (dgsem_cache.hpp:16-26)."""
    _fields_ = [(n, C.c_double) for n in ("mass", "ke", "en", "max_press", "min_press", "max_temp", "min_temp",
                                          "max_dens", "min_dens")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class EntropyProduction(C.Structure):
    """theseus_entropy_production: sum w J S.dU and sum w J |S.dU|."""
    _fields_ = [("production", C.c_double), ("magnitude", C.c_double)]


class Clock(C.Structure):
    """theseus_clock: the device-resident variable-step clock (64 bytes)."""
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("dt_real", C.c_double), ("t_final", C.c_double),
                ("cfl", C.c_double), ("max_dt_growth", C.c_double), ("rate", C.c_double), ("steps", C.c_longlong)]


_lib = None


def _init_torch_device() -> None:
    """When torch is present and sees a GPU, initialise its device before
    the library makes its first HIP call.  In one process the two share the
    HIP runtime, and torch's device initialisation has been seen to report
    no GPU when the library had already created a stream of its own."""
    try:
        import torch
    except ImportError:
        return
    if getattr(torch.version, "hip", None) and torch.cuda.is_available():
        torch.cuda.init()


def load(path: Path | str | None = None) -> C.CDLL:
    """Load libtheseus_rhs.so (raises FileNotFoundError when it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built; run `python -m theseus_amd.build` (or __graft_entry__.build())")
    _init_torch_device()
    lib = C.CDLL(str(p))
    lib.theseus_rhs_create.argtypes = [C.POINTER(CaseDesc), C.POINTER(C.c_void_p)]
    lib.theseus_rhs_create.restype = C.c_int
    lib.theseus_rhs_create_lte.argtypes = [C.POINTER(CaseDesc), C.POINTER(LteTable), C.POINTER(C.c_void_p)]
    lib.theseus_rhs_create_lte.restype = C.c_int
    lib.theseus_rhs_destroy.argtypes = [C.c_void_p]
    lib.theseus_rhs_destroy.restype = C.c_int
    lib.theseus_rhs_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
    lib.theseus_rhs_sizes.restype = C.c_int
    lib.theseus_rhs_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.theseus_rhs_eval.restype = C.c_int
    lib.theseus_rhs_estimate_stability.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Stability), C.c_void_p]
    lib.theseus_rhs_estimate_stability.restype = C.c_int
    lib.theseus_rhs_step_ssprk3.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p]
    lib.theseus_rhs_step_ssprk3.restype = C.c_int
    lib.theseus_rhs_get_alpha.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.theseus_rhs_get_alpha.restype = C.c_int
    lib.theseus_rhs_enable_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    lib.theseus_rhs_enable_kernel_timing.restype = C.c_int
    lib.theseus_rhs_kernel_timing.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    lib.theseus_rhs_kernel_timing.restype = C.c_int
    lib.theseus_rhs_integral_measures.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(IntegralMeasures), C.c_void_p]
    lib.theseus_rhs_integral_measures.restype = C.c_int
    lib.theseus_rhs_integral_measures_baseline.argtypes = [C.c_void_p, C.POINTER(IntegralMeasures)]
    lib.theseus_rhs_integral_measures_baseline.restype = C.c_int
    lib.theseus_rhs_integral_measures_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.theseus_rhs_integral_measures_noalloc.restype = C.c_int
    lib.theseus_rhs_entropy_production.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(EntropyProduction),
                                                   C.c_void_p]
    lib.theseus_rhs_entropy_production.restype = C.c_int
    lib.theseus_orient_elements.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p]
    lib.theseus_orient_elements.restype = C.c_int
    lib.theseus_orient_node.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.theseus_orient_node.restype = C.c_int
    lib.theseus_rhs_step_ssprk3_cfl.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.theseus_rhs_step_ssprk3_cfl.restype = C.c_int
    lib.theseus_last_error.argtypes = []
    lib.theseus_last_error.restype = C.c_char_p
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    lib.theseus_pool_create.argtypes = [C.POINTER(PoolDesc), C.POINTER(C.c_void_p)]
    lib.theseus_pool_destroy.argtypes = [vp]
    lib.theseus_pool_info.argtypes = [vp, ip, ip, ip, ip, ip, ip]
    lib.theseus_pool_admit.argtypes = [vp, C.POINTER(CaseDesc), vp, C.c_double, ip, dp, vp]
    lib.theseus_pool_step.argtypes = [vp, vp]
    lib.theseus_pool_read.argtypes = [vp, C.c_int, vp, vp]
    lib.theseus_pool_release.argtypes = [vp, C.c_int, vp]
    lib.theseus_pool_integral_measures.argtypes = [vp, C.c_int, C.POINTER(IntegralMeasures), vp]
    lib.theseus_pool_group_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
    lib.theseus_pool_group_step.argtypes = [vp, vp]
    lib.theseus_pool_group_destroy.argtypes = [vp]
    for name in ("theseus_pool_create", "theseus_pool_destroy", "theseus_pool_info", "theseus_pool_admit",
                 "theseus_pool_step", "theseus_pool_read", "theseus_pool_release", "theseus_pool_integral_measures",
                 "theseus_pool_group_create", "theseus_pool_group_step", "theseus_pool_group_destroy"):
        getattr(lib, name).restype = C.c_int
    if path is None:
        _lib = lib
    return lib


class TheseusError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"theseus error {code}: {msg}")
        self.code = code


def check(lib: C.CDLL, rc: int) -> None:
    if rc != OK:
        raise TheseusError(rc, lib.theseus_last_error().decode(errors="replace"))
