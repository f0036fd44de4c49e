"""CPU-side checks of the drop-in boundary: the library loads, exports every
entry point include/theseus_rhs.h declares, and rejects invalid case
descriptors before touching a GPU; the host mirror parses the reference's
runTime strings like the reference factory."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from theseus_amd import _abi as A
from theseus_amd.config import RunTime, is_viscous, numerical_flux_name
from theseus_amd.mesh import box_mesh
from theseus_amd.operator import build_descriptor

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "theseus_rhs.h").read_text()
    return sorted(set(re.findall(r"THESEUS_API\s+[\w\s\*]+?\b(theseus_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = A.load()
    decl = declared_symbols()
    assert decl and set(decl) == set(A.EXPORTED_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", str(A.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (theseus_\w+)", out))
    assert exported == set(decl), (exported ^ set(decl))
    for name in decl:
        getattr(lib, name)


def _desc(**over):
    rt = RunTime(dim=2, order=3, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_RANOCHA, surface_flux=A.SURFACE_LLF)
    mesh = box_mesh(2, (3, 2), 0.0, 1.0)
    d, keep = build_descriptor(rt, mesh)
    for k, v in over.items():
        setattr(d, k, v)
    return d, keep


@pytest.mark.parametrize("field,value,code", [
    ("dim", 4, A.ERR_INVALID_ARGUMENT),
    ("order", 0, A.ERR_INVALID_ARGUMENT),
    ("num_elements", 0, A.ERR_INVALID_ARGUMENT),
    ("flow_model", 7, A.ERR_INVALID_ARGUMENT),
    ("gamma", 1.0, A.ERR_INVALID_ARGUMENT),
    ("precision", 5, A.ERR_UNSUPPORTED),
    ("volume_flux", 9, A.ERR_UNSUPPORTED),
])
def test_create_rejects_invalid_descriptors(field, value, code):
    lib = A.load()
    d, keep = _desc(**{field: value})
    h = C.c_void_p()
    rc = lib.theseus_rhs_create(C.byref(d), C.byref(h))
    assert rc == code and not h.value
    assert lib.theseus_last_error().decode()


def test_create_rejects_inconsistent_connectivity():
    lib = A.load()
    rt = RunTime(dim=2, order=3, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_RANOCHA, surface_flux=A.SURFACE_LLF)
    mesh = box_mesh(2, (3, 2), 0.0, 1.0)
    mesh.face_neighbor[0, 1] = 2  # element 0's +x neighbour should be 1
    d, keep = build_descriptor(rt, mesh)
    h = C.c_void_p()
    assert lib.theseus_rhs_create(C.byref(d), C.byref(h)) == A.ERR_INVALID_ARGUMENT
    assert "symmetric" in lib.theseus_last_error().decode()


_GOOD_STATE = (1.2, 0.3, -0.1, 2.5 / 0.4 + 0.5 * (0.09 + 0.01) / 1.2)


@pytest.mark.parametrize("side,code,msg", [
    ((A.BC_AXIS, ()), A.ERR_UNSUPPORTED, "axisymmetric"),
    ((9, ()), A.ERR_INVALID_ARGUMENT, "unknown boundary type"),
    ((A.BC_PRESCRIBED_STATE, (1.0, 0.0, 0.0, -1.0)), A.ERR_INVALID_ARGUMENT, "positive density and pressure"),
    ((A.BC_SUPERSONIC_INFLOW, (-1.0, 0.0, 0.0, 3.0)), A.ERR_INVALID_ARGUMENT, "positive density and pressure"),
    ((A.BC_NO_SLIP_ISOTHERMAL, (0.0, 0.0, 0.0)), A.ERR_INVALID_ARGUMENT, "temperature must be positive"),
    ((A.BC_NO_SLIP_ADIABATIC, (0.0, float("nan"), 0.0)), A.ERR_INVALID_ARGUMENT, "finite"),
])
def test_create_validates_boundary_conditions(side, code, msg):
    """The reference's configuration checks (Simulation.cpp:555-820) at create time."""
    lib = A.load()
    rt = RunTime(dim=2, order=3, flow_model=A.FLOW_NAVIER_STOKES, volume_flux=A.VOLUME_RANOCHA,
                 surface_flux=A.SURFACE_LLF)
    sides = [A.BC_SLIP_WALL] * 4
    sides[2] = side
    mesh = box_mesh(2, (3, 2), 0.0, 1.0, periodic=(True, False), side_bc=sides)
    d, keep = build_descriptor(rt, mesh)
    h = C.c_void_p()
    assert lib.theseus_rhs_create(C.byref(d), C.byref(h)) == code and not h.value
    assert msg in lib.theseus_last_error().decode()


def test_boundary_payload_is_required_and_navier_stokes_boundaries_pass_validation():
    lib = A.load()
    rt = RunTime(dim=2, order=3, flow_model=A.FLOW_NAVIER_STOKES, volume_flux=A.VOLUME_RANOCHA,
                 surface_flux=A.SURFACE_LLF)
    sides = [A.BC_SLIP_WALL, A.BC_SUPERSONIC_OUTFLOW, (A.BC_PRESCRIBED_STATE, _GOOD_STATE),
             (A.BC_NO_SLIP_ISOTHERMAL, (0.1, 0.0, 300.0))]
    mesh = box_mesh(2, (3, 2), 0.0, 1.0, periodic=False, side_bc=sides)
    assert list(mesh.bc_type) == [0, 2, 3, 7] and mesh.bc_data.shape == (4, A.BC_DATA_STRIDE)
    assert (mesh.face_neighbor[:, 2] == -3).sum() == 3  # y-low side of the bottom row: boundary index 2
    d, keep = build_descriptor(rt, mesh)
    d.bc_data = C.POINTER(C.c_double)()
    h = C.c_void_p()
    assert lib.theseus_rhs_create(C.byref(d), C.byref(h)) == A.ERR_INVALID_ARGUMENT
    assert "needs bc_data" in lib.theseus_last_error().decode()
    d, keep = build_descriptor(rt, mesh)
    rc = lib.theseus_rhs_create(C.byref(d), C.byref(h))
    # without a GPU the device allocation is what fails; validation passed
    assert rc not in (A.ERR_INVALID_ARGUMENT, A.ERR_UNSUPPORTED), lib.theseus_last_error().decode()
    if h.value:
        lib.theseus_rhs_destroy(h)


def test_null_arguments():
    lib = A.load()
    assert lib.theseus_rhs_create(None, None) == A.ERR_INVALID_ARGUMENT
    assert lib.theseus_rhs_eval(None, None, None, None) == A.ERR_INVALID_ARGUMENT
    assert lib.theseus_rhs_destroy(None) == A.OK


@pytest.mark.parametrize("s,fam", [("", "chandrashekar"), ("Chandrashekar", "chandrashekar"), ("chan", "chandrashekar"),
                                   ("RANOCHA", "ranocha"), ("piflux", "pirozzoli"), ("pi", "pirozzoli"),
                                   ("hll", "hll"), ("LaxFriedrichs", "llf"), ("lfr", "llf"), ("lax_x", "llf"),
                                   ("Roe", "roe")])
def test_numerical_flux_aliases(s, fam):
    """SimFactory_impl.hpp:78-118"""
    assert numerical_flux_name(s) == fam


def test_numerical_flux_errors():
    with pytest.raises(ValueError):
        numerical_flux_name("upwind")
    assert numerical_flux_name("LTEECFlux") == "lteec"


def test_lte_gas_model_selection():
    """SimFactory_impl.hpp:75-95,140-145: the Helmholtz LTE gas is gas_model
    "lte_helmholtz", or "lte" with the LTEEC flux, and takes the LTEEC,
    Pirozzoli, HLL and LLF fluxes; the legacy LTE gas ("lte" with another
    flux) is not implemented; LTEEC needs the LTE gas."""
    rt = RunTime.from_json({"dim": 3, "order": 3, "gas_model": "lte", "numerical_flux": "lteec"},
                           {"SUBCELL_FV_BLENDING": False})
    assert (rt.volume_flux, rt.surface_flux, rt.blending) == (A.VOLUME_LTEEC, A.SURFACE_LTEEC, False)
    with pytest.raises(NotImplementedError):
        RunTime.from_json({"dim": 2, "order": 3, "gas_model": "lte", "numerical_flux": "lfr"})
    # SimFactory_impl.hpp:140-145: the Helmholtz gas also takes Pirozzoli, HLL and LLF, nothing else
    for name, pair in (("Pirozzoli", (A.VOLUME_PIROZZOLI, A.SURFACE_PIROZZOLI)),
                       ("HLL", (A.VOLUME_CENTRAL, A.SURFACE_HLL)), ("LLF", (A.VOLUME_CENTRAL, A.SURFACE_LLF))):
        rt = RunTime.from_json({"dim": 3, "order": 3, "gas_model": "lte_helmholtz", "numerical_flux": name},
                               {"SUBCELL_FV_BLENDING": False})
        assert (rt.volume_flux, rt.surface_flux, rt.lte) == (*pair, True)
    assert not RunTime.from_json({"dim": 2, "order": 3, "numerical_flux": "hll"}).lte
    for name in ("Ranocha", "Roe", "Chandrashekar"):
        with pytest.raises(ValueError, match="Helmholtz LTE supports"):
            RunTime.from_json({"dim": 2, "order": 3, "gas_model": "lte_helmholtz", "numerical_flux": name})
    with pytest.raises(ValueError):
        RunTime.from_json({"dim": 2, "order": 3, "numerical_flux": "lteec"})
    assert RunTime.from_json({"dim": 2, "order": 3}, {"SUBCELL_FV_BLENDING": False}).blending is False
    assert RunTime.from_json({"dim": 2, "order": 3}).blending is True


def test_flow_model_strings():
    for s in ("viscous", "Navier-Stokes", "cns", "ns"):
        assert is_viscous(s)
    for s in ("euler", "", None, "inviscid"):
        assert not is_viscous(s)


def test_runtime_json_defaults_and_overrides():
    rt = RunTime.from_json({"order": 3, "dim": 3, "numerical_flux": "Ranocha", "flow_model": "navier-stokes"})
    assert (rt.volume_flux, rt.surface_flux) == (A.VOLUME_RANOCHA, A.SURFACE_RANOCHA)
    assert rt.flow_model == A.FLOW_NAVIER_STOKES
    assert (rt.gamma, rt.Pr, rt.R_gas, rt.mu, rt.alpha_max) == (1.4, 0.72, 287.05, 0.02, 0.5)
    rt = RunTime.from_json({"order": 3, "dim": 3, "numerical_flux": "ranocha", "surface_flux": "llf", "mu": 1e-3,
                            "Pr": 0.71, "cfl": 0.4, "variable_dt": False, "dt": 1e-4})
    assert rt.surface_flux == A.SURFACE_LLF and rt.mu == 1e-3 and rt.dt == 1e-4 and not rt.variable_dt
    with pytest.raises(NotImplementedError):
        RunTime.from_json({"order": 3, "dim": 2, "gas_model": "lte"})


def test_box_mesh_connectivity_and_geometry():
    m = box_mesh(3, (4, 3, 2), (0, 0, 0), (2.0, 1.5, 1.0), periodic=(True, False, True))
    ne = m.num_elements
    nb = m.face_neighbor
    for e in range(ne):
        for f in range(6):
            o = nb[e, f]
            if o >= 0:
                assert nb[o, f ^ 1] == e
    assert (nb[:, 2:4] < 0).any() and (nb[:, [0, 1, 4, 5]] >= 0).all()
    h = np.array([0.5, 0.5, 0.5])
    np.testing.assert_allclose(m.elem_jac, np.prod(h))
    np.testing.assert_allclose(m.elem_adj.reshape(ne, 3, 3)[0], np.diag(np.prod(h) / h))


def _lte_arrays(nx=4, ny=4):
    rho = np.linspace(0.5, 2.0, nx)
    T = np.linspace(200.0, 400.0, ny)
    e = np.linspace(0.0, 1e5, ny)
    planes = np.ones(6 * nx * ny)
    inv = np.full(nx * ny, 300.0)
    return rho, T, e, planes, inv


@pytest.mark.parametrize("mutate,code", [
    ("blending", A.ERR_UNSUPPORTED),
    ("ideal_flux", A.ERR_INVALID_ARGUMENT),
    ("central_roe", A.ERR_INVALID_ARGUMENT),
    ("central_pirozzoli", A.ERR_INVALID_ARGUMENT),
    ("bad_grid", A.ERR_INVALID_ARGUMENT),
    ("ns_with_boundaries", A.ERR_UNSUPPORTED),
    ("bad_viscosity_model", A.ERR_INVALID_ARGUMENT),
    ("bad_sutherland", A.ERR_INVALID_ARGUMENT),
])
def test_create_lte_rejects_invalid_inputs(mutate, code):
    """theseus_rhs_create_lte validates before any device work; Navier-Stokes
    with the tabulated gas is limited to periodic meshes."""
    lib = A.load()
    rt = RunTime(dim=3, order=3, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_LTEEC,
                 surface_flux=A.SURFACE_LTEEC, blending=mutate == "blending")
    if mutate == "ideal_flux":
        rt.volume_flux = A.VOLUME_RANOCHA
    if mutate == "central_roe":  # not one of the gas's pairings (SimFactory_impl.hpp:140-245)
        rt.volume_flux, rt.surface_flux = A.VOLUME_CENTRAL, A.SURFACE_ROE
    if mutate == "central_pirozzoli":
        rt.volume_flux, rt.surface_flux = A.VOLUME_CENTRAL, A.SURFACE_PIROZZOLI
    if mutate == "ns_with_boundaries":
        rt.flow_model = A.FLOW_NAVIER_STOKES
    mesh = box_mesh(3, (2, 2, 2), 0.0, 1.0, periodic=mutate != "ns_with_boundaries")
    d, keep = build_descriptor(rt, mesh)
    arrays = [np.ascontiguousarray(a) for a in _lte_arrays()]
    if mutate == "bad_grid":
        arrays[1][2] = arrays[1][1]  # T grid not strictly increasing
    t = A.LteTable()
    t.nx, t.ny = 4, 4
    t.rho_grid, t.T_grid, t.e_grid, t.planes, t.inverse = (a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrays)
    if mutate == "bad_viscosity_model":
        t.viscosity_model = 2
    if mutate == "bad_sutherland":
        t.viscosity_model, t.mu0, t.T0, t.Ts = 1, 1.716e-5, 0.0, 110.4  # T0 must be positive
    h = C.c_void_p()
    assert lib.theseus_rhs_create_lte(C.byref(d), C.byref(t), C.byref(h)) == code
    assert not h.value and lib.theseus_last_error().decode()
    if mutate == "ns_with_boundaries":
        assert "periodic" in lib.theseus_last_error().decode()
