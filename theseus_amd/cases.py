"""Initial conditions and the workload definitions the engine is measured on.

ICs produce the byNODES L-vector (numpy float64, shape (neq*ndof_total,))
by nodal interpolation at the GLL nodes, as the reference's
``ProjectCoefficient`` does for a DG GLL space (src/Simulation.cpp:1076-1109).
"""
from __future__ import annotations

import numpy as np

from . import _abi as A
from .config import RunTime
from .mesh import BoxMesh, box_mesh, gll_nodes


def _pack(rho, mom, E):
    return np.concatenate([rho[None], np.asarray(mom), E[None]], axis=0).reshape(-1)


def taylor_green(mesh: BoxMesh, order: int, gamma: float = 1.4, Ma: float = 0.1) -> np.ndarray:
    """TaylorGreenVortexIC (StateInit.hpp:139-162)."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    x, y, z = X[0], X[1], X[2]
    p0 = 1.0 / (gamma * Ma * Ma)
    den = np.ones_like(x)
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    p = p0 + 1.0 / 16.0 * (np.cos(2.0 * x) + np.cos(2.0 * y)) * (np.cos(2.0 * z) + 2)
    E = p / (gamma - 1.0) + 0.5 * den * (u * u + v * v + w * w)
    return _pack(den, [den * u, den * v, den * w], E)


def taylor_green_2d(mesh: BoxMesh, order: int, gamma: float = 1.4, Ma: float = 0.1) -> np.ndarray:
    """TaylorGreenVortex2DIC (StateInit.hpp:218-246)."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    x, y = X[0], X[1]
    p0 = 1.0 / (gamma * Ma * Ma)
    den = np.ones_like(x)
    u = np.sin(x) * np.cos(y)
    v = -np.cos(x) * np.sin(y)
    p = p0 + 0.25 * (np.cos(2.0 * x) + np.cos(2.0 * y))
    E = p / (gamma - 1.0) + 0.5 * den * (u * u + v * v)
    return _pack(den, [den * u, den * v], E)


def lid_driven_cavity(mesh, order: int, Ma: float = 0.1, gamma: float = 1.4) -> np.ndarray:
    """LidDrivenCavityIC (StateInit.hpp:258-268): fluid at rest, rho = 1,
    p = 1/(Ma^2 gamma)."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    one = np.ones(X.shape[1])
    p = 1.0 / (Ma * Ma * gamma)
    return _pack(one, [0.0 * one for _ in range(mesh.dim)], one * (p / (gamma - 1.0)))


def compressible_taylor_green(mesh: BoxMesh, order: int, Ma: float = 0.1) -> np.ndarray:
    """CompressibleTaylorGreenVortexIC (StateInit.hpp:176-205): p0 = 1e5,
    T0 = 273.15, R = 287.05, gamma = 1.4, U0 = Ma sqrt(gamma R T0), rho = p/(R T0)."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    x, y, z = X[0], X[1], X[2]
    p0, T0, R, gamma = 1.0e5, 273.15, 287.05, 1.4
    rho0 = p0 / (R * T0)
    U0 = Ma * np.sqrt(gamma * R * T0)
    u = U0 * np.sin(x) * np.cos(y) * np.cos(z)
    v = -U0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    p = p0 + (rho0 * U0 * U0) / 16.0 * (np.cos(2.0 * x) + np.cos(2.0 * y)) * (np.cos(2.0 * z) + 2)
    rho = p / (R * T0)
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return _pack(rho, [rho * u, rho * v, rho * w], E)


def kelvin_helmholtz(mesh: BoxMesh, order: int, gammaM1Inverse: float = 2.5) -> np.ndarray:
    """KelvinHelmholtzInstabilyIC (StateInit.hpp:662-695)."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    x, y = X[0], X[1]
    B = np.tanh(15.0 * y + 7.5) - np.tanh(15.0 * y - 7.5)
    rho = 0.5 + 0.75 * B
    u = 0.5 * (B - 1.0)
    v = 0.1 * np.sin(2.0 * np.pi * x)
    p = np.ones_like(x)
    E = p * gammaM1Inverse + rho * 0.5 * (u * u + v * v)
    return _pack(rho, [rho * u, rho * v], E)


def forward_facing_step(mesh, order: int, gammaM1Inverse: float = 2.5) -> np.ndarray:
    """ForwardFacingStepIC (StateInit.hpp:805-816): uniform Mach-3 flow,
    rho = 1.4, u = (3, 0), p = 1; the energy uses the parameter 1/(gamma-1)."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    one = np.ones(X.shape[1])
    E = 1.0 * gammaM1Inverse + 0.5 * 1.4 * 3.0 * 3.0
    return _pack(1.4 * one, [1.4 * 3.0 * one, 0.0 * one], E * one)


def shu_vortex(mesh: BoxMesh, order: int, gamma: float = 1.4, strength: float = 5.0,
               center=(0.0, 0.0), u_inf=(1.0, 1.0)) -> np.ndarray:
    """Textbook isentropic vortex (Shu): rho_inf = p_inf = 1, u_inf = (1,1)."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    dx, dy = X[0] - center[0], X[1] - center[1]
    r2 = dx * dx + dy * dy
    b = strength
    du = b / (2 * np.pi) * np.exp(0.5 * (1 - r2))
    T = 1.0 - (gamma - 1) * b * b / (8 * gamma * np.pi**2) * np.exp(1 - r2)
    rho = T ** (1.0 / (gamma - 1))
    p = rho**gamma
    u = u_inf[0] - du * dy
    v = u_inf[1] + du * dx
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return _pack(rho, [rho * u, rho * v], E)


def reference_isentropic_vortex(mesh: BoxMesh, order: int, radius=0.2, Minf=0.5, beta=0.2,
                                R_gas=287.05, gamma=1.4) -> np.ndarray:
    """IsentropicVortexIC (StateInit.hpp:392-434), the reference's own vortex."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    x, y = X[0], X[1]
    vel_inf, den_inf = 1.0, 1.0
    pres_inf = (den_inf / gamma) * (vel_inf / Minf) * (vel_inf / Minf)
    temp_inf = pres_inf / (den_inf * R_gas)
    r2rad = (x * x + y * y) / (radius * radius)
    shrinv1 = 1.0 / (gamma - 1.0)
    velX = vel_inf * (1 - beta * y / radius * np.exp(-0.5 * r2rad))
    velY = vel_inf * beta * x / radius * np.exp(-0.5 * r2rad)
    vel2 = velX * velX + velY * velY
    specific_heat = R_gas * gamma * shrinv1
    temp = temp_inf - 0.5 * (vel_inf * beta) ** 2 / specific_heat * np.exp(-r2rad)
    den = den_inf * (temp / temp_inf) ** shrinv1
    pres = den * R_gas * temp
    energy = shrinv1 * pres / den + 0.5 * vel2
    return _pack(den, [den * velX, den * velY], den * energy)


def lte_vortex(mesh: BoxMesh, order: int, radius=0.2, vel_inf=10.0, beta=1.0, rho_inf=0.1,
               temp_inf=275.0) -> np.ndarray:
    """LTEVortexIC (StateInit.hpp:15-69): a vortex centred at the origin whose
    field is shaped with the calorically perfect gas's relations (gamma 1.4,
    R 287.05, whatever the case's gas): Gaussian temperature dip
    T = T_inf - (vel_inf beta)^2/(2 cp) exp(-r^2/radius^2), clamped below at
    0.2 T_inf, isentropic density and pressure from it, rhoE = p/(gamma-1) +
    rho |u|^2/2."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    x, y = X[0], X[1]
    gamma, R_gas = 1.4, 287.05
    gm1 = gamma - 1.0
    cp = gamma * R_gas / gm1
    pres_inf = rho_inf * R_gas * temp_inf
    r2rad = (x * x + y * y) / (radius * radius)
    velX = vel_inf * (1.0 - beta * y / radius * np.exp(-0.5 * r2rad))
    velY = vel_inf * (beta * x / radius * np.exp(-0.5 * r2rad))
    temp = temp_inf - 0.5 * (vel_inf * beta) ** 2 / cp * np.exp(-r2rad)
    temp = np.maximum(temp, 0.2 * temp_inf)
    den = rho_inf * (temp / temp_inf) ** (1.0 / gm1)
    pres = pres_inf * (temp / temp_inf) ** (gamma / gm1)
    rhoE = pres / gm1 + 0.5 * den * (velX * velX + velY * velY)
    return _pack(den, [den * velX, den * velY], rhoE)


def four_quadrant_riemann(mesh: BoxMesh, order: int, gamma: float = 1.4, split=(0.5, 0.5)) -> np.ndarray:
    """2-D four-quadrant Riemann problem (configuration with quadrant states
    NE (1.5,0,0,1.5), NW (0.5323,1.206,0,0.3), SW (0.138,1.206,1.206,0.029),
    SE (0.5323,0,1.206,0.3))."""
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    east = X[0] >= split[0]
    north = X[1] >= split[1]
    states = {
        (True, True): (1.5, 0.0, 0.0, 1.5),
        (False, True): (0.5323, 1.206, 0.0, 0.3),
        (False, False): (0.138, 1.206, 1.206, 0.029),
        (True, False): (0.5323, 0.0, 1.206, 0.3),
    }
    rho = np.empty_like(X[0])
    u = np.empty_like(X[0])
    v = np.empty_like(X[0])
    p = np.empty_like(X[0])
    for (ee, nn), (r, uu, vv, pp) in states.items():
        m = (east == ee) & (north == nn)
        rho[m], u[m], v[m], p[m] = r, uu, vv, pp
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return _pack(rho, [rho * u, rho * v], E)


# ---------------------------------------------------------------------------
# Workloads (sizes as the engine is benchmarked on)
# ---------------------------------------------------------------------------

def c1_vortex2d(n: int = 16, order: int = 3):
    """2-D Euler isentropic vortex, periodic [-5,5]^2, Chandrashekar volume + Roe face."""
    mesh = box_mesh(2, n, -5.0, 5.0, periodic=True)
    rt = RunTime(dim=2, order=order, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_CHANDRASHEKAR,
                 surface_flux=A.SURFACE_ROE, cfl=0.5, blending=True)
    return rt, mesh, shu_vortex(mesh, order, rt.gamma)


def c2_tgv3d(n: int = 32, order: int = 3, Ma: float = 0.1, Re: float = 1600.0, Pr: float = 0.71):
    """3-D compressible Navier-Stokes Taylor-Green vortex, periodic [0,2pi]^3,
    Ranocha volume + Lax-Friedrichs face, BR1, constant viscosity mu = 1/Re
    (rho0 = 1, U0 = 1, L = 1)."""
    mesh = box_mesh(3, n, 0.0, 2 * np.pi, periodic=True)
    rt = RunTime(dim=3, order=order, flow_model=A.FLOW_NAVIER_STOKES, volume_flux=A.VOLUME_RANOCHA,
                 surface_flux=A.SURFACE_LLF, Pr=Pr, mu=1.0 / Re, blending=True)
    return rt, mesh, taylor_green(mesh, order, rt.gamma, Ma)


def c3_riemann2d(n: int = 128, order: int = 4):
    """2-D four-quadrant Riemann problem on [0,1]^2, transmissive boundaries,
    central volume + HLL face (reference pairing 'hll'), blending on."""
    mesh = box_mesh(2, n, 0.0, 1.0, periodic=False, bc=A.BC_SUPERSONIC_OUTFLOW)
    rt = RunTime(dim=2, order=order, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_CENTRAL,
                 surface_flux=A.SURFACE_HLL, blending=True)
    return rt, mesh, four_quadrant_riemann(mesh, order, rt.gamma)


def lte_taylor_green(mesh: BoxMesh, order: int, table, rho0: float = 0.1, T0: float = 3500.0, mach: float = 0.3):
    """Taylor-Green vortex of the tabulated LTE gas: rho = rho0, velocity
    amplitude U = mach * c(rho0, T0) with c^2 = gamma_eff R T of the synthetic
    gas, pressure p0 + rho0 U^2/16 (cos 2x + cos 2y)(cos 2z + 2) with
    p0 = rho0 R T0 (on a 2-D mesh the planar vortex, p0 + rho0 U^2/4
    (cos 2x + cos 2y)); T = p/(rho R) (the synthetic gas's pressure
    H_rho rho^2 = rho R T) and the energy from the table's interpolant."""
    from .lte import R_AIR, SyntheticGas, _energy_at

    gas = SyntheticGas()
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    x, y = X[0], X[1]
    z = X[2] if mesh.dim == 3 else np.zeros_like(x)
    gam = 1.0 + R_AIR / gas.cv(T0)
    U = mach * np.sqrt(gam * R_AIR * T0)
    u = U * np.sin(x) * np.cos(y) * np.cos(z)
    v = -U * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    rho = np.full_like(x, rho0)
    if mesh.dim == 3:
        p = rho0 * R_AIR * T0 + rho0 * U * U / 16.0 * (np.cos(2.0 * x) + np.cos(2.0 * y)) * (np.cos(2.0 * z) + 2)
    else:
        p = rho0 * R_AIR * T0 + rho0 * U * U / 4.0 * (np.cos(2.0 * x) + np.cos(2.0 * y))
    T = p / (rho0 * R_AIR)
    i = int(np.searchsorted(table.rho, rho0, side="right") - 1)
    e = _energy_at(table, min(i, table.nx - 2), rho0, T)
    E = rho * e + 0.5 * rho * (u * u + v * v + w * w)
    mom = [rho * u, rho * v, rho * w] if mesh.dim == 3 else [rho * u, rho * v]
    return _pack(rho, mom, E)


def c4_lte_tgv(n: int = 24, order: int = 3, table=None, viscous: bool = False, dim: int = 3, flux: str = "lteec"):
    """3-D LTE Taylor-Green vortex, periodic [0,2pi]^3, LTEEC flux, the
    synthetic dissociating-gas table (256 x 256, log rho in [-4,1], log T in
    [2.3,4.2]), rho0 = 0.1, T0 = 3500 K, amplitude Mach 0.3.  viscous: the
    Navier-Stokes form (BR1, viscosity and conductivity from the table's
    mu and lambda planes); dim = 2: the planar vortex on [0,2pi]^2; flux:
    another pairing of the tabulated gas ("pirozzoli", "hll", "llf")."""
    from .config import LTE_FAMILIES, PAIRINGS
    from .lte import synthetic_table

    if flux not in LTE_FAMILIES:
        raise ValueError(f"the tabulated gas takes the fluxes {LTE_FAMILIES}, not {flux!r}")
    table = table if table is not None else synthetic_table()
    mesh = box_mesh(dim, n, 0.0, 2 * np.pi, periodic=True)
    vol, surf = PAIRINGS[flux]
    rt = RunTime(dim=dim, order=order, flow_model=A.FLOW_NAVIER_STOKES if viscous else A.FLOW_EULER,
                 volume_flux=vol, surface_flux=surf, blending=False, lte=True)
    return rt, mesh, lte_taylor_green(mesh, order, table), table
