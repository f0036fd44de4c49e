"""CPU oracle: integral measures (RHSOperator_impl.hpp:530-717) and the
variable-step loop (src/Simulation.cpp:1136-1158,1240-1301), pinned by closed
forms: GLL quadrature is exact for polynomials of degree <= 2p-1 per
direction, nodal extremes of monotone fields sit at known mesh corners, and a
free stream has a constant stability rate, so the loop's step count and its
last (shortened) step follow from t_final / (cfl / rate)."""
import numpy as np
import pytest

from theseus_amd import _abi as A
from theseus_amd.config import RunTime
from theseus_amd.mesh import box_mesh, gll_nodes, tensor_mesh
from theseus_amd.operator import build_descriptor

GAMMA, R_GAS = 1.4, 287.0


def _case(oracle, mesh, p, flow=A.FLOW_EULER):
    rt = RunTime(dim=mesh.dim, order=p, flow_model=flow, volume_flux=A.VOLUME_RANOCHA, surface_flux=A.SURFACE_LLF,
                 gamma=GAMMA, R_gas=R_GAS)
    desc, keep = build_descriptor(rt, mesh)
    c = oracle.OracleCase(desc)
    c._keep = keep
    return c


def _pack(rho, mom, rhoE):
    return np.concatenate([rho[None], np.asarray(mom), rhoE[None]], axis=0).reshape(-1)


def _mesh3():
    # uneven element sizes: breaks differ per direction
    return tensor_mesh([np.array([0.0, 0.3, 0.55, 1.0]), np.array([0.0, 0.5, 1.0]), np.array([0.0, 0.2, 0.7, 1.0])],
                       periodic=True)


def test_constant_state_measures(oracle):
    mesh = _mesh3()
    p = 3
    c = _case(oracle, mesh, p)
    N = mesh.num_elements * (p + 1) ** 3
    rho, u, pr = 1.3, np.array([0.4, -0.2, 0.7]), 2.5
    rhoE = pr / (GAMMA - 1) + 0.5 * rho * u @ u
    U = _pack(np.full(N, rho), [np.full(N, rho * v) for v in u], np.full(N, rhoE))
    m = c.integral_measures(U)
    V = 1.0
    assert m["mass"] == pytest.approx(rho * V, rel=1e-14)
    assert m["ke"] == pytest.approx(0.5 * rho * (u @ u) * V, rel=1e-14)
    assert m["en"] == pytest.approx(rhoE * V, rel=1e-14)
    for k in ("max_press", "min_press"):
        assert m[k] == pytest.approx(pr, rel=1e-14)
    for k in ("max_temp", "min_temp"):
        assert m[k] == pytest.approx(pr / (rho * R_GAS), rel=1e-14)
    assert m["max_dens"] == rho and m["min_dens"] == rho


@pytest.mark.parametrize("p", [2, 3, 4])
def test_polynomial_integrals_and_corner_extremes(oracle, p):
    """rho = 2 + x + y z and rhoE = 10 + x y z, then rho = rho0, m = (x, y, z):
    every integrand has degree <= 2 per direction, so the GLL node rule
    (exact to degree 2p - 1) integrates mass and energy exactly; with a
    constant density the kinetic energy |m|^2 / (2 rho0) is also exact."""
    mesh = _mesh3()
    c = _case(oracle, mesh, p)
    X = mesh.node_coordinates(gll_nodes(p)[0])
    x, y, z = X
    # mass and energy: polynomial density and total energy
    rho = 2 + x + y * z
    rhoE = 10 + x * y * z
    U = _pack(rho, [0 * x, 0 * x, 0 * x], rhoE)
    m = c.integral_measures(U)
    assert m["mass"] == pytest.approx(2 + 0.5 + 0.25, rel=1e-14)
    assert m["en"] == pytest.approx(10 + 0.125, rel=1e-14)
    assert m["ke"] == 0.0
    assert m["min_dens"] == pytest.approx(2.0, rel=1e-15)   # node at the origin
    assert m["max_dens"] == pytest.approx(4.0, rel=1e-15)   # node at (1, 1, 1)
    # kinetic energy and pressure extremes: constant density rho0
    rho0 = 2.0
    rhoE = 10 + x
    U = _pack(np.full_like(x, rho0), [x, y, z], rhoE)
    m = c.integral_measures(U)
    assert m["ke"] == pytest.approx((1 / 3) * 3 / (2 * rho0), rel=1e-14)
    # p = 0.4 (10 + x - (x^2 + y^2 + z^2) / 4): largest at (1, 0, 0), smallest at (0, 1, 1)
    pmax = (GAMMA - 1) * (10 + 1 - 0.25)
    pmin = (GAMMA - 1) * (10 - 0.5)
    assert m["max_press"] == pytest.approx(pmax, rel=1e-14)
    assert m["min_press"] == pytest.approx(pmin, rel=1e-14)
    assert m["max_temp"] == pytest.approx(pmax / (rho0 * R_GAS), rel=1e-14)
    assert m["min_temp"] == pytest.approx(pmin / (rho0 * R_GAS), rel=1e-14)


def test_measures_2d_and_reference_initial_values(oracle):
    """2-D: maxima start at 0 and minima at 1e32 (RHSOperator_impl.hpp:655-663),
    so a negative pressure leaves max_press at 0."""
    mesh = box_mesh(2, (3, 2), 0.0, (2.0, 1.0))
    p = 3
    c = _case(oracle, mesh, p)
    X = mesh.node_coordinates(gll_nodes(p)[0])
    x, y = X
    rho = 1 + 0.25 * x * y
    U = _pack(rho, [0 * x, 0 * x], -1.0 + 0 * x)  # rhoE < 0: negative pressure everywhere
    m = c.integral_measures(U)
    assert m["mass"] == pytest.approx(2.0 + 0.25 * 2.0 * 0.5, rel=1e-14)
    assert m["en"] == pytest.approx(-2.0, rel=1e-14)
    assert m["max_press"] == 0.0 and m["max_temp"] == 0.0
    assert m["min_press"] == pytest.approx(-(GAMMA - 1), rel=1e-15)


def test_variable_step_loop_on_a_free_stream(oracle):
    """Uniform flow: the rate is constant and the state stays put, so
    dt = cfl / rate every step, the loop takes ceil(t_final / dt) steps and the
    last one is shortened to land on t_final."""
    mesh = box_mesh(2, (3, 3), 0.0, 1.0)
    p = 3
    c = _case(oracle, mesh, p)
    N = mesh.num_elements * (p + 1) ** 2
    q = np.array([1.1, 0.3, -0.2, 2.6])
    U = np.repeat(q, N)
    rate = c.stability(U)
    total = max(rate[0], rate[1]) + rate[2]
    cfl = 0.4
    dt = cfl / total
    t_final = 3.37 * dt
    U1, t, dtn, dt_real, steps = c.adaptive_steps(U, cfl, 1.2, t_final, nsteps=100)
    assert steps == 4
    assert t == pytest.approx(t_final, rel=1e-15)
    assert dt_real == pytest.approx(0.37 * dt, rel=1e-12)
    assert dtn == pytest.approx(dt, rel=1e-12)
    np.testing.assert_allclose(U1, U, atol=1e-13)
    # a step cap ends the loop early
    U2, t2, _, _, s2 = c.adaptive_steps(U, cfl, 1.2, t_final, nsteps=2)
    assert s2 == 2 and t2 == pytest.approx(2 * dt, rel=1e-15)


def test_lte_measures_follow_the_gas_temperature(oracle):
    """Tabulated gas: the temperature comes from recovery on the table and the
    pressure from the table.  A state built from a known temperature field
    T = 3000 + 1000 y (with the synthetic gas's closed-form e(T)) recovers
    extremes 3000 and 4000, and the pressure extremes are rho R T of the same
    nodes (the ideal-gas branch would give (gamma - 1) rho e instead)."""
    from theseus_amd.lte import R_AIR, SyntheticGas

    mesh = box_mesh(3, (2, 2, 2), 0.0, 1.0)
    p = 2
    rt = RunTime(dim=3, order=p, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_LTEEC, surface_flux=A.SURFACE_LTEEC,
                 R_gas=R_AIR)
    desc, keep = build_descriptor(rt, mesh)
    c = oracle.OracleCase(desc)
    c._keep = keep
    table = oracle.LteTable()
    table.attach(c)
    X = mesh.node_coordinates(gll_nodes(p)[0])
    x, y, z = X
    gas = SyntheticGas()
    T0 = 10 ** 2.3
    rho = 0.1 * (1 + 0.2 * x)
    T = 3000 + 1000 * y
    e = np.array([gas.energy(t, T0) for t in T])
    U = _pack(rho, [0 * x, 0 * x, 0 * x], rho * e)
    m = c.integral_measures(U)
    # e = H - T H_T of the bicubic interpolant: H_T carries an O(h^3) error
    # (h ~ 70 K near 4000 K, where cv varies fastest), ~1e-3 K in T
    assert m["min_temp"] == pytest.approx(3000.0, rel=1e-6)
    assert m["max_temp"] == pytest.approx(4000.0, rel=1e-6)
    assert m["min_press"] == pytest.approx(0.1 * R_AIR * 3000.0, rel=1e-5)
    assert m["max_press"] == pytest.approx(0.12 * R_AIR * 4000.0, rel=1e-5)
    assert m["mass"] == pytest.approx(0.11, rel=1e-14)
    assert m["ke"] == 0.0
