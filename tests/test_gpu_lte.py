"""GPU parity of the tabulated LTE gas with the LTEEC flux (the LTE
Taylor-Green workload's physics) against the CPU FP64 oracle on the same
synthetic Helmholtz table and states: RHS relative L2 <= 1e-12, the
stability estimate, and the state after SSP-RK3 steps <= 1e-10."""
import sys

import numpy as np
import pytest

from .gpu_helpers import per_field_rel, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def table():
    from theseus_amd.lte import synthetic_table

    return synthetic_table()


def _oracle(oracle, rt, mesh, table):
    from theseus_amd.operator import build_descriptor

    desc, keep = build_descriptor(rt, mesh)
    c = oracle.OracleCase(desc)
    c._keep = keep
    t = oracle.LteTable(arrays=(table.rho, table.T, table.e, table.planes.reshape(-1), table.inv.reshape(-1)))
    if getattr(table, "sutherland", None) is not None:
        t.set_sutherland(*table.sutherland)
    t.attach(c)
    c._table = t
    return c


@pytest.mark.parametrize("n,order", [(2, 3), (3, 2), (2, 4)])
def test_lte_rhs_parity(oracle, table, n, order):
    import torch

    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import MakeRHSOperator

    _say(f"case lte rhs n={n} p={order}")
    rt, mesh, U, _ = c4_lte_tgv(n=n, order=order, table=table)
    ref = _oracle(oracle, rt, mesh, table).rhs(U)
    op = MakeRHSOperator(rt, mesh, lte_table=table)
    u = torch.from_numpy(U).cuda()
    du = torch.empty_like(u)
    op.Mult(u, du)
    torch.cuda.synchronize()
    got = du.cpu().numpy()
    op.close()
    assert np.isfinite(got).all()
    err = rel_l2(got, ref)
    assert err <= 1e-12, (err, per_field_rel(got, ref, 5))


def test_lte_stability_and_steps_parity(oracle, table):
    import torch

    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import MakeRHSOperator

    _say("case lte steps n=2 p=3")
    rt, mesh, U, _ = c4_lte_tgv(n=2, order=3, table=table)
    ora = _oracle(oracle, rt, mesh, table)
    rates = ora.stability(U)
    op = MakeRHSOperator(rt, mesh, lte_table=table)
    u = torch.from_numpy(U).cuda()
    s = op.EstimateStability(u)
    got = np.array([s.advective_rate, s.surface_rate, s.diffusive_rate])
    np.testing.assert_allclose(got, rates, rtol=1e-12, atol=0)
    dt = 0.5 / (max(rates[0], rates[1]) + rates[2])
    ref = U.copy()
    for _ in range(20):
        ref = ora.step(ref, dt)
    t = 0.0
    for _ in range(20):
        t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    res = u.cpu().numpy()
    op.close()
    err = rel_l2(res, ref)
    assert err <= 1e-10, (err, per_field_rel(res, ref, 5))


# ---------------------------------------------------------------------------
# Navier-Stokes with the tabulated gas (BR1; viscosity and conductivity from
# the table's mu and lambda planes, LTETransport.hpp:32-65)
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def viscous_table(table):
    """The workload's table with viscosity and conductivity scaled by 1e3
    (Reynolds number ~300 instead of ~3e5), so that the viscous terms are a
    sizeable part of the right-hand side."""
    import dataclasses

    planes = table.planes.copy()
    planes[4:6] *= 1e3
    return dataclasses.replace(table, planes=planes)


def _gpu_rhs(rt, mesh, table, U):
    import torch

    from theseus_amd.operator import MakeRHSOperator

    op = MakeRHSOperator(rt, mesh, lte_table=table)
    u = torch.from_numpy(U).cuda()
    du = torch.empty_like(u)
    op.Mult(u, du)
    torch.cuda.synchronize()
    got = du.cpu().numpy()
    op.close()
    return got


@pytest.mark.parametrize("dim,n,order", [(3, 2, 3), (3, 3, 2), (3, 2, 4), (3, 5, 1), (2, 3, 3), (2, 4, 2), (2, 3, 4)])
def test_lte_navier_stokes_rhs_parity(oracle, viscous_table, dim, n, order):
    """Full RHS relative L2 <= 1e-12; the viscous part alone (Navier-Stokes
    minus Euler, on both sides) within 1e-10 of its own norm."""
    import dataclasses

    from theseus_amd import _abi as A
    from theseus_amd.cases import c4_lte_tgv

    _say(f"case lte navier-stokes rhs dim={dim} n={n} p={order}")
    rt, mesh, U, _ = c4_lte_tgv(n=n, order=order, table=viscous_table, viscous=True, dim=dim)
    rt_eu = dataclasses.replace(rt, flow_model=A.FLOW_EULER)
    ref = _oracle(oracle, rt, mesh, viscous_table).rhs(U)
    ref_eu = _oracle(oracle, rt_eu, mesh, viscous_table).rhs(U)
    got = _gpu_rhs(rt, mesh, viscous_table, U)
    got_eu = _gpu_rhs(rt_eu, mesh, viscous_table, U)
    assert np.isfinite(got).all()
    err = rel_l2(got, ref)
    assert err <= 1e-12, (err, per_field_rel(got, ref, dim + 2))
    vis_ref, vis_got = ref - ref_eu, got - got_eu
    assert np.abs(vis_ref).max() > 5e-4 * np.abs(ref).max()  # the viscous terms matter here
    verr = rel_l2(vis_got, vis_ref)
    assert verr <= 1e-10, (verr, per_field_rel(vis_got, vis_ref, dim + 2))


def test_lte_navier_stokes_stability_and_steps_parity(oracle, viscous_table):
    """The stability estimate (with the tabulated diffusivity) to 1e-12 and
    the state after 20 SSP-RK3 steps to 1e-10."""
    import torch

    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import MakeRHSOperator

    _say("case lte navier-stokes steps n=2 p=3")
    rt, mesh, U, _ = c4_lte_tgv(n=2, order=3, table=viscous_table, viscous=True)
    ora = _oracle(oracle, rt, mesh, viscous_table)
    rates = ora.stability(U)
    assert rates[2] > 0
    op = MakeRHSOperator(rt, mesh, lte_table=viscous_table)
    u = torch.from_numpy(U).cuda()
    s = op.EstimateStability(u)
    got = np.array([s.advective_rate, s.surface_rate, s.diffusive_rate])
    np.testing.assert_allclose(got, rates, rtol=1e-12, atol=0)
    dt = 0.5 / (max(rates[0], rates[1]) + rates[2])
    ref = U.copy()
    for _ in range(20):
        ref = ora.step(ref, dt)
    t = 0.0
    for _ in range(20):
        t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    res = u.cpu().numpy()
    op.close()
    err = rel_l2(res, ref)
    assert err <= 1e-10, (err, per_field_rel(res, ref, 5))


# ---------------------------------------------------------------------------
# The mixed-precision form of the tabulated gas: FP32 pair and face fluxes
# (the Itoh-Abe means of the interpolant on an FP32 copy of the table);
# temperature recovery, per-node pressure and sound speed, stage combination
# in FP64.  Errors per field against the FP64 oracle, bounds set from
# measured values with about 3x margin.
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("viscous", [False, True])
@pytest.mark.parametrize("dim,n", [(3, 3), (2, 4)])
def test_lte_mixed_rhs_and_steps_per_field(oracle, table, viscous_table, viscous, dim, n):
    """RHS per field against the oracle (density and momenta <= 1e-4; the
    energy RHS, a small difference of fluxes carrying the internal energy of
    about 3.5e6 J/kg, <= 1e-2) and the state after 20 SSP-RK3 steps (<= 5e-4
    per field; the 3-D z-momentum, zero at the start, <= 5e-3)."""
    import dataclasses

    import torch

    from theseus_amd import _abi as A
    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import MakeRHSOperator

    tab = viscous_table if viscous else table
    _say(f"case lte mixed dim={dim} n={n} p=3 viscous={viscous}")
    rt, mesh, U, _ = c4_lte_tgv(n=n, order=3, table=tab, viscous=viscous, dim=dim)
    rtm = dataclasses.replace(rt, precision=A.PRECISION_MIXED)
    ora = _oracle(oracle, rt, mesh, tab)
    ref = ora.rhs(U)
    got = _gpu_rhs(rtm, mesh, tab, U)
    nq = dim + 2
    errs = per_field_rel(got, ref, nq)
    _say(f"  RHS per-field rel L2 {errs}")
    assert max(errs[:nq - 1]) <= 1e-4 and errs[nq - 1] <= 1e-2, errs
    rates = ora.stability(U)
    dt = 0.5 / (max(rates[0], rates[1]) + rates[2])
    ref_u = U.copy()
    for _ in range(20):
        ref_u = ora.step(ref_u, dt)
    op = MakeRHSOperator(rtm, mesh, lte_table=tab)
    u = torch.from_numpy(U).cuda()
    t = 0.0
    for _ in range(20):
        t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    res = u.cpu().numpy()
    op.close()
    serr = per_field_rel(res, ref_u, nq)
    _say(f"  state after 20 steps per-field rel L2 {serr}")
    zmom = [3] if dim == 3 else []
    assert max(e for q, e in enumerate(serr) if q not in zmom) <= 5e-4, serr
    assert all(serr[q] <= 5e-3 for q in zmom), serr


@pytest.mark.parametrize("viscous", [False, True])
def test_lte_mixed_full_size_state_per_field(table, viscous):
    """The LTE Taylor-Green workload at full size (24^3, p=3), Euler and
    Navier-Stokes: the mixed state after 50 SSP-RK3 steps against the FP64
    engine's (itself inside the oracle's own rounding spread over steps):
    <= 1e-5 per field but the z-momentum (zero at the start: <= 5e-4)."""
    import dataclasses

    import torch

    from theseus_amd import _abi as A
    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import MakeRHSOperator

    _say(f"case lte mixed full size 24^3 p=3, 50 steps, viscous={viscous}")
    rt, mesh, U, _ = c4_lte_tgv(n=24, order=3, table=table, viscous=viscous)
    out = {}
    dt = None
    for name, r in (("fp64", rt), ("mixed", dataclasses.replace(rt, precision=A.PRECISION_MIXED))):
        op = MakeRHSOperator(r, mesh, lte_table=table)
        u = torch.from_numpy(U).cuda()
        if dt is None:
            dt = 0.5 / op.EstimateStability(u).TotalRate()
        t = 0.0
        for _ in range(50):
            t = op.Step(u, t, dt)
        torch.cuda.synchronize()
        out[name] = u.cpu().numpy()
        op.close()
    serr = per_field_rel(out["mixed"], out["fp64"], 5)
    _say(f"  state after 50 steps per-field rel L2 {serr}")
    assert max(serr[0], serr[1], serr[2], serr[4]) <= 1e-5, serr
    assert serr[3] <= 5e-4, serr


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lte_navier_stokes_sutherland_viscosity(oracle, table, precision):
    """The SUTHERLAND build of LTETransport (LTETransport.hpp:22-43): viscosity
    mu0 (T0 + Ts) (T/T0)^1.5 / (T + Ts) at each node's recovered temperature,
    T0 = 273.15, Ts = 110.4 (Physics.hpp:40-42), mu0 = 1.716e-2 (the
    reference's default 1.716e-5 scaled by 1e3 so that the viscous terms
    matter), the conductivity still the lambda plane (scaled alike): RHS
    <= 1e-12 in FP64 (its viscous part <= 1e-10), per field <= 1e-4 but the
    energy (<= 1e-2) in mixed precision."""
    import dataclasses

    from theseus_amd import _abi as A
    from theseus_amd.cases import c4_lte_tgv

    planes = table.planes.copy()
    planes[5] *= 1e3
    st = dataclasses.replace(table, planes=planes, sutherland=(1.716e-2, 273.15, 110.4))
    _say(f"case lte navier-stokes sutherland {precision}")
    rt, mesh, U, _ = c4_lte_tgv(n=2, order=3, table=st, viscous=True)
    rt_eu = dataclasses.replace(rt, flow_model=A.FLOW_EULER)
    if precision == "mixed":
        rt = dataclasses.replace(rt, precision=A.PRECISION_MIXED)
    ora = _oracle(oracle, dataclasses.replace(rt, precision=A.PRECISION_FP64), mesh, st)
    ref = ora.rhs(U)
    got = _gpu_rhs(rt, mesh, st, U)
    if precision == "fp64":
        err = rel_l2(got, ref)
        assert err <= 1e-12, (err, per_field_rel(got, ref, 5))
        ref_eu = _oracle(oracle, rt_eu, mesh, st).rhs(U)
        vis_ref = ref - ref_eu
        assert np.abs(vis_ref).max() > 1e-3 * np.abs(ref).max()
        vis_got = got - _gpu_rhs(rt_eu, mesh, st, U)
        verr = rel_l2(vis_got, vis_ref)
        assert verr <= 1e-10, verr
    else:
        errs = per_field_rel(got, ref, 5)
        _say(f"  RHS per-field rel L2 {errs}")
        assert max(errs[:4]) <= 1e-4 and errs[4] <= 1e-2, errs


@pytest.mark.parametrize("dim", [2, 3])
def test_lte_transmissive_boundaries(oracle, table, dim):
    """The tabulated gas on a box with transmissive (supersonic-outflow)
    boundaries on every side (bc_kernels.hpp:236-238: F*(q, q, n)): FP64 RHS
    <= 1e-12, the stability estimate to 1e-12 and 10 SSP-RK3 steps <= 1e-10;
    the mixed form per field as on periodic boxes."""
    import dataclasses

    import torch

    from theseus_amd import _abi as A
    from theseus_amd.cases import lte_taylor_green
    from theseus_amd.config import RunTime
    from theseus_amd.mesh import box_mesh
    from theseus_amd.operator import MakeRHSOperator

    _say(f"case lte transmissive boundaries dim={dim}")
    mesh = box_mesh(dim, 3, 0.0, 2 * np.pi, periodic=False, bc=A.BC_SUPERSONIC_OUTFLOW)
    U = lte_taylor_green(mesh, 3, table)
    rt = RunTime(dim=dim, order=3, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_LTEEC,
                 surface_flux=A.SURFACE_LTEEC, blending=False)
    ora = _oracle(oracle, rt, mesh, table)
    ref = ora.rhs(U)
    got = _gpu_rhs(rt, mesh, table, U)
    err = rel_l2(got, ref)
    assert err <= 1e-12, (err, per_field_rel(got, ref, dim + 2))
    mx = per_field_rel(_gpu_rhs(dataclasses.replace(rt, precision=A.PRECISION_MIXED), mesh, table, U), ref, dim + 2)
    _say(f"  mixed RHS per-field rel L2 {mx}")
    assert max(mx[:dim + 1]) <= 1e-4 and mx[dim + 1] <= 1e-2, mx
    rates = ora.stability(U)
    op = MakeRHSOperator(rt, mesh, lte_table=table)
    u = torch.from_numpy(U).cuda()
    s = op.EstimateStability(u)
    np.testing.assert_allclose([s.advective_rate, s.surface_rate, s.diffusive_rate], rates, rtol=1e-12, atol=0)
    dt = 0.5 / (max(rates[0], rates[1]) + rates[2])
    ref_u = U.copy()
    for _ in range(10):
        ref_u = ora.step(ref_u, dt)
    t = 0.0
    for _ in range(10):
        t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    res = u.cpu().numpy()
    op.close()
    serr = rel_l2(res, ref_u)
    assert serr <= 1e-10, (serr, per_field_rel(res, ref_u, dim + 2))
