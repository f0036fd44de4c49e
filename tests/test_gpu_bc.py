"""Boundary conditions on the GPU against the CPU FP64 oracle
(bc_kernels.hpp:32-375, Flow.hpp:55-170): every boundary type on the sides
of 2-D and 3-D boxes, Euler and Navier-Stokes, several flux pairings, with
and without blending.  Bars: FP64 right-hand side relative L2 <= 1e-12,
state after 100 SSP-RK3 steps <= 1e-10, stability rates to 1e-13; mixed
precision <= 1e-5 per field; a pooled case with its own boundary table
equals the single-case operator."""
import dataclasses
import sys

import numpy as np
import pytest

from theseus_amd import _abi as A
from theseus_amd.config import RunTime
from theseus_amd.mesh import box_mesh

from .gpu_helpers import PAIRINGS, gpu_rhs, oracle_case, per_field_rel, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

G = 1.4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _state(rho, u, p):
    u = np.asarray(u, float)
    return tuple(np.r_[rho, rho * u, p / (G - 1) + 0.5 * rho * u @ u])


def _sides(dim, kind):
    """Boundary sets: 'euler' mixes slip wall, inflow, outflow, prescribed
    state and symmetry; 'walls' puts a moving adiabatic wall, an isothermal
    wall and slip/symmetry/prescribed/outflow sides around a NS box."""
    z = [0.0] * dim
    if kind == "euler":
        s = [(A.BC_SUPERSONIC_INFLOW, _state(1.1, [0.4] + z[1:], 0.8)), A.BC_SUPERSONIC_OUTFLOW,
             A.BC_SLIP_WALL, A.BC_SYMMETRY, (A.BC_PRESCRIBED_STATE, _state(0.9, [0.3] + [0.05] * (dim - 1), 0.7)),
             A.BC_SLIP_WALL]
    elif kind == "walls":
        vw = [0.1] + z[1:]
        s = [(A.BC_PRESCRIBED_STATE, _state(1.0, [0.3] + z[1:], 1.0 / G)), A.BC_SUPERSONIC_OUTFLOW,
             (A.BC_NO_SLIP_ADIABATIC, tuple(vw) + (0.02,)), (A.BC_NO_SLIP_ISOTHERMAL, tuple(z) + (1.0 / (G * 287.05),)),
             A.BC_SYMMETRY, A.BC_SLIP_WALL]
    else:  # Euler with no-slip sides: the reference's inviscid switch gives a zero flux there
        s = [(A.BC_NO_SLIP_ADIABATIC, tuple(z) + (0.0,)), A.BC_SLIP_WALL, A.BC_SLIP_WALL,
             (A.BC_NO_SLIP_ISOTHERMAL, tuple(z) + (300.0,)), A.BC_SYMMETRY, A.BC_SYMMETRY]
    return s[: 2 * dim]


def _case(dim, order, n, pairing, flow, kind, blending=False, periodic=False, mu=0.02):
    vol, surf = PAIRINGS[pairing]
    mesh = box_mesh(dim, n, 0.0, 1.0, periodic=periodic, side_bc=_sides(dim, kind))
    rt = RunTime(dim=dim, order=order, flow_model=flow, volume_flux=vol, surface_flux=surf, blending=blending, mu=mu)
    return rt, mesh


def _parity(oracle, rt, mesh, U, tol=1e-12):
    ref = oracle_case(oracle, rt, mesh).rhs(U)
    got, _ = gpu_rhs(rt, mesh, U)
    assert np.isfinite(got).all()
    err = rel_l2(got, ref)
    assert err <= tol, (err, per_field_rel(got, ref, rt.dim + 2))


@pytest.mark.parametrize("pairing", ["ranocha_llf", "chandrashekar", "roe", "hll", "pirozzoli"])
@pytest.mark.parametrize("dim,order,n", [(2, 3, (4, 3)), (3, 2, (3, 2, 2)), (3, 3, (2, 3, 2))])
def test_euler_boundaries_rhs_parity(oracle, pairing, dim, order, n):
    _say(f"case bc euler {pairing} dim={dim} p={order} n={n}")
    rt, mesh = _case(dim, order, n, pairing, A.FLOW_EULER, "euler")
    _parity(oracle, rt, mesh, smooth_state(mesh, order, seed=31 + dim, amp=0.2))


@pytest.mark.parametrize("dim,order,n", [(2, 4, (5, 4)), (3, 3, (3, 2, 3))])
def test_euler_boundaries_with_blending(oracle, dim, order, n):
    _say(f"case bc euler blending dim={dim} p={order}")
    rt, mesh = _case(dim, order, n, "hll", A.FLOW_EULER, "euler", blending=True)
    U = smooth_state(mesh, order, seed=41, amp=0.4)
    _parity(oracle, rt, mesh, U)


@pytest.mark.parametrize("dim", [2, 3])
def test_euler_no_slip_sides_give_zero_flux(oracle, dim):
    _say(f"case bc euler no-slip dim={dim}")
    rt, mesh = _case(dim, 3, (3,) * dim, "ranocha_llf", A.FLOW_EULER, "euler_noslip")
    _parity(oracle, rt, mesh, smooth_state(mesh, 3, seed=43))


@pytest.mark.parametrize("pairing", ["ranocha_llf", "chandrashekar_roe"])
@pytest.mark.parametrize("dim,order,n", [(2, 3, (4, 3)), (3, 3, (2, 3, 2)), (3, 2, (3, 3, 2))])
def test_navier_stokes_boundaries_rhs_parity(oracle, pairing, dim, order, n):
    _say(f"case bc ns {pairing} dim={dim} p={order} n={n}")
    rt, mesh = _case(dim, order, n, pairing, A.FLOW_NAVIER_STOKES, "walls", mu=0.05)
    _parity(oracle, rt, mesh, smooth_state(mesh, order, seed=51 + dim, amp=0.2))


def test_navier_stokes_walls_with_blending(oracle):
    _say("case bc ns blending 2d p=4")
    rt, mesh = _case(2, 4, (4, 4), "ranocha_llf", A.FLOW_NAVIER_STOKES, "walls", blending=True, mu=0.01)
    _parity(oracle, rt, mesh, smooth_state(mesh, 4, seed=53, amp=0.4))


@pytest.mark.parametrize("flow,kind", [(A.FLOW_EULER, "euler"), (A.FLOW_NAVIER_STOKES, "walls")])
def test_mixed_precision_with_boundaries(oracle, flow, kind):
    _say(f"case bc mixed flow={flow}")
    rt, mesh = _case(3, 3, (2, 3, 2), "ranocha_llf", flow, kind, mu=0.05)
    U = smooth_state(mesh, 3, seed=61)
    ref = oracle_case(oracle, rt, mesh).rhs(U)
    got, _ = gpu_rhs(dataclasses.replace(rt, precision=A.PRECISION_MIXED), mesh, U)
    errs = per_field_rel(got, ref, 5)
    _say(f"  per-field rel L2 {errs}")
    assert max(errs) <= 1e-5, errs


@pytest.mark.parametrize("flow,kind", [(A.FLOW_EULER, "euler"), (A.FLOW_NAVIER_STOKES, "walls")])
def test_stability_with_boundaries(oracle, flow, kind):
    import torch

    from theseus_amd.operator import MakeRHSOperator

    _say(f"case bc stability flow={flow}")
    rt, mesh = _case(3, 3, (2, 2, 3), "ranocha_llf", flow, kind)
    U = smooth_state(mesh, 3, seed=63)
    ref = oracle_case(oracle, rt, mesh).stability(U)
    op = MakeRHSOperator(rt, mesh)
    s = op.EstimateStability(torch.from_numpy(U).cuda())
    op.close()
    got = np.array([s.advective_rate, s.surface_rate, s.diffusive_rate])
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("flow,kind,pairing,dim,order,n", [
    (A.FLOW_EULER, "euler", "chandrashekar_roe", 2, 3, (4, 4)),
    (A.FLOW_NAVIER_STOKES, "walls", "ranocha_llf", 3, 2, (2, 3, 2)),
])
def test_ssprk3_100_steps_with_boundaries(oracle, flow, kind, pairing, dim, order, n):
    import torch

    from theseus_amd.operator import MakeRHSOperator

    _say(f"case bc rk3-100 flow={flow} dim={dim} p={order}")
    rt, mesh = _case(dim, order, n, pairing, flow, kind, blending=True)
    U = smooth_state(mesh, order, seed=65)
    ora = oracle_case(oracle, rt, mesh)
    rates = ora.stability(U)
    dt = 0.3 / (max(rates[0], rates[1]) + rates[2])
    ref = U.copy()
    for _ in range(100):
        ref = ora.step(ref, dt)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U).cuda()
    t = 0.0
    for _ in range(100):
        t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    op.close()
    got = u.cpu().numpy()
    assert np.isfinite(got).all()
    err = rel_l2(got, ref)
    assert err <= 1e-10, (err, per_field_rel(got, ref, dim + 2))


def test_pooled_cases_keep_their_own_boundary_tables():
    """Two cases with different boundary tables in one pool, each equal to
    its single-case operator (the pool remaps boundary indices per case)."""
    import torch

    from theseus_amd.operator import MakeRHSOperator
    from theseus_amd.serving import Pool

    _say("case bc pool two boundary tables")
    rt, m1 = _case(3, 3, (2, 2, 2), "ranocha_llf", A.FLOW_EULER, "euler")
    _, m2 = _case(3, 3, (3, 2, 2), "ranocha_llf", A.FLOW_EULER, "euler_noslip")
    u1 = torch.from_numpy(smooth_state(m1, 3, seed=71)).cuda()
    u2 = torch.from_numpy(smooth_state(m2, 3, seed=72)).cuda()
    pool = Pool(rt, 64 << 20, 4, tile=0, xcd_remap=1)
    c1, dt1 = pool.admit(m1, u1, 0.3)
    c2, dt2 = pool.admit(m2, u2, 0.3)
    for _ in range(5):
        pool.step()
    got = []
    for cid, u in ((c1, u1), (c2, u2)):
        out = torch.empty_like(u)
        pool.read(cid, out)
        got.append(out)
    torch.cuda.synchronize()
    pool.close()
    for (m, u0, dt, g) in ((m1, u1, dt1, got[0]), (m2, u2, dt2, got[1])):
        op = MakeRHSOperator(rt, m)
        u = u0.clone()
        t = 0.0
        for _ in range(5):
            t = op.Step(u, t, dt)
        torch.cuda.synchronize()
        op.close()
        err = rel_l2(g.cpu().numpy(), u.cpu().numpy())
        assert err <= 1e-14, err


@pytest.mark.parametrize("dim,order,n,flow,kind", [
    (2, 1, (5, 4), A.FLOW_NAVIER_STOKES, "walls"),
    (2, 5, (3, 2), A.FLOW_EULER, "euler"),
    (3, 1, (3, 2, 3), A.FLOW_EULER, "euler"),
    (3, 5, (2, 1, 2), A.FLOW_NAVIER_STOKES, "walls"),
    (3, 4, (2, 2, 1), A.FLOW_NAVIER_STOKES, "walls"),
])
def test_boundaries_at_extreme_orders(oracle, dim, order, n, flow, kind):
    """Orders 1 and 5 (and 4): boundary faces on partially filled workgroups
    (p = 1 packs many elements per workgroup) and on the largest elements."""
    _say(f"case bc extreme order dim={dim} p={order} n={n} flow={flow}")
    rt, mesh = _case(dim, order, n, "ranocha_llf", flow, kind, mu=0.05)
    _parity(oracle, rt, mesh, smooth_state(mesh, order, seed=81 + order, amp=0.2))


def test_pooled_navier_stokes_cases_with_walls():
    """Navier-Stokes cases with walls in one pool (gradient pass and second
    pass of the boundary variant), each equal to its single-case operator."""
    import torch

    from theseus_amd.operator import MakeRHSOperator
    from theseus_amd.serving import Pool

    _say("case bc pool navier-stokes walls")
    rt, m1 = _case(3, 3, (2, 2, 2), "ranocha_llf", A.FLOW_NAVIER_STOKES, "walls", mu=0.05)
    _, m2 = _case(3, 3, (2, 3, 2), "ranocha_llf", A.FLOW_NAVIER_STOKES, "walls", mu=0.05)
    us = [torch.from_numpy(smooth_state(m, 3, seed=s)).cuda() for m, s in ((m1, 91), (m2, 92))]
    pool = Pool(rt, 64 << 20, 4, tile=0, xcd_remap=1)
    ids = [pool.admit(m, u, 0.3) for m, u in ((m1, us[0]), (m2, us[1]))]
    for _ in range(4):
        pool.step()
    got = []
    for (cid, _), u in zip(ids, us):
        out = torch.empty_like(u)
        pool.read(cid, out)
        got.append(out)
    torch.cuda.synchronize()
    pool.close()
    for m, u0, (_, dt), g in zip((m1, m2), us, ids, got):
        op = MakeRHSOperator(rt, m)
        u = u0.clone()
        t = 0.0
        for _ in range(4):
            t = op.Step(u, t, dt)
        torch.cuda.synchronize()
        op.close()
        err = rel_l2(g.cpu().numpy(), u.cpu().numpy())
        assert err <= 1e-14, err
