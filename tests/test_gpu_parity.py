"""GPU parity: the HIP engine (through the C-ABI) against the CPU FP64 oracle
on identical meshes and states.  Bar: RHS relative L2 <= 1e-12, state after
100 SSP-RK3 steps <= 1e-10 (FP64)."""
import sys

import numpy as np
import pytest

from theseus_amd import _abi as A

from .gpu_helpers import PAIRINGS, gpu_rhs, make_case, oracle_case, per_field_rel, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("pairing", sorted(PAIRINGS))
@pytest.mark.parametrize("dim,order,n", [(2, 3, (4, 3)), (3, 2, (3, 2, 2)), (3, 3, (2, 3, 2))])
def test_euler_rhs_parity(oracle, pairing, dim, order, n):
    _say(f"case euler {pairing} dim={dim} p={order} n={n}")
    rt, mesh = make_case(dim, order, n, pairing, blending=False)
    U = smooth_state(mesh, order, rt.gamma, seed=dim * 10 + order)
    ref = oracle_case(oracle, rt, mesh).rhs(U)
    got, _ = gpu_rhs(rt, mesh, U)
    err = rel_l2(got, ref)
    assert err <= RHS_TOL, (err, per_field_rel(got, ref, dim + 2))


@pytest.mark.parametrize("pairing", ["ranocha_llf", "chandrashekar", "llf"])
@pytest.mark.parametrize("dim,order,n", [(2, 2, (3, 4)), (3, 3, (2, 2, 3))])
def test_ns_rhs_parity(oracle, pairing, dim, order, n):
    _say(f"case ns {pairing} dim={dim} p={order} n={n}")
    rt, mesh = make_case(dim, order, n, pairing, flow=A.FLOW_NAVIER_STOKES, blending=False, mu=0.05, Pr=0.71)
    U = smooth_state(mesh, order, rt.gamma, seed=7)
    ref = oracle_case(oracle, rt, mesh).rhs(U)
    got, _ = gpu_rhs(rt, mesh, U)
    err = rel_l2(got, ref)
    assert err <= RHS_TOL, (err, per_field_rel(got, ref, dim + 2))


@pytest.mark.parametrize("dim,order,n,pairing", [(2, 4, (6, 5), "hll"), (2, 3, (5, 5), "chandrashekar"),
                                                 (3, 2, (3, 3, 2), "ranocha_llf")])
def test_blending_parity(oracle, dim, order, n, pairing):
    """Persson-Peraire indicator + subcell FV blend on a discontinuous state,
    with transmissive boundaries in 2-D."""
    _say(f"case blending {pairing} dim={dim} p={order} n={n}")
    rt, mesh = make_case(dim, order, n, pairing, blending=True, periodic=(dim == 3))
    from theseus_amd.cases import four_quadrant_riemann

    if dim == 2:
        U = four_quadrant_riemann(mesh, order, rt.gamma, split=(0.43, 0.57))
    else:
        U = smooth_state(mesh, order, rt.gamma, amp=0.4, seed=3)
        U = U.reshape(dim + 2, -1)
        U[:, : U.shape[1] // 3] *= 1.6  # a jump inside the first elements
        U = U.reshape(-1)
    ora = oracle_case(oracle, rt, mesh)
    ref = ora.rhs(U)
    got, alpha = gpu_rhs(rt, mesh, U)
    assert np.array_equal(alpha > 0, ora.alpha > 0), "blending switched on in different elements"
    assert (ora.alpha > 0).any(), "test state must trigger blending"
    np.testing.assert_allclose(alpha, ora.alpha, rtol=1e-13, atol=1e-15)
    err = rel_l2(got, ref)
    assert err <= RHS_TOL, (err, per_field_rel(got, ref, dim + 2))


def test_transmissive_boundary_parity(oracle):
    _say("case transmissive roe 2d p=3")
    rt, mesh = make_case(2, 3, (4, 3), "roe", blending=False, periodic=False)
    U = smooth_state(mesh, 3, rt.gamma, seed=5)
    ref = oracle_case(oracle, rt, mesh).rhs(U)
    got, _ = gpu_rhs(rt, mesh, U)
    assert rel_l2(got, ref) <= RHS_TOL


@pytest.mark.parametrize("flow", [A.FLOW_EULER, A.FLOW_NAVIER_STOKES])
def test_stability_parity(oracle, flow):
    import torch

    from theseus_amd.operator import MakeRHSOperator

    _say(f"case stability flow={flow}")
    rt, mesh = make_case(3, 3, (3, 2, 2), "ranocha_llf", flow=flow, blending=False)
    U = smooth_state(mesh, 3, rt.gamma, seed=11)
    ref = oracle_case(oracle, rt, mesh).stability(U)
    op = MakeRHSOperator(rt, mesh)
    s = op.EstimateStability(torch.from_numpy(U).cuda())
    got = np.array([s.advective_rate, s.surface_rate, s.diffusive_rate])
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("flow,pairing,dim,order,n", [
    (A.FLOW_EULER, "chandrashekar_roe", 2, 3, (4, 4)),
    (A.FLOW_NAVIER_STOKES, "ranocha_llf", 3, 3, (2, 2, 2)),
])
def test_ssprk3_100_steps_parity(oracle, flow, pairing, dim, order, n):
    """State after 100 SSP-RK3 steps (fixed dt from the stability estimate) <= 1e-10."""
    import torch

    from theseus_amd.operator import MakeRHSOperator

    _say(f"case rk3-100 {pairing} dim={dim} p={order}")
    rt, mesh = make_case(dim, order, n, pairing, flow=flow, blending=True)
    U = smooth_state(mesh, order, rt.gamma, seed=2)
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
    got = u.cpu().numpy()
    err = rel_l2(got, ref)
    assert err <= 1e-10, (err, per_field_rel(got, ref, dim + 2))


def test_engine_reproduces_reference_golden_vortex():
    """The HIP engine on the reference's golden isentropic-vortex case
    (500 SSP-RK3 steps, Chandrashekar, blending on) against the reference's
    own output fields (tests/golden/isentropic_vortex_cycle500.npz)."""
    import torch

    from theseus_amd.operator import MakeRHSOperator

    from .golden_vortex import DT, NSTEPS, golden_permutation, primitive_fields, vortex_case

    _say("case golden vortex 500 steps")
    rt, mesh, U0 = vortex_case()
    d, perm = golden_permutation(mesh)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U0).cuda()
    t = 0.0
    for _ in range(NSTEPS):
        t = op.Step(u, t, DT)
    torch.cuda.synchronize()
    f = primitive_fields(u.cpu().numpy())
    for k in ("rho", "u", "v", "p"):
        g = d[k][perm]
        rel = np.linalg.norm(f[k] - g) / np.linalg.norm(g)
        assert rel < 1e-12, (k, rel)
    alpha = op.Alpha().cpu().numpy()
    np.testing.assert_array_equal(alpha, d["alpha"][perm].reshape(-1, 16)[:, 0])
