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
    t.attach(c)
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
