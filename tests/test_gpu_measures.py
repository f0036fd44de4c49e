"""GPU integral measures and the device-resident variable-step clock against
the CPU FP64 oracle: measures of Euler, Navier-Stokes, blended and LTE
states (sums to 1e-13 relative, extremes to 1e-14), the baseline kept from
the first call, per-case measures in the serving pool equal to the
single-case operator's bit for bit; variable SSP-RK3 steps with dt chosen on
the device equal the oracle's host loop (state <= 1e-10, t and dt <= 1e-12),
land on t_final and then leave the state untouched, honour the growth
limit, and replay identically from a captured graph."""
import sys

import numpy as np
import pytest

from theseus_amd import _abi as A

from .gpu_helpers import make_case, oracle_case, per_field_rel, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

SUMS = ("mass", "ke", "en")
EXTREMES = ("max_press", "min_press", "max_temp", "min_temp", "max_dens", "min_dens")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check_measures(got, ref, rel_sum=1e-13, rel_ext=1e-14):
    for k in SUMS:
        assert getattr(got, k) == pytest.approx(ref[k], rel=rel_sum, abs=1e-300), (k, getattr(got, k), ref[k])
    for k in EXTREMES:
        assert getattr(got, k) == pytest.approx(ref[k], rel=rel_ext, abs=1e-300), (k, getattr(got, k), ref[k])


@pytest.mark.parametrize("dim,order,n,pairing,flow,blending", [
    (2, 3, (5, 4), "ranocha_llf", A.FLOW_EULER, False),
    (2, 4, (4, 4), "chandrashekar_roe", A.FLOW_EULER, True),
    (3, 3, (3, 3, 3), "ranocha_llf", A.FLOW_NAVIER_STOKES, False),
    (3, 2, (4, 3, 2), "pirozzoli", A.FLOW_EULER, True),
])
def test_integral_measures_parity(oracle, dim, order, n, pairing, flow, blending):
    import torch

    from theseus_amd.operator import MakeRHSOperator

    _say(f"case measures dim={dim} p={order} n={n} {pairing} flow={flow}")
    rt, mesh = make_case(dim, order, n, pairing, flow=flow, blending=blending)
    U = smooth_state(mesh, order, seed=11, amp=0.3)
    ref = oracle_case(oracle, rt, mesh).integral_measures(U)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U).cuda()
    base0 = op.GetIntegralMeasuresBaseline()
    assert base0.mass == 0.0
    got = op.ComputeIntegralMeasures(u)
    _check_measures(got, ref)
    # the first call is the baseline; later calls do not move it
    op.Step(u, 0.0, 1e-3)
    later = op.ComputeIntegralMeasures(u)
    assert later.ke != got.ke
    assert op.GetIntegralMeasuresBaseline() == got
    op.close()


def test_integral_measures_lte(oracle):
    import torch

    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.lte import synthetic_table
    from theseus_amd.operator import MakeRHSOperator

    _say("case measures lte n=2 p=3")
    table = synthetic_table()
    rt, mesh, U, _ = c4_lte_tgv(n=2, order=3, table=table)
    c = oracle_case(oracle, rt, mesh)
    t = oracle.LteTable(arrays=(table.rho, table.T, table.e, table.planes.reshape(-1), table.inv.reshape(-1)))
    t.attach(c)
    ref = c.integral_measures(U)
    op = MakeRHSOperator(rt, mesh, lte_table=table)
    got = op.ComputeIntegralMeasures(torch.from_numpy(U).cuda())
    op.close()
    _check_measures(got, ref, rel_ext=1e-13)


def test_pool_case_measures_equal_single_case():
    import torch

    from theseus_amd.operator import MakeRHSOperator
    from theseus_amd.serving import CaseFactory, Pool, euler_runtime

    _say("case pool measures p=3")
    f = CaseFactory()
    rt = euler_runtime(3, A.PRECISION_FP64)
    pool = Pool(rt, 256 << 20, 8, tile=0, xcd_remap=1)
    ids = []
    for n, amp, ph in [(3, 1.0, (0.0, 0.2, 0.4)), (4, 2.5, (1.0, 0.3, 0.0))]:
        cid, _ = pool.admit(f.mesh(n), f.state(rt, n, amp, ph), 0.5)
        ids.append((cid, n, amp, ph))
    for cid, n, amp, ph in ids:
        m = pool.integral_measures(cid)
        op = MakeRHSOperator(rt, f.mesh(n))
        u0 = f.state(rt, n, amp, ph)
        u = u0 if isinstance(u0, torch.Tensor) else torch.from_numpy(u0).cuda()
        ref = op.ComputeIntegralMeasures(u)
        op.close()
        assert m == {k: getattr(ref, k) for k in m}
    pool.close()


def _tgv2d(oracle, n=4, order=3):
    rt, mesh = make_case(2, order, (n, n), "ranocha_llf", blending=False)
    U = smooth_state(mesh, order, seed=5, amp=0.25)
    return rt, mesh, U


def test_variable_steps_match_the_oracle_loop(oracle):
    import torch

    from theseus_amd.operator import DeviceClock, MakeRHSOperator

    _say("case clock: variable steps vs oracle loop")
    rt, mesh, U = _tgv2d(oracle)
    c = oracle_case(oracle, rt, mesh)
    cfl, growth, t_final = 0.5, 1.2, 1e9
    nsteps = 25
    Uref, tref, dtref, dtr_ref, sref = c.adaptive_steps(U, cfl, growth, t_final, nsteps)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U).cuda()
    clk = DeviceClock(cfl, growth, t_final)
    for _ in range(nsteps):
        op.StepCFL(u, clk)
    torch.cuda.synchronize()
    st = clk.read()
    got = u.cpu().numpy()
    op.close()
    assert st["steps"] == sref == nsteps
    assert st["t"] == pytest.approx(tref, rel=1e-12)
    assert st["dt_real"] == pytest.approx(dtr_ref, rel=1e-12)
    err = rel_l2(got, Uref)
    assert err <= 1e-10, (err, per_field_rel(got, Uref, 4))


def test_clock_lands_on_t_final_then_holds(oracle):
    import torch

    from theseus_amd.operator import DeviceClock, MakeRHSOperator

    _say("case clock: t_final landing")
    rt, mesh, U = _tgv2d(oracle)
    c = oracle_case(oracle, rt, mesh)
    r = c.stability(U)
    dt0 = 0.5 / (max(r[0], r[1]) + r[2])
    t_final = 3.6 * dt0
    Uref, tref, _, dtr_ref, sref = c.adaptive_steps(U, 0.5, 1.2, t_final, 100)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U).cuda()
    clk = DeviceClock(0.5, 1.2, t_final)
    for _ in range(sref):
        op.StepCFL(u, clk)
    torch.cuda.synchronize()
    st = clk.read()
    at_end = u.cpu().numpy()
    assert st["steps"] == sref
    assert st["t"] == pytest.approx(t_final, rel=1e-15)
    assert st["dt_real"] == pytest.approx(dtr_ref, rel=1e-12)
    assert rel_l2(at_end, Uref) <= 1e-10
    # further calls take no step and leave u bit for bit as it is
    for _ in range(3):
        op.StepCFL(u, clk)
    torch.cuda.synchronize()
    st2 = clk.read()
    op.close()
    assert st2["steps"] == sref and st2["dt_real"] == 0.0 and st2["t"] == st["t"]
    assert np.array_equal(u.cpu().numpy(), at_end)


def test_clock_growth_limit():
    """A previous step of 1e-9 caps the next at max_dt_growth * 1e-9."""
    import torch

    from theseus_amd.operator import DeviceClock, MakeRHSOperator

    _say("case clock: growth limit")
    rt, mesh = make_case(2, 3, (3, 3), "ranocha_llf", blending=False)
    U = smooth_state(mesh, 3, seed=2)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U).cuda()
    clk = DeviceClock(0.5, 1.5)
    # seed the previous nominal step
    raw = clk.buf.cpu().numpy().copy()
    c = A.Clock.from_buffer_copy(raw.tobytes())
    c.dt = 1e-9
    clk.buf.copy_(torch.from_numpy(np.frombuffer(bytes(c), dtype=np.uint8).copy()))
    op.StepCFL(u, clk)
    op.StepCFL(u, clk)
    torch.cuda.synchronize()
    st = clk.read()
    op.close()
    assert st["dt"] == pytest.approx(1e-9 * 1.5 * 1.5, rel=1e-15)
    assert st["t"] == pytest.approx(1e-9 * 1.5 + 1e-9 * 1.5 * 1.5, rel=1e-15)
    assert st["steps"] == 2


def test_clock_steps_replay_from_a_graph():
    """Three variable steps captured once and replayed: every replay equals
    the same steps run eagerly (bitwise), with eager work between replays."""
    import torch

    from theseus_amd.operator import DeviceClock, MakeRHSOperator

    _say("case clock: graph replay")
    rt, mesh = make_case(2, 3, (4, 4), "ranocha_llf", blending=False)
    U = smooth_state(mesh, 3, seed=9, amp=0.2)
    op = MakeRHSOperator(rt, mesh)
    # eager reference: 9 steps, snapshots after 3, 6, 9
    ue = torch.from_numpy(U).cuda()
    ce = DeviceClock(0.5, 1.2)
    snaps = []
    for k in range(9):
        op.StepCFL(ue, ce)
        if k % 3 == 2:
            torch.cuda.synchronize()
            snaps.append((ue.clone(), ce.read()))
    ug = torch.from_numpy(U).cuda()
    cg = DeviceClock(0.5, 1.2)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for _ in range(3):
            op.StepCFL(ug, cg)
    torch.cuda.synchronize()
    for r in range(3):
        g.replay()
        torch.cuda.synchronize()
        scratch = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        assert scratch == 1000.0
        ref_u, ref_c = snaps[r]
        assert torch.equal(ug, ref_u), (r, (ug - ref_u).abs().max().item())
        assert cg.read() == ref_c
    op.close()
