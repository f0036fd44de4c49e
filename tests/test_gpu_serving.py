"""Serving pool and continuous-batching server on the GPU: a case's result
is bitwise the same alone and in a mix of other cases (and at a different
slot offset), equals the single-case operator's, matches the FP64 oracle;
admission respects the HBM-budget capacity and reuses released slots; the
server drains a Poisson stream and reports p50/p99."""
import math
import sys

import numpy as np
import pytest

from theseus_amd import _abi as A

from .gpu_helpers import oracle_case, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run_pool_case(pool, factory, cases, target, steps):
    """Admit `cases` (list of (n, amp, phases)) in order, step, return the state of case index `target`."""
    import torch

    ids = []
    for n, amp, ph in cases:
        u0 = factory.state(pool.rt, n, amp, ph)
        cid, dt = pool.admit(factory.mesh(n), u0, 0.5)
        ids.append((cid, dt, n))
    for _ in range(steps):
        pool.step()
    torch.cuda.synchronize()
    cid, dt, n = ids[target]
    out = torch.empty(5 * n ** 3 * (pool.rt.order + 1) ** 3, dtype=torch.float64, device="cuda")
    pool.read(cid, out)
    return out.cpu().numpy(), dt


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
@pytest.mark.parametrize("order", [2, 3])
def test_case_alone_equals_case_in_mix(precision, order):
    from theseus_amd.serving import CaseFactory, Pool, euler_runtime

    _say(f"case pool alone-vs-mix p={order} precision={precision}")
    f = CaseFactory()
    rt = euler_runtime(order, precision)
    target = (4, 2.0, (0.3, 1.1, 2.0))
    pool = Pool(rt, 256 << 20, 16, tile=0, xcd_remap=1)
    alone, dt_alone = _run_pool_case(pool, f, [target], 0, 6)
    pool.close()
    pool = Pool(rt, 256 << 20, 16, tile=0, xcd_remap=1)
    mix = [(3, 4.0, (0.0, 0.5, 1.0)), (5, 0.7, (2.0, 0.1, 0.4)), target, (4, 3.3, (1.0, 1.0, 1.0))]
    mixed, dt_mix = _run_pool_case(pool, f, mix, 2, 6)
    pool.close()
    assert dt_alone == dt_mix
    assert np.isfinite(alone).all()
    assert np.array_equal(alone, mixed), np.abs(alone - mixed).max()


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
@pytest.mark.parametrize("order", [2, 3])
def test_case_sharing_workgroups_with_another_case(precision, order):
    """A case placed at an odd slot after a 27-element case shares its first
    workgroup (two elements each at p = 2 and 3) with that case: its result
    must still equal the case alone (each element's face points use its own
    case's background pressure in the mixed mode)."""
    from theseus_amd.serving import CaseFactory, Pool, euler_runtime

    _say(f"case pool straddling workgroup p={order} precision={precision}")
    f = CaseFactory()
    rt = euler_runtime(order, precision)
    target = (4, 2.0, (0.3, 1.1, 2.0))
    pool = Pool(rt, 256 << 20, 16, tile=0, xcd_remap=1)
    alone, dt_alone = _run_pool_case(pool, f, [target], 0, 6)
    pool.close()
    pool = Pool(rt, 256 << 20, 16, tile=0, xcd_remap=1)
    mixed, dt_mix = _run_pool_case(pool, f, [(3, 4.0, (0.0, 0.5, 1.0)), target, (3, 0.7, (2.0, 0.1, 0.4))], 1, 6)
    pool.close()
    assert dt_alone == dt_mix
    assert np.array_equal(alone, mixed), np.abs(alone - mixed).max()


def test_pool_matches_single_case_operator_and_oracle(oracle):
    import torch

    from theseus_amd.operator import MakeRHSOperator
    from theseus_amd.serving import CaseFactory, Pool, euler_runtime

    _say("case pool vs single-case operator vs oracle p=3")
    f = CaseFactory()
    rt = euler_runtime(3, A.PRECISION_FP64)
    n, amp, ph = 3, 1.5, (0.2, 0.4, 0.6)
    pool = Pool(rt, 128 << 20, 8, tile=0, xcd_remap=1)
    got, dt = _run_pool_case(pool, f, [(5, 1.0, (0, 0, 0)), (n, amp, ph)], 1, 10)
    pool.close()
    mesh = f.mesh(n)
    u0 = f.state(rt, n, amp, ph)
    op = MakeRHSOperator(rt, mesh)
    u = u0.clone()
    t = 0.0
    for _ in range(10):
        t = op.Step(u, t, dt)
    torch.cuda.synchronize()
    single = u.cpu().numpy()
    assert rel_l2(got, single) <= 1e-14, rel_l2(got, single)
    ora = oracle_case(oracle, rt, mesh)
    ref = u0.cpu().numpy().copy()
    for _ in range(10):
        ref = ora.step(ref, dt)
    assert rel_l2(got, ref) <= 1e-10, rel_l2(got, ref)


def test_pool_capacity_release_and_reuse():
    from theseus_amd.serving import CapacityError, CaseFactory, Pool, euler_runtime

    _say("case pool capacity")
    f = CaseFactory()
    rt = euler_runtime(2, A.PRECISION_MIXED)
    # budget for 64 element slots
    per = 3 * 5 * 27 * 8 + 8 * 10 + 4 * 7 + 8
    pool = Pool(rt, 64 * per + per // 2, 8, tile=0, xcd_remap=0)
    info = pool.info()
    assert info["capacity_elements"] == 64, info
    a, _ = pool.admit(f.mesh(3), f.state(rt, 3, 1.0, (0, 0, 0)), 0.5)  # 27 elements
    b, _ = pool.admit(f.mesh(3), f.state(rt, 3, 1.0, (1, 0, 0)), 0.5)  # 27 elements
    with pytest.raises(CapacityError):
        pool.admit(f.mesh(3), f.state(rt, 3, 1.0, (2, 0, 0)), 0.5)
    pool.release(a)
    info = pool.info()
    assert info["active_cases"] == 1 and info["free_elements"] == 64 - 27
    c, _ = pool.admit(f.mesh(3), f.state(rt, 3, 1.0, (2, 0, 0)), 0.5)
    for _ in range(2):
        pool.step()
    pool.close()
    assert c == a  # the freed case slot is reused


def test_autotuner_picks_a_tile():
    from theseus_amd.serving import Pool, euler_runtime

    _say("case pool autotune p=4")
    pool = Pool(euler_runtime(4, A.PRECISION_MIXED), 512 << 20, 8, tile=-1, xcd_remap=-1)
    info = pool.info()
    pool.close()
    assert info["tile"] in (0, 1) and info["xcd_remap"] in (0, 1), info


def test_server_drains_a_poisson_stream():
    from theseus_amd.serving import Server, make_requests

    _say("case server stream")
    reqs = make_requests(num=24, rate_per_s=400.0, n_range=(3, 6), steps=8)
    srv = Server(hbm_budget_bytes=1 << 30, max_cases=8, keep_results=True)
    res = srv.run(reqs, time_limit_s=120.0)
    srv.close()
    _say(f"  {res}")
    assert res["cases_done"] == 24
    assert math.isfinite(res["tts_p50_s"]) and res["tts_p99_s"] >= res["tts_p50_s"] > 0
    assert all(r.result is not None and np.isfinite(r.result).all() for r in reqs)


def test_async_admission_across_launch_ranges():
    """Admissions that only enqueue (no dt requested) and releases, with no
    host synchronisation between them and the steps, move the launch range
    over 2048 -> 4096 slots and back through kept graphs; the case in the
    middle ends bitwise as it does alone, where its dt came back to the host."""
    import torch

    from theseus_amd.serving import CaseFactory, Pool, euler_runtime

    _say("case pool asynchronous admission across launch ranges p=2")
    f = CaseFactory()
    rt = euler_runtime(2, A.PRECISION_FP64)
    per = 3 * 5 * 27 * 8 + 8 * 10 + 4 * 7 + 8
    target = (4, 2.0, (0.3, 1.1, 2.0))
    pool = Pool(rt, 4608 * per + per // 2, 8, tile=0, xcd_remap=1)
    assert pool.info()["capacity_elements"] == 4608
    alone, _ = _run_pool_case(pool, f, [target], 0, 6)
    pool.close()

    pool = Pool(rt, 4608 * per + per // 2, 8, tile=0, xcd_remap=1)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        big = f.state(rt, 12, 1.0, (0.0, 0.0, 0.0))   # 1728 elements: range 2048
        tgt = f.state(rt, *target)                    # 64 elements after it
        big2 = f.state(rt, 13, 0.8, (1.0, 0.0, 0.0))  # 2197 elements
        a, dt_a = pool.admit(f.mesh(12), big, 0.5, s, want_dt=False)
        b, _ = pool.admit(f.mesh(4), tgt, 0.5, s, want_dt=False)
        assert dt_a is None
        for _ in range(3):
            pool.step(s)
        pool.release(a, s)                            # range stays 2048 (the target ends at 1792)
        c, _ = pool.admit(f.mesh(13), big2, 0.5, s, want_dt=False)  # after the target: range 4096
        for _ in range(2):
            pool.step(s)
        pool.release(c, s)                            # back to 2048, its graph reused
        pool.step(s)
    s.synchronize()
    out = torch.empty(5 * 4 ** 3 * 27, dtype=torch.float64, device="cuda")
    pool.read(b, out, s)
    mixed = out.cpu().numpy()
    info = pool.info()
    pool.close()
    assert info["active_cases"] == 1 and info["free_elements"] == 4608 - 64, info
    assert np.array_equal(alone, mixed), np.abs(alone - mixed).max()


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
def test_group_launch_equals_per_order_pools(precision):
    """Pools of orders 2, 3 and 4 stepped by one launch per stage end
    bitwise where the same pools stepped one launch each end (p = 2 at four
    elements per workgroup in both)."""
    import torch

    from theseus_amd.serving import CaseFactory, Pool, PoolGroup, euler_runtime

    _say(f"pool group vs per-order pools precision={precision}")
    f = CaseFactory()
    cases = {2: [(4, 2.0, (0.3, 1.1, 2.0)), (3, 0.7, (2.0, 0.1, 0.4))],
             3: [(3, 1.5, (0.2, 0.4, 0.6)), (4, 3.3, (1.0, 1.0, 1.0))],
             4: [(2, 1.0, (0.5, 0.0, 0.1)), (3, 2.5, (0.0, 0.9, 0.3))]}
    tile = {2: 1, 3: 0, 4: 0}

    def run(grouped):
        pools = {p: Pool(euler_runtime(p, precision), 256 << 20, 8, tile=tile[p], xcd_remap=1) for p in cases}
        ids = {p: [pools[p].admit(f.mesh(n), f.state(pools[p].rt, n, a, ph), 0.5)[0] for n, a, ph in cases[p]]
               for p in cases}
        group = PoolGroup([pools[p] for p in cases]) if grouped else None
        for _ in range(5):
            if grouped:
                group.step()
            else:
                for p in cases:
                    pools[p].step()
        torch.cuda.synchronize()
        out = {}
        for p in cases:
            for (n, _, _), cid in zip(cases[p], ids[p]):
                buf = torch.empty(5 * n ** 3 * (p + 1) ** 3, dtype=torch.float64, device="cuda")
                pools[p].read(cid, buf)
                out[(p, cid)] = buf.cpu().numpy()
        if group is not None:
            group.close()
        for pool in pools.values():
            pool.close()
        return out

    alone = run(False)
    grouped = run(True)
    assert alone.keys() == grouped.keys()
    for key in alone:
        assert np.isfinite(alone[key]).all()
        assert np.array_equal(alone[key], grouped[key]), (key, np.abs(alone[key] - grouped[key]).max())


def test_group_rejects_mismatched_pools():
    from theseus_amd.serving import Pool, PoolGroup, euler_runtime

    _say("pool group argument checks")
    a = Pool(euler_runtime(2, A.PRECISION_MIXED), 64 << 20, 4, tile=0, xcd_remap=1)
    b = Pool(euler_runtime(2, A.PRECISION_MIXED), 64 << 20, 4, tile=0, xcd_remap=1)
    c = Pool(euler_runtime(3, A.PRECISION_FP64), 64 << 20, 4, tile=0, xcd_remap=1)
    d = Pool(euler_runtime(5, A.PRECISION_MIXED), 64 << 20, 4, tile=0, xcd_remap=1)
    try:
        for pools, msg in (([a, b], "one pool per order"), ([a, c], "precision"), ([d], "order 2, 3 or 4")):
            with pytest.raises(A.TheseusError, match=msg):
                PoolGroup(pools)
    finally:
        for p in (a, b, c, d):
            p.close()


def test_server_single_launch_drains_a_poisson_stream():
    from theseus_amd.serving import Server, make_requests

    _say("case server stream, one launch for all orders")
    reqs = make_requests(num=24, rate_per_s=400.0, n_range=(3, 6), steps=8)
    srv = Server(hbm_budget_bytes=1 << 30, max_cases=8, keep_results=True, single_launch=True)
    res = srv.run(reqs, time_limit_s=120.0)
    srv.close()
    _say(f"  {res}")
    assert res["cases_done"] == 24
    assert all(r.result is not None and np.isfinite(r.result).all() for r in reqs)
