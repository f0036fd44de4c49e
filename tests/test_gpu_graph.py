"""SSP-RK3 steps recorded into a HIP graph (torch.cuda.CUDAGraph) give the
same state, bit for bit, as the same steps launched one by one: the step
path launches kernels only (no allocation, memset, copy or synchronisation
on the stream).  Each replay restarts from the initial state and is checked
against the eager result, with eager steps of another state and a
synchronisation between replays."""
import dataclasses
import sys

import numpy as np
import pytest

from theseus_amd import _abi as A

from .gpu_helpers import make_case, smooth_state

pytestmark = pytest.mark.gpu

STEPS_PER_GRAPH = 4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dim,order,n,flow,blending,precision", [
    (2, 3, (6, 5), A.FLOW_EULER, True, A.PRECISION_FP64),
    (3, 3, (3, 2, 3), A.FLOW_NAVIER_STOKES, False, A.PRECISION_FP64),
    (3, 3, (2, 3, 2), A.FLOW_NAVIER_STOKES, False, A.PRECISION_MIXED),
])
def test_captured_steps_match_eager(dim, order, n, flow, blending, precision):
    import torch

    from theseus_amd.operator import MakeRHSOperator

    _say(f"case graph capture dim={dim} p={order} n={n} flow={flow} blending={blending} precision={precision}")
    rt, mesh = make_case(dim, order, n, "chandrashekar", flow=flow, blending=blending)
    rt = dataclasses.replace(rt, precision=precision)
    U0 = smooth_state(mesh, order, seed=31, amp=0.2)
    op = MakeRHSOperator(rt, mesh)
    dt = 2e-4
    u0 = torch.from_numpy(U0).cuda()

    eager = u0.clone()
    for _ in range(STEPS_PER_GRAPH):
        op.Step(eager, 0.0, dt)
    torch.cuda.synchronize()
    ref = eager.cpu().numpy()
    assert np.isfinite(ref).all() and not np.array_equal(ref, U0)

    state = u0.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        for _ in range(STEPS_PER_GRAPH):
            op.Step(state, 0.0, dt)
    torch.cuda.synchronize()

    other = torch.from_numpy(smooth_state(mesh, order, seed=32, amp=0.1)).cuda()
    for replay in range(3):
        state.copy_(u0)
        graph.replay()
        torch.cuda.synchronize()
        got = state.cpu().numpy()
        assert np.array_equal(got, ref), (replay, np.abs(got - ref).max())
        op.Step(other, 0.0, dt)
        torch.cuda.synchronize()
    assert np.isfinite(other.cpu().numpy()).all()


def test_mixed_first_evaluation_under_capture_is_rejected():
    """The mixed mode fixes its background pressure from the first state a
    handle evaluates; recorded into a graph, that computation would rerun on
    every replay, so a first evaluation under capture fails loudly and leaves
    the handle usable: eager steps afterwards, then a capture, replay as the
    eager steps do."""
    import torch

    from theseus_amd.operator import MakeRHSOperator

    _say("case graph capture: mixed first evaluation under capture")
    rt, mesh = make_case(3, 3, (2, 2, 2), "chandrashekar", flow=A.FLOW_NAVIER_STOKES, blending=False)
    rt = dataclasses.replace(rt, precision=A.PRECISION_MIXED)
    U0 = smooth_state(mesh, 3, seed=41, amp=0.2)
    op = MakeRHSOperator(rt, mesh)
    dt = 2e-4
    u0 = torch.from_numpy(U0).cuda()
    state = u0.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with pytest.raises(A.TheseusError, match="captured"):
        with torch.cuda.graph(graph, stream=side):
            op.Step(state, 0.0, dt)
    torch.cuda.synchronize()
    eager = u0.clone()
    for _ in range(2):
        op.Step(eager, 0.0, dt)
    torch.cuda.synchronize()
    ref = eager.cpu().numpy()
    state.copy_(u0)
    graph2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph2, stream=side):
        for _ in range(2):
            op.Step(state, 0.0, dt)
    torch.cuda.synchronize()
    state.copy_(u0)
    graph2.replay()
    torch.cuda.synchronize()
    assert np.array_equal(state.cpu().numpy(), ref)
    op.close()
