"""theseus_rhs_integral_measures_noalloc: the integral measures computed
without any allocation (workgroup partials in the library's fixed device
array), written to a caller's device buffer.  Checked through the C-ABI
against theseus_rhs_integral_measures on the same states (same kernels, same
fixed reduction order: bit-identical), on two calls in a row whose second
input is smaller, and for the size limit's error."""
import ctypes as C
import sys

import numpy as np
import pytest

from theseus_amd import _abi as A

from .gpu_helpers import make_case, smooth_state

pytestmark = pytest.mark.gpu

FIELDS = ("mass", "ke", "en", "max_press", "min_press", "max_temp", "min_temp", "max_dens", "min_dens")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _measures_both(op, U):
    import torch

    u = torch.from_numpy(U).cuda()
    out = torch.full((9,), float("nan"), dtype=torch.float64, device="cuda")
    lib = A.load()
    rc = lib.theseus_rhs_integral_measures_noalloc(op._h, C.c_void_p(u.data_ptr()), C.c_void_p(out.data_ptr()),
                                                   C.c_void_p(0))
    assert rc == 0, lib.theseus_last_error()
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    m = A.IntegralMeasures()
    rc = lib.theseus_rhs_integral_measures(op._h, C.c_void_p(u.data_ptr()), C.byref(m), C.c_void_p(0))
    assert rc == 0, lib.theseus_last_error()
    ref = np.array([getattr(m, f) for f in FIELDS])
    return got, ref


def test_noalloc_measures_equal_the_allocating_entry_point_on_two_calls():
    from theseus_amd.operator import MakeRHSOperator

    # first a larger 3-D Navier-Stokes case (4,096 elements), then a smaller
    # 2-D Euler case: the second call reuses the array the first one filled
    cases = [(3, 3, (16, 16, 16), "ranocha_llf", A.FLOW_NAVIER_STOKES), (2, 4, (5, 3), "hll", A.FLOW_EULER)]
    for dim, order, n, pairing, flow in cases:
        _say(f"case noalloc measures dim={dim} p={order} n={n}")
        rt, mesh = make_case(dim, order, n, pairing, flow=flow, blending=False)
        U = smooth_state(mesh, order, seed=dim + order)
        op = MakeRHSOperator(rt, mesh)
        got, ref = _measures_both(op, U)
        op.close()
        assert np.isfinite(ref).all() and ref[0] > 0
        np.testing.assert_array_equal(got, ref)


def test_noalloc_measures_reject_a_case_beyond_the_fixed_scratch():
    import torch

    from theseus_amd.operator import MakeRHSOperator

    # 3-D p = 4: one element per workgroup, so 41^3 = 68,921 elements need
    # more than the 65,536 workgroups the fixed array holds
    _say("case noalloc measures size limit (41^3 p=4, no launch)")
    rt, mesh = make_case(3, 4, (41, 41, 41), "hll", blending=False)
    op = MakeRHSOperator(rt, mesh)
    u = torch.ones(op.Height, dtype=torch.float64, device="cuda")
    out = torch.zeros(9, dtype=torch.float64, device="cuda")
    lib = A.load()
    rc = lib.theseus_rhs_integral_measures_noalloc(op._h, C.c_void_p(u.data_ptr()), C.c_void_p(out.data_ptr()),
                                                   C.c_void_p(0))
    msg = lib.theseus_last_error().decode()
    op.close()
    assert rc == A.ERR_CAPACITY, (rc, msg)
    assert "65536" in msg
    assert (out.cpu().numpy() == 0).all()
