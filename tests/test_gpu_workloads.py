"""The measured workloads at their full sizes, stepped on the engine and by
the oracle from the same initial state:

* the 2-D Euler isentropic vortex (Shu) on the periodic square [-5,5]^2,
  16x16 quads, p=3, Chandrashekar volume + Roe face flux, blending on,
  100 SSP-RK3 steps at CFL 0.5 with the step chosen on the device;
* the 2-D four-quadrant Riemann problem, 128x128 quads, p=4, central volume
  + HLL face flux, subcell FV blending, transmissive boundaries, 100 steps
  at the fixed dt of CFL 0.5 on the initial state, with the state and the
  blending coefficients compared every 10 steps.

Bar: state after 100 steps within 1e-10 relative L2 of the oracle; the same
set of blended elements, and blending coefficients within 1e-10: alpha is a
sigmoid of the indicator eta with slope up to 9.21/(4 theta) ~ 1.6e3 at p=4,
and eta = 1 - |M1|^2/|m|^2 cancels, so state differences of ~1e-14 move alpha
by ~1e-12."""
import sys

import numpy as np
import pytest

from .gpu_helpers import oracle_case, per_field_rel, rel_l2

pytestmark = pytest.mark.gpu

ALPHA_TOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_c1_vortex_100_cfl_steps_match_the_oracle(oracle):
    import torch

    from theseus_amd.cases import c1_vortex2d
    from theseus_amd.operator import DeviceClock, MakeRHSOperator

    _say("case C1 vortex 16^2 p=3 Chandrashekar+Roe, 100 StepCFL steps at CFL 0.5")
    rt, mesh, U0 = c1_vortex2d()
    c = oracle_case(oracle, rt, mesh)
    cfl, growth, nsteps = 0.5, 1.2, 100
    Uref, tref, dtref, dtr_ref, sref = c.adaptive_steps(U0, cfl, growth, float("inf"), nsteps)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U0).cuda()
    clk = DeviceClock(cfl, growth)
    for _ in range(nsteps):
        op.StepCFL(u, clk)
    torch.cuda.synchronize()
    st = clk.read()
    got = u.cpu().numpy()
    op.close()
    assert st["steps"] == sref == nsteps
    assert st["t"] == pytest.approx(tref, rel=1e-12)
    err = rel_l2(got, Uref)
    _say(f"  t = {st['t']:.6f}, state rel L2 {err:.2e}")
    assert err <= 1e-10, (err, per_field_rel(got, Uref, 4))


def test_c3_riemann_full_size_100_steps_match_the_oracle(oracle):
    import torch

    from theseus_amd.cases import c3_riemann2d
    from theseus_amd.operator import MakeRHSOperator

    _say("case C3 Riemann 128^2 p=4 HLL, blending, 100 fixed steps")
    rt, mesh, U0 = c3_riemann2d()
    c = oracle_case(oracle, rt, mesh)
    oracle._lib.oracle_set_num_threads(16)
    r = c.stability(U0)
    dt = 0.5 / (max(r[0], r[1]) + r[2])
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U0).cuda()
    U = U0.copy()
    blended = 0
    for s in range(1, 101):
        if s % 10 == 0:
            # the stage-3 input of this step: its coefficients are what the
            # engine reports after the step
            u1 = U + dt * c.rhs(U)
            u2 = 0.75 * U + 0.25 * (u1 + dt * c.rhs(u1))
            aref, _ = c.blending(u2)
        U = c.step(U, dt)
        op.Step(u, 0.0, dt)
        if s % 10 == 0:
            a = op.Alpha().cpu().numpy()
            got = u.cpu().numpy()
            err = rel_l2(got, U)
            blended = int((aref > 0).sum())
            _say(f"  step {s}: state rel L2 {err:.2e}, alpha max diff {np.abs(a - aref).max():.2e}, "
                 f"{blended} blended elements")
            assert err <= 1e-10, (s, err, per_field_rel(got, U, 4))
            assert ((a > 0) == (aref > 0)).all()
            assert np.abs(a - aref).max() <= ALPHA_TOL
    op.close()
    assert blended > 0
