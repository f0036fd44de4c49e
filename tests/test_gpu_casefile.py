"""Reference case files run on the GPU engine: the isentropic-vortex case
loaded from its own config.json and MFEM mesh (three uniform refinements)
reproduces the reference's golden output after 500 SSP-RK3 steps
(tests/golden/isentropic_vortex_cycle500.npz), and the Navier-Stokes 2-D
Taylor-Green case loaded from its Gmsh file (Hilbert-ordered) matches the
oracle's RHS on the same loaded mesh."""
import sys
from pathlib import Path

import numpy as np
import pytest

from .gpu_helpers import oracle_case, per_field_rel, rel_l2

pytestmark = pytest.mark.gpu

CASES = Path(__file__).resolve().parent / "golden" / "cases"


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_loaded_vortex_case_reproduces_the_reference_golden_output():
    import torch

    from theseus_amd.casefile import load_case
    from theseus_amd.operator import MakeRHSOperator

    from .golden_vortex import DT, NSTEPS, golden_permutation, primitive_fields

    _say("case casefile vortex 500 steps")
    cf = load_case(CASES / "IsentropicVortex" / "config.json")
    d, perm = golden_permutation(cf.mesh)
    op = MakeRHSOperator(cf.runtime, cf.mesh)
    u = torch.from_numpy(cf.U0).cuda()
    t = 0.0
    for _ in range(NSTEPS):
        t = op.Step(u, t, DT)
    torch.cuda.synchronize()
    op.close()
    f = primitive_fields(u.cpu().numpy())
    for k in ("rho", "u", "v", "p"):
        g = d[k][perm]
        rel = np.linalg.norm(f[k] - g) / np.linalg.norm(g)
        assert rel < 1e-11, (k, rel)


def test_loaded_tgv2d_navier_stokes_rhs_parity(oracle):
    import torch

    from theseus_amd.casefile import load_case
    from theseus_amd.operator import MakeRHSOperator

    _say("case casefile tgv2d NS rhs")
    cf = load_case(CASES / "TaylorGreenVortex2D" / "config.json")
    ref = oracle_case(oracle, cf.runtime, cf.mesh).rhs(cf.U0)
    op = MakeRHSOperator(cf.runtime, cf.mesh)
    u = torch.from_numpy(cf.U0).cuda()
    du = torch.empty_like(u)
    op.Mult(u, du)
    torch.cuda.synchronize()
    op.close()
    got = du.cpu().numpy()
    err = rel_l2(got, ref)
    # Ma 0.08 on a 32^2 mesh: as for the full-size 3-D Taylor-Green case, the
    # RHS is fixed by FP64 only to ~7e-12 (the oracle's own builds with and
    # without fused multiply-add differ by 7.0e-12 here)
    assert err <= 2e-11, (err, per_field_rel(got, ref, 4))


def test_loaded_lid_driven_cavity_reproduces_the_reference_golden_output():
    """The reference's Navier-Stokes lid-driven cavity case file (no-slip
    adiabatic walls, moving lid, named boundary data), loaded and run on the
    engine for 4000 SSP-RK3 steps, against the reference's own output."""
    import torch

    from theseus_amd.operator import MakeRHSOperator

    from .golden_cavity import DT, NSTEPS, cavity_case, check_fields, golden_permutation

    _say("case golden lid-driven cavity 16^2 p=3 NS, 4000 steps")
    rt, mesh, U0 = cavity_case()
    d, perm = golden_permutation(mesh)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U0).cuda()
    t = 0.0
    for _ in range(NSTEPS):
        t = op.Step(u, t, DT)
    torch.cuda.synchronize()
    op.close()
    errs = check_fields(u.cpu().numpy(), d, perm)
    _say(f"  rel L2 per field {errs}")


def test_forward_facing_step_reproduces_the_reference_golden_output():
    """The reference's forward-facing-step case file (13,600 quads p=3,
    Chandrashekar, subcell FV blending, slip walls, supersonic inflow and
    outflow), 100 fixed steps on the engine, against the reference's own
    output: fields to 2e-12 of each field's scale, and the blending
    coefficients of the last stage (nonzero in 50 elements) to 1e-12."""
    import torch

    from theseus_amd.operator import MakeRHSOperator

    from .golden_ffs import DT, NSTEPS, check, ffs_case

    _say("case golden forward-facing step 13600 quads p=3, blending, 100 steps")
    rt, mesh, U0 = ffs_case()
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U0).cuda()
    t = 0.0
    for _ in range(NSTEPS):
        t = op.Step(u, t, DT)
    alpha = op.Alpha().cpu().numpy()
    torch.cuda.synchronize()
    op.close()
    errs = check(u.cpu().numpy(), alpha, mesh)
    _say(f"  largest node error / field scale {errs}")


@pytest.mark.parametrize("name", ["cpg-roe-cavity", "cpg-hll-cavity"])
def test_numerical_flux_cavity_reproduces_the_reference_golden_output(name):
    """The reference's Roe and HLL numerical-flux goldens (p=2 Navier-Stokes
    cavity, no-slip walls, moving lid) at cycles 10 and 20, at the reference
    checker's rtol 1e-10 and atol 1e-12."""
    import torch

    from theseus_amd.operator import MakeRHSOperator

    from .golden_numflux import CYCLES, DT, compare, golden_rows, numflux_case

    _say(f"case golden {name} p=2 NS, 20 steps")
    rt, mesh, U0 = numflux_case(name)
    op = MakeRHSOperator(rt, mesh)
    u = torch.from_numpy(U0).cuda()
    t = 0.0
    for s in range(1, max(CYCLES) + 1):
        t = op.Step(u, t, DT)
        if s in CYCLES:
            d, rows = golden_rows(mesh, name, s)
            res = compare(u.cpu().numpy(), d, rows)
            _say(f"  cycle {s}: {res}")
            assert all(res[k] <= 0 for k in ("Density", "Pressure", "Velocity")), (s, res)
    op.close()
