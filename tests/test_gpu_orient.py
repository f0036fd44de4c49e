"""Meshes whose elements keep their own storage frames, through the C-ABI:
the reference's forward-facing-step Gmsh mesh read without re-orienting
its 800 rotated quads, and a 3-D Navier-Stokes box whose every element was
given a random one of the 48 frames of the cube.  The handle with
elem_orient takes and returns states in the storage frames; its results,
brought back to the aligned frame, equal the aligned handle's bit for bit
(same kernels on the same aligned data) and the oracle's to 1e-12."""
import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from theseus_amd import _abi as A

from .gpu_helpers import make_case, oracle_case, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

CASES = Path(__file__).resolve().parent / "golden" / "cases"


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ffs_unrefined():
    import json

    from theseus_amd.casefile import load_case_config

    cfg = json.loads((CASES / "ForwardFacingStep" / "config.json").read_text())
    cfg["runTime"]["ser_ref_levels"] = 0
    cf = load_case_config(cfg, CASES / "ForwardFacingStep" / "ForwardFacingStep.msh")
    from theseus_amd.mesh import gll_nodes

    X = cf.mesh.node_coordinates(gll_nodes(cf.runtime.order)[0])
    U = cf.U0.reshape(4, -1).copy()
    bump = 0.05 * np.sin(3.0 * X[0]) * np.cos(5.0 * X[1])
    U[0] *= 1.0 + bump
    U[3] *= 1.0 + 0.5 * bump
    return cf.runtime, cf.mesh, U.reshape(-1)


def _run(rt, mesh, U, orient, nsteps, dt, lte_table=None):
    import torch

    from theseus_amd.operator import MakeRHSOperator

    op = MakeRHSOperator(rt, mesh, lte_table=lte_table, elem_orient=orient)
    u = torch.from_numpy(U).cuda()
    du = torch.empty_like(u)
    op.Mult(u, du)
    stab = op.EstimateStability(u)
    meas = op.ComputeIntegralMeasures(u)
    for _ in range(nsteps):
        op.Step(u, 0.0, dt)
    torch.cuda.synchronize()
    out = (du.cpu().numpy(), u.cpu().numpy(), (stab.advective_rate, stab.surface_rate, stab.diffusive_rate), meas)
    op.close()
    return out


def _check(rt, mesh, U_al, nsteps, dt, oracle):
    from theseus_amd.meshio import to_aligned, to_storage

    p = rt.order
    ref_rhs = oracle_case(oracle, rt, mesh).rhs(U_al)
    du_a, u_a, st_a, me_a = _run(rt, mesh, U_al, None, nsteps, dt)
    du_s, u_s, st_s, me_s = _run(rt, mesh, to_storage(U_al, mesh, p), mesh.elem_orient, nsteps, dt)
    np.testing.assert_array_equal(to_aligned(du_s, mesh, p), du_a)
    np.testing.assert_array_equal(to_aligned(u_s, mesh, p), u_a)
    assert st_s == st_a and me_s == me_a
    err = rel_l2(to_aligned(du_s, mesh, p), ref_rhs)
    _say(f"  rhs rel L2 vs oracle {err:.2e}")
    assert err <= 1e-12


def test_forward_facing_step_mesh_in_its_file_frames(oracle):
    _say("case orientation: forward-facing step 3400 quads, 800 stored rotated, 20 steps")
    rt, mesh, U = _ffs_unrefined()
    assert (mesh.elem_orient != 0).sum() == 800
    _check(rt, mesh, U, 20, 1e-4, oracle)


def test_navier_stokes_box_with_random_element_frames(oracle):
    _say("case orientation: 3-D NS box 3^3 p=3, random frames of the cube, 10 steps")
    rt, box = make_case(3, 3, (3, 3, 3), "ranocha_llf", flow=A.FLOW_NAVIER_STOKES, blending=True)
    rng = np.random.default_rng(11)
    codes = rng.integers(0, 48, size=box.num_elements).astype(np.uint8)
    codes[0] = 5  # not all aligned
    mesh = SimpleNamespace(**{k: getattr(box, k) for k in ("dim", "num_elements", "face_neighbor", "bc_type",
                                                           "elem_jac", "elem_adj", "lower", "upper")})
    mesh.elem_orient = codes
    mesh.node_coordinates = box.node_coordinates
    U = smooth_state(box, 3, seed=3, amp=0.2)
    _check(rt, mesh, U, 10, 2e-4, oracle)


@pytest.mark.parametrize("viscous,mixed", [(False, False), (True, False), (False, True)])
def test_tabulated_gas_box_with_random_element_frames(oracle, viscous, mixed):
    """The tabulated gas (LTE Taylor-Green vortex, 3^3 p=3; Euler and
    Navier-Stokes in FP64, Euler in mixed precision) on a box whose elements
    keep random frames of the cube: the node temperatures are recovered on
    the aligned copy, so the results equal the aligned handle's bit for bit,
    and the FP64 RHS the oracle's to 1e-12."""
    import dataclasses

    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.lte import synthetic_table
    from theseus_amd.meshio import to_aligned, to_storage
    from theseus_amd.operator import build_descriptor

    _say(f"case orientation: LTE box 3^3 p=3 viscous={viscous} mixed={mixed}, random frames, 5 steps")
    table = synthetic_table(128, 128)
    rt, box, U, _ = c4_lte_tgv(n=3, order=3, table=table, viscous=viscous)
    if mixed:
        rt = dataclasses.replace(rt, precision=A.PRECISION_MIXED)
    rng = np.random.default_rng(17)
    codes = rng.integers(0, 48, size=box.num_elements).astype(np.uint8)
    codes[0] = 7
    mesh = SimpleNamespace(**{k: getattr(box, k) for k in ("dim", "num_elements", "face_neighbor", "bc_type",
                                                           "elem_jac", "elem_adj", "lower", "upper")})
    mesh.elem_orient = codes
    mesh.node_coordinates = box.node_coordinates
    dt = 2e-6
    du_a, u_a, st_a, me_a = _run(rt, mesh, U, None, 5, dt, lte_table=table)
    du_s, u_s, st_s, me_s = _run(rt, mesh, to_storage(U, mesh, 3), mesh.elem_orient, 5, dt, lte_table=table)
    np.testing.assert_array_equal(to_aligned(du_s, mesh, 3), du_a)
    np.testing.assert_array_equal(to_aligned(u_s, mesh, 3), u_a)
    assert st_s == st_a and me_s == me_a
    if not mixed:
        desc, keep = build_descriptor(dataclasses.replace(rt, precision=A.PRECISION_FP64), box)
        c = oracle.OracleCase(desc)
        c._keep = keep
        t = oracle.LteTable(arrays=(table.rho, table.T, table.e, table.planes.reshape(-1), table.inv.reshape(-1)))
        t.attach(c)
        err = rel_l2(du_a, c.rhs(U))
        _say(f"  rhs rel L2 vs oracle {err:.2e}")
        assert err <= 1e-12
