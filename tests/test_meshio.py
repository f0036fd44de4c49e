"""Mesh readers and the case-file loader on the CPU.

Fixtures under tests/golden/cases/ are the reference's own case inputs
(TestCases/Euler/2D/IsentropicVortex, TestCases/NavierStokes/2D/TaylorGreenVortex,
TestCases/Euler/2D/KelvinHelmholtzInstability: config.json and mesh file).
Checks: the periodic MFEM mesh refined three times is the golden vortex's
24x24 tensor mesh (same geometry per element, same neighbours); the Gmsh
2.2 periodic mesh is the uniform 32x32 grid on [-pi, pi]^2; elements listed
with rotated or reflected corner orders give the same engine descriptor;
boundary physical tags become boundary indices; a non-affine element is
rejected; a Hilbert order on a 2^k grid visits face neighbours one after the
other; and the oracle's RHS on a loaded case equals the RHS on the
hand-built mesh of the same case."""
from pathlib import Path

import numpy as np
import pytest

from theseus_amd import _abi as A
from theseus_amd.mesh import box_mesh, gll_nodes
from theseus_amd.meshio import (AffineMesh, hilbert_order, read_gmsh22, read_mesh, read_mfem, reorder,
                                uniform_refine)
from theseus_amd.operator import build_descriptor

CASES = Path(__file__).resolve().parent / "golden" / "cases"
VORTEX = CASES / "IsentropicVortex"
TGV2D = CASES / "TaylorGreenVortex2D"
KH = CASES / "KelvinHelmholtzInstability"
TGV3D_REF = Path("/root/reference/TestCases/NavierStokes/3D/TaylorGreenVortex/TaylorGreenVortex.msh")


def _check_symmetric(m):
    nb = m.face_neighbor
    for e in range(m.num_elements):
        for f in range(2 * m.dim):
            o = nb[e, f]
            if o >= 0:
                assert nb[o, f ^ 1] == e, (e, f, o)


def _match(a, b):
    """Element map a -> b by element centres (periodic-safe: exact centres)."""
    from scipy.spatial import cKDTree

    ca = a.corners.mean(axis=1) if isinstance(a, AffineMesh) else _centres(a)
    cb = b.corners.mean(axis=1) if isinstance(b, AffineMesh) else _centres(b)
    dist, idx = cKDTree(cb).query(ca)
    assert dist.max() < 1e-9 and len(np.unique(idx)) == len(idx)
    return idx


def _centres(m):
    x01 = np.array([0.0, 1.0])
    X = m.node_coordinates(x01).reshape(m.dim, m.num_elements, -1)
    return X.mean(axis=2).T


def test_mfem_periodic_vortex_mesh_refines_to_the_golden_tensor_mesh():
    from .golden_vortex import vortex_case

    m = read_mfem(VORTEX / "IsentropicVortex.mesh")
    assert m.num_elements == 9 and m.dim == 2 and m.periodic == (True, True)
    assert m.bc_type.size == 0
    _check_symmetric(m)
    w = np.array([0.666666667, 0.666666666, 0.666666667])
    np.testing.assert_allclose(np.sort(m.elem_jac), np.sort(np.outer(w, w).ravel()), rtol=1e-15)
    r = uniform_refine(m, 3)
    _check_symmetric(r)
    _, tm, _ = vortex_case()
    idx = _match(r, tm)
    np.testing.assert_allclose(r.elem_jac, tm.elem_jac[idx], rtol=1e-13)
    np.testing.assert_allclose(r.elem_adj, tm.elem_adj[idx], rtol=1e-13, atol=1e-17)
    for f in range(4):
        assert (idx[r.face_neighbor[:, f]] == tm.face_neighbor[idx, f]).all()


def test_gmsh_periodic_tgv2d_is_the_uniform_grid():
    m = read_gmsh22(TGV2D / "TaylorGreenVortex2D.msh")
    assert m.num_elements == 1024 and m.periodic == (True, True)
    _check_symmetric(m)
    h = 2 * np.pi / 32
    np.testing.assert_allclose(m.elem_jac, h * h, rtol=1e-9)
    np.testing.assert_allclose(m.elem_adj.reshape(-1, 2, 2), np.broadcast_to(np.eye(2) * h, (1024, 2, 2)),
                               rtol=1e-9, atol=1e-12)
    c = m.corners.mean(axis=1)
    for d in range(2):
        step = np.zeros(2)
        step[d] = h
        dx = (c[m.face_neighbor[:, 2 * d + 1]] - c - step + np.pi) % (2 * np.pi) - np.pi
        assert np.abs(dx).max() < 1e-9


@pytest.mark.skipif(not TGV3D_REF.exists(), reason="the reference's 3-D TGV mesh is not present")
def test_gmsh_periodic_tgv3d():
    m = read_gmsh22(TGV3D_REF)
    assert m.num_elements == 32 ** 3 and m.periodic == (True, True, True)
    _check_symmetric(m)
    np.testing.assert_allclose(m.elem_jac, (2 * np.pi / 32) ** 3, rtol=1e-9)


def _grid_mfem_text(n, rotate, reflect, dim=2):
    """A periodic n x n MFEM mesh (L2 nodes) whose element corner lists are
    rotated by rotate[e] quarter turns and optionally mirrored."""
    lines = ["MFEM mesh v1.0", "", "dimension", "2", "", "elements", str(n * n)]
    nodes = []
    for j in range(n):
        for i in range(n):
            e = j * n + i
            v = [j * n + i, j * n + (i + 1) % n, ((j + 1) % n) * n + (i + 1) % n, ((j + 1) % n) * n + i]  # ccw
            xy = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            k = rotate[e] % 4
            v, xy = v[k:] + v[:k], xy[k:] + xy[:k]
            if reflect[e]:
                v, xy = [v[0]] + v[1:][::-1], [xy[0]] + xy[1:][::-1]
            lines.append("1 3 " + " ".join(map(str, v)))
            # L2 nodes in the element's lexicographic corner order: ccw 0,1,3,2
            lex = [xy[0], xy[1], xy[3], xy[2]]
            nodes += [f"{0.5 * a} {0.25 * b}" for a, b in lex]
    lines += ["", "boundary", "0", "", "vertices", str(n * n), "", "nodes", "FiniteElementSpace",
              "FiniteElementCollection: L2_T1_2D_P1", "VDim: 2", "Ordering: 1", ""] + nodes
    return "\n".join(lines) + "\n"


def test_rotated_and_mirrored_corner_orders_give_the_same_descriptor(tmp_path):
    n = 4
    rng = np.random.default_rng(3)
    base = tmp_path / "base.mesh"
    base.write_text(_grid_mfem_text(n, [0] * n * n, [0] * n * n))
    rot = tmp_path / "rot.mesh"
    rot.write_text(_grid_mfem_text(n, rng.integers(0, 4, n * n), rng.integers(0, 2, n * n)))
    a, b = read_mfem(base), read_mfem(rot)
    _check_symmetric(b)
    assert (a.elem_jac > 0).all() and (b.elem_jac > 0).all()
    np.testing.assert_allclose(a.elem_jac, 0.125)
    # every element of b, once oriented, has a's geometry up to a symmetry of
    # the square, and its neighbours are the same elements
    np.testing.assert_allclose(np.sort(np.abs(b.elem_adj), axis=1), np.sort(np.abs(a.elem_adj), axis=1))
    for e in range(n * n):
        assert set(b.face_neighbor[e]) == set(a.face_neighbor[e])


def test_gmsh_boundary_tags_and_affinity(tmp_path):
    """2 x 1 quads on [0,2] x [0,1], physical lines 7 (left), 8 (right), 9 (bottom/top)."""
    txt = """$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
6
1 0 0 0
2 1 0 0
3 2 0 0
4 0 1 0
5 1 1 0
6 2 1 0
$EndNodes
$Elements
8
1 1 2 7 1 4 1
2 1 2 8 2 3 6
3 1 2 9 3 1 2
4 1 2 9 3 2 3
5 1 2 9 4 5 4
6 1 2 9 4 6 5
7 3 2 1 1 1 2 5 4
8 3 2 1 1 2 3 6 5
$EndElements
"""
    p = tmp_path / "box.msh"
    p.write_text(txt)
    m = read_mesh(p, bc_types={7: 2, 8: 2, 9: 2})
    assert m.bc_attributes == (7, 8, 9) and list(m.bc_type) == [2, 2, 2]
    assert m.face_neighbor[0, 0] == -1 and m.face_neighbor[1, 1] == -2
    assert (m.face_neighbor[:, 2:] == -3).all()
    assert m.face_neighbor[0, 1] == 1 and m.face_neighbor[1, 0] == 0
    # a trapezoid is not affine
    bad = txt.replace("6 2 1 0", "6 2.5 1 0")
    q = tmp_path / "bad.msh"
    q.write_text(bad)
    with pytest.raises(ValueError, match="not affine"):
        read_mesh(q)


def test_uniform_refine_of_a_box_is_the_finer_box():
    coarse = box_mesh(3, (2, 1, 2), 0.0, (2.0, 1.0, 3.0))
    corners = np.stack([coarse.node_coordinates(np.array([0.0, 1.0])).reshape(3, coarse.num_elements, 8)[:, :, c].T
                        for c in range(8)], axis=1)
    am = AffineMesh(3, corners, coarse.face_neighbor.copy(), coarse.bc_type)
    fine = uniform_refine(am, 1)
    ref = box_mesh(3, (4, 2, 4), 0.0, (2.0, 1.0, 3.0))
    idx = _match(fine, ref)
    np.testing.assert_allclose(fine.elem_jac, ref.elem_jac[idx], rtol=1e-14)
    for f in range(6):
        assert (idx[fine.face_neighbor[:, f]] == ref.face_neighbor[idx, f]).all()


def test_hilbert_order_on_a_power_of_two_grid_walks_face_neighbours():
    m = read_gmsh22(TGV2D / "TaylorGreenVortex2D.msh")
    order = hilbert_order(m)
    assert sorted(order) == list(range(m.num_elements))
    r = reorder(m, order)
    _check_symmetric(r)
    for k in range(r.num_elements - 1):
        assert k + 1 in set(r.face_neighbor[k]), k
    np.testing.assert_array_equal(r.elem_jac, m.elem_jac[order])


def _oracle_rhs(oracle, rt, mesh, U):
    desc, keep = build_descriptor(rt, mesh)
    c = oracle.OracleCase(desc)
    c._keep = keep
    return c.rhs(U)


def _gather(U, idx, neq, ndof):
    """State on mesh b (element idx[e] of b for element e of a) -> a's order."""
    return U.reshape(neq, -1, ndof)[:, idx, :].reshape(-1)


def test_loaded_vortex_case_equals_the_hand_built_case(oracle):
    from theseus_amd.casefile import load_case

    from .golden_vortex import vortex_case

    cf = load_case(VORTEX / "config.json")
    assert cf.mesh.num_elements == 576 and cf.final_time == 2.0
    rt, tm, U0 = vortex_case()
    assert (cf.runtime.volume_flux, cf.runtime.surface_flux, cf.runtime.blending) == \
        (rt.volume_flux, rt.surface_flux, rt.blending)
    idx = _match(cf.mesh, tm)
    np.testing.assert_allclose(cf.U0, _gather(U0, idx, 4, 16), rtol=1e-14, atol=1e-14)
    r_loaded = _oracle_rhs(oracle, cf.runtime, cf.mesh, cf.U0)
    r_hand = _gather(_oracle_rhs(oracle, rt, tm, U0), idx, 4, 16)
    # same geometry, rounded differently (corner differences of the refined
    # file mesh vs the refined breaks: J and adj agree to ~1e-13, and the
    # divergence amplifies that about twentyfold)
    assert np.linalg.norm(r_loaded - r_hand) <= 1e-11 * np.linalg.norm(r_hand)


def test_loaded_tgv2d_ns_case_equals_the_uniform_box(oracle):
    from theseus_amd.cases import taylor_green_2d
    from theseus_amd.casefile import load_case

    cf = load_case(TGV2D / "config.json")
    rt = cf.runtime
    assert rt.flow_model == A.FLOW_NAVIER_STOKES and rt.mu == 0.000625 and cf.mesh.num_elements == 1024
    box = box_mesh(2, 32, -np.pi, np.pi)
    Ub = taylor_green_2d(box, 3, 1.4, 0.08)
    idx = _match(cf.mesh, box)
    np.testing.assert_allclose(cf.U0, _gather(Ub, idx, 4, 16), rtol=1e-9, atol=1e-10)
    r_loaded = _oracle_rhs(oracle, rt, cf.mesh, cf.U0)
    r_box = _gather(_oracle_rhs(oracle, rt, box, Ub), idx, 4, 16)
    # the Gmsh nodes sit up to 8e-12 off the uniform grid (4e-11 of an
    # element); against the background pressure 1/(gamma Ma^2) ~ 112 those
    # metric differences move the RHS by a few 1e-8 relative
    assert np.linalg.norm(r_loaded - r_box) <= 1e-7 * np.linalg.norm(r_box)


@pytest.mark.parametrize("flux,pair", [("HLL", ("CENTRAL", "HLL")), ("LLF", ("CENTRAL", "LLF")),
                                       ("Pirozzoli", ("PIROZZOLI", "PIROZZOLI")), ("LTEEC", ("LTEEC", "LTEEC"))])
def test_loaded_case_with_the_helmholtz_gas_builds_its_table(flux, pair):
    """gas_model "lte_helmholtz" with each flux the reference's factory takes
    for it (SimFactory_impl.hpp:140-245): the runtime selects the pair and the
    tabulated gas, and the loader builds the table from the case's N_rho, N_T
    and bounds (SimFactory_impl.hpp:146-180)."""
    import json

    from theseus_amd.casefile import load_case_config

    cfg = json.loads((TGV2D / "config.json").read_text())
    cfg["compileTime"] = {"SUBCELL_FV_BLENDING": False}
    cfg["runTime"].update(gas_model="lte_helmholtz", numerical_flux=flux, N_rho=12, N_T=16, rho_min=0.5,
                          rho_max=2.0, T_min=200.0, T_max=2000.0, rho_dist="log", T_dist="log")
    cf = load_case_config(cfg, TGV2D / "TaylorGreenVortex2D.msh")
    rt = cf.runtime
    assert rt.lte and not rt.blending
    assert (rt.volume_flux, rt.surface_flux) == (getattr(A, "VOLUME_" + pair[0]), getattr(A, "SURFACE_" + pair[1]))
    t = cf.lte_table
    assert t is not None and t.rho.size == 12 and t.T.size == 16
    np.testing.assert_allclose([t.rho[0], t.rho[-1], t.T[0], t.T[-1]], [0.5, 2.0, 200.0, 2000.0], rtol=1e-12)
    assert np.allclose(np.diff(np.log(t.T)), np.log(10.0) / 15)  # T_dist "log"


def test_lte_vortex_initial_condition_invariants():
    """LTEVortexIC (StateInit.hpp:15-69) with the LTEVortex case's parameters
    (radius 0.2, vel_inf 10, beta 1, rho_inf 0.1, T_inf 275): p/rho^1.4 is the
    free stream's everywhere (isentropic), p = rho R T with the Gaussian dip
    of depth (vel_inf beta)^2/(2 cp) at the centre, the far field is the free
    stream, and the velocity at the centre is (vel_inf, 0)."""
    from theseus_amd.casefile import initial_condition

    params = {"x1": 0.2, "x2": 10.0, "x3": 1.0, "x4": 0.1, "x5": 275.0}
    mesh = box_mesh(2, 8, -3.0, 3.0)
    U = initial_condition("LTEVortexIC", params, mesh, 3).reshape(4, -1)
    rho, mx, my, E = U
    u, v = mx / rho, my / rho
    p = 0.4 * (E - 0.5 * rho * (u * u + v * v))
    R, cp = 287.05, 1.4 * 287.05 / 0.4
    p_inf = 0.1 * R * 275.0
    np.testing.assert_allclose(p / rho ** 1.4, p_inf / 0.1 ** 1.4, rtol=1e-12)
    T = p / (rho * R)
    from theseus_amd.mesh import gll_nodes

    X = mesh.node_coordinates(gll_nodes(3)[0])
    r2 = (X[0] ** 2 + X[1] ** 2) / 0.04
    np.testing.assert_allclose(T, 275.0 - 50.0 / cp * np.exp(-r2), rtol=1e-12)
    far = r2 > 200
    assert far.any()
    np.testing.assert_allclose(rho[far], 0.1, rtol=1e-12)
    np.testing.assert_allclose(u[far], 10.0, rtol=1e-12)
    np.testing.assert_allclose(v[far], 0.0, atol=1e-10)
    i0 = np.argmin(r2)
    assert r2[i0] < 1e-20
    assert abs(u[i0] - 10.0) < 1e-12 and abs(v[i0]) < 1e-12


def test_loaded_kelvin_helmholtz_case():
    from theseus_amd.casefile import load_case

    cf = load_case(KH / "config.json")
    assert cf.mesh.num_elements == 9 * 4 ** 5 and cf.mesh.periodic == (True, True)
    rho = cf.U0.reshape(4, -1)[0]
    assert rho.min() >= 0.5 - 1e-12 and rho.max() <= 2.0 + 1e-12
