"""Element orientation through the C-ABI (host code, no GPU):
theseus_orient_elements aligns the reference's own forward-facing-step Gmsh
mesh (800 of its 3,400 quads stored in a rotated frame) exactly as the case
reader does, and aligns a hex mesh whose every element was given a random
storage frame; theseus_orient_node matches the reader's node tables."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from theseus_amd import _abi as A
from theseus_amd.meshio import _symmetries, node_storage_index, read_mesh, storage_frames

CASES = Path(__file__).resolve().parent / "golden" / "cases"


def _orient(dim, vids, xyz):
    lib = A.load()
    ne = vids.shape[0]
    vids = np.ascontiguousarray(vids, dtype=np.int64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    code = np.zeros(ne, dtype=np.uint8)
    nbr = np.zeros((ne, 2 * dim), dtype=np.int32)
    jac = np.zeros(ne)
    adj = np.zeros((ne, dim * dim))
    rc = lib.theseus_orient_elements(dim, ne, vids.ctypes.data, xyz.ctypes.data, code.ctypes.data, nbr.ctypes.data,
                                     jac.ctypes.data, adj.ctypes.data)
    if rc:
        raise A.TheseusError(rc, lib.theseus_last_error().decode())
    return code, nbr, jac, adj


def test_forward_facing_step_mesh_is_aligned_as_the_reader_aligns_it():
    mesh = read_mesh(CASES / "ForwardFacingStep" / "ForwardFacingStep.msh", 2)
    assert (mesh.elem_orient != 0).sum() == 800
    vids, xyz = storage_frames(mesh)
    code, nbr, jac, adj = _orient(2, vids, xyz)
    np.testing.assert_array_equal(code, mesh.elem_orient)
    np.testing.assert_array_equal(nbr, np.where(mesh.face_neighbor >= 0, mesh.face_neighbor, -1))
    np.testing.assert_allclose(jac, mesh.elem_jac, rtol=1e-13)
    np.testing.assert_allclose(adj, mesh.elem_adj, rtol=1e-12, atol=1e-15 * np.abs(mesh.elem_adj).max())


def test_hex_mesh_with_random_storage_frames_is_realigned():
    from theseus_amd.mesh import box_mesh

    rng = np.random.default_rng(7)
    n = 4
    box = box_mesh(3, (n, n, n), 0.0, 1.0, periodic=False)
    ne = box.num_elements
    # vertex ids of a structured grid, element corners in lexicographic order
    ijk = np.stack(np.unravel_index(np.arange(ne), (n, n, n), order="F"), axis=1)
    corners = np.array([[(c >> d) & 1 for d in range(3)] for c in range(8)])
    vid = lambda p: p[..., 0] + (n + 1) * (p[..., 1] + (n + 1) * p[..., 2])  # noqa: E731
    lex = vid(ijk[:, None, :] + corners[None])
    xyz = (ijk[:, None, :] + corners[None]) / n
    syms = _symmetries(3)
    codes = rng.integers(0, len(syms), size=ne)
    st_vids = np.empty_like(lex)
    st_xyz = np.empty_like(xyz)
    for e in range(ne):
        cm = syms[codes[e]]
        st_vids[e, cm] = lex[e]  # storage corner cm[c] holds aligned corner c
        st_xyz[e, cm] = xyz[e]
    code, nbr, jac, adj = _orient(3, st_vids, st_xyz)
    np.testing.assert_allclose(jac, 1.0 / n ** 3, rtol=1e-13)
    # every interior face is seen with the same corner order from both sides
    faces = [[[c for c in range(8) if ((c >> d) & 1) == s] for s in (0, 1)] for d in range(3)]
    aligned = np.array([st_vids[e, syms[code[e]]] for e in range(ne)])
    inner = 0
    for e in range(ne):
        for d in range(3):
            for s in (0, 1):
                f = nbr[e, 2 * d + s]
                if f >= 0:
                    inner += 1
                    assert list(aligned[e, faces[d][s]]) == list(aligned[f, faces[d][1 - s]])
    assert inner == 2 * 3 * n * n * (n - 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_orient_node_matches_the_reader_tables(dim):
    lib = A.load()
    for code in range(len(_symmetries(dim))):
        for order in (1, 3):
            want = node_storage_index(dim, order, code)
            got = [lib.theseus_orient_node(dim, order, code, n) for n in range(want.size)]
            np.testing.assert_array_equal(got, want)
    assert lib.theseus_orient_node(dim, 3, len(_symmetries(dim)), 0) < 0
