"""Reference golden case: isentropic vortex on periodic [-1,1]^2 (base
vertices +-0.333333333, 3 uniform refinements -> 24x24 quads), p=3,
numerical_flux 'Chandrashekar', blending on, gamma 1.4, R 287.05, IC
IsentropicVortexIC(0.2, 0.5, 0.2, 287.05, 1.4), dt = 1e-3 fixed, 500 SSP-RK3
steps (TestCases/Euler/2D/IsentropicVortex/config.json,
tests/recipes/regression/golden-vortex.sh).  Expected fields come from the
reference's own output (tests/golden/isentropic_vortex_cycle500.npz)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from theseus_amd import _abi as A
from theseus_amd.cases import reference_isentropic_vortex
from theseus_amd.config import RunTime
from theseus_amd.mesh import gll_nodes, refine_breaks, tensor_mesh

GOLDEN = Path(__file__).resolve().parent / "golden" / "isentropic_vortex_cycle500.npz"
DT = 1e-3
NSTEPS = 500


def vortex_case():
    b = refine_breaks([-1.0, -0.333333333, 0.333333333, 1.0], 3)
    mesh = tensor_mesh([b, b], periodic=True)
    rt = RunTime(dim=2, order=3, flow_model=A.FLOW_EULER, volume_flux=A.VOLUME_CHANDRASHEKAR,
                 surface_flux=A.SURFACE_CHANDRASHEKAR, blending=True, gamma=1.4, R_gas=287.05)
    U0 = reference_isentropic_vortex(mesh, 3)
    return rt, mesh, U0


def golden_permutation(mesh):
    """Index into the golden arrays for every L-vector dof of `mesh`."""
    d = np.load(GOLDEN)
    gx, gy = d["x"], d["y"]
    ndof = 16
    ng = gx.size // ndof
    gcen = np.c_[gx.reshape(ng, ndof).mean(1), gy.reshape(ng, ndof).mean(1)]
    x01, _ = gll_nodes(3)
    X = mesh.node_coordinates(x01)
    mx = X[0].reshape(-1, ndof)
    my = X[1].reshape(-1, ndof)
    mcen = np.c_[mx.mean(1), my.mean(1)]
    from scipy.spatial import cKDTree

    _, gel = cKDTree(gcen).query(mcen)
    assert len(np.unique(gel)) == mesh.num_elements
    perm = np.empty(mx.size, dtype=np.int64)
    for e in range(mesh.num_elements):
        gpts = np.c_[gx[gel[e] * ndof:(gel[e] + 1) * ndof], gy[gel[e] * ndof:(gel[e] + 1) * ndof]]
        dist, loc = cKDTree(gpts).query(np.c_[mx[e], my[e]])
        assert dist.max() < 1e-12 and len(np.unique(loc)) == ndof
        perm[e * ndof:(e + 1) * ndof] = gel[e] * ndof + loc
    return d, perm


def primitive_fields(U, gamma=1.4):
    U = U.reshape(4, -1)
    rho = U[0]
    u = U[1] / rho
    v = U[2] / rho
    p = (gamma - 1) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / rho)
    return {"rho": rho, "u": u, "v": v, "p": p}
