"""Reference golden case: forward-facing step
(TestCases/Euler/2D/ForwardFacingStep/config.json with its Gmsh mesh, copied
to tests/golden/cases/ForwardFacingStep): 3,400 quads refined once to 13,600,
p=3, Chandrashekar volume and surface flux, subcell FV blending on
(alpha_max 0.5), supersonic inflow on the left (the named state
ForwardFacingStepLeftBCVector), slip walls on the top and bottom including the
step, supersonic outflow on the right, Mach-3 uniform initial state
(ForwardFacingStepIC), fixed dt 1e-4 and 100 SSP-RK3 steps, as the reference's
regression recipe runs it (tests/recipes/regression/golden-step.sh:18-27).

Expected fields: the reference's own output at cycle 100
(tests/golden/forward_facing_step_cycle100.npz, decoded by
tests/golden/make_ffs_golden.py).  The fixture holds every node of the 377
elements that deviate from the inflow state by more than 1e-13, with their
blending coefficients (nonzero in 50 of them, up to alpha_max), and the
largest deviation of the other elements per field.

The coefficient the reference writes is the one its last RHS evaluation
computed, i.e. that of the third SSP-RK3 stage of step 100, evaluated on the
second stage's state (RHSOperator_impl.hpp:411-437 runs inside Mult)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from theseus_amd.casefile import load_case
from theseus_amd.mesh import gll_nodes

HERE = Path(__file__).resolve().parent
CASE = HERE / "golden" / "cases" / "ForwardFacingStep"
GOLDEN = HERE / "golden" / "forward_facing_step_cycle100.npz"
DT = 1e-4
NSTEPS = 100
NDOF = 16
FREESTREAM = {"rho": 1.4, "u": 3.0, "v": 0.0, "p": 1.0}
# The reference's recipe compares two runs of one binary at atol 3e-13,
# rtol 1e-13.  A restatement with its own summation orders (GLL nodes, D,
# the inverse Vandermonde, flux and face orders) and its own element
# orientation (the reader re-orients 800 of the 3,400 Gmsh quads so that
# neighbours share the in-face point order, which permutes the volume and
# face sums of those elements) agrees after 100 steps across a shock to a
# few 1e-12 of each field's largest magnitude: the bound held is
# |ours - golden| <= ABS_TOL * max|golden field| at every node.
ABS_TOL = 2e-12
# Blending coefficients: a sigmoid of the modal indicator eta with slope up
# to 9.21/(4 theta) ~ 1.1e3 at p=3, and eta = 1 - |M1|^2/|m|^2 cancels, so
# state differences at the 1e-14 level move alpha by ~1e-12 (the oracle is
# 4e-14 from the reference's values here).
ALPHA_TOL = 1e-10


def ffs_case():
    cf = load_case(CASE / "config.json", CASE / "ForwardFacingStep.msh")
    return cf.runtime, cf.mesh, cf.U0


def golden_match(mesh, order=3):
    """(fixture, element ids of `mesh` that the fixture keeps, node index
    into the fixture arrays for each of their L-vector nodes)."""
    from scipy.spatial import cKDTree

    d = np.load(GOLDEN)
    assert int(d["num_elements"]) == mesh.num_elements
    gx, gy = d["x"], d["y"]
    ng = gx.size // NDOF
    gcen = np.c_[gx.reshape(ng, NDOF).mean(1), gy.reshape(ng, NDOF).mean(1)]
    X = mesh.node_coordinates(gll_nodes(order)[0])
    mx, my = X[0].reshape(-1, NDOF), X[1].reshape(-1, NDOF)
    dist, ours = cKDTree(np.c_[mx.mean(1), my.mean(1)]).query(gcen)
    assert dist.max() < 1e-12 and len(np.unique(ours)) == ng
    perm = np.empty((ng, NDOF), dtype=np.int64)
    for k, e in enumerate(ours):
        gpts = np.c_[gx[k * NDOF:(k + 1) * NDOF], gy[k * NDOF:(k + 1) * NDOF]]
        dd, loc = cKDTree(gpts).query(np.c_[mx[e], my[e]])
        assert dd.max() < 1e-12 and len(np.unique(loc)) == NDOF
        perm[k] = k * NDOF + loc
    return d, ours, perm


def primitive(U, gamma=1.4):
    U = U.reshape(4, -1)
    rho = U[0]
    return {"rho": rho, "u": U[1] / rho, "v": U[2] / rho,
            "p": (gamma - 1) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / rho)}


def check(U, alpha, mesh):
    """Compare a final state and the last stage's blending coefficients with
    the fixture; returns the largest normalised node error per field."""
    d, ours, perm = golden_match(mesh)
    f = primitive(U)
    ne = mesh.num_elements
    rest = np.ones(ne, dtype=bool)
    rest[ours] = False
    errs = {}
    for k, val in f.items():
        v = val.reshape(ne, NDOF)
        g = d[k][perm]
        scale = np.abs(d[k]).max() if k != "v" else np.abs(d["u"]).max()
        err = np.abs(v[ours] - g).max() / scale
        # elements the fixture does not keep: the inflow state to the bound the fixture records
        rest_err = np.abs(v[rest] - FREESTREAM[k]).max()
        errs[k] = float(err)
        assert err <= ABS_TOL, (k, err)
        assert rest_err <= float(d["rest_dev_" + k]) + ABS_TOL * scale, (k, rest_err)
    a = np.asarray(alpha)
    assert np.abs(a[ours] - d["alpha"]).max() <= ALPHA_TOL, np.abs(a[ours] - d["alpha"]).max()
    assert (a[rest] == 0).all()
    assert ((a[ours] > 0) == (d["alpha"] > 0)).all()
    return errs
