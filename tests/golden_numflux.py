"""Reference golden cases of the numerical-flux regression
(TestCases/GoldenData/NumericalFluxes, copied to tests/golden/numflux with the
CPG part of its manifest): the Navier-Stokes lid-driven cavity (the
LidDrivenCavity mesh, 16x16 quads, Hilbert order) at p=2 with the reference's
"Roe" and "HLL" pairings (central volume flux; Roe or HLL face flux), no-slip
adiabatic walls, a lid moving at (1, 0), fixed dt 1e-4; cycles 10 and 20.
The reference's checker (tests/numflux_regression_test.py:79-112,182-211)
keys each row by (subcell centre, node coordinates) of the gll_subcells
output and compares density, pressure and velocity at rtol 1e-10,
atol 1e-12; the same tolerances are used here."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from theseus_amd.casefile import load_case_config
from theseus_amd.mesh import gll_nodes

HERE = Path(__file__).resolve().parent
DIR = HERE / "golden" / "numflux"
MESH = HERE / "golden" / "cases" / "LidDrivenCavity" / "LidDrivenCavity.msh"
MANIFEST = json.loads((DIR / "manifest.json").read_text())
RTOL = MANIFEST["rtol"]
ATOL = MANIFEST["atol"]
DT = 1e-4
CYCLES = (10, 20)
CASES = ("cpg-roe-cavity", "cpg-hll-cavity")


def numflux_case(name):
    c = MANIFEST["cases"][name]
    cf = load_case_config({"compileTime": c["compileTime"], "runTime": c["runTime"]}, MESH)
    assert cf.runtime.dt == DT and not cf.runtime.variable_dt
    return cf.runtime, cf.mesh, cf.U0


def golden_rows(mesh, name, cycle, order=2):
    """(fixture, L-vector node index of each fixture row)."""
    from scipy.spatial import cKDTree

    d = np.load(DIR / f"{name}-{cycle:06d}.npz", allow_pickle=False)
    nd = (order + 1) ** 2
    X = mesh.node_coordinates(gll_nodes(order)[0])
    mx, my = X[0].reshape(-1, nd), X[1].reshape(-1, nd)
    key = d["coordinates"]
    _, el = cKDTree(np.c_[mx.mean(1), my.mean(1)]).query(key[:, :2])
    rows = np.empty(key.shape[0], dtype=np.int64)
    for r in range(key.shape[0]):
        e = el[r]
        dist = np.hypot(mx[e] - key[r, 3], my[e] - key[r, 4])
        k = int(np.argmin(dist))
        assert dist[k] < 1e-12
        rows[r] = e * nd + k
    return d, rows


def compare(U, d, rows, gamma=1.4):
    """Largest |ours - golden| - (atol + rtol |golden|) per field (<= 0 passes)."""
    U = U.reshape(4, -1)
    rho = U[0][rows]
    u, v = U[1][rows] / rho, U[2][rows] / rho
    p = (gamma - 1) * (U[3][rows] - 0.5 * (U[1][rows] ** 2 + U[2][rows] ** 2) / rho)
    got = {"Density": rho, "Pressure": p, "Velocity": np.c_[u, v]}
    out = {}
    for k, val in got.items():
        g = d[k]
        out[k] = float((np.abs(val - g) - (ATOL + RTOL * np.abs(g))).max())
        out[k + "_maxabs"] = float(np.abs(val - g).max())
    return out
