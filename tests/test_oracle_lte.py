"""CPU oracle of the tabulated LTE gas (oracle/oracle_lte.hpp), pinned by:
the reference's 80-digit discrete-gradient golden vectors (one-cell tables,
tests/helmholtz_difference_cases.hpp and helmholtz_channel_difference_cases.hpp,
tolerance 4096 eps as in helmholtz_eos_tests.cpp:32-45 and
helmholtz_mean_stress_tests.cpp:116-118), exact reversal symmetry, the
polynomial table H = 2r + r^2/2 + rT/4 - T^2 that the bicubic interpolant
reproduces exactly (helmholtz_fixture.hpp, helmholtz_eos_tests.cpp:95-98),
the LTEEC flux against its closed form on that table
(lteec_flux_fixture.hpp:47-107, 2e-10), temperature recovery round trips on
the synthetic dissociating-gas table, and the synthetic table's
thermodynamics (p = rho R T, e(T) from its closed form)."""
from pathlib import Path

import numpy as np
import pytest

from oracle.pyoracle import LteTable

GOLD = Path(__file__).resolve().parent / "golden" / "helmholtz_means.npz"
EPS = np.finfo(float).eps


def one_cell_table(bounds, jets):
    rho = np.array(bounds[:2])
    T = np.array(bounds[2:])
    e = np.array([-1e9, 1e9])
    planes = np.zeros(24)
    planes[:16] = jets
    inv = np.full(4, 0.5 * (bounds[2] + bounds[3]))
    return LteTable(arrays=(rho, T, e, planes, inv))


@pytest.mark.parametrize("key", ["difference", "channel"])
def test_thermodynamic_means_match_reference_golden_vectors(key):
    g = np.load(GOLD)
    n = len(g[f"{key}_expected"])
    assert n >= 7
    for c in range(n):
        t = one_cell_table(g[f"{key}_bounds"][c], g[f"{key}_jets"][c])
        r1, t1, r2, t2 = g[f"{key}_states"][c]
        exp = g[f"{key}_expected"][c]
        fwd = t.means(r1, t1, r2, t2)
        rev = t.means(r2, t2, r1, t1)
        for q in range(2):
            tol = 4096 * EPS * max(1.0, abs(exp[q]))
            assert abs(fwd[q] - exp[q]) <= tol, (key, c, q, fwd[q], exp[q])
        assert np.array_equal(fwd, rev), (key, c, fwd, rev)  # exact reversal symmetry


def polynomial_table():
    """helmholtz_fixture.hpp: 3x3 grid, rho in [0.5,2], T in [2,5]."""
    rho = np.linspace(0.5, 2, 3)
    T = np.linspace(2, 5, 3)
    e = np.linspace(0, 32, 3)
    planes = np.zeros((6, 3, 3))
    for j in range(3):
        for i in range(3):
            r, t = rho[i], T[j]
            planes[0, j, i] = 2 * r + r * r / 2 + r * t / 4 - t * t
            planes[1, j, i] = 2 + r + t / 4
            planes[2, j, i] = r / 4 - 2 * t
            planes[3, j, i] = 0.25
            planes[4, j, i] = 2 + r + t
            planes[5, j, i] = 4 + 2 * r + t
    return LteTable(arrays=(rho, T, e, planes.reshape(-1), np.full(9, 5.0)))


def test_polynomial_table_is_reproduced_exactly():
    t = polynomial_table()
    rng = np.random.default_rng(1)
    for _ in range(50):
        r, T = rng.uniform(0.5, 2), rng.uniform(2, 5)
        v = t.eval(r, T)
        tol = 4096 * EPS
        assert abs(v["H"] - (2 * r + r * r / 2 + r * T / 4 - T * T)) <= tol * 40
        assert abs(v["energy"] - (2 * r + r * r / 2 + T * T)) <= tol * 40
        assert abs(v["entropy"] - (2 * T - r / 4)) <= tol * 10
        assert abs(v["pressure"] - r * r * (2 + r + T / 4)) <= tol * 40
        # c^2 = p_rho - rho^2 H_rhoT^2 / H_TT with H_TT = -2
        c2 = 4 * r + 3 * r * r + r * T / 2 + r * r / 32
        assert abs(v["c2"] - c2) <= tol * 100


def closed_form_lteec(dim, r1, t1, r2, t2, q1, q2, n, face):
    """Independent closed-form divided differences for H = 2r + r^2/2 + rT/4 - T^2
    (lteec_flux_fixture.hpp:47-107)."""
    b1, b2 = 1 / t1, 1 / t2
    dr, db = r2 - r1, b2 - b1
    diff_r = abs(dr) >= 2.22e-15 + 1.49e-8 * max(r1, r2)
    diff_b = abs(db) >= 2.22e-15 + 1.49e-8 * max(b1, b2)
    pr = (0.5 * (b1 + b2) * (2 * (r1 + r2) + r1 * r1 + r1 * r2 + r2 * r2) + (r1 + r2) / 4 if diff_r else
          0.5 * ((4 * r1 + 3 * r1 * r1 + r1 * t1 / 2) * b1 + (4 * r2 + 3 * r2 * r2 + r2 * t2 / 2) * b2))
    gr = (0.5 * (b1 + b2) * (4 + 1.5 * (r1 + r2)) + 0.5 if diff_r else
          0.5 * ((4 + 3 * r1 + t1 / 2) * b1 + (4 + 3 * r2 + t2 / 2) * b2))
    pb = 0.5 * (2 * r1 * r1 + r1 ** 3 + 2 * r2 * r2 + r2 ** 3)
    gb = 0.5 * (4 * r1 + 1.5 * r1 * r1 + 4 * r2 + 1.5 * r2 * r2) + (t1 * t2 if diff_b else 0.5 * (t1 * t1 + t2 * t2))
    rho = pr / gr
    energy = gb - pb / rho
    p1 = r1 * r1 * (2 + r1 + t1 / 4)
    p2 = r2 * r2 * (2 + r2 + t2 / 4)
    u1, u2 = q1[1:dim + 1] / r1, q2[1:dim + 1] / r2
    un = 0.5 * (u1 + u2) @ n
    kinetic = 0.5 * (u1 @ u2)
    work = 0.5 * (p1 * u2 + p2 * u1) @ n
    c1 = np.sqrt(4 * r1 + 3 * r1 * r1 + r1 * t1 / 2 + r1 * r1 / 32)
    c2 = np.sqrt(4 * r2 + 3 * r2 * r2 + r2 * t2 / 2 + r2 * r2 / 32)
    nn = np.sqrt(n @ n)
    lam = 0.5 * max(abs(u1 @ n) + c1 * nn, abs(u2 @ n) + c2 * nn) if face else 0.0
    out = np.zeros(dim + 2)
    out[0] = rho * un - lam * (q2[0] - q1[0])
    out[1:dim + 1] = rho * un * 0.5 * (u1 + u2) + 0.5 * (p1 + p2) * n - lam * (q2[1:dim + 1] - q1[1:dim + 1])
    out[dim + 1] = rho * un * (energy + kinetic) + work - lam * (q2[dim + 1] - q1[dim + 1])
    return out


def fixture_state(dim, r, t, shift):
    """lteec_flux_fixture.hpp:45-57"""
    q = np.zeros(dim + 2)
    q[0] = r
    kin = 0.0
    for d in range(dim):
        u = 0.5 * (d + 1) + shift
        q[d + 1] = r * u
        kin += 0.5 * u * u
    q[dim + 1] = r * (2 * r + r * r / 2 + t * t + kin)
    return q


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("face", [False, True])
def test_lteec_flux_matches_closed_form(dim, face):
    t = polynomial_table()
    rng = np.random.default_rng(7 + dim + 2 * face)
    for k in range(40):
        r1, r2 = rng.uniform(0.6, 1.9, size=2)
        t1, t2 = rng.uniform(2.2, 4.8, size=2)
        if k % 5 == 0:
            r2, t2 = r1 * (1 + 1e-10), t1  # near-coincident densities: the derivative branch
        q1 = fixture_state(dim, r1, t1, 0.1)
        q2 = fixture_state(dim, r2, t2, -0.2)
        n = rng.normal(size=dim)
        F = t.flux(face, dim, q1, q2, n)
        ref = closed_form_lteec(dim, r1, t1, r2, t2, q1, q2, n, face)
        np.testing.assert_allclose(F, ref, rtol=2e-10, atol=2e-10)


@pytest.fixture(scope="module")
def synth():
    return LteTable()


def test_synthetic_table_thermodynamics(synth):
    """p = rho R T exactly at nodes and to interpolation accuracy between; e(T)
    independent of rho and increasing; c^2 > 0."""
    R = 287.05
    r_, T_, e_, planes, inv = synth.arrays()
    assert np.all(np.diff(e_) > 0) and np.all(np.diff(r_) > 0) and np.all(np.diff(T_) > 0)
    np.testing.assert_allclose(r_[[0, -1]], [1e-4, 10.0], rtol=1e-15)
    np.testing.assert_allclose(T_[[0, -1]], [10 ** 2.3, 10 ** 4.2], rtol=1e-14)
    for i, j in [(3, 7), (128, 200), (250, 17)]:  # at the table's nodes the interpolant is exact
        v = synth.eval(r_[i], T_[j])
        assert abs(v["pressure"] - r_[i] * R * T_[j]) <= 1e-13 * r_[i] * R * T_[j]
    for rho, T in [(0.1, 3500.0), (1e-3, 500.0), (5.0, 12000.0), (0.1, 3437.2)]:  # bicubic accuracy between
        v = synth.eval(rho, T)
        assert abs(v["pressure"] - rho * R * T) <= 1e-5 * rho * R * T
        assert v["c2"] > 0 and v["cv"] > 0
    e1 = synth.eval(0.01, 3500.0)["energy"]
    e2 = synth.eval(3.0, 3500.0)["energy"]
    assert abs(e1 - e2) <= 1e-8 * abs(e1)
    # cv against gamma_eff(T) = 1.4 - 0.2 tanh((T-4000)/1000)
    for T in (800.0, 3500.0, 4000.0, 7000.0):
        cv = R / (0.4 - 0.2 * np.tanh((T - 4000) / 1000))
        assert abs(synth.eval(0.1, T)["cv"] - cv) <= 2e-3 * cv  # second derivative of the bicubic: O(h^2)


def test_temperature_recovery_round_trip(synth):
    rng = np.random.default_rng(3)
    for _ in range(200):
        rho = 10 ** rng.uniform(-3.5, 0.8)
        T = 10 ** rng.uniform(2.35, 4.15)
        e = synth.eval(rho, T)["energy"]
        Tr, status, iters = synth.recover(rho, e)
        assert status == 0 and iters <= 20
        assert abs(Tr - T) <= 1e-9 + 1e-11 * T, (rho, T, Tr)


def test_means_reduce_to_point_values_for_equal_states(synth):
    """Equal endpoints: the mean density is rho and the mean energy is e."""
    for rho, T in [(0.1, 3500.0), (2.0, 900.0)]:
        m = synth.means(rho, T, rho, T)
        v = synth.eval(rho, T)
        np.testing.assert_allclose(m[0], rho, rtol=1e-12)
        np.testing.assert_allclose(m[1], v["energy"], rtol=1e-10)


def test_lteec_is_entropy_conservative_on_the_synthetic_table(synth):
    """Tadmor's condition for the volume flux, (w2 - w1).F = psi2 - psi1 with
    entropy variables w = (g - |u|^2/2, u, -1)/T and potential psi = p u.n/T,
    holds for the discrete-gradient means (one cell and across cells)."""
    dim = 3
    rng = np.random.default_rng(11)
    for k in range(30):
        rho1 = 0.1 * (1 + 0.05 * rng.normal())
        rho2 = rho1 * (1 + (0.002 if k % 2 else 0.05) * rng.normal())
        T1 = 3500 * (1 + 0.05 * rng.normal())
        T2 = T1 * (1 + (0.002 if k % 2 else 0.05) * rng.normal())
        u1, u2 = rng.normal(size=3) * 300, rng.normal(size=3) * 300
        n = rng.normal(size=3)
        qs, ws, psis = [], [], []
        for rho, T, u in ((rho1, T1, u1), (rho2, T2, u2)):
            v = synth.eval(rho, T)
            q = np.concatenate([[rho], rho * u, [rho * (v["energy"] + 0.5 * u @ u)]])
            w = np.concatenate([[(v["gibbs"] - 0.5 * u @ u) / T], u / T, [-1 / T]])
            qs.append(q)
            ws.append(w)
            psis.append(v["pressure"] * (u @ n) / T)
        F = synth.flux(False, dim, qs[0], qs[1], n)
        lhs = (ws[1] - ws[0]) @ F
        rhs = psis[1] - psis[0]
        scale = np.abs(ws[1] - ws[0]) @ np.abs(F) + abs(psis[0]) + abs(psis[1])
        assert abs(lhs - rhs) <= 1e-9 * scale, (k, lhs, rhs)


def _lte_case(n, order, table):
    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import build_descriptor
    from oracle.pyoracle import OracleCase

    rt, mesh, U, _ = c4_lte_tgv(n=n, order=order, table=table)
    desc, keep = build_descriptor(rt, mesh)
    c = OracleCase(desc)
    c._keep = keep
    ot = LteTable(arrays=(table.rho, table.T, table.e, table.planes.reshape(-1), table.inv.reshape(-1)))
    ot.attach(c)
    return rt, mesh, U, c


@pytest.fixture(scope="module")
def small_table():
    from theseus_amd.lte import synthetic_table

    return synthetic_table(64, 64)


def test_product_table_builder_matches_the_oracle_builder():
    """theseus_amd.lte.synthetic_table (the engine's input) and the oracle's
    builder agree: grids and forward planes bitwise, inverse guesses to 1e-12."""
    from theseus_amd.lte import synthetic_table

    t = synthetic_table(48, 40)
    o = LteTable(48, 40)
    r, T, e, planes, inv = o.arrays()
    assert np.array_equal(t.rho, r) and np.array_equal(t.T, T) and np.array_equal(t.e, e)
    np.testing.assert_allclose(t.planes.reshape(-1), planes, rtol=1e-15, atol=0)
    np.testing.assert_allclose(t.inv.reshape(-1), inv, rtol=1e-12)


def test_lte_operator_free_stream_and_conservation(small_table):
    """Uniform flow gives dU = 0 to round-off; on a periodic box the
    w_i J_i-weighted sum of dU vanishes (LTEEC is conservative)."""
    from theseus_amd.mesh import gll_nodes

    rt, mesh, U, c = _lte_case(2, 3, small_table)
    N = mesh.num_elements * 64
    q = U.reshape(5, -1)[:, 1].copy()
    Uc = np.repeat(q, N)
    r = c.rhs(Uc)
    assert np.abs(r).max() < 1e-9 * np.abs(q).max()
    r = c.rhs(U).reshape(5, -1)
    _, w = gll_nodes(3)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)
    Wall = np.tile(W, mesh.num_elements) * np.repeat(mesh.elem_jac, 64)
    for k in range(5):
        assert abs(r[k] @ Wall) <= 1e-11 * (np.abs(r[k]) @ Wall + 1e-300), k


def test_lte_operator_conserves_entropy_in_the_volume(small_table):
    """Entropy stability of the assembled LTEEC operator on a periodic box:
    with entropy variables w = ((g - |u|^2/2)/T, u/T, -1/T) of the table,
    the discrete entropy production sum_i w_i J_i (w(U_i) . dU_i) is <= 0
    (the volume terms conserve entropy, the Lax-Friedrichs faces dissipate)."""
    from theseus_amd.mesh import gll_nodes

    rt, mesh, U, c = _lte_case(2, 3, small_table)
    t = LteTable(arrays=(small_table.rho, small_table.T, small_table.e, small_table.planes.reshape(-1),
                         small_table.inv.reshape(-1)))
    Uq = U.reshape(5, -1)
    r = c.rhs(U).reshape(5, -1)
    _, w = gll_nodes(3)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1)
    Wall = np.tile(W, mesh.num_elements) * np.repeat(mesh.elem_jac, 64)
    prod = 0.0
    for nidx in range(Uq.shape[1]):
        q = Uq[:, nidx]
        u = q[1:4] / q[0]
        e = (q[4] - 0.5 * q[0] * u @ u) / q[0]
        T, st, _ = t.recover(q[0], e)
        assert st == 0
        v = t.eval(q[0], T)
        wv = np.concatenate([[(v["gibbs"] - 0.5 * u @ u) / T], u / T, [-1 / T]])
        prod += Wall[nidx] * (wv @ r[:, nidx])
    scale = np.abs(Wall) @ np.abs(r).sum(axis=0)
    assert prod <= 1e-10 * scale, (prod, scale)
