"""CPU oracle: numerical fluxes pinned by the reference's own flux tests
(tests/test_riemann_fluxes.cpp:64-801) and by independent closed forms."""
import math

import numpy as np
import pytest

GAMMA = 1.4
VOLUME = {"central": 0, "chandrashekar": 1, "ranocha": 2, "pirozzoli": 3}
SURFACE = {"chandrashekar": 0, "ranocha": 1, "pirozzoli": 2, "roe": 3, "hll": 4, "llf": 5}


def state(rho, vel, p, gamma=GAMMA):
    vel = np.asarray(vel, float)
    E = p / (gamma - 1) + 0.5 * rho * vel @ vel
    return np.concatenate([[rho], rho * vel, [E]])


def normal_flux(q, n, gamma=GAMMA):
    """Independent physical Euler flux f(q).n."""
    dim = len(n)
    rho = q[0]
    u = q[1:1 + dim] / rho
    E = q[-1]
    p = (gamma - 1) * (E - 0.5 * rho * u @ u)
    un = u @ n
    return np.concatenate([[rho * un], rho * u * un + p * np.asarray(n), [(E + p) * un]])


def entropy_vars(q, gamma=GAMMA):
    dim = len(q) - 2
    rho = q[0]
    u = q[1:1 + dim] / rho
    p = (gamma - 1) * (q[-1] - 0.5 * rho * u @ u)
    s = np.log(p) - gamma * np.log(rho)
    beta = rho / p
    return np.concatenate([[(gamma - s) / (gamma - 1) - 0.5 * beta * u @ u], beta * u, [-beta]])


STATES_2D = [state(1.2, [31.0, -7.0], 101325.0), state(0.8, [-10.0, -2.0], 90000.0)]
STATES_3D = [state(1.1, [12.0, -4.0, 3.0], 95000.0), state(0.7, [-5.0, 2.0, 9.0], 70000.0)]


@pytest.mark.parametrize("name", sorted(SURFACE))
@pytest.mark.parametrize("dim", [2, 3])
def test_face_flux_consistency(oracle, name, dim):
    """F*(q,q,n) = f(q).n (test_riemann_fluxes.cpp:64-83), oblique non-unit normal."""
    q = STATES_2D[0] if dim == 2 else STATES_3D[0]
    n = np.array([0.6, 0.8]) if dim == 2 else np.array([0.3, -0.5, 0.9])
    F = oracle.face_flux(SURFACE[name], dim, GAMMA, q, q, n)
    np.testing.assert_allclose(F, normal_flux(q, n), rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("name", sorted(VOLUME))
@pytest.mark.parametrize("dim", [2, 3])
def test_volume_flux_consistency(oracle, name, dim):
    q = STATES_2D[1] if dim == 2 else STATES_3D[1]
    m = np.array([0.25, -1.5]) if dim == 2 else np.array([0.25, -1.5, 0.75])
    F = oracle.volume_flux(VOLUME[name], dim, GAMMA, q, q, m, m)
    np.testing.assert_allclose(F, normal_flux(q, m), rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("name", sorted(SURFACE))
def test_face_flux_normal_reversal(oracle, name):
    """F*(qL,qR,n) = -F*(qR,qL,-n) at 1e-12 (test_riemann_fluxes.cpp:85-114)."""
    qL = state(1.0, [20.0, 3.0], 100000.0)
    qR = state(0.8, [-10.0, -2.0], 90000.0)
    n = np.array([1.0, 0.0])
    a = oracle.face_flux(SURFACE[name], 2, GAMMA, qL, qR, n)
    b = oracle.face_flux(SURFACE[name], 2, GAMMA, qR, qL, -n)
    scale = np.maximum(1.0, np.maximum(abs(a), abs(b)))
    assert np.all(abs(a + b) <= 1e-12 * scale)


@pytest.mark.parametrize("name", sorted(SURFACE))
def test_zero_normal_velocity_pressure_flux(oracle, name):
    """test_riemann_fluxes.cpp:116-135"""
    q = state(1.0, [0.0, 12.0], 100000.0)
    F = oracle.face_flux(SURFACE[name], 2, GAMMA, q, q, np.array([1.0, 0.0]))
    assert abs(F[0]) <= 1e-12 and abs(F[2]) <= 1e-12
    assert abs(F[1] - 100000.0) <= 1e-8 and abs(F[3]) <= 1e-8


@pytest.mark.parametrize("name", sorted(SURFACE))
def test_face_flux_finite_for_strong_states(oracle, name):
    qL = state(1.0, [800.0, 50.0], 101325.0)
    qR = state(0.2, [-300.0, -20.0], 20000.0)
    assert np.all(np.isfinite(oracle.face_flux(SURFACE[name], 2, GAMMA, qL, qR, np.array([0.6, 0.8]))))


@pytest.mark.parametrize("name", ["chandrashekar", "ranocha"])
def test_tadmor_entropy_conservation(oracle, name):
    """(w_R - w_L).F# = psi_R - psi_L with psi = rho u_d (test_riemann_fluxes.cpp:158-193,357-373)."""
    qL = state(1.0, [20.0, 3.0], 100000.0)
    qR = state(1.000001, [-10.0, -2.0], 90000.0)
    wL, wR = entropy_vars(qL), entropy_vars(qR)
    for d in range(2):
        m = np.zeros(2)
        m[d] = 1.0
        F = oracle.volume_flux(VOLUME[name], 2, GAMMA, qL, qR, m, m)
        assert abs(F @ (wR - wL) - (qR[1 + d] - qL[1 + d])) <= 1e-13
    # a strongly different pair in 3-D, oblique metric, relative tolerance
    qL, qR = STATES_3D
    m = np.array([0.3, -1.2, 0.5])
    F = oracle.volume_flux(VOLUME[name], 3, GAMMA, qL, qR, m, m)
    lhs = F @ (entropy_vars(qR) - entropy_vars(qL))
    rhs = (qR[1:4] - qL[1:4]) @ m
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(rhs))


def test_tadmor_fails_for_central_flux(oracle):
    """Sanity of the Tadmor check itself: the central flux is not EC."""
    qL, qR = STATES_3D
    m = np.array([0.3, -1.2, 0.5])
    F = oracle.volume_flux(VOLUME["central"], 3, GAMMA, qL, qR, m, m)
    lhs = F @ (entropy_vars(qR) - entropy_vars(qL))
    assert abs(lhs - (qR[1:4] - qL[1:4]) @ m) > 1e-6


def test_chandrashekar_dissipation_ignores_common_tangential_boost(oracle):
    """test_riemann_fluxes.cpp:304-331"""
    n = np.array([1.0, 0.0])
    a = oracle.face_flux(0, 2, GAMMA, state(1.1, [12.0, 0.0], 95000.0), state(0.9, [-7.0, 0.0], 88000.0), n)
    b = oracle.face_flux(0, 2, GAMMA, state(1.1, [12.0, 1200.0], 95000.0), state(0.9, [-7.0, 1200.0], 88000.0), n)
    assert abs(a[0] - b[0]) <= 1e-11


def test_roe_resolves_stationary_contact(oracle):
    """test_riemann_fluxes.cpp:333-355"""
    F = oracle.face_flux(3, 2, GAMMA, state(1.0, [0.0, 7.0], 100000.0), state(2.0, [0.0, 7.0], 100000.0),
                         np.array([1.0, 0.0]))
    assert abs(F[0]) <= 1e-12 and abs(F[2]) <= 1e-12
    assert abs(F[1] - 100000.0) <= 1e-8 and abs(F[3]) <= 1e-8


def _random_pairs(rng, dim, k=12):
    out = []
    for _ in range(k):
        qL = state(rng.uniform(0.3, 2), rng.uniform(-80, 80, dim), rng.uniform(2e4, 2e5))
        qR = state(rng.uniform(0.3, 2), rng.uniform(-80, 80, dim), rng.uniform(2e4, 2e5))
        n = rng.normal(size=dim) * rng.uniform(0.2, 3)
        out.append((qL, qR, n))
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_llf_and_hll_against_independent_formulas(oracle, dim):
    """LLF and HLL against direct restatements from f(q).n and the wave-speed
    bounds (test_riemann_fluxes.cpp:378-447)."""
    rng = np.random.default_rng(3 + dim)
    for qL, qR, n in _random_pairs(rng, dim):
        fL, fR = normal_flux(qL, n), normal_flux(qR, n)
        nn = np.linalg.norm(n)

        def prim(q):
            u = q[1:1 + dim] / q[0]
            p = (GAMMA - 1) * (q[-1] - 0.5 * q[0] * u @ u)
            return u @ n, math.sqrt(GAMMA * p / q[0]) * nn

        unL, cL = prim(qL)
        unR, cR = prim(qR)
        lam = max(abs(unL) + cL, abs(unR) + cR)
        expected = 0.5 * (fL + fR) - 0.5 * lam * (qR - qL)
        got = oracle.face_flux(SURFACE["llf"], dim, GAMMA, qL, qR, n)
        np.testing.assert_allclose(got, expected, rtol=1e-12, atol=1e-12 * np.abs(expected).max())
        sl, sr = min(unL - cL, unR - cR), max(unL + cL, unR + cR)
        if sl >= 0:
            hll = fL
        elif sr <= 0:
            hll = fR
        else:
            hll = (sr * fL - sl * fR + sl * sr * (qR - qL)) / (sr - sl)
        got = oracle.face_flux(SURFACE["hll"], dim, GAMMA, qL, qR, n)
        np.testing.assert_allclose(got, hll, rtol=1e-12, atol=1e-12 * np.abs(hll).max())


def _euler_jacobian(q, n):
    """Flux Jacobian d(f.n)/dq by complex-step differentiation (independent)."""
    m = len(q)
    J = np.zeros((m, m))
    h = 1e-30
    for k in range(m):
        qc = q.astype(complex)
        qc[k] += 1j * h
        dim = len(n)
        rho = qc[0]
        u = qc[1:1 + dim] / rho
        E = qc[-1]
        p = (GAMMA - 1) * (E - 0.5 * rho * (u @ u))
        un = u @ n
        f = np.concatenate([[rho * un], rho * u * un + p * np.asarray(n), [(E + p) * un]])
        J[:, k] = f.imag / h
    return J


@pytest.mark.parametrize("dim", [2, 3])
def test_roe_against_eigendecomposition(oracle, dim):
    """Roe flux = 0.5(fL+fR) - 0.5|A(q_roe)|(qR-qL) with |A| from the
    eigen-decomposition of the flux Jacobian at the Roe-averaged state
    (independent of the reference's characteristic formulas)."""
    rng = np.random.default_rng(11 + dim)
    for qL, qR, n in _random_pairs(rng, dim, k=8):
        rL, rR = math.sqrt(qL[0]), math.sqrt(qR[0])
        uL, uR = qL[1:1 + dim] / qL[0], qR[1:1 + dim] / qR[0]

        def H(q, u):
            p = (GAMMA - 1) * (q[-1] - 0.5 * q[0] * u @ u)
            return (q[-1] + p) / q[0]

        u = (rL * uL + rR * uR) / (rL + rR)
        Hh = (rL * H(qL, uL) + rR * H(qR, uR)) / (rL + rR)
        rho = rL * rR
        p = (GAMMA - 1) / GAMMA * rho * (Hh - 0.5 * u @ u)
        qroe = state(rho, u, p)
        A = _euler_jacobian(qroe, n)
        lam, R = np.linalg.eig(A)
        absA = (R @ np.diag(np.abs(lam)) @ np.linalg.inv(R)).real
        expected = 0.5 * (normal_flux(qL, n) + normal_flux(qR, n)) - 0.5 * absA @ (qR - qL)
        got = oracle.face_flux(SURFACE["roe"], dim, GAMMA, qL, qR, n)
        np.testing.assert_allclose(got, expected, rtol=1e-9, atol=1e-9 * np.abs(expected).max())


def test_log_mean_against_high_precision(oracle):
    """Both branches of ComputeLogMean (theseus_kernels.hpp:106-123) against
    (y-x)/ln(y/x) evaluated in 50-digit arithmetic."""
    import mpmath as mp

    mp.mp.dps = 50
    for x, y in [(1.0, 1.0 + 1e-9), (1.0, 1.00001), (2.0, 2.0 * 1.01), (1.0, 1.05), (0.3, 0.9), (5e-3, 7.0)]:
        exact = float((mp.mpf(y) - mp.mpf(x)) / mp.log(mp.mpf(y) / mp.mpf(x)))
        got = oracle.log_mean(x, y)
        tol = 1e-15 if abs(y / x - 1) < 0.02 else 5e-15
        assert abs(got - exact) <= tol * exact, (x, y, got, exact)
    assert oracle.log_mean(1.7, 1.7) == 1.7


def test_entropy_variables_and_gradient_conversion(oracle):
    """entropy_state (EOS.hpp:206-234) against the closed form, and
    grad_entropy_to_grad_prim (EOS.hpp:236-263) inverts the chain rule
    dS = (dS/dprim) dprim (complex-step Jacobian, independent)."""
    rng = np.random.default_rng(5)
    for dim in (2, 3):
        for _ in range(5):
            rho, p = rng.uniform(0.5, 2), rng.uniform(0.5, 3)
            u = rng.uniform(-1, 1, dim)
            q = state(rho, u, p)
            np.testing.assert_allclose(oracle.entropy_state(dim, GAMMA, q), entropy_vars(q), rtol=1e-13, atol=1e-14)
            prim = np.concatenate([[rho], u, [p]])

            def S_of_prim(pr):
                r, uu, pp = pr[0], pr[1:1 + dim], pr[-1]
                s = np.log(pp) - GAMMA * np.log(r)
                beta = r / pp
                return np.concatenate([[(GAMMA - s) / (GAMMA - 1) - 0.5 * beta * (uu @ uu)], beta * uu, [-beta]])

            Jac = np.zeros((dim + 2, dim + 2))
            for k in range(dim + 2):
                pc = prim.astype(complex)
                pc[k] += 1e-30j
                Jac[:, k] = S_of_prim(pc).imag / 1e-30
            dprim = rng.normal(size=dim + 2)
            dS = Jac @ dprim
            got = oracle.grad_entropy_to_grad_prim(dim, GAMMA, q, dS)
            np.testing.assert_allclose(got, dprim, rtol=1e-10, atol=1e-12)


def test_viscous_flux_against_stress_and_heat_flux(oracle):
    """NavierStokesFlux.hpp:389-519 against tau = mu(grad u + grad u^T - 2/3 div u I),
    q = kappa grad T, energy = u.tau + q (cf. tests/normal_viscous_face_tests.cpp:55-121)."""
    rng = np.random.default_rng(9)
    Pr, R, mu, mub = 0.71, 287.05, 0.013, 2.0 / 3.0
    cp = GAMMA / (GAMMA - 1) * R
    for dim in (2, 3):
        rho, p = 1.3, 2.0e5
        u = rng.uniform(-5, 5, dim)
        q = state(rho, u, p)
        G = rng.normal(size=(dim, dim + 2))  # G[d] = gradient of (rho,u,p) along x_d
        rows = [G[d] for d in range(dim)] + [np.zeros(dim + 2)] * (3 - dim)
        got = oracle.viscous_flux(dim, GAMMA, Pr, R, mu, mub, q, *rows)
        gu = np.array([[G[d][1 + i] for d in range(dim)] for i in range(dim)])  # gu[i,d] = du_i/dx_d
        div = np.trace(gu)
        tau = mu * (gu + gu.T - mub * div * np.eye(dim))
        gradT = np.array([(G[d][-1] - p / rho * G[d][0]) / (rho * R) for d in range(dim)])
        kappa = mu * cp / Pr
        exp = np.zeros((dim + 2, dim))
        exp[1:1 + dim] = tau
        exp[-1] = u @ tau + kappa * gradT
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12 * np.abs(exp).max())
