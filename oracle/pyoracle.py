"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU FP64 oracle
(liboracle.so, built from oracle_capi.cpp by oracle/Makefile).

Only tests/, the smoke check and the bench's cpu_baseline leg may use it, and
only as the checker."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

_dp = C.POINTER(C.c_double)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        raise FileNotFoundError(f"{LIB} not built (make -C oracle)")
    lib = C.CDLL(str(LIB))
    lib.oracle_last_error.restype = C.c_char_p
    lib.oracle_num_threads.restype = C.c_int
    lib.oracle_set_num_threads.argtypes = [C.c_int]
    lib.oracle_log_mean.argtypes = [C.c_double, C.c_double, C.c_double]
    lib.oracle_log_mean.restype = C.c_double
    lib.oracle_volume_flux.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_face_flux.argtypes = [C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    lib.oracle_physical_flux.argtypes = [C.c_int, C.c_double, _dp, _dp]
    lib.oracle_entropy_state.argtypes = [C.c_int, C.c_double, _dp, _dp]
    lib.oracle_grad_entropy_to_grad_prim.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp]
    lib.oracle_viscous_flux.argtypes = [C.c_int] + [C.c_double] * 5 + [_dp] * 5
    lib.oracle_basis.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_slipwall_pstar.argtypes = [C.c_int, C.c_double, _dp, _dp]
    lib.oracle_slipwall_pstar.restype = C.c_double
    lib.oracle_bc_inviscid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    lib.oracle_bc_viscous.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 5 + [_dp] * 7
    lib.oracle_bc_grad_jump.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
    lib.oracle_case_create.argtypes = [C.c_void_p]
    lib.oracle_case_create.restype = C.c_void_p
    lib.oracle_case_destroy.argtypes = [C.c_void_p]
    lib.oracle_rhs.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.oracle_rhs.restype = C.c_int
    lib.oracle_blending.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.oracle_estimate_stability.argtypes = [C.c_void_p, _dp, _dp]
    lib.oracle_estimate_stability.restype = C.c_int
    lib.oracle_integral_measures.argtypes = [C.c_void_p, _dp, _dp]
    lib.oracle_integral_measures.restype = C.c_int
    lib.oracle_entropy_production.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.oracle_entropy_production.restype = C.c_int
    lib.oracle_ssprk3_step.argtypes = [C.c_void_p, _dp, C.c_double]
    lib.oracle_ssprk3_step.restype = C.c_int
    _lib = lib
    return lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def arr(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def log_mean(x, y, eps=1e-4):
    return load().oracle_log_mean(float(x), float(y), float(eps))


def volume_flux(fid, dim, gamma, q1, q2, m1, m2):
    q1, q2, m1, m2 = arr(q1), arr(q2), arr(m1), arr(m2)
    F = np.zeros(dim + 2)
    load().oracle_volume_flux(fid, dim, gamma, _p(q1), _p(q2), _p(m1), _p(m2), _p(F))
    return F


def face_flux(fid, dim, gamma, q1, q2, n):
    q1, q2, n = arr(q1), arr(q2), arr(n)
    F = np.zeros(dim + 2)
    load().oracle_face_flux(fid, dim, gamma, _p(q1), _p(q2), _p(n), _p(F))
    return F


def physical_flux(dim, gamma, q):
    q = arr(q)
    F = np.zeros(15)
    load().oracle_physical_flux(dim, gamma, _p(q), _p(F))
    return F.reshape(5, 3)[: dim + 2, :dim]


def entropy_state(dim, gamma, q):
    q = arr(q)
    E = np.zeros(dim + 2)
    load().oracle_entropy_state(dim, gamma, _p(q), _p(E))
    return E


def grad_entropy_to_grad_prim(dim, gamma, q, dE):
    q, dE = arr(q), arr(dE)
    dP = np.zeros(dim + 2)
    load().oracle_grad_entropy_to_grad_prim(dim, gamma, _p(q), _p(dE), _p(dP))
    return dP


def viscous_flux(dim, gamma, Pr, R, mu, mu_bulk, q, dx, dy, dz):
    q, dx, dy, dz = arr(q), arr(dx), arr(dy), arr(dz)
    F = np.zeros(15)
    load().oracle_viscous_flux(dim, gamma, Pr, R, mu, mu_bulk, _p(q), _p(dx), _p(dy), _p(dz), _p(F))
    return F.reshape(5, 3)[: dim + 2, :dim]


def _bc_data(data):
    d = np.zeros(8)
    if data is not None:
        v = np.asarray(data, dtype=np.float64).ravel()
        d[: v.size] = v
    return d


def slipwall_pstar(dim, gamma, q, n):
    q, n = arr(q), arr(n)
    return load().oracle_slipwall_pstar(dim, gamma, _p(q), _p(n))


def bc_inviscid(face_id, bc_type, dim, gamma, q, n, data=None):
    q, n, d = arr(q), arr(n), _bc_data(data)
    F = np.zeros(dim + 2)
    load().oracle_bc_inviscid(face_id, bc_type, dim, gamma, _p(d), _p(q), _p(n), _p(F))
    return F


def bc_viscous(face_id, bc_type, dim, gamma, Pr, R, mu, mu_bulk, q, dx, dy, dz, n, data=None):
    q, dx, dy, dz, n, d = arr(q), arr(dx), arr(dy), arr(dz), arr(n), _bc_data(data)
    F = np.zeros(dim + 2)
    load().oracle_bc_viscous(face_id, bc_type, dim, gamma, Pr, R, mu, mu_bulk, _p(d), _p(q), _p(dx), _p(dy), _p(dz),
                             _p(n), _p(F))
    return F


def bc_grad_jump(bc_type, dim, gamma, R, S, data=None):
    S, d = arr(S), _bc_data(data)
    J = np.zeros(dim + 2)
    load().oracle_bc_grad_jump(bc_type, dim, gamma, R, _p(d), _p(S), _p(J))
    return J


def basis(p):
    Np = p + 1
    x, w = np.zeros(Np), np.zeros(Np)
    D, Dh, Dh2 = np.zeros(Np * Np), np.zeros(Np * Np), np.zeros(Np * Np)
    load().oracle_basis(p, _p(x), _p(w), _p(D), _p(Dh), _p(Dh2))
    return x, w, D.reshape(Np, Np), Dh.reshape(Np, Np), Dh2.reshape(Np, Np)


class OracleCase:
    """The oracle's operator for one case descriptor (theseus_case_desc)."""

    def __init__(self, desc):
        lib = load()
        self._lib = lib
        self._h = lib.oracle_case_create(C.byref(desc))
        if not self._h:
            raise RuntimeError(lib.oracle_last_error().decode())
        self.ne = desc.num_elements

    def rhs(self, U):
        U = arr(U)
        out = np.zeros_like(U)
        alpha = np.zeros(self.ne)
        rc = self._lib.oracle_rhs(self._h, _p(U), _p(out), _p(alpha))
        if rc:
            raise RuntimeError(self._lib.oracle_last_error().decode())
        self.alpha = alpha
        return out

    def blending(self, U):
        U = arr(U)
        alpha, eta = np.zeros(self.ne), np.zeros(self.ne)
        self._lib.oracle_blending(self._h, _p(U), _p(alpha), _p(eta))
        return alpha, eta

    def stability(self, U):
        U = arr(U)
        out = np.zeros(3)
        rc = self._lib.oracle_estimate_stability(self._h, _p(U), _p(out))
        if rc:
            raise RuntimeError(self._lib.oracle_last_error().decode())
        return out

    MEASURES = ("mass", "ke", "en", "max_press", "min_press", "max_temp", "min_temp", "max_dens", "min_dens")

    def integral_measures(self, U) -> dict:
        """RHSOperator::ComputeIntegralMeasures (RHSOperator_impl.hpp:530-717)."""
        U = arr(U)
        out = np.zeros(9)
        rc = self._lib.oracle_integral_measures(self._h, _p(U), _p(out))
        if rc:
            raise RuntimeError(self._lib.oracle_last_error().decode())
        return dict(zip(self.MEASURES, out.tolist()))

    def entropy_production(self, U, dU):
        """(sum w J S(U).dU, sum w J |S(U).dU|) with the case gas's entropy variables."""
        U, dU = arr(U), arr(dU)
        out = np.zeros(2)
        if self._lib.oracle_entropy_production(self._h, _p(U), _p(dU), _p(out)):
            raise RuntimeError(self._lib.oracle_last_error().decode())
        return out[0], out[1]

    def adaptive_steps(self, U, cfl, max_dt_growth, t_final, nsteps, t=0.0):
        """The reference's variable-step loop (src/Simulation.cpp:1136-1158,
        1240-1301; TimeStepControl.hpp:15-48): dt = cfl / TotalRate on the
        initial state, then per step dt_real = min(dt, time to t_final),
        step, t += dt_real, and dt = min(cfl / TotalRate, max_dt_growth * dt)
        on the new state; the loop ends once t >= t_final - 1e-8 dt.
        Returns (U, t, dt, dt_real, steps)."""
        eps = np.finfo(float).eps

        def total(r):
            return max(r[0], r[1]) + r[2]

        U = arr(U).copy()
        rate = total(self.stability(U))
        if not rate > 0.0:
            raise RuntimeError("The stability rate must be positive for variable timestepping.")
        dt = cfl / rate
        dt_real = 0.0
        steps = 0
        done = t >= t_final - 1e-8 * dt
        while not done and steps < nsteps:
            tol = 16.0 * eps * max(1.0, abs(t))
            delta = t_final - t
            dt_real = min(dt, delta) if (np.isfinite(t_final) and delta > tol) else dt
            U = self.step(U, dt_real)
            t += dt_real
            steps += 1
            rate = total(self.stability(U))
            if not rate > 0.0:
                raise RuntimeError("The stability rate must be positive for variable timestepping.")
            dt = min(cfl / rate, max_dt_growth * dt)
            done = t >= t_final - 1e-8 * dt
        return U, t, dt, dt_real, steps

    def step(self, U, dt):
        U = arr(U).copy()
        rc = self._lib.oracle_ssprk3_step(self._h, _p(U), float(dt))
        if rc:
            raise RuntimeError(self._lib.oracle_last_error().decode())
        return U

    def __del__(self):
        try:
            if self._h:
                self._lib.oracle_case_destroy(self._h)
        except Exception:
            pass


# ---- tabulated LTE gas ------------------------------------------------------
def _lte_sigs(lib):
    if getattr(lib, "_lte_ready", False):
        return lib
    vp = C.c_void_p
    lib.oracle_lte_synthetic.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4
    lib.oracle_lte_synthetic.restype = vp
    lib.oracle_lte_from_arrays.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_lte_from_arrays.restype = vp
    lib.oracle_lte_free.argtypes = [vp]
    lib.oracle_lte_arrays.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_lte_eval.argtypes = [vp, C.c_double, C.c_double, _dp]
    lib.oracle_lte_eval.restype = C.c_int
    lib.oracle_lte_recover.argtypes = [vp, C.c_double, C.c_double, _dp, C.POINTER(C.c_int)]
    lib.oracle_lte_recover.restype = C.c_int
    lib.oracle_lte_means.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    lib.oracle_lte_means.restype = C.c_int
    lib.oracle_lte_flux.argtypes = [vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_lte_flux.restype = C.c_int
    lib.oracle_case_set_lte.argtypes = [vp, vp]
    lib.oracle_lte_gas_point.argtypes = [vp, C.c_int, C.c_double, _dp, _dp, _dp]
    lib.oracle_lte_gas_point.restype = C.c_int
    lib.oracle_lte_viscous_flux.argtypes = [vp, C.c_int, C.c_double] + [_dp] * 5
    lib.oracle_lte_viscous_flux.restype = C.c_int
    lib.oracle_lte_set_sutherland.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    lib._lte_ready = True
    return lib


class LteTable:
    """The oracle's Helmholtz table: the synthetic dissociating gas, or given arrays."""

    EVAL_KEYS = ("pressure", "energy", "entropy", "gibbs", "cv", "c2", "p_rho", "p_T", "g_rho", "g_T", "H", "H_rho",
                 "H_T")

    def __init__(self, nx=256, ny=256, rho=(1e-4, 10.0), T=(10 ** 2.3, 10 ** 4.2), arrays=None):
        lib = _lte_sigs(load())
        self._lib = lib
        if arrays is None:
            self._h = lib.oracle_lte_synthetic(nx, ny, rho[0], rho[1], T[0], T[1])
        else:
            r, t, e, planes, inv = (arr(a) for a in arrays)
            nx, ny = r.size, t.size
            self._h = lib.oracle_lte_from_arrays(nx, ny, _p(r), _p(t), _p(e), _p(planes), _p(inv))
        if not self._h:
            raise RuntimeError(lib.oracle_last_error().decode())
        self.nx, self.ny = nx, ny

    def arrays(self):
        nx, ny = self.nx, self.ny
        r, t, e = np.zeros(nx), np.zeros(ny), np.zeros(ny)
        planes, inv = np.zeros(6 * nx * ny), np.zeros(nx * ny)
        self._lib.oracle_lte_arrays(self._h, _p(r), _p(t), _p(e), _p(planes), _p(inv))
        return r, t, e, planes, inv

    def eval(self, rho, T):
        out = np.zeros(13)
        if self._lib.oracle_lte_eval(self._h, float(rho), float(T), _p(out)) != 0:
            return None
        return dict(zip(self.EVAL_KEYS, out))

    def recover(self, rho, e):
        T = np.zeros(1)
        it = C.c_int()
        st = self._lib.oracle_lte_recover(self._h, float(rho), float(e), _p(T), C.byref(it))
        return float(T[0]), st, it.value

    def means(self, r1, t1, r2, t2):
        out = np.zeros(2)
        if self._lib.oracle_lte_means(self._h, r1, t1, r2, t2, _p(out)) != 0:
            return None
        return out

    def flux(self, face, dim, q1, q2, d1, d2=None):
        q1, q2, d1 = arr(q1), arr(q2), arr(d1)
        d2 = d1 if d2 is None else arr(d2)
        F = np.zeros(dim + 2)
        if self._lib.oracle_lte_flux(self._h, int(face), dim, _p(q1), _p(q2), _p(d1), _p(d2), _p(F)) != 0:
            raise RuntimeError(self._lib.oracle_last_error().decode())
        return F

    def gas_point(self, dim, q, dS=None, mu_bulk=2.0 / 3.0):
        """The tabulated gas at the conservative state q: dict of T, p, mu,
        lambda, cp, cv, the entropy variables S and the primitive gradient dP
        that the entropy-variable gradient dS converts to."""
        neq = dim + 2
        q = arr(q)
        dS = np.zeros(neq) if dS is None else arr(dS)
        out = np.zeros(6 + 2 * neq)
        if self._lib.oracle_lte_gas_point(self._h, dim, mu_bulk, _p(q), _p(dS), _p(out)) != 0:
            raise RuntimeError(self._lib.oracle_last_error().decode())
        keys = ("T", "p", "mu", "lambda", "cp", "cv")
        r = {k: out[i] for i, k in enumerate(keys)}
        r["S"] = out[6:6 + neq]
        r["dP"] = out[6 + neq:]
        return r

    def viscous_flux(self, dim, q, dx, dy, dz, mu_bulk=2.0 / 3.0):
        """Physical viscous flux [dim+2][dim] of the tabulated gas from the
        primitive gradients (energy slot = grad T)."""
        q, dx, dy, dz = arr(q), arr(dx), arr(dy), arr(dz)
        F = np.zeros(15)
        if self._lib.oracle_lte_viscous_flux(self._h, dim, mu_bulk, _p(q), _p(dx), _p(dy), _p(dz), _p(F)) != 0:
            raise RuntimeError(self._lib.oracle_last_error().decode())
        return F.reshape(5, 3)[: dim + 2, :dim]

    def set_sutherland(self, mu0, T0, Ts):
        """Viscosity by Sutherland's law instead of the mu plane."""
        self._lib.oracle_lte_set_sutherland(self._h, float(mu0), float(T0), float(Ts))

    def attach(self, case: "OracleCase"):
        """Make the oracle case use this table with the LTEEC flux."""
        self._lib.oracle_case_set_lte(case._h, self._h)
        case._lte = self

    def __del__(self):
        try:
            self._lib.oracle_lte_free(self._h)
        except Exception:
            pass
