"""Tabulated LTE gas inputs: the six-plane Helmholtz table the engine reads
(include/theseus_rhs.h, theseus_lte_table) and the synthetic dissociating
gas of the LTE Taylor-Green workload.

Table layout (LTEHelmholtzTable.hpp:13-31): planes H, H_rho, H_T, H_rhoT,
mu, lambda, each nx*ny with index T_idx*nx + rho_idx; grids rho (nx), T (ny),
and the inverse table's energy axis e (ny) with a temperature guess
inv[e_idx*nx + rho_idx] (LTEHelmholtzTableBuilder.hpp:148-216).

Synthetic gas: gamma_eff(T) = 1.4 - 0.2 tanh((T - 4000)/1000),
cv = R/(gamma_eff - 1), e(T) = int_{T0}^T cv (closed form),
s(rho, T) = int_{T0}^T cv/T dT - R ln rho (16-point Gauss-Legendre per grid
interval), H = e - T s, H_rho = R T/rho, H_T = -s, H_rhoT = R/rho.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

R_AIR = 287.05


@dataclass
class HelmholtzTable:
    rho: np.ndarray      # (nx,)
    T: np.ndarray        # (ny,)
    e: np.ndarray        # (ny,) inverse-table energy axis
    planes: np.ndarray   # (6, ny, nx)
    inv: np.ndarray      # (ny, nx) temperature guess at (rho_i, e_j)
    # viscosity by Sutherland's law (mu0, T0, Ts) instead of the mu plane
    # (LTETransport.hpp:22-43, the reference's SUTHERLAND build); None: the plane
    sutherland: tuple | None = None

    @property
    def nx(self) -> int:
        return self.rho.size

    @property
    def ny(self) -> int:
        return self.T.size


# Physics.hpp:41-42: the Sutherland constants of the reference's SUTHERLAND build
SUTHERLAND_T0 = 273.15
SUTHERLAND_TS = 110.4


def with_sutherland(table: HelmholtzTable, mu0: float = 1.716e-5, T0: float = SUTHERLAND_T0,
                    Ts: float = SUTHERLAND_TS) -> HelmholtzTable:
    """The table with its viscosity from Sutherland's law (LTETransport.hpp:22-43,
    the reference's SUTHERLAND build; mu0 is the runTime key "mu0")."""
    import dataclasses

    return dataclasses.replace(table, sutherland=(mu0, T0, Ts))


def log_grid(n: int, lo: float, hi: float) -> np.ndarray:
    """LTETable.hpp:94-102"""
    step = np.log10(hi / lo) / (n - 1)
    return np.array([lo * 10.0 ** (i * step) for i in range(n)])


def uniform_grid(n: int, lo: float, hi: float) -> np.ndarray:
    """LTETable.hpp:85-93"""
    step = (hi - lo) / (n - 1)
    return np.array([lo + i * step for i in range(n)])


class SyntheticGas:
    def __init__(self, R=R_AIR, Tc=4000.0, w=1000.0, g0=1.4, g1=0.2):
        self.R, self.Tc, self.w, self.g0, self.g1 = R, Tc, w, g0, g1

    def cv(self, T):
        return self.R / (self.g0 - 1 - self.g1 * np.tanh((T - self.Tc) / self.w))

    def energy(self, T, T0):
        """int_{T0}^T cv dT: with x = (T - Tc)/w, cv = R/(a - b tanh x) and
        int dx/(a - b tanh x) = (a x + b ln(a cosh x - b sinh x))/(a^2 - b^2)."""
        a, b = self.g0 - 1, self.g1

        def F(TT):
            x = (TT - self.Tc) / self.w
            ax = abs(x)
            sg = 1.0 if x >= 0 else -1.0
            lg = ax + np.log(0.5 * ((a - b * sg) + (a + b * sg) * np.exp(-2 * ax)))
            return self.R * self.w * (a * x + b * lg) / (a * a - b * b)

        return F(T) - F(T0)


_GL16_X = np.array([0.0950125098376374, 0.2816035507792589, 0.4580167776572274, 0.6178762444026438,
                    0.7554044083550030, 0.8656312023878318, 0.9445750230732326, 0.9894009349916499])
_GL16_W = np.array([0.1894506104550685, 0.1826034150449236, 0.1691565193950025, 0.1495959888165767,
                    0.1246289712555339, 0.0951585116824928, 0.0622535239386479, 0.0271524594117541])


def entropy_integral(gas: SyntheticGas, T: np.ndarray) -> np.ndarray:
    s = np.zeros_like(T)
    for j in range(1, T.size):
        a, b = T[j - 1], T[j]
        c, h = 0.5 * (a + b), 0.5 * (b - a)
        acc = 0.0
        for k in range(8):
            t1, t2 = c - h * _GL16_X[k], c + h * _GL16_X[k]
            acc += _GL16_W[k] * (gas.cv(t1) / t1 + gas.cv(t2) / t2)
        s[j] = s[j - 1] + h * acc
    return s


def _hermite(x):
    """Cubic Hermite position and slope bases (value and first derivative) on [0,1]."""
    x2 = x * x
    x3 = x2 * x
    p0, p1 = 2 * x3 - 3 * x2 + 1, -2 * x3 + 3 * x2
    s0, s1 = x3 - 2 * x2 + x, x3 - x2
    dp0, dp1 = 6 * x2 - 6 * x, -6 * x2 + 6 * x
    ds0, ds1 = 3 * x2 - 4 * x + 1, 3 * x2 - 2 * x
    return (p0, p1, s0, s1), (dp0, dp1, ds0, ds1)


def _energy_at(t: HelmholtzTable, i: int, rho: float, T: np.ndarray) -> np.ndarray:
    """Interpolated e = H - T H_T at density rho (in cell i) for temperatures T
    (vectorised over T): LTEHelmholtzInterpolation.hpp:136-172, LTEHelmholtzEOS.hpp:33-36."""
    j = np.clip(np.searchsorted(t.T, T, side="right") - 1, 0, t.ny - 2)
    dr = t.rho[i + 1] - t.rho[i]
    dt = t.T[j + 1] - t.T[j]
    xi = (rho - t.rho[i]) / dr
    eta = (T - t.T[j]) / dt
    (rp0, rp1, rs0, rs1), _ = _hermite(np.float64(xi))
    (tp0, tp1, ts0, ts1), (dtp0, dtp1, dts0, dts1) = _hermite(eta)
    H, Hr, HT, HrT = t.planes[0], t.planes[1], t.planes[2], t.planes[3]
    h = np.zeros_like(T)
    hT = np.zeros_like(T)
    for ds, (rv, rs) in enumerate(((rp0, rs0), (rp1, rs1))):
        for ts, (tv, tsl, dtv, dtsl) in enumerate(((tp0, ts0, dtp0, dts0), (tp1, ts1, dtp1, dts1))):
            jj, ii = j + ts, i + ds
            c0 = H[jj, ii]
            c1 = Hr[jj, ii] * dr
            c2 = HT[jj, ii] * dt
            c3 = HrT[jj, ii] * dr * dt
            h += c0 * rv * tv + c1 * rs * tv + c2 * rv * tsl + c3 * rs * tsl
            hT += c0 * rv * dtv + c1 * rs * dtv + c2 * rv * dtsl + c3 * rs * dtsl
    hT /= dt
    return h - T * hT


def synthetic_table(nx: int = 256, ny: int = 256, rho=(1e-4, 10.0), T=(10 ** 2.3, 10 ** 4.2),
                    gas: SyntheticGas | None = None, rho_dist: str = "log", T_dist: str = "log") -> HelmholtzTable:
    """The LTE Taylor-Green workload's table: log-spaced rho and T grids, the
    forward planes from the synthetic gas, and the inverse table built as the
    reference builds it (energy axis uniform over the padded range of
    H - T H_T, 60 bisection halvings of the interpolated energy per density
    column, endpoint temperatures outside a column's range)."""
    gas = gas or SyntheticGas()
    rg = (log_grid if rho_dist == "log" else uniform_grid)(nx, *rho)
    Tg = (log_grid if T_dist == "log" else uniform_grid)(ny, *T)
    sT = entropy_integral(gas, Tg)
    e_of_T = np.array([gas.energy(t, Tg[0]) for t in Tg])
    s = sT[:, None] - gas.R * np.log(rg)[None, :]
    H = e_of_T[:, None] - Tg[:, None] * s
    planes = np.empty((6, ny, nx))
    planes[0] = H
    planes[1] = gas.R * Tg[:, None] / rg[None, :]
    planes[2] = -s
    planes[3] = gas.R / rg[None, :] * np.ones((ny, 1))
    mu = 1.8e-5 * (Tg / 300.0) ** 0.7
    planes[4] = mu[:, None] * np.ones((1, nx))
    planes[5] = (mu * (gas.cv(Tg) + gas.R) / 0.72)[:, None] * np.ones((1, nx))
    tht = Tg[:, None] * planes[2]
    en = H - tht
    scale = max(1.0, float(np.max(np.abs(H) + np.abs(tht))))
    pad = 64 * np.finfo(float).eps * scale
    emin, emax = float(en.min()) - pad, float(en.max()) + pad
    step = (emax - emin) / (ny - 1)
    e = np.array([emin + j * step for j in range(ny)])
    t = HelmholtzTable(rg, Tg, e, planes, np.zeros((ny, nx)))
    inv = np.empty((ny, nx))
    for i in range(nx):
        ic = min(i, nx - 2)
        lowT, highT = Tg[0], Tg[-1]
        lowE = _energy_at(t, ic, rg[i], np.array([lowT]))[0]
        highE = _energy_at(t, ic, rg[i], np.array([highT]))[0]
        lo = np.full(ny, lowT)
        hi = np.full(ny, highT)
        below = e <= lowE
        above = e >= highE
        hi[below] = lo[below]
        lo[above] = hi[above]
        mid_mask = ~(below | above)
        for _ in range(60):
            mid = lo + (hi - lo) / 2
            en_mid = _energy_at(t, ic, rg[i], mid)
            go_up = (en_mid < e) & mid_mask
            go_dn = (~(en_mid < e)) & mid_mask
            lo = np.where(go_up, mid, lo)
            hi = np.where(go_dn, mid, hi)
        inv[:, i] = lo + (hi - lo) / 2
    t.inv = inv
    return t
