"""Entropy production of a right-hand side (the FP64 check that the engine's
theseus_rhs_entropy_production computes, restated in oracle/oracle_operator.hpp),
pinned by the chain rule: the entropy variables are the gradient of the
mathematical entropy eta(U) = -rho s/(gamma-1), s = ln p - gamma ln rho
(EOS.hpp:206-234), so sum w J S(U).V equals d/de sum w J eta(U + eV), here
by a central difference of eta written independently in numpy.  An
entropy-stable scheme (Ranocha volume flux, Lax-Friedrichs faces, BR1 viscous
terms) on a periodic mesh produces no entropy: the production is <= 0.
For the tabulated gas the entropy density is rho H_T(rho, T) = -rho s, with T
recovered from the internal energy and s from the table's interpolant."""
import numpy as np
import pytest

from oracle.pyoracle import OracleCase
from theseus_amd import _abi as A
from theseus_amd.config import RunTime
from theseus_amd.mesh import box_mesh, gll_nodes
from theseus_amd.operator import build_descriptor


def _case(dim, n, order, flow, vol=A.VOLUME_RANOCHA, surf=A.SURFACE_LLF):
    rt = RunTime(dim=dim, order=order, flow_model=flow, volume_flux=vol, surface_flux=surf, mu=0.01,
                 blending=False)
    mesh = box_mesh(dim, n, 0.0, 2 * np.pi, periodic=True)
    desc, keep = build_descriptor(rt, mesh)
    c = OracleCase(desc)
    c._keep = keep
    return rt, mesh, c


def _state(mesh, order, dim, rng):
    x01, _ = gll_nodes(order)
    X = mesh.node_coordinates(x01)
    x, y = X[0], X[1]
    z = X[2] if dim == 3 else 0.0 * x
    rho = 1.0 + 0.2 * np.sin(x + 0.3) * np.cos(y) + 0.1 * np.cos(z)
    vel = [np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z), 0.3 * np.sin(z)][:dim]
    p = 10.0 + np.cos(2 * x) * np.cos(y) + 0.5 * np.sin(z)
    E = p / 0.4 + 0.5 * rho * sum(v * v for v in vel)
    return np.concatenate([rho[None], np.array([rho * v for v in vel]), E[None]], axis=0).reshape(-1)


def _weights(mesh, order, dim):
    _, w = gll_nodes(order)
    W = np.einsum("i,j,k->kji", w, w, w).reshape(-1) if dim == 3 else np.einsum("i,j->ji", w, w).reshape(-1)
    return np.tile(W, mesh.num_elements) * np.repeat(mesh.elem_jac, W.size)


def _total_entropy(U, Wall, dim, gamma=1.4):
    q = U.reshape(dim + 2, -1)
    rho = q[0]
    ke = 0.5 * (q[1:dim + 1] ** 2).sum(axis=0) / rho
    p = (gamma - 1) * (q[dim + 1] - ke)
    s = np.log(p) - gamma * np.log(rho)
    return Wall @ (-rho * s / (gamma - 1))


@pytest.mark.parametrize("dim,n,order", [(2, 3, 3), (3, 2, 3), (3, 2, 2)])
def test_entropy_production_is_the_derivative_of_the_total_entropy(dim, n, order):
    rng = np.random.default_rng(3)
    rt, mesh, c = _case(dim, n, order, A.FLOW_NAVIER_STOKES)
    U = _state(mesh, order, dim, rng)
    Wall = _weights(mesh, order, dim)
    for V in (c.rhs(U), rng.uniform(-1, 1, U.size) * np.repeat(np.abs(U.reshape(dim + 2, -1)).max(axis=1),
                                                                  U.size // (dim + 2))):
        P, M = c.entropy_production(U, V)
        eps = 1e-5
        fd = (_total_entropy(U + eps * V, Wall, dim) - _total_entropy(U - eps * V, Wall, dim)) / (2 * eps)
        assert M > 0
        assert abs(P - fd) <= 1e-7 * M, (P, fd, M)


@pytest.mark.parametrize("dim", [2, 3])
def test_entropy_conservative_volume_terms_conserve_entropy(dim):
    """A continuous state (shared face nodes equal) has no face jumps, so the
    Lax-Friedrichs dissipation vanishes and the Ranocha volume terms conserve
    the total entropy to rounding (Tadmor's condition, summed)."""
    rt, mesh, c = _case(dim, 3 if dim == 2 else 2, 3, A.FLOW_EULER)
    U = _state(mesh, 3, dim, None)
    P, M = c.entropy_production(U, c.rhs(U))
    assert abs(P) <= 1e-12 * M, (P, M)


@pytest.mark.parametrize("flow", [A.FLOW_EULER, A.FLOW_NAVIER_STOKES])
@pytest.mark.parametrize("dim", [2, 3])
def test_entropy_stable_scheme_dissipates(flow, dim):
    """With jumps across the faces (each node perturbed) the faces dissipate,
    and with viscosity so do the viscous terms: production < 0."""
    rng = np.random.default_rng(5)
    rt, mesh, c = _case(dim, 3 if dim == 2 else 2, 3, flow)
    U = _state(mesh, 3, dim, rng).reshape(dim + 2, -1)
    U = (U * (1 + 0.01 * rng.uniform(-1, 1, U.shape[1]))).reshape(-1)
    P, M = c.entropy_production(U, c.rhs(U))
    assert P < -1e-6 * M, (P, M)


def test_tabulated_gas_entropy_production_is_the_derivative_of_its_total_entropy():
    """The same chain rule for the tabulated gas (LTEHelmholtzGasEOS.hpp:144-157):
    eta = -rho s(rho, T) with T recovered from e and s = -H_T of the
    interpolant, node by node through the oracle's table (a 2^3 p=2 case of
    the LTE Taylor-Green vortex on a 64x64 synthetic table); and the LTEEC
    operator on that periodic box dissipates (production < 0)."""
    from oracle.pyoracle import LteTable
    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.lte import synthetic_table

    table = synthetic_table(64, 64)
    rt, mesh, U, _ = c4_lte_tgv(n=2, order=2, table=table)
    desc, keep = build_descriptor(rt, mesh)
    c = OracleCase(desc)
    c._keep = keep
    t = LteTable(arrays=(table.rho, table.T, table.e, table.planes.reshape(-1), table.inv.reshape(-1)))
    t.attach(c)
    Wall = _weights(mesh, 2, 3)

    def total(Ux):
        q = Ux.reshape(5, -1)
        tot = 0.0
        for k in range(q.shape[1]):
            rho = q[0, k]
            e = (q[4, k] - 0.5 * (q[1:4, k] @ q[1:4, k]) / rho) / rho
            T, st, _ = t.recover(rho, e)
            assert st == 0
            tot += Wall[k] * (-rho * t.eval(rho, T)["entropy"])
        return tot

    rng = np.random.default_rng(11)
    V = rng.uniform(-1, 1, U.size) * np.repeat(np.abs(U.reshape(5, -1)).max(axis=1), U.size // 5) * 1e-3
    P, M = c.entropy_production(U, V)
    eps = 1e-4
    fd = (total(U + eps * V) - total(U - eps * V)) / (2 * eps)
    assert abs(P - fd) <= 1e-6 * M, (P, fd, M)
    Up = (U.reshape(5, -1) * (1 + 1e-3 * rng.uniform(-1, 1, U.size // 5))).reshape(-1)
    P, M = c.entropy_production(Up, c.rhs(Up))
    assert P < -1e-6 * M, (P, M)
