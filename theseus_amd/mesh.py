"""Affine box meshes: the geometry the reference's MFEM setup would produce
for a Cartesian box of quads/hexes (dgsem_cache_utilities.hpp:469-639),
restated in closed form.

Elements are numbered lexicographically (x fastest).  The reference
coordinate of every element is [0,1]^dim (MFEM's reference cube), so for an
element of size h the Jacobian determinant is prod(h) and the adjugate of
dx/dxi = diag(h) is diag(prod(h)/h_d).  Local faces are ordered
(-x, +x, -y, +y, -z, +z); a periodic direction wraps, a non-periodic one
gets boundary code -(bc+1).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BC_DATA_STRIDE = 8  # doubles of payload per boundary index (THESEUS_BC_DATA_STRIDE)


@dataclass
class BoxMesh:
    dim: int
    n: tuple
    lower: tuple
    upper: tuple
    periodic: tuple
    elem_jac: np.ndarray = field(repr=False)
    elem_adj: np.ndarray = field(repr=False)
    face_neighbor: np.ndarray = field(repr=False)
    bc_type: np.ndarray = field(repr=False)
    bc_data: np.ndarray = field(default=None, repr=False)  # (nbc, BC_DATA_STRIDE) payload, None = zeros

    @property
    def num_elements(self) -> int:
        return int(np.prod(self.n))

    @property
    def h(self) -> np.ndarray:
        return (np.asarray(self.upper, float) - np.asarray(self.lower, float)) / np.asarray(self.n, float)

    def element_index(self, idx) -> int:
        e = 0
        for d in reversed(range(self.dim)):
            e = e * self.n[d] + idx[d]
        return e

    def node_coordinates(self, x01: np.ndarray) -> np.ndarray:
        """Physical coordinates of every GLL node, shape (dim, ne*ndof) in the
        L-vector dof order e*ndof + (i + Np*(j + Np*k))."""
        Np = len(x01)
        dim = self.dim
        h = self.h
        lo = np.asarray(self.lower, float)
        ne = self.num_elements
        ndof = Np**dim
        coords = np.empty((dim, ne, ndof))
        eidx = np.indices(self.n[::-1]).reshape(dim, -1)[::-1]  # element multi-index, x fastest
        nidx = np.indices((Np,) * dim).reshape(dim, -1)[::-1]    # node multi-index, i fastest
        for d in range(dim):
            coords[d] = lo[d] + h[d] * (eidx[d][:, None] + x01[nidx[d]][None, :])
        return coords.reshape(dim, ne * ndof)


@dataclass
class TensorMesh(BoxMesh):
    """Cartesian tensor-product mesh with per-direction break points
    (element sizes may vary per direction and index)."""

    breaks: tuple = ()

    @property
    def h(self):
        raise AttributeError("TensorMesh has per-element sizes; use breaks")

    def node_coordinates(self, x01: np.ndarray) -> np.ndarray:
        Np = len(x01)
        dim = self.dim
        ne = self.num_elements
        ndof = Np**dim
        coords = np.empty((dim, ne, ndof))
        eidx = np.indices(self.n[::-1]).reshape(dim, -1)[::-1]
        nidx = np.indices((Np,) * dim).reshape(dim, -1)[::-1]
        for d in range(dim):
            b = np.asarray(self.breaks[d], float)
            x0 = b[eidx[d]]
            hh = b[eidx[d] + 1] - b[eidx[d]]
            coords[d] = x0[:, None] + hh[:, None] * x01[nidx[d]][None, :]
        return coords.reshape(dim, ne * ndof)


def tensor_mesh(breaks, periodic=True, bc: int = 2) -> TensorMesh:
    """Affine tensor-product mesh; breaks[d] are the element boundaries along d."""
    dim = len(breaks)
    breaks = tuple(np.asarray(b, dtype=np.float64) for b in breaks)
    n = tuple(len(b) - 1 for b in breaks)
    base = box_mesh(dim, n, [b[0] for b in breaks], [b[-1] for b in breaks], periodic=periodic, bc=bc)
    ne = base.num_elements
    eidx = np.indices(n[::-1]).reshape(dim, -1)[::-1]
    h = np.stack([np.diff(breaks[d])[eidx[d]] for d in range(dim)])  # (dim, ne)
    jac = np.prod(h, axis=0)
    adj = np.zeros((ne, dim, dim))
    for d in range(dim):
        adj[:, d, d] = jac / h[d]
    return TensorMesh(dim, n, base.lower, base.upper, base.periodic, jac, adj.reshape(ne, dim * dim).copy(),
                      base.face_neighbor, base.bc_type, breaks=breaks)


def refine_breaks(base, levels: int) -> np.ndarray:
    """Uniform refinement of each base interval by midpoint bisection,
    `levels` times (what uniform mesh refinement does to an affine segment)."""
    b = np.asarray(base, dtype=np.float64)
    for _ in range(levels):
        mid = 0.5 * (b[:-1] + b[1:])
        out = np.empty(2 * len(b) - 1)
        out[0::2] = b
        out[1::2] = mid
        b = out
    return b


def box_mesh(dim: int, n, lower, upper, periodic=True, bc: int = 2, side_bc=None) -> BoxMesh:
    """Cartesian box of n[0] x n[1] (x n[2]) affine elements.

    periodic: bool or per-direction tuple.  Non-periodic faces get boundary
    index 0 whose type is `bc` (default: supersonic outflow = transmissive).
    side_bc: optional sequence of 2*dim entries, one per box side in the
    order (x low, x high, y low, y high, z low, z high), each a boundary
    type or a (type, data) pair with up to BC_DATA_STRIDE payload values;
    every non-periodic side then gets a boundary index of its own (numbered
    in that order) and the entries of periodic sides are ignored.
    """
    n = tuple(int(v) for v in (n if np.iterable(n) else [n] * dim))
    lower = tuple(float(v) for v in (lower if np.iterable(lower) else [lower] * dim))
    upper = tuple(float(v) for v in (upper if np.iterable(upper) else [upper] * dim))
    per = tuple(bool(v) for v in (periodic if np.iterable(periodic) else [periodic] * dim))
    assert len(n) == dim and len(lower) == dim and len(upper) == dim and len(per) == dim
    ne = int(np.prod(n))
    h = (np.asarray(upper) - np.asarray(lower)) / np.asarray(n)
    J = float(np.prod(h))
    jac = np.full(ne, J)
    adj = np.zeros((ne, dim, dim))
    for d in range(dim):
        adj[:, d, d] = J / h[d]
    idx = np.indices(n[::-1]).reshape(dim, -1)[::-1]  # (dim, ne), x fastest
    strides = np.cumprod((1,) + n[:-1])
    nbr = np.empty((ne, 2 * dim), dtype=np.int32)
    any_bnd = False
    types, payload = [], []
    for d in range(dim):
        for s, step in ((0, -1), (1, +1)):
            j = idx[d] + step
            if per[d]:
                jj = j % n[d]
                nbr[:, 2 * d + s] = np.arange(ne) + (jj - idx[d]) * strides[d]
            else:
                inside = (j >= 0) & (j < n[d])
                code = -1
                if side_bc is not None:
                    entry = side_bc[2 * d + s]
                    t, data = (entry, ()) if np.isscalar(entry) else (entry[0], entry[1])
                    row = np.zeros(BC_DATA_STRIDE)
                    row[: len(data)] = data
                    types.append(int(t))
                    payload.append(row)
                    code = -len(types)
                val = np.where(inside, np.arange(ne) + step * strides[d], code)
                any_bnd |= bool((~inside).any())
                nbr[:, 2 * d + s] = val
    if side_bc is not None:
        bc_type = np.array(types, dtype=np.int32)
        bc_data = np.array(payload).reshape(-1, BC_DATA_STRIDE)
    else:
        bc_type = np.array([bc] if any_bnd else [], dtype=np.int32)
        bc_data = None
    return BoxMesh(dim, n, lower, upper, per, jac, adj.reshape(ne, dim * dim).copy(), nbr, bc_type, bc_data)


def gll_nodes(p: int):
    """GLL nodes and weights on [0,1] (float64), as the engine builds them."""
    N = p
    x = np.zeros(p + 1)
    w = np.zeros(p + 1)
    from numpy.polynomial import legendre as L

    c = np.zeros(N + 1)
    c[N] = 1
    dc = L.legder(c)
    inner = np.sort(L.legroots(dc)) if N > 1 else np.array([])
    z = np.concatenate([[-1.0], inner, [1.0]])
    PN = L.legval(z, c)
    wz = 2.0 / (N * (N + 1) * PN**2)
    x[:] = (1 + z) / 2
    w[:] = wz / 2
    x[0], x[-1] = 0.0, 1.0
    return x, w
