"""Mesh readers for the reference's case files, producing the engine's
per-element affine geometry (Jacobian, adjugate rows) and face
connectivity.

Formats (the two the reference's test cases use, src/Simulation.cpp:315-331):
  * MFEM mesh v1.0 with quadrilaterals or hexahedra, optionally periodic
    through an L2 ``nodes`` section (one coordinate set per element, in the
    element's lexicographic corner order), e.g.
    TestCases/Euler/2D/IsentropicVortex/IsentropicVortex.mesh;
  * Gmsh 2.2 ASCII with quadrilaterals or hexahedra, physical tags on the
    boundary elements, and a ``$Periodic`` section (slave -> master node
    pairs), e.g. TestCases/NavierStokes/*/TaylorGreenVortex/*.msh.

The engine evaluates every interior face from both sides with the same
tangential point order, so every element is re-oriented (a symmetry of the
reference square/cube applied to its corner list) such that its face (d, 1)
and its neighbour's face (d, 0) carry the same vertices in the same
tangential order.  Orientations propagate breadth-first from one element
per connected component; a mesh that cannot be oriented that way (an
unstructured vertex) raises ValueError.  Elements must be affine
(parallelograms / parallelepipeds) to a relative tolerance.

``uniform_refine`` splits every element into 2^dim children
(``ser_ref_levels``, src/Simulation.cpp:343-353), ``hilbert_order`` gives an
element permutation along a Hilbert curve of the element centres
(``mesh_ordering: Hilbert``, :363-374; the permutation is this module's own
and need not equal MFEM's), ``reorder`` applies one.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .mesh import BC_DATA_STRIDE

# Gmsh 2.2 / MFEM corner order -> lexicographic corner order (x fastest)
_QUAD_TO_LEX = (0, 1, 3, 2)               # ccw (0,0),(1,0),(1,1),(0,1)
_HEX_TO_LEX = (0, 1, 3, 2, 4, 5, 7, 6)    # bottom ccw, then top ccw


@dataclass
class AffineMesh:
    """Unstructured conforming mesh of affine quads/hexes, oriented for the
    engine.  Duck-types theseus_amd.mesh.BoxMesh where the operator and the
    initial conditions use it."""

    dim: int
    corners: np.ndarray            # (ne, 2^dim, dim) lexicographic corner coordinates
    face_neighbor: np.ndarray      # (ne, 2*dim) int32: element, or -(bc+1) on a boundary
    bc_type: np.ndarray            # (nbc,) int32 engine boundary type per boundary index
    bc_attributes: tuple = ()      # file attribute (physical tag) of each boundary index
    elem_attr: np.ndarray = field(default=None, repr=False)
    vertex_ids: np.ndarray = field(default=None, repr=False)  # (ne, 2^dim) after periodic identification
    bc_data: np.ndarray = field(default=None, repr=False)     # (nbc, BC_DATA_STRIDE) boundary payload
    # per element, the code (index into _symmetries, the engine's
    # theseus_case_desc.elem_orient convention) of the symmetry that maps the
    # file's corner order onto the aligned one; None after refinement
    elem_orient: np.ndarray = field(default=None, repr=False)

    @property
    def num_elements(self) -> int:
        return int(self.corners.shape[0])

    @property
    def n(self):
        return (self.num_elements,)

    @property
    def origin(self) -> np.ndarray:
        return self.corners[:, 0, :]

    @property
    def edges(self) -> np.ndarray:
        """(ne, dim, dim): edge vector of local axis d (row d)."""
        dim = self.dim
        return np.stack([self.corners[:, 1 << d, :] - self.corners[:, 0, :] for d in range(dim)], axis=1)

    @property
    def elem_jac(self) -> np.ndarray:
        return np.linalg.det(self.edges)

    @property
    def elem_adj(self) -> np.ndarray:
        """Row rd = J * d(xi_rd)/dx: the adjugate of the Jacobian matrix whose
        columns are the edge vectors, flattened (ne, dim*dim)."""
        A = np.transpose(self.edges, (0, 2, 1))  # columns = edges
        J = np.linalg.det(A)
        inv = np.linalg.inv(A)                   # row rd = d(xi_rd)/dx
        return (inv * J[:, None, None]).reshape(self.num_elements, self.dim * self.dim)

    @property
    def lower(self):
        return tuple(self.corners.reshape(-1, self.dim).min(axis=0))

    @property
    def upper(self):
        return tuple(self.corners.reshape(-1, self.dim).max(axis=0))

    @property
    def periodic(self):
        return tuple(bool(not (self.face_neighbor[:, 2 * d:2 * d + 2] < 0).any()) for d in range(self.dim))

    def node_coordinates(self, x01: np.ndarray) -> np.ndarray:
        """Physical coordinates of every GLL node, (dim, ne*ndof), L-vector order."""
        Np = len(x01)
        dim = self.dim
        nidx = np.indices((Np,) * dim).reshape(dim, -1)[::-1]  # node multi-index, i fastest
        xi = x01[nidx]                                           # (dim, ndof)
        E = self.edges                                           # (ne, dim_ref, dim_phys)
        X = self.origin[:, None, :] + np.einsum("rn,erp->enp", xi, E)
        return np.transpose(X, (2, 0, 1)).reshape(dim, -1)


# ---------------------------------------------------------------------------
# hypercube symmetries as corner permutations
# ---------------------------------------------------------------------------
def _symmetries(dim: int):
    """Every symmetry of [0,1]^dim as (perm, flip): new local axis k is old
    axis perm[k], reversed when flip[k]; returned with the corner map
    new lexicographic corner -> old lexicographic corner."""
    out = []
    for perm in itertools.permutations(range(dim)):
        for flip in itertools.product((0, 1), repeat=dim):
            cmap = []
            for c in range(1 << dim):
                bits_new = [(c >> k) & 1 for k in range(dim)]
                old = 0
                for k in range(dim):
                    b = bits_new[k] ^ flip[k]
                    old |= b << perm[k]
                cmap.append(old)
            out.append(np.array(cmap))
    return out


def _face(dim: int, d: int, s: int):
    """Lexicographic corners of face (d, s), in tangential lexicographic order."""
    return [c for c in range(1 << dim) if ((c >> d) & 1) == s]


def _orient(dim: int, vids: np.ndarray):
    """Re-order each element's corner list so that neighbouring elements see
    their shared face with the same tangential order.  Returns (corner
    permutation per element (ne, 2^dim), face_neighbor with -1 on boundary
    faces)."""
    ne = vids.shape[0]
    syms = _symmetries(dim)
    faces = [[_face(dim, d, s) for s in (0, 1)] for d in range(dim)]
    # face key -> list of (element, file-local face given as sorted vertex tuple)
    owners: dict = {}
    for e in range(ne):
        for d in range(dim):
            for s in (0, 1):
                key = tuple(sorted(vids[e, faces[d][s]]))
                owners.setdefault(key, []).append(e)
    for key, els in owners.items():
        if len(els) > 2:
            raise ValueError(f"non-conforming or non-manifold face {key}: {len(els)} elements")
    perm = np.full((ne, 1 << dim), -1, dtype=np.int64)
    nbr = np.full((ne, 2 * dim), -1, dtype=np.int64)
    comp = np.full(ne, -1, dtype=np.int64)
    done = np.zeros(ne, dtype=bool)

    def oriented(e):
        return vids[e, perm[e]]

    for start in range(ne):
        if done[start]:
            continue
        perm[start] = syms[0]
        done[start] = True
        queue = [start]
        while queue:
            e = queue.pop()
            comp[e] = start
            ov = oriented(e)
            for d in range(dim):
                for s in (0, 1):
                    fv = ov[faces[d][s]]
                    key = tuple(sorted(fv))
                    others = [x for x in owners[key] if x != e]
                    if not others:
                        if len(owners[key]) == 2:  # an element touching itself across a periodic face
                            others = [e]
                        else:
                            continue
                    f = others[0]
                    if f == e and len(owners[key]) == 1:
                        continue
                    target = list(fv)
                    if done[f]:
                        got = list(oriented(f)[faces[d][1 - s]])
                        if got != target:
                            raise ValueError(f"elements {e} and {f} cannot be oriented consistently "
                                             "(the engine needs a structured face orientation)")
                    else:
                        for cm in syms:
                            if list(vids[f, cm][faces[d][1 - s]]) == target:
                                perm[f] = cm
                                break
                        else:
                            raise ValueError(f"no orientation of element {f} matches face ({d},{s}) of {e}")
                        done[f] = True
                        queue.append(f)
                    nbr[e, 2 * d + s] = f
    return perm, nbr, comp


def storage_frames(mesh: AffineMesh):
    """The file's own element frames: (corner vertex ids (ne, 2^dim) and
    corner coordinates (ne, 2^dim, dim), both in each element's file corner
    order, lexicographic)."""
    syms = _symmetries(mesh.dim)
    ne, nc = mesh.num_elements, 1 << mesh.dim
    vids = np.empty((ne, nc), dtype=np.int64)
    xyz = np.empty_like(mesh.corners)
    for e in range(ne):
        cm = syms[int(mesh.elem_orient[e])]
        vids[e, cm] = mesh.vertex_ids[e]
        xyz[e, cm] = mesh.corners[e]
    return vids, xyz


def node_storage_index(dim: int, order: int, code: int) -> np.ndarray:
    """Storage index within its element of every aligned node for a symmetry
    code (the engine's theseus_orient_node)."""
    perm, flip = list(itertools.permutations(range(dim)))[code >> dim], [(code >> (dim - 1 - k)) & 1 for k in range(dim)]
    Np = order + 1
    a = np.indices((Np,) * dim).reshape(dim, -1)[::-1]  # aligned multi-index, axis 0 fastest
    s = np.empty_like(a)
    for k in range(dim):
        s[perm[k]] = np.where(flip[k], Np - 1 - a[k], a[k])
    return sum(s[k] * Np ** k for k in range(dim))


def to_storage(U: np.ndarray, mesh: AffineMesh, order: int) -> np.ndarray:
    """An aligned L-vector rearranged into the elements' file frames."""
    return _permute_state(U, mesh, order, True)


def to_aligned(U: np.ndarray, mesh: AffineMesh, order: int) -> np.ndarray:
    return _permute_state(U, mesh, order, False)


def _permute_state(U, mesh, order, to_st):
    dim, ne = mesh.dim, mesh.num_elements
    nd = (order + 1) ** dim
    V = np.asarray(U).reshape(-1, ne, nd)
    out = np.empty_like(V)
    for c in np.unique(mesh.elem_orient):
        sel = mesh.elem_orient == c
        idx = node_storage_index(dim, order, int(c))
        if to_st:
            tmp = np.empty_like(V[:, sel])
            tmp[:, :, idx] = V[:, sel]
            out[:, sel] = tmp
        else:
            out[:, sel] = V[:, sel][:, :, idx]
    return out.reshape(np.shape(U))


def _finish(dim, vids, coords_lex, boundary, bc_types, attrs, rtol=1e-9):
    """Orient, check affinity, attach boundary indices."""
    perm, nbr, comp = _orient(dim, vids)
    ne = vids.shape[0]
    rows = np.arange(ne)[:, None]
    # a component oriented from a mirrored first element comes out with
    # negative Jacobians throughout: reversing local axis 0 everywhere in it
    # keeps every face pairing and makes them positive
    c0 = coords_lex[rows, perm]
    det = np.linalg.det(np.stack([c0[:, 1 << d] - c0[:, 0] for d in range(dim)], axis=1))
    for c in np.unique(comp):
        sel = comp == c
        if (det[sel] < 0).all():
            perm[sel] = perm[sel][:, [k ^ 1 for k in range(1 << dim)]]
            nbr[np.ix_(sel, [0, 1])] = nbr[np.ix_(sel, [1, 0])]
        elif (det[sel] < 0).any():
            raise ValueError("element orientations disagree with the geometry (mixed Jacobian signs)")
    corners = coords_lex[rows, perm]
    ov = vids[rows, perm]
    # affinity: every corner is origin + sum of its edge vectors
    E = np.stack([corners[:, 1 << d] - corners[:, 0] for d in range(dim)], axis=1)
    h = np.linalg.norm(E, axis=2).max(axis=1)
    for c in range(1 << dim):
        pred = corners[:, 0] + sum(((c >> d) & 1) * E[:, d] for d in range(dim))
        dev = np.linalg.norm(corners[:, c] - pred, axis=1) / h
        if (dev > rtol).any():
            e = int(np.argmax(dev))
            raise ValueError(f"element {e} is not affine (corner {c} off by {dev[e]:.2e} of its size)")
    if (np.linalg.det(E) <= 0).any():
        raise ValueError("an element has non-positive Jacobian after orientation")
    # boundary faces: file attribute -> boundary index
    bc_attr = sorted({a for a in boundary.values()})
    faces = [[_face(dim, d, s) for s in (0, 1)] for d in range(dim)]
    for e in range(ne):
        for d in range(dim):
            for s in (0, 1):
                if nbr[e, 2 * d + s] >= 0:
                    continue
                key = tuple(sorted(ov[e, faces[d][s]]))
                a = boundary.get(key)
                if a is None:
                    if not bc_attr or 0 not in bc_attr:
                        bc_attr.append(0)
                    a = 0
                nbr[e, 2 * d + s] = -(bc_attr.index(a) + 1)
    # bc_types: one type for every boundary, or attribute -> type or (type, payload)
    entries = [bc_types.get(a, 2) if isinstance(bc_types, dict) else bc_types for a in bc_attr]
    types = np.array([int(v if np.isscalar(v) else v[0]) for v in entries], dtype=np.int32)
    data = np.zeros((len(entries), BC_DATA_STRIDE))
    for i, v in enumerate(entries):
        if not np.isscalar(v):
            data[i, : len(v[1])] = v[1]
    if not (nbr < 0).any():
        types = np.zeros(0, dtype=np.int32)
        data = np.zeros((0, BC_DATA_STRIDE))
    index = {tuple(cm): k for k, cm in enumerate(_symmetries(dim))}
    codes = np.array([index[tuple(perm[e])] for e in range(ne)], dtype=np.uint8)
    return AffineMesh(dim, np.ascontiguousarray(corners, dtype=np.float64), nbr.astype(np.int32), types,
                      tuple(bc_attr), np.asarray(attrs), ov, data, codes)


# ---------------------------------------------------------------------------
# MFEM mesh v1.0
# ---------------------------------------------------------------------------
def _tokens(path):
    for line in Path(path).read_text().splitlines():
        line = line.split("#", 1)[0].strip()
        if line:
            yield line


def read_mfem(path, bc_types=2) -> AffineMesh:
    """MFEM mesh v1.0 (quads: geometry 3, hexes: 5); `bc_types` maps a
    boundary attribute to the engine's boundary type (or one type for all)."""
    lines = list(_tokens(path))
    if not lines or not lines[0].startswith("MFEM mesh v1.0"):
        raise ValueError("not an MFEM mesh v1.0 file")
    it = iter(lines[1:])
    dim = None
    elems, attrs, bnd, verts, nodes = [], [], {}, None, None
    for line in it:
        if line == "dimension":
            dim = int(next(it))
        elif line == "elements":
            for _ in range(int(next(it))):
                v = [int(x) for x in next(it).split()]
                geom = v[1]
                if geom not in (3, 5):
                    raise ValueError(f"element geometry {geom} is not a quadrilateral or hexahedron")
                order = _QUAD_TO_LEX if geom == 3 else _HEX_TO_LEX
                attrs.append(v[0])
                elems.append([v[2:][k] for k in order])
        elif line == "boundary":
            for _ in range(int(next(it))):
                v = [int(x) for x in next(it).split()]
                bnd[tuple(sorted(v[2:]))] = v[0]
        elif line == "vertices":
            nv = int(next(it))
            nxt = next(it)
            if nxt == "nodes":  # coordinates come from the nodes section
                nodes = _mfem_nodes(it, len(elems), dim)
            else:
                sdim = int(nxt)  # space dimension, then one vertex per line
                verts = np.array([[float(x) for x in next(it).split()][:sdim] for _ in range(nv)])
        elif line == "nodes":
            nodes = _mfem_nodes(it, len(elems), dim)
    if dim not in (2, 3):
        raise ValueError(f"dimension {dim} not supported (2 or 3)")
    vids = np.array(elems, dtype=np.int64)
    if nodes is not None:
        coords_lex = nodes
    else:
        coords_lex = verts[vids][..., :dim]
    return _finish(dim, vids, coords_lex, bnd, bc_types, attrs)


def _mfem_nodes(it, ne, dim):
    """L2 (discontinuous) order-1 nodes: one corner set per element in its
    lexicographic order (FiniteElementCollection L2_T*_{2,3}D_P1)."""
    header = {}
    for _ in range(4):
        line = next(it)
        if ":" in line:
            k, v = line.split(":", 1)
            header[k.strip()] = v.strip()
        else:
            header[line] = ""
    fec = header.get("FiniteElementCollection", "")
    if not fec.startswith("L2_") or not fec.endswith("_P1"):
        raise ValueError(f"nodes collection {fec!r}: only L2 order-1 (periodic) nodes are read")
    ordering = int(header.get("Ordering", "1"))
    nc = 1 << dim
    vals = []
    while len(vals) < ne * nc * dim:
        vals.extend(float(x) for x in next(it).split())
    a = np.array(vals[:ne * nc * dim])
    if ordering == 1:   # byVDIM: x y (z) per node
        return a.reshape(ne, nc, dim)
    return a.reshape(dim, ne, nc).transpose(1, 2, 0)  # byNODES


# ---------------------------------------------------------------------------
# Gmsh 2.2 ASCII
# ---------------------------------------------------------------------------
def read_gmsh22(path, bc_types=2) -> AffineMesh:
    """Gmsh 2.2 ASCII with quads (type 3) or hexes (type 5); boundary
    elements (lines, type 1; quads in 3-D) carry the physical tag used as
    the boundary attribute; $Periodic node pairs identify slave nodes with
    their masters."""
    text = Path(path).read_text().split("\n")
    sec = {}
    i = 0
    while i < len(text):
        t = text[i].strip()
        if t.startswith("$") and not t.startswith("$End"):
            name = t[1:]
            j = i + 1
            while not text[j].strip().startswith("$End" + name):
                j += 1
            sec[name] = [s for s in text[i + 1:j] if s.strip()]
            i = j
        i += 1
    fmt = sec.get("MeshFormat", ["0"])[0].split()
    if not fmt or not fmt[0].startswith("2.2") or (len(fmt) > 1 and fmt[1] != "0"):
        raise ValueError("not a Gmsh 2.2 ASCII file")
    nn = int(sec["Nodes"][0])
    ids = np.empty(nn, dtype=np.int64)
    xyz = np.empty((nn, 3))
    for k, line in enumerate(sec["Nodes"][1:nn + 1]):
        v = line.split()
        ids[k] = int(v[0])
        xyz[k] = [float(x) for x in v[1:4]]
    index = {int(n): k for k, n in enumerate(ids)}
    # periodic identification (slave -> master, followed transitively)
    master = {}
    if "Periodic" in sec:
        lines = sec["Periodic"]
        ne_p = int(lines[0])
        p = 1
        for _ in range(ne_p):
            p += 1                                   # "dim slave master"
            if lines[p].split()[0] == "Affine":
                p += 1
            npairs = int(lines[p])
            p += 1
            for q in range(npairs):
                a, b = (int(x) for x in lines[p + q].split()[:2])
                master[a] = b
            p += npairs

    def root(n):
        seen = 0
        while n in master and master[n] != n:
            n = master[n]
            seen += 1
            if seen > 8:
                raise ValueError("cyclic periodic node map")
        return n

    rows = [[int(x) for x in line.split()] for line in sec["Elements"][1:]]
    types = {v[1] for v in rows}
    dim = 3 if 5 in types else 2
    vol_t, bnd_t = (5, 3) if dim == 3 else (3, 1)
    if vol_t not in types:
        raise ValueError("no quadrilateral or hexahedral elements")
    other = types - {vol_t, bnd_t, 15} - ({1} if dim == 3 else set())
    if other:
        raise ValueError(f"element types {sorted(other)} are not supported (quadrilaterals / hexahedra only)")
    order = _HEX_TO_LEX if dim == 3 else _QUAD_TO_LEX
    elems, attrs, bnd = [], [], {}
    for v in rows:
        ntag = v[2]
        phys = v[3] if ntag else 0
        nodes = v[3 + ntag:]
        if v[1] == vol_t:
            elems.append([nodes[k] for k in order])
            attrs.append(phys)
        elif v[1] == bnd_t:
            bnd[tuple(sorted(root(n) for n in nodes))] = phys
    raw = np.array(elems, dtype=np.int64)
    coords = xyz[np.vectorize(index.__getitem__)(raw)][..., :dim]
    vids = np.vectorize(root)(raw)
    return _finish(dim, vids, coords, bnd, bc_types, attrs)


def gmsh_physical_names(path) -> dict:
    """{name: physical tag} from a Gmsh 2.2 file's $PhysicalNames section (the
    names a case file's boundary_conditions refer to)."""
    lines = Path(path).read_text().split("\n")
    out = {}
    for i, line in enumerate(lines):
        if line.strip() == "$PhysicalNames":
            n = int(lines[i + 1])
            for row in lines[i + 2:i + 2 + n]:
                dim_, tag, name = row.split(maxsplit=2)
                out[name.strip().strip('"')] = int(tag)
            break
    return out


def read_mesh(path, bc_types=2) -> AffineMesh:
    p = Path(path)
    head = p.read_text()[:64]
    if head.startswith("MFEM mesh"):
        return read_mfem(p, bc_types)
    if head.startswith("$MeshFormat"):
        return read_gmsh22(p, bc_types)
    raise ValueError(f"{p}: unknown mesh format")


# ---------------------------------------------------------------------------
# refinement and ordering
# ---------------------------------------------------------------------------
def uniform_refine(mesh: AffineMesh, levels: int = 1) -> AffineMesh:
    """Split every element into 2^dim affine children, child c (lexicographic
    bits) of element e becoming element e*2^dim + c."""
    for _ in range(levels):
        dim = mesh.dim
        nc = 1 << dim
        ne = mesh.num_elements
        o = mesh.origin
        E = mesh.edges
        corners = np.empty((ne * nc, nc, dim))
        nbr = np.empty((ne * nc, 2 * dim), dtype=np.int32)
        for c in range(nc):
            cb = [(c >> d) & 1 for d in range(dim)]
            for k in range(nc):
                kb = [(k >> d) & 1 for d in range(dim)]
                corners[c::nc, k] = o + sum((cb[d] + kb[d]) * 0.5 * E[:, d] for d in range(dim))
            for d in range(dim):
                for s in (0, 1):
                    child = np.arange(ne) * nc
                    if cb[d] != s:   # sibling inside the parent
                        nbr[c::nc, 2 * d + s] = child + (c ^ (1 << d))
                    else:            # across the parent's face: the neighbour's child with bit d flipped
                        pn = mesh.face_neighbor[:, 2 * d + s]
                        nbr[c::nc, 2 * d + s] = np.where(pn >= 0, pn * nc + (c ^ (1 << d)), pn)
        mesh = AffineMesh(dim, corners, nbr, mesh.bc_type, mesh.bc_attributes,
                          np.repeat(np.asarray(mesh.elem_attr), nc) if mesh.elem_attr is not None else None, None,
                          mesh.bc_data)
    return mesh


def _hilbert_index(coords: np.ndarray, bits: int = 16) -> np.ndarray:
    """Skilling's transpose algorithm (AIP Conf. Proc. 707, 381 (2004)):
    integer coordinates (n, dim) in [0, 2^bits) -> Hilbert index."""
    X = coords.astype(np.int64).T.copy()  # (dim, n)
    n = X.shape[0]
    M = 1 << (bits - 1)
    Q = M
    while Q > 1:
        P = Q - 1
        for i in range(n):
            hi = (X[i] & Q) != 0
            X[0] = np.where(hi, X[0] ^ P, X[0])
            t = np.where(hi, 0, (X[0] ^ X[i]) & P)
            X[0] ^= t
            X[i] ^= t
        Q >>= 1
    for i in range(1, n):
        X[i] ^= X[i - 1]
    t = np.zeros_like(X[0])
    Q = M
    while Q > 1:
        t = np.where((X[n - 1] & Q) != 0, t ^ (Q - 1), t)
        Q >>= 1
    for i in range(n):
        X[i] ^= t
    # interleave the transposed bits, most significant first
    h = np.zeros(X.shape[1], dtype=object)
    for b in range(bits - 1, -1, -1):
        for i in range(n):
            h = h * 2 + ((X[i] >> b) & 1)
    return h


def hilbert_order(mesh: AffineMesh, bits: int = 16) -> np.ndarray:
    """Element permutation (new position -> old element) along a Hilbert
    curve through the element centres."""
    cen = mesh.corners.mean(axis=1)
    lo, hi = cen.min(axis=0), cen.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)
    q = np.floor((cen - lo) / span * ((1 << bits) - 1)).astype(np.int64)
    key = _hilbert_index(q, bits)
    return np.array(sorted(range(len(key)), key=lambda e: key[e]), dtype=np.int64)


def reorder(mesh: AffineMesh, order: np.ndarray) -> AffineMesh:
    """Elements in the order `order` (new position -> old element)."""
    order = np.asarray(order, dtype=np.int64)
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    nbr = mesh.face_neighbor[order].astype(np.int64)
    nbr = np.where(nbr >= 0, inv[np.maximum(nbr, 0)], nbr).astype(np.int32)
    attr = np.asarray(mesh.elem_attr)[order] if mesh.elem_attr is not None else None
    vids = mesh.vertex_ids[order] if mesh.vertex_ids is not None else None
    codes = mesh.elem_orient[order] if mesh.elem_orient is not None else None
    return AffineMesh(mesh.dim, mesh.corners[order], nbr, mesh.bc_type, mesh.bc_attributes, attr, vids, mesh.bc_data,
                      codes)
