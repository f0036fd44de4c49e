"""The reference's output formats, so that its tooling reads this engine's
results and this engine restarts from its checkpoints.

Checkpoints (src/Simulation.cpp:1477-1603, CheckpointConfig.hpp:30-175):
``<dir>/Cycle<N>/checkpoint_cycle_<N>.<rank:08d>.chk`` holds the line
``THESEUS_CHECKPOINT_RAW_V1``, a little-endian uint64 entry count and the
raw FP64 state (the L-vector of the rank's elements, the engine's layout);
``checkpoint_cycle_<N>.json`` holds the metadata (format_version 2,
state_format raw_vector_v1, conservative, cartesian, real_bytes 8, time,
cycle, mpi_ranks, order, dimension, num_equations, global_elements,
global_dofs) that ``ValidateMetadata`` checks on load.

Visualization (VisualizationConfig.hpp, VisualizationFields.hpp; MFEM's
ParaView output in the ``gll_subcells`` mesh mode): one VTU point per GLL
node of every element (x fastest), p^dim linear subcells per element (VTK
quads / hexahedra), point data under the reference's output names, cell
data ``attribute``; arrays in the "binary" format (base64 of a UInt32 byte
count followed by base64 of the little-endian data) and a ``.pvtu`` that
lists the pieces.  Fields are written in name order.
"""
from __future__ import annotations

import base64
import json
import re
import struct
from pathlib import Path

import numpy as np

MAGIC = b"THESEUS_CHECKPOINT_RAW_V1\n"


# ---------------------------------------------------------------------------
# checkpoints
# ---------------------------------------------------------------------------
def _cycle_dir(directory, cycle):
    return Path(directory) / f"Cycle{int(cycle)}"


def checkpoint_metadata(time: float, cycle: int, order: int, dim: int, num_elements: int, ndof_total: int,
                        mpi_ranks: int = 1) -> dict:
    """CheckpointConfig::Metadata (CheckpointConfig.hpp:94-110); global_dofs is
    the vector size of the state (num_equations x scalar dofs)."""
    neq = dim + 2
    return {"format_version": 2, "state_format": "raw_vector_v1", "state_representation": "conservative",
            "geometry": "cartesian", "real_bytes": 8, "time": float(time), "cycle": int(cycle),
            "mpi_ranks": int(mpi_ranks), "order": int(order), "dimension": int(dim), "num_equations": neq,
            "global_elements": int(num_elements), "global_dofs": int(ndof_total) * neq}


def save_checkpoint(directory, cycle: int, time: float, U: np.ndarray, order: int, dim: int,
                    num_elements: int) -> Path:
    """Simulation::SaveCheckpoint for one rank."""
    U = np.ascontiguousarray(U, dtype="<f8").reshape(-1)
    d = _cycle_dir(directory, cycle)
    d.mkdir(parents=True, exist_ok=True)
    rank_file = d / f"checkpoint_cycle_{int(cycle)}.{0:08d}.chk"
    with open(rank_file, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<Q", U.size))
        f.write(U.tobytes())
    meta = checkpoint_metadata(time, cycle, order, dim, num_elements, U.size // (dim + 2))
    (d / f"checkpoint_cycle_{int(cycle)}.json").write_text(json.dumps(meta, indent=2) + "\n")
    return d


def validate_metadata(meta: dict, expected: dict) -> bool:
    """CheckpointConfig::ValidateMetadata (CheckpointConfig.hpp:112-175):
    False for legacy metadata (time and cycle only); raises ValueError for
    an incompatible or unsupported checkpoint."""
    if "format_version" not in meta:
        return False
    fv = meta.get("format_version", 0)
    if fv not in (1, 2):
        raise ValueError("Unsupported checkpoint metadata format version")
    if meta.get("state_format") != "raw_vector_v1":
        raise ValueError("Unsupported checkpoint state format")
    if fv == 2:
        if meta.get("state_representation") != "conservative":
            raise ValueError("Unsupported checkpoint state representation")
        if meta.get("geometry") != "cartesian":
            raise ValueError("Checkpoint is incompatible with the current run: geometry differs")
    if meta.get("real_bytes", 0) != 8:
        raise ValueError("Checkpoint floating-point representation differs")
    for k in ("mpi_ranks", "order", "dimension", "num_equations", "global_elements", "global_dofs"):
        if k not in meta or int(meta[k]) != int(expected[k]):
            raise ValueError(f"Checkpoint is incompatible with the current run: {k} differs")
    return True


def load_checkpoint(directory, cycle: int, order: int, dim: int, num_elements: int, ndof_total: int):
    """Simulation::LoadCheckpoint for one rank: returns (U, time, cycle)."""
    d = _cycle_dir(directory, cycle)
    meta = json.loads((d / f"checkpoint_cycle_{int(cycle)}.json").read_text())
    if int(meta.get("cycle", 0)) != int(cycle):
        raise ValueError("checkpoint cycle in the metadata differs from the requested one")
    expected = checkpoint_metadata(0.0, cycle, order, dim, num_elements, ndof_total)
    current = validate_metadata(meta, expected)
    raw = (d / f"checkpoint_cycle_{int(cycle)}.{0:08d}.chk").read_bytes()
    n = ndof_total * (dim + 2)
    if current:
        if not raw.startswith(MAGIC):
            raise ValueError("Invalid checkpoint state header")
        (stored,) = struct.unpack_from("<Q", raw, len(MAGIC))
        if stored != n:
            raise ValueError(f"Checkpoint state size {stored} does not match current space size {n}")
        U = np.frombuffer(raw, dtype="<f8", count=n, offset=len(MAGIC) + 8).copy()
    else:
        raise ValueError("legacy (MFEM grid-function) checkpoints are not read")
    return U, float(meta["time"]), int(meta["cycle"])


# ---------------------------------------------------------------------------
# visualization fields (VisualizationFields.hpp:55-77, EOS.hpp)
# ---------------------------------------------------------------------------
OUTPUT_NAMES = {
    "density": "Density", "velocity": "Velocity", "pressure": "Pressure", "blending_coefficient": "Blending Coeff",
    "temperature": "Temperature", "sound_speed": "Sound Speed", "mach_number": "Mach Number",
    "specific_internal_energy": "Specific Internal Energy", "internal_energy_density": "Internal Energy Density",
    "specific_total_energy": "Specific Total Energy", "viscosity": "Viscosity",
    "thermal_conductivity": "Thermal Conductivity",
}


def point_fields(U: np.ndarray, dim: int, ndof: int, fields=("density", "velocity", "pressure"), gamma=1.4,
                 R_gas=287.05, mu=0.02, Pr=0.72, alpha=None) -> dict:
    """Calorically perfect gas fields at every GLL node, keyed by output name."""
    q = U.reshape(dim + 2, -1)
    rho, m, rhoE = q[0], q[1:dim + 1], q[dim + 1]
    ke = 0.5 * (m * m).sum(axis=0) / rho
    p = (gamma - 1.0) * (rhoE - ke)
    out = {}
    for f in fields:
        if f == "density":
            v = rho
        elif f == "velocity":
            v = (m / rho).T.copy()
        elif f == "pressure":
            v = p
        elif f == "temperature":
            v = p / (rho * R_gas)
        elif f == "sound_speed":
            v = np.sqrt(gamma * p / rho)
        elif f == "mach_number":
            v = np.sqrt(2 * ke / rho) / np.sqrt(gamma * p / rho)
        elif f == "specific_internal_energy":
            v = (rhoE - ke) / rho
        elif f == "internal_energy_density":
            v = rhoE - ke
        elif f == "specific_total_energy":
            v = rhoE / rho
        elif f == "viscosity":
            v = np.full_like(rho, mu)
        elif f == "thermal_conductivity":
            v = np.full_like(rho, mu * gamma / (gamma - 1.0) * R_gas / Pr)
        elif f == "blending_coefficient":
            if alpha is None:
                raise ValueError("blending_coefficient needs the per-element alpha")
            v = np.repeat(np.asarray(alpha, dtype=np.float64), ndof)
        else:
            raise ValueError(f"Unknown visualization field {f!r}. Supported fields: {', '.join(OUTPUT_NAMES)}")
        out[OUTPUT_NAMES[f]] = v
    return out


# ---------------------------------------------------------------------------
# VTU (gll_subcells)
# ---------------------------------------------------------------------------
_VTK_TYPE = {"Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}


def _b64(arr: np.ndarray) -> str:
    data = np.ascontiguousarray(arr).tobytes()
    return (base64.b64encode(struct.pack("<I", len(data))) + base64.b64encode(data)).decode()


def subcell_connectivity(dim: int, order: int, num_elements: int):
    """(connectivity, offsets, types) of the p^dim linear subcells of every element."""
    Np = order + 1
    nd = Np ** dim
    if dim == 2:
        loc = []
        for j in range(order):
            for i in range(order):
                b = j * Np + i
                loc.append([b, b + 1, b + Np + 1, b + Np])
        vtk = 9
    else:
        loc = []
        for k in range(order):
            for j in range(order):
                for i in range(order):
                    b = (k * Np + j) * Np + i
                    lo = [b, b + 1, b + Np + 1, b + Np]
                    loc.append(lo + [x + Np * Np for x in lo])
        vtk = 12
    loc = np.array(loc, dtype=np.int64)
    conn = (np.arange(num_elements)[:, None, None] * nd + loc[None]).reshape(-1)
    nv = loc.shape[1]
    ncell = num_elements * loc.shape[0]
    offsets = np.arange(1, ncell + 1) * nv
    return conn.astype("<i4"), offsets.astype("<i4"), np.full(ncell, vtk, dtype="u1")


def write_vtu(path, points: np.ndarray, dim: int, order: int, num_elements: int, point_data: dict,
              attribute=None) -> None:
    """One piece: points (dim, npts) or (npts, dim) GLL coordinates, point data
    {name: (npts,) or (npts, ncomp)}."""
    P = np.asarray(points, dtype=np.float64)
    if P.shape[0] == dim and P.shape[1] != dim:
        P = P.T
    npts = P.shape[0]
    P3 = np.zeros((npts, 3))
    P3[:, :dim] = P
    conn, offs, types = subcell_connectivity(dim, order, num_elements)
    ncell = offs.size
    attr = np.ones(ncell, dtype="<i4") if attribute is None else np.repeat(
        np.asarray(attribute, dtype="<i4"), order ** dim)
    lines = ['<VTKFile type="UnstructuredGrid" version="2.2" byte_order="LittleEndian">', "<UnstructuredGrid>",
             f'<Piece NumberOfPoints="{npts}" NumberOfCells="{ncell}">', "<Points>",
             '<DataArray type="Float64" NumberOfComponents="3" format="binary">', _b64(P3.astype("<f8")),
             "</DataArray>", "</Points>", "<Cells>",
             '<DataArray type="Int32" Name="connectivity" format="binary">', _b64(conn), "</DataArray>",
             '<DataArray type="Int32" Name="offsets" format="binary">', _b64(offs), "</DataArray>",
             '<DataArray type="UInt8" Name="types" format="binary">', _b64(types), "</DataArray>", "</Cells>",
             '<CellData Scalars="attribute">', '<DataArray type="Int32" Name="attribute" format="binary">',
             _b64(attr), "</DataArray>", "</CellData>", "<PointData >"]
    for name in sorted(point_data):
        v = np.asarray(point_data[name], dtype="<f8")
        nc = 1 if v.ndim == 1 else v.shape[1]
        comps = "".join(f' ComponentName{c}="{c}"' for c in range(nc)) if nc > 1 else ""
        lines += [f'<DataArray type="Float64" Name="{name}" NumberOfComponents="{nc}"{comps} format="binary" >',
                  _b64(v), "</DataArray>"]
    lines += ["</PointData>", "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    Path(path).write_text("\n".join(lines) + "\n")


def write_pvtu(path, pieces, point_data: dict) -> None:
    """The .pvtu index of the pieces (file names relative to it)."""
    lines = ['<?xml version="1.0"?>', '<VTKFile type="PUnstructuredGrid" version ="2.2" byte_order="LittleEndian">',
             '<PUnstructuredGrid GhostLevel="0">', "<PPoints>",
             '\t<PDataArray type="Float64"  Name="Points" NumberOfComponents="3" format="binary"/>', "</PPoints>",
             "<PCells>", '\t<PDataArray type="Int32"  Name="connectivity" NumberOfComponents="1" format="binary"/>',
             '\t<PDataArray type="Int32"  Name="offsets"      NumberOfComponents="1" format="binary"/>',
             '\t<PDataArray type="UInt8"  Name="types"        NumberOfComponents="1" format="binary"/>',
             "</PCells>", "<PPointData>"]
    for name in sorted(point_data):
        v = np.asarray(point_data[name])
        nc = 1 if v.ndim == 1 else v.shape[1]
        comps = "".join(f' ComponentName{c}="{c}"' for c in range(nc)) if nc > 1 else ""
        lines.append(f'<PDataArray type="Float64" Name="{name}" NumberOfComponents="{nc}"{comps} format="binary" />')
    lines += ["</PPointData>", "<PCellData>",
              '\t<PDataArray type="Int32" Name="attribute" NumberOfComponents="1" format="binary"/>', "</PCellData>"]
    lines += [f'<Piece Source="{Path(p).name}"/>' for p in pieces]
    lines += ["</PUnstructuredGrid>", "</VTKFile>"]
    Path(path).write_text("\n".join(lines) + "\n")


def read_vtu(path) -> dict:
    """Arrays of one piece, keyed by Name ("Points" for the coordinates)."""
    text = Path(path).read_text()
    out = {}
    for m in re.finditer(r"<DataArray([^>]*)>\s*([^<]*)</DataArray>", text):
        attrs, payload = m.group(1), m.group(2).strip()
        name = re.search(r'Name="([^"]*)"', attrs)
        name = name.group(1) if name else "Points"
        typ = re.search(r'type="([^"]*)"', attrs).group(1)
        ncomp = int((re.search(r'NumberOfComponents="(\d+)"', attrs) or [None, "1"])[1])
        (nbytes,) = struct.unpack("<I", base64.b64decode(payload[:8]))
        data = base64.b64decode(payload[8:])[:nbytes]
        arr = np.frombuffer(data, dtype=_VTK_TYPE[typ])
        out[name] = arr.reshape(-1, ncomp) if ncomp > 1 else arr
    return out


def read_pvtu(path) -> dict:
    """Every piece of a .pvtu, concatenated in the order listed."""
    text = Path(path).read_text()
    pieces = [read_vtu(Path(path).parent / s) for s in re.findall(r'<Piece Source="([^"]+)"', text)]
    out = {}
    for k in pieces[0]:
        if k == "connectivity":
            shift = np.cumsum([0] + [p["Points"].shape[0] for p in pieces[:-1]])
            out[k] = np.concatenate([p[k] + s for p, s in zip(pieces, shift)])
        elif k == "offsets":
            shift = np.cumsum([0] + [p["connectivity"].size for p in pieces[:-1]])
            out[k] = np.concatenate([p[k] + s for p, s in zip(pieces, shift)])
        else:
            out[k] = np.concatenate([p[k] for p in pieces])
    return out


def save_visualization(directory, cycle: int, mesh, order: int, U: np.ndarray, fields=None, gamma=1.4,
                       R_gas=287.05, mu=0.02, Pr=0.72, alpha=None) -> Path:
    """``<dir>/Cycle<NNNNNN>/proc000000.vtu`` + ``data.pvtu`` (one rank);
    default fields as VisualizationConfig: density, velocity, pressure and,
    with blending, the blending coefficient."""
    from .mesh import gll_nodes

    if fields is None:
        fields = ["density", "velocity", "pressure"] + (["blending_coefficient"] if alpha is not None else [])
    dim = mesh.dim
    ndof = (order + 1) ** dim
    X = mesh.node_coordinates(gll_nodes(order)[0])
    pdata = point_fields(U, dim, ndof, fields, gamma, R_gas, mu, Pr, alpha)
    d = Path(directory) / f"Cycle{int(cycle):06d}"
    d.mkdir(parents=True, exist_ok=True)
    piece = d / "proc000000.vtu"
    write_vtu(piece, X, dim, order, mesh.num_elements, pdata)
    write_pvtu(d / "data.pvtu", [piece], pdata)
    return d
