"""Output formats on the CPU: the raw checkpoint (byte layout, metadata,
round trip, the reference's compatibility checks) and the gll_subcells
VTU/PVTU (subcell connectivity as in the reference's golden output, write ->
read round trip, derived fields against closed forms).  When the
reference's golden PVTU is present, the reader reproduces the decoded golden
arrays (tests/golden/isentropic_vortex_cycle500.npz) exactly."""
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from theseus_amd.mesh import box_mesh, gll_nodes
from theseus_amd.output import (MAGIC, load_checkpoint, point_fields, read_pvtu, read_vtu, save_checkpoint,
                                save_visualization, subcell_connectivity, validate_metadata)

GOLDEN_PVTU = Path("/root/reference/TestCases/GoldenData/IsentropicVortex/Cycle000500/data.pvtu")
GOLDEN_NPZ = Path(__file__).resolve().parent / "golden" / "isentropic_vortex_cycle500.npz"


def test_checkpoint_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    dim, order, ne = 3, 2, 5
    ndof = ne * (order + 1) ** dim
    U = rng.standard_normal(5 * ndof)
    d = save_checkpoint(tmp_path, 40, 0.125, U, order, dim, ne)
    raw = (d / "checkpoint_cycle_40.00000000.chk").read_bytes()
    assert raw.startswith(b"THESEUS_CHECKPOINT_RAW_V1\n") and MAGIC == b"THESEUS_CHECKPOINT_RAW_V1\n"
    assert struct.unpack_from("<Q", raw, len(MAGIC))[0] == U.size
    assert len(raw) == len(MAGIC) + 8 + 8 * U.size
    meta = json.loads((d / "checkpoint_cycle_40.json").read_text())
    assert meta == {"format_version": 2, "state_format": "raw_vector_v1", "state_representation": "conservative",
                    "geometry": "cartesian", "real_bytes": 8, "time": 0.125, "cycle": 40, "mpi_ranks": 1,
                    "order": 2, "dimension": 3, "num_equations": 5, "global_elements": 5, "global_dofs": U.size}
    V, t, c = load_checkpoint(tmp_path, 40, order, dim, ne, ndof)
    assert np.array_equal(U, V) and t == 0.125 and c == 40


def test_checkpoint_compatibility_checks(tmp_path):
    dim, order, ne = 2, 3, 4
    ndof = ne * 16
    save_checkpoint(tmp_path, 3, 1.0, np.zeros(4 * ndof), order, dim, ne)
    with pytest.raises(ValueError, match="order"):
        load_checkpoint(tmp_path, 3, 4, dim, ne, ne * 25)
    with pytest.raises(ValueError, match="global_elements"):
        load_checkpoint(tmp_path, 3, order, dim, ne + 1, (ne + 1) * 16)
    meta = {"time": 1.0, "cycle": 3}
    assert validate_metadata(meta, {}) is False
    bad = {"format_version": 3}
    with pytest.raises(ValueError, match="format version"):
        validate_metadata(bad, {})
    d = tmp_path / "Cycle3"
    (d / "checkpoint_cycle_3.00000000.chk").write_bytes(b"NOT_A_CHECKPOINT\n" + bytes(8))
    with pytest.raises(ValueError, match="header"):
        load_checkpoint(tmp_path, 3, order, dim, ne, ndof)


def test_subcell_connectivity_matches_the_reference_layout():
    # the reference's golden vortex output (p=3 quads): first subcells 0 1 5 4 | 1 2 6 5 | 2 3 7 6
    conn, offs, types = subcell_connectivity(2, 3, 2)
    assert list(conn[:12]) == [0, 1, 5, 4, 1, 2, 6, 5, 2, 3, 7, 6]
    assert list(offs[:3]) == [4, 8, 12] and set(types) == {9} and offs.size == 18
    assert list(conn[36:40]) == [16, 17, 21, 20]  # second element's first subcell
    conn3, offs3, types3 = subcell_connectivity(3, 2, 1)
    assert list(conn3[:8]) == [0, 1, 4, 3, 9, 10, 13, 12] and set(types3) == {12} and offs3.size == 8


def test_visualization_round_trip_and_derived_fields(tmp_path):
    mesh = box_mesh(2, (3, 2), 0.0, (1.5, 1.0))
    order = 3
    X = mesh.node_coordinates(gll_nodes(order)[0])
    rho = 1.0 + 0.1 * X[0]
    u, v = 0.3 + 0 * X[0], -0.2 * X[1]
    p = 2.0 + X[1]
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    U = np.concatenate([rho, rho * u, rho * v, E])
    alpha = np.linspace(0, 0.5, mesh.num_elements)
    d = save_visualization(tmp_path, 7, mesh, order, U, alpha=alpha,
                           fields=["density", "velocity", "pressure", "blending_coefficient", "temperature",
                                   "mach_number"])
    assert (d / "data.pvtu").exists() and d.name == "Cycle000007"
    a = read_pvtu(d / "data.pvtu")
    np.testing.assert_allclose(a["Points"][:, :2], X.T, rtol=0, atol=0)
    np.testing.assert_allclose(a["Density"], rho, rtol=1e-15)
    np.testing.assert_allclose(a["Velocity"], np.c_[u, v], rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(a["Pressure"], p, rtol=1e-13)
    np.testing.assert_allclose(a["Temperature"], p / (rho * 287.05), rtol=1e-13)
    np.testing.assert_allclose(a["Mach Number"], np.sqrt(u * u + v * v) / np.sqrt(1.4 * p / rho), rtol=1e-13)
    np.testing.assert_array_equal(a["Blending Coeff"], np.repeat(alpha, 16))
    assert a["connectivity"].size == mesh.num_elements * 9 * 4 and (a["attribute"] == 1).all()
    text = (d / "proc000000.vtu").read_text()
    names = [n for n in ("Blending Coeff", "Density", "Mach Number", "Pressure", "Temperature", "Velocity")]
    assert [text.index(f'Name="{n}"') for n in names] == sorted(text.index(f'Name="{n}"') for n in names)


@pytest.mark.skipif(not GOLDEN_PVTU.exists(), reason="the reference's golden PVTU is not present")
def test_reader_decodes_the_reference_golden_output():
    a = read_pvtu(GOLDEN_PVTU)
    g = np.load(GOLDEN_NPZ)
    np.testing.assert_array_equal(a["Points"][:, 0], g["x"])
    np.testing.assert_array_equal(a["Density"], g["rho"])
    np.testing.assert_array_equal(a["Pressure"], g["p"])
    np.testing.assert_array_equal(a["Velocity"][:, 1], g["v"])
    np.testing.assert_array_equal(a["Blending Coeff"], g["alpha"])
    one = read_vtu(GOLDEN_PVTU.parent / "proc000000.vtu")
    conn, offs, types = subcell_connectivity(2, 3, 288)
    np.testing.assert_array_equal(one["connectivity"], conn)
    np.testing.assert_array_equal(one["offsets"], offs)


def test_point_fields_closed_forms():
    U = np.array([2.0, 2.0 * 3.0, 2.0 * 4.0, 10.0 / 0.4 + 0.5 * 2.0 * 25.0])  # one node: rho 2, v (3,4), p 10
    f = point_fields(U, 2, 1, ["pressure", "sound_speed", "specific_internal_energy", "internal_energy_density",
                               "specific_total_energy", "viscosity", "thermal_conductivity"], mu=0.5, Pr=0.5)
    assert f["Pressure"][0] == pytest.approx(10.0, rel=1e-15)
    assert f["Sound Speed"][0] == pytest.approx(np.sqrt(1.4 * 10 / 2), rel=1e-15)
    assert f["Internal Energy Density"][0] == pytest.approx(25.0, rel=1e-15)
    assert f["Specific Internal Energy"][0] == pytest.approx(12.5, rel=1e-15)
    assert f["Specific Total Energy"][0] == pytest.approx((25.0 + 25.0) / 2.0, rel=1e-15)
    assert f["Viscosity"][0] == 0.5
    assert f["Thermal Conductivity"][0] == pytest.approx(0.5 * 3.5 * 287.05 / 0.5, rel=1e-15)
    with pytest.raises(ValueError, match="Unknown visualization field"):
        point_fields(U, 2, 1, ["vorticity"])
