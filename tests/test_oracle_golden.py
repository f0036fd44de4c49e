"""The CPU oracle against the reference's own golden output: isentropic
vortex, 500 SSP-RK3 steps (see tests/golden_vortex.py), and the
Navier-Stokes lid-driven cavity with no-slip walls, 4000 steps (see
tests/golden_cavity.py).  The reference's
regression compares at atol=rtol=1e-13 between runs of one binary
(tests/recipes/regression/golden-vortex.sh); a restatement with its own GLL
nodes and summation orders is held to 1e-12 relative."""
import numpy as np

from .golden_vortex import DT, NSTEPS, golden_permutation, primitive_fields, vortex_case


def test_oracle_reproduces_reference_golden_vortex(oracle):
    from theseus_amd.operator import build_descriptor

    rt, mesh, U0 = vortex_case()
    d, perm = golden_permutation(mesh)
    desc, keep = build_descriptor(rt, mesh)
    case = oracle.OracleCase(desc)
    U = U0.copy()
    for _ in range(NSTEPS):
        U = case.step(U, DT)
    f = primitive_fields(U)
    for k in ("rho", "u", "v", "p"):
        g = d[k][perm]
        rel = np.linalg.norm(f[k] - g) / np.linalg.norm(g)
        assert rel < 1e-12, (k, rel)
        np.testing.assert_allclose(f[k], g, rtol=0, atol=1e-12 * np.abs(g).max())
    alpha, eta = case.blending(U)
    np.testing.assert_array_equal(alpha, d["alpha"][perm].reshape(-1, 16)[:, 0])


def test_golden_detects_a_planted_mistake(oracle):
    """A plausible slip (gamma entered as 1.4001 in the operator) moves the
    500-step result far outside the 1e-12 band, so the fixture pins the
    arithmetic rather than just the initial condition."""
    from theseus_amd.operator import build_descriptor

    rt, mesh, U0 = vortex_case()
    d, perm = golden_permutation(mesh)
    rt.gamma = 1.4001
    desc, keep = build_descriptor(rt, mesh)
    case = oracle.OracleCase(desc)
    U = U0.copy()
    for _ in range(NSTEPS):
        U = case.step(U, DT)
    f = primitive_fields(U, gamma=1.4)
    g = d["p"][perm]
    assert np.linalg.norm(f["p"] - g) / np.linalg.norm(g) > 1e-9


def test_oracle_reproduces_reference_golden_lid_driven_cavity(oracle):
    """Pins the wall boundary conditions end to end: no-slip adiabatic walls
    and a moving lid (wall pressure, BR1 boundary jumps, viscous wall flux)."""
    from theseus_amd.operator import build_descriptor

    from .golden_cavity import DT, NSTEPS, cavity_case, check_fields, golden_permutation

    rt, mesh, U0 = cavity_case()
    d, perm = golden_permutation(mesh)
    desc, keep = build_descriptor(rt, mesh)
    case = oracle.OracleCase(desc)
    oracle._lib.oracle_set_num_threads(8)
    U = U0.copy()
    for _ in range(NSTEPS):
        U = case.step(U, DT)
    errs = check_fields(U, d, perm)
    print("cavity rel L2 per field", errs)
    alpha, _ = case.blending(U)
    assert (alpha == 0).all() and (d["alpha"] == 0).all()


def test_cavity_golden_detects_a_planted_lid_speed(oracle):
    """A lid moving at 1.0001 instead of 1 leaves the 4000-step result far
    outside the band: the fixture pins the wall data, not just the flow."""
    from theseus_amd.operator import build_descriptor

    from .golden_cavity import DT, cavity_case, golden_permutation

    rt, mesh, U0 = cavity_case()
    d, perm = golden_permutation(mesh)
    lid = [i for i in range(mesh.bc_data.shape[0]) if mesh.bc_data[i, 0] == 1.0]
    assert len(lid) == 1
    mesh.bc_data[lid[0], 0] = 1.0001
    desc, keep = build_descriptor(rt, mesh)
    case = oracle.OracleCase(desc)
    oracle._lib.oracle_set_num_threads(8)
    U = U0.copy()
    for _ in range(400):
        U = case.step(U, DT)
    ref_desc, ref_keep = build_descriptor(rt, cavity_case()[1])
    ref = oracle.OracleCase(ref_desc)
    V = U0.copy()
    for _ in range(400):
        V = ref.step(V, DT)
    u_planted, u_ref = U.reshape(4, -1)[1], V.reshape(4, -1)[1]
    assert np.linalg.norm(u_planted - u_ref) / np.linalg.norm(u_ref) > 1e-6


def _ffs_run(oracle, rt, mesh, U0):
    """100 fixed steps; returns the final state and the blending
    coefficients of the last RHS evaluation (third stage of the last step)."""
    from theseus_amd.operator import build_descriptor

    from .golden_ffs import DT, NSTEPS

    desc, keep = build_descriptor(rt, mesh)
    case = oracle.OracleCase(desc)
    oracle._lib.oracle_set_num_threads(8)
    U = U0.copy()
    for _ in range(NSTEPS - 1):
        U = case.step(U, DT)
    u1 = U + DT * case.rhs(U)
    u2 = 0.75 * U + 0.25 * (u1 + DT * case.rhs(u1))
    alpha, _ = case.blending(u2)
    return case.step(U, DT), alpha


def test_oracle_reproduces_reference_golden_forward_facing_step(oracle):
    """Pins the shock-capturing path end to end against the reference's own
    output: the Persson-Peraire indicator and its eta -> alpha map (nonzero
    in 50 elements, up to alpha_max), the subcell FV residual and the blend,
    slip walls at the step and supersonic inflow/outflow."""
    from .golden_ffs import check, ffs_case

    rt, mesh, U0 = ffs_case()
    U, alpha = _ffs_run(oracle, rt, mesh, U0)
    errs = check(U, alpha, mesh)
    print("forward-facing step: largest node error / field scale", errs)


def test_ffs_golden_detects_a_planted_blending_limit(oracle):
    """alpha_max entered as 0.49 instead of 0.5 changes only the elements
    whose blending saturates; the 100-step result leaves the band, so the
    fixture pins the FV blend and not just the smooth part of the flow."""
    from .golden_ffs import ABS_TOL, golden_match, ffs_case, primitive

    rt, mesh, U0 = ffs_case()
    rt.alpha_max = 0.49
    U, _ = _ffs_run(oracle, rt, mesh, U0)
    d, ours, perm = golden_match(mesh)
    f = primitive(U)["rho"].reshape(mesh.num_elements, -1)
    assert np.abs(f[ours] - d["rho"][perm]).max() / np.abs(d["rho"]).max() > 1e3 * ABS_TOL


def test_oracle_reproduces_reference_numerical_flux_cavities(oracle):
    """The reference's Roe and HLL goldens (p=2 Navier-Stokes cavity with
    no-slip walls, cycles 10 and 20) at the reference checker's rtol 1e-10,
    atol 1e-12; the Roe run also fails the HLL golden (the two pin different
    face fluxes)."""
    from theseus_amd.operator import build_descriptor

    from .golden_numflux import CASES, CYCLES, DT, compare, golden_rows, numflux_case

    oracle._lib.oracle_set_num_threads(8)
    finals = {}
    for name in CASES:
        rt, mesh, U0 = numflux_case(name)
        desc, keep = build_descriptor(rt, mesh)
        case = oracle.OracleCase(desc)
        U = U0.copy()
        for s in range(1, max(CYCLES) + 1):
            U = case.step(U, DT)
            if s in CYCLES:
                d, rows = golden_rows(mesh, name, s)
                res = compare(U, d, rows)
                assert all(res[k] <= 0 for k in ("Density", "Pressure", "Velocity")), (name, s, res)
        finals[name] = (U, mesh)
    U, mesh = finals["cpg-roe-cavity"]
    d, rows = golden_rows(mesh, "cpg-hll-cavity", max(CYCLES))
    assert compare(U, d, rows)["Velocity"] > 1e-4
