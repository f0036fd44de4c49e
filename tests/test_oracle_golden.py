"""The CPU oracle against the reference's own golden output: isentropic
vortex, 500 SSP-RK3 steps (see tests/golden_vortex.py).  The reference's
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
