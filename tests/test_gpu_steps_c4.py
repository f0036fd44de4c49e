"""The full-size LTE Taylor-Green workload (24^3 hexes, p=3, LTEEC flux, the
synthetic 256x256 Helmholtz table, 4,423,680 values), Euler and
Navier-Stokes, stepped 20 times on the engine in FP64 at the fixed dt of CFL
0.5 on the initial state, against the oracle's run of the same steps
(tests/golden/c4_lte_tgv24_p3_20steps.npz, made by
tests/golden/make_c4_steps_fixture.py: every 509th value of the final
state, per-equation sums and L2 norms).

Bar.  This workload starts with a uniform density, and the reference's
discrete-gradient means (LTEHelmholtzEOS.hpp:246-300) divide differences of
table values by density differences just above their switch to the
derivative form when a pair straddles a table cell: rounding-level
differences grow to about 4e-10 per field in one step, in the oracle itself
(its run from the initial state perturbed by a relative 1e-15, stored in the
fixture, departs from its unperturbed run by 1.7e-10 (rho), 1.1e-9
(momenta), 1.3e-7 (z-momentum, zero at the start) and 2.6e-10 (energy)
after the 20 steps).  So the engine's sampled values must lie within three
times the oracle's own spread per equation (and within 1e-10 where that is
smaller); the conservation sums, which only rounding moves, within 4e-12 of
the equation's L1 scale (for the momenta, which sum to about zero, the L1
scale of all three)."""
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "c4_lte_tgv24_p3_20steps.npz"


@pytest.fixture(scope="module")
def table():
    from theseus_amd.lte import synthetic_table

    return synthetic_table()


@pytest.mark.parametrize("viscous", [False, True])
def test_c4_full_size_steps_match_the_oracle_run(table, viscous):
    import torch

    from theseus_amd.cases import c4_lte_tgv
    from theseus_amd.operator import MakeRHSOperator

    key = "ns" if viscous else "euler"
    print(f"case C4 LTE TGV 24^3 p=3 {key} FP64, 20 fixed steps vs oracle fixture", flush=True)
    sys.stdout.flush()
    f = np.load(FIXTURE)
    rt, mesh, U0, _ = c4_lte_tgv(n=24, order=3, table=table, viscous=viscous)
    dt = float(f[f"{key}_dt"])
    op = MakeRHSOperator(rt, mesh, lte_table=table)
    u = torch.from_numpy(U0).cuda()
    for _ in range(int(f["nsteps"])):
        op.Step(u, 0.0, dt)
    torch.cuda.synchronize()
    op.close()
    U = u.cpu().numpy()
    stride = int(f["stride"])
    sample = U[::stride]
    ref = f[f"{key}_sample"]
    Ntot = U.size // 5
    idx = np.arange(0, U.size, stride)
    pert = f[f"{key}_perturbed_sample"]
    errs, spread = [], []
    for q in range(5):
        m = (idx >= q * Ntot) & (idx < (q + 1) * Ntot)
        nrm = np.linalg.norm(ref[m])
        errs.append(float(np.linalg.norm(sample[m] - ref[m]) / nrm))
        spread.append(float(np.linalg.norm(pert[m] - ref[m]) / nrm))
    bound = [max(1e-10, 3 * sp) for sp in spread]
    Uq = U.reshape(5, -1)
    norms = np.linalg.norm(Uq, axis=1)
    sums = Uq.sum(axis=1)
    l1 = np.abs(Uq).sum(axis=1)
    print(f"  sampled rel L2 per equation {errs}; the oracle's own spread {spread}", flush=True)
    assert all(e <= b for e, b in zip(errs, bound)), (errs, bound)
    np.testing.assert_allclose(norms, f[f"{key}_norms"], rtol=max(bound), atol=0)
    scale = np.array([l1[0], l1[1:4].sum(), l1[1:4].sum(), l1[1:4].sum(), l1[4]])
    assert (np.abs(sums - f[f"{key}_sums"]) <= 4e-12 * scale).all(), (sums - f[f"{key}_sums"], scale)
