"""Oracle fixture for the full-size 3-D Navier-Stokes Taylor-Green workload
(32^3 hexes, p=3, Ranocha volume + Lax-Friedrichs face flux, BR1, Ma 0.1,
Re 1600, Pr 0.71): 100 SSP-RK3 steps at the fixed dt of CFL 0.5 on the
initial state, run by the CPU oracle (test infrastructure).  The state has
10,485,760 values, too many to keep, so the fixture holds dt, every
STRIDE-th value of the final state and per-equation sums and L2 norms of the
whole final state; tests/test_gpu_steps_c2.py compares the engine's state
with them.  Takes a few minutes on 8 cores."""
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import pyoracle  # noqa: E402
from theseus_amd.cases import c2_tgv3d  # noqa: E402
from theseus_amd.operator import build_descriptor  # noqa: E402

OUT = HERE / "c2_tgv32_p3_100steps.npz"
STRIDE = 1009
NSTEPS = 100
CFL = 0.5


def main():
    if not pyoracle.LIB.exists():
        pyoracle.build()
    lib = pyoracle.load()
    lib.oracle_set_num_threads(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    rt, mesh, U0 = c2_tgv3d()
    desc, keep = build_descriptor(rt, mesh)
    case = pyoracle.OracleCase(desc)
    r = case.stability(U0)
    dt = CFL / (max(r[0], r[1]) + r[2])
    U = U0.copy()
    t0 = time.time()
    for s in range(NSTEPS):
        U = case.step(U, dt)
        if s % 10 == 9:
            print(f"step {s + 1} ({time.time() - t0:.0f} s)", flush=True)
    Uq = U.reshape(5, -1)
    np.savez_compressed(OUT, dt=dt, nsteps=NSTEPS, stride=STRIDE, sample=U[::STRIDE],
                        sums=Uq.sum(axis=1), norms=np.linalg.norm(Uq, axis=1))
    print(OUT, "dt", dt)


if __name__ == "__main__":
    sys.exit(main())
