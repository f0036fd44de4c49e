"""Convert the reference's golden forward-facing-step output
(TestCases/GoldenData/ForwardFacingStep.tgz: Cycle000100/proc00000{0,1}.vtu,
VTU point data written by the reference's own run of
TestCases/Euler/2D/ForwardFacingStep/config.json with fixed dt 1e-4 and 100
SSP-RK3 steps, tests/recipes/regression/golden-step.sh:18-27) into
tests/golden/forward_facing_step_cycle100.npz.

Most of the 13,600 elements still hold the uniform Mach-3 inflow state
(rho 1.4, u (3, 0), p 1) to round-off after 100 steps.  The fixture keeps
every node of each element that deviates from that state by more than
KEEP_ABOVE in any field, and for the remaining elements only the largest
deviation per field (so a test can still bound them).  The blending
coefficient is constant per element and is kept per kept element.  This
script only decodes data the reference wrote (inline base64, UInt32
byte-count header).
"""
import io
import sys
import tarfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from make_vortex_golden import _arrays  # noqa: E402

SRC = Path("/root/reference/TestCases/GoldenData/ForwardFacingStep.tgz")
OUT = HERE / "forward_facing_step_cycle100.npz"
NDOF = 16  # p = 3 quads
KEEP_ABOVE = 1e-13
FREESTREAM = {"rho": 1.4, "u": 3.0, "v": 0.0, "p": 1.0}


def main():
    parts = []
    with tarfile.open(SRC, "r:gz") as tf:
        for r in (0, 1):
            f = tf.extractfile(f"ForwardFacingStep/Cycle000100/proc00000{r}.vtu")
            parts.append(_arrays(io.TextIOWrapper(f, encoding="ascii").read()))
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("Points", "Density", "Pressure", "Velocity",
                                                              "Blending Coeff")}
    fields = {"rho": cat["Density"], "u": cat["Velocity"][:, 0], "v": cat["Velocity"][:, 1],
              "p": cat["Pressure"]}
    ne = fields["rho"].size // NDOF
    dev = {k: np.abs(v - FREESTREAM[k]).reshape(ne, NDOF).max(1) for k, v in fields.items()}
    keep = np.maximum.reduce(list(dev.values())) > KEEP_ABOVE
    alpha = cat["Blending Coeff"].reshape(ne, NDOF)
    assert (alpha == alpha[:, :1]).all()
    assert (alpha[~keep] == 0).all()
    nodes = np.repeat(keep, NDOF)
    out = {"x": cat["Points"][nodes, 0], "y": cat["Points"][nodes, 1],
           "alpha": alpha[keep, 0], "num_elements": np.int64(ne)}
    for k, v in fields.items():
        out[k] = v[nodes]
        out["rest_dev_" + k] = dev[k][~keep].max()
    np.savez_compressed(OUT, **out)
    print(OUT, "elements kept", int(keep.sum()), "of", ne, "alpha>0 in", int((alpha[:, 0] > 0).sum()),
          {k: float(out["rest_dev_" + k]) for k in fields})


if __name__ == "__main__":
    sys.exit(main())
