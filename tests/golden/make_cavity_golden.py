"""Convert the reference's golden lid-driven cavity output (VTU point data,
TestCases/GoldenData/LidDrivenCavity/Cycle004000/proc00000{0,1}.vtu) into
tests/golden/lid_driven_cavity_cycle4000.npz.

The VTU files are data written by the reference's own run of
TestCases/NavierStokes/2D/LidDrivenCavity/config.json (16x16 quads on
[0,2]^2, p=3, Chandrashekar flux, BR1, no-slip adiabatic walls with a lid
moving at (1, 0), dt=1e-4, 4000 SSP-RK3 steps, 2 ranks); this script only
decodes them (inline base64, UInt32 byte-count header).
"""
import base64
import re
import sys
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/TestCases/GoldenData/LidDrivenCavity/Cycle004000")
OUT = Path(__file__).resolve().parent / "lid_driven_cavity_cycle4000.npz"


def _arrays(text):
    out = {}
    for m in re.finditer(r'<DataArray([^>]*)>\s*([^<]*)</DataArray>', text):
        attrs, payload = m.group(1), m.group(2).strip()
        name = re.search(r'Name="([^"]*)"', attrs)
        name = name.group(1) if name else "Points"
        typ = re.search(r'type="([^"]*)"', attrs).group(1)
        ncomp = int((re.search(r'NumberOfComponents="(\d+)"', attrs) or [None, "1"])[1])
        nbytes = int(np.frombuffer(base64.b64decode(payload[:8]), dtype="<u4")[0])
        data = base64.b64decode(payload[8:])[:nbytes]
        arr = np.frombuffer(data, dtype={"Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}[typ])
        out[name] = arr.reshape(-1, ncomp) if ncomp > 1 else arr
    return out


def main():
    parts = [_arrays((SRC / f"proc00000{r}.vtu").read_text()) for r in (0, 1)]
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("Points", "Density", "Pressure", "Velocity",
                                                              "Blending Coeff")}
    np.savez_compressed(OUT, x=cat["Points"][:, 0], y=cat["Points"][:, 1], rho=cat["Density"],
                        p=cat["Pressure"], u=cat["Velocity"][:, 0], v=cat["Velocity"][:, 1],
                        alpha=cat["Blending Coeff"])
    print(OUT, cat["Points"].shape, "alpha max", cat["Blending Coeff"].max())


if __name__ == "__main__":
    sys.exit(main())
