"""Convert the reference's golden isentropic-vortex output (VTU point data,
TestCases/GoldenData/IsentropicVortex/Cycle000500/proc00000{0,1}.vtu) into
tests/golden/isentropic_vortex_cycle500.npz.

The VTU files are data written by the reference's own run (24x24 quads,
p=3, Chandrashekar flux, blending on, dt=1e-3, 500 SSP-RK3 steps, 2 ranks);
this script only decodes them (inline base64, UInt32 byte-count header).
"""
import base64
import re
import sys
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/TestCases/GoldenData/IsentropicVortex/Cycle000500")
OUT = Path(__file__).resolve().parent / "isentropic_vortex_cycle500.npz"


def _arrays(text):
    out = {}
    for m in re.finditer(r'<DataArray([^>]*)>\s*([^<]*)</DataArray>', text):
        attrs, payload = m.group(1), m.group(2).strip()
        name = re.search(r'Name="([^"]*)"', attrs)
        name = name.group(1) if name else "Points"
        typ = re.search(r'type="([^"]*)"', attrs).group(1)
        ncomp = int((re.search(r'NumberOfComponents="(\d+)"', attrs) or [None, "1"])[1])
        head = base64.b64decode(payload[:8])
        nbytes = int(np.frombuffer(head, dtype="<u4")[0])
        data = base64.b64decode(payload[8:])[:nbytes]
        dt = {"Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}[typ]
        arr = np.frombuffer(data, dtype=dt)
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
