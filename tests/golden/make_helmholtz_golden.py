"""Extract the Helmholtz discrete-gradient golden vectors that the
reference's tests hold (tests/helmholtz_difference_cases.hpp and
tests/helmholtz_channel_difference_cases.hpp: one-cell tables, state pairs
and 80-digit expected means) into tests/golden/helmholtz_means.npz.

Only the numbers are taken (inputs and expected outputs); run from the
repository root with the reference at /root/reference:
    python tests/golden/make_helmholtz_golden.py
"""
import re
from pathlib import Path

import numpy as np

REF = Path("/root/reference/tests")
OUT = Path(__file__).resolve().parent / "helmholtz_means.npz"
NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def cases(path):
    text = path.read_text()
    body = text[text.index("cases[] = {"):]
    out = []
    # each case: four brace groups of numbers: bounds(4), states(4), jets(16), expected(2)
    for m in re.finditer(r"\{\s*(\{[^{}]*\}\s*,\s*\{[^{}]*\}\s*,\s*\{[^{}]*\}\s*,\s*\{[^{}]*\})\s*,?\s*\}", body):
        groups = re.findall(r"\{([^{}]*)\}", m.group(1))
        vals = [[float(v) for v in re.findall(NUM, g)] for g in groups]
        assert [len(v) for v in vals] == [4, 4, 16, 2], [len(v) for v in vals]
        out.append(vals)
    return out


def main():
    data = {}
    for key, fname in (("difference", "helmholtz_difference_cases.hpp"),
                       ("channel", "helmholtz_channel_difference_cases.hpp")):
        cs = cases(REF / fname)
        data[f"{key}_bounds"] = np.array([c[0] for c in cs])
        data[f"{key}_states"] = np.array([c[1] for c in cs])
        data[f"{key}_jets"] = np.array([c[2] for c in cs])
        data[f"{key}_expected"] = np.array([c[3] for c in cs])
        print(key, len(cs), "cases")
    np.savez(OUT, **data)


if __name__ == "__main__":
    main()
