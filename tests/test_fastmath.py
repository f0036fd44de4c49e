"""Accuracy of the engine's own FP64 logarithm (theseus_amd/csrc/fastmath.hpp),
compiled for the host and compared with 40-digit arithmetic."""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def ln_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("fastmath")
    src = d / "ln.cpp"
    src.write_text('#include "fastmath.hpp"\n'
                   'extern "C" void ln_many(const double* x, double* y, long n)'
                   '{ for (long i = 0; i < n; ++i) y[i] = thx::ln_pos(x[i]); }\n')
    so = d / "ln.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", f"-I{ROOT / 'theseus_amd' / 'csrc'}", str(src),
                    "-o", str(so)], check=True)
    return ctypes.CDLL(str(so))


def test_ln_pos_accuracy(ln_lib):
    import mpmath as mp

    rng = np.random.default_rng(0)
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 3000)), 1 + rng.uniform(-1e-6, 1e-6, 1000),
                        rng.uniform(0.5, 2, 3000), [1.0, 2.0, 0.5, np.sqrt(0.5), 1e-300, 1e300]])
    y = np.empty_like(x)
    ln_lib.ln_many(x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(x.size))
    mp.mp.dps = 40
    worst = 0.0
    for xi, yi in zip(x, y):
        ex = mp.log(mp.mpf(xi))
        if ex == 0:
            assert yi == 0.0
            continue
        worst = max(worst, abs(float(mp.mpf(yi) - ex)) / np.spacing(abs(float(ex))))
    print("worst ulp", worst)
    assert worst < 1.6, worst
