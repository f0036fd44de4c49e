"""Accuracy of the engine's own FP64 logarithm and logarithmic mean
(theseus_amd/csrc/fastmath.hpp), compiled for the host and compared with
40-digit arithmetic."""
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
                   '{ for (long i = 0; i < n; ++i) y[i] = thx::ln_pos(x[i]); }\n'
                   'extern "C" void lm_many(const double* x, const double* y, double* lm, double* inv, long n)'
                   '{ for (long i = 0; i < n; ++i) { auto r = thx::log_mean_pair<double, true, true>(x[i], y[i]);'
                   ' lm[i] = r.lm; inv[i] = r.inv; } }\n'
                   'extern "C" void lm_many_f(const float* x, const float* y, float* lm, float* inv, long n)'
                   '{ for (long i = 0; i < n; ++i) { auto r = thx::log_mean_pair<float, true, true>(x[i], y[i]);'
                   ' lm[i] = r.lm; inv[i] = r.inv; } }\n')
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


def _lm_exact(x, y):
    import mpmath as mp

    x, y = mp.mpf(x), mp.mpf(y)
    return (y - x) / mp.log(y / x)


def test_log_mean_pair_accuracy(ln_lib):
    """Both branches (the atanh series for y/x in [0.707, 1.414], the logarithm
    beyond) within a few ulp of the exact logarithmic mean and its inverse;
    the reference's own formula (theseus_kernels.hpp:106-123) is within an ulp
    or two of the same value, so the two agree to rounding."""
    import mpmath as mp

    mp.mp.dps = 40
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-5, 5, 4000))
    ratio = np.concatenate([1 + rng.uniform(-1e-8, 1e-8, 500), 1 + rng.uniform(-0.02, 0.02, 1000),
                            np.exp(rng.uniform(-0.4, 0.4, 1500)), np.exp(rng.uniform(-6, 6, 1000))])
    y = x * ratio
    y[:3] = x[:3]  # equal arguments: LM = x
    lm = np.empty_like(x)
    inv = np.empty_like(x)
    ln_lib.lm_many(x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p),
                   lm.ctypes.data_as(ctypes.c_void_p), inv.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(x.size))
    worst_lm = worst_inv = 0.0
    for xi, yi, a, b in zip(x, y, lm, inv):
        ex = mp.mpf(xi) if xi == yi else _lm_exact(xi, yi)
        worst_lm = max(worst_lm, abs(float((mp.mpf(a) - ex) / ex)))
        worst_inv = max(worst_inv, abs(float((mp.mpf(b) - 1 / ex) * ex)))
    print("worst relative error", worst_lm, worst_inv)
    assert worst_lm < 1e-15 and worst_inv < 1e-15, (worst_lm, worst_inv)
    # FP32 variant: within a few FP32 ulp
    xf, yf = x.astype(np.float32), y.astype(np.float32)
    lmf = np.empty_like(xf)
    invf = np.empty_like(xf)
    ln_lib.lm_many_f(xf.ctypes.data_as(ctypes.c_void_p), yf.ctypes.data_as(ctypes.c_void_p),
                     lmf.ctypes.data_as(ctypes.c_void_p), invf.ctypes.data_as(ctypes.c_void_p), ctypes.c_long(xf.size))
    worst = 0.0
    for xi, yi, a in zip(xf[::7], yf[::7], lmf[::7]):
        ex = mp.mpf(float(xi)) if xi == yi else _lm_exact(float(xi), float(yi))
        worst = max(worst, abs(float((mp.mpf(float(a)) - ex) / ex)))
    assert worst < 8 * 2.0 ** -24, worst
