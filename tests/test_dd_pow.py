"""The device's pow (double-double, nebulasem_amd/csrc/hip/dd_math.hpp),
built for the host from the same source, against exact arithmetic and
against glibc's pow, which the reference's equation of state calls.

* correctly rounded on every sample: the exact x^y from 60-digit decimal
  arithmetic (Python's decimal ln / exp are correctly rounded), rounded to
  the nearest double, equals pow_dd(x, y);
* against glibc: the same result wherever glibc rounds correctly; the count
  of samples where they differ is reported (glibc's pow is documented below
  0.52 ulp, so it misrounds a small fraction of inputs lying within a few
  thousandths of an ulp of a rounding boundary);
* the equation-of-state inputs: (rho theta R / P0) in [0.5, 1.5] with
  gamma = cp/cv, 1/gamma and R/cp as the Euler path uses them;
* the two evaluations of Ziv's strategy: the first one's error stays within
  the bound its rounding test assumes (with a factor 4 to spare), and
  pow_dd returns exactly what the second evaluation alone returns."""
import ctypes as C
import os
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CP, CV = 1004.67, 715.5


@pytest.fixture(scope="module")
def ddlib(tmp_path_factory):
    out = tmp_path_factory.mktemp("ddpow") / "libddpow.so"
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(out),
                           os.path.join(HERE, "native", "dd_pow_host.cpp")])
    lib = C.CDLL(str(out))
    for f in (lib.dd_pow_many, lib.libm_pow_many, lib.dd_pow_accurate_many):
        f.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int64]
    lib.dd_pow_first_eval.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_int64]
    return lib


def run(lib, fn, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    getattr(lib, fn)(x.ctypes.data, C.c_double(y), out.ctypes.data, x.size)
    return out


def exact_pow(x, y):
    getcontext().prec = 60
    return float((Decimal(float(x)).ln() * Decimal(float(y))).exp())


@pytest.mark.parametrize("y", [CP / CV, CV / CP, (CP - CV) / CP, 1.0 / 3.0, 2.5, -0.7])
def test_pow_dd_correctly_rounded(ddlib, y):
    rng = np.random.default_rng(12345)
    x = np.concatenate([rng.uniform(0.5, 1.5, 1500), np.exp(rng.uniform(-30, 30, 500))])
    got = run(ddlib, "dd_pow_many", x, y)
    want = np.array([exact_pow(v, y) for v in x])
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(x[i], got[i], want[i]) for i in bad[:5]]


def test_pow_dd_against_glibc_on_eos_inputs(ddlib):
    """Per exponent, the fraction of 200000 equation-of-state inputs where the
    correctly rounded pow_dd and glibc disagree (each disagreement is one ulp,
    glibc's misrounding, checked on the first few exactly)."""
    rng = np.random.default_rng(7)
    x = rng.uniform(0.6, 1.2, 200000)
    for y in (CP / CV, CV / CP, (CP - CV) / CP):
        a = run(ddlib, "dd_pow_many", x, y)
        b = run(ddlib, "libm_pow_many", x, y)
        diff = np.nonzero(a != b)[0]
        print(f"y = {y:.6f}: pow_dd vs glibc differ on {diff.size} of {x.size}")
        assert diff.size <= 2e-3 * x.size
        for i in diff[:5]:
            assert a[i] == exact_pow(x[i], y)
            assert abs(np.spacing(a[i])) >= abs(a[i] - b[i])


def test_pow_dd_special_cases(ddlib):
    x = np.array([0.0, -1.0, np.inf, np.nan, 1.0, 5e-324, 1e308])
    got = run(ddlib, "dd_pow_many", x, 1.4)
    want = run(ddlib, "libm_pow_many", x, 1.4)
    assert np.array_equal(got, want, equal_nan=True)


@pytest.mark.parametrize("y", [CP / CV, CV / CP, (CP - CV) / CP, 1.0 / 3.0, 2.5, -0.7, 30.0])
def test_pow_dd_first_evaluation(ddlib, y):
    """The first evaluation's relative error, measured against the 2^-90
    evaluation, is below (1 + |y|) 2^-73 on 1.5 M inputs over three ranges;
    its rounding test assumes (1 + |y|) 2^-71, so a slip of the series, the
    reduction or the Newton step that moved the error by a factor 4 fails
    here.  The rounding test passes on all but a small fraction, and
    wherever it passes the two evaluations give the same double."""
    rng = np.random.default_rng(2024)
    x = np.concatenate([rng.uniform(0.5, 1.5, 800000), np.exp(rng.uniform(-30, 30, 400000)),
                        np.exp(rng.uniform(-690, 690, 300000))])
    rel = np.empty_like(x)
    passed = np.zeros(x.size, np.int32)
    ddlib.dd_pow_first_eval(x.ctypes.data, C.c_double(y), rel.ctypes.data, passed.ctypes.data, x.size)
    inr = rel >= 0
    assert inr[:800000].all()  # |y log x| <= 680 there
    assert rel[inr].max() <= (1.0 + abs(y)) * 2.0 ** -73, np.log2(rel[inr].max())
    assert 1.0 - passed[inr].mean() <= (1e-3 if abs(y) > 3 else 1e-4)
    a = run(ddlib, "dd_pow_many", x, y)
    b = run(ddlib, "dd_pow_accurate_many", x, y)
    assert np.array_equal(a, b)
