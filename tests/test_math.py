"""The FP64 exp and log of the device kernels, built for the host and checked
against the host libm, which the reference's numba kernels call for np.exp and
np.log (glibc on Linux).

exp restates the table-driven method glibc has used since 2.28 with the
multiply-adds fused as in its FMA build, so it must agree in every bit on any
argument; log (Tang's table-driven method) is correctly rounded except near
halfway cases and may differ from glibc's in its last bit on a small fraction
of arguments."""
from __future__ import annotations

import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "finitewave_amd" / "csrc"


def _host_has_fma() -> bool:
    try:
        return " fma " in Path("/proc/cpuinfo").read_text().replace("\n", " ")
    except OSError:
        return False


@pytest.fixture(scope="module")
def mathlib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("fwmath") / "libfwmath.so"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I", str(CSRC), str(ROOT / "tests" / "native" / "fw_math_check.cpp"),
                    "-o", str(out), "-lm"], check=True)
    lib = ctypes.CDLL(str(out))
    for f in (lib.fw_check_exp, lib.fw_check_log):
        f.restype = ctypes.c_int64
        f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    return lib


def _run(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    mine = np.empty_like(x)
    ref = np.empty_like(x)
    bad = fn(x.ctypes.data, x.size, mine.ctypes.data, ref.ctypes.data)
    return bad, mine, ref


SPECIAL = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-300, -1e-300, 2.0 ** -54, -2.0 ** -54,
           2.0 ** -55, 512.0, -512.0, 709.782712893384, 709.7827128933841,
           -708.3964185322641, -745.1332191019411, -745.1332191019412, -746.0, 710.0, 1.0,
           -1.0, 0.6931471805599453, 5e-324, -5e-324]


@pytest.mark.skipif(not _host_has_fma(), reason="the host libm's non-FMA exp rounds its "
                    "polynomial differently (about 7e-4 of arguments)")
def test_exp_equals_host_libm_bit_for_bit(mathlib):
    rng = np.random.default_rng(3)
    xs = [np.array(SPECIAL),
          rng.uniform(-100, 100, 400_000),          # ionic-model rate arguments
          rng.uniform(-1, 1, 200_000),
          rng.uniform(-745.2, -700, 100_000),        # results in the subnormal range
          rng.uniform(512, 709.8, 50_000),
          rng.uniform(-2e-16, 2e-16, 50_000)]
    # random bit patterns with |x| < 2048 (every reduction index and scale)
    u = rng.integers(0, 2 ** 63, 400_000, dtype=np.uint64)
    e = (u >> np.uint64(52)) & np.uint64(0x7FF)
    u = (u & ~(np.uint64(0x7FF) << np.uint64(52))) | ((np.uint64(900) + e % np.uint64(134))
                                                      << np.uint64(52))
    u |= rng.integers(0, 2, u.size, dtype=np.uint64) << np.uint64(63)
    xs.append(u.view(np.float64))
    for x in xs:
        bad, mine, ref = _run(mathlib.fw_check_exp, x)
        assert bad == 0, x[np.flatnonzero(mine.view(np.uint64) != ref.view(np.uint64))[:5]]


def test_exp_accuracy_against_closed_forms(mathlib):
    """Independent of the host libm: exact powers of e are not representable,
    but exp(k ln 2) for the double nearest k ln 2 is within 1 ulp of 2^k, exp
    of a sum equals the product within 2 ulp, and exp(0) = 1."""
    k = np.arange(-1020, 1020, 7, dtype=np.float64)
    x = k * np.log(2.0)
    _, mine, _ = _run(mathlib.fw_check_exp, x)
    # x differs from k ln 2 by at most half an ulp of x; exp carries that
    # relative error plus its own rounding
    dx = np.abs(x.astype(np.longdouble) - k.astype(np.longdouble) *
                np.log(np.longdouble(2))).astype(np.float64)
    assert np.all(np.abs(mine / np.exp2(k) - 1) <= dx * 1.001 + 2.3e-16)
    rng = np.random.default_rng(5)
    a = rng.uniform(-20, 20, 10_000)
    b = rng.uniform(-20, 20, 10_000)
    _, ea, _ = _run(mathlib.fw_check_exp, a)
    _, eb, _ = _run(mathlib.fw_check_exp, b)
    _, eab, _ = _run(mathlib.fw_check_exp, a + b)
    # a + b is rounded: |d exp| <= ulp(a + b) * exp(a + b)
    tol = 4.5e-16 + np.spacing(np.abs(a + b))
    assert np.all(np.abs(ea * eb / eab - 1) <= tol)
    _, one, _ = _run(mathlib.fw_check_exp, np.array([0.0, -0.0]))
    assert np.all(one == 1.0)


def test_log_is_within_rounding_of_host_libm(mathlib):
    """Measured: 9e-5 of arguments in (0, 200] and 2e-3 near 1 (where glibc
    switches to a polynomial in x - 1) differ, each by one ulp."""
    rng = np.random.default_rng(4)
    for x, rate in ((rng.uniform(1e-6, 200.0, 400_000), 2e-4),
                    (rng.uniform(0.9, 1.1, 200_000), 4e-3),
                    (np.exp(rng.uniform(-700, 700, 100_000)), 2e-4)):
        bad, mine, ref = _run(mathlib.fw_check_log, x)
        assert bad <= rate * x.size
        assert np.all(np.abs(mine - ref) <= np.spacing(np.abs(ref)))
    _, s, _ = _run(mathlib.fw_check_log, np.array([1.0, 0.0, -1.0, np.inf, np.nan]))
    assert s[0] == 0.0 and s[1] == -np.inf and np.isnan(s[2]) and s[3] == np.inf
    assert np.isnan(s[4])
