"""Row mean / standard deviation (netcalc/include/math_utils.h:24-101) on the
CPU: the oracle pinned by closed forms and a known answer that fixes the
reference's summation order (math_utils_gpu.cu:23-156), and the CPU platform
against the oracle bit for bit."""
import numpy as np
import pytest

import oracle
from netsci_amd import cuarray, netcalc


def test_oracle_against_float64_statistics():
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((6, 3001)) * 3 + 7).astype(np.float32)
    u, s = oracle.row_std(a)
    ref = a.astype(np.float64)
    np.testing.assert_allclose(u, ref.mean(1), rtol=2e-6)
    np.testing.assert_allclose(s, ref.std(1), rtol=2e-5)  # population deviation (/ n)
    np.testing.assert_array_equal(oracle.row_mean(a), u)


def test_oracle_closed_forms():
    a = np.full((2, 3000), 2.0, np.float32)  # partial sums stay exact
    u, s = oracle.row_std(a)
    assert np.all(u == 2.0) and np.all(s == 0.0)
    b = np.tile(np.array([1.0, 3.0], np.float32), (1, 500))  # mean 2, deviation 1
    u, s = oracle.row_std(b)
    assert u[0] == 2.0 and s[0] == 1.0
    one = np.array([[5.5]], np.float32)
    u, s = oracle.row_std(one)
    assert u[0] == 5.5 and s[0] == 0.0


def test_oracle_summation_order_known_answer():
    """n = 1025: lane 0 of the reference's 1024 adds a[0] + a[1024] = 1e8 -
    1e8 = 0 and the tree adds the other 1023 ones exactly, so the mean is
    1023 / 1025. A sequential sum would lose every +1 against 1e8 (ulp 8)
    and give 0."""
    a = np.ones((1, 1025), np.float32)
    a[0, 0], a[0, 1024] = 1e8, -1e8
    u = oracle.row_mean(a)
    assert u[0] == np.float32(1023.0) / np.float32(1025.0)
    seq = np.float32(0.0)
    for v in a[0]:
        seq = np.float32(seq + v)
    assert seq == 0.0  # the order matters for this input


def _arr(a):
    c = cuarray.FloatCuArray()
    c.fromNumpy2D(np.ascontiguousarray(a, np.float32))
    return c


@pytest.mark.parametrize("shape", [(1, 1), (3, 7), (4, 1024), (5, 1025), (2, 40000)])
def test_cpu_platform_equals_oracle(shape):
    rng = np.random.default_rng(shape[1])
    a = (rng.standard_normal(shape) * 10 + 3).astype(np.float32)
    u, s = oracle.row_std(a)
    cu, cs, cm = cuarray.FloatCuArray(), cuarray.FloatCuArray(), cuarray.FloatCuArray()
    netcalc.standardDeviation(a=_arr(a), u=cu, sigma=cs, m=shape[0], n=shape[1],
                              platform=netcalc.CPU_PLATFORM)
    netcalc.mean(a=_arr(a), u=cm, m=shape[0], n=shape[1], platform=netcalc.CPU_PLATFORM)
    np.testing.assert_array_equal(cu.toNumpy1D(), u)
    np.testing.assert_array_equal(cs.toNumpy1D(), s)
    np.testing.assert_array_equal(cm.toNumpy1D(), u)


def test_rows_follow_the_array_size_and_bad_arguments_raise():
    """m is recomputed as a.size() / n (math_utils.cpp:14,32)."""
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    u = cuarray.FloatCuArray()
    netcalc.mean(a=_arr(a), u=u, m=99, n=6, platform=netcalc.CPU_PLATFORM)
    np.testing.assert_array_equal(u.toNumpy1D(), oracle.row_mean(a.reshape(2, 6)))
    with pytest.raises(ValueError):
        netcalc.mean(a=_arr(a), u=u, m=3, n=4, platform=7)
    with pytest.raises(ValueError):
        netcalc.mean(a=_arr(a), u=u, m=3, n=0, platform=netcalc.CPU_PLATFORM)
