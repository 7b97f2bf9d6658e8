"""Row mean / standard deviation on the GPU equal the oracle's restatement
of math_utils_gpu.cu:23-156 bit for bit (GPU required)."""
import numpy as np
import pytest

import oracle
from netsci_amd import abi, cuarray, netcalc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(1, 1), (3, 7), (4, 1024), (5, 1025), (1000, 6000),
                                   (2, 100003)])
def test_device_rows_bitexact(shape):
    print(f"row stats {shape}", flush=True)
    rng = np.random.default_rng(shape[0] + shape[1])
    a = (rng.standard_normal(shape) * 10 + 3).astype(np.float32)
    u, s = oracle.row_std(a)
    da = abi.DeviceArray.from_numpy(a)
    du = abi.DeviceArray((shape[0],), np.float32)
    ds = abi.DeviceArray((shape[0],), np.float32)
    abi.check(abi.lib().nsx_row_std(da.ptr, shape[0], shape[1], du.ptr, ds.ptr, None), "std")
    abi.sync()
    np.testing.assert_array_equal(du.numpy(), u)
    np.testing.assert_array_equal(ds.numpy(), s)
    dm = abi.DeviceArray((shape[0],), np.float32)
    abi.check(abi.lib().nsx_row_mean(da.ptr, shape[0], shape[1], dm.ptr, None), "mean")
    abi.sync()
    np.testing.assert_array_equal(dm.numpy(), u)


def test_summation_order_known_answer_gpu():
    a = np.ones((1, 1025), np.float32)
    a[0, 0], a[0, 1024] = 1e8, -1e8
    u = cuarray.FloatCuArray()
    A = cuarray.FloatCuArray()
    A.fromNumpy2D(a)
    netcalc.mean(a=A, u=u, m=1, n=1025, platform=netcalc.GPU_PLATFORM)
    assert u.toNumpy1D()[0] == np.float32(1023.0) / np.float32(1025.0)


def test_netcalc_gpu_equals_cpu_platform():
    rng = np.random.default_rng(5)
    a = (rng.standard_normal((17, 2049)) * 4).astype(np.float32)
    A = cuarray.FloatCuArray()
    A.fromNumpy2D(a)
    out = {}
    for plat in (netcalc.GPU_PLATFORM, netcalc.CPU_PLATFORM):
        u, s = cuarray.FloatCuArray(), cuarray.FloatCuArray()
        netcalc.standardDeviation(a=A, u=u, sigma=s, m=17, n=2049, platform=plat)
        out[plat] = (u.toNumpy1D(), s.toNumpy1D())
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
