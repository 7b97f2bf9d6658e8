"""CuArray device transfers (GPU required): the reference's
test_FloatCuArrayNoData_toDevicePositive
(tests/cuarray/python/test_FloatCuArray.py:46-50) and a host -> device ->
host round trip through libnsx's allocator and pinned staging."""
import numpy as np
import pytest

from netsci_amd import cuarray

pytestmark = pytest.mark.gpu


def test_to_device_positive():
    print("cuarray device", flush=True)
    a = cuarray.FloatCuArray()
    a.init(20, 100)
    assert a.allocateDevice() == 0
    assert a.toDevice() == 0
    assert a.allocatedDevice() == 1
    assert a.deallocateDevice() == 0
    assert a.allocatedDevice() == 0


@pytest.mark.parametrize("cls,dtype", [(cuarray.FloatCuArray, np.float32),
                                       (cuarray.IntCuArray, np.int32)])
def test_round_trip_restores_the_device_copy(cls, dtype):
    """toHost() brings back what toDevice() sent, after the host copy was
    changed in between; 10 MB crosses the 8 MB pinned staging buffer in two
    pieces each way."""
    rng = np.random.default_rng(1)
    data = (rng.standard_normal((2560, 1024)) * 100).astype(dtype)
    a = cls()
    a.fromNumpy2D(data)
    a.allocateDevice()
    assert a.toDevice() == 0
    a.set(dtype(7), 0, 0)
    a.set(dtype(9), 2559, 1023)
    assert a.toHost() == 0
    np.testing.assert_array_equal(a.toNumpy2D(), data)
    a.deallocateDevice()
