"""CuArray semantics, ported from the reference's cuarray tests
(tests/cuarray/python/test_FloatCuArray.py:10-73, conftest.py:6-21,
tests/cuarray/cpp/FloatCuArray_test.cpp:8-279). Host-side only: device
transfers need a GPU and are covered by the GPU tests."""
import numpy as np
import pytest

from netsci_amd import cuarray


@pytest.fixture(scope="module")
def no_data():
    a = cuarray.FloatCuArray()
    a.init(20, 100)
    return a


@pytest.fixture(scope="module")
def with_data():
    a = cuarray.FloatCuArray()
    a.fromNumpy2D(np.random.default_rng(0).random((20, 100)).astype(np.float32))
    return a


def test_shape_and_bytes(no_data):
    assert no_data.m() == 20
    assert no_data.n() == 100
    assert no_data.bytes() == 8000
    assert no_data.size() == 2000
    assert len(no_data) == 20


def test_allocation_flags(no_data):
    assert no_data.allocatedHost() == 1
    assert no_data.allocatedDevice() == 0
    assert no_data.owner() == 1


def test_zero_initialised(no_data):
    assert (no_data.toNumpy2D() == 0).all()
    assert all(no_data.get(i, j) == 0.0 for i in range(20) for j in range(0, 100, 7))


def test_to_device_without_allocation_is_noop(no_data):
    """toDevice/toHost return 1 and change nothing unless both sides exist
    (cuarray.cu:124-146)."""
    assert no_data.toDevice() == 1
    assert no_data.allocatedDevice() == 0
    assert no_data.toHost() == 1
    assert no_data.deallocateDevice() == 0


def test_from_cuarray_deep_copy(with_data):
    """test_FloatCuArray.py:53-73 (keyword arguments)."""
    f = cuarray.FloatCuArray()
    f.fromCuArray(cuArray=with_data, start=18, end=19, m=2, n=100)
    assert (f.m(), f.n(), f.bytes(), f.size()) == (2, 100, 800, 200)
    assert f.allocatedHost() == 1 and f.allocatedDevice() == 0 and f.owner() == 1
    for i in range(2):
        for j in range(100):
            assert f.get(i, j) == with_data.get(18 + i, j)


def test_from_cuarray_shape_mismatch_returns_1(with_data):
    f = cuarray.FloatCuArray()
    assert f.fromCuArray(cuArray=with_data, start=0, end=1, m=3, n=100) == 1


def test_set_is_bounds_checked():
    a = cuarray.IntCuArray()
    a.init(2, 3)
    assert a.set(7, 1, 2) == 0
    assert a.get(1, 2) == 7
    assert a.set(9, 2, 0) == 1
    assert a.set(9, 0, 3) == 1


def test_numpy_round_trip():
    x = np.arange(12, dtype=np.float32).reshape(3, 4)
    a = cuarray.FloatCuArray()
    a.fromNumpy2D(x)
    np.testing.assert_array_equal(a.toNumpy2D(), x)
    np.testing.assert_array_equal(a.toNumpy1D(), x.ravel())
    b = cuarray.IntCuArray()
    b.fromNumpy1D(np.array([3, 1, 2], np.int32))
    assert (b.m(), b.n()) == (1, 3)


def test_rows_index_and_assign():
    a = cuarray.IntCuArray()
    a.init(2, 3)
    a[1][2] = 5
    assert a[1][2] == 5 and len(a[1]) == 3
    with pytest.raises(IndexError):
        a[2]
    with pytest.raises(IndexError):
        a[0][3]


def test_sort_and_argsort_descending():
    a = cuarray.FloatCuArray()
    a.fromNumpy2D(np.array([[0.5, 2.0, -1.0, 1.5]], np.float32))
    np.testing.assert_array_equal(a.sort(0).toNumpy1D(), [2.0, 1.5, 0.5, -1.0])
    np.testing.assert_array_equal(a.argsort(0).toNumpy1D(), [1, 3, 0, 2])


def test_save_load_npy(tmp_path):
    x = np.random.default_rng(1).random((5, 7)).astype(np.float32)
    a = cuarray.FloatCuArray()
    a.fromNumpy2D(x)
    f = str(tmp_path / "a.npy")
    a.save(f)
    np.testing.assert_array_equal(np.load(f), x)
    b = cuarray.FloatCuArray()
    b.load(f)
    np.testing.assert_array_equal(b.toNumpy2D(), x)
    np.save(str(tmp_path / "d.npy"), x.astype(np.float64).ravel())
    c = cuarray.FloatCuArray()
    c.load(str(tmp_path / "d.npy"))
    assert (c.m(), c.n()) == (1, 35)
    np.testing.assert_array_equal(c.toNumpy1D(), x.ravel())


def test_repr():
    a = cuarray.IntCuArray()
    a.fromNumpy2D(np.array([[1, 22], [3, 4]], np.int32))
    assert repr(a).splitlines()[0].startswith("1")
    assert repr(cuarray.FloatCuArray()) == "null"
