"""Checkpointed GC on the GPU platform (GPU required): the file protocol of
generalized_correlation.cpp:8-101, the values equal to the oracle's pair
loop (the reference CPU path restated) and to the CPU platform's run of the
same call."""
import os

import numpy as np
import pytest

import oracle
from netsci_amd import cuarray, netcalc, synth

pytestmark = pytest.mark.gpu


def _fa(x):
    c = cuarray.FloatCuArray()
    c.fromNumpy2D(np.ascontiguousarray(x, np.float32))
    return c


def _ia(x):
    c = cuarray.IntCuArray()
    c.fromNumpy2D(np.ascontiguousarray(x, np.int32))
    return c


def test_gpu_checkpoints_equal_cpu(tmp_path):
    print("checkpoint gpu", flush=True)
    N, n, d, k = 7, 180, 3, 4
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=12)
    ab = synth.all_pairs(N)  # 21 pairs
    out = {}
    for plat, sub in ((netcalc.GPU_PLATFORM, "gpu"), (netcalc.CPU_PLATFORM, "cpu")):
        os.makedirs(tmp_path / sub)
        base = str(tmp_path / sub / "ck.npy")
        R = cuarray.FloatCuArray()
        rc = netcalc.generalizedCorrelationWithCheckpointing(
            X=_fa(X), R=R, ab=_ia(ab), k=k, n=n, xd=2, d=d, platform=plat,
            checkpointFrequency=5, checkpointFileName=base)
        assert rc == plat
        out[sub] = (R.toNumpy1D(), sorted(os.listdir(tmp_path / sub)),
                    np.load(str(tmp_path / sub / "ck_20.npy")))
    want = oracle.pairs(X, ab, k, n, d, gc=True)
    np.testing.assert_array_equal(out["gpu"][0], want)
    np.testing.assert_array_equal(out["gpu"][0], out["cpu"][0])
    # The last checkpoint (after pair 20) holds every value.
    np.testing.assert_array_equal(out["gpu"][2].ravel(), want)
    assert out["gpu"][1] == out["cpu"][1] == ["ck_20.npy", "ck_ab.npy"]
    np.testing.assert_array_equal(out["gpu"][2], out["cpu"][2])
