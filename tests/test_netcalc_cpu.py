"""netcalc with platform=CPU_PLATFORM against the oracle, plus the API
contract (return values, errors, kwargs, Settings, checkpoint/restart).
Ports the CPU half of tests/netcalc/python/test_{MutualInformation,
GeneralizedCorrelation}.py and adds what those leave open."""
import os

import numpy as np
import pytest

import oracle
from netsci_amd import abi, cuarray, netcalc, synth
from tests.conftest import curve_inputs


def _fa(x):
    a = cuarray.FloatCuArray()
    x = np.asarray(x, np.float32)
    (a.fromNumpy2D if x.ndim == 2 else a.fromNumpy1D)(x)
    return a


def _ia(x):
    a = cuarray.IntCuArray()
    a.fromNumpy2D(np.asarray(x, np.int32))
    return a


def test_golden_cpu(golden_1000, golden_2000):
    """test_MutualInformation.py:13-64, CPU half."""
    for X, n, want in ((golden_1000, 1000, 0.8449974060058594),
                       (golden_2000, 2000, 0.8485465049743652)):
        got = netcalc.mutualInformationCpu(Xa=_fa(X[:n]), Xb=_fa(X[n:]), k=4, n=n, xd=2, d=1)
        assert got == want


@pytest.mark.parametrize("d,n,variant", [(1, 600, "gc"), (2, 700, "mi"), (3, 500, "gc")])
def test_curves_cpu_equals_oracle(d, n, variant):
    xa, xb = curve_inputs(n, d, variant)
    got = netcalc.generalizedCorrelationCpu(Xa=_fa(xa.reshape(d, n)), Xb=_fa(xb.reshape(d, n)),
                                            k=4, n=n, xd=2, d=d)
    assert got == oracle.gc_from_mi(oracle.mi(xa, xb, 4, n, d), d)


def test_pair_loop_cpu_equals_oracle_with_ties():
    """Quantised coordinates (frequent k-th-neighbour ties): the product CPU
    path's O(n) libstdc++-order selection equals std::sort's."""
    rng = np.random.default_rng(3)
    n, d, k = 300, 3, 6
    X = (np.round(rng.standard_normal((5, d * n)) * 3) / 3).astype(np.float32)
    ab = synth.all_pairs(5)
    got = abi.cpu_ksg_mi(X, ab, k, n, d)
    np.testing.assert_array_equal(got, oracle.pairs(X, ab, k, n, d, gc=False))


def test_platform_return_values():
    """test_MutualInformation.py:157-187 / test_GeneralizedCorrelation.py:229-259."""
    n = 1000
    xa, _ = curve_inputs(n, 2, "mi")
    X = _fa(xa.reshape(2, n))
    ab = cuarray.IntCuArray()
    ab.init(1, 2)
    ab.set(0, 0, 0)
    ab.set(1, 0, 1)
    I, R = cuarray.FloatCuArray(), cuarray.FloatCuArray()
    assert netcalc.mutualInformation(X=X, I=I, ab=ab, k=4, n=n, xd=2, d=1,
                                     platform=netcalc.CPU_PLATFORM) == 1
    assert netcalc.generalizedCorrelation(X=X, R=R, ab=ab, k=4, n=n, xd=2, d=1,
                                          platform=netcalc.CPU_PLATFORM) == 1
    assert (I.m(), I.n()) == (1, 1) and (R.m(), R.n()) == (1, 1)


def test_invalid_platform_and_arguments_raise():
    X = _fa(np.zeros((2, 100)))
    ab = _ia([[0, 1]])
    R = cuarray.FloatCuArray()
    with pytest.raises(RuntimeError, match="Invalid platform"):
        netcalc.generalizedCorrelation(X=X, R=R, ab=ab, k=4, n=100, xd=2, d=1, platform=7)
    with pytest.raises(ValueError):
        netcalc.generalizedCorrelation(X=X, R=R, ab=ab, k=100, n=100, xd=2, d=1, platform=1)
    with pytest.raises(ValueError):
        netcalc.generalizedCorrelation(X=X, R=R, ab=_ia([[0, 5]]), k=4, n=100, xd=2, d=1,
                                       platform=1)
    with pytest.raises(ValueError):
        netcalc.generalizedCorrelation(X=X, R=R, ab=ab, k=4, n=100, xd=2, d=2, platform=1)
    # xd: the reference's GPU kernels exist only for two variables
    # (mutual_information_gpu.cu:1015-1030); the check runs before any device
    # call. Its CPU path ignores xd, and so does the CPU platform.
    for gpu_call in (
            lambda: netcalc.generalizedCorrelation(X=X, R=R, ab=ab, k=4, n=100, xd=3, d=1,
                                                   platform=netcalc.GPU_PLATFORM),
            lambda: netcalc.mutualInformation(X=X, I=R, ab=ab, k=4, n=100, xd=1, d=1,
                                              platform=netcalc.GPU_PLATFORM),
            lambda: netcalc.mutualInformationGpu(Xa=_fa(np.zeros(100)), Xb=_fa(np.zeros(100)),
                                                 k=4, n=100, xd=3, d=1)):
        with pytest.raises(ValueError, match="xd must be 2"):
            gpu_call()
    Xr = _fa(np.random.default_rng(1).standard_normal((2, 100)))
    vals = []
    for xd in (2, 3):
        netcalc.mutualInformation(X=Xr, I=R, ab=ab, k=4, n=100, xd=xd, d=1,
                                  platform=netcalc.CPU_PLATFORM)
        vals.append(R.toNumpy1D()[0])
    assert vals[0] == vals[1]
    A, H, P = _fa(np.zeros((3, 3))), cuarray.FloatCuArray(), cuarray.IntCuArray()
    with pytest.raises(ValueError, match="Invalid platform"):
        netcalc.hedetniemiAllShortestPaths(A=A, H=H, paths=P, maxPathLength=2, tolerance=0.1,
                                           platform=3)


def test_gpu_platform_without_device_fails_loudly():
    if abi.device_count() > 0:
        pytest.skip("a GPU is present")
    X = _fa(np.zeros((2, 100)))
    with pytest.raises(RuntimeError, match="HIP device"):
        netcalc.generalizedCorrelation(X=X, R=cuarray.FloatCuArray(), ab=_ia([[0, 1]]), k=4,
                                       n=100, xd=2, d=1, platform=netcalc.GPU_PLATFORM)


def test_settings_autofill_and_splat():
    """netcalc/__init__.py:74-128: empty ab -> all N^2 ordered pairs; the
    settings splat into the call as keyword arguments."""
    N, n = 4, 200
    Xn, _ = synth.calpha_trajectory(N=N, n=n, seed=5)
    s = netcalc.GeneralizedCorrelationSettings(X=_fa(Xn), R=cuarray.FloatCuArray(),
                                               ab=cuarray.IntCuArray(), k=4, n=n, xd=2, d=3,
                                               platform=netcalc.CPU_PLATFORM)
    ab = s.ab.toNumpy2D()
    assert ab.shape == (16, 2)
    assert ab[5].tolist() == [1, 1] and ab[6].tolist() == [1, 2]
    assert sorted(s.keys()) == sorted(["X", "R", "ab", "k", "n", "xd", "d", "platform"])
    assert netcalc.generalizedCorrelation(**s) == 1
    got = s.R.toNumpy1D()
    np.testing.assert_array_equal(got, oracle.pairs(Xn, ab, 4, n, 3, gc=True))


def test_checkpoint_files_and_restart(tmp_path):
    """generalized_correlation.cpp:8-101 file protocol and
    mutual_information.cpp:174-201 restart: an interrupted run resumed from
    its last checkpoint gives the uninterrupted result."""
    N, n, d, k = 6, 150, 3, 4
    Xn, _ = synth.calpha_trajectory(N=N, n=n, seed=6)
    ab = synth.all_pairs(N)  # 15 pairs
    base = str(tmp_path / "ckpt.npy")
    R = cuarray.FloatCuArray()
    assert netcalc.generalizedCorrelationWithCheckpointing(
        X=_fa(Xn), R=R, ab=_ia(ab), k=k, n=n, xd=2, d=d, platform=netcalc.CPU_PLATFORM,
        checkpointFrequency=4, checkpointFileName=base) == 1
    files = sorted(os.listdir(tmp_path))
    assert files == ["ckpt_12.npy", "ckpt_ab.npy"]  # checkpoints 0,4,8 removed
    full = R.toNumpy1D()
    np.testing.assert_array_equal(full, oracle.pairs(Xn, ab, k, n, d, gc=True))
    last = np.load(str(tmp_path / "ckpt_12.npy")).ravel()
    np.testing.assert_array_equal(last[:13], full[:13])
    assert (last[13:] == 0).all()
    np.testing.assert_array_equal(np.load(str(tmp_path / "ckpt_ab.npy")), ab)
    # Resume after the checkpoint as the reference's restart helper does.
    rest = cuarray.IntCuArray()
    netcalc.generateRestartAbFromCheckpointFile(ab=_ia(ab), restartAb=rest,
                                                checkpointFileName=str(tmp_path / "ckpt_12.npy"))
    np.testing.assert_array_equal(rest.toNumpy2D(), ab[13:])
    R2 = cuarray.FloatCuArray()
    netcalc.generalizedCorrelation(X=_fa(Xn), R=R2, ab=rest, k=k, n=n, xd=2, d=d,
                                   platform=netcalc.CPU_PLATFORM)
    np.testing.assert_array_equal(np.concatenate([last[:13], R2.toNumpy1D()]), full)
    with pytest.raises(ValueError):
        netcalc.mutualInformationWithCheckpointing(
            X=_fa(Xn), I=cuarray.FloatCuArray(), ab=_ia(ab), k=k, n=n, xd=2, d=d,
            platform=1, checkpointFrequency=-1, checkpointFileName=base)


def _graph(N, density, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.1, 3.0, size=(N, N)).astype(np.float32)
    A[rng.random((N, N)) > density] = np.inf
    np.fill_diagonal(A, 0.0)
    return A


def test_hedetniemi_cpu_platform_equals_oracle():
    N, L = 45, 7
    A = _graph(N, 0.1, 4)
    H, P = cuarray.FloatCuArray(), cuarray.IntCuArray()
    netcalc.hedetniemiAllShortestPaths(A=_fa(A), H=H, paths=P, maxPathLength=L,
                                       tolerance=1e-4, platform=netcalc.CPU_PLATFORM)
    Ho, _, Po = oracle.hedetniemi(A, L, 1e-4)
    np.testing.assert_array_equal(H.toNumpy2D(), Ho)
    np.testing.assert_array_equal(P.toNumpy1D(), Po)
    for i, j in [(0, 3), (7, 44), (9, 9)]:
        np.testing.assert_array_equal(
            netcalc.recoverSingleShortestPath(paths=P, maxPathLength=L, i=i, j=j),
            oracle.recover_path(Po, N, L, i, j))
    H2 = cuarray.FloatCuArray()
    netcalc.hedetniemiAllShortestPathLengths(A=_fa(A), H=H2, maxPathLength=L,
                                             platform=netcalc.CPU_PLATFORM)
    np.testing.assert_array_equal(H2.toNumpy2D(), Ho)


def test_corr_to_adj_cpu():
    C = np.random.default_rng(2).random((9, 9)).astype(np.float32)
    C[2, 3] = 0
    A = cuarray.FloatCuArray()
    netcalc.correlationToAdjacency(A=A, C=_fa(C), n=9, platform=netcalc.CPU_PLATFORM)
    np.testing.assert_array_equal(A.toNumpy2D(), oracle.corr_to_adj(C))


def test_gc_transform_threaded_chunks_equal_oracle():
    """The host GC transform (generalized_correlation_cpu.cpp:8-27) runs in
    chunks of 32,768 values on several threads for long inputs; every value
    must equal the oracle's, at the chunk boundaries too, with MI <= 0
    (including -0.0) mapped to 0."""
    rng = np.random.default_rng(5)
    P = 3 * 32768 + 123
    mi = rng.uniform(-0.2, 2.0, P).astype(np.float32)
    mi[::997] = 0.0
    mi[5] = -0.0
    idx = np.unique(np.concatenate([np.arange(0, P, 331), [0, 5, 32767, 32768, 65535, 65536,
                                                            98303, 98304, P - 1]]))
    for d in (1, 2, 3):
        gc = abi.gc_from_mi(mi, d)
        want = np.array([oracle.gc_from_mi(float(mi[p]), d) for p in idx], np.float32)
        np.testing.assert_array_equal(gc[idx].view(np.uint32), want.view(np.uint32))
        np.testing.assert_array_equal(abi.gc_from_mi(mi[:1000], d), gc[:1000])
