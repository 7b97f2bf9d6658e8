"""Run-to-run determinism (GPU required): every GPU result is computed in a
fixed order — KSG terms summed in frame order, ties resolved by the same
selection, Jacobi min-plus with no in-place race, the screener's gradient
partials summed in wave order — so two runs give the same bits, including
across launch shapes. (The reference's in-place Hedetniemi update and its
warp-synchronous reductions without __syncwarp do not promise this,
hedetniemi_gpu.cu:59-66, mutual_information_gpu.cu:18-59.)"""
import numpy as np
import pytest

from netsci_amd import abi, screen, synth
from oracle import screen_oracle as so

pytestmark = pytest.mark.gpu


def test_ksg_twice_with_ties_and_shapes():
    print("determinism ksg", flush=True)
    n, d, k = 500, 3, 6
    rng = np.random.default_rng(3)
    X = (np.round(rng.standard_normal((8, d * n)) * 4) / 4).astype(np.float32)  # tie-heavy
    ab = synth.all_pairs(8)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    runs = []
    try:
        # All-frames kernel in several launch shapes, then the sorted-list
        # path twice and with another walk shape.
        abi.set_method(abi.METHOD_ALL_FRAMES)
        for shape in [(2, 128, 4, 0, 0), (2, 128, 4, 0, 0), (1, 64, 8, 0, 1), (2, 256, 1, 2, 0)]:
            abi.set_config(*shape)
            dmi = abi.DeviceArray((len(ab),), np.float32)
            abi.ksg_mi(dX, X.shape[1], n, d, n, dab, len(ab), k, dmi)
            runs.append((dmi.numpy(), abi.last_tie_frames()))
        for waves, batch in [(4, 16), (4, 16), (8, 5)]:
            abi.set_method(abi.METHOD_LISTS, waves, batch)
            dmi = abi.DeviceArray((len(ab),), np.float32)
            abi.ksg_mi(dX, X.shape[1], n, d, n, dab, len(ab), k, dmi)
            runs.append((dmi.numpy(), abi.last_tie_frames()))
    finally:
        abi.set_config(0)
        abi.set_method(abi.METHOD_AUTO)
    assert runs[0][1] > 0
    for mi, ties in runs[1:]:
        np.testing.assert_array_equal(mi, runs[0][0])
        assert ties == runs[0][1]


def test_apsp_twice():
    print("determinism apsp", flush=True)
    N, L = 200, 8
    rng = np.random.default_rng(4)
    A = rng.uniform(0.1, 3.0, (N, N)).astype(np.float32)
    A[rng.random((N, N)) > 0.05] = np.inf
    np.fill_diagonal(A, 0.0)
    outs = []
    for _ in range(2):
        dA = abi.DeviceArray.from_numpy(A)
        dH = abi.DeviceArray((N, N), np.float32)
        dHi = abi.DeviceArray((L, N, N), np.float32)
        dP = abi.DeviceArray((N * N * L,), np.int32)
        abi.hedetniemi(dA, N, L, 1e-4, dH, dHi, dP)
        outs.append((dH.numpy(), dP.numpy()))
    np.testing.assert_array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_screener_training_twice():
    """Same data, seed and schedule: identical parameters after training."""
    print("determinism screener", flush=True)
    rng = np.random.default_rng(6)
    P = 6000
    x = rng.standard_normal((P, 16)).astype(np.float32)
    x[:, 15] = 1.0
    y = (np.tanh(x[:, 0]) * 0.5 + 0.5).astype(np.float32)
    feat = (so.bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)
    params = []
    for _ in range(2):
        s = screen.Screener(seed=0)
        s.fit(feat, y, epochs=3, batch=1024, lr=3e-3, seed=0)
        params.append(s.params.copy())
    np.testing.assert_array_equal(params[0].view(np.uint32), params[1].view(np.uint32))
