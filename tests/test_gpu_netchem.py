"""The tutorial pipeline on the GPU (reference tutorial/
tutorial_2_gen_corr_protein.ipynb): protein G's PDB + DCD through netchem, then
generalized correlation of all 56 x 56 ordered node pairs with the trajectory's
100 frames (k = 6, d = 3). GPU values equal the CPU platform's and the
oracle's bit for bit."""
import os

import numpy as np
import pytest

import oracle
from netsci_amd import cuarray, netcalc, netchem

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "netchem")


def test_protein_g_generalized_correlation():
    print("netchem protein G -> GC", flush=True)
    graph = netchem.Graph()
    graph.init(trajectoryFile=os.path.join(GOLD, "protG_short.dcd"),
               topologyFile=os.path.join(GOLD, "protG_short.pdb"), firstFrame=0, lastFrame=99)
    N, n = graph.numNodes(), graph.numFrames()
    assert (N, n) == (56, 100)
    ab = cuarray.IntCuArray()
    ab.fromNumpy2D(np.array([[i, j] for i in range(N) for j in range(N)], np.int32))
    out = {}
    for plat in (netcalc.GPU_PLATFORM, netcalc.CPU_PLATFORM):
        R = cuarray.FloatCuArray()
        assert netcalc.generalizedCorrelation(X=graph.nodeCoordinates(), R=R, ab=ab, k=6, n=n,
                                              d=3, xd=2, platform=plat) == plat
        out[plat] = R.toNumpy1D()
    np.testing.assert_array_equal(out[0], out[1])
    X = graph.nodeCoordinates().toNumpy2D()
    pairs = ab.toNumpy2D()
    want = oracle.pairs(X, pairs, 6, n, 3, gc=True)
    np.testing.assert_array_equal(out[0], want)
    gc = out[0].reshape(N, N)
    assert np.allclose(np.diag(gc), gc.max())  # a node with itself: the largest GC
    np.testing.assert_array_equal(gc, gc.T)
