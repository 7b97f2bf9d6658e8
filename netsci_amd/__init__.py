"""netsci_amd — an MI355X (gfx950) engine for NetSci's KSG mutual-information /
generalized-correlation pair loop and Hedetniemi all-pairs shortest paths.

The user-facing modules mirror the reference's SWIG modules:
``netsci_amd.cuarray`` (FloatCuArray, IntCuArray), ``netsci_amd.netcalc``
(mutualInformation, generalizedCorrelation, hedetniemiAllShortestPaths, ...)
and ``netsci_amd.netchem`` (Network: node coordinates from PDB + DCD).
Both sit on ``_netsci`` (C++ CuArray / netcalc host layer, pybind11), which
calls ``libnsx.so`` (HIP kernels behind the C ABI in include/netsci_hip.h).
There is no pure-Python fallback: importing without the built extension
raises ImportError naming the build command.
"""
import os

_HERE = os.path.dirname(os.path.abspath(__file__))


def _load():
    try:
        from . import _netsci  # noqa: F401
    except ImportError as e:  # pragma: no cover - exercised only when unbuilt
        raise ImportError(
            "netsci_amd native extension is not built; run "
            "`python -m netsci_amd.build` (or __graft_entry__.build())") from e
    return _netsci


_netsci = _load()
