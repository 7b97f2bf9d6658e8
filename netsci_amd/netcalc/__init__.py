"""netcalc — KSG mutual information, generalized correlation and Hedetniemi
shortest paths with the reference's Python surface
(netcalc/python/netcalc/__init__.py:7-128, netcalc/swig/python_netcalc_rename.i).

``platform=GPU_PLATFORM`` runs the gfx950 kernels, ``CPU_PLATFORM`` the host
implementation; both give the reference CPU path's values. Keyword names are
the reference's (callers pass them by name, e.g.
docs/tutorial2_protein_structure_gc.rst:89-98).
"""
from dataclasses import dataclass, fields
from pathlib import Path
from typing import ItemsView, KeysView, Mapping, Union, ValuesView

from ..cuarray import FloatCuArray, IntCuArray
from .._netsci import (  # noqa: F401
    CPU_PLATFORM,
    GPU_PLATFORM,
    correlationToAdjacency,
    correlationToAdjacencyGpu,
    generalizedCorrelation,
    generalizedCorrelationCpu,
    generalizedCorrelationGpu,
    generalizedCorrelationWithCheckpointing,
    generateRestartAbFromCheckpointFile,
    hedetniemiAllShortestPathLengths,
    hedetniemiAllShortestPathLengthsGpu,
    hedetniemiAllShortestPaths,
    hedetniemiAllShortestPathsCpu,
    hedetniemiAllShortestPathsGpu,
    mean,
    meanGpu,
    mutualInformation,
    mutualInformationCpu,
    mutualInformationGpu,
    mutualInformationWithCheckpointing,
    recoverSingleShortestPath,
    standardDeviation,
    standardDeviationGpu,
    windowedCorrelation,
)


class _SettingsMapping(Mapping):
    """Mapping behaviour shared by the Settings dataclasses
    (netcalc/__init__.py:47-68): splatting yields the call's keyword
    arguments, without the checkpoint pair when checkpointFrequency == -1."""

    def __iter__(self):
        names = [f.name for f in fields(self)]
        if self.checkpointFrequency == -1:
            for key in names:
                if key not in ("checkpointFrequency", "checkpointFileName"):
                    yield key
        else:
            yield from names

    def items(self) -> ItemsView:
        return ItemsView({k: self[k] for k in self})

    def keys(self) -> KeysView:
        return KeysView(self)

    def values(self) -> ValuesView:
        return ValuesView({k: self[k] for k in self})

    def __len__(self):
        return sum(1 for _ in self)

    def __getitem__(self, key: str):
        return getattr(self, key)


@dataclass
class MutualInformationSettings(_SettingsMapping):
    """netcalc/__init__.py:30-71."""
    X: FloatCuArray
    I: FloatCuArray  # noqa: E741
    ab: IntCuArray
    k: int
    n: int
    xd: int
    d: int
    platform: int = GPU_PLATFORM
    checkpointFrequency: int = -1
    checkpointFileName: Union[str, Path] = None

    def __post_init__(self):
        if self.checkpointFileName is not None:
            self.checkpointFileName = str(self.checkpointFileName)

    def __repr__(self):
        return (f"MutualInformationSettings(k={self.k}, n={self.n}, xd={self.xd}, d={self.d}, "
                f"platform={self.platform}, checkpointFrequency={self.checkpointFrequency}, "
                f"checkpointFileName={self.checkpointFileName})")


@dataclass
class GeneralizedCorrelationSettings(_SettingsMapping):
    """netcalc/__init__.py:74-128, including the empty-ab auto-fill of all
    N^2 ordered pairs (i, j), self-pairs included (:90-102). The reference
    takes N = X.n() / d, which is the frame count of a node-row X, and so
    writes node indices past the last row; here N = X.m(), the node count."""
    X: FloatCuArray
    R: FloatCuArray
    ab: IntCuArray
    k: int
    n: int
    xd: int
    d: int
    platform: int = GPU_PLATFORM
    checkpointFrequency: int = -1
    checkpointFileName: Union[str, Path] = None

    def __post_init__(self):
        if self.checkpointFileName is not None:
            self.checkpointFileName = str(self.checkpointFileName)
        if not self.ab.size():
            print("Initializing ab")
            import numpy as np
            num = int(self.X.m())
            ii, jj = np.meshgrid(np.arange(num, dtype=np.int32),
                                 np.arange(num, dtype=np.int32), indexing="ij")
            self.ab.fromNumpy2D(np.stack([ii.ravel(), jj.ravel()], axis=1))

    def __repr__(self):
        return (f"GeneralizedCorrelationSettings(k={self.k}, n={self.n}, xd={self.xd}, "
                f"d={self.d}, platform={self.platform}, "
                f"checkpointFrequency={self.checkpointFrequency}, "
                f"checkpointFileName={self.checkpointFileName})")
