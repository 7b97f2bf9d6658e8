"""netchem — node coordinates from a PDB topology and a DCD trajectory
(reference: netchem/python/netchem/__init__.py:1-47, netchem/swig/python_netchem*.i).

``Network`` (alias ``Graph``) builds one node per residue tag with its
mass-weighted coordinates per frame; ``Network.nodeCoordinates()`` is the
FloatCuArray that ``netcalc.generalizedCorrelation`` takes as ``X``.
"""
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import ItemsView, KeysView, Mapping, Union, ValuesView

from .._netsci import Atom, Atoms, Network, Node, atomicWeight, dcdAtomCount, dcdFrameCount  # noqa: F401

Graph = Network


@dataclass
class NetworkSettings(Mapping):
    """Keyword arguments of ``Network.init`` (netchem/__init__.py:17-47)."""
    trajectoryFile: Union[str, Path]
    topologyFile: Union[str, Path]
    firstFrame: int
    lastFrame: int
    stride: int = 1

    def __post_init__(self):
        self.trajectoryFile = str(self.trajectoryFile)
        self.topologyFile = str(self.topologyFile)

    def __iter__(self):
        return iter(asdict(self))

    def __len__(self):
        return len(asdict(self))

    def __getitem__(self, key: str):
        return getattr(self, key)

    def items(self) -> ItemsView:
        return ItemsView(self)

    def keys(self) -> KeysView:
        return KeysView(self)

    def values(self) -> ValuesView:
        return ValuesView(self)


__all__ = ["Atom", "Atoms", "Node", "Network", "Graph", "NetworkSettings", "atomicWeight",
           "dcdAtomCount", "dcdFrameCount"]
