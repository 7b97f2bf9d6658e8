"""The C-ABI library loads and exports every function include/netsci_hip.h
declares (no GPU needed; no compute calls)."""
import ctypes
import os
import re

from netsci_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(nsx_\w+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(abi.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 25
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert _declared() == set(abi.SIGNATURES)


def test_status_calls_without_gpu():
    assert isinstance(abi.device_count(), int)
    assert abi.lib().nsx_last_error() is not None
