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


def test_argument_checks_run_before_any_device_call():
    """Invalid arguments are rejected with a message and status 1 before the
    library touches a device, so they behave the same with or without a GPU.
    The pointers are placeholders that are never dereferenced."""
    L = abi.lib()
    fake = ctypes.c_void_p(0x1000)
    # paths need the Hi snapshots
    assert L.nsx_hedetniemi(fake, 10, 3, 1e-4, fake, None, fake, None) == 1
    assert b"Hi snapshot" in L.nsx_last_error()
    # list-group indices of the path backtrack are 32-bit
    assert L.nsx_hedetniemi(fake, 92681, 2, 1e-4, fake, fake, fake, None) == 1
    assert b"92680" in L.nsx_last_error()
    assert L.nsx_hedetniemi(fake, 0, 3, 1e-4, fake, fake, fake, None) == 1
