"""CPU oracle — test infrastructure only.

A restatement of the reference's KSG / GC / Hedetniemi semantics, used as the
checker by tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg. Product code under ``netsci_amd/`` never imports it.
"""
from .oracle import *  # noqa: F401,F403
