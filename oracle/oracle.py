"""Loader for the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time against the GPU library.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
# IBMDEM_ORACLE_LIB names another build of the oracle (a copy with a planted
# mistake, to check that the tests catch it); the default is the Makefile's
DEFAULT_LIB = os.path.join(HERE, "_build", "libibmdem_oracle.so")
LIB = os.environ.get("IBMDEM_ORACLE_LIB", DEFAULT_LIB)
sys.path.insert(0, os.path.join(ROOT, "ibm-dem_amd"))

from ibmdem.abi import Simulation, load_library  # noqa: E402

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def library():
    global _lib
    if _lib is None:
        if LIB == DEFAULT_LIB and not os.path.exists(LIB):
            build()
        _lib = load_library(LIB)
    return _lib


def simulation(case) -> Simulation:
    return Simulation(library(), case)
