"""The C-ABI boundary: the shared library loads without a GPU and exports
every entry point include/neospice_amd.h declares (no compute call here);
host-only entry points behave as documented (status codes, errors)."""
import ctypes as C
import os
import re

import numpy as np

import neospice_amd as na
from neospice_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "neospice_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(nsa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) > 20
    L = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing


def test_error_reporting_without_gpu():
    """Argument errors come back as negative status with a message."""
    L = _lib.lib()
    cp = np.array([0, 1], np.int32)
    ri = np.array([0], np.int32)
    perm = np.zeros(1, np.int32)
    assert L.nsa_min_degree_order(1, _lib.i32p(cp), _lib.i32p(ri), _lib.i32p(perm)) == 0
    rc = L.nsa_min_degree_order(-1, _lib.i32p(cp), _lib.i32p(ri), _lib.i32p(perm))
    assert rc < 0 and L.nsa_last_error()


def test_default_options_match_reference_defaults():
    """SimOptions defaults (options.hpp) the Newton loop uses."""
    o = na.default_options()
    assert o.reltol == 1e-3 and o.abstol == 1e-12 and o.vntol == 1e-6
    assert o.gmin == 1e-12 and o.itl1 == 100 and o.itl4 == 50
