"""The C-ABI libraries load on a machine without a GPU and export every symbol their headers
declare (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

from _gpu import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "ephmc.h")


def declared(header):
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(ephmc_\w+)\s*\(", text, re.M)))


def test_header_lists_entry_points():
    names = declared(HEADER)
    assert "ephmc_create" in names and "ephmc_force" in names and len(names) >= 25


def test_engine_exports_every_declared_symbol():
    lib = C.CDLL(os.path.join(PKG, "libephmc.so"))
    missing = [n for n in declared(HEADER) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_imports():
    import sys
    sys.path.insert(0, PKG)
    import ephmc
    assert ephmc.LIB is not None
    t, kappa, mu, omega = ephmc.scale_params(40.0, (0.10, -0.03, 0.02), (0.3, 0.3, 0.3), 0.0, 0.006, 128)
    assert omega == pytest.approx(0.006 * 40 / 128)
    assert mu == pytest.approx(-2 * 0.15 * 40 / 128)
