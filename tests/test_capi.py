"""The C-ABI library loads without a GPU and exports every symbol that
include/*.h declares."""
import ctypes
import glob
import os
import re

from nebulasem_amd._lib import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(nsem_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_declared_symbols():
    assert os.path.exists(LIB_PATH)
    L = lib()
    names = declared_symbols()
    assert len(names) > 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_error_reporting():
    L = lib()
    h = ctypes.c_void_p()
    rc = L.nsem_mesh_load(b"/nonexistent/grid_0.txt", (ctypes.c_int32 * 3)(1, 1, 1), 3, ctypes.byref(h))
    assert rc == 3
    assert b"cannot open" in L.nsem_last_error()
