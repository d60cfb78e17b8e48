"""The C-ABI library loads (no GPU needed) and exports every symbol that
include/pdes.h declares."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "pdes.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pdes_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from pdes_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "build the library first (__graft_entry__.build())"
    L = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, f"missing exports: {missing}"
    assert len(declared_symbols()) >= 20


def test_binding_covers_header():
    from pdes_amd import _lib
    declared = set(declared_symbols())
    bound = set(_lib.exported_symbols())
    assert declared <= bound, f"unbound: {sorted(declared - bound)}"
