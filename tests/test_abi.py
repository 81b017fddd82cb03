"""The C ABI library loads and exports every function the headers declare
(no device calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "libfastlib_amd", "lib", "libfastbit_hip.so")


def declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**(fbh_[a-z0-9_]+)\s*\(",
                            txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_declared_symbols():
    assert os.path.exists(LIB), "libfastbit_hip.so not built"
    lib = C.CDLL(LIB)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_loads():
    import libfastlib_amd
    assert libfastlib_amd.lib() is not None
