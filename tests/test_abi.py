"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import glob
import os
import re
import ctypes

from ppf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in [os.path.join(ROOT, "include", "ppf_mi355x.h")]:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ppf_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert "ppf_pcg_solve" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    names = declared_functions()
    assert names <= set(_lib._PROTOS), names - set(_lib._PROTOS)
    _lib.load()  # prototypes bind


def test_error_string_available_without_gpu():
    lib = _lib.load()
    assert isinstance(lib.ppf_last_error(), bytes)
