"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import re

from conftest import ROOT

from picoquant_hip import _lib


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(pq_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert _declared() == set(_lib.SIGNATURES)


def test_load_binds_signatures():
    lib = _lib.load()
    for name in _lib.SIGNATURES:
        assert getattr(lib, name).argtypes is not None


def test_python_constants_match_header():
    txt = (ROOT / "include" / "pq_hip.h").read_text()
    defs = {k: int(v) for k, v in re.findall(r"#define\s+(PQ_[A-Z0-9_]+)\s+(\d+)", txt)}
    kc = {"PQ_KC_" + n.upper(): v for v, n in _lib.KC_NAMES.items()}
    checked = 0
    for name, val in defs.items():
        if hasattr(_lib, name):
            assert getattr(_lib, name) == val, name
            checked += 1
        elif name in kc:
            assert kc[name] == val, name
            checked += 1
    assert checked >= 10
