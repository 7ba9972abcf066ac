"""The C-ABI library loads without a GPU and exports every entry point that
include/llg/*.h declares (sundials_compat.h holds only inline dispatchers
and types).  No compute call is made here."""
import ctypes as C
import re
from pathlib import Path

import pytest

import llg_amd as la

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted(p for p in (ROOT / "include" / "llg").glob("*.h") if p.name != "sundials_compat.h")
DECL = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?[\s\*])([A-Za-z_]\w*)\s*\(", re.M)


def declared(path):
    text = path.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//.*", "", text)
    names = []
    for m in DECL.finditer(text):
        line = text[text.rfind("\n", 0, m.start()) + 1: m.end()]
        if "typedef" in line or "#define" in line or line.strip().startswith("static"):
            continue
        names.append(m.group(1))
    return names


def test_headers_declare_functions():
    total = sum(len(declared(h)) for h in HEADERS)
    assert total > 100


@pytest.mark.parametrize("header", HEADERS, ids=[h.name for h in HEADERS])
def test_library_exports_header_symbols(header):
    lib = C.CDLL(str(la.native.LIB_PATH))
    missing = [n for n in declared(header) if not hasattr(lib, n)]
    assert not missing, f"{header.name}: not exported: {missing}"


def test_python_binding_covers_exports():
    names = set()
    for h in HEADERS:
        names.update(declared(h))
    assert names <= set(la.native.SIGNATURES), sorted(names - set(la.native.SIGNATURES))
    la.native.load()     # prototypes attach without a device


def test_load_fails_loudly_when_missing(tmp_path):
    with pytest.raises(RuntimeError):
        la.native.load(tmp_path / "nope.so")
