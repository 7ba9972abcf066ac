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


OPS_CHECK = r"""
#include <stddef.h>
#include <stdio.h>
#include "llg/sundials_compat.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(struct _generic_N_Vector_Ops),
         offsetof(struct _generic_N_Vector_Ops, nvdotprodmulti),
         offsetof(struct _generic_N_Vector_Ops, nvdotprodmultiallreduce),
         offsetof(struct _generic_N_Vector_Ops, nvbufsize),
         offsetof(struct _generic_N_Vector_Ops, nvbufunpack),
         offsetof(struct _generic_N_Vector_Ops, nvprint),
         offsetof(struct _generic_N_Vector_Ops, nvprintfile));
  return 0;
}
"""


@pytest.mark.parametrize("compiler,lang", [("gcc", "c"), ("g++", "c++")])
def test_ops_table_has_sundials7_layout(tmp_path, compiler, lang):
    """The N_Vector ops table has SUNDIALS 7's 56 slots in order: an
    unmodified CVODE copies it with sizeof and calls N_VPrint / N_VBufSize
    through these offsets.  The header's static asserts must hold in C and
    C++, and the offsets are checked here against the public field order
    (11 utility, 19 standard, 3 fused, 7 array, 9 local-reduction, 2
    single-buffer, 3 XBraid buffer and 2 print ops)."""
    import subprocess
    src = tmp_path / ("ops." + ("c" if lang == "c" else "cpp"))
    src.write_text(OPS_CHECK)
    exe = tmp_path / "ops"
    r = subprocess.run([compiler, "-x", lang, "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    p = C.sizeof(C.c_void_p)
    assert vals == [56 * p, 32 * p, 50 * p, 51 * p, 53 * p, 54 * p, 55 * p]
