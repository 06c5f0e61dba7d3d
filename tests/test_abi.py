"""The C-ABI library loads without a GPU and exports every entry point its header declares;
the ctypes binding declares a signature for each of them. No compute call is made."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hashdag_hip.h")

DECL = re.compile(r"^(?:const\s+)?[A-Za-z_][\w]*\s*\*?\s*(hd_[a-z0-9_]+)\s*\(", re.M)


def declared():
    names = DECL.findall(open(HEADER).read())
    assert len(names) >= 30
    return names


def test_library_exports_every_declared_symbol():
    from hashdag_amd._lib import LIB_PATH
    lib = C.CDLL(LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_has_a_signature_for_every_symbol():
    from hashdag_amd import _lib
    missing = [n for n in declared() if n not in _lib._SIGS]
    assert not missing, f"no ctypes signature: {missing}"


def test_header_entry_points_cite_the_reference():
    text = open(HEADER).read()
    # every declaration is preceded by a comment naming the reference file:line it replaces
    # or stating that it has no counterpart
    blocks = re.split(r"\n\s*\n", text)
    for b in blocks:
        for n in DECL.findall(b):
            assert re.search(r"[a-z_]+\.(h|cpp|cu):\d+|no counterpart|reference", b), n
