"""CPU-side checks of the drop-in boundary: the C-ABI library loads and exports
every entry point that include/tcym.h, include/tcym_batch.h and
include/tcym_io.h declare (no compute calls here)."""
import os
import re

import topologicalcymsolver_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="tcym.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tcym_[a-z0-9_]+)\s*\(", src)))


def all_declared():
    out = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            out |= set(declared_symbols(h))
    return sorted(out)


def test_header_declares_what_the_wrapper_binds():
    assert declared_symbols("tcym.h") == sorted(T.EXPORTS)
    assert declared_symbols("tcym_batch.h") == sorted(T.BATCH_EXPORTS)
    assert declared_symbols("tcym_io.h") == sorted(T.IO_EXPORTS)


def test_library_exports_every_declared_symbol():
    L = T.load_library()
    missing = [s for s in all_declared() if not hasattr(L, s)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in T.load_library().tcym_version()


def test_library_is_gfx950_only():
    # the code object bundled in the library targets gfx950 and nothing else
    data = open(T.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-f]+)", data))
    assert targets == {b"gfx950"}, targets
