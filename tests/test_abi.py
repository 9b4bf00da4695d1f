"""The C ABI library loads and exports every entry point include/jagua_amd.h declares; the
Python mirror binds all of them.  No compute calls (no GPU needed)."""
import ctypes
import os
import re

import jagua_amd as J
from jagua_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "jagua_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(jg_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(N.lib_path())
    names = declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n


def test_bindings_cover_the_header():
    assert set(declared()) == set(N.EXPORTED)


def test_version():
    assert "gfx950" in J.version()
