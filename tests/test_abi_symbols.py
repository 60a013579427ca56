"""The C-ABI library loads without a GPU and exports every entry point that
include/anqi.h declares."""
import ctypes
import os
import re

import pytest

from anqi_testlib import REPO

LIB = os.path.join(REPO, "anqi-implementation_amd", "libanqi_hip.so")


def declared():
    src = open(os.path.join(REPO, "include", "anqi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(anqi_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import sys
    sys.path.insert(0, os.path.join(REPO, "anqi-implementation_amd"))
    import pyanqi
    missing = [n for n in declared() if n not in pyanqi.SIGS]
    assert not missing, missing
