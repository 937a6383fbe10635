"""The C-ABI library loads and exports every symbol include/*.h declares (no
compute calls without a GPU), and status codes map to the reference's
exception types."""
import ctypes
import glob
import os
import re

import pytest

from kymatio_amd import _lib
from tests.helpers import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        txt = open(h).read()
        txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
        for m in re.finditer(r'^\s*(?:const\s+)?\w+\s*\*?\s*(kym_\w+)\s*\(', txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_functions():
    assert 'kym_scat2d_forward' in _declared()


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binds_only_declared_symbols():
    declared = set(_declared())
    assert set(_lib.exported_symbols()) <= declared


def test_version_and_error_mapping():
    L = _lib.lib()
    assert L.kym_version() == 1
    assert _lib.STATUS_EXC[1] is TypeError
    assert _lib.STATUS_EXC[2] is RuntimeError
    assert _lib.STATUS_EXC[3] is ValueError
