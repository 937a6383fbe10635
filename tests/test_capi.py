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


def _prototypes():
    protos = {}
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        txt = re.sub(r'/\*.*?\*/', '', open(h).read(), flags=re.S)
        for m in re.finditer(r'(kym_\w+)\s*\(([^)]*)\)\s*;', txt):
            params = [p.strip() for p in m.group(2).split(',') if p.strip() and p.strip() != 'void']
            protos[m.group(1)] = params
    return protos


def test_binding_arity_matches_header():
    """Every ctypes binding has exactly as many arguments as the C prototype."""
    protos = _prototypes()
    for name, argtypes in _lib.signatures().items():
        assert name in protos, name
        assert len(argtypes) == len(protos[name]), (name, len(argtypes), protos[name])


def test_binding_kinds_match_header():
    """Pointer parameters bind as pointers, 64-bit integers as c_longlong."""
    protos = _prototypes()
    for name, argtypes in _lib.signatures().items():
        for decl, at in zip(protos[name], argtypes):
            if '*' in decl:
                assert at is ctypes.c_void_p or issubclass(at, ctypes._Pointer), (name, decl, at)
            elif 'long long' in decl:
                assert at in (ctypes.c_longlong, ctypes.c_ulonglong), (name, decl, at)
            elif decl.startswith('float'):
                assert at is ctypes.c_float, (name, decl, at)
            elif decl.startswith('int'):
                assert at is ctypes.c_int, (name, decl, at)
