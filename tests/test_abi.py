"""The C-ABI library loads and exports every entry point include/ptmf.h declares
(no compute calls: these run without a GPU)."""
import ctypes as C
import os
import re

from ptatin3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ptmf_\w+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = C.CDLL(_lib.LIBPATH)
    names = declared("ptmf.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared("ptmf.h")) == set(_lib.SIGNATURES)


def test_no_gpu_is_not_an_error_for_queries():
    lib = _lib.load()
    n = C.c_int(-1)
    assert lib.ptmf_device_count(C.byref(n)) == 0
    assert n.value >= 0
    maj, mnr = C.c_int(), C.c_int()
    assert lib.ptmf_version(C.byref(maj), C.byref(mnr)) == 0


def test_bad_arguments_rejected_without_device():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.ptmf_op_create(0, 4, 4, C.byref(h)) == 1  # PTMF_ERR_ARG
    assert b"bad box" in lib.ptmf_last_error()
    assert lib.ptmf_a11_mult(None, None, None) == 1
