"""The C-ABI libraries load on a host without a GPU and export every symbol
their headers declare (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    src = open(header).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(mps2_\w+|gqmps2_\w+)\s*\(", src, flags=re.M)))


def test_engine_exports_every_declared_symbol():
    from mps2_amd import engine
    lib = ctypes.CDLL(engine.LIB_PATH)
    names = declared(os.path.join(ROOT, "include", "mps2_amd.h"))
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole ABI
    assert set(names) <= set(engine.SIGNATURES)


def test_device_count_without_gpu_is_not_an_error():
    from mps2_amd import engine
    n = engine.device_count()
    assert n >= 0


def test_errors_are_reported_not_fatal():
    from mps2_amd import engine
    lib = engine.lib()
    rc = lib.mps2_ten_free(None)
    assert rc == 0
    rc = lib.mps2_contract(None, None, 0, None, None, None)
    assert rc == engine.lib().mps2_version() * 0 + 1  # MPS2_ERR_ARG
    assert b"NULL" in lib.mps2_last_error()


def test_host_layer_exports_every_declared_symbol():
    from mps2_amd import gqmps2
    lib = ctypes.CDLL(gqmps2.LIB_PATH)
    names = declared(os.path.join(ROOT, "include", "gqmps2_capi.h"))
    assert "gqmps2_measure_multi_site" in names and len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(gqmps2.SIGNATURES)
