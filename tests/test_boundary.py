"""CPU-side checks of the drop-in boundary: libmadam.so loads and exports
every symbol include/madam.h declares (no compute calls without a GPU)."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "madam.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", src)) - {"if", "sizeof"})


def test_header_declares_reference_abi():
    names = declared_symbols()
    for n in ("destripe", "destripe_with_cache", "clear_caches"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import libmadam_amd
    lib = ct.CDLL(libmadam_amd.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_wrapper_matches_reference_signatures():
    import inspect
    from libmadam_amd import libmadam_wrapper as w
    assert list(inspect.signature(w.destripe).parameters) == [
        "comm", "pars", "dets", "weights", "timestamps", "pixels", "pixweights", "signal", "periods", "npsd",
        "psdstarts", "psdfreqs", "psdvals"]
    assert list(inspect.signature(w.destripe_with_cache).parameters) == [
        "comm", "timestamps", "pixels", "pixweights", "signal", "outpath"]
    assert w.dict2parstring({"a": 1, "survey": ["x : 0 - 1", "y : 1 - 2"]}) == \
        b"a = 1;survey = x : 0 - 1;survey = y : 1 - 2;"
    assert w.dets2detstring(["a", "b"]) == b"a;b;"
