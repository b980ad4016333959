"""The C ABI library loads and exports every entry point declared in
include/gta/*.h (no device calls)."""
import ctypes as C
import os
import re

import pytest

from gtalign_alpha_amd import batch

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "gta")
HEADERS = [os.path.join(INCLUDE, f) for f in sorted(os.listdir(INCLUDE)) if f.endswith(".h")]


def declared_functions():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(gta_[a-z_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_batch_interface():
    names = declared_functions()
    for n in ("gta_create", "gta_destroy", "gta_set_queries", "gta_set_references",
              "gta_process", "gta_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(batch.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_defaults_and_section_size_without_device():
    from gtalign_alpha_amd.capi import Options, alignment_section_size
    lib = batch.lib()
    o = Options()
    lib.gta_default_options(C.byref(o))
    assert abs(o.prescore - 0.4) < 1e-7 and abs(o.scorethld - 0.5) < 1e-7
    assert o.convergence == 18 and o.depth == 2 and o.refinement == 2
    assert lib.gta_alignment_section_size(2, 300, 5, 1000) == alignment_section_size(2, 300, 5, 1000)


def test_no_cpu_fallback_when_library_missing(monkeypatch):
    monkeypatch.setattr(batch, "LIB_PATH", "/nonexistent/libgta_amd.so")
    monkeypatch.setattr(batch, "_lib", None)
    with pytest.raises(RuntimeError):
        batch.lib()
