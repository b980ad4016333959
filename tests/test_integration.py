"""The reference-side adapter of INTEGRATION.md compiles against the
reference's own headers (TdFinalizer, PMBatchStrData, CLOptions, gproc) and
this library's header: the ```cpp block is extracted and checked with
`g++ -std=c++14 -fsyntax-only`, as the reference's CPU build compiles
(src/CMakeLists.txt:1-12, without GPUINUSE).  Skipped where the reference
tree is absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFSRC = "/root/reference/src"


def adapter_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    assert blocks, "no C++ block in INTEGRATION.md"
    return blocks[0]


def test_adapter_compiles_against_reference_headers(tmp_path):
    if not os.path.isdir(os.path.join(REFSRC, "libgenp", "goutp")):
        pytest.skip("reference sources not present")
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "adapter.cpp"
    src.write_text(adapter_source() + "\nint main() { return sizeof(AmdBatch) > 0 ? 0 : 1; }\n")
    r = subprocess.run([gxx, "-std=c++14", "-fsyntax-only", "-I" + REFSRC, "-I" + os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_adapter_uses_every_stage_entry_point():
    s = adapter_source()
    for name in ("gta_create", "gta_destroy", "gta_set_queries", "gta_set_references", "gta_process",
                 "gta_last_error", "gta_default_options"):
        assert name + "(" in s, name
