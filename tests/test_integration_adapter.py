"""The maintainer-side adapter of the integration notes (include/femgpu_lmtpp_adapter.hpp)
compiles against the reference's own mesh headers (the LMT++ sources, present in the build
container only): every member it touches -- node_list, elem_list, pos, dep, epsilon, sigma,
tr_epsilon, apply, make_rect -- exists with the type it assumes.  Compiled, not run."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(REF), reason="the LMT++ sources are not on this machine")
def test_adapter_compiles_against_reference_headers(tmp_path):
    src = os.path.join(ROOT, "tests", "adapter", "adapter_check.cpp")
    out = os.path.join(str(tmp_path), "adapter_check.o")
    cmd = ["g++", "-std=gnu++11", "-fpermissive", "-w", "-O0", "-I" + REF, "-I/usr/include/libxml2",
           "-I" + os.path.join(ROOT, "include"), "-c", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert os.path.getsize(out) > 0
