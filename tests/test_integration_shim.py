"""The reference-side drop-in (desert-search_amd/shim/gproxy_dsx.cpp, the gproxy.cu replacement INTEGRATION.md
describes) compiles against the reference's own src/gpu/gproxy.h and links with libdsx.so, alone and together
with the reference's dsearch_desert.cpp built with -DDRIP_GPU (Makefile:39-42): every gproxy function the
reference's GPU build calls resolves to the shim. Needs the reference tree; skipped where it is absent."""
import os
import subprocess
import tempfile

import pytest

from conftest import REPO

REF_SRC = "/root/reference/src"
SHIM = os.path.join(REPO, "desert-search_amd", "shim", "gproxy_dsx.cpp")
LIBDIR = os.path.join(REPO, "desert-search_amd")


def _need():
    if not os.path.exists(os.path.join(REF_SRC, "gpu", "gproxy.h")):
        pytest.skip("reference tree not present")
    if not os.path.exists(os.path.join(LIBDIR, "libdsx.so")):
        pytest.skip("libdsx.so not built")


def _gxx(args, td):
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-pthread", "-I" + REF_SRC, "-I" + os.path.join(REPO, "include")]
                       + args + ["-L" + LIBDIR, "-ldsx", "-Wl,--no-undefined"], capture_output=True, text=True,
                       cwd=td, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]


def test_shim_compiles_against_gproxy_h():
    _need()
    with tempfile.TemporaryDirectory() as td:
        _gxx(["-shared", "-fPIC", SHIM, "-o", os.path.join(td, "libgproxy_dsx.so")], td)


def test_reference_gpu_build_links_with_shim():
    _need()
    with tempfile.TemporaryDirectory() as td:
        _gxx(["-DDRIP_GPU", "-ffp-contract=off", os.path.join(REF_SRC, "dsearch_desert.cpp"), SHIM,
              "-o", os.path.join(td, "dsearch_desert_gpu")], td)
        assert os.path.getsize(os.path.join(td, "dsearch_desert_gpu")) > 0
