"""The C ABI library: builds for gfx950, exports every symbol include/*.h declares, and fails loudly
(an error code, no crash, no CPU fallback) where no gfx950 device is present."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from conftest import REPO

LIB = os.path.join(REPO, "desert-search_amd", "libdsx.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(dsx_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    assert os.path.exists(LIB), "libdsx.so not built"
    names = declared_functions()
    assert {"dsx_create", "dsx_destroy", "dsx_gate", "dsx_strerror"} <= names
    lib = ctypes.CDLL(LIB)
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_library_has_gfx950_code_object():
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", LIB], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout


def test_no_device_is_an_error_not_a_fallback(pkg):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(pkg.dsx.DsxError) as e:
        pkg.dsx.Engine(0)
    assert e.value.code == pkg.dsx.DSX_E_NODEVICE


def test_strerror_names_every_code(pkg):
    lib = pkg.dsx.load()
    for code in range(5):
        assert lib.dsx_strerror(code) and lib.dsx_strerror(code) != b"unknown error"
