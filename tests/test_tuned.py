"""tuned.json (the autotuner's chosen settings beside libdsx.so): dsx_create applies its "env" object as
defaults of the environment before its first HIP call; a variable already set wins, DSX_TUNED=0 skips the
file, a malformed file is ignored. Runs without a GPU: dsx_create reads the file, then fails for want of
a device. Also checks the committed tuned.json, when present, has the form the library reads."""
import json
import os
import shutil
import subprocess
import sys

import pytest

from conftest import REPO

LIB = os.path.join(REPO, "desert-search_amd", "libdsx.so")

PROBE = r"""
import ctypes, sys
lib = ctypes.CDLL(sys.argv[1])
h = ctypes.c_void_p()
rc = lib.dsx_create(0, ctypes.byref(h))
libc = ctypes.CDLL(None)
libc.getenv.restype = ctypes.c_char_p
for k in sys.argv[2:]:
    v = libc.getenv(k.encode())
    print(k, "-" if v is None else v.decode())
"""


def _probe(tmp_path, text, env_extra, keys):
    d = tmp_path / "lib"
    d.mkdir()
    shutil.copy(LIB, d / "libdsx.so")
    if text is not None:
        (d / "tuned.json").write_text(text)
    env = {k: v for k, v in os.environ.items() if not k.startswith("DSX_")}
    env.update(env_extra)
    r = subprocess.run([sys.executable, "-c", PROBE, str(d / "libdsx.so")] + keys, env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return dict(l.split(" ", 1) for l in r.stdout.splitlines()), r.stderr


pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libdsx.so not built")

DOC = json.dumps({"env": {"DSX_WORKERS": "20", "DSX_FLOOD_NT": 128, "DSX_TUNED": "1", "OTHER": "x"},
                  "trials": [{"env": {"DSX_GROUP": "7"}}]}, indent=1)


def test_tuned_env_applied(tmp_path):
    got, _ = _probe(tmp_path, DOC, {}, ["DSX_WORKERS", "DSX_FLOOD_NT", "DSX_GROUP", "OTHER", "DSX_TUNED"])
    # string and number values; only the "env" object, only DSX_ variables, never DSX_TUNED itself
    assert got == {"DSX_WORKERS": "20", "DSX_FLOOD_NT": "128", "DSX_GROUP": "-", "OTHER": "-", "DSX_TUNED": "-"}


def test_tuned_env_set_variable_wins(tmp_path):
    got, _ = _probe(tmp_path, DOC, {"DSX_WORKERS": "12"}, ["DSX_WORKERS", "DSX_FLOOD_NT"])
    assert got == {"DSX_WORKERS": "12", "DSX_FLOOD_NT": "128"}


def test_tuned_off_and_absent(tmp_path):
    got, _ = _probe(tmp_path, DOC, {"DSX_TUNED": "0"}, ["DSX_WORKERS"])
    assert got == {"DSX_WORKERS": "-"}


def test_tuned_absent(tmp_path):
    got, _ = _probe(tmp_path, None, {}, ["DSX_WORKERS"])
    assert got == {"DSX_WORKERS": "-"}


@pytest.mark.parametrize("text", ['{"env": {"DSX_WORKERS": "20", "DSX_GROUP": }}', '{"env": {"DSX_WORKERS": "2',
                                  '{"nothing": 1}'])
def test_tuned_malformed_ignored(tmp_path, text):
    got, err = _probe(tmp_path, text, {}, ["DSX_WORKERS"])
    assert got == {"DSX_WORKERS": "-"} and "ignored" in err


def test_committed_tuned_json():
    path = os.path.join(REPO, "desert-search_amd", "tuned.json")
    if not os.path.exists(path):
        pytest.skip("no tuned.json committed")
    doc = json.load(open(path))
    assert isinstance(doc["env"], dict)
    assert all(k.startswith("DSX_") and isinstance(v, str) for k, v in doc["env"].items())
    # the reader takes the first "env" key of the file: the top-level one must come first
    assert json.loads(open(path).read(), object_pairs_hook=lambda p: p)[0][0] == "env"
