"""The C ABI library and the CLI tools load and export what include/gstore_amd.h
declares (no device calls).  CPU only."""
import os
import re
import subprocess

import pytest

import gstore_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gstore_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gs_[a-z_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = _declared()
    for n in ("gs_open", "gs_open_arrays", "gs_open_kron", "gs_pagerank", "gs_bfs", "gs_wcc", "gs_kcore",
              "gs_close", "gs_last_error", "gs_convert", "gs_save"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = gstore_amd.load_library()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_without_device():
    import ctypes
    lib = gstore_amd.load_library()
    buf = ctypes.create_string_buffer(64)
    assert lib.gs_last_error(buf, 64) >= 0
    assert lib.gs_kernel_name(0) == b"pr_scatter"


@pytest.mark.parametrize("tool", ["gstoreu", "gstored"])
def test_cli_help(tool):
    exe = os.path.join(ROOT, "gstore_amd", "bin", tool)
    assert os.path.exists(exe), f"{exe} not built"
    r = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "-j" in r.stderr
