"""The C-ABI library: it loads, exports every entry point include/ib3d.h
declares, and fails loudly (no CPU fallback) where no GPU is visible."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ib3d.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ib3d_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from ib3d_amd import native
    lib = native.lib()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_is_plain_c():
    txt = open(HEADER).read()
    assert "extern \"C\"" in txt
    for bad in ["torch", "std::", "hipStream_t", "template"]:
        assert bad not in txt


def test_no_gpu_means_an_error_not_a_fallback():
    from ib3d_amd import native
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from ib3d_amd import native\n"
            "try:\n"
            "    native.Solver(8, 8, 8)\n"
            "    print('INIT_OK')\n"
            "except native.IB3DError as e:\n"
            "    print('ERR', e)\n") % os.path.join(ROOT, "ib3d-gpu_amd")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert "ERR" in out.stdout, out.stdout + out.stderr
