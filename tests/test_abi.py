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


def _err(lib):
    lib.ib3d_last_error.restype = ctypes.c_char_p
    return lib.ib3d_last_error().decode()


def test_solver_option_entry_points_validate_their_arguments():
    """The Poisson-mode options are checked before anything touches a GPU,
    so their error paths run here: a bad value returns nonzero with a
    message and leaves the last good setting in place."""
    from ib3d_amd import native
    lib = native.lib()
    D, I = ctypes.c_double, ctypes.c_int
    lib.ib3d_set_poisson_mode.argtypes = [I, D, D, I, I, ctypes.c_int64]
    lib.ib3d_set_krylov.argtypes = [I]
    lib.ib3d_set_multigrid.argtypes = [I, I, D, I, D, I]
    lib.ib3d_cnn_set_accept.argtypes = [D]
    assert lib.ib3d_set_poisson_mode(4, 1.9, 1e-6, 20, 1, 0) == 0
    assert lib.ib3d_set_poisson_mode(5, 1.9, 1e-6, 20, 1, 0) != 0
    assert "mode must be" in _err(lib)
    assert lib.ib3d_set_poisson_mode(3, 2.0, 1e-6, 20, 1, 0) != 0
    assert "omega" in _err(lib)
    assert lib.ib3d_set_krylov(6) == 0
    for bad in (0, 24):
        assert lib.ib3d_set_krylov(bad) != 0
        assert "restart" in _err(lib)
    assert lib.ib3d_set_multigrid(2, 2, 1.55, 8, 1.5, 0) == 0
    cases = [((0, 2, 1.55, 8, 1.5, 0), "pre-smoothing"), ((2, -1, 1.55, 8, 1.5, 0), "post-smoothing"),
             ((2, 2, 2.0, 8, 1.5, 0), "relaxation"), ((2, 2, 1.55, 0, 1.5, 0), "coarse sweeps"),
             ((2, 2, 1.55, 8, 1.5, 1), "max_levels")]
    for args, word in cases:
        assert lib.ib3d_set_multigrid(*args) != 0
        assert word in _err(lib), (args, _err(lib))
    lib.ib3d_set_multigrid_level.argtypes = [I, I, I]
    assert lib.ib3d_set_multigrid_level(1, 3, 2) == 0
    assert lib.ib3d_set_multigrid_level(1, -1, -1) == 0
    for args, word in [((0, 3, 2), "level"), ((8, 3, 2), "level"), ((1, 0, 2), "sweeps"), ((1, 3, -1), "sweeps"),
                       ((1, 65, 2), "sweeps")]:
        assert lib.ib3d_set_multigrid_level(*args) != 0
        assert word in _err(lib), (args, _err(lib))
    lib.ib3d_set_multigrid_small.argtypes = [I, I]
    assert lib.ib3d_set_multigrid_small(5, 4) == 0
    assert lib.ib3d_set_multigrid_small(-1, -1) == 0
    for args in ((0, 4), (5, -1), (65, 4), (-1, 4)):
        assert lib.ib3d_set_multigrid_small(*args) != 0
        assert "sweeps" in _err(lib), (args, _err(lib))
    assert lib.ib3d_cnn_set_accept(0.5) == 0
    for bad in (0.0, 1.5, -1.0):
        assert lib.ib3d_cnn_set_accept(bad) != 0
        assert "ratio" in _err(lib)
    assert lib.ib3d_cnn_set_accept(1.0) == 0
