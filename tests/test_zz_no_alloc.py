"""ib3d_computeDiv_noalloc (no device allocation: its partial reductions use
a static per-GPU device array) against ib3d_computeDiv, both through the C
ABI, on two calls in a row, the second on a smaller grid.  The divergence
maximum must be bit-identical, and so must the stored b/dt (checked through
its norm and through one red-black SOR iteration from zero, which reads b/dt
at every fluid cell).  Before each allocation-free call the right-hand side
holds another field's divergence, so a call that wrote nothing would fail."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _one_sor_iteration(s):
    nx, ny, nz = s.n
    s.upload("pc", np.zeros((nx + 2) * (ny + 2) * (nz + 2)))
    s.REDBLACKSOR(1e300)
    return s.download("pc")


def _check(case):
    from ib3d_amd.native import Solver
    s = Solver.from_case(case)
    try:
        s.step()
        s.nsMomentum()
        s.velocityBC()
        s.solidCellBC()
        s.velocityForcing1()
        s.velocityBC()
        ut_a = s.download("ut")
        ut_b = ut_a * 1.25 + 0.01 * np.sin(np.arange(ut_a.size) * 0.37)
        L = s.L
        out = {}
        for name in ("noalloc", "alloc"):
            # the right-hand side first holds the divergence of another field
            s.upload("ut", ut_a)
            d0 = ctypes.c_double()
            assert L.ib3d_computeDiv(ctypes.c_double(s.deltat), ctypes.byref(d0)) == 0
            s.upload("ut", ut_b)
            d = ctypes.c_double()
            fn = L.ib3d_computeDiv_noalloc if name == "noalloc" else L.ib3d_computeDiv
            rc = fn(ctypes.c_double(s.deltat), ctypes.byref(d))
            assert rc == 0, L.ib3d_last_error().decode()
            assert d.value != d0.value
            rn, bn = s.residual_norm()
            out[name] = (d.value, bn, _one_sor_iteration(s))
        a, b = out["noalloc"], out["alloc"]
        print(f"[gpu] no-alloc computeDiv {s.n}: divmax {a[0]!r} vs {b[0]!r}, ||b/dt|| {a[1]!r} vs {b[1]!r}",
              flush=True)
        assert np.float64(a[0]).view(np.int64) == np.float64(b[0]).view(np.int64)
        assert np.float64(a[1]).view(np.int64) == np.float64(b[1]).view(np.int64)
        assert np.array_equal(a[2].view(np.int64), b[2].view(np.int64))
    finally:
        s.close()


def test_noalloc_computeDiv_matches_computeDiv_on_two_calls(case_dir):
    _check(case_dir("small"))   # 48x32x32
    _check(case_dir("tiny"))    # 32x24x24: the second call is smaller
