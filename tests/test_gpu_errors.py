"""Error paths of the library on the GPU, driven from the host with valid
inputs: a sweep cap below what the tolerance needs ends the solve with an
error (as the reference rule's cap does), and uploads that change the surface
or the grid invalidate what was built on the old ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_sweep_cap_is_an_error_in_every_l2_mode(case_dir):
    from ib3d_amd.native import IB3DError, Solver
    d = case_dir("tiny")
    for mode in (1, 2, 3, 4):
        s = Solver.from_case(d)
        try:
            s.set_poisson_mode(mode, omega=1.9, tol=1e-12, check=2, warm=False, max_sweeps=4)
            if mode >= 3:
                s.set_krylov(2)
            with pytest.raises(IB3DError, match="sweep cap"):
                s.step()
        finally:
            s.close()


def test_surface_and_grid_uploads_invalidate_the_plans(case_dir):
    from ib3d_amd.native import IB3DError, Solver
    d = case_dir("tiny")
    s = Solver.from_case(d)
    try:
        s.step()
        xc = s.download("xcent")
        s.upload("xcent", xc)
        with pytest.raises(IB3DError, match="computeNormDistance"):
            s.velocityForcing1()
        with pytest.raises(IB3DError):
            s.download("pNormDis")
        s.L.ib3d_computeNormDistance()
        s.velocityForcing1()
        x1 = s.download("x1")
        s.upload("x1", x1)
        with pytest.raises(IB3DError, match="computeNormDistance"):
            s.pressureForcing1()
    finally:
        s.close()
