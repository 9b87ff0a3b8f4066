"""Poisson-solver presets shared by the benchmark and the GPU tests.

Every L2 preset stops at ||b/dt - A pc||_2 <= tol ||b/dt||_2, checked in FP64.

  reference  -- the reference's rule (subPcorVcor.f90:148-212): FP64
                red-black SOR, omega 1.99, stop when max|pc - pco| < eps1,
                zero guess
  fp64-cold  -- FP64 red-black SOR at `omega` from a zero guess
  fp64       -- the same from the previous step's correction
  mixed-dc   -- FP32 red-black SOR sweeps on the defect equation inside an
                FP64 defect-correction loop, previous-step guess
  gmres-sor  -- FP64 refinement, FP32 flexible GMRES(12), every Arnoldi step
                preconditioned by 20 FP32 red-black SOR sweeps at omega 1.9
  mixed      -- FP64 refinement, FP32 flexible GMRES(6), every Arnoldi step
                preconditioned by one FP32 multigrid V-cycle (4+4 red-black
                SOR sweeps per level at omega 1.55, 5+4 on the levels cycled
                inside one workgroup, 4 on the coarsest)
"""

# FP32 GMRES with SOR preconditioning: Arnoldi steps per refinement step,
# FP32 SOR sweeps per preconditioner application and their omega
GMRES_SOR = {"m": 12, "k": 20, "omega": 1.9, "target": 1.0}
# the multigrid preconditioner: Arnoldi steps per refinement step, pre- and
# post-smoothing sweeps per level and their omega, sweeps and omega on the
# coarsest level
MULTIGRID = {"m": 6, "nu1": 4, "nu2": 4, "omega": 1.55, "coarse": 4, "coarse_omega": 1.5, "target": 1.0}
# "target": a Krylov cycle hands back to the FP64 residual check once its own
# residual estimate is below target x tol x ||b/dt|| (the check still decides
# when the solve ends).  Poisson ms per step, 0.3 -> 1.0, one box: 4.264 ->
# 4.098 at 256x192x192 (steps 6-25), 4.605 -> 4.385 over its 500-step run,
# 10.34 -> 9.70 at 384x256x256, 29.78 -> 26.16 at 512x320x320.  The SOR-
# preconditioned comparison mode gains too (55.995 -> 53.647 ms at
# 256x192x192, steps 6-25), so it runs at its best exit as well.
# sweeps before and after on the second level (one eighth of the finest),
# where each sweep is still a launch of real work; None: the coarse levels'
# counts.  At 256x192x192, (3, 2) gives the whole 500-step run 1.4% more
# steps/s but the steps right after the impulsive start 1.4% fewer, and
# (3, 3) costs 1% at 512x320x320, so the default keeps 4+4.
LEVEL1 = None
# sweeps before and after on the coarsest levels, those cycled inside one
# workgroup (16x12x12 and below at 256x192x192), where a sweep costs a
# barrier rather than a launch.  Poisson ms per step against 4+4, one box:
# 4.296 / 4.346 at 256x192x192 (steps 21-100), 4.620 / 4.669 over its
# 500-step run, 10.381 / 10.406 at 384x256x256, 29.85 / 29.99 at
# 512x320x320; 6+6 was better on the last two but not at 256x192x192.
SMALL = (5, 4)
# how the smoothing is launched (identical results either way): whole sweeps
# in one launch on the first `fused` levels, two temporally blocked launches
# per smoothing phase from level `tiles` on
LAUNCH = {"fused": 2, "tiles": 2}

MODES = ("reference", "fp64-cold", "fp64", "mixed-dc", "gmres-sor", "mixed")


def apply(solver, mode, omega=1.99, tol=1e-6, inner=20, gmres=None, multigrid=None):
    """Configure `solver` (an ib3d_amd.native.Solver) for one preset."""
    g = dict(GMRES_SOR, **(gmres or {}))
    mg = dict(MULTIGRID, **(multigrid or {}))
    if mode == "reference":
        solver.set_poisson_mode(0)
    elif mode == "fp64-cold":
        solver.set_poisson_mode(1, omega=omega, tol=tol, check=inner, warm=False)
    elif mode == "fp64":
        solver.set_poisson_mode(1, omega=omega, tol=tol, check=inner, warm=True)
    elif mode == "mixed-dc":
        solver.set_poisson_mode(2, omega=omega, tol=tol, check=inner, warm=True)
    elif mode == "gmres-sor":
        solver.set_poisson_mode(3, omega=g["omega"], tol=tol, check=g["k"], warm=True)
        solver.set_krylov(g["m"])
        solver.set_krylov_target(g["target"])
    elif mode == "mixed":
        solver.set_poisson_mode(4, tol=tol, warm=True)
        solver.set_krylov(mg["m"])
        solver.set_krylov_target(mg["target"])
        solver.set_multigrid(mg["nu1"], mg["nu2"], mg["omega"], mg["coarse"], mg["coarse_omega"], 0)
        if LEVEL1 is not None:
            solver.set_multigrid_level(1, *LEVEL1)
        if SMALL is not None:
            solver.set_multigrid_small(*SMALL)
        solver.set_fused_sweeps(LAUNCH["fused"])
        solver.set_coarse_tiles(LAUNCH["tiles"])
    else:
        raise ValueError(f"unknown Poisson preset {mode!r}; known: {', '.join(MODES)}")
