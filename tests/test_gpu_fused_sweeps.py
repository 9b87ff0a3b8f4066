"""Multigrid levels smoothed by whole red-black sweeps in one launch
(k_sweep2_pz: red and black of a sweep in one pass, written to a second
buffer) against the same solver with two in-place half-sweep launches per
sweep: the same FP32 operations in the same order, so every field, every
Krylov count and every residual must be bit-identical.  Covers the first
sweep from zero, an odd sweep count (its first sweep in place), chunks of one
and several planes, and the finest level alone or with the coarse levels."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(case, fused, nsteps, nu=(4, 4), planes=0):
    """fused: the number of levels smoothed by whole sweeps (0: none)"""
    from ib3d_amd import presets
    from ib3d_amd.native import Solver
    s = Solver.from_case(case)
    try:
        presets.apply(s, "mixed")
        mg = presets.MULTIGRID
        s.set_multigrid(nu[0], nu[1], mg["omega"], mg["coarse"], mg["coarse_omega"], 0)
        s.set_coarse_tiles(0)  # the fused sweep on every level it is asked for
        s.set_fused_sweeps(fused, planes=planes)
        stats = []
        for _ in range(nsteps):
            s.step()
            st = s.poisson_stats()
            stats.append((st["sweeps"], st["outer"], st["relres"]))
        return stats, {f: s.download(f) for f in ("u", "v", "w", "p", "pc")}
    finally:
        s.set_fused_sweeps()
        s.set_coarse_tiles()
        s.close()


@pytest.mark.parametrize("name,nsteps,nu,planes,levels", [("small", 6, (4, 4), 0, 1), ("small", 4, (3, 2), 0, 3),
                                                          ("cfg1", 4, (4, 4), 0, 1), ("cfg1", 3, (4, 4), 0, 4),
                                                          ("cfg2", 3, (4, 4), 0, 1), ("cfg2", 2, (3, 3), 4, 4)])
def test_fused_sweeps_are_bitwise_the_half_sweeps(case_dir, name, nsteps, nu, planes, levels):
    d = case_dir(name)
    print(f"[gpu] fused sweeps {name} nu={nu} planes={planes} levels={levels}", flush=True)
    ref_stats, ref = _run(d, 0, nsteps, nu, planes)
    got_stats, got = _run(d, levels, nsteps, nu, planes)
    assert got_stats == ref_stats
    for f in ref:
        assert np.array_equal(got[f].view(np.int64), ref[f].view(np.int64)), f
    print(f"[gpu] fused sweeps {name}: identical over {nsteps} steps, Krylov steps {[x[0] for x in got_stats]}",
          flush=True)
    sys.stdout.flush()
