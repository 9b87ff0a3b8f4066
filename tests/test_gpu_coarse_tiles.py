"""Coarse multigrid levels smoothed by temporally blocked launches
(k_smooth_tile: an 8x8x8 cube plus a halo of 2M cells swept M times in LDS,
two launches per smoothing phase) against the same solver with two in-place
half-sweep launches per sweep: the same FP32 operations on the same operands,
so every field, every Krylov count and every residual must be bit-identical.
Covers the first phase from zero, odd sweep counts (launches of 2 and 1
sweeps), one sweep per launch, counts above the kernel's limit (the
half-sweep launches are kept), the fused finest-level sweep beside it, and
the 16 x 16 x 8 blocks of a level of half a million cells or more (level 1
at 256x192x192)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(case, tiles, nsteps, nu, fused):
    from ib3d_amd import presets
    from ib3d_amd.native import Solver
    s = Solver.from_case(case)
    try:
        presets.apply(s, "mixed")
        mg = presets.MULTIGRID
        s.set_multigrid(4, 4, mg["omega"], mg["coarse"], mg["coarse_omega"], 0)
        s.set_multigrid_coarse(nu[0], nu[1])
        s.set_fused_sweeps(fused)
        s.set_coarse_tiles(tiles)
        stats = []
        for _ in range(nsteps):
            s.step()
            st = s.poisson_stats()
            stats.append((st["sweeps"], st["outer"], st["relres"]))
        return stats, {f: s.download(f) for f in ("u", "v", "w", "p", "pc")}
    finally:
        s.set_coarse_tiles()
        s.set_fused_sweeps()
        s.close()


@pytest.mark.parametrize("name,nsteps,nu,fused,first", [("small", 5, (4, 4), 0, 1), ("small", 4, (3, 2), 0, 1),
                                                        ("small", 3, (1, 5), 0, 1), ("cfg1", 4, (4, 4), 0, 1),
                                                        ("cfg1", 3, (3, 3), 1, 2), ("cfg2", 3, (4, 4), 1, 2),
                                                        ("multi", 3, (4, 4), 0, 1), ("cfg2", 2, (4, 4), 1, 1)])
def test_coarse_tiles_are_bitwise_the_half_sweeps(case_dir, name, nsteps, nu, fused, first):
    """first: the first level smoothed by the tiled launches"""
    d = case_dir(name)
    print(f"[gpu] coarse tiles {name} nu={nu} fused={fused} from level {first}", flush=True)
    ref_stats, ref = _run(d, 0, nsteps, nu, fused)
    got_stats, got = _run(d, first, nsteps, nu, fused)
    assert got_stats == ref_stats
    for f in ref:
        assert np.array_equal(got[f].view(np.int64), ref[f].view(np.int64)), f
    print(f"[gpu] coarse tiles {name}: identical over {nsteps} steps, Krylov steps {[x[0] for x in got_stats]}",
          flush=True)
    sys.stdout.flush()
