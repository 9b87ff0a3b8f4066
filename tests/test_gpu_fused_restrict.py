"""The multigrid residual restricted in one z-marching launch
(k_resid_restrict_pz: each thread forms its coarse cell's eight children's
residuals from coalesced rows and sums them) against the two-pass forms (the
finest level's residual stored by the matrix-vector kernel and averaged by a
second launch; below it, one thread per coarse slot gathering every child's
stencil): the same FP32 operations in the same order, so every field, every
Krylov count and every residual must be bit-identical over whole solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(case, levels, nsteps):
    from ib3d_amd import presets
    from ib3d_amd.native import Solver
    s = Solver.from_case(case)
    try:
        presets.apply(s, "mixed")
        s.set_fused_restrict(levels)
        stats = []
        for _ in range(nsteps):
            s.step()
            st = s.poisson_stats()
            stats.append((st["sweeps"], st["outer"], st["relres"]))
        return stats, {f: s.download(f) for f in ("u", "v", "w", "p", "pc")}
    finally:
        s.set_fused_restrict()
        s.close()


@pytest.mark.parametrize("name,nsteps,levels", [("small", 6, 1), ("small", 4, 64), ("multi", 4, 2), ("cfg1", 4, 1),
                                                ("cfg1", 3, 64), ("cfg2", 3, 1), ("cfg2", 2, 64)])
def test_one_pass_restriction_is_bitwise_the_two_pass_one(case_dir, name, nsteps, levels):
    d = case_dir(name)
    print(f"[gpu] fused restriction {name} levels={levels}", flush=True)
    ref_stats, ref = _run(d, 0, nsteps)
    got_stats, got = _run(d, levels, nsteps)
    assert got_stats == ref_stats
    for f in ref:
        assert np.array_equal(got[f].view(np.int64), ref[f].view(np.int64)), f
    print(f"[gpu] fused restriction {name}: identical over {nsteps} steps, Krylov steps {[x[0] for x in got_stats]}",
          flush=True)
