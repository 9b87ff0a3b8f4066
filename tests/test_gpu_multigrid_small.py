"""ib3d_set_multigrid_small acts on exactly the levels cycled inside the
one-workgroup kernel: the coarse counts set to (5, 4) with every level cycled
by launches held at 4+4 by ib3d_set_multigrid_level must give, bit for bit,
what set_multigrid_small(5, 4) gives with the coarse counts at 4+4.  The
in-kernel levels are the coarsest whose arrays together fit in 64 KB of LDS:
from level 3 (16x8x8) at 128x64x64 and from level 4 (16x12x12) at
256x192x192.  The counts must also change the solve (the fields differ from
4+4 there)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(case, nsteps, setup):
    from ib3d_amd import presets
    from ib3d_amd.native import Solver
    s = Solver.from_case(case)
    try:
        presets.apply(s, "mixed")
        setup(s)
        stats = []
        for _ in range(nsteps):
            s.step()
            st = s.poisson_stats()
            stats.append((st["sweeps"], st["outer"], st["relres"]))
        return stats, {f: s.download(f) for f in ("u", "v", "w", "p", "pc")}
    finally:
        presets.apply(s, "mixed")
        s.close()


def _same(a, b):
    return a[0] == b[0] and all(np.array_equal(a[1][f].view(np.int64), b[1][f].view(np.int64)) for f in a[1])


@pytest.mark.parametrize("name,nsteps,launched", [("cfg1", 4, (1, 2)), ("cfg2", 3, (1, 2, 3))])
def test_small_level_sweeps_reach_exactly_the_in_kernel_levels(case_dir, name, nsteps, launched):
    d = case_dir(name)
    print(f"[gpu] in-kernel sweep counts {name}", flush=True)

    def by_knob(s):
        s.set_multigrid_coarse(4, 4)
        s.set_multigrid_small(5, 4)

    def by_levels(s):
        s.set_multigrid_small(-1, -1)
        s.set_multigrid_coarse(5, 4)
        for lv in launched:
            s.set_multigrid_level(lv, 4, 4)

    def plain(s):
        s.set_multigrid_coarse(4, 4)
        s.set_multigrid_small(-1, -1)

    a = _run(d, nsteps, by_knob)
    b = _run(d, nsteps, by_levels)
    c = _run(d, nsteps, plain)
    assert _same(a, b)
    assert not _same(a, c)
    for st in a[0]:
        assert st[2] <= 1e-6
    print(f"[gpu] in-kernel sweep counts {name}: identical over {nsteps} steps, Krylov steps "
          f"{[x[0] for x in a[0]]} (4+4: {[x[0] for x in c[0]]})", flush=True)
