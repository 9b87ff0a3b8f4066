"""Each GMRES-IR refinement iteration (Krylov cycle, FP64 residual and its
norm) can be replayed as one captured HIP graph.  The graph runs the same
kernels in the same order as the uncaptured launches, so the fields agree
bit for bit.  Four steps replay the graph several times per solve, with the
host's own synchronisations and eager work (predictor, forcing, the
first residual) between the replays."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(case, mode, graphs):
    from ib3d_amd.native import Solver
    s = Solver.from_case(case)
    try:
        s.set_poisson_mode(mode, tol=1e-6, warm=True)
        s.set_krylov(6)
        if mode == 4:
            s.set_multigrid(2, 2, 1.55, 8, 1.5, 0)
        s.set_graphs(graphs)
        stats = []
        for st in range(4):
            _say(f"[gpu] mode={mode} graphs={graphs} step={st + 1}")
            s.step()
            ps = s.poisson_stats()
            stats.append((ps["sweeps"], ps["outer"]))
        fields = [s.download(f).copy() for f in ("u", "v", "w", "p")]
    finally:
        s.close()
    return fields, stats


@pytest.mark.parametrize("mode", [3, 4])
def test_graph_replay_matches_launches(case_dir, mode):
    case = case_dir("small")
    f_eager, s_eager = _run(case, mode, False)
    f_graph, s_graph = _run(case, mode, True)
    assert s_graph == s_eager
    for name, a, b in zip(("u", "v", "w", "p"), f_eager, f_graph):
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), f"{name} differs with the graph"
