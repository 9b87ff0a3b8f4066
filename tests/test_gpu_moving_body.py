"""Moving body on the GPU.

- Geometry: the tiny case's sphere shifted by s*0.03 in y at steps 1..8;
  after the reference's own computeSurfaceNorm, selectiveRetagging,
  cellCount and computeNormDistance (run from its sources, fixtures in
  tests/golden/move_reference.json) the GPU's tags, nearest elements,
  lists and normal distances agree bit for bit.
- Flow: a prescribed transverse oscillation y_c = A sin(2 pi f t) with the
  wall velocity and acceleration in the forcing, solid cells and surface
  forces; every step's fields agree bit for bit with the oracle (which
  shares the reference's arithmetic; the reference has no moving wall).
"""
import hashlib
import json
import math
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dig(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


RT = {"cell": "cell.i64", "minElemcell": "minElemcell.i64", "ycent": "ycent.f64", "cosBeta": "cosBeta.f64"}
ND = {"interceptedIndexPtr": "interceptedIndexPtr.i64", "fluidIndexPtr": "fluidIndexPtr.i64",
      "solidIndexPtr": "solidIndexPtr.i64", "redCellIndexPtr": "redCellIndexPtr.i64",
      "blackCellIndexPtr": "blackCellIndexPtr.i64", "nelp": "nelp.i64", "pNormDis": "pNormDis.f64",
      "u1NormDis": "u1NormDis.f64", "u2NormDis": "u2NormDis.f64", "v1NormDis": "v1NormDis.f64",
      "v2NormDis": "v2NormDis.f64", "w1NormDis": "w1NormDis.f64", "w2NormDis": "w2NormDis.f64"}


def test_retagging_matches_reference(case_dir):
    from ib3d_amd.native import Solver
    with open(os.path.join(HERE, "golden", "move_reference.json")) as fh:
        g = json.load(fh)
    s = Solver.from_case(case_dir("tiny"))
    for st in g["steps"]:
        k = st["step"]
        _say(f"[gpu] moving geometry step {k}")
        s.move_body((0.0, st["yc"], 0.0))
        for key in ["ib", "fluid", "solid", "red", "black"]:
            assert s.counts[key] == st["counts"][key], (k, key)
        for name, f in RT.items():
            assert dig(s.download(name)) == g["arrays"][f"m{k:04d}_rt_{f}"], (k, name)
        for name, f in ND.items():
            assert dig(s.download(name)) == g["arrays"][f"m{k:04d}_nd_{f}"], (k, name)
    s.close()


def motion(t, amp, freq, dt):
    w = 2.0 * math.pi * freq
    y = amp * math.sin(w * t)
    v = amp * w * math.cos(w * t)
    vp = amp * w * math.cos(w * (t - dt))
    return y, v, (v - vp) / dt


def test_moving_steps_match_oracle(case_dir):
    import pyoracle
    from ib3d_amd.native import Solver
    d = case_dir("tiny")
    s = Solver.from_case(d)
    o = pyoracle.Oracle.from_case(d)
    dt = s.deltat
    amp, freq = 0.1, 2.0          # ~0.4 cell per step at this resolution
    for k in range(1, 5):
        t = k * dt
        y, v, a = motion(t, amp, freq, dt)
        _say(f"[gpu] moving step {k}: y_c={y:.5f} v={v:.4f} a={a:.3f}")
        s.move_body((0.0, y, 0.0), (0.0, v, 0.0), (0.0, a, 0.0))
        o.move_body((0.0, y, 0.0), (0.0, v, 0.0), (0.0, a, 0.0))
        s.first_step = True
        r = s.step()
        ro = o.step(True)
        assert r["niter"] == ro["niter"], (k, r["niter"], ro["niter"])
        for f in ("u", "v", "w", "p", "cell"):
            a_, b_ = s.download(f), o.array(f)
            assert np.array_equal(a_.view(np.int64), np.asarray(b_).view(np.int64)), (k, f)
        np.testing.assert_allclose(r["coef"], ro["coef"], rtol=1e-12, atol=1e-14)
    s.close()


def test_warm_started_modes_agree_with_cold_fp64_on_a_moving_body(case_dir):
    """Cells that stop being fluid keep a stale correction under a warm start
    unless it is cleared; the mixed and FP64 warm-started solves must match
    the cold FP64 solve to the solver tolerance after several moving steps."""
    from ib3d_amd.native import Solver
    d = case_dir("tiny")
    runs = {}
    for name, mode, warm in [("cold", 1, False), ("fp64", 1, True), ("mixed", 2, True), ("gmres", 3, True),
                             ("multigrid", 4, True)]:
        s = Solver.from_case(d)
        if mode >= 3:
            s.set_poisson_mode(mode, omega=1.9, tol=1e-9, check=20, warm=warm)
            s.set_krylov(12)
        else:
            s.set_poisson_mode(mode, omega=1.99, tol=1e-9, check=10, warm=warm)
        dt = s.deltat
        for k in range(1, 7):
            y, v, a = motion(k * dt, 0.1, 2.0, dt)
            s.move_body((0.0, y, 0.0), (0.0, v, 0.0), (0.0, a, 0.0))
            s.first_step = True
            s.step()
            assert s.poisson_stats()["relres"] <= 1e-9
        runs[name] = {f: s.download(f) for f in ("u", "p")}
        s.close()
    for name in ("fp64", "mixed", "gmres", "multigrid"):
        for f in ("u", "p"):
            e = float(np.max(np.abs(runs[name][f] - runs["cold"][f])) / np.max(np.abs(runs["cold"][f])))
            _say(f"[gpu] moving body, {name} warm vs cold FP64: {f} rel diff {e:.2e}")
            assert e < 1e-6, (name, f, e)


def test_forcing_plans_follow_the_latest_geometry(case_dir):
    """computeNormDistance only queues the forcing stencils; the plans are
    finished by the next forcing call.  A cellCount after it (new lists)
    drops the queued plans, so the forcing call fails as it would without
    computeNormDistance; computeNormDistance again, and the forcing matches
    a run without the extra calls, bit for bit."""
    import ctypes as C
    from ib3d_amd.native import IB3DError, Solver
    d = case_dir("tiny")

    def run(extra):
        s = Solver.from_case(d)
        try:
            s.move_body((0.0, 0.02, 0.0))
            s.nsMomentum()
            s.velocityBC()
            s.solidCellBC()
            if extra:
                cnt5 = np.zeros(5, dtype=np.int32)
                assert s.L.ib3d_cellCount(cnt5.ctypes.data_as(C.POINTER(C.c_int32))) == 0
                with pytest.raises(IB3DError, match="computeNormDistance has not run"):
                    s.velocityForcing1()
                assert s.L.ib3d_computeNormDistance() == 0
                assert s.L.ib3d_computeNormDistance() == 0   # a second queue replaces the first
            s.velocityForcing1()
            s.pressureForcing1()
            return {f: dig(s.download(f)) for f in ("ut", "vt", "wt", "p")}
        finally:
            s.close()

    plain = run(False)
    assert run(True) == plain
