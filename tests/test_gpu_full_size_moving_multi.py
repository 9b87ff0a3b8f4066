"""Configurations 3 and 5 at their stated sizes.

* 384x256x256, the oscillating sphere (20480 triangles): setup, one
  geometry rebuild (selective retagging, lists, normal distances) and the
  predictor and velocity forcing on the moved body agree with the oracle bit
  for bit; then three mixed-precision steps of the oscillation each meet
  ||b/dt - A pc|| <= 1e-6 ||b/dt|| (recomputed in FP64 from the final pc), and
  their fields agree with FP64 SOR from a zero guess to what that tolerance
  allows.
* 512x320x320, eight ellipsoids (seed 2026, 4096 triangles each): the tags
  and nearest elements of a three-plane slab through the cluster agree with
  the oracle bit for bit (the oracle's brute-force search over all 32768
  elements is too slow for the whole window on one core), three
  mixed-precision steps meet the residual bound, and the per-body drag and
  lift coefficients, weighted by each body's projected area, add up to the
  global ones."""
import math
import os
import shutil
import sys

import numpy as np
import pytest

from test_gpu_parity import SETUP_ARRAYS, assert_bits

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(900)
def test_cfg3_full_size_rebuild_bitwise_and_mixed_steps(case_dir):
    from ib3d_amd import presets
    from ib3d_amd.native import Solver
    import pyoracle
    d = case_dir("cfg3")
    _say("[gpu] cfg3 384x256x256: oracle setup")
    o = pyoracle.Oracle.from_case(d)
    s = Solver.from_case(d)
    try:
        assert s.n == (384, 256, 256)
        for a in SETUP_ARRAYS:
            assert_bits(f"cfg3 setup {a}", s.download(a), o.array(a))
        # the oscillation's displacement after 20 steps, its wall velocity and
        # the finite-difference acceleration of the step before
        a0, om, dt = 0.25, 2.0 * math.pi * 0.2, 0.002
        t = 20 * dt
        yc, vel = a0 * math.sin(om * t), a0 * om * math.cos(om * t)
        acc = (vel - a0 * om * math.cos(om * (t - dt))) / dt
        _say(f"[gpu] cfg3: geometry rebuild at y_c = {yc:.6f} (oracle about a minute)")
        s.move_body((0.0, yc, 0.0), (0.0, vel, 0.0), (0.0, acc, 0.0))
        o.move_body((0.0, yc, 0.0), (0.0, vel, 0.0), (0.0, acc, 0.0))
        for a in ["cell", "minElemcell", "fluidIndexPtr", "solidIndexPtr", "interceptedIndexPtr",
                  "redCellIndexPtr", "blackCellIndexPtr", "nelp", "pNormDis", "u1NormDis", "u2NormDis",
                  "v1NormDis", "v2NormDis", "w1NormDis", "w2NormDis", "ycent", "cosBeta"]:
            assert_bits(f"cfg3 moved {a}", s.download(a), o.array(a))
        s.nsMomentum(); o.call("ns_momentum")
        s.velocityBC(); o.call("velocity_bc")
        s.solidCellBC(); o.call("solid_cell_bc")
        s.velocityForcing1(); o.call("velocity_forcing")
        for f in ["ut", "vt", "wt"]:
            assert_bits(f"cfg3 moved forcing {f}", s.download(f), o.array(f))
    finally:
        s.close()
    del o

    def run(preset):
        s = Solver.from_case(d)
        try:
            presets.apply(s, preset)
            vprev = 0.0
            for n in range(1, 4):
                t = n * dt
                yc, vel = a0 * math.sin(om * t), a0 * om * math.cos(om * t)
                s.move_body((0.0, yc, 0.0), (0.0, vel, 0.0), (0.0, (vel - vprev) / dt, 0.0))
                vprev = vel
                s.first_step = True
                s.step()
                rel = s.poisson_stats()["relres"]
                rn, bn = s.residual_norm()
                _say(f"[gpu] cfg3 {preset} step {n}: relres {rel:.3e}, recomputed {rn / bn:.3e}")
                assert rel <= 1e-6 and rn <= 1e-6 * bn * (1 + 1e-9)
            return {f: s.download(f) for f in ("u", "v", "w", "p")}
        finally:
            s.close()

    mixed = run("mixed")
    cold = run("fp64-cold")
    err = {k: float(np.max(np.abs(mixed[k] - cold[k])) / np.max(np.abs(cold[k]))) for k in mixed}
    _say("[gpu] cfg3 mixed vs FP64 zero-guess after 3 moving steps: " + ", ".join(f"{k}={v:.3e}" for k, v in err.items()))
    assert err["u"] < 1e-5 and err["v"] < 1e-4 and err["w"] < 1e-4 and err["p"] < 1e-3


def _slab_case(src, dst, k_plane):
    """The case `src` with the search window cut to the planes k_plane-1..k_plane+1."""
    os.makedirs(dst, exist_ok=True)
    for f in ("gridx.dat", "gridy.dat", "gridz.dat"):
        shutil.copy(os.path.join(src, f), dst)
    shutil.copytree(os.path.join(src, "geometries"), os.path.join(dst, "geometries"), dirs_exist_ok=True)
    with open(os.path.join(src, "inputdata")) as fh:
        lines = fh.read().splitlines()
    lines[-1] = f"      {k_plane - 1}, {k_plane + 1}"
    with open(os.path.join(dst, "inputdata"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


@pytest.mark.timeout(900)
def test_cfg5_full_size_slab_bitwise_steps_and_per_body(case_dir, tmp_path):
    from ib3d_amd import cases, presets
    from ib3d_amd.native import Solver
    import pyoracle
    d = case_dir("cfg5")
    inp = cases.read_inputdata(os.path.join(d, "inputdata"))
    kmid = (inp["k_startSearch"] + inp["k_endSearch"]) // 2
    slab = str(tmp_path / "slab")
    _slab_case(d, slab, kmid)
    _say(f"[gpu] cfg5 512x320x320: oracle tagging of the planes {kmid - 1}..{kmid + 1}")
    o = pyoracle.Oracle.from_case(slab)
    s = Solver.from_case(d)
    try:
        assert s.n == (512, 320, 320)
        nx, ny, nz = s.n
        sl = (slice(inp["i_startSearch"] - 1, inp["i_endSearch"]), slice(inp["j_startSearch"] - 1, inp["j_endSearch"]),
              slice(kmid - 2, kmid + 1))
        for a in ("cell", "minElemcell"):
            g = s.download(a).reshape((nx + 2, ny + 2, nz + 2), order="F")[sl]
            r = o.array(a).reshape((nx + 2, ny + 2, nz + 2), order="F")[sl]
            assert np.array_equal(g, r), a
        tags = s.download("cell").reshape((nx + 2, ny + 2, nz + 2), order="F")[sl]
        _say(f"[gpu] cfg5 slab tags bit-identical: {int((tags == 2).sum())} intercepted, "
             f"{int((tags == 1).sum())} solid cells in the slab")
        del o
        body = cases.read_gmsh_bodies(os.path.join(d, "geometries", "sphere.msh"), inp["surGeoPoints"])
        s.set_bodies(body)
        presets.apply(s, "mixed")
        for n in range(1, 4):
            r = s.step()
            rel = s.poisson_stats()["relres"]
            rn, bn = s.residual_norm()
            _say(f"[gpu] cfg5 mixed step {n}: relres {rel:.3e}, recomputed {rn / bn:.3e}")
            assert rel <= 1e-6 and rn <= 1e-6 * bn * (1 + 1e-9)
        c = r["coef"]
        per = s.stress_bodies()
        a3, b3 = s.download("alpha3"), s.download("beta3")
        ax = np.array([0.5 * np.sum(0.5 * np.abs(b3[body == k])) for k in range(1, 9)])
        ay = np.array([0.5 * np.sum(0.5 * np.abs(a3[body == k])) for k in range(1, 9)])
        for k in range(8):
            _say(f"[gpu] cfg5 body {k + 1}: Cd_visc {per[k, 0]:.5f} Cd_press {per[k, 1]:.5f} "
                 f"Cl_visc {per[k, 2]:.5f} Cl_press {per[k, 3]:.5f}")
        np.testing.assert_allclose(np.sum(per[:, 0] * ax) / ax.sum(), c[0], rtol=1e-10)
        np.testing.assert_allclose(np.sum(per[:, 1] * ay) / ay.sum(), c[1], rtol=1e-10)
        np.testing.assert_allclose(np.sum(per[:, 2] * ay) / ay.sum(), c[2], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(np.sum(per[:, 3] * ax) / ax.sum(), c[3], rtol=1e-10, atol=1e-12)
    finally:
        s.close()
