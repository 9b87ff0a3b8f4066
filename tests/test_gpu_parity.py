"""GPU parity: the HIP path against the oracle (the C restatement, itself
pinned bit for bit to the reference built from its own sources).

FP64 fields, tags, lists and stencil cells must agree bit for bit; the
surface-force coefficients are sums whose order differs (a tree on the
GPU, serial in the reference), so they are held to 1e-12 relative."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        return bool(np.all(a.view(np.int64) == b.view(np.int64)))
    return bool(np.all(a == b))


def assert_bits(name, a, b):
    if not bits_equal(a, b):
        a = np.asarray(a, dtype=np.float64)
        b = np.asarray(b, dtype=np.float64)
        nd = int(np.sum(a != b)) if a.shape == b.shape else -1
        md = float(np.max(np.abs(a - b))) if a.shape == b.shape and a.size else -1
        raise AssertionError(f"{name}: not bitwise equal (n_diff={nd}, max_abs={md:.3e})")


SETUP_ARRAYS = ["cell", "fluidIndexPtr", "solidIndexPtr", "interceptedIndexPtr", "redCellIndexPtr",
                "blackCellIndexPtr", "nelp", "pNormDis", "u1NormDis", "u2NormDis", "v1NormDis",
                "v2NormDis", "w1NormDis", "w2NormDis", "xcent", "ycent", "zcent", "cosAlpha",
                "cosBeta", "cosGamma", "alpha3", "beta3", "gamma3", "x1", "y1", "z1", "deltax",
                "xu", "yv", "zw", "xp", "yp", "zp", "Acx", "Acy", "Acz"]


@pytest.fixture
def pair(case_dir):
    from ib3d_amd.native import Solver
    import pyoracle

    made = []

    def make(name):
        d = case_dir(name)
        _say(f"[gpu] setup case={name} dir={d}")
        o = pyoracle.Oracle.from_case(d)
        s = Solver.from_case(d)
        s.case_dir = d
        made.append(s)
        return s, o
    yield make
    for s in made:
        s.close()


@pytest.mark.parametrize("name", ["tiny", "small", "multi"])
def test_setup_bitwise(pair, name):
    s, o = pair(name)
    assert_bits("minElemcell", s.download("minElemcell"), o.minElemcell)
    for a in SETUP_ARRAYS:
        assert_bits(a, s.download(a), o.array(a))
    for setn, n in [("uu", s.n[0]), ("uv", s.n[0]), ("vv", s.n[1]), ("vu", s.n[1]), ("ww", s.n[2]),
                    ("wu", s.n[2])]:
        t = o.array("tab_" + setn).reshape(12, n + 3)
        for q in range(12):
            nm = ("ca" if q < 6 else "ck") + str(q % 6 + 1) + "_" + setn
            assert_bits(nm, s.download(nm)[1:n], t[q, 2:n + 1])


@pytest.mark.parametrize("name,nsteps", [("tiny", 3), ("small", 2), ("multi", 2)])
def test_steps_stagewise_bitwise(pair, name, nsteps):
    s, o = pair(name)
    for st in range(1, nsteps + 1):
        first = st == 1
        _say(f"[gpu] case={name} step={st} nsMomentum")
        s.nsMomentum(); o.call("ns_momentum")
        for f in ["ut", "vt", "wt", "resi_u", "resi_v", "resi_w"]:
            assert_bits(f"step{st} nsMomentum {f}", s.download(f), o.array(f))
        s.velocityBC(); o.call("velocity_bc")
        s.solidCellBC(); o.call("solid_cell_bc")
        for f in ["ut", "vt", "wt", "p"]:
            assert_bits(f"step{st} BC {f}", s.download(f), o.array(f))
        _say(f"[gpu] case={name} step={st} velocityForcing1")
        s.velocityForcing1(); o.call("velocity_forcing")
        for f in ["ut", "vt", "wt"]:
            assert_bits(f"step{st} forcing {f}", s.download(f), o.array(f))
        s.velocityBC(); o.call("velocity_bc")
        _say(f"[gpu] case={name} step={st} poissonSolver")
        it, dm, ds, _ = s.poissonSolver()
        it_o, dm_o, ds_o = o.poisson()
        assert it == it_o, f"step{st}: nIterPcor {it} vs {it_o}"
        assert dm == dm_o and ds == ds_o
        for f in ["u", "v", "w", "ut", "vt", "wt", "p", "pc"]:
            assert_bits(f"step{st} poisson {f}", s.download(f), o.array(f))
        _say(f"[gpu] case={name} step={st} pressureForcing1")
        s.pressureForcing1(); o.call("pressure_forcing")
        assert_bits(f"step{st} pressureForcing p", s.download("p"), o.array("p"))
        c = s.stressCal(first)
        co = o.stress(first)
        np.testing.assert_allclose(c, co, rtol=1e-12, atol=1e-14)
        if first:
            for nm in ["Elemcell", "ucell", "vcell", "wcell", "pcell"]:
                assert_bits(nm, s.download(nm), o.array(nm))


def test_whole_step_entry_matches_stagewise(pair):
    s, o = pair("tiny")
    for st in range(1, 4):
        _say(f"[gpu] case=tiny ib3d_step {st}")
        r = s.step()
        ro = o.step(st == 1)
        assert r["niter"] == ro["niter"]
        assert r["divmax"] == ro["divmax"]
        np.testing.assert_allclose(r["coef"], ro["coef"], rtol=1e-12, atol=1e-14)
    for f in ["u", "v", "w", "p"]:
        assert_bits(f, s.download(f), o.array(f))


def test_per_body_forces_match_oracle(pair):
    """Two ellipsoids: the per-body coefficients of the GPU (tree sums)
    agree with the oracle's serial sums to 1e-12 relative."""
    import os
    from ib3d_amd import cases
    s, o = pair("multi")
    d = s.case_dir
    body = cases.read_gmsh_bodies(os.path.join(d, "geometries", "sphere.msh"), s.inp["surGeoPoints"])
    s.set_bodies(body)
    o.set_bodies(body)
    for k in range(2):
        s.step()
        o.step(k == 0)
        np.testing.assert_allclose(s.stress_bodies(), o.stress_bodies(), rtol=1e-12, atol=1e-14)


def test_momentum_bitwise_on_extreme_values(pair):
    """The predictor divides field values by grid quantities through
    precomputed reciprocals and FMA corrections; numerators outside the
    range where that is exact (zeros of either sign, values near the
    underflow threshold, values whose products exceed 2^900) must still give
    the correctly rounded quotient.  Fields filled with such values agree
    with the oracle bit for bit."""
    s, o = pair("small")
    nx, ny, nz = s.n
    shape = (nx + 2, ny + 2, nz + 2)
    rng = np.random.default_rng(11)
    for name, base in [("u", 1.0), ("v", 0.0), ("w", 0.0), ("p", 0.0)]:
        a = base + 0.3 * rng.standard_normal(shape)
        m = rng.random(shape)
        a[m < 0.10] = 0.0
        a[(m >= 0.10) & (m < 0.15)] = -0.0
        a[(m >= 0.15) & (m < 0.17)] = 1e-300 * np.sign(a[(m >= 0.15) & (m < 0.17)] + 0.5)
        a[(m >= 0.17) & (m < 0.18)] = 5e-324
        a[(m >= 0.18) & (m < 0.19)] = 1e150
        a[:, :, 3:6] = a[:, :, 2:3]          # equal planes: zero differences
        s.upload(name, a)
        o.field(name)[...] = a
    _say("[gpu] case=small momentum on extreme values")
    s.nsMomentum(); o.call("ns_momentum")
    for f in ["ut", "vt", "wt", "resi_u", "resi_v", "resi_w"]:
        assert_bits(f"extreme nsMomentum {f}", s.download(f), o.array(f))
