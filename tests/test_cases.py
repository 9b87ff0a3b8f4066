"""Case generators and the reference's file formats."""
import os

import numpy as np
import pytest

from ib3d_amd import cases

REFERENCE = "/root/reference"


@pytest.mark.parametrize("name", ["tiny", "small", "cfg1", "cfg2", "cfg3"])
def test_stretched_axes(name):
    cfg = cases.CONFIGS[name]
    for d in range(3):
        nodes, (rl, rr, ml, mr) = cases.stretched_axis(cfg.lo[d], cfg.hi[d], cfg.ulo[d], cfg.uhi[d], cfg.h, cfg.n[d])
        assert nodes.size == cfg.n[d] + 1
        assert nodes[0] == cfg.lo[d] and nodes[-1] == cfg.hi[d]
        w = np.diff(nodes)
        assert np.all(w > 0)
        inner = (nodes[:-1] >= cfg.ulo[d] - 1e-12) & (nodes[1:] <= cfg.uhi[d] + 1e-12)
        np.testing.assert_allclose(w[inner], cfg.h, rtol=1e-9)
        assert 1.0 <= rl < 1.3 and 1.0 <= rr < 1.3


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_icosphere_is_a_closed_outward_surface(level):
    v, f = cases.icosphere(level, 0.5)
    assert len(f) == 20 * 4 ** level
    assert len(v) == 10 * 4 ** level + 2
    assert len(v) - (3 * len(f)) // 2 + len(f) == 2          # Euler characteristic of a sphere
    p1, p2, p3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    nrm = np.cross(p2 - p1, p3 - p1)
    cen = (p1 + p2 + p3) / 3
    assert np.all(np.einsum("ij,ij->i", nrm, cen) > 0)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 0.5, rtol=1e-14)
    edges = {}
    for a, b, c in f:
        for e in [(a, b), (b, c), (c, a)]:
            edges[e] = edges.get(e, 0) + 1
    assert all((b, a) in edges for (a, b) in edges)            # every edge shared with opposite orientation


def test_case_files_round_trip(tmp_path):
    cfg = cases.CONFIGS["tiny"]
    meta = cases.write_case(cfg, str(tmp_path))
    inp = cases.read_inputdata(os.path.join(tmp_path, "inputdata"))
    assert (inp["nx"], inp["ny"], inp["nz"]) == cfg.n
    assert inp["dt_order"] * 1e-3 == cfg.dt
    assert inp["surGeoPoints"] == cfg.n_points
    gx = cases.read_grid(os.path.join(tmp_path, "gridx.dat"), cfg.n[0] + 1)
    nodes, _ = cases.stretched_axis(cfg.lo[0], cfg.hi[0], cfg.ulo[0], cfg.uhi[0], cfg.h, cfg.n[0])
    assert np.array_equal(gx, nodes)
    v, f = cases.icosphere(cfg.level, cfg.radius)
    nv, conn = cases.read_gmsh(os.path.join(tmp_path, "geometries", "sphere.msh"), inp["surGeoPoints"])
    assert np.array_equal(nv, v)
    assert np.array_equal(conn, f + 1)
    assert meta["window"][0][0] == inp["i_startSearch"]


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference not mounted")
def test_reads_the_reference_shipped_case():
    """The shipped inputdata (27 list-directed values, inputdata:1-10) and
    grid files (282 x 242 x 242 cells)."""
    inp = cases.read_inputdata(os.path.join(REFERENCE, "inputdata"))
    assert (inp["nx"], inp["ny"], inp["nz"]) == (282, 242, 242)
    assert inp["re"] == 100.0 and inp["itamax"] == 52000 and inp["eps1"] == 1e-5
    assert inp["dt_order"] == 0.75 and inp["inor"] == 1 and inp["surGeoPoints"] == 162
    assert (inp["i_startSearch"], inp["k_endSearch"]) == (78, 166)
    gx = cases.read_grid(os.path.join(REFERENCE, "gridx.dat"), 283)
    assert gx[0] == 0.0 and np.all(np.diff(gx) > 0)


def test_result_file_round_trip(tmp_path):
    """write_result/read_result keep every double (shortest round-trip text)
    and the reference's record order (i fastest, subWriteOutput.f90:65-69)."""
    from ib3d_amd import cases
    rng = np.random.default_rng(7)
    n = (5, 4, 3)
    shape = (n[0] + 2, n[1] + 2, n[2] + 2)
    f = [rng.standard_normal(shape) * 10.0 ** rng.integers(-12, 12, size=shape) for _ in range(4)]
    path = str(tmp_path / "result")
    cases.write_result(path, *f, 0.125, 7, 9)
    with open(path) as fh:
        first = fh.readline().split()
        second = fh.readline().split()
    assert float(first[0]) == f[0][0, 0, 0] and float(second[0]) == f[0][1, 0, 0]
    got = cases.read_result(path, n)
    for a, b in zip(got[:4], f):
        assert np.array_equal(a, b)
    assert got[4:] == (0.125, 7, 9)


def test_iter_dat_columns(tmp_path):
    """FORMAT(' ',I8,I10,4E15.6) of subPcorVcor.f90:62, as Fortran writes it."""
    from ib3d_amd import cases
    path = str(tmp_path / "iter.dat")
    with open(path, "w") as fh:
        fh.write(" " + "%8d" % 12 + "%10d" % 2380 + "%15s" % "0.707429E+01" + "%15s" % "0.123457E-02"
                 + "%15s" % "0.456000E+00" + "\n")
    (ita, nit, dm, ds, t), = cases.read_iter_dat(path)
    assert (ita, nit) == (12, 2380)
    assert dm == 7.07429 and ds == 0.00123457 and t == 0.456


def test_fortran_driver_fails_loudly_without_a_gpu(tmp_path):
    """bin/ib3d_main stops with the library's message and a non-zero status
    where no GPU is visible (there is no CPU path behind the drop-in)."""
    import os
    import subprocess
    from ib3d_amd import cases
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    driver = os.path.join(root, "ib3d-gpu_amd", "bin", "ib3d_main")
    if not os.path.exists(driver):
        pytest.skip("Fortran driver not built")
    cases.write_case(cases.CONFIGS["tiny"], str(tmp_path))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([driver, "1"], cwd=str(tmp_path), capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode != 0
    assert "ib3d: ib3d_init" in r.stdout + r.stderr
