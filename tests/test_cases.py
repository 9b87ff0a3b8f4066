"""Case generators and the reference's file formats."""
import os

import numpy as np
import pytest

from ib3d_amd import cases

REFERENCE = "/root/reference"


@pytest.mark.parametrize("name", ["tiny", "small", "cfg1", "cfg2"])
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
