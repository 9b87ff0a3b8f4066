"""nsMomentum with its interior columns from LDS tiles marching in z (64 x 8
or 64 x 4 columns; each plane of u, v, w staged once with a two-cell ring,
each column's planes k-2..k+2 in registers) and the direct kernel with
64-bit indices (variant 4) against the default direct kernel (32-bit byte
offsets): the same arithmetic in the same order, so ut, vt, wt and the
residual arrays must be bit-identical, also against the oracle, and on
fields with extreme values."""
import numpy as np
import pytest

from test_gpu_parity import assert_bits

pytestmark = pytest.mark.gpu
OUT = ("ut", "vt", "wt", "resi_u", "resi_v", "resi_w")


def _predict(s, tiles):
    s.set_momentum_tiles(tiles)
    s.nsMomentum()
    return {f: s.download(f) for f in OUT}


@pytest.mark.parametrize("name", ["tiny", "small", "multi", "cfg1", "cfg2"])
def test_tiled_predictor_is_bitwise_the_direct_one(case_dir, name):
    from ib3d_amd.native import Solver
    s = Solver.from_case(case_dir(name))
    try:
        for step in range(3):
            a = _predict(s, 0)
            for variant in (1, 2, 3, 4):
                b = _predict(s, variant)
                for f in OUT:
                    assert_bits(f"{name} step {step} variant {variant} {f}", b[f], a[f])
            s.set_momentum_tiles(0)
            s.step()
        print(f"[gpu] tiled predictor bit-identical on {name}", flush=True)
    finally:
        s.set_momentum_tiles(0)
        s.close()


def test_tiled_predictor_matches_oracle_on_extreme_values(case_dir):
    from ib3d_amd.native import Solver
    import pyoracle
    d = case_dir("small")
    s = Solver.from_case(d)
    o = pyoracle.Oracle.from_case(d)
    try:
        rng = np.random.default_rng(99)
        n = s.download("u").size
        vals = np.array([0.0, -0.0, 1e-300, 5e-324, 1e150, -1e150, 1.0, -2.5])
        for f in ("u", "v", "w", "p"):
            x = rng.standard_normal(n) * rng.choice([1.0, 1e-200, 1e100], n)
            x[rng.integers(0, n, n // 20)] = rng.choice(vals, n // 20)
            s.upload(f, x)
            o.array(f)[:] = x
        o.call("ns_momentum")
        for variant in (0, 1, 2, 3, 4):
            s.set_momentum_tiles(variant)
            s.nsMomentum()
            for f in OUT:
                assert_bits(f"extreme variant {variant} {f}", s.download(f), o.array(f))
    finally:
        s.set_momentum_tiles(0)
        s.close()
