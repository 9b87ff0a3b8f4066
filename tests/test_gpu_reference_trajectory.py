"""GPU against the reference's own 200-step run of the 128x64x64 sphere
(tests/golden/cfg1_reference.json, produced by the reference built from its
sources): SOR iteration count and divmax of every step, drag and lift of
every step, and the bitwise digests of the fields after step 1 and step 200.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def dig(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_cfg1_200_steps_match_the_reference(case_dir):
    from ib3d_amd.native import Solver
    with open(os.path.join(HERE, "golden", "cfg1_reference.json")) as fh:
        g = json.load(fh)
    d = case_dir("cfg1")
    print("[gpu] cfg1 reference trajectory, 200 steps", flush=True)
    s = Solver.from_case(d)
    try:
        for name, key in [("cell", "cell.i64"), ("interceptedIndexPtr", "interceptedIndexPtr.i64"),
                          ("fluidIndexPtr", "fluidIndexPtr.i64"), ("pNormDis", "pNormDis.f64")]:
            assert dig(s.download(name)) == g["arrays"][key], name
        for st in range(1, 201):
            r = s.step()
            ref = g["steps"][st - 1]
            assert r["niter"] == ref["nIterPcor"], (st, r["niter"], ref["nIterPcor"])
            assert r["divmax"] == ref["divmax"], (st, r["divmax"], ref["divmax"])
            np.testing.assert_allclose(r["coef"][:2], g["drag"][st - 1][:2], rtol=1e-11, atol=1e-13)
            np.testing.assert_allclose(r["coef"][2:], g["lift"][st - 1][:2], rtol=1e-9, atol=1e-13)
            if st in (1, 200):
                pre = f"s{st:04d}_"
                for f in ["u", "v", "w"]:
                    assert dig(s.download(f)) == g["arrays"][pre + f"poi_{f}.f64"], (st, f)
                assert dig(s.download("p")) == g["arrays"][pre + "prf_p.f64"], (st, "p")
            if st % 50 == 0:
                print(f"[gpu] cfg1 step {st} ok (nIterPcor {r['niter']})", flush=True)
        cd = r["coef"][0] + r["coef"][1]
        ref_cd = g["drag"][199][0] + g["drag"][199][1]
        assert abs(cd - ref_cd) <= 0.005 * abs(ref_cd)
    finally:
        s.close()
