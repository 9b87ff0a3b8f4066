"""The OpenMP build of the C restatement (the all-cores CPU baseline) gives
the serial build's results bit for bit: its parallel loops update distinct
cells and reduce only with max.  Two steps of the 32x24x24 sphere in each
build, in child processes (the thread count is fixed at load)."""
import hashlib
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OMP_LIB = os.path.join(ROOT, "oracle", "_build", "libib3d_oracle_omp.so")

CHILD = r"""
import hashlib, json, sys
sys.path.insert(0, {oracle!r}); sys.path.insert(0, {pkg!r})
import pyoracle
if {omp}:
    pyoracle.use_openmp()
o = pyoracle.Oracle.from_case({case!r})
its = [o.step(first=(n == 0))["niter"] for n in range(2)]
dig = {{f: hashlib.sha256(o.field(f).tobytes()).hexdigest() for f in ("u", "v", "w", "p", "pc")}}
print(json.dumps({{"its": its, "dig": dig}}))
"""


def _run(case, omp, threads):
    code = CHILD.format(oracle=os.path.join(ROOT, "oracle"), pkg=os.path.join(ROOT, "ib3d-gpu_amd"), omp=omp,
                        case=case)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=600,
                         check=True).stdout
    return json.loads(out.strip().splitlines()[-1])


@pytest.mark.skipif(not os.path.exists(OMP_LIB), reason="OpenMP oracle not built (make -C oracle)")
def test_openmp_build_is_bitwise_the_serial_one(case_dir):
    d = case_dir("tiny")
    serial = _run(d, False, 1)
    par = _run(d, True, 4)
    assert serial["its"] == par["its"]
    assert serial["dig"] == par["dig"]
