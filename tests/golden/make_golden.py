#!/usr/bin/env python3
"""Writes the golden fixtures from the reference itself.

The reference's Fortran is compiled from its own sources into
oracle/_ref/ref_harness (oracle/Makefile.ref; only where the reference is
mounted).  This script generates a case, runs the harness on it and stores
compact fixtures:

* SHA-256 digests of every dumped array's exact bytes (a digest match is a
  bitwise match of the whole array);
* per-step SOR iteration counts, divmax and the Cd/Cl lines the reference
  appended to dragcoff.dat / liftcoff.dat;
* a handful of small arrays verbatim (for readable failures).

Usage:
  python tests/golden/make_golden.py tiny            # ~1 s
  python tests/golden/make_golden.py cfg1 [RUN_DIR]  # 200 steps, ~25 min
      (RUN_DIR: an existing harness run of the same case to reuse)
  python tests/golden/make_golden.py multi           # two ellipsoids, 3 steps, ~2 s
  python tests/golden/make_golden.py move            # moving-body geometry, ~2 s
      (tiny case, body shifted by s*0.03 in y at step s = 1..8; the
       reference's computeSurfaceNorm, selectiveRetagging, cellCount and
       computeNormDistance after every shift)
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ib3d-gpu_amd"))
from ib3d_amd import cases  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def digest(path):
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def input_digests(case):
    out = {}
    for f in ["inputdata", "gridx.dat", "gridy.dat", "gridz.dat", os.path.join("geometries", "sphere.msh")]:
        out[f] = digest(os.path.join(case, f))
    return out


def collect(case, dump, nsteps, dump_steps):
    g = {"inputs": input_digests(case), "arrays": {}, "steps": [], "drag": [], "lift": []}
    for fn in sorted(os.listdir(dump)):
        if fn.endswith((".f64", ".i64")):
            g["arrays"][fn] = digest(os.path.join(dump, fn))
    with open(os.path.join(dump, "counts.txt")) as fh:
        a = fh.readline().split()
        g["counts"] = dict(zip(["nx", "ny", "nz", "ib", "fluid", "solid", "red", "black", "ibElems"],
                               [int(x) for x in a]))
    with open(os.path.join(dump, "steps.txt")) as fh:
        for line in fh:
            p = line.split()
            g["steps"].append({"ita": int(p[0]), "nIterPcor": int(p[1]), "divmax": float(p[2]),
                               "totime": float(p[3])})
    for name, key in [("dragcoff.dat", "drag"), ("liftcoff.dat", "lift")]:
        with open(os.path.join(case, name)) as fh:
            g[key] = [[float(x) for x in line.split()] for line in fh if line.strip()]
    g["dump_steps"] = dump_steps
    g["nsteps"] = nsteps
    return g


def run_harness(case, nsteps, dump_steps):
    dump = os.path.join(case, "dump")
    os.makedirs(dump, exist_ok=True)
    for f in ["dragcoff.dat", "liftcoff.dat", "iter.dat"]:
        if os.path.exists(os.path.join(case, f)):
            os.remove(os.path.join(case, f))
    subprocess.run([HARNESS, "run", str(nsteps), dump] + [str(s) for s in dump_steps], cwd=case, check=True,
                   stdout=subprocess.DEVNULL)
    return dump


MOVE_DY = 0.03
MOVE_STEPS = 8


def make_move():
    tmp = tempfile.mkdtemp(prefix="golden_move_")
    try:
        case = tmp
        cases.write_case(cases.CONFIGS["tiny"], case)
        dump = os.path.join(case, "dump")
        os.makedirs(dump)
        subprocess.run([HARNESS, "move", str(MOVE_STEPS), dump, repr(MOVE_DY)], cwd=case, check=True,
                       stdout=subprocess.DEVNULL)
        g = {"inputs": input_digests(case), "dy": MOVE_DY, "nsteps": MOVE_STEPS, "arrays": {}, "steps": []}
        for fn in sorted(os.listdir(dump)):
            if fn.endswith((".f64", ".i64")):
                g["arrays"][fn] = digest(os.path.join(dump, fn))
        with open(os.path.join(dump, "move_steps.txt")) as fh:
            for line in fh:
                p = line.split()
                g["steps"].append({"step": int(p[0]), "yc": float(p[1]),
                                   "counts": dict(zip(["ib", "fluid", "solid", "red", "black"],
                                                      [int(x) for x in p[2:7]]))})
        with open(os.path.join(HERE, "move_reference.json"), "w") as fh:
            json.dump(g, fh, indent=0, sort_keys=True)
        print(f"wrote move fixtures: {len(g['arrays'])} array digests, {len(g['steps'])} steps")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    which = sys.argv[1] if len(sys.argv) > 1 else "tiny"
    if not os.path.exists(HARNESS):
        raise SystemExit(f"{HARNESS} missing: build it with make -C oracle -f Makefile.ref")
    if which == "move":
        return make_move()
    cfg = cases.CONFIGS[which]
    if which in ("tiny", "multi"):
        nsteps, dump_steps = 3, [1, 2, 3]
    elif which == "cfg1":
        nsteps, dump_steps = 200, [1, 200]
    else:
        raise SystemExit("tiny, multi, cfg1 or move")
    reuse = sys.argv[2] if len(sys.argv) > 2 else None
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        if reuse:
            case = reuse
            dump = os.path.join(case, "dump")
        else:
            case = tmp
            cases.write_case(cfg, case)
            dump = run_harness(case, nsteps, dump_steps)
        # the case written now must be the one the run used
        fresh = os.path.join(tmp, "fresh")
        cases.write_case(cfg, fresh)
        if input_digests(fresh) != input_digests(case):
            raise SystemExit("the reused run was made from different case files")
        g = collect(case, dump, nsteps, dump_steps)
        with open(os.path.join(HERE, f"{which}_reference.json"), "w") as fh:
            json.dump(g, fh, indent=0, sort_keys=True)
        if which == "tiny":
            keep = {}
            for name in ["cell.i64", "interceptedIndexPtr.i64", "nelp.i64", "pNormDis.f64", "u2NormDis.f64",
                         "Acx.f64", "Acy.f64", "Acz.f64", "ca_uu.f64", "ca_vu.f64", "Elemcell.i64", "pcell.i64"]:
                dt = np.int64 if name.endswith(".i64") else np.float64
                a = np.fromfile(os.path.join(dump, name), dtype=dt)
                keep[name.split(".")[0]] = a.astype(np.int8) if name == "cell.i64" else a
            np.savez_compressed(os.path.join(HERE, "tiny_reference_arrays.npz"), **keep)
        print(f"wrote fixtures for {which}: {len(g['arrays'])} array digests, {len(g['steps'])} steps")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
