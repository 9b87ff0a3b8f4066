#!/usr/bin/env python3
"""Benchmark: steps/s and Poisson ms/step of the IB3D hot path on one MI355X.

Workload: the fixed sphere at Re = 100 on the 256x192x192 stretched grid
(h = D/40 near the body, 20480-triangle icosphere, dt = 0.002), started from
the reference's initial condition (u = 1 on fluid cells).  A "step" is one
full time step of the hot path (momentum predictor, boundary conditions,
IB velocity forcing, Poisson solve + projection, IB pressure forcing,
surface forces), all on the GPU with inputs resident in HBM.

Poisson modes (same stopping rule: ||b/dt - A p||_2 <= 1e-6 ||b/dt||_2):
  fp64-cold     -- plain FP64 red-black SOR (omega 1.99, the reference's)
                   from a zero guess: the reference's method with the L2
                   stopping rule (omega 1.7 needs more than 20000 sweeps per
                   solve on this grid)
  fp64          -- the same from the previous step's correction
  mixed-dc      -- FP32 red-black SOR sweeps (omega 1.99) on the defect
                   equation inside an FP64 defect-correction loop,
                   previous-step guess
  gmres-sor     -- FP64 refinement whose defect equation is solved by FP32
                   flexible GMRES(12), every Arnoldi step preconditioned by
                   20 FP32 red-black SOR sweeps at omega 1.9 (GMRES-IR),
                   previous-step guess
  mixed         -- the same FP64 refinement and FP32 flexible GMRES(6),
                   every Arnoldi step preconditioned by one FP32 geometric
                   multigrid V-cycle (red-black SOR smoothing, 4+4 sweeps per
                   level at omega 1.55, 5+4 on the levels cycled inside one
                   workgroup, trilinear prolongation),
                   previous-step guess (headline)
  mixed+learned -- mixed, starting from the CNN's guess when its residual
                   is below half the previous-step guess's
The headline `value` is steps/s of --mode; the --compare modes are measured
in the same run and reported beside it, with the Poisson speedup of --mode
over each.  The network is trained first on 1000 recorded solves of the
128x64x64 case (not timed).  The learned guess does not pay with the
multigrid solver (it is kept on no solve and costs its inference), so the
headline is the same solver without it; `learned_guess` in the line reports
the learned form's measurement beside it.  After the timed steps, the
configuration's whole stated run (500 steps, t = 1.0) of the headline mode
and of the learned form is timed on its own (`full_run`).

Roofline: the dominant kernel of mode 4 is the finest multigrid level's
FP32 red-black sweep in one launch (k_sweep2_pz); algorithmic bytes per
launch = 12 B per interior cell (old value and right-hand side read, new
value written, both colours).  For the SOR modes it is the half-sweep:
16 B (FP64) or 8 B (FP32) per interior cell per launch.  The average launch
duration is measured with HIP events on the library's stream
(ib3d_bench_sor).

CPU baseline: the reference's own Fortran (built from its sources with
amdflang -O3, oracle/_ref) timed on this host, one thread, on a bounded
sample: setup, one predictor/forcing/projection pass and a fixed number of
SOR iterations on the same 256x192x192 case; the per-step time is the sum of
the stage times with the SOR part scaled to the FP64 path's measured sweeps
per step.
"""
import argparse
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ib3d-gpu_amd"))

HBM_PEAK_GBS = 8000.0
from ib3d_amd import presets  # noqa: E402

GMRES = dict(presets.GMRES_SOR)
MG = presets.MULTIGRID
SMALL = presets.SMALL


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# The learned guess is kept when its residual is below this fraction of the
# previous-step guess's.  With the multigrid solver a guess that is only
# 10-40% better in residual norm was measured to double the Krylov steps of
# the solve, so a 2x margin is asked for.
LEARNED_ACCEPT = 0.5


class Motion:
    """Prescribed transverse oscillation y_c = a0 sin(2 pi f t): per step the
    body is moved, its geometry rebuilt (selective retagging) and the wall
    velocity and acceleration (finite difference, as the reference's
    disabled form) passed to the forcing and the surface forces."""

    def __init__(self, a0, freq, dt):
        self.a0, self.w, self.dt = a0, 2.0 * math.pi * freq, dt
        self.t, self.v_prev = 0.0, 0.0

    def advance(self, s):
        self.t = self.t + self.dt
        y = self.a0 * math.sin(self.w * self.t)
        v = self.a0 * self.w * math.cos(self.w * self.t)
        s.move_body((0.0, y, 0.0), (0.0, v, 0.0), (0.0, (v - self.v_prev) / self.dt, 0.0))
        self.v_prev = v
        s.first_step = True


def run_mode(solver_cls, case, mode, steps, warmup, omega, tol, inner, net=None, moving=None):
    s = solver_cls.from_case(case)
    motion = Motion(moving[0], moving[1], s.deltat) if moving else None
    from ib3d_amd import cases as _cases
    body = _cases.read_gmsh_bodies(os.path.join(case, "geometries", "sphere.msh"), s.inp["surGeoPoints"])
    multi = body.max() > 1
    if multi:
        s.set_bodies(body)
    if net is not None:
        net.config(record=False, guess=mode == "mixed+learned")
        net.set_accept(LEARNED_ACCEPT if mode == "mixed+learned" else 1.0)
    presets.apply(s, "mixed" if mode == "mixed+learned" else mode, omega=omega, tol=tol, inner=inner, gmres=GMRES)
    for q in range(warmup):
        if motion:
            motion.advance(s)
        r = s.step()
        st = s.poisson_stats()
        log(f"[bench] {mode} warmup {q + 1}/{warmup}: sweeps={st['sweeps']} relres={st['relres']:.2e} "
            f"poisson={st['poisson_ms']:.1f} ms Cd={r['coef'][0] + r['coef'][1]:.4f}")
    s.sync()
    sweeps, sor_ms, poi_ms, cds = [], [], [], []
    t0 = time.perf_counter()
    geo_s = 0.0
    for q in range(steps):
        if motion:
            tg = time.perf_counter()
            motion.advance(s)
            s.sync()
            geo_s += time.perf_counter() - tg
        r = s.step()
        st = s.poisson_stats()
        if mode != "reference" and not st["relres"] <= tol:
            raise SystemExit(f"bench.py: {mode} step {q + 1} ended at relres {st['relres']:.3e} > {tol}")
        sweeps.append(st["sweeps"])
        sor_ms.append(st["sor_ms"])
        poi_ms.append(st["poisson_ms"])
        cds.append(float(r["coef"][0] + r["coef"][1]))
        log(f"[bench] {mode} step {q + 1}/{steps}: sweeps={st['sweeps']} outer={st['outer']} "
            f"relres={st['relres']:.2e} poisson={st['poisson_ms']:.1f} ms")
    s.sync()
    t1 = time.perf_counter()
    if net is not None:
        net.config(False, False)
    out = {
        "mode": mode,
        "s_total": t1 - t0,
        "ms_per_step": 1e3 * (t1 - t0) / steps,
        "steps_per_s": steps / (t1 - t0),
        "sweeps_per_step": sum(sweeps) / steps,
        "sor_ms_per_step": sum(sor_ms) / steps,
        "poisson_ms_per_step": sum(poi_ms) / steps,
        "cd_last": cds[-1],
    }
    if motion:
        out["geometry_ms_per_step"] = 1e3 * geo_s / steps
    if multi:
        out["per_body_coef"] = [[round(float(x), 6) for x in row] for row in s.stress_bodies()]
    if mode == "mixed+learned":
        st = net.stats()
        out["learned_guess"] = {"solves": st["solves"], "kept": st["kept"], "guess_ms": round(st["guess_ms"], 3)}
    return s, out


def full_run(solver_cls, case, mode, nsteps, tol, omega, inner, net=None, moving=None):
    """The headline mode over the configuration's whole stated run (steps
    1..nsteps from the initial condition), timed on its own: whole-run
    steps/s, the settled second half's Poisson ms and Krylov steps, Cd at the
    end and the worst final residual of any solve."""
    s = solver_cls.from_case(case)
    motion = Motion(moving[0], moving[1], s.deltat) if moving else None
    if net is not None:
        net.config(record=False, guess=mode == "mixed+learned")
        net.set_accept(LEARNED_ACCEPT if mode == "mixed+learned" else 1.0)
    base = net.stats() if net is not None else None
    presets.apply(s, "mixed" if mode == "mixed+learned" else mode, omega=omega, tol=tol, inner=inner, gmres=GMRES)
    poi, sw, worst = [], [], 0.0
    s.sync()
    t0 = time.perf_counter()
    for q in range(nsteps):
        if motion:
            motion.advance(s)
        r = s.step()
        st = s.poisson_stats()
        if not st["relres"] <= tol:
            raise SystemExit(f"bench.py: full run {mode} step {q + 1} ended at relres {st['relres']:.3e} > {tol}")
        worst = max(worst, st["relres"])
        poi.append(st["poisson_ms"])
        sw.append(st["sweeps"])
        if (q + 1) % 100 == 0:
            log(f"[bench] full run {mode}: step {q + 1}/{nsteps}, Cd {r['coef'][0] + r['coef'][1]:.5f}")
    s.sync()
    t1 = time.perf_counter()
    half = nsteps // 2
    out = {"mode": mode, "steps": nsteps, "t_end": round(nsteps * s.deltat, 6),
           "steps_per_s": round(nsteps / (t1 - t0), 3), "ms_per_step": round(1e3 * (t1 - t0) / nsteps, 3),
           "poisson_ms_per_step_settled": round(float(np.mean(poi[half:])), 3),
           "sweeps_per_step_settled": round(float(np.mean(sw[half:])), 2),
           "settled_steps": f"{half + 1}-{nsteps}",
           "cd_end": round(float(r["coef"][0] + r["coef"][1]), 6), "max_relres": float(worst)}
    if net is not None:
        if mode == "mixed+learned":
            st = net.stats()
            out["learned_guess"] = {"solves": st["solves"] - base["solves"], "kept": st["kept"] - base["kept"],
                                    "guess_ms": round(st["guess_ms"], 3)}
        net.config(False, False)
    s.close()
    return out


def train_guess(cfg_name, rec_steps, epochs, batch, lr, omega, tol, inner):
    """Record (input, converged correction) pairs from `rec_steps` FP64
    solves of the given case, then train the network on 64^3 tiles of them."""
    from ib3d_amd import cases
    from ib3d_amd.native import Net, Solver
    case = tempfile.mkdtemp(prefix="ib3d_train_")
    try:
        cases.write_case(cases.CONFIGS[cfg_name], case)
        t0 = time.perf_counter()
        s = Solver.from_case(case)
        s.set_poisson_mode(1, omega=omega, tol=tol, check=inner, warm=True)
        net = Net(seed=20251)
        net.config(record=True, guess=False)
        for q in range(rec_steps):
            s.step()
            if (q + 1) % 100 == 0:
                log(f"[bench] recorded {q + 1}/{rec_steps} solves")
        net.config(False, False)
        s.close()
        t1 = time.perf_counter()
        pairs = net.pair_count()
        loss = net.train(epochs=epochs, batch=batch, lr=lr, seed=7)
        t2 = time.perf_counter()
        net.clear_pairs()
        log(f"[bench] trained on {pairs} solves in {t2 - t1:.1f} s, loss {loss[0]:.3e} -> {loss[-1]:.3e}")
        info = {"train_case": cfg_name, "recorded_solves": pairs, "epochs": epochs, "batch": batch, "lr": lr,
                "record_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1),
                "loss_first": float(loss[0]), "loss_last": float(loss[-1])}
        return net, info
    finally:
        shutil.rmtree(case, ignore_errors=True)


def kernel_roofline(s, precision, omega, n_cells, sweeps=200):
    """precision 64 / 32: the half-sweep (two launches per sweep, 16 / 8 B per
    cell each); 33: the FP32 whole sweep in one launch (12 B per cell)."""
    s.bench_sor(precision, 5, omega)  # warm the tables and caches
    ms = s.bench_sor(precision, sweeps, omega)
    if precision == 33:
        launch_s = ms * 1e-3
        bytes_per_launch = 12.0 * n_cells
    else:
        launch_s = 0.5 * ms * 1e-3         # red and black launches per sweep
        bytes_per_launch = 0.5 * (32.0 if precision == 64 else 16.0) * n_cells
    achieved = bytes_per_launch / launch_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "launch_us": round(launch_s * 1e6, 3),
            "bytes_per_launch": bytes_per_launch}


def host_cpu():
    """CPU model and the logical CPUs this process may use (lscpu, else
    /proc/cpuinfo)."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    return {"model": model, "logical_cpus_available": len(os.sched_getaffinity(0))}


def port_sample(case, nsweeps):
    """Stage times of the C restatement's OpenMP build (the oracle, as the
    all-cores CPU baseline) on the case in `case`: one predictor / forcing /
    projection pass and `nsweeps` reference-rule SOR iterations.  Runs in
    the process that calls it; the thread count is OMP_NUM_THREADS's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.use_openmp()
    o = pyoracle.Oracle.from_case(case)
    t = {}

    def timed(name, fn):
        t0 = time.perf_counter()
        fn()
        t[name] = t.get(name, 0.0) + time.perf_counter() - t0
    timed("nsMomentum", lambda: o.call("ns_momentum"))
    timed("velocityBC", lambda: o.call("velocity_bc"))
    timed("solidCellBC", lambda: o.call("solid_cell_bc"))
    timed("velocityForcing1+BC", lambda: (o.call("velocity_forcing"), o.call("velocity_bc")))
    timed("computeDiv", lambda: o.compute_div())
    t0 = time.perf_counter()
    for _ in range(nsweeps):
        o.redblack_sor(1e300)   # one iteration each
    t["sor_per_sweep"] = (time.perf_counter() - t0) / nsweeps
    timed("correct+BC", lambda: (o.call("correct_pressure"), o.call("correct_velocity"), o.call("velocity_bc")))
    timed("pressureForcing1", lambda: o.call("pressure_forcing"))
    timed("stressCal1", lambda: o.stress(True))
    return t


def cpu_baseline_all_cores(case, gpu_sweeps_per_step, nsweeps):
    """The all-cores leg: port_sample in a child process with OMP_NUM_THREADS
    set to the logical CPUs this process may use."""
    threads = len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.perf_counter()
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--port-sample", case, str(nsweeps)],
                         capture_output=True, text=True, env=env, timeout=900, check=True).stdout
    wall = time.perf_counter() - t0
    vals = json.loads(out.strip().splitlines()[-1])
    per_step = (vals["nsMomentum"] + 2 * vals["velocityBC"] + vals["solidCellBC"] + vals["velocityForcing1+BC"] +
                vals["computeDiv"] + gpu_sweeps_per_step * vals["sor_per_sweep"] + vals["correct+BC"] +
                vals["pressureForcing1"] + vals["stressCal1"])
    return {"value": round(1.0 / per_step, 6), "unit": "steps/s", "cores": threads, "kind": "port",
            "extrapolated": True,
            "sample": (f"the C restatement (gcc -O2, OpenMP on its independent loops: predictor, divergence, "
                       f"red-black SOR colours, corrections; {threads} threads) on the same case: one "
                       f"predictor/forcing/projection pass and {nsweeps} SOR iterations "
                       f"({vals['sor_per_sweep'] * 1e3:.1f} ms each), {wall:.1f} s wall with setup; the value is the "
                       f"stage times with the SOR part scaled to {gpu_sweeps_per_step:.0f} sweeps/step = "
                       f"{per_step:.2f} s per step")}


def cpu_baseline(case, gpu_sweeps_per_step, nsweeps, cfg1_steps):
    """The reference's own Fortran (oracle/_ref), one thread: a measured
    whole-step time on the 128x64x64 sphere, and an extrapolated per-step
    time on the benchmark's case from a timed sample of its stages."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    from ib3d_amd import cases
    t0 = time.perf_counter()
    subprocess.run([exe, "sample", str(nsweeps)], cwd=case, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, timeout=900)
    wall = time.perf_counter() - t0
    vals = {}
    with open(os.path.join(case, "harness_sample.txt")) as fh:
        for line in fh:
            k, v = line.split()
            vals[k] = float(v)
    per_step = (vals["nsMomentum"] + 2 * vals["velocityBC"] + vals["solidCellBC"] + vals["velocityForcing1+BC"] +
                vals["computeDiv"] + gpu_sweeps_per_step * vals["sor_per_sweep"] + vals["correct+BC"] +
                vals["pressureForcing1"] + vals["stressCal1"])
    # whole steps of the 128x64x64 sphere, timed step by step
    c1 = tempfile.mkdtemp(prefix="ib3d_cpu_cfg1_")
    measured = None
    try:
        cases.write_case(cases.CONFIGS["cfg1"], c1)
        for sub in ("out", "tag"):
            os.makedirs(os.path.join(c1, sub), exist_ok=True)
        subprocess.run([exe, "time", str(cfg1_steps)], cwd=c1, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=900)
        steps = []
        with open(os.path.join(c1, "harness_time.txt")) as fh:
            for line in fh:
                w = line.split()
                if w[0] == "step":
                    steps.append((int(w[2]), float(w[3])))
        measured = {"workload": "fixed sphere 128x64x64, 5120 triangles, dt=0.005, reference rule",
                    "steps_timed": len(steps), "s_per_step": round(sum(t for _, t in steps) / len(steps), 3),
                    "steps_per_s": round(len(steps) / sum(t for _, t in steps), 5),
                    "nIterPcor": [n for n, _ in steps]}
    finally:
        shutil.rmtree(c1, ignore_errors=True)
    cpu = host_cpu()
    try:
        all_cores = cpu_baseline_all_cores(case, gpu_sweeps_per_step, 4 * nsweeps)
    except (subprocess.SubprocessError, OSError, ValueError, KeyError) as e:
        all_cores = {"error": str(e)[:200]}
    return {"value": round(1.0 / per_step, 6), "unit": "steps/s", "cores": 1, "kind": "reference",
            "extrapolated": True, "cpu_model": cpu["model"], "logical_cpus_available": cpu["logical_cpus_available"],
            "sample": (f"reference Fortran (amdflang -O3, 1 thread) on the same 256x192x192 case: setup, one "
                       f"predictor/forcing/projection pass and {nsweeps} SOR iterations "
                       f"({vals['sor_per_sweep'] * 1e3:.1f} ms each), {wall:.1f} s wall; the value is an "
                       f"extrapolation: the stage times with the SOR part scaled to {gpu_sweeps_per_step:.0f} "
                       f"sweeps/step = {per_step:.2f} s per step"),
            "measured_whole_steps": measured, "all_cores": all_cores}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", default="mixed",
                    choices=["mixed+learned", "mixed", "gmres-sor", "mixed-dc", "fp64", "fp64-cold", "reference"])
    ap.add_argument("--compare", default="fp64-cold,fp64,mixed-dc,gmres-sor,mixed+learned",
                    help="comma-separated modes measured in the same run for comparison ('none' to skip)")
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--omega", type=float, default=1.99)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--inner", type=int, default=20)
    ap.add_argument("--cpu-sweeps", type=int, default=20)
    ap.add_argument("--cpu-cfg1-steps", type=int, default=2, help="whole reference steps timed on 128x64x64")
    ap.add_argument("--full-steps", type=int, default=None,
                    help="steps of the untimed-by-contract whole run (default: the configuration's stated run)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-config", default="cfg1", help="case whose solves are recorded for training")
    ap.add_argument("--rec-steps", type=int, default=1000, help="recorded solves (two 64^3 tiles each)")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--gmres-m", type=int, default=GMRES["m"])
    ap.add_argument("--gmres-k", type=int, default=GMRES["k"])
    ap.add_argument("--gmres-omega", type=float, default=GMRES["omega"])
    ap.add_argument("--port-sample", nargs=2, metavar=("CASE_DIR", "SWEEPS"), help=argparse.SUPPRESS)
    ap.add_argument("--profile-momentum", action="store_true",
                    help="set the case up, run one step, then three nsMomentum calls with the direct kernel and three "
                         "with the LDS-tiled one (a short program for rocprofv3 counter passes)")
    ap.add_argument("--profile-sor", nargs=2, type=int, metavar=("PRECISION", "SWEEPS"),
                    help="profiling aid: set up the case, run SWEEPS RB-SOR sweeps and one copy probe, exit")
    a = ap.parse_args()
    GMRES.update(m=a.gmres_m, k=a.gmres_k, omega=a.gmres_omega)
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              f"(the hot path's multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        raise SystemExit(2)
    if a.port_sample:
        print(json.dumps(port_sample(a.port_sample[0], int(a.port_sample[1]))), flush=True)
        return
    from ib3d_amd import cases
    from ib3d_amd.native import Solver

    cfg = cases.CONFIGS[a.config]
    case = tempfile.mkdtemp(prefix="ib3d_bench_")
    if a.profile_momentum:
        try:
            cases.write_case(cfg, case)
            s = Solver.from_case(case)
            s.step()
            for tiles in (0, 1):
                s.set_momentum_tiles(tiles)
                for _ in range(3):
                    s.nsMomentum()
            s.set_momentum_tiles(0)
            s.sync()
            print(json.dumps({"profile": "momentum", "calls": 3, "variants": ["direct", "lds tiles 64x8"]}), flush=True)
            s.close()
        finally:
            shutil.rmtree(case, ignore_errors=True)
        return
    if a.profile_sor:
        try:
            cases.write_case(cfg, case)
            s = Solver.from_case(case)
            s.step()  # one step so the right-hand side is the physical one
            prec, nsw = a.profile_sor
            ms = s.bench_sor(prec, nsw, a.omega)
            nx, ny, nz = cfg.n
            nbytes = (8 if prec == 64 else 4) * nx * ny * nz
            cms = s.probe_copy(nbytes)
            print(json.dumps({"precision": prec, "sweeps": nsw, "ms_per_sweep": ms, "copy_bytes": nbytes,
                              "copy_ms": cms}), flush=True)
            s.close()
        finally:
            shutil.rmtree(case, ignore_errors=True)
        return
    try:
        meta = cases.write_case(cfg, case)
        n_tris = meta["n_tris"]
        nx, ny, nz = cfg.n
        n_cells = nx * ny * nz
        modes = [a.mode] + [x for x in a.compare.split(",") if x and x != "none" and x != a.mode]
        net, net_info = None, None
        # the network is trained before the timed run only when the headline
        # uses it; otherwise after the headline run and its roofline probe,
        # which then see the device memory as a plain solver run does
        if a.mode == "mixed+learned":
            net, net_info = train_guess(a.train_config, a.rec_steps, a.epochs, 4, 1e-3, a.omega, a.tol, a.inner)
        t0 = time.perf_counter()
        t_train = 0.0
        moving = (cfg.extra["a0"], cfg.extra["freq"]) if cfg.extra.get("a0") else None
        s, main_res = run_mode(Solver, case, a.mode, a.steps, a.warmup, a.omega, a.tol, a.inner, net, moving)
        main_fields = {f: s.download(f) for f in ("u", "p")}
        log(f"[bench] {a.mode}: {main_res}")
        prec = 32 if a.mode.startswith("mixed") or a.mode == "gmres-sor" else 64
        # the dominant kernel: the finest level's whole FP32 sweep in mode 4,
        # the half-sweep otherwise
        rprec = 33 if a.mode.startswith("mixed") else prec
        roof = kernel_roofline(s, rprec, a.omega, n_cells)
        roof32 = kernel_roofline(s, 32, a.omega, n_cells) if rprec != 32 and prec == 32 else None
        roof64 = kernel_roofline(s, 64, a.omega, n_cells) if prec == 32 else roof
        s.close()
        if "mixed+learned" in modes and net is None:
            tt = time.perf_counter()
            net, net_info = train_guess(a.train_config, a.rec_steps, a.epochs, 4, 1e-3, a.omega, a.tol, a.inner)
            t_train = time.perf_counter() - tt
        cmp_all = []
        errors = {}
        for m in modes[1:]:
            s2, r2 = run_mode(Solver, case, m, a.steps, a.warmup, a.omega, a.tol, a.inner, net, moving)
            if m.startswith("fp64"):
                # field difference of the headline mode against this FP64 run after the same steps
                errors["vs_" + m] = {f: float(np.max(np.abs(main_fields[f] - s2.download(f))) /
                                          np.max(np.abs(s2.download(f)))) for f in ("u", "p")}
            s2.close()
            log(f"[bench] {m}: {r2}")
            cmp_all.append(r2)
        # the configuration's whole stated run for the headline mode (and for
        # the same solver without the network when the headline uses it)
        nfull = a.full_steps if a.full_steps is not None else cfg.steps
        full = []
        if nfull > 0:
            extra = ["mixed"] if a.mode == "mixed+learned" else (["mixed+learned"] if "mixed+learned" in modes else [])
            for m in [a.mode] + extra:
                full.append(full_run(Solver, case, m, nfull, a.tol, a.omega, a.inner, net, moving))
                log(f"[bench] full run: {full[-1]}")
        gpu_wall = time.perf_counter() - t0 - t_train
        by_mode = {r["mode"]: r for r in cmp_all + [main_res]}
        fp64_sweeps = by_mode.get("fp64-cold", by_mode.get("fp64", main_res))["sweeps_per_step"]
        cpu = None if a.no_cpu else cpu_baseline(case, fp64_sweeps, a.cpu_sweeps, a.cpu_cfg1_steps)
        traffic = None
        tf = os.path.join(ROOT, "profiles", "sor_traffic.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                t = json.load(fh)
            traffic = t.get({33: "fp32_sweep", 32: "fp32", 64: "fp64"}[rprec])
        roof_out = {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac")}
        roof_out["traffic"] = traffic
        roof_out["kernel"] = ("k_sweep2_pz: FP32 red-black sweep in one launch (12 B per cell)" if rprec == 33 else
                              f"k_sor_pz<{'float' if prec == 32 else 'double'}> half-sweep")
        roof_out["launch_us"] = roof["launch_us"]
        # what a plain streaming copy reaches on this device (604 MB read and
        # written), for reading `frac` against a practical ceiling
        s_cp = Solver.from_case(case)
        cp_ms = min(s_cp.probe_copy(603979776) for _ in range(3))
        s_cp.close()
        roof_out["copy_gbs"] = round(2 * 603979776 / cp_ms / 1e6, 1)
        line = {
            "metric": "steps/s + Poisson ms/step to 1e-6 residual: FP64 SOR vs mixed+learned guess",
            "value": round(main_res["steps_per_s"], 4),
            "unit": "steps/s",
            "n_gpus": 1,
            "steps": a.steps,
            "warmup": a.warmup,
            "ms_per_step": round(main_res["ms_per_step"], 3),
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "fp32+fp64" if prec == 32 else "fp64",
            "data": "synthetic (generated grid and icosphere, reference initial condition)",
            "config": {"workload": (f"sphere oscillating as y_c = {moving[0]} sin(2 pi {moving[1]} t) "
                                    f"(geometry rebuilt every step)" if moving else
                                    f"{cfg.extra['bodies']} ellipsoids (seed {cfg.extra['seed']})"
                                    if cfg.extra.get("bodies") else "fixed sphere")
                                   + f" Re={cfg.extra.get('re_d', 100.0)}, {nx}x{ny}x{nz} stretched grid, "
                                   f"{n_tris} triangles, dt={cfg.dt}",
                       "poisson": (f"{a.mode}: FP64 refinement, FP32 flexible GMRES({MG['m']}) preconditioned by "
                                   f"one FP32 multigrid V-cycle (RB-SOR {MG['nu1']}+{MG['nu2']} sweeps per level at "
                                   f"omega={MG['omega']}"
                                   + (f", {SMALL[0]}+{SMALL[1]} on the levels cycled inside one workgroup"
                                      if SMALL else "")
                                   + f", {MG['coarse']} on the coarsest, trilinear prolongation; a Krylov cycle "
                                   f"hands back to the FP64 check at {MG['target']} x tol)"
                                   if a.mode in ("mixed", "mixed+learned") else
                                   f"{a.mode}: FP64 refinement, FP32 flexible GMRES({GMRES['m']}) preconditioned by "
                                   f"{GMRES['k']} FP32 RB-SOR sweeps at omega={GMRES['omega']}"
                                   if a.mode == "gmres-sor" else
                                   f"{a.mode}: RB-SOR omega={a.omega}, {a.inner} sweeps between residual checks")
                                  + f", ||r||/||b|| <= {a.tol}, "
                                  + (f"CNN guess, kept when its residual is below {LEARNED_ACCEPT} x the previous-step "
                                     f"guess's, else the previous-step guess" if a.mode == "mixed+learned"
                                     else "previous-step guess" if a.mode in ("mixed", "gmres-sor", "mixed-dc", "fp64")
                                     else "zero guess"),
                       "parallelism": "1 GPU"},
            "poisson_ms_per_step": round(main_res["poisson_ms_per_step"], 3),
            "sor_ms_per_step": round(main_res["sor_ms_per_step"], 3),
            "sweeps_per_step": round(main_res["sweeps_per_step"], 1),
            "cd": round(main_res["cd_last"], 5),
            "roofline": roof_out,
            "roofline_fp64_sweep": {k: roof64[k] for k in ("achieved", "frac", "launch_us")},
            **({"roofline_fp32_half_sweep": {k: roof32[k] for k in ("achieved", "frac", "launch_us")}}
               if roof32 else {}),
            "cpu_baseline": cpu,
            "gpu_wall_s": round(gpu_wall, 1),
            "field_error": errors,
        }
        if full:
            line["full_run"] = full
        for k in ("per_body_coef", "geometry_ms_per_step"):
            if k in main_res:
                line[k] = main_res[k]
        if net_info:
            lr = main_res if a.mode == "mixed+learned" else by_mode.get("mixed+learned")
            plain = by_mode.get("mixed")
            if lr is not None:
                net_info.update(lr.get("learned_guess", {}))
                if plain is not None:
                    net_info["poisson_ms_per_step"] = round(lr["poisson_ms_per_step"], 3)
                    net_info["poisson_speedup_vs_mixed"] = round(plain["poisson_ms_per_step"] /
                                                                 lr["poisson_ms_per_step"], 3)
                    net_info["accept_ratio"] = LEARNED_ACCEPT
            line["learned_guess"] = net_info
        if cmp_all:
            line["compare"] = [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()
                                if k != "learned_guess"} for r in cmp_all]
            for r in cmp_all:
                line["poisson_speedup_vs_" + r["mode"]] = round(
                    r["poisson_ms_per_step"] / main_res["poisson_ms_per_step"], 3)
        print(json.dumps(line), flush=True)
    finally:
        shutil.rmtree(case, ignore_errors=True)


if __name__ == "__main__":
    main()
