"""ctypes binding of the HIP library (include/ib3d.h).

The library is the product path: there is no CPU fallback.  If
``lib/libib3d_hip.so`` is missing or no GPU is visible, every call raises.
Routine names follow the reference's subroutines (main.f90:5-67).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import cases

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libib3d_hip.so")

_lib = None


class IB3DError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IB3DError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    I, I64, D, P = C.c_int, C.c_int64, C.c_double, C.c_void_p
    PD, PI64, PI32 = C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    PF = C.POINTER(C.c_float)
    sig = {
        "ib3d_init": [I64, I64, I64],
        "ib3d_finalize": [],
        "ib3d_sync": [],
        "ib3d_upload": [C.c_char_p, P, I64],
        "ib3d_download": [C.c_char_p, P, I64],
        "ib3d_set_grid": [PD, PD, PD],
        "ib3d_set_mesh": [I64, PD, PD, PD, I64, PI64, PI64, PI64],
        "ib3d_shiftSurfaceNodesInitial": [D, D, D],
        "ib3d_computeSurfaceNorm": [I64],
        "ib3d_tagging": [PI64, PI32],
        "ib3d_cellCount": [PI32],
        "ib3d_computeNormDistance": [],
        "ib3d_initialConditions": [],
        "ib3d_coefficientMatrix": [],
        "ib3d_non_uni_coeff": [],
        "ib3d_nsMomentum": [D, D],
        "ib3d_velocityBC": [],
        "ib3d_solidCellBC": [],
        "ib3d_velocityForcing1": [],
        "ib3d_pressureForcing1": [],
        "ib3d_computeDiv": [D, PD],
        "ib3d_computeDiv_noalloc": [D, PD],
        "ib3d_REDBLACKSOR": [D, PI64, PD, PD],
        "ib3d_correctPressure": [],
        "ib3d_correctVelocity": [D],
        "ib3d_poissonSolver": [D, D, PI64, PD, PD, PD],
        "ib3d_stressCal1": [D, PD],
        "ib3d_stressCal2": [D, PD],
        "ib3d_set_poisson_mode": [I, D, D, I, I, I64],
        "ib3d_set_krylov": [I],
        "ib3d_set_multigrid": [I, I, D, I, D, I],
        "ib3d_set_multigrid_coarse": [I, I],
        "ib3d_set_multigrid_level": [I, I, I],
        "ib3d_set_multigrid_small": [I, I],
        "ib3d_set_krylov_target": [D],
        "ib3d_set_fused_sweeps": [I, I, I, I],
        "ib3d_set_coarse_tiles": [I],
        "ib3d_set_fused_restrict": [I],
        "ib3d_selfcheck_resid_guard": [I64, I64, I64, I, PI64],
        "ib3d_set_momentum_tiles": [I],
        "ib3d_set_graphs": [I],
        "ib3d_poisson_stats": [PI64, PI64, PD, PD, PD, PD],
        "ib3d_residual_norm": [PD, PD],
        "ib3d_step": [I, D, D, D, D, PD],
        "ib3d_bench_sor": [I, I, D, PD],
        "ib3d_probe_copy": [I64, PD],
        "ib3d_shiftSurfaceNodesMovingbody": [D, D, D, D, D, D],
        "ib3d_selectiveRetagging": [PI32],
        "ib3d_set_wall_motion": [PD, PD],
        "ib3d_set_element_bodies": [PI32, I64, C.c_int32],
        "ib3d_stress_bodies": [PD],
        "ib3d_cnn_init": [C.c_uint64],
        "ib3d_cnn_release": [],
        "ib3d_cnn_set_params": [PF, I64],
        "ib3d_cnn_get_params": [PF, I64],
        "ib3d_cnn_forward": [PF, I64, I64, I64, I64, PF],
        "ib3d_cnn_gradient": [PF, PF, I64, I64, I64, I64, PF],
        "ib3d_cnn_adam": [PF, D, D, D, D],
        "ib3d_cnn_config": [I, I],
        "ib3d_cnn_set_accept": [D],
        "ib3d_cnn_clear_pairs": [],
        "ib3d_cnn_train": [I, I, D, C.c_uint64, PD],
        "ib3d_cnn_stats": [PI64, PI64, PD, PD, PD],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.ib3d_last_error.restype = C.c_char_p
    L.ib3d_last_error.argtypes = []
    L.ib3d_array_count.restype = I64
    L.ib3d_array_count.argtypes = [C.c_char_p]
    for name in ("ib3d_cnn_param_count", "ib3d_cnn_pair_count"):
        getattr(L, name).restype = I64
        getattr(L, name).argtypes = []
    _lib = L
    return L


def _ok(rc):
    if rc != 0:
        raise IB3DError(lib().ib3d_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


INT_ARRAYS = {"cell", "minElemcell", "fluidIndexPtr", "solidIndexPtr", "interceptedIndexPtr",
              "redCellIndexPtr", "blackCellIndexPtr", "nelp", "nelu1", "nelu2", "nelv1", "nelv2",
              "nelw1", "nelw2", "Elemcell", "ucell", "vcell", "wcell", "pcell"}


def resid_guard_selfcheck(nx, ny, nz, drop_interior=False):
    """Reads outside the arrays that the residual kernel's guard admits on an
    nx x ny x nz grid when every slot holds a finite b/dt (host check, no GPU
    needed): 0 with the interior test, more without it."""
    out = C.c_int64()
    _ok(lib().ib3d_selfcheck_resid_guard(int(nx), int(ny), int(nz), int(bool(drop_interior)), C.byref(out)))
    return out.value


class Solver:
    """The device-resident module of one case, driven like main.f90."""

    def __init__(self, nx, ny, nz):
        self.L = lib()
        self.n = (nx, ny, nz)
        _ok(self.L.ib3d_init(nx, ny, nz))
        self.first_step = True

    def close(self):
        self.L.ib3d_finalize()

    # ----------------------------------------------------------- transfers
    def download(self, name):
        cnt = self.L.ib3d_array_count(name.encode())
        if cnt <= 0:
            raise IB3DError(f"array {name} is not available")
        dt = np.int64 if name in INT_ARRAYS else np.float64
        out = np.empty(cnt, dtype=dt)
        _ok(self.L.ib3d_download(name.encode(), out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def field(self, name):
        nx, ny, nz = self.n
        return self.download(name).reshape((nx + 2, ny + 2, nz + 2), order="F")

    def upload(self, name, arr):
        dt = np.int64 if name in INT_ARRAYS else np.float64
        a = np.ascontiguousarray(np.asarray(arr, dtype=dt).reshape(-1, order="F"))
        _ok(self.L.ib3d_upload(name.encode(), a.ctypes.data_as(C.c_void_p), a.nbytes))

    # ----------------------------------------------------------- setup
    @classmethod
    def from_case(cls, directory, window=None):
        """readInput + readSurfaceMeshGmsh (host I/O) followed by the setup
        calls of main.f90:9-18 on the device."""
        inp = cases.read_inputdata(os.path.join(directory, "inputdata"))
        nx, ny, nz = inp["nx"], inp["ny"], inp["nz"]
        s = cls(nx, ny, nz)
        s.inp = inp
        s.re = inp["re"]
        s.rev = 1.0 / inp["re"]
        s.deltat = inp["dt_order"] * 1e-3
        s.eps1 = inp["eps1"]
        gx = cases.read_grid(os.path.join(directory, "gridx.dat"), nx + 1)
        gy = cases.read_grid(os.path.join(directory, "gridy.dat"), ny + 1)
        gz = cases.read_grid(os.path.join(directory, "gridz.dat"), nz + 1)
        _ok(s.L.ib3d_set_grid(_dp(gx), _dp(gy), _dp(gz)))
        nodes, conn = cases.read_gmsh(os.path.join(directory, "geometries", "sphere.msh"), inp["surGeoPoints"])
        xn, yn, zn = (np.ascontiguousarray(nodes[:, q]) for q in range(3))
        p1, p2, p3 = (np.ascontiguousarray(conn[:, q]) for q in range(3))
        _ok(s.L.ib3d_set_mesh(len(nodes), _dp(xn), _dp(yn), _dp(zn), len(conn), _ip(p1), _ip(p2), _ip(p3)))
        _ok(s.L.ib3d_shiftSurfaceNodesInitial(inp["xShift"], inp["yShift"], inp["zShift"]))
        _ok(s.L.ib3d_computeSurfaceNorm(inp["inor"]))
        win = np.array(window or [inp["i_startSearch"], inp["i_endSearch"], inp["j_startSearch"],
                                  inp["j_endSearch"], inp["k_startSearch"], inp["k_endSearch"]], dtype=np.int64)
        cnt3 = np.zeros(3, dtype=np.int32)
        _ok(s.L.ib3d_tagging(_ip(win), cnt3.ctypes.data_as(C.POINTER(C.c_int32))))
        s.tag_counts = cnt3.copy()
        cnt5 = np.zeros(5, dtype=np.int32)
        _ok(s.L.ib3d_cellCount(cnt5.ctypes.data_as(C.POINTER(C.c_int32))))
        s.counts = dict(zip(["ib", "fluid", "solid", "red", "black"], cnt5.tolist()))
        _ok(s.L.ib3d_computeNormDistance())
        _ok(s.L.ib3d_initialConditions())
        _ok(s.L.ib3d_coefficientMatrix())
        _ok(s.L.ib3d_non_uni_coeff())
        return s

    def set_bodies(self, body):
        """Body id (1..n) of every surface element, in mesh order."""
        b = np.ascontiguousarray(body, dtype=np.int32)
        self.nbody = int(b.max())
        _ok(self.L.ib3d_set_element_bodies(b.ctypes.data_as(C.POINTER(C.c_int32)), b.size, self.nbody))

    def stress_bodies(self):
        """Per-body {Cd_visc, Cd_press, Cl_visc, Cl_press} of the last step."""
        out = np.zeros(4 * self.nbody)
        _ok(self.L.ib3d_stress_bodies(_dp(out)))
        return out.reshape(self.nbody, 4)

    def move_body(self, disp, vel=(0.0, 0.0, 0.0), acc=(0.0, 0.0, 0.0)):
        """Prescribed rigid motion: nodes at (initial placement + disp), then
        computeSurfaceNorm, selectiveRetagging, cellCount and
        computeNormDistance; the wall moves with velocity `vel` and
        acceleration `acc` in the forcing, solid cells and surface forces."""
        inp = self.inp
        v = np.array(vel, dtype=np.float64)
        a = np.array(acc, dtype=np.float64)
        _ok(self.L.ib3d_set_wall_motion(_dp(v), _dp(a)))
        _ok(self.L.ib3d_shiftSurfaceNodesMovingbody(inp["xShift"], inp["yShift"], inp["zShift"], float(disp[0]),
                                                    float(disp[1]), float(disp[2])))
        _ok(self.L.ib3d_computeSurfaceNorm(inp["inor"]))
        c3 = np.zeros(3, dtype=np.int32)
        _ok(self.L.ib3d_selectiveRetagging(c3.ctypes.data_as(C.POINTER(C.c_int32))))
        cnt5 = np.zeros(5, dtype=np.int32)
        _ok(self.L.ib3d_cellCount(cnt5.ctypes.data_as(C.POINTER(C.c_int32))))
        self.counts = dict(zip(["ib", "fluid", "solid", "red", "black"], cnt5.tolist()))
        _ok(self.L.ib3d_computeNormDistance())

    # ----------------------------------------------------------- routines
    def nsMomentum(self):
        _ok(self.L.ib3d_nsMomentum(self.rev, self.deltat))

    def velocityBC(self):
        _ok(self.L.ib3d_velocityBC())

    def solidCellBC(self):
        _ok(self.L.ib3d_solidCellBC())

    def velocityForcing1(self):
        _ok(self.L.ib3d_velocityForcing1())

    def pressureForcing1(self):
        _ok(self.L.ib3d_pressureForcing1())

    def poissonSolver(self):
        it = C.c_int64()
        dm, ds, ts = C.c_double(), C.c_double(), C.c_double()
        _ok(self.L.ib3d_poissonSolver(self.deltat, self.eps1, C.byref(it), C.byref(dm), C.byref(ds), C.byref(ts)))
        return it.value, dm.value, ds.value, ts.value

    def computeDiv(self):
        dm = C.c_double()
        _ok(self.L.ib3d_computeDiv(self.deltat, C.byref(dm)))
        return dm.value

    def computeDiv_noalloc(self):
        """computeDiv with no device allocation (static scratch per GPU)."""
        dm = C.c_double()
        _ok(self.L.ib3d_computeDiv_noalloc(self.deltat, C.byref(dm)))
        return dm.value

    def REDBLACKSOR(self, epsi):
        it = C.c_int64()
        d1, d2 = C.c_double(), C.c_double()
        _ok(self.L.ib3d_REDBLACKSOR(epsi, C.byref(it), C.byref(d1), C.byref(d2)))
        return it.value, d2.value

    def stressCal(self, first):
        out = np.zeros(4)
        fn = self.L.ib3d_stressCal1 if first else self.L.ib3d_stressCal2
        _ok(fn(self.re, _dp(out)))
        return out

    def set_poisson_mode(self, mode, omega=1.7, tol=1e-6, check=10, warm=False, max_sweeps=0):
        _ok(self.L.ib3d_set_poisson_mode(int(mode), float(omega), float(tol), int(check), int(bool(warm)),
                                         int(max_sweeps)))

    def set_krylov(self, restart):
        _ok(self.L.ib3d_set_krylov(int(restart)))

    def set_multigrid(self, nu1=4, nu2=4, omega=1.55, coarse_sweeps=4, coarse_omega=1.5, max_levels=0):
        _ok(self.L.ib3d_set_multigrid(int(nu1), int(nu2), float(omega), int(coarse_sweeps), float(coarse_omega),
                                      int(max_levels)))

    def set_fused_sweeps(self, levels=2, planes=0, blocks=480, coarse_blocks=192):
        """Multigrid smoothing of the first `levels` levels with one launch
        per red-black sweep (0: two half-sweep launches everywhere);
        identical results.  True/False mean 1/0."""
        _ok(self.L.ib3d_set_fused_sweeps(int(levels), int(planes), int(blocks), int(coarse_blocks)))

    def set_coarse_tiles(self, from_level=2):
        """Multigrid levels from `from_level` on (0: none) smoothed by two
        temporally blocked launches per phase instead of half-sweep
        launches; identical results.  True/False mean 2/0."""
        if from_level is True:
            from_level = 2
        _ok(self.L.ib3d_set_coarse_tiles(int(from_level)))

    def set_fused_restrict(self, levels=64):
        """Residual and restriction in one z-marching launch on the first
        `levels` multigrid levels (0: none); identical results."""
        _ok(self.L.ib3d_set_fused_restrict(int(levels)))

    def set_momentum_tiles(self, on=True):
        """nsMomentum's interior kernel: 0 straight from HBM (the default), 1-3
        from LDS tiles marching in z, 4 straight from HBM with 64-bit indices;
        identical results."""
        _ok(self.L.ib3d_set_momentum_tiles(int(on)))

    def set_krylov_target(self, target):
        """Early exit of a Krylov cycle at target x tol x ||b/dt|| (default 0.3)."""
        _ok(self.L.ib3d_set_krylov_target(float(target)))

    def set_multigrid_coarse(self, nu1, nu2):
        """Smoothing sweeps on the levels below the finest (after set_multigrid)."""
        _ok(self.L.ib3d_set_multigrid_coarse(int(nu1), int(nu2)))

    def set_multigrid_level(self, level, nu1=-1, nu2=-1):
        """Smoothing sweeps on one multigrid level (1..7) cycled by launches,
        overriding the coarse levels' counts; (-1, -1) removes the override."""
        _ok(self.L.ib3d_set_multigrid_level(int(level), int(nu1), int(nu2)))

    def set_multigrid_small(self, nu1=-1, nu2=-1):
        """Smoothing sweeps on the coarsest levels, those smoothed inside the
        one-workgroup kernel; (-1, -1): the coarse levels' counts."""
        _ok(self.L.ib3d_set_multigrid_small(int(nu1), int(nu2)))

    def set_graphs(self, on=True):
        """Replay each GMRES-IR iteration as a captured HIP graph or launch its
        kernels one by one (the default); the results are identical."""
        _ok(self.L.ib3d_set_graphs(1 if on else 0))

    def poisson_stats(self):
        sw, out = C.c_int64(), C.c_int64()
        rr, r0, ts, tp = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _ok(self.L.ib3d_poisson_stats(C.byref(sw), C.byref(out), C.byref(rr), C.byref(r0), C.byref(ts),
                                      C.byref(tp)))
        return {"sweeps": sw.value, "outer": out.value, "relres": rr.value, "relres0": r0.value,
                "sor_ms": ts.value, "poisson_ms": tp.value}

    def residual_norm(self):
        r, b = C.c_double(), C.c_double()
        _ok(self.L.ib3d_residual_norm(C.byref(r), C.byref(b)))
        return r.value, b.value

    def step(self):
        """One time step (main.f90:26-39 for the first, 41-66 after) on the device."""
        out = np.zeros(8)
        _ok(self.L.ib3d_step(int(self.first_step), self.rev, self.re, self.deltat, self.eps1, _dp(out)))
        self.first_step = False
        return {"niter": int(out[0]), "divmax": out[1], "derr_std": out[2], "coef": out[3:7].copy(),
                "sor_s": out[7]}

    def bench_sor(self, precision, sweeps, omega):
        ms = C.c_double()
        _ok(self.L.ib3d_bench_sor(int(precision), int(sweeps), float(omega), C.byref(ms)))
        return ms.value

    def probe_copy(self, nbytes):
        ms = C.c_double()
        _ok(self.L.ib3d_probe_copy(int(nbytes), C.byref(ms)))
        return ms.value

    def sync(self):
        _ok(self.L.ib3d_sync())


class Net:
    """The learned initial guess held by the library (one per process).

    Parameters are a flat FP32 vector of 15601 values: per layer
    W[tap][cin][cout] then bias[cout], layers 3->16, 16->16, 16->16, 16->1."""

    N_PARAM = 15601

    def __init__(self, seed=1234):
        self.L = lib()
        _ok(self.L.ib3d_cnn_init(int(seed)))

    def release(self):
        _ok(self.L.ib3d_cnn_release())

    def params(self):
        out = np.empty(self.N_PARAM, dtype=np.float32)
        _ok(self.L.ib3d_cnn_get_params(_fp(out), out.size))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        _ok(self.L.ib3d_cnn_set_params(_fp(a), a.size))

    def forward(self, x):
        """x: (nb, vz, vy, vx, 3) -> (nb, vz, vy, vx)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        nb, vz, vy, vx, _ = x.shape
        y = np.empty((nb, vz, vy, vx), dtype=np.float32)
        _ok(self.L.ib3d_cnn_forward(_fp(x), nb, vx, vy, vz, _fp(y)))
        return y

    def gradient(self, x, dy):
        x = np.ascontiguousarray(x, dtype=np.float32)
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        nb, vz, vy, vx, _ = x.shape
        g = np.empty(self.N_PARAM, dtype=np.float32)
        _ok(self.L.ib3d_cnn_gradient(_fp(x), _fp(dy), nb, vx, vy, vz, _fp(g)))
        return g

    def adam(self, grad, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        _ok(self.L.ib3d_cnn_adam(_fp(g), lr, beta1, beta2, eps))

    def config(self, record=False, guess=False):
        _ok(self.L.ib3d_cnn_config(int(bool(record)), int(bool(guess))))

    def set_accept(self, ratio):
        _ok(self.L.ib3d_cnn_set_accept(float(ratio)))

    def pair_count(self):
        return int(self.L.ib3d_cnn_pair_count())

    def clear_pairs(self):
        _ok(self.L.ib3d_cnn_clear_pairs())

    def train(self, epochs=20, batch=4, lr=1e-3, seed=7):
        loss = np.zeros(epochs)
        _ok(self.L.ib3d_cnn_train(int(epochs), int(batch), float(lr), int(seed), _dp(loss)))
        return loss

    def stats(self):
        n, k = C.c_int64(), C.c_int64()
        rp, rg, ms = C.c_double(), C.c_double(), C.c_double()
        _ok(self.L.ib3d_cnn_stats(C.byref(n), C.byref(k), C.byref(rp), C.byref(rg), C.byref(ms)))
        return {"solves": n.value, "kept": k.value, "rel_prev": rp.value, "rel_guess": rg.value,
                "guess_ms": ms.value}
