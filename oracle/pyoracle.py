"""ctypes front end of the C restatement (the oracle).

TEST INFRASTRUCTURE: imported only by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py``, always as the checker.  The product path
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libib3d_oracle.so")
_REF = os.path.join(_HERE, "_ref", "ref_harness")

sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "ib3d-gpu_amd"))
from ib3d_amd import cases  # noqa: E402  (host-side readers of the case files)

_lib = None


def use_openmp():
    """Load the OpenMP build of the restatement (the all-cores CPU baseline)
    instead of the serial one; call before the first use.  Both give the
    same results bit for bit; the thread count is OMP_NUM_THREADS's."""
    global _LIB
    if _lib is not None:
        raise RuntimeError("use_openmp: the oracle library is already loaded")
    _LIB = os.path.join(_HERE, "_build", "libib3d_oracle_omp.so")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"oracle library missing: {_LIB} (run make -C oracle)")
        L = C.CDLL(_LIB)
        P, I64, D = C.c_void_p, C.c_int64, C.c_double
        PD, PI = C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.orc_create.restype = P
        L.orc_create.argtypes = [I64, I64, I64]
        L.orc_free.argtypes = [P]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [P]
        for name in ["orc_compute_surface_norm", "orc_cell_count", "orc_compute_norm_distance",
                     "orc_initial_conditions", "orc_coefficient_matrix", "orc_non_uni_coeff",
                     "orc_ns_momentum", "orc_velocity_bc", "orc_solid_cell_bc",
                     "orc_velocity_forcing", "orc_pressure_forcing", "orc_correct_pressure",
                     "orc_correct_velocity"]:
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = C.c_int
        L.orc_set_params.argtypes = [P, D, D, D, I64]
        L.orc_set_grid.argtypes = [P, PD, PD, PD]
        L.orc_set_mesh.argtypes = [P, I64, PD, I64, PI]
        L.orc_shift_nodes_initial.argtypes = [P, D, D, D]
        L.orc_tagging.argtypes = [P, PI]
        L.orc_compute_div.argtypes = [P, PD]
        L.orc_redblack_sor.argtypes = [P, D, PI, PD]
        L.orc_poisson_solver.argtypes = [P, PI, PD, PD]
        L.orc_stress.argtypes = [P, C.c_int, PD]
        L.orc_step.argtypes = [P, C.c_int, PI, PD, PD, PD]
        L.orc_residual_l2.argtypes = [P, PD, PD]
        L.orc_sor_l2.argtypes = [P, D, D, I64, I64, PI, PD]
        L.orc_array.argtypes = [P, C.c_char_p, C.POINTER(C.c_void_p), PI, C.POINTER(C.c_int)]
        L.orc_counts.argtypes = [P, PI]
        L.orc_shift_nodes_moving.argtypes = [P, D, D, D, D, D, D]
        L.orc_selective_retagging.argtypes = [P]
        L.orc_set_wall_motion.argtypes = [P, PD, PD]
        L.orc_set_bodies.argtypes = [P, C.POINTER(C.c_int32), C.c_int32]
        L.orc_stress_bodies.argtypes = [P, PD]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Oracle:
    """One serial FP64 solver state, driven in the reference's call order."""

    def __init__(self, nx, ny, nz):
        self.L = lib()
        self.h = self.L.orc_create(nx, ny, nz)
        if not self.h:
            raise ValueError("bad grid size")
        self.n = (nx, ny, nz)

    def __del__(self):
        try:
            if self.h:
                self.L.orc_free(self.h)
                self.h = None
        except Exception:
            pass

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error(self.h).decode())

    # ---------------------------------------------------------------- setup
    @classmethod
    def from_case(cls, directory):
        inp = cases.read_inputdata(os.path.join(directory, "inputdata"))
        nx, ny, nz = inp["nx"], inp["ny"], inp["nz"]
        o = cls(nx, ny, nz)
        o.inp = inp
        gx = cases.read_grid(os.path.join(directory, "gridx.dat"), nx + 1)
        gy = cases.read_grid(os.path.join(directory, "gridy.dat"), ny + 1)
        gz = cases.read_grid(os.path.join(directory, "gridz.dat"), nz + 1)
        nodes, conn = cases.read_gmsh(os.path.join(directory, "geometries", "sphere.msh"),
                                      inp["surGeoPoints"])
        o._ok(o.L.orc_set_params(o.h, inp["re"], inp["dt_order"], inp["eps1"], inp["inor"]))
        o._ok(o.L.orc_set_grid(o.h, _dp(gx), _dp(gy), _dp(gz)))
        nodes = np.ascontiguousarray(nodes)
        conn = np.ascontiguousarray(conn)
        o._ok(o.L.orc_set_mesh(o.h, len(nodes), _dp(nodes), len(conn), _ip(conn)))
        o._ok(o.L.orc_shift_nodes_initial(o.h, inp["xShift"], inp["yShift"], inp["zShift"]))
        o._ok(o.L.orc_compute_surface_norm(o.h))
        win = np.array([inp["i_startSearch"], inp["i_endSearch"], inp["j_startSearch"],
                        inp["j_endSearch"], inp["k_startSearch"], inp["k_endSearch"]], dtype=np.int64)
        o._ok(o.L.orc_tagging(o.h, _ip(win)))
        o.cell_tagging = o.array("cell").copy()
        o.minElemcell = o.array("minElemcell").copy()
        o._ok(o.L.orc_cell_count(o.h))
        o._ok(o.L.orc_compute_norm_distance(o.h))
        o._ok(o.L.orc_initial_conditions(o.h))
        o._ok(o.L.orc_coefficient_matrix(o.h))
        o._ok(o.L.orc_non_uni_coeff(o.h))
        return o

    # ---------------------------------------------------------------- access
    def set_bodies(self, body):
        self.body = np.ascontiguousarray(body, dtype=np.int32)
        self.nbody = int(self.body.max())
        self._ok(self.L.orc_set_bodies(self.h, self.body.ctypes.data_as(C.POINTER(C.c_int32)), self.nbody))

    def stress_bodies(self):
        out = np.zeros(4 * self.nbody)
        self._ok(self.L.orc_stress_bodies(self.h, _dp(out)))
        return out.reshape(self.nbody, 4)

    def set_wall_motion(self, vel, acc):
        """Wall velocity and body acceleration used by the forcing, the solid
        cells and the surface forces (zero: the reference's fixed body)."""
        v = np.array(vel, dtype=np.float64)
        a = np.array(acc, dtype=np.float64)
        self._ok(self.L.orc_set_wall_motion(self.h, _dp(v), _dp(a)))

    def move_body(self, disp, vel=(0.0, 0.0, 0.0), acc=(0.0, 0.0, 0.0)):
        """Prescribed rigid motion: nodes at (initial placement + disp), then
        computeSurfaceNorm, selectiveRetagging, cellCount and
        computeNormDistance; the wall moves with velocity `vel` and
        acceleration `acc` in the forcing, solid cells and surface forces."""
        inp = self.inp
        v = np.array(vel, dtype=np.float64)
        a = np.array(acc, dtype=np.float64)
        self._ok(self.L.orc_set_wall_motion(self.h, _dp(v), _dp(a)))
        self._ok(self.L.orc_shift_nodes_moving(self.h, inp["xShift"], inp["yShift"], inp["zShift"],
                                               float(disp[0]), float(disp[1]), float(disp[2])))
        self._ok(self.L.orc_compute_surface_norm(self.h))
        self._ok(self.L.orc_selective_retagging(self.h))
        self._ok(self.L.orc_cell_count(self.h))
        self._ok(self.L.orc_compute_norm_distance(self.h))

    def array(self, name):
        ptr = C.c_void_p()
        cnt = C.c_int64()
        typ = C.c_int()
        self._ok(self.L.orc_array(self.h, name.encode(), C.byref(ptr), C.byref(cnt), C.byref(typ)))
        if cnt.value == 0:
            return np.zeros(0)
        ct = C.c_double if typ.value == 8 else C.c_int64
        buf = (ct * cnt.value).from_address(ptr.value)
        return np.ctypeslib.as_array(buf)

    def field(self, name):
        nx, ny, nz = self.n
        return self.array(name).reshape((nx + 2, ny + 2, nz + 2), order="F")

    def counts(self):
        out = np.zeros(7, dtype=np.int64)
        self.L.orc_counts(self.h, _ip(out))
        return dict(zip(["ib", "fluid", "solid", "red", "black", "ibElems", "ibNodes"], out.tolist()))

    # ---------------------------------------------------------------- steps
    def call(self, name):
        self._ok(getattr(self.L, "orc_" + name)(self.h))

    def poisson(self):
        it = C.c_int64()
        dm = C.c_double()
        ds = C.c_double()
        self._ok(self.L.orc_poisson_solver(self.h, C.byref(it), C.byref(dm), C.byref(ds)))
        return it.value, dm.value, ds.value

    def stress(self, first):
        out = np.zeros(4)
        self._ok(self.L.orc_stress(self.h, int(first), _dp(out)))
        return out

    def step(self, first):
        it = C.c_int64()
        dm = C.c_double()
        ds = C.c_double()
        coef = np.zeros(4)
        self._ok(self.L.orc_step(self.h, int(first), C.byref(it), C.byref(dm), C.byref(ds), _dp(coef)))
        return {"niter": it.value, "divmax": dm.value, "derr_std": ds.value, "coef": coef}

    def compute_div(self):
        dm = C.c_double()
        self._ok(self.L.orc_compute_div(self.h, C.byref(dm)))
        return dm.value

    def redblack_sor(self, epsi):
        it = C.c_int64()
        d2 = C.c_double()
        self._ok(self.L.orc_redblack_sor(self.h, epsi, C.byref(it), C.byref(d2)))
        return it.value, d2.value

    def residual_l2(self):
        r = C.c_double()
        b = C.c_double()
        self._ok(self.L.orc_residual_l2(self.h, C.byref(r), C.byref(b)))
        return r.value, b.value

    def sor_l2(self, omega, tol, max_sweeps, check=1):
        it = C.c_int64()
        rel = C.c_double()
        self._ok(self.L.orc_sor_l2(self.h, omega, tol, max_sweeps, check, C.byref(it), C.byref(rel)))
        return it.value, rel.value


def ref_harness_path():
    """The reference built from its own sources (oracle/_ref), if present."""
    return _REF if os.path.exists(_REF) else None
