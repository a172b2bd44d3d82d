"""TEST INFRASTRUCTURE ONLY: ctypes view of the CPU oracle (oracle/mhd_oracle.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  Arrays are exposed as numpy views shaped
(ktot, jtot, nv, itot) -- the reference FieldArray layout, element (n,k,j,i) at
[k, j, n, i] (src_cpp_kokkos_host/mhd.hpp:164-175).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

NPRIM, NCONS, NRIEM = 12, 10, 23


class OrcConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("bc", C.c_int * 6),
        ("xmin", C.c_double * 3), ("xmax", C.c_double * 3),
        ("gam", C.c_double), ("cfl", C.c_double), ("alphabp", C.c_double),
        ("energy_flux", C.c_int), ("grid_shift", C.c_int), ("stages", C.c_int),
    ]


class OrcIcParams(C.Structure):
    _fields_ = [
        ("rho0", C.c_double), ("p0", C.c_double), ("v0", C.c_double), ("a", C.c_double),
        ("bx0", C.c_double), ("amp", C.c_double), ("sigma", C.c_double), ("seed", C.c_uint64),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcConfig)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        for name in ("orc_prim", "orc_cons"):
            getattr(L, name).restype = dp
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_flux.restype = dp
        L.orc_flux.argtypes = [C.c_void_p, C.c_int]
        L.orc_grid.restype = dp
        L.orc_grid.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_generate_problem_kh.argtypes = [C.c_void_p, C.c_double, C.c_double]
        for name in ("orc_generate_grid", "orc_prim_to_cons_all", "orc_set_boundary", "orc_update_primitive",
                     "orc_save_cons", "orc_average_saved"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_generate_problem_shear.argtypes = [C.c_void_p, C.POINTER(OrcIcParams)]
        L.orc_control_timestep.restype = C.c_double
        L.orc_control_timestep.argtypes = [C.c_void_p]
        L.orc_evaluate_ch.restype = C.c_double
        L.orc_evaluate_ch.argtypes = [C.c_void_p]
        L.orc_numerical_flux.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_update_conserv.argtypes = [C.c_void_p, C.c_double]
        L.orc_damp_psi.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_run.restype = C.c_double
        L.orc_run.argtypes = [C.c_void_p, C.c_int, C.c_double, dp]
        L.orc_diagnostics.argtypes = [C.c_void_p, dp]
        L.orc_vanleer.restype = C.c_double
        L.orc_vanleer.argtypes = [C.c_double, C.c_double]
        L.orc_prim2riemann.argtypes = [dp, C.c_int, C.c_int, dp]
        L.orc_hlld.argtypes = [dp, dp, dp]
        L.orc_face_flux.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int, C.c_double, dp]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(nx, ny, nz, bc=(1, 1, 2, 2, 1, 1), xmin=(-0.5, -1.0, -0.5), xmax=(0.5, 1.0, 0.5),
                gam=1.4, cfl=0.05, alphabp=0.1, energy_flux=0, grid_shift=1, stages=1) -> OrcConfig:
    cfg = OrcConfig()
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    for i, v in enumerate(bc):
        cfg.bc[i] = v
    for i in range(3):
        cfg.xmin[i], cfg.xmax[i] = xmin[i], xmax[i]
    cfg.gam, cfg.cfl, cfg.alphabp = gam, cfl, alphabp
    cfg.energy_flux, cfg.grid_shift, cfg.stages = energy_flux, grid_shift, stages
    return cfg


class Oracle:
    """One box of the CPU restatement, with the reference's step functions as methods."""

    def __init__(self, cfg: OrcConfig | None = None, **kw):
        self.cfg = cfg if cfg is not None else make_config(**kw)
        self.L = lib()
        self.h = self.L.orc_create(C.byref(self.cfg))
        it, jt, kt = C.c_int(), C.c_int(), C.c_int()
        self.L.orc_sizes(self.h, C.byref(it), C.byref(jt), C.byref(kt))
        self.itot, self.jtot, self.ktot = it.value, jt.value, kt.value

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.orc_destroy(h)
            self.h = None

    def _view(self, ptr, nv):
        n = self.ktot * self.jtot * nv * self.itot
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.ktot, self.jtot, nv, self.itot)

    @property
    def P(self) -> np.ndarray:
        return self._view(self.L.orc_prim(self.h), NPRIM)

    @property
    def U(self) -> np.ndarray:
        return self._view(self.L.orc_cons(self.h), NCONS)

    def F(self, d: int) -> np.ndarray:
        return self._view(self.L.orc_flux(self.h, d), NCONS)

    def grid(self, axis: int, centre: bool = False) -> np.ndarray:
        n = (self.itot, self.jtot, self.ktot)[axis - 1]
        return np.ctypeslib.as_array(self.L.orc_grid(self.h, axis, int(centre)), shape=(n,))

    # setup
    def generate_grid(self):
        self.L.orc_generate_grid(self.h)

    def generate_problem_kh(self, vy_amp: float = 1.0e-3, rrv: float = 1.0e-2):
        self.L.orc_generate_problem_kh(self.h, vy_amp, rrv)

    def generate_problem_shear(self, rho0=1.0, p0=1.0, v0=0.5, a=0.05, bx0=0.1, amp=1e-2, sigma=0.1,
                               seed=20261203):
        p = OrcIcParams(rho0, p0, v0, a, bx0, amp, sigma, seed)
        self.L.orc_generate_problem_shear(self.h, C.byref(p))

    def prim_to_cons_all(self):
        self.L.orc_prim_to_cons_all(self.h)

    # step functions (reference names in comments)
    def control_timestep(self) -> float:        # ControlTimestep
        return self.L.orc_control_timestep(self.h)

    def set_boundary(self):                      # SetBoundaryCondition
        self.L.orc_set_boundary(self.h)

    def evaluate_ch(self) -> float:              # EvaluateCh
        return self.L.orc_evaluate_ch(self.h)

    def numerical_flux(self, d: int, chg: float):  # GetNumericalFluxD
        self.L.orc_numerical_flux(self.h, d, chg)

    def update_conserv(self, dt: float):         # UpdateConservU
        self.L.orc_update_conserv(self.h, dt)

    def damp_psi(self, dt: float, chg: float):   # DampPsi
        self.L.orc_damp_psi(self.h, dt, chg)

    def update_primitive(self):                  # UpdatePrimitvP
        self.L.orc_update_primitive(self.h)

    # the two-stage Runge-Kutta step's own parts (no reference counterpart)
    def save_cons(self):
        self.L.orc_save_cons(self.h)

    def average_saved(self):
        self.L.orc_average_saved(self.h)

    def run(self, nsteps: int, t0: float = 0.0):
        log = np.zeros(max(nsteps, 1))
        t = self.L.orc_run(self.h, nsteps, t0, _dp(log))
        return t, log[:nsteps]

    def diagnostics(self) -> np.ndarray:
        out = np.zeros(4)
        self.L.orc_diagnostics(self.h, _dp(out))
        return out

    def interior(self, arr: np.ndarray) -> np.ndarray:
        """(nv, nz, ny, nx) copy of the interior cells of a (ktot, jtot, nv, itot) array."""
        s = slice(2, -3)
        return np.ascontiguousarray(arr[2:self.ktot - 3, 2:self.jtot - 3, :, 2:self.itot - 3].transpose(2, 0, 1, 3))


def vanleer(a: float, b: float) -> float:
    return lib().orc_vanleer(a, b)


def prim2riemann(prim: np.ndarray, d: int, energy_flux: int = 0) -> np.ndarray:
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.zeros(NRIEM)
    lib().orc_prim2riemann(_dp(prim), d, energy_flux, _dp(out))
    return out


def hlld(left: np.ndarray, right: np.ndarray) -> np.ndarray:
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    out = np.zeros(NCONS)
    lib().orc_hlld(_dp(left), _dp(right), _dp(out))
    return out


def face_flux(pm2, pm1, p0, pp1, d: int, chg: float, energy_flux: int = 0) -> np.ndarray:
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (pm2, pm1, p0, pp1)]
    out = np.zeros(NCONS)
    lib().orc_face_flux(*[_dp(a) for a in arrs], d, energy_flux, chg, _dp(out))
    return out


def splitmix64(x: int) -> int:
    return lib().orc_splitmix64(x)
