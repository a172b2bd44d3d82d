"""Python mirror of the src_cpp_kokkos_host step interface over libmhd_amd.so.

`Box` is one MHD box resident in HBM.  Its methods carry the reference function names
(src_cpp_kokkos_host/mhd.hpp:277-287, boundary.hpp:139-144) and argument meanings;
host arrays use the reference FieldArray layout, numpy shape (ktot, jtot, nv, itot),
element (n,k,j,i) at [k, j, n, i].
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi

NPRIM, NCONS = 12, 10
PERIODIC, REFLECTION, OUTFLOW = 1, 2, 3

# primitive / conserved variable indices (src_cpp_kokkos_host/mhd.hpp:251-268)
nden, nve1, nve2, nve3, nene, npre, ncsp, nbm1, nbm2, nbm3, nbps, nst = range(12)
mden, mrv1, mrv2, mrv3, meto, mbm1, mbm2, mbm3, mbps, mst = range(10)


@dataclass
class Config:
    """Runtime form of src_cpp_kokkos_host/config.hpp plus the constants of mhd.cpp."""
    nx: int = 64
    ny: int = 64
    nz: int = 64
    bc: tuple = (PERIODIC, PERIODIC, REFLECTION, REFLECTION, PERIODIC, PERIODIC)
    xmin: tuple = (-0.5, -1.0, -0.5)
    xmax: tuple = (0.5, 1.0, 0.5)
    gam: float = 1.4
    cfl: float = 0.05
    alphabp: float = 0.1
    energy_flux: int = 0     # 0: |B|^2 B_n (src_cpp_kokkos_host); 1: (v.B) B_n
    grid_shift: int = 1      # 1: C++ variants' face formula; 0: unshifted
    precision: int = 0       # 0: FP64; 1: FP32 reconstruction + HLLD; 2: also the divergence stored in FP32
    device: int = 0
    stages: int = 1          # 1: forward Euler (the reference); 2: two-stage SSP Runge-Kutta

    def to_c(self) -> capi.MhdConfig:
        c = capi.MhdConfig()
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        for i, v in enumerate(self.bc):
            c.bc[i] = v
        for i in range(3):
            c.xmin[i], c.xmax[i] = self.xmin[i], self.xmax[i]
        c.gam, c.cfl, c.alphabp = self.gam, self.cfl, self.alphabp
        c.energy_flux, c.grid_shift, c.precision, c.device = (self.energy_flux, self.grid_shift,
                                                              self.precision, self.device)
        c.stages = self.stages
        return c


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Box:
    """One box of the Kelvin-Helmholtz MHD problem in HBM (AllocateHydroVariables)."""

    def __init__(self, cfg: Config | None = None, **kw):
        self.cfg = cfg if cfg is not None else Config(**kw)
        self.L = capi.load()
        self._c = self.cfg.to_c()
        h = C.c_void_p()
        capi.check(self.L.mhd_create(C.byref(self._c), C.byref(h)))
        self.h = h
        it, jt, kt = C.c_int(), C.c_int(), C.c_int()
        capi.check(self.L.mhd_get_dims(self.h, C.byref(it), C.byref(jt), C.byref(kt)))
        self.itot, self.jtot, self.ktot = it.value, jt.value, kt.value

    def close(self):
        if getattr(self, "h", None):
            self.L.mhd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- data movement
    def _shape(self, nv):
        return (self.ktot, self.jtot, nv, self.itot)

    def device_bytes(self) -> int:
        b = C.c_uint64()
        capi.check(self.L.mhd_device_bytes(self.h, C.byref(b)))
        return b.value

    def set_grid(self, x1a, x1b, x2a, x2b, x3a, x3b):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x1a, x1b, x2a, x2b, x3a, x3b)]
        capi.check(self.L.mhd_set_grid(self.h, *[_dp(a) for a in arrs]))

    def generate_grid(self):
        capi.check(self.L.mhd_generate_grid(self.h))

    def get_grid(self):
        arrs = [np.zeros(n) for n in (self.itot, self.itot, self.jtot, self.jtot, self.ktot, self.ktot)]
        capi.check(self.L.mhd_get_grid(self.h, *[_dp(a) for a in arrs]))
        return arrs

    def upload_prim(self, P: np.ndarray):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == self._shape(NPRIM), P.shape
        capi.check(self.L.mhd_upload_prim(self.h, _dp(P)))

    def upload_cons(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == self._shape(NCONS), U.shape
        capi.check(self.L.mhd_upload_cons(self.h, _dp(U)))

    def download_prim(self, out: np.ndarray = None) -> np.ndarray:
        """P in the host layout; out (C-contiguous float64 of that shape) is filled and returned
        when given, so a loop that downloads often reuses one array, as the reference's P does"""
        P = self._out(out, NPRIM)
        capi.check(self.L.mhd_download_prim(self.h, _dp(P)))
        return P

    def _out(self, out, nv):
        if out is None:
            return np.zeros(self._shape(nv))
        if out.shape != self._shape(nv) or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError(f"out must be C-contiguous float64 of shape {self._shape(nv)}")
        return out

    def snapshot_begin(self):
        """queue a copy of the current primitive state to host memory; later steps overlap it"""
        capi.check(self.L.mhd_snapshot_begin(self.h))

    def snapshot_end(self, copy: bool = True) -> np.ndarray:
        """wait for the snapshot begun last; copy=False returns a view of the library's pinned
        buffer, valid until the next snapshot_begin or close"""
        ptr = C.POINTER(C.c_double)()
        capi.check(self.L.mhd_snapshot_end(self.h, C.byref(ptr)))
        shape = self._shape(NPRIM)
        view = np.ctypeslib.as_array(ptr, shape=(int(np.prod(shape)),)).reshape(shape)
        return view.copy() if copy else view

    def download_cons(self, out: np.ndarray = None) -> np.ndarray:
        U = self._out(out, NCONS)
        capi.check(self.L.mhd_download_cons(self.h, _dp(U)))
        return U

    def download_flux(self, d: int) -> np.ndarray:
        F = np.zeros(self._shape(NCONS))
        capi.check(self.L.mhd_download_flux(self.h, d, _dp(F)))
        return F

    # ---- the reference step functions
    def control_timestep(self) -> float:          # ControlTimestep
        dt = C.c_double()
        capi.check(self.L.mhd_control_timestep(self.h, C.byref(dt)))
        return dt.value

    def set_boundary_condition(self):              # SetBoundaryCondition
        capi.check(self.L.mhd_set_boundary(self.h))

    def evaluate_ch(self) -> float:                # EvaluateCh
        ch = C.c_double()
        capi.check(self.L.mhd_evaluate_ch(self.h, C.byref(ch)))
        return ch.value

    def get_numerical_flux(self, d: int, chg: float):  # GetNumericalFluxD
        capi.check(self.L.mhd_numerical_flux(self.h, d, chg))

    def update_conserv_u(self, dt: float):         # UpdateConservU
        capi.check(self.L.mhd_update_conserv(self.h, dt))

    def damp_psi(self, dt: float, chg: float):     # DampPsi
        capi.check(self.L.mhd_damp_psi(self.h, dt, chg))

    def update_primitiv_p(self):                   # UpdatePrimitvP
        capi.check(self.L.mhd_update_primitive(self.h))

    def step_reference_order(self, time: float) -> tuple[float, float]:
        """One loop body of src_cpp_kokkos_host/main.cpp:229-259, one call per function."""
        dt = self.control_timestep()
        self.set_boundary_condition()
        chg = self.evaluate_ch()
        for d in (1, 2, 3):
            self.get_numerical_flux(d, chg)
        self.update_conserv_u(dt)
        self.damp_psi(dt, chg)
        self.update_primitiv_p()
        return time + dt, dt

    def run(self, nsteps: int, time: float = 0.0, tmax: float = 0.0, mode: int = 0):
        """nsteps loop bodies on the device; returns (time, dt_log, steps_done)."""
        t = C.c_double(time)
        log = np.zeros(max(nsteps, 1))
        done = C.c_int()
        capi.check(self.L.mhd_run(self.h, nsteps, tmax, mode, C.byref(t), _dp(log), C.byref(done)))
        return t.value, log[:done.value].copy(), done.value

    def run_noalloc(self, nsteps: int, time: float = 0.0, tmax: float = 0.0):
        """run() of the fused step through mhd_run_noalloc (no device allocation; scratch in
        the library's static device array); returns (time, dt_log, steps_done)."""
        t = C.c_double(time)
        log = np.zeros(max(nsteps, 1))
        done = C.c_int()
        capi.check(self.L.mhd_run_noalloc(self.h, nsteps, tmax, C.byref(t), _dp(log), C.byref(done)))
        return t.value, log[:done.value].copy(), done.value

    def profile_kernels(self, nsteps: int, mode: int = 0) -> dict:
        """Average ms per launch position of one loop iteration (advances the state)."""
        ms = np.zeros(32)
        n = C.c_int()
        capi.check(self.L.mhd_profile_kernels(self.h, nsteps, mode, _dp(ms), 32, C.byref(n)))
        names = self.L.mhd_kernel_names(mode).decode().split(",")
        return dict(zip(names, ms[:n.value].tolist()))

    def set_tiling(self, sw: int, ly: int, lz: int):
        """x-strip width (16/32/64) and y / z march chunk lengths of the fused kernels"""
        capi.check(self.L.mhd_set_tiling(self.h, sw, ly, lz))

    def set_xy_waves_per_simd(self, n: int):
        """FP64 xy kernel: 2 = up to two wavefronts per SIMD with block-shared strip edges
        (where the box allows it), 1 = one wavefront per SIMD with the strip-edge pre-pass,
        0 = the default rule"""
        capi.check(self.L.mhd_set_xy_waves_per_simd(self.h, n))

    def get_xy_waves_per_simd(self) -> int:
        n = C.c_int()
        capi.check(self.L.mhd_get_xy_waves_per_simd(self.h, C.byref(n)))
        return n.value

    def get_tiling(self) -> tuple[int, int, int]:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        capi.check(self.L.mhd_get_tiling(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def last_run_timing(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int()
        capi.check(self.L.mhd_last_run_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---- problems and diagnostics
    def generate_problem_kh(self, vy_amp: float = 1.0e-3, rrv: float = 1.0e-2):
        capi.check(self.L.mhd_generate_problem_kh(self.h, vy_amp, rrv))

    def generate_problem_shear(self, rho0=1.0, p0=1.0, v0=0.5, a=0.05, bx0=0.1, amp=1e-2, sigma=0.1,
                               seed=20261203):
        p = capi.MhdIcParams(rho0, p0, v0, a, bx0, amp, sigma, seed)
        capi.check(self.L.mhd_generate_problem_shear(self.h, C.byref(p)))

    def diagnostics(self) -> np.ndarray:
        """(volume, mix, sqrt(<vy^2>), max|divB|) of the current primitive state."""
        out = np.zeros(4)
        capi.check(self.L.mhd_diagnostics(self.h, _dp(out)))
        return out

    def status(self) -> int:
        s = C.c_int()
        capi.check(self.L.mhd_status(self.h, C.byref(s)))
        return s.value

    def interior(self, arr: np.ndarray) -> np.ndarray:
        """(nv, nz, ny, nx) copy of the interior cells of a (ktot, jtot, nv, itot) array."""
        return np.ascontiguousarray(arr[2:self.ktot - 3, 2:self.jtot - 3, :, 2:self.itot - 3].transpose(2, 0, 1, 3))


class Batch:
    """Many boxes advanced together (mhd_batch_* of include/mhd_amd.h).

    All boxes share the physics of `cfg` (its nx/ny/nz are ignored); sizes and initial
    states are per box.  Storage is one HBM arena of `arena_bytes`."""

    def __init__(self, cfg: Config, max_boxes: int, arena_bytes: int):
        self.cfg = cfg
        self.L = capi.load()
        self._c = cfg.to_c()
        h = C.c_void_p()
        capi.check(self.L.mhd_batch_create(C.byref(self._c), max_boxes, arena_bytes, C.byref(h)))
        self.h = h
        self.max_boxes = max_boxes
        self.shapes: dict[int, tuple[int, int, int]] = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.mhd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def box_bytes(self, nx: int, ny: int, nz: int) -> int:
        b = C.c_uint64()
        capi.check(self.L.mhd_batch_box_bytes(self.h, nx, ny, nz, C.byref(b)))
        return b.value

    def free_bytes(self) -> tuple[int, int]:
        a, t = C.c_uint64(), C.c_uint64()
        capi.check(self.L.mhd_batch_free_bytes(self.h, C.byref(a), C.byref(t)))
        return a.value, t.value

    def set_tiling(self, nx, ny, nz, sw, ly, lz):
        capi.check(self.L.mhd_batch_set_tiling(self.h, nx, ny, nz, sw, ly, lz))

    def add_box(self, nx: int, ny: int, nz: int, tmax: float = 0.0, ic: dict | None = None,
                generate: bool = True) -> int:
        """returns the box id; raises capi.MhdError (MHD_ERR_NOMEM) when it does not fit"""
        bid = C.c_int()
        p = None
        if generate:
            d = dict(rho0=1.0, p0=1.0, v0=0.5, a=0.05, bx0=0.1, amp=1e-2, sigma=0.1, seed=20261203)
            d.update(ic or {})
            p = capi.MhdIcParams(d["rho0"], d["p0"], d["v0"], d["a"], d["bx0"], d["amp"], d["sigma"], d["seed"])
        capi.check(self.L.mhd_batch_add_box(self.h, nx, ny, nz, C.byref(p) if p is not None else None, tmax,
                                            C.byref(bid)))
        self.shapes[bid.value] = (nx, ny, nz)
        return bid.value

    def try_add_box(self, *a, **kw) -> int | None:
        try:
            return self.add_box(*a, **kw)
        except capi.MhdError as e:
            if "(-3)" in str(e):
                return None
            raise

    def upload_state(self, bid: int, P: np.ndarray, U: np.ndarray):
        P = np.ascontiguousarray(P, dtype=np.float64)
        U = np.ascontiguousarray(U, dtype=np.float64)
        capi.check(self.L.mhd_batch_upload_state(self.h, bid, _dp(P), _dp(U)))

    def remove_box(self, bid: int):
        capi.check(self.L.mhd_batch_remove_box(self.h, bid))
        self.shapes.pop(bid, None)

    def step(self, nsteps: int) -> float:
        """advances every live box nsteps iterations; returns the GPU ms they took"""
        capi.check(self.L.mhd_batch_step(self.h, nsteps))
        ms = C.c_double()
        capi.check(self.L.mhd_batch_last_step_ms(self.h, C.byref(ms)))
        return ms.value

    def box_state(self, bid: int) -> dict:
        t, n, s, st = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        capi.check(self.L.mhd_batch_box_state(self.h, bid, C.byref(t), C.byref(n), C.byref(s), C.byref(st)))
        return {"time": t.value, "steps": n.value, "stopped": bool(s.value), "status": st.value}

    def poll(self) -> list[int]:
        ids = (C.c_int * self.max_boxes)()
        n = C.c_int()
        capi.check(self.L.mhd_batch_poll(self.h, ids, self.max_boxes, C.byref(n)))
        return list(ids[:min(n.value, self.max_boxes)])

    def finish_box(self, bid: int, want_prim: bool = False):
        """(diagnostics[4], primitive array or None) of the box's current state"""
        d = np.zeros(4)
        P = None
        if want_prim:
            nx, ny, nz = self.shapes[bid]
            P = np.zeros((nz + 5, ny + 5, NPRIM, nx + 5))
        capi.check(self.L.mhd_batch_finish_box(self.h, bid, _dp(d), _dp(P) if P is not None else None))
        return d, P
