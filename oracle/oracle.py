"""ctypes wrapper of liboracle.so, the C restatement in mhd2d_oracle.c.

Test infrastructure only: the product library never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
TWIN_PATH = os.path.join(HERE, "libthrust_twin.so")
# the same restatement compiled with FMA contraction (a measurement of what
# contraction alone changes; never the reference arithmetic)
FMA_PATH = os.path.join(HERE, "liboracle_fma.so")
BC_PERIODIC, BC_SYMMETRIC = 0, 1
SCHEME_SINGLE, SCHEME_MULTI_SHARE = 0, 1

_libs = {}


def build():
    """Compiles liboracle.so with the recipe in this directory's Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return LIB_PATH


def load(path=LIB_PATH):
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", HERE, os.path.basename(path)], check=True)
    lib = ctypes.CDLL(path)
    D = ctypes.POINTER(ctypes.c_double)
    P = ctypes.c_void_p
    c_d, c_i = ctypes.c_double, ctypes.c_int
    sig = {
        "oracle_create": (P, [c_i, c_i, c_d, c_d, c_d, c_d, c_d, c_d, c_i, c_i]),
        "oracle_destroy": (None, [P]),
        "oracle_set_state": (None, [P, D]),
        "oracle_get_state": (None, [P, D]),
        "oracle_get_stage1_face_b": (None, [P, D, D]),
        "oracle_face_regions": (None, [P, ctypes.c_void_p, ctypes.c_void_p]),
        "oracle_ct_contact_emf": (c_d, [D, D, D, D, D, D]),
        "oracle_set_eta_field": (c_i, [P, D]),
        "oracle_set_eta_staggered": (c_i, [P, D, D, D]),
        "oracle_apply_boundary": (None, [P]),
        "oracle_apply_boundary_axis": (None, [P, c_i, c_i]),
        "oracle_calculate_dt": (c_d, [P]),
        "oracle_step": (None, [P, c_i]),
        "oracle_get_dt": (c_d, [P]),
        "oracle_set_fixed_dt": (None, [P, c_d]),
        "oracle_get_time": (c_d, [P]),
        "oracle_check_crashed": (c_i, [P]),
        "oracle_get_stage1_fluxes": (None, [P, D, D]),
        "oracle_compute_fluxes": (None, [P, D, D, D]),
        "oracle_minmod": (c_d, [c_d, c_d, c_d]),
        "oracle_hlld_flux": (None, [D, D, c_d, c_d, D]),
        "oracle_cell_dt": (c_d, [D, c_d, c_d, c_d, c_d]),
        "oracle_num_threads": (c_i, []),
        "oracle_set_scheme": (c_i, [P, c_i]),
        "oracle_check_crashed_multi": (c_i, [P]),
        "oracle_apply_halo": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleSim:
    """The reference's ResistiveMHD2D on the CPU (same array shapes as the
    HIP binding: (nx, ny, 8) float64, ghosts included)."""

    def __init__(self, nx, ny, dx, dy, gamma=5.0 / 3.0, cfl=0.7, eps=1e-20, eta=0.0,
                 bc_x=BC_PERIODIC, bc_y=BC_PERIODIC, lib_path=LIB_PATH):
        self._lib = load(lib_path)
        self.nx, self.ny = nx, ny
        self.gamma, self.eps, self.dx, self.dy = gamma, eps, dx, dy
        self._s = self._lib.oracle_create(nx, ny, dx, dy, gamma, cfl, eps, eta, bc_x, bc_y)
        if not self._s:
            raise ValueError("oracle_create failed (nx, ny must be >= 12)")

    @classmethod
    def from_config(cls, cfg, lib_path=LIB_PATH):
        return cls(cfg.nx, cfg.ny, cfg.dx, cfg.dy, cfg.gamma, cfg.cfl, cfg.eps, cfg.eta, cfg.bc_x, cfg.bc_y,
                   lib_path=lib_path)

    def __del__(self):
        if getattr(self, "_s", None):
            self._lib.oracle_destroy(self._s)
            self._s = None

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.nx, self.ny, 8)
        self._lib.oracle_set_state(self._s, _p(U))

    def get_state(self):
        out = np.empty((self.nx, self.ny, 8))
        self._lib.oracle_get_state(self._s, _p(out))
        return out

    def stage1_face_b(self):
        """Face B (bX, bY) of UBar right after the stage-1 CT of the last
        single-GPU step."""
        bx = np.empty((self.nx, self.ny))
        by = np.empty((self.nx, self.ny))
        self._lib.oracle_get_stage1_face_b(self._s, _p(bx), _p(by))
        return bx, by

    def face_regions(self):
        """HLLD region bits of every face of the current state, (rx, ry) as
        (nx, ny) uint8 arrays: rx[i, j] the face (i+1/2, j), ry[i, j] the
        face (i, j+1/2); bits outer-left, single-star left, double-star left,
        outer-right, single-star right, double-star right."""
        rx = np.zeros((self.nx, self.ny), dtype=np.uint8)
        ry = np.zeros((self.nx, self.ny), dtype=np.uint8)
        self._lib.oracle_face_regions(self._s, rx.ctypes.data, ry.ctypes.data)
        return rx, ry

    def set_eta_field(self, eta):
        if eta is None:
            self._lib.oracle_set_eta_field(self._s, None)
            return
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        assert eta.shape == (self.nx, self.ny)
        if self._lib.oracle_set_eta_field(self._s, _p(eta)) != 0:
            raise MemoryError("oracle_set_eta_field")

    def set_eta_staggered(self, xface, yface, cell):
        """The MPI library's getEta at x-faces ((i + 1/2) dx, j dy), y-faces
        (i dx, (j + 1/2) dy) and cells (i dx, j dy), one (nx, ny) plane each;
        all None: the uniform eta."""
        if xface is None and yface is None and cell is None:
            self._lib.oracle_set_eta_staggered(self._s, None, None, None)
            return
        planes = [np.ascontiguousarray(a, dtype=np.float64) for a in (xface, yface, cell)]
        for a in planes:
            assert a.shape == (self.nx, self.ny)
        self._eta_m = planes
        if self._lib.oracle_set_eta_staggered(self._s, *(_p(a) for a in planes)) != 0:
            raise MemoryError("oracle_set_eta_staggered")

    def apply_boundary(self):
        self._lib.oracle_apply_boundary(self._s)

    def apply_boundary_axis(self, axis, kind):
        self._lib.oracle_apply_boundary_axis(self._s, axis, kind)

    def set_fixed_dt(self, dt):
        """dt > 0: every later step uses dt instead of the CFL value."""
        self._lib.oracle_set_fixed_dt(self._s, dt)

    def calculate_dt(self):
        return self._lib.oracle_calculate_dt(self._s)

    def step(self, n=1):
        self._lib.oracle_step(self._s, int(n))

    @property
    def dt(self):
        return self._lib.oracle_get_dt(self._s)

    @property
    def time(self):
        return self._lib.oracle_get_time(self._s)

    def crashed(self):
        if getattr(self, "scheme", SCHEME_SINGLE) == SCHEME_MULTI_SHARE:
            return bool(self._lib.oracle_check_crashed_multi(self._s))
        return bool(self._lib.oracle_check_crashed(self._s))

    def set_scheme(self, scheme):
        """SCHEME_SINGLE (single-GPU library) or SCHEME_MULTI_SHARE (the MPI
        library's step on one rank: periodic halo; eta uniform or from
        set_eta_staggered, not from the per-cell plane)."""
        if self._lib.oracle_set_scheme(self._s, scheme) != 0:
            raise ValueError("the multi-GPU scheme needs periodic x and y and a uniform eta")
        self.scheme = scheme

    def apply_halo(self):
        """The MPI library's sendrecv_U on one rank."""
        self._lib.oracle_apply_halo(self._s)

    def stage1_fluxes(self):
        F = np.empty((self.nx, self.ny, 8))
        G = np.empty((self.nx, self.ny, 8))
        self._lib.oracle_get_stage1_fluxes(self._s, _p(F), _p(G))
        return F, G

    def compute_fluxes(self, U):
        """Fluxes F, G of state U (shift to centre, HLLD + resistive, back)."""
        U = np.array(U, dtype=np.float64, order="C", copy=True)
        F = np.empty((self.nx, self.ny, 8))
        G = np.empty((self.nx, self.ny, 8))
        self._lib.oracle_compute_fluxes(self._s, _p(U), _p(F), _p(G))
        return F, G


def minmod(x, y, eps=1e-20):
    return load().oracle_minmod(float(x), float(y), float(eps))


def hlld_flux(left, right, gamma=5.0 / 3.0, eps=1e-20):
    """One HLLD interface flux; left/right = (rho, u, v, w, bX, bY, bZ, p)."""
    L = np.ascontiguousarray(left, dtype=np.float64)
    R = np.ascontiguousarray(right, dtype=np.float64)
    out = np.empty(8)
    load().oracle_hlld_flux(_p(L), _p(R), gamma, eps, _p(out))
    return out


def ct_contact_emf(NF5, NG4, F5, G4, NF0, NG0):
    """CT-contact corner EMF of one flux set on a 2 x 2 patch of cells
    (i, j), (i, j+1), (i+1, j), (i+1, j+1) (each argument: 4 values)."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (NF5, NG4, F5, G4, NF0, NG0)]
    return load().oracle_ct_contact_emf(*[_p(x) for x in a])


def cell_dt(cell, gamma, dx, dy, eps=1e-20):
    c = np.ascontiguousarray(cell, dtype=np.float64)
    return load().oracle_cell_dt(_p(c), gamma, dx, dy, eps)


def num_threads():
    """OpenMP threads the oracle's loops run on."""
    return load().oracle_num_threads()


_twin = None


def load_twin():
    """libthrust_twin.so: the reference's pass structure on rocThrust's
    OpenMP backend (the benchmark's CPU baseline)."""
    global _twin
    if _twin is not None:
        return _twin
    if not os.path.exists(TWIN_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "libthrust_twin.so"], check=True)
    lib = ctypes.CDLL(TWIN_PATH)
    D = ctypes.POINTER(ctypes.c_double)
    P = ctypes.c_void_p
    c_d, c_i = ctypes.c_double, ctypes.c_int
    sig = {
        "twin_create": (P, [c_i, c_i, c_d, c_d, c_d, c_d, c_d, c_d, c_i, c_i]),
        "twin_destroy": (None, [P]),
        "twin_set_state": (None, [P, D]),
        "twin_get_state": (None, [P, D]),
        "twin_set_eta_field": (None, [P, D]),
        "twin_apply_boundary": (None, [P]),
        "twin_step": (c_i, [P, c_i]),
        "twin_get_dt": (c_d, [P]),
        "twin_get_time": (c_d, [P]),
        "twin_set_threads": (None, [c_i]),
        "twin_max_threads": (c_i, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _twin = lib
    return lib


class TwinSim:
    """The Thrust twin of the reference's single-GPU step (same array shapes
    as OracleSim)."""

    def __init__(self, nx, ny, dx, dy, gamma=5.0 / 3.0, cfl=0.7, eps=1e-20, eta=0.0,
                 bc_x=BC_PERIODIC, bc_y=BC_PERIODIC):
        self._lib = load_twin()
        self.nx, self.ny = nx, ny
        self._s = self._lib.twin_create(nx, ny, dx, dy, gamma, cfl, eps, eta, bc_x, bc_y)
        if not self._s:
            raise MemoryError("twin_create failed")

    @classmethod
    def from_config(cls, cfg):
        return cls(cfg.nx, cfg.ny, cfg.dx, cfg.dy, cfg.gamma, cfg.cfl, cfg.eps, cfg.eta, cfg.bc_x, cfg.bc_y)

    def __del__(self):
        if getattr(self, "_s", None):
            self._lib.twin_destroy(self._s)
            self._s = None

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.nx, self.ny, 8)
        self._lib.twin_set_state(self._s, _p(U))

    def get_state(self):
        out = np.empty((self.nx, self.ny, 8))
        self._lib.twin_get_state(self._s, _p(out))
        return out

    def set_eta_field(self, eta):
        if eta is None:
            self._lib.twin_set_eta_field(self._s, None)
            return
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        assert eta.shape == (self.nx, self.ny)
        self._lib.twin_set_eta_field(self._s, _p(eta))

    def apply_boundary(self):
        self._lib.twin_apply_boundary(self._s)

    def step(self, n=1):
        """n steps of the examples' loop (step, then the NaN check); True if
        a step crashed."""
        return bool(self._lib.twin_step(self._s, int(n)))

    @property
    def dt(self):
        return self._lib.twin_get_dt(self._s)

    @property
    def time(self):
        return self._lib.twin_get_time(self._s)

    @staticmethod
    def set_threads(n):
        load_twin().twin_set_threads(int(n))

    @staticmethod
    def max_threads():
        return load_twin().twin_max_threads()
