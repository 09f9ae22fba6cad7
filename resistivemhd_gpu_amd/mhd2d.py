"""ctypes binding of libmhd2d.so (the C ABI in include/mhd2d.h) and a Python
mirror of the reference's single-GPU class ``ResistiveMHD2D``
(ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY/resistiveMHD_2D.hpp:9-52).

The HIP library is required: loading raises if it is missing or fails to
load; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmhd2d.so")

MHD_BC_PERIODIC = 0
MHD_BC_SYMMETRIC = 1
GHOST = 3
NVAR = 8

# every entry point include/mhd2d.h declares
EXPORTS = (
    "mhd_default_config", "mhd_create", "mhd_destroy", "mhd_last_error", "mhd_get_config",
    "mhd_set_state", "mhd_get_state", "mhd_set_eta_field", "mhd_set_eta", "mhd_apply_boundary",
    "mhd_set_boundary", "mhd_calculate_dt", "mhd_one_step_rk2", "mhd_step_rk2", "mhd_get_dt",
    "mhd_get_time", "mhd_set_time", "mhd_check_crashed", "mhd_save", "mhd_load", "mhd_synchronize",
    "mhd_fill_orszag_tang", "mhd_fill_harris", "mhd_fill_harris_eta", "mhd_init_orszag_tang",
    "mhd_init_harris", "mhd_set_sweep_length", "mhd_get_sweep_length", "mhd_profile_enable",
    "mhd_profile_read", "mhd_set_numerics", "mhd_get_numerics", "mhd_face_regions", "mhd_fill_turbulence",
    "mhd_init_turbulence", "mhd_set_scheme", "mhd_get_scheme",
    "mhd_batch_create", "mhd_batch_destroy", "mhd_batch_last_error", "mhd_batch_size", "mhd_batch_set_state",
    "mhd_batch_get_state", "mhd_batch_set_eta_field", "mhd_batch_apply_boundary", "mhd_batch_init_orszag_tang",
    "mhd_batch_init_harris", "mhd_batch_step", "mhd_batch_synchronize", "mhd_batch_get_dt", "mhd_batch_get_time",
    "mhd_batch_check_crashed", "mhd_batch_set_numerics", "mhd_batch_set_sweep_length",
    "mhd_batch_get_sweep_length", "mhd_batch_set_graph", "mhd_batch_autotune", "mhd_batch_set_waves_per_cu",
    "mhd_batch_get_waves_per_cu", "mhd_batch_save", "mhd_batch_load", "mhd_batch_set_strip_packing",
    "mhd_batch_get_packed_items",
    "mhd_adjoint_create", "mhd_adjoint_destroy", "mhd_adjoint_last_error", "mhd_adjoint_set_eta_field",
    "mhd_adjoint_set_state", "mhd_adjoint_set_dts", "mhd_adjoint_get_dts", "mhd_adjoint_forward",
    "mhd_adjoint_get_stride",
    "mhd_adjoint_get_state", "mhd_adjoint_set_seed", "mhd_adjoint_backward", "mhd_adjoint_get_grad_state",
    "mhd_adjoint_get_grad_eta", "mhd_adjoint_synchronize", "mhd_adjoint_set_direction", "mhd_adjoint_tangent",
    "mhd_adjoint_get_tangent", "mhd_adjoint_set_states", "mhd_adjoint_get_states", "mhd_adjoint_set_seeds",
    "mhd_adjoint_get_grad_states", "mhd_save_subdomain", "mhd_load_subdomain", "mhd_adjoint_get_grad_etas",
    "mhd_adjoint_set_states_device", "mhd_adjoint_get_states_device", "mhd_adjoint_set_seeds_device",
    "mhd_adjoint_get_grad_states_device", "mhd_adjoint_get_grad_etas_device", "mhd_adjoint_set_face_kernel",
    "mhd_adjoint_autotune", "mhd_adjoint_set_degenerate_tol", "mhd_adjoint_replay", "mhd_step_rk2_noalloc",
    "mhd_adjoint_set_eta_fields_device", "mhd_set_eta_staggered",
)

MHD_NUMERICS_FAST = 0
MHD_NUMERICS_EXACT = 1
MHD_SCHEME_SINGLE = 0
MHD_SCHEME_MULTI_SHARE = 1


class MhdConfig(ctypes.Structure):
    """struct mhd_config of include/mhd2d.h."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("xmin", ctypes.c_double), ("ymin", ctypes.c_double),
        ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
        ("eps", ctypes.c_double), ("eta", ctypes.c_double),
        ("bc_x", ctypes.c_int), ("bc_y", ctypes.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class MHDError(RuntimeError):
    pass


_lib = None


def load_library(path=LIB_PATH):
    """Loads libmhd2d.so once and declares argument types."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MHDError(f"{path} is missing: build it with `python -m resistivemhd_gpu_amd.build`")
    try:
        # PyTorch ships its own HIP runtime; loading it first lets the library
        # bind to the same HSA runtime, so torch tensors and this library can
        # share the GPU in one process whichever initialises the device first.
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    C = ctypes.POINTER(MhdConfig)
    I = ctypes.c_int
    sig = {
        "mhd_default_config": (I, [C, I, I, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "mhd_create": (I, [C, ctypes.POINTER(P)]),
        "mhd_destroy": (I, [P]),
        "mhd_last_error": (ctypes.c_char_p, [P]),
        "mhd_get_config": (I, [P, C]),
        "mhd_set_state": (I, [P, D]),
        "mhd_get_state": (I, [P, D]),
        "mhd_set_eta_field": (I, [P, D]),
        "mhd_set_eta_staggered": (I, [P, D, D, D]),
        "mhd_set_eta": (I, [P, ctypes.c_double]),
        "mhd_apply_boundary": (I, [P, I, I]),
        "mhd_set_boundary": (I, [P, I, I]),
        "mhd_calculate_dt": (I, [P, D]),
        "mhd_one_step_rk2": (I, [P]),
        "mhd_step_rk2": (I, [P, I]),
        "mhd_get_dt": (I, [P, D]),
        "mhd_get_time": (I, [P, D]),
        "mhd_set_time": (I, [P, ctypes.c_double]),
        "mhd_check_crashed": (I, [P, ctypes.POINTER(I)]),
        "mhd_save": (I, [P, ctypes.c_char_p, ctypes.c_char_p, I]),
        "mhd_load": (I, [P, ctypes.c_char_p]),
        "mhd_synchronize": (I, [P]),
        "mhd_fill_orszag_tang": (I, [C, D]),
        "mhd_fill_harris": (I, [C, ctypes.c_ulonglong, ctypes.c_double, D]),
        "mhd_fill_harris_eta": (I, [C, ctypes.c_double, ctypes.c_double, D]),
        "mhd_init_orszag_tang": (I, [P]),
        "mhd_init_harris": (I, [P, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "mhd_set_sweep_length": (I, [P, I]),
        "mhd_get_sweep_length": (I, [P, ctypes.POINTER(I)]),
        "mhd_profile_enable": (I, [P, I]),
        "mhd_profile_read": (I, [P, D, D, ctypes.POINTER(I)]),
        "mhd_set_numerics": (I, [P, I]),
        "mhd_get_numerics": (I, [P, ctypes.POINTER(I)]),
        "mhd_face_regions": (I, [P, I, P, P]),
        "mhd_fill_turbulence": (I, [C, ctypes.c_ulonglong, I, ctypes.c_double, ctypes.c_double, D]),
        "mhd_init_turbulence": (I, [P, ctypes.c_ulonglong, I, ctypes.c_double, ctypes.c_double]),
        "mhd_set_scheme": (I, [P, I]),
        "mhd_get_scheme": (I, [P, ctypes.POINTER(I)]),
        "mhd_batch_create": (I, [C, I, ctypes.POINTER(P)]),
        "mhd_batch_destroy": (I, [P]),
        "mhd_batch_last_error": (ctypes.c_char_p, [P]),
        "mhd_batch_size": (I, [P, ctypes.POINTER(I)]),
        "mhd_batch_set_state": (I, [P, I, D]),
        "mhd_batch_get_state": (I, [P, I, D]),
        "mhd_batch_save": (I, [P, I, ctypes.c_char_p, ctypes.c_char_p, I]),
        "mhd_batch_load": (I, [P, I, ctypes.c_char_p]),
        "mhd_batch_set_eta_field": (I, [P, I, D]),
        "mhd_batch_apply_boundary": (I, [P]),
        "mhd_batch_init_orszag_tang": (I, [P, I]),
        "mhd_batch_init_harris": (I, [P, I, ctypes.c_ulonglong, ctypes.c_double]),
        "mhd_batch_step": (I, [P, I]),
        "mhd_batch_synchronize": (I, [P]),
        "mhd_batch_get_dt": (I, [P, I, D]),
        "mhd_batch_get_time": (I, [P, I, D]),
        "mhd_batch_check_crashed": (I, [P, I, ctypes.POINTER(I)]),
        "mhd_batch_set_numerics": (I, [P, I]),
        "mhd_batch_set_sweep_length": (I, [P, I]),
        "mhd_batch_get_sweep_length": (I, [P, ctypes.POINTER(I)]),
        "mhd_batch_set_graph": (I, [P, I]),
        "mhd_batch_autotune": (I, [P, I, ctypes.POINTER(I)]),
        "mhd_batch_set_waves_per_cu": (I, [P, I]),
        "mhd_batch_get_waves_per_cu": (I, [P, ctypes.POINTER(I)]),
        "mhd_batch_set_strip_packing": (I, [P, I]),
        "mhd_batch_get_packed_items": (I, [P, ctypes.POINTER(I)]),
        "mhd_adjoint_create": (I, [C, I, ctypes.POINTER(P)]),
        "mhd_adjoint_destroy": (I, [P]),
        "mhd_adjoint_last_error": (ctypes.c_char_p, [P]),
        "mhd_adjoint_set_eta_field": (I, [P, I, D]),
        "mhd_adjoint_set_state": (I, [P, I, D]),
        "mhd_adjoint_set_dts": (I, [P, I, D, I]),
        "mhd_adjoint_get_dts": (I, [P, I, D, I]),
        "mhd_adjoint_forward": (I, [P, I, I, I]),
        "mhd_adjoint_get_stride": (I, [P, ctypes.POINTER(I)]),
        "mhd_adjoint_get_state": (I, [P, I, D]),
        "mhd_adjoint_set_seed": (I, [P, I, D]),
        "mhd_adjoint_backward": (I, [P]),
        "mhd_adjoint_get_grad_state": (I, [P, I, D]),
        "mhd_adjoint_get_grad_eta": (I, [P, I, D]),
        "mhd_adjoint_synchronize": (I, [P]),
        "mhd_adjoint_set_direction": (I, [P, I, D, D]),
        "mhd_adjoint_tangent": (I, [P, I]),
        "mhd_adjoint_replay": (I, [P, I]),
        "mhd_step_rk2_noalloc": (I, [C, P, ctypes.c_longlong, ctypes.c_longlong, P, I, P, P]),
        "mhd_adjoint_set_eta_fields_device": (I, [P, P]),
        "mhd_adjoint_get_tangent": (I, [P, I, D]),
        "mhd_adjoint_set_states": (I, [P, D]),
        "mhd_adjoint_get_states": (I, [P, D]),
        "mhd_adjoint_set_seeds": (I, [P, D]),
        "mhd_adjoint_get_grad_states": (I, [P, D]),
        "mhd_save_subdomain": (I, [P, ctypes.c_char_p, ctypes.c_char_p, I, I]),
        "mhd_load_subdomain": (I, [P, ctypes.c_char_p]),
        "mhd_adjoint_get_grad_etas": (I, [P, D]),
        "mhd_adjoint_set_states_device": (I, [P, P]),
        "mhd_adjoint_get_states_device": (I, [P, P]),
        "mhd_adjoint_set_seeds_device": (I, [P, P]),
        "mhd_adjoint_get_grad_states_device": (I, [P, P]),
        "mhd_adjoint_get_grad_etas_device": (I, [P, P]),
        "mhd_adjoint_set_face_kernel": (I, [P, I]),
        "mhd_adjoint_autotune": (I, [P, I, ctypes.POINTER(I)]),
        "mhd_adjoint_set_degenerate_tol": (I, [P, ctypes.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc, handle=None):
    if rc != 0:
        msg = load_library().mhd_last_error(handle)
        raise MHDError(f"libmhd2d error {rc}: {msg.decode() if msg else ''}")


def default_config(nx_int, ny_int, xmin, lx, ymin, ly, **overrides):
    """Reference example defaults (gamma 5/3, CFL 0.7, EPS 1e-20) for an
    interior of nx_int x ny_int cells; nx, ny of the result include ghosts."""
    cfg = MhdConfig()
    _check(load_library().mhd_default_config(ctypes.byref(cfg), nx_int, ny_int, xmin, lx, ymin, ly))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


def orszag_tang_config(n=256, eta=1e-3):
    """Orszag-Tang vortex on [0, 2pi]^2, periodic x and y."""
    return default_config(n, n, 0.0, 2.0 * np.pi, 0.0, 2.0 * np.pi, eta=eta)


def harris_config(nx=4096, ny=2048, eta_max=6e-3, ly=100.0):
    """Harris sheet on [-100, 100] x [-ly/2, ly/2] (default [-50, 50]),
    periodic x, symmetric y.  eta_max is the dt diffusion limit (the peak of
    the resistivity field)."""
    return default_config(nx, ny, -100.0, 200.0, -0.5 * ly, ly, eta=eta_max,
                          bc_x=MHD_BC_PERIODIC, bc_y=MHD_BC_SYMMETRIC)


def turbulence_config(n=8192, eta=1e-4):
    """Decaying turbulence on [0, 2pi]^2, periodic x and y (one GPU's
    subdomain of the MPI library's decomposition)."""
    return default_config(n, n, 0.0, 2.0 * np.pi, 0.0, 2.0 * np.pi, eta=eta)


def fill_orszag_tang(cfg):
    out = np.empty((cfg.nx, cfg.ny, NVAR), dtype=np.float64)
    _check(load_library().mhd_fill_orszag_tang(ctypes.byref(cfg), _dptr(out)))
    return out


def fill_harris(cfg, seed=0, noise_amp=1e-3):
    out = np.empty((cfg.nx, cfg.ny, NVAR), dtype=np.float64)
    _check(load_library().mhd_fill_harris(ctypes.byref(cfg), seed, noise_amp, _dptr(out)))
    return out


def fill_turbulence(cfg, seed=0, kmax=8, rms_v=0.5, rms_b=0.5):
    out = np.empty((cfg.nx, cfg.ny, NVAR), dtype=np.float64)
    _check(load_library().mhd_fill_turbulence(ctypes.byref(cfg), seed, kmax, rms_v, rms_b, _dptr(out)))
    return out


def fill_harris_eta(cfg, eta0=1e-3, eta1=5e-3):
    out = np.empty((cfg.nx, cfg.ny), dtype=np.float64)
    _check(load_library().mhd_fill_harris_eta(ctypes.byref(cfg), eta0, eta1, _dptr(out)))
    return out


class ResistiveMHD2D:
    """Mirror of the reference's ``class ResistiveMHD2D`` over the C ABI.

    State arrays are numpy float64 of shape (nx, ny, 8) in the reference's
    ConservationParameter order (rho, rhoU, rhoV, rhoW, bX, bY, bZ, e), which
    is the reference's linear layout k = j + i*ny with 8 doubles per cell.
    """

    def __init__(self, cfg):
        self._lib = load_library()
        self.cfg = MhdConfig.from_buffer_copy(cfg)
        h = ctypes.c_void_p()
        _check(self._lib.mhd_create(ctypes.byref(self.cfg), ctypes.byref(h)))
        self._h = h

    # -- lifetime -------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.mhd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _ok(self, rc):
        _check(rc, self._h)

    @property
    def shape(self):
        return (self.cfg.nx, self.cfg.ny, NVAR)

    # -- reference API --------------------------------------------------
    def initializeU(self, U, apply_boundary=True):
        """Uploads a full state (ghosts included) and, as the examples'
        initializeU does, applies the configured boundary."""
        self.setU(U)
        if apply_boundary:
            self.applyBoundary()

    def setU(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.shape:
            raise ValueError(f"state shape {U.shape} != {self.shape}")
        self._ok(self._lib.mhd_set_state(self._h, _dptr(U)))

    def getU(self):
        out = np.empty(self.shape, dtype=np.float64)
        self._ok(self._lib.mhd_get_state(self._h, _dptr(out)))
        return out

    def oneStepRK2(self):
        self._ok(self._lib.mhd_one_step_rk2(self._h))

    def steps(self, n):
        """n RK2 steps with no host synchronisation in between."""
        self._ok(self._lib.mhd_step_rk2(self._h, int(n)))

    def calculateDt(self):
        dt = ctypes.c_double()
        self._ok(self._lib.mhd_calculate_dt(self._h, ctypes.byref(dt)))
        return dt.value

    def checkCalculationIsCrashed(self):
        c = ctypes.c_int()
        self._ok(self._lib.mhd_check_crashed(self._h, ctypes.byref(c)))
        return bool(c.value)

    def save(self, directoryname, filenameWithoutStep, step):
        self._ok(self._lib.mhd_save(self._h, os.fsencode(directoryname), os.fsencode(filenameWithoutStep),
                                    int(step)))

    def saveSubdomain(self, directoryname, filenameWithoutStep, step, rank=0):
        """The MPI library's per-rank save: interior cells only,
        {dir}/{name}_{step}_{rank}.bin."""
        self._ok(self._lib.mhd_save_subdomain(self._h, directoryname.encode(), filenameWithoutStep.encode(),
                                              int(step), int(rank)))

    def loadSubdomain(self, path):
        self._ok(self._lib.mhd_load_subdomain(self._h, str(path).encode()))

    def load(self, path):
        self._ok(self._lib.mhd_load(self._h, os.fsencode(path)))

    # Boundary::periodicBoundaryX2nd etc. (P/boundary.hpp, S/boundary.hpp)
    def periodicBoundaryX2nd(self):
        self._ok(self._lib.mhd_apply_boundary(self._h, 0, MHD_BC_PERIODIC))

    def periodicBoundaryY2nd(self):
        self._ok(self._lib.mhd_apply_boundary(self._h, 1, MHD_BC_PERIODIC))

    def symmetricBoundaryX2nd(self):
        self._ok(self._lib.mhd_apply_boundary(self._h, 0, MHD_BC_SYMMETRIC))

    def symmetricBoundaryY2nd(self):
        self._ok(self._lib.mhd_apply_boundary(self._h, 1, MHD_BC_SYMMETRIC))

    def applyBoundary(self):
        """The configured boundary, x first then y."""
        self._ok(self._lib.mhd_apply_boundary(self._h, 0, self.cfg.bc_x))
        self._ok(self._lib.mhd_apply_boundary(self._h, 1, self.cfg.bc_y))

    def setBoundary(self, bc_x, bc_y):
        self._ok(self._lib.mhd_set_boundary(self._h, bc_x, bc_y))
        self.cfg.bc_x, self.cfg.bc_y = bc_x, bc_y

    # -- resistivity, time ----------------------------------------------
    def setEtaField(self, eta):
        if eta is None:
            self._ok(self._lib.mhd_set_eta_field(self._h, None))
            return
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        if eta.shape != self.shape[:2]:
            raise ValueError(f"eta shape {eta.shape} != {self.shape[:2]}")
        self._ok(self._lib.mhd_set_eta_field(self._h, _dptr(eta)))

    def setEtaStaggered(self, xface, yface, cell):
        """MHD_SCHEME_MULTI_SHARE: the MPI library's getEta sampled at
        x-faces ((i + 1/2) dx, j dy), y-faces (i dx, (j + 1/2) dy) and cells
        (i dx, j dy), one (nx, ny) plane each, ghost entries equal to their
        periodic images; all None: the uniform eta."""
        if xface is None and yface is None and cell is None:
            self._ok(self._lib.mhd_set_eta_staggered(self._h, None, None, None))
            return
        planes = [np.ascontiguousarray(a, dtype=np.float64) for a in (xface, yface, cell)]
        for a in planes:
            if a.shape != self.shape[:2]:
                raise ValueError(f"eta shape {a.shape} != {self.shape[:2]}")
        self._ok(self._lib.mhd_set_eta_staggered(self._h, *(_dptr(a) for a in planes)))

    def setEta(self, eta):
        self._ok(self._lib.mhd_set_eta(self._h, float(eta)))
        self.cfg.eta = float(eta)

    @property
    def dt(self):
        v = ctypes.c_double()
        self._ok(self._lib.mhd_get_dt(self._h, ctypes.byref(v)))
        return v.value

    @property
    def totalTime(self):
        v = ctypes.c_double()
        self._ok(self._lib.mhd_get_time(self._h, ctypes.byref(v)))
        return v.value

    @totalTime.setter
    def totalTime(self, t):
        self._ok(self._lib.mhd_set_time(self._h, float(t)))

    def synchronize(self):
        self._ok(self._lib.mhd_synchronize(self._h))

    def set_numerics(self, mode):
        """MHD_NUMERICS_EXACT (default) or MHD_NUMERICS_FAST ("exact"/"fast")."""
        if isinstance(mode, str):
            mode = {"fast": MHD_NUMERICS_FAST, "exact": MHD_NUMERICS_EXACT}[mode]
        self._ok(self._lib.mhd_set_numerics(self._h, int(mode)))

    def numerics(self):
        v = ctypes.c_int()
        self._ok(self._lib.mhd_get_numerics(self._h, ctypes.byref(v)))
        return "exact" if v.value == MHD_NUMERICS_EXACT else "fast"

    def face_regions(self, mode):
        """HLLD region bits of every face of the current state under the
        arithmetic `mode` ("exact"/"fast"): (rx, ry) as (nx, ny) uint8 arrays,
        rx[i, j] the face (i+1/2, j), ry[i, j] the face (i, j+1/2); 0xFF where
        no update reads the face."""
        if isinstance(mode, str):
            mode = {"fast": MHD_NUMERICS_FAST, "exact": MHD_NUMERICS_EXACT}[mode]
        rx = np.empty((self.cfg.nx, self.cfg.ny), dtype=np.uint8)
        ry = np.empty((self.cfg.nx, self.cfg.ny), dtype=np.uint8)
        self._ok(self._lib.mhd_face_regions(self._h, int(mode), rx.ctypes.data, ry.ctypes.data))
        return rx, ry

    def set_sweep_length(self, columns):
        self._ok(self._lib.mhd_set_sweep_length(self._h, int(columns)))

    def sweep_length(self):
        v = ctypes.c_int()
        self._ok(self._lib.mhd_get_sweep_length(self._h, ctypes.byref(v)))
        return v.value

    def profile_enable(self, on=True):
        self._ok(self._lib.mhd_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        """(ms of stage-1 kernels, ms of stage-2 kernels, steps timed)."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._ok(self._lib.mhd_profile_read(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    # -- setups ---------------------------------------------------------
    def initOrszagTang(self):
        self._ok(self._lib.mhd_init_orszag_tang(self._h))

    def initHarris(self, seed=0, noise_amp=1e-3, eta0=1e-3, eta1=5e-3):
        self._ok(self._lib.mhd_init_harris(self._h, seed, noise_amp, eta0, eta1))

    def initTurbulence(self, seed=0, kmax=8, rms_v=0.5, rms_b=0.5):
        self._ok(self._lib.mhd_init_turbulence(self._h, seed, kmax, rms_v, rms_b))

    def set_scheme(self, scheme):
        """MHD_SCHEME_SINGLE (default) or MHD_SCHEME_MULTI_SHARE ("single"/"multi")."""
        if isinstance(scheme, str):
            scheme = {"single": MHD_SCHEME_SINGLE, "multi": MHD_SCHEME_MULTI_SHARE}[scheme]
        self._ok(self._lib.mhd_set_scheme(self._h, int(scheme)))

    def scheme(self):
        v = ctypes.c_int()
        self._ok(self._lib.mhd_get_scheme(self._h, ctypes.byref(v)))
        return "multi" if v.value == MHD_SCHEME_MULTI_SHARE else "single"


class ResistiveMHD2DBatch:
    """Many independent runs advanced together (mhd_batch_* of the C ABI):
    one launch per RK stage over a work list of every run's items, steps
    replayed from a hipGraph.  Each run has its own configuration, state,
    resistivity and dt; per run the result is bitwise that of
    ResistiveMHD2D with the same numerics."""

    def __init__(self, cfgs):
        self._lib = load_library()
        self.cfgs = [MhdConfig.from_buffer_copy(c) for c in cfgs]
        arr = (MhdConfig * len(self.cfgs))(*self.cfgs)
        h = ctypes.c_void_p()
        rc = self._lib.mhd_batch_create(arr, len(self.cfgs), ctypes.byref(h))
        if rc != 0:
            msg = self._lib.mhd_batch_last_error(None)
            raise MHDError(f"libmhd2d error {rc}: {msg.decode() if msg else ''}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mhd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.cfgs)

    def _ok(self, rc):
        if rc != 0:
            msg = self._lib.mhd_batch_last_error(self._h)
            raise MHDError(f"libmhd2d error {rc}: {msg.decode() if msg else ''}")

    def shape(self, run):
        c = self.cfgs[run]
        return (c.nx, c.ny, NVAR)

    def setU(self, run, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self.shape(run):
            raise ValueError(f"state shape {U.shape} != {self.shape(run)}")
        self._ok(self._lib.mhd_batch_set_state(self._h, run, _dptr(U)))

    def getU(self, run):
        out = np.empty(self.shape(run), dtype=np.float64)
        self._ok(self._lib.mhd_batch_get_state(self._h, run, _dptr(out)))
        return out

    def save(self, run, directoryname, filenameWithoutStep, step):
        """ResistiveMHD2D.save of one run: {dir}/{name}_{step}.bin."""
        self._ok(self._lib.mhd_batch_save(self._h, run, os.fsencode(directoryname),
                                          os.fsencode(filenameWithoutStep), int(step)))

    def load(self, run, path):
        self._ok(self._lib.mhd_batch_load(self._h, run, os.fsencode(path)))

    def setEtaField(self, run, eta):
        if eta is None:
            self._ok(self._lib.mhd_batch_set_eta_field(self._h, run, None))
            return
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        c = self.cfgs[run]
        if eta.shape != (c.nx, c.ny):
            raise ValueError(f"eta shape {eta.shape} != {(c.nx, c.ny)}")
        self._ok(self._lib.mhd_batch_set_eta_field(self._h, run, _dptr(eta)))

    def applyBoundary(self):
        self._ok(self._lib.mhd_batch_apply_boundary(self._h))

    def initOrszagTang(self, run):
        self._ok(self._lib.mhd_batch_init_orszag_tang(self._h, run))

    def initHarris(self, run, seed=0, noise_amp=1e-3):
        self._ok(self._lib.mhd_batch_init_harris(self._h, run, seed, noise_amp))

    def steps(self, n):
        self._ok(self._lib.mhd_batch_step(self._h, int(n)))

    def synchronize(self):
        self._ok(self._lib.mhd_batch_synchronize(self._h))

    def dt(self, run):
        v = ctypes.c_double()
        self._ok(self._lib.mhd_batch_get_dt(self._h, run, ctypes.byref(v)))
        return v.value

    def totalTime(self, run):
        v = ctypes.c_double()
        self._ok(self._lib.mhd_batch_get_time(self._h, run, ctypes.byref(v)))
        return v.value

    def checkCalculationIsCrashed(self, run):
        c = ctypes.c_int()
        self._ok(self._lib.mhd_batch_check_crashed(self._h, run, ctypes.byref(c)))
        return bool(c.value)

    def set_numerics(self, mode):
        if isinstance(mode, str):
            mode = {"fast": MHD_NUMERICS_FAST, "exact": MHD_NUMERICS_EXACT}[mode]
        self._ok(self._lib.mhd_batch_set_numerics(self._h, int(mode)))

    def set_sweep_length(self, columns):
        self._ok(self._lib.mhd_batch_set_sweep_length(self._h, int(columns)))

    def sweep_length(self):
        v = ctypes.c_int()
        self._ok(self._lib.mhd_batch_get_sweep_length(self._h, ctypes.byref(v)))
        return v.value

    def set_graph(self, on=True):
        self._ok(self._lib.mhd_batch_set_graph(self._h, 1 if on else 0))

    def set_waves_per_cu(self, waves):
        """Launch shape of the stage kernels: 0 = one wavefront per work
        item; w > 0 = w wavefronts per CU looping over the work list."""
        self._ok(self._lib.mhd_batch_set_waves_per_cu(self._h, int(waves)))

    def waves_per_cu(self):
        v = ctypes.c_int()
        self._ok(self._lib.mhd_batch_get_waves_per_cu(self._h, ctypes.byref(v)))
        return v.value

    STRIP_PACK_AUTO, STRIP_PACK_ALWAYS, STRIP_PACK_NEVER = 0, 1, 2

    def set_strip_packing(self, mode):
        """Packing of short last strips (several column segments per
        wavefront): STRIP_PACK_AUTO (when the launch takes several rounds),
        _ALWAYS or _NEVER; results are bitwise the same."""
        self._ok(self._lib.mhd_batch_set_strip_packing(self._h, int(mode)))

    def packed_items(self):
        v = ctypes.c_int()
        self._ok(self._lib.mhd_batch_get_packed_items(self._h, ctypes.byref(v)))
        return v.value

    def autotune(self, steps=2):
        """Sweep length, then launch shape, timed on the batch itself;
        returns the sweep length (waves_per_cu() gives the shape)."""
        v = ctypes.c_int()
        self._ok(self._lib.mhd_batch_autotune(self._h, int(steps), ctypes.byref(v)))
        return v.value


class ResistiveMHD2DAdjoint:
    """Forward rollout and discrete adjoint of the RK2 step for a batch of
    runs (mhd_adjoint_* of the C ABI): dt of every step frozen at its forward
    value, sqrt(N) checkpointing.  Gradients are of a loss of the final
    states with respect to the initial states (before their boundary fill)
    and the per-cell resistivity."""

    def __init__(self, cfgs):
        self._lib = load_library()
        self.cfgs = [MhdConfig.from_buffer_copy(c) for c in cfgs]
        arr = (MhdConfig * len(self.cfgs))(*self.cfgs)
        h = ctypes.c_void_p()
        rc = self._lib.mhd_adjoint_create(arr, len(self.cfgs), ctypes.byref(h))
        if rc != 0:
            msg = self._lib.mhd_adjoint_last_error(None)
            raise MHDError(f"libmhd2d error {rc}: {msg.decode() if msg else ''}")
        self._h = h
        self.nsteps = 0
        self._eta_set = {}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mhd_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.cfgs)

    def _ok(self, rc):
        if rc != 0:
            msg = self._lib.mhd_adjoint_last_error(self._h)
            raise MHDError(f"libmhd2d error {rc}: {msg.decode() if msg else ''}")

    def shape(self, run):
        c = self.cfgs[run]
        return (c.nx, c.ny, NVAR)

    def setEtaField(self, run, eta):
        if eta is None:
            self._ok(self._lib.mhd_adjoint_set_eta_field(self._h, run, None))
            self._eta_set[run] = False
            return
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        self._ok(self._lib.mhd_adjoint_set_eta_field(self._h, run, _dptr(eta)))
        self._eta_set[run] = True

    def setU0(self, run, U0):
        U0 = np.ascontiguousarray(U0, dtype=np.float64)
        if U0.shape != self.shape(run):
            raise ValueError(f"state shape {U0.shape} != {self.shape(run)}")
        self._ok(self._lib.mhd_adjoint_set_state(self._h, run, _dptr(U0)))

    def set_dts(self, run, dts):
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        self._ok(self._lib.mhd_adjoint_set_dts(self._h, run, _dptr(dts), len(dts)))

    def dts(self, run, nsteps=None):
        n = nsteps or self.nsteps
        out = np.empty(n)
        self._ok(self._lib.mhd_adjoint_get_dts(self._h, run, _dptr(out), n))
        return out

    STRIDE_FIT = -1

    def forward(self, nsteps, stride=0, frozen_dt=False):
        """Rollout of nsteps steps with a checkpoint every `stride` steps
        (0: ceil(sqrt(nsteps)); STRIDE_FIT: the largest stride up to nsteps
        whose storage fits in half of the free device memory, so that the
        backward pass recomputes nothing when the whole rollout fits)."""
        self._ok(self._lib.mhd_adjoint_forward(self._h, int(nsteps), int(stride), 1 if frozen_dt else 0))
        self.nsteps = int(nsteps)

    def stride(self):
        """Checkpoint stride of the last forward."""
        k = ctypes.c_int(0)
        self._ok(self._lib.mhd_adjoint_get_stride(self._h, ctypes.byref(k)))
        return k.value

    def getU(self, run):
        out = np.empty(self.shape(run))
        self._ok(self._lib.mhd_adjoint_get_state(self._h, run, _dptr(out)))
        return out

    def set_seed(self, run, lam):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        if lam.shape != self.shape(run):
            raise ValueError(f"seed shape {lam.shape} != {self.shape(run)}")
        self._ok(self._lib.mhd_adjoint_set_seed(self._h, run, _dptr(lam)))

    def backward(self):
        self._ok(self._lib.mhd_adjoint_backward(self._h))

    def grad_state(self, run):
        out = np.empty(self.shape(run))
        self._ok(self._lib.mhd_adjoint_get_grad_state(self._h, run, _dptr(out)))
        return out

    def grad_eta(self, run):
        c = self.cfgs[run]
        out = np.empty((c.nx, c.ny))
        self._ok(self._lib.mhd_adjoint_get_grad_eta(self._h, run, _dptr(out)))
        return out

    def set_direction(self, run, dU0=None, deta=None):
        """Direction of the tangent-linear model (None: zero)."""
        a = None if dU0 is None else np.ascontiguousarray(dU0, dtype=np.float64)
        b = None if deta is None else np.ascontiguousarray(deta, dtype=np.float64)
        self._ok(self._lib.mhd_adjoint_set_direction(self._h, run, _dptr(a) if a is not None else None,
                                                     _dptr(b) if b is not None else None))

    def tangent(self, nsteps=None):
        self._ok(self._lib.mhd_adjoint_tangent(self._h, int(nsteps or self.nsteps)))

    def replay(self, nsteps=None):
        """Decision replay: the rollout of the perturbed state set by
        set_replay_state with every branch decision and dt of the recorded
        rollout of U0 (read the result with get_tangent)."""
        self._ok(self._lib.mhd_adjoint_replay(self._h, int(nsteps or self.nsteps)))

    def set_replay_state(self, run, P0, eta_p=None):
        """Perturbed initial state (before its boundary fill) and perturbed
        resistivity of the decision replay.  A run with an eta field needs
        eta_p (the whole perturbed field); a run on the uniform cfg.eta keeps
        it."""
        if self._eta_set.get(run) and eta_p is None:
            raise ValueError("this run has an eta field: pass the perturbed field as eta_p")
        self.set_direction(run, P0, eta_p)

    def get_tangent(self, run):
        out = np.empty(self.shape(run))
        self._ok(self._lib.mhd_adjoint_get_tangent(self._h, run, _dptr(out)))
        return out

    # -- whole-batch transfers (lists of per-run arrays) --------------------
    def _concat(self, arrays):
        if len(arrays) != len(self.cfgs):
            raise ValueError("one array per run")
        for r, x in enumerate(arrays):
            if np.shape(x) != self.shape(r):
                raise ValueError(f"run {r}: shape {np.shape(x)} != {self.shape(r)}")
        return np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in arrays]))

    def _split(self, flat):
        out, o = [], 0
        for r in range(len(self.cfgs)):
            n = int(np.prod(self.shape(r)))
            out.append(flat[o:o + n].reshape(self.shape(r)))
            o += n
        return out

    def _total(self):
        return sum(int(np.prod(self.shape(r))) for r in range(len(self.cfgs)))

    def setU0_all(self, arrays):
        flat = self._concat(arrays)
        self._ok(self._lib.mhd_adjoint_set_states(self._h, _dptr(flat)))

    def getU_all(self):
        flat = np.empty(self._total())
        self._ok(self._lib.mhd_adjoint_get_states(self._h, _dptr(flat)))
        return self._split(flat)

    def set_seed_all(self, arrays):
        flat = self._concat(arrays)
        self._ok(self._lib.mhd_adjoint_set_seeds(self._h, _dptr(flat)))

    def grad_state_all(self):
        flat = np.empty(self._total())
        self._ok(self._lib.mhd_adjoint_get_grad_states(self._h, _dptr(flat)))
        return self._split(flat)

    def autotune(self, steps=2):
        """Sweep length of the forward stage kernels by the batch autotuner
        (on the initial states); returns the chosen column count."""
        cols = ctypes.c_int(0)
        self._ok(self._lib.mhd_adjoint_autotune(self._h, int(steps), ctypes.byref(cols)))
        return cols.value

    FACE_AUTO, FACE_PER_FACE, FACE_SWEEP = 0, 1, 2

    def set_face_kernel(self, mode):
        """Kernel form of the face adjoints: FACE_AUTO, FACE_PER_FACE or
        FACE_SWEEP (results agree to the rounding of atomic sums)."""
        self._ok(self._lib.mhd_adjoint_set_face_kernel(self._h, int(mode)))

    def set_degenerate_tol(self, tol):
        """Derivative rule at degenerate HLLD star states.  tol = 0 (the
        default) differentiates the step exactly.  tol > 0: on faces where
        |rho (S - u)(S - SM) - Bx^2| < tol rho (S - u)^2 the tangent and the
        adjoint take the derivative of Miyoshi & Kusano's degenerate limit
        (tangential star components equal to the outer ones); the forward
        step is unchanged."""
        self._ok(self._lib.mhd_adjoint_set_degenerate_tol(self._h, float(tol)))

    def grad_eta_all(self):
        n = [self.cfgs[r].nx * self.cfgs[r].ny for r in range(len(self.cfgs))]
        flat = np.empty(sum(n))
        self._ok(self._lib.mhd_adjoint_get_grad_etas(self._h, _dptr(flat)))
        out, o = [], 0
        for r, k in enumerate(n):
            out.append(flat[o:o + k].reshape(self.cfgs[r].nx, self.cfgs[r].ny))
            o += k
        return out

    # -- device-side transfers: flat float64 torch tensors on the GPU -------
    def _dev(self, t, numel):
        import torch
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise TypeError("expects a contiguous float64 tensor on the GPU")
        if t.numel() != numel:
            raise ValueError(f"{t.numel()} elements, expected {numel}")
        torch.cuda.current_stream().synchronize()
        return ctypes.c_void_p(t.data_ptr())

    def setEtaField_all_device(self, t):
        """Every run's eta plane (nx * ny per run, run order) from a flat
        float64 tensor on the GPU."""
        n = sum(c.nx * c.ny for c in self.cfgs)
        self._ok(self._lib.mhd_adjoint_set_eta_fields_device(self._h, self._dev(t, n)))
        for r in range(len(self.cfgs)):
            self._eta_set[r] = True

    def setU0_all_device(self, t):
        self._ok(self._lib.mhd_adjoint_set_states_device(self._h, self._dev(t, self._total())))

    def getU_all_device(self, t):
        self._ok(self._lib.mhd_adjoint_get_states_device(self._h, self._dev(t, self._total())))

    def set_seed_all_device(self, t):
        self._ok(self._lib.mhd_adjoint_set_seeds_device(self._h, self._dev(t, self._total())))

    def grad_state_all_device(self, t):
        self._ok(self._lib.mhd_adjoint_get_grad_states_device(self._h, self._dev(t, self._total())))

    def grad_eta_all_device(self, t):
        n = sum(c.nx * c.ny for c in self.cfgs)
        self._ok(self._lib.mhd_adjoint_get_grad_etas_device(self._h, self._dev(t, n)))
