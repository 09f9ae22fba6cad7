"""MI355X-native 2D resistive-MHD RK2 step.

The compute path is the HIP library ``libmhd2d.so`` (C ABI in
``include/mhd2d.h``); this package holds its sources (``csrc/``), the build
recipe and the Python mirror of the reference's ``ResistiveMHD2D`` class.
"""
from .mhd2d import (  # noqa: F401
    GHOST, MHD_BC_PERIODIC, MHD_BC_SYMMETRIC, MHD_NUMERICS_EXACT, MHD_NUMERICS_FAST, MHD_SCHEME_MULTI_SHARE,
    MHD_SCHEME_SINGLE, NVAR, EXPORTS, MHDError, MhdConfig, ResistiveMHD2D, ResistiveMHD2DAdjoint, ResistiveMHD2DBatch, default_config, fill_harris,
    fill_harris_eta, fill_orszag_tang, fill_turbulence, harris_config, load_library, orszag_tang_config,
    turbulence_config,
)
