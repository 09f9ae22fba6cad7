"""GPU parity: the HIP step against the CPU oracle on the same inputs.

Tolerance: relative L2 error per conserved variable <= 1e-10 after the
stated number of steps (the FP64 bar), and the dt sequence to 1e-12.
"""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from tests.helpers import NAMES, boundary_applied, gpu_run, oracle_run, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _report(tag, err):
    print(tag, " ".join(f"{n}={e:.2e}" for n, e in zip(NAMES, err)), flush=True)


def _compare(cfg, U0, steps, eta=None, sweep=None, tag=""):
    print(f"case {tag}: nx={cfg.nx} ny={cfg.ny} steps={steps} sweep={sweep}", flush=True)
    o = oracle_run(cfg, U0, steps, eta)
    g = gpu_run(cfg, U0, steps, eta, sweep)
    Uo, Ug = o.get_state(), g.getU()
    err = rel_l2(Ug, Uo)
    _report(tag, err)
    assert np.all(np.isfinite(Ug))
    assert np.all(err <= TOL), err
    assert abs(g.dt - o.dt) <= 1e-12 * abs(o.dt)
    assert abs(g.totalTime - o.time) <= 1e-12 * abs(o.time)
    return Ug, Uo


def test_orszag_tang_64_periodic():
    cfg = mhd.orszag_tang_config(64, eta=1e-3)
    U0 = boundary_applied(cfg, mhd.fill_orszag_tang(cfg))
    _compare(cfg, U0, 100, tag="ot64")


def test_orszag_tang_ragged_symmetric_xy():
    cfg = mhd.orszag_tang_config(70, eta=1e-3)
    cfg.ny = 70 + 6 + 13  # ragged strip count along y
    cfg.bc_x = mhd.MHD_BC_SYMMETRIC
    cfg.bc_y = mhd.MHD_BC_SYMMETRIC
    U0 = boundary_applied(cfg, mhd.fill_orszag_tang(cfg))
    _compare(cfg, U0, 40, sweep=16, tag="ot-sym")


def test_harris_small_eta_field():
    cfg = mhd.harris_config(128, 96)
    U0 = boundary_applied(cfg, mhd.fill_harris(cfg, seed=3, noise_amp=1e-2))
    eta = mhd.fill_harris_eta(cfg)
    _compare(cfg, U0, 60, eta=eta, tag="harris")
