"""The Thrust twin of the reference's pass structure (oracle/thrust_twin.cpp,
the benchmark's CPU baseline) against the oracle: the same states, dt and
time bit for bit, for every boundary combination and a per-cell eta field,
on one and on several OpenMP threads."""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from oracle.oracle import OracleSim, TwinSim
from tests.helpers import boundary_applied


def _cases():
    c1 = mhd.orszag_tang_config(48, eta=1e-3)
    yield "ot48", c1, boundary_applied(c1, mhd.fill_orszag_tang(c1)), None
    c2 = mhd.harris_config(64, 40, eta_max=6e-3)
    yield "harris64x40", c2, boundary_applied(c2, mhd.fill_harris(c2, seed=2, noise_amp=1e-2)), mhd.fill_harris_eta(c2)
    c3 = mhd.orszag_tang_config(30, eta=3e-4)
    c3.ny = 30 + 6 + 7
    c3.bc_x = c3.bc_y = mhd.MHD_BC_SYMMETRIC
    yield "ot-sym", c3, boundary_applied(c3, mhd.fill_orszag_tang(c3)), None


@pytest.mark.parametrize("threads", [1, 4])
def test_twin_equals_oracle(threads):
    TwinSim.set_threads(threads)
    for name, cfg, U0, eta in _cases():
        o = OracleSim.from_config(cfg)
        o.set_state(U0)
        t = TwinSim.from_config(cfg)
        t.set_state(U0)
        if eta is not None:
            o.set_eta_field(eta)
            t.set_eta_field(eta)
        o.step(12)
        assert not t.step(12), name
        assert np.array_equal(t.get_state(), o.get_state()), name
        assert t.dt == o.dt and t.time == o.time, name


def test_twin_boundary_equals_oracle():
    for name, cfg, _, _ in _cases():
        rng = np.random.default_rng(3)
        U = rng.normal(size=(cfg.nx, cfg.ny, 8))
        o = OracleSim.from_config(cfg)
        o.set_state(U)
        o.apply_boundary()
        t = TwinSim.from_config(cfg)
        t.set_state(U)
        t.apply_boundary()
        assert np.array_equal(t.get_state(), o.get_state()), name
