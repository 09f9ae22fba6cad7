"""Shared helpers of the test suite: run the same setup through the HIP
library and through the CPU oracle, and compare states."""
import numpy as np

from oracle.oracle import OracleSim
import resistivemhd_gpu_amd as mhd

NAMES = ("rho", "rhoU", "rhoV", "rhoW", "bX", "bY", "bZ", "e")


def rel_l2(a, b):
    """Relative L2 error of a against b, per conserved variable (last axis)."""
    num = np.sqrt(((a - b) ** 2).reshape(-1, a.shape[-1]).sum(axis=0))
    den = np.sqrt((b ** 2).reshape(-1, b.shape[-1]).sum(axis=0))
    den = np.where(den == 0.0, 1.0, den)
    return num / den


def oracle_run(cfg, U0, steps, eta_field=None):
    o = OracleSim.from_config(cfg)
    o.set_state(U0)
    if eta_field is not None:
        o.set_eta_field(eta_field)
    o.step(steps)
    return o


def gpu_run(cfg, U0, steps, eta_field=None, sweep=None, numerics="fast"):
    sim = mhd.ResistiveMHD2D(cfg)
    sim.set_numerics(numerics)
    sim.setU(U0)
    if eta_field is not None:
        sim.setEtaField(eta_field)
    if sweep:
        sim.set_sweep_length(sweep)
    sim.steps(steps)
    return sim


def boundary_applied(cfg, U0):
    """U0 with the configured boundary applied by the oracle (as the
    examples' initializeU does)."""
    o = OracleSim.from_config(cfg)
    o.set_state(U0)
    o.apply_boundary()
    return o.get_state()


def interior(U):
    g = mhd.GHOST
    return U[g:-g, g:-g]
