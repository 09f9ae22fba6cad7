"""GPU parity: the HIP step against the CPU oracle on the same inputs.

Two arithmetic modes are checked:
  exact -- the reference's expressions operation for operation: the state
           after N steps must be bitwise identical to the oracle's;
  fast  -- shared reciprocals / FMA: the reference's HLLD flux jumps where a
           rotational wave speed crosses zero, so rounding-level differences
           can flip a region; checked at relative L2 <= 1e-8 per variable and
           dt, time to 1e-10.  Exact (the default) carries the 1e-10 bar.
"""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from tests.helpers import NAMES, boundary_applied, gpu_run, oracle_run, rel_l2

pytestmark = pytest.mark.gpu

TOL_FAST = 1e-8


def _compare(cfg, U0, steps, numerics, eta=None, sweep=None, tag=""):
    print(f"case {tag}: nx={cfg.nx} ny={cfg.ny} bc=({cfg.bc_x},{cfg.bc_y}) steps={steps} "
          f"numerics={numerics} sweep={sweep}", flush=True)
    o = oracle_run(cfg, U0, steps, eta)
    g = gpu_run(cfg, U0, steps, eta, sweep, numerics)
    Uo, Ug = o.get_state(), g.getU()
    err = rel_l2(Ug, Uo)
    print(tag, numerics, " ".join(f"{n}={e:.2e}" for n, e in zip(NAMES, err)), flush=True)
    assert np.all(np.isfinite(Ug))
    assert not g.checkCalculationIsCrashed()
    if numerics == "exact":
        assert np.array_equal(Ug, Uo), err
        assert g.dt == o.dt and g.totalTime == o.time
    else:
        assert np.all(err <= TOL_FAST), err
        assert abs(g.dt - o.dt) <= 1e-10 * abs(o.dt)
        assert abs(g.totalTime - o.time) <= 1e-10 * abs(o.time)
    return Ug, Uo


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_orszag_tang_64_periodic(numerics):
    cfg = mhd.orszag_tang_config(64, eta=1e-3)
    U0 = boundary_applied(cfg, mhd.fill_orszag_tang(cfg))
    _compare(cfg, U0, 100, numerics, tag="ot64")


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_orszag_tang_ragged_symmetric_xy(numerics):
    cfg = mhd.orszag_tang_config(70, eta=1e-3)
    cfg.ny = 70 + 6 + 13  # ragged strip count along y
    cfg.bc_x = mhd.MHD_BC_SYMMETRIC
    cfg.bc_y = mhd.MHD_BC_SYMMETRIC
    U0 = boundary_applied(cfg, mhd.fill_orszag_tang(cfg))
    _compare(cfg, U0, 40, numerics, sweep=16, tag="ot-sym")


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_harris_small_eta_field(numerics):
    cfg = mhd.harris_config(128, 96)
    U0 = boundary_applied(cfg, mhd.fill_harris(cfg, seed=3, noise_amp=1e-2))
    eta = mhd.fill_harris_eta(cfg)
    _compare(cfg, U0, 60, numerics, eta=eta, tag="harris")


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_orszag_tang_256_100_steps(numerics):
    """The plain-method case at its full size: Orszag-Tang 256^2, eta 1e-3,
    CFL 0.7, checked after 100 steps."""
    cfg = mhd.orszag_tang_config(256, eta=1e-3)
    U0 = boundary_applied(cfg, mhd.fill_orszag_tang(cfg))
    _compare(cfg, U0, 100, numerics, tag="ot256")


@pytest.mark.parametrize("sweep", [0, 8, 33, 1000])
def test_sweep_lengths_agree(sweep):
    """The result does not depend on how columns are split between
    wavefronts (segment warm-up recomputes the same values)."""
    cfg = mhd.harris_config(150, 70)
    U0 = boundary_applied(cfg, mhd.fill_harris(cfg, seed=5, noise_amp=1e-2))
    eta = mhd.fill_harris_eta(cfg)
    _compare(cfg, U0, 20, "exact", eta=eta, sweep=sweep, tag=f"sweep{sweep}")


def _div_b_interior(cfg, U):
    g = mhd.GHOST
    bx, by = U[..., 4], U[..., 5]
    return (bx[g:-g, g:-g] - bx[g - 1:-g - 1, g:-g]) / cfg.dx + (by[g:-g, g:-g] - by[g:-g, g - 1:-g - 1]) / cfg.dy


def test_harris_4096x2048_full_size_properties():
    """The headline case at its full size (Harris sheet 4096 x 2048,
    periodic x, symmetric y, per-cell eta, 20 steps): the discrete
    divergence of B stays at round-off, and shifting the start (state and
    eta plane) along the periodic x axis shifts the result bitwise."""
    cfg = mhd.harris_config(4096, 2048)
    steps, g = 20, mhd.GHOST
    print(f"case harris4096x2048-properties: steps={steps}", flush=True)
    sim = mhd.ResistiveMHD2D(cfg)
    sim.initHarris(seed=0, noise_amp=1e-3, eta0=1e-3, eta1=5e-3)
    U0 = sim.getU()
    eta = mhd.fill_harris_eta(cfg, 1e-3, 5e-3)
    sim.steps(steps)
    assert not sim.checkCalculationIsCrashed()
    U1 = sim.getU()
    bscale = np.abs(U1[..., 4:6]).max() / min(cfg.dx, cfg.dy)
    div = np.abs(_div_b_interior(cfg, U1)).max()
    print(f"max|divB| {div:.2e} (B/dx {bscale:.2e})", flush=True)
    assert div < 1e-12 * bscale
    sx = 1024
    Us = U0.copy()
    Us[g:-g] = np.roll(U0[g:-g], sx, axis=0)
    es = eta.copy()
    es[g:-g] = np.roll(eta[g:-g], sx, axis=0)
    del U0
    sim2 = mhd.ResistiveMHD2D(cfg)
    sim2.initializeU(Us)
    sim2.setEtaField(es)
    del Us, es
    sim2.steps(steps)
    assert np.array_equal(np.roll(U1[g:-g, g:-g], sx, axis=0), sim2.getU()[g:-g, g:-g])
    assert sim2.dt == sim.dt


def test_harris_4096x2048_configured_length():
    """The headline run at its configured length (Harris sheet 4096 x 2048,
    2000 steps): it does not crash, every value stays finite, and the
    discrete divergence of B stays at round-off throughout."""
    cfg = mhd.harris_config(4096, 2048)
    print("case harris4096x2048-2000-steps", flush=True)
    sim = mhd.ResistiveMHD2D(cfg)
    sim.initHarris(seed=0, noise_amp=1e-3, eta0=1e-3, eta1=5e-3)
    t = 0.0
    for chunk in range(4):
        sim.steps(500)
        assert not sim.checkCalculationIsCrashed(), chunk
        U = sim.getU()
        assert np.all(np.isfinite(U)), chunk
        bscale = np.abs(U[..., 4:6]).max() / min(cfg.dx, cfg.dy)
        div = np.abs(_div_b_interior(cfg, U)).max()
        print(f"steps {500 * (chunk + 1)}: time {sim.totalTime:.4f} max|divB| {div:.2e}", flush=True)
        assert div < 1e-12 * bscale, chunk
        assert sim.totalTime > t
        t = sim.totalTime


def _golden_hashes():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "full_size_hashes.json")
    return json.load(open(path))


@pytest.mark.parametrize("name", ["harris_4096x2048", "harris_4096x4096"])
def test_full_size_headline_bitwise_against_oracle_hash(name):
    """The headline workloads at full size (Harris sheet 4096 x 2048 and
    4096 x 4096, 100 steps, exact numerics): the state is bitwise the
    oracle's, compared through the SHA-256 the oracle's run wrote
    (tests/golden/make_full_size_hashes.py), with dt and time."""
    from tests.golden.make_full_size_hashes import state_hash
    ref = _golden_hashes()[name]
    print(f"case {name}: {ref['steps']} steps against the oracle's hash", flush=True)
    cfg = mhd.harris_config(ref["nx"], ref["ny"], ly=ref["ly"])
    sim = mhd.ResistiveMHD2D(cfg)
    sim.initHarris(seed=ref["seed"], noise_amp=ref["noise_amp"], eta0=ref["eta0"], eta1=ref["eta1"])
    assert state_hash(sim.getU()) == ref["initial_state_sha256"]
    sim.steps(ref["steps"])
    assert not sim.checkCalculationIsCrashed()
    assert state_hash(sim.getU()) == ref["state_sha256"]
    assert sim.dt == ref["dt"] and sim.totalTime == ref["time"]
