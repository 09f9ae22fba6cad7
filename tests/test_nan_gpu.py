"""GPU parity on states that drive the step's rare paths, against the oracle
with NaN-aware comparison:

  - a cell whose energy is NaN and a cell with negative pressure: the HLLD
    star states and speeds turn non-finite, which the exact kernel hands to
    the reference's literal six-candidate sum (P/hlld.cu:35-107), and the
    crash check of P/resistiveMHD_2D.cu:294-309 (isnan(U.e) anywhere) must
    fire in both;
  - a band of flow faster than the fast speed towards -x: the clamped SR is
    exactly 0, so only the outer-right region (SR <= 0, P/hlld.cu:64-69) is
    active at its faces;
  - a band towards +x: the clamped SL = 0 puts its faces in the left star
    region built with a zero wave speed (P/hlld.cu:315-318).

The same states run as runs of a batch next to an ordinary run."""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from oracle.oracle import OracleSim
from tests.helpers import boundary_applied

pytestmark = pytest.mark.gpu

STEPS = 3


def _energy(rho, u, v, w, bx, by, bz, p, gamma):
    return p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)


def special_state(cfg, U, nan_cell, negp_cell, left_band, right_band):
    """U with a NaN energy, a negative-pressure cell and two bands of
    super-fast flow along x (speed 12 against fast speeds of order 2)."""
    U = U.copy()
    g = cfg.gamma
    for (i0, i1), speed in ((left_band, -12.0), (right_band, 12.0)):
        rho = U[i0:i1, :, 0]
        v, w = U[i0:i1, :, 2] / rho, U[i0:i1, :, 3] / rho
        bx, by, bz = U[i0:i1, :, 4], U[i0:i1, :, 5], U[i0:i1, :, 6]
        p = (g - 1.0) * (U[i0:i1, :, 7] - 0.5 * U[i0:i1, :, 1] ** 2 / rho - 0.5 * rho * (v * v + w * w)
                         - 0.5 * (bx * bx + by * by + bz * bz))
        U[i0:i1, :, 1] = rho * speed
        U[i0:i1, :, 7] = _energy(rho, speed, v, w, bx, by, bz, p, g)
    i, j = nan_cell
    U[i, j, 7] = np.nan
    i, j = negp_cell
    rho = U[i, j, 0]
    kin = 0.5 * (U[i, j, 1] ** 2 + U[i, j, 2] ** 2 + U[i, j, 3] ** 2) / rho
    U[i, j, 7] = kin + 0.5 * (U[i, j, 4] ** 2 + U[i, j, 5] ** 2 + U[i, j, 6] ** 2) - 0.3   # p = -0.3 (gamma - 1)
    return boundary_applied(cfg, U)


def cases():
    c1 = mhd.orszag_tang_config(64, eta=1e-3)
    U1 = special_state(c1, boundary_applied(c1, mhd.fill_orszag_tang(c1)), (40, 40), (20, 50), (8, 16), (48, 56))
    c2 = mhd.harris_config(96, 70, eta_max=6e-3)
    U2 = special_state(c2, boundary_applied(c2, mhd.fill_harris(c2, seed=3, noise_amp=1e-2)), (60, 30), (30, 44),
                       (10, 20), (70, 80))
    eta2 = mhd.fill_harris_eta(c2)
    return [("orszag_tang_64", c1, U1, None), ("harris_96x70", c2, U2, eta2)]


def _oracle(cfg, U0, eta):
    o = OracleSim.from_config(cfg)
    o.set_state(U0)
    if eta is not None:
        o.set_eta_field(eta)
    o.step(STEPS)
    return o


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def _check(name, Ug, Uo, crashed_g, crashed_o, dt_g, dt_o):
    nan_cells = int(np.isnan(Uo[..., 7]).sum())
    print(f"{name}: NaN energies {nan_cells}, crashed gpu={crashed_g} oracle={crashed_o}, "
          f"dt gpu={dt_g!r} oracle={dt_o!r}", flush=True)
    assert nan_cells > 0
    assert crashed_o and crashed_g
    assert _same(Ug, Uo), np.argwhere(~((Ug == Uo) | (np.isnan(Ug) & np.isnan(Uo))))[:10]
    assert _same(np.array(dt_g), np.array(dt_o))


@pytest.mark.parametrize("case", [0, 1])
def test_nan_and_superfast_single(case):
    name, cfg, U0, eta = cases()[case]
    print(f"case {name}: single handle, {STEPS} steps", flush=True)
    o = _oracle(cfg, U0, eta)
    s = mhd.ResistiveMHD2D(cfg)
    s.setU(U0)
    if eta is not None:
        s.setEtaField(eta)
    s.steps(STEPS)
    _check(name, s.getU(), o.get_state(), s.checkCalculationIsCrashed(), o.crashed(), s.dt, o.dt)


def test_nan_and_superfast_batch():
    print(f"case batch of the special states and an ordinary run, {STEPS} steps", flush=True)
    cs = cases()
    c0 = mhd.orszag_tang_config(48, eta=1e-3)
    U00 = boundary_applied(c0, mhd.fill_orszag_tang(c0))
    runs = [(cs[0][1], cs[0][2], None), (c0, U00, None), (cs[1][1], cs[1][2], cs[1][3])]
    b = mhd.ResistiveMHD2DBatch([c for c, _, _ in runs])
    for r, (c, U0, eta) in enumerate(runs):
        b.setU(r, U0)
        if eta is not None:
            b.setEtaField(r, eta)
    b.steps(STEPS)
    for r, (c, U0, eta) in enumerate(runs):
        o = _oracle(c, U0, eta)
        if r == 1:
            assert not b.checkCalculationIsCrashed(r) and not o.crashed()
            assert np.array_equal(b.getU(r), o.get_state())
            assert b.dt(r) == o.dt
        else:
            _check(f"batch run {r}", b.getU(r), o.get_state(), b.checkCalculationIsCrashed(r), o.crashed(),
                   b.dt(r), o.dt)
