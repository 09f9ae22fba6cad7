"""HLLD region diagnostic (mhd_face_regions) on the GPU.

The diagnostic evaluates the region indicators of P/hlld.cu:64-69 for every
face of a state, in either arithmetic, from the stage kernel's
reconstruction.  Pinned by the uniform-state closed forms (the same cases
as the oracle's, tests/test_oracle.py) in both arithmetics, and by equality
with the oracle's regions (oracle_face_regions, the reference's pass
structure) on evolved Harris and Orszag-Tang states in exact arithmetic.
The count of faces whose fast-arithmetic region differs from the exact one
is then a measurement (bench.py's fast_vs_exact block reports it at the
benchmark sizes)."""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from oracle.oracle import OracleSim
from tests.helpers import boundary_applied, oracle_run
from tests.test_oracle import face_interiors, region_case_states

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_face_regions_uniform_closed_forms(numerics):
    for n, (U, ex, ey) in enumerate(region_case_states()):
        print(f"case uniform-{n} numerics={numerics} expect x={ex} y={ey}", flush=True)
        cfg = mhd.orszag_tang_config(U.shape[0] - 2 * mhd.GHOST)
        cfg.ny = U.shape[1]
        sim = mhd.ResistiveMHD2D(cfg)
        sim.setU(U)
        rx, ry = sim.face_regions(numerics)
        fx, fy = face_interiors(rx, ry)
        assert np.all(fx == ex), (ex, np.unique(fx))
        assert np.all(fy == ey), (ey, np.unique(fy))


def _unread_faces_marked(rx, ry):
    """0xFF exactly outside the faces the update reads."""
    mx = np.full(rx.shape, True)
    my = np.full(ry.shape, True)
    mx[2:-3, 3:-3] = False
    my[3:-3, 2:-3] = False
    return np.all(rx[mx] == 0xFF) and np.all(ry[my] == 0xFF)


def _cases():
    cfg = mhd.harris_config(128, 96)
    U0 = boundary_applied(cfg, mhd.fill_harris(cfg, seed=3, noise_amp=1e-2))
    yield "harris128x96-20", cfg, U0, mhd.fill_harris_eta(cfg), 20
    cfg = mhd.orszag_tang_config(64, eta=1e-3)
    yield "ot64-30", cfg, boundary_applied(cfg, mhd.fill_orszag_tang(cfg)), None, 30


def test_exact_face_regions_match_oracle():
    for tag, cfg, U0, eta, steps in _cases():
        print(f"case {tag}", flush=True)
        o = oracle_run(cfg, U0, steps, eta)
        U = o.get_state()
        sim = mhd.ResistiveMHD2D(cfg)
        sim.setU(U)
        rx, ry = sim.face_regions("exact")
        ox, oy = face_interiors(*o.face_regions())
        gx, gy = face_interiors(rx, ry)
        assert _unread_faces_marked(rx, ry)
        assert np.array_equal(gx, ox), int((gx != ox).sum())
        assert np.array_equal(gy, oy), int((gy != oy).sum())
        # the states exercise more than one region
        assert len(np.unique(np.concatenate([ox.ravel(), oy.ravel()]))) >= 3


def test_fast_region_flips_are_rare():
    """Fast arithmetic takes the exact HLLD branch on all but a small
    fraction of faces of the same state (a measurement; the bound only
    catches a fast path that reconstructs different faces)."""
    for tag, cfg, U0, eta, steps in _cases():
        o = oracle_run(cfg, U0, steps, eta)
        sim = mhd.ResistiveMHD2D(cfg)
        sim.setU(o.get_state())
        ex, ey = face_interiors(*sim.face_regions("exact"))
        fx, fy = face_interiors(*sim.face_regions("fast"))
        flips = int((ex != fx).sum() + (ey != fy).sum())
        print(f"case {tag}: {flips} of {ex.size + ey.size} faces flip", flush=True)
        assert flips <= 1e-3 * (ex.size + ey.size)
