"""GPU checks of the batched form (mhd_batch_*): every run of a batch must
end bitwise where a single handle with the same configuration and numerics
ends (same kernels body, one launch per stage over all runs' items), and so
where the oracle ends in exact numerics; graph replay, eager steps, sweep
lengths and the autotuner must not change any result."""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from oracle.oracle import OracleSim
from tests.helpers import boundary_applied

pytestmark = pytest.mark.gpu


def _mixed_runs():
    """Uneven grids, all boundary combinations, uniform and per-cell eta."""
    runs = []
    c = mhd.orszag_tang_config(64, eta=1e-3)
    runs.append((c, boundary_applied(c, mhd.fill_orszag_tang(c)), None))
    c = mhd.orszag_tang_config(70, eta=3e-4)
    c.ny = 70 + 6 + 47
    c.bc_x = c.bc_y = mhd.MHD_BC_SYMMETRIC
    runs.append((c, boundary_applied(c, mhd.fill_orszag_tang(c)), None))
    c = mhd.harris_config(150, 70)
    eta = mhd.fill_harris_eta(c)
    runs.append((c, boundary_applied(c, mhd.fill_harris(c, seed=4, noise_amp=1e-2)), eta))
    c = mhd.harris_config(33, 130, eta_max=1e-2)
    c.eta = 1e-2
    runs.append((c, boundary_applied(c, mhd.fill_harris(c, seed=5, noise_amp=1e-2)), None))
    c = mhd.orszag_tang_config(12, eta=1e-4)     # smallest interior (one strip, few columns)
    runs.append((c, boundary_applied(c, mhd.fill_orszag_tang(c)), None))
    c = mhd.orszag_tang_config(200, eta=5e-3)
    c.bc_x = mhd.MHD_BC_SYMMETRIC
    runs.append((c, boundary_applied(c, mhd.fill_orszag_tang(c)), None))
    return runs


def _batch(runs, numerics="exact"):
    b = mhd.ResistiveMHD2DBatch([c for c, _, _ in runs])
    b.set_numerics(numerics)
    for r, (c, U0, eta) in enumerate(runs):
        b.setU(r, U0)
        if eta is not None:
            b.setEtaField(r, eta)
    return b


def _single(c, U0, eta, steps, numerics):
    s = mhd.ResistiveMHD2D(c)
    s.set_numerics(numerics)
    s.setU(U0)
    if eta is not None:
        s.setEtaField(eta)
    s.steps(steps)
    return s


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_batch_matches_single_handles(numerics):
    runs = _mixed_runs()
    steps = 31   # graph replays plus one eager step
    print(f"case batch-vs-single numerics={numerics} runs={len(runs)} steps={steps}", flush=True)
    b = _batch(runs, numerics)
    b.steps(steps)
    for r, (c, U0, eta) in enumerate(runs):
        s = _single(c, U0, eta, steps, numerics)
        assert np.array_equal(b.getU(r), s.getU()), r
        assert b.dt(r) == s.dt and b.totalTime(r) == s.totalTime, r
        assert not b.checkCalculationIsCrashed(r)


def test_batch_matches_oracle_exact():
    runs = _mixed_runs()[:3]
    steps = 20
    print(f"case batch-vs-oracle runs={len(runs)} steps={steps}", flush=True)
    b = _batch(runs)
    b.steps(steps)
    for r, (c, U0, eta) in enumerate(runs):
        o = OracleSim.from_config(c)
        o.set_state(U0)
        if eta is not None:
            o.set_eta_field(eta)
        o.step(steps)
        assert np.array_equal(b.getU(r), o.get_state()), r
        assert b.dt(r) == o.dt and b.totalTime(r) == o.time


def test_batch_graph_eager_sweep_and_autotune_agree():
    runs = _mixed_runs()
    print("case batch-graph-eager-autotune", flush=True)
    a = _batch(runs)
    a.set_graph(False)
    a.steps(7)
    a.steps(6)
    b = _batch(runs)
    b.autotune(steps=1)          # must leave the states where they were
    b.steps(4)
    b.set_sweep_length(9)
    b.steps(9)
    for r in range(len(runs)):
        assert np.array_equal(a.getU(r), b.getU(r)), r
        assert a.totalTime(r) == b.totalTime(r)


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_batch_persistent_launch_shapes_agree(numerics):
    """Persistent stage launches (w wavefronts per CU looping over the work
    list, including fewer wavefronts than one XCD's items) end bitwise where
    one wavefront per item ends; the shape the autotuner picks is valid."""
    runs = _mixed_runs()
    print(f"case batch-persistent numerics={numerics}", flush=True)
    ref = _batch(runs, numerics)
    ref.steps(9)
    for w in (1, 3, 8):
        b = _batch(runs, numerics)
        b.set_sweep_length(11)
        b.set_waves_per_cu(w)
        assert b.waves_per_cu() == w
        b.steps(5)
        b.set_sweep_length(24)      # a new work list keeps the shape
        b.steps(4)
        for r in range(len(runs)):
            assert np.array_equal(ref.getU(r), b.getU(r)), (w, r)
            assert ref.totalTime(r) == b.totalTime(r)
    b = _batch(runs, numerics)
    b.autotune(steps=1)
    assert b.waves_per_cu() >= 0
    b.steps(9)
    for r in range(len(runs)):
        assert np.array_equal(ref.getU(r), b.getU(r)), r


def _packing_runs():
    """Last strips of every packable height class: 4 rows (16 segments per
    wavefront), 16, 28 (two per wavefront), a run that is all last strip
    (20 rows, symmetric x, whose bottom stale face goes through the packed
    path), and one that is not packable (40 rows)."""
    runs = []
    c = mhd.orszag_tang_config(64, eta=1e-3)                 # 60 + 4
    runs.append((c, boundary_applied(c, mhd.fill_orszag_tang(c)), None))
    c = mhd.orszag_tang_config(136, eta=3e-4)                # 120 + 16
    c.bc_y = mhd.MHD_BC_SYMMETRIC
    runs.append((c, boundary_applied(c, mhd.fill_orszag_tang(c)), None))
    c = mhd.harris_config(150, 88)                           # 60 + 28
    eta = mhd.fill_harris_eta(c)
    runs.append((c, boundary_applied(c, mhd.fill_harris(c, seed=4, noise_amp=1e-2)), eta))
    c = mhd.harris_config(97, 20, eta_max=1e-2)              # 20 rows only
    c.bc_x = mhd.MHD_BC_SYMMETRIC
    runs.append((c, boundary_applied(c, mhd.fill_harris(c, seed=5, noise_amp=1e-2)), None))
    c = mhd.orszag_tang_config(100, eta=5e-3)                # 60 + 40: not packed
    runs.append((c, boundary_applied(c, mhd.fill_orszag_tang(c)), None))
    return runs


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_batch_packed_strips_match_single_handles(numerics):
    """Packed last strips (several column segments per wavefront, a launch
    of their own) end bitwise where single handles end, with segments that
    do and do not fill a packed wavefront and a last segment shorter than
    the rest; AUTO packs the batched workload's shape and leaves a small
    batch unpacked."""
    runs = _packing_runs()
    steps = 12
    print(f"case batch-packed numerics={numerics} runs={len(runs)} steps={steps}", flush=True)
    singles = [_single(c, U0, eta, steps, numerics) for c, U0, eta in runs]
    # (sweep length, persistent waves per CU): a persistent launch runs the
    # packed items as a launch of their own
    for L, w in ((7, 0), (16, 0), (7, 3)):
        b = _batch(runs, numerics)
        b.set_sweep_length(L)
        b.set_waves_per_cu(w)
        b.set_strip_packing(b.STRIP_PACK_ALWAYS)
        assert b.packed_items() > 0
        b.steps(steps)
        for r, s in enumerate(singles):
            assert np.array_equal(b.getU(r), s.getU()), (L, w, r)
            assert b.dt(r) == s.dt and b.totalTime(r) == s.totalTime, (L, w, r)
    b = _batch(runs, numerics)
    b.set_strip_packing(b.STRIP_PACK_NEVER)
    assert b.packed_items() == 0
    b.set_strip_packing(b.STRIP_PACK_AUTO)
    assert b.packed_items() == 0       # a few hundred items: one round


def test_batch_init_and_boundary():
    cfgs = [mhd.orszag_tang_config(48), mhd.harris_config(64, 32)]
    b = mhd.ResistiveMHD2DBatch(cfgs)
    b.initOrszagTang(0)
    b.initHarris(1, seed=2, noise_amp=1e-3)
    b.applyBoundary()
    assert np.array_equal(b.getU(0), boundary_applied(cfgs[0], mhd.fill_orszag_tang(cfgs[0])))
    assert np.array_equal(b.getU(1), boundary_applied(cfgs[1], mhd.fill_harris(cfgs[1], seed=2, noise_amp=1e-3)))
    with pytest.raises(mhd.MHDError):
        b.initOrszagTang(5)


def test_batch_save_matches_single_handle_and_loads_back(tmp_path):
    """A run's snapshot is byte for byte the file a single handle writes for
    the same state (the reference's save layout), and loading it restores
    the run; a file of the wrong size is refused."""
    runs = _mixed_runs()[:3]
    print("case batch-save-load", flush=True)
    b = _batch(runs)
    b.steps(5)
    for r, (c, U0, eta) in enumerate(runs):
        s = _single(c, U0, eta, 5, "exact")
        b.save(r, str(tmp_path), f"batch{r}", 5)
        s.save(str(tmp_path), f"single{r}", 5)
        assert (tmp_path / f"batch{r}_5.bin").read_bytes() == (tmp_path / f"single{r}_5.bin").read_bytes(), r
    U1 = b.getU(1)
    b.steps(3)
    b.load(1, str(tmp_path / "batch1_5.bin"))
    assert np.array_equal(b.getU(1), U1)
    with pytest.raises(mhd.MHDError):
        b.load(0, str(tmp_path / "batch1_5.bin"))   # another run's grid size
