"""GPU checks at the configured workloads of the two forms beside the plain
step (the shapes bench.py times):

  batched -- 256 independent runs, Orszag-Tang / Harris 50/50, N drawn from
             {64, 128, 192, 256, 384, 512}^2, eta log-uniform in [1e-4, 1e-2],
             autotuned sweep and launch shape, packed last strips (AUTO),
             500 steps: sampled runs of every size end bitwise where single
             handles end, and the smallest ones where the oracle ends;
  adjoint -- Harris sheet 1024 x 512, batch of 32, 2000 steps, the
             configured gradient check (bench.gradient_check).
"""
import numpy as np
import pytest

import bench
import resistivemhd_gpu_amd as mhd
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def _init_single(c, kind, seed):
    s = mhd.ResistiveMHD2D(c)
    if kind == "harris":
        s.initializeU(mhd.fill_harris(c, seed=seed, noise_amp=1e-3))
    else:
        s.initOrszagTang()
    return s


def test_batched_workload_256_runs_500_steps():
    runs = bench.batch_configs(mhd, 256)
    steps = 500
    sizes = sorted({c.nx - 6 for c, _, _ in runs})
    print(f"case batched workload: {len(runs)} runs, sizes {sizes}, {steps} steps", flush=True)
    assert sizes == list(bench.BATCH_SIZES)
    b = mhd.ResistiveMHD2DBatch([c for c, _, _ in runs])
    for r, (c, kind, seed) in enumerate(runs):
        if kind == "harris":
            b.initHarris(r, seed=seed, noise_amp=1e-3)
        else:
            b.initOrszagTang(r)
    b.applyBoundary()
    cols = b.autotune(steps=2)
    packed = b.packed_items()
    print(f"autotuned sweep {cols}, waves per CU {b.waves_per_cu()}, packed items {packed}", flush=True)
    assert packed > 0                      # AUTO packs the short last strips of this shape
    b.steps(steps)
    # two runs of each (size, setup) pair, which includes runs whose last
    # strip is packed (every size but 512: its last strip has 32 rows)
    sample, seen = [], {}
    for r, (c, kind, _) in enumerate(runs):
        k = (c.nx, kind)
        if seen.get(k, 0) < 2:
            seen[k] = seen.get(k, 0) + 1
            sample.append(r)
    for r in sample:
        c, kind, seed = runs[r]
        s = _init_single(c, kind, seed)
        s.steps(steps)
        assert not b.checkCalculationIsCrashed(r), r
        assert np.array_equal(b.getU(r), s.getU()), (r, c.nx, kind)
        assert b.dt(r) == s.dt and b.totalTime(r) == s.totalTime, r
    print(f"{len(sample)} sampled runs bitwise equal to single handles", flush=True)
    # the smallest runs against the oracle itself
    for r in [r for r in sample if runs[r][0].nx == 64 + 6][:2]:
        c, kind, seed = runs[r]
        o = OracleSim.from_config(c)
        o.set_state(mhd.fill_harris(c, seed=seed, noise_amp=1e-3) if kind == "harris" else mhd.fill_orszag_tang(c))
        o.apply_boundary()
        o.step(steps)
        assert np.array_equal(b.getU(r), o.get_state()), r
        assert b.dt(r) == o.dt


def test_gradient_check_configured_case():
    """The configured differentiation case: Harris sheet 1024 x 512, batch
    of 32, 2000 steps, observed states from seeded 6-Gaussian resistivity
    fields, loss on rho, Bx, By from eta = 1e-3, degenerate-limit rule at
    1e-6.  On run 0 along 16 random directions in (eta, vx0, vy0): the
    adjoint agrees with the tangent-linear model to 1e-10 and with central
    differences (h = 1e-5) of the decision replay, the map the adjoint
    differentiates, to 1e-6 (measured: below 5e-8)."""
    print("case gradient check: harris 1024x512, batch 32, 2000 steps, 16 directions", flush=True)
    res = bench.gradient_check(32, 1024, 512, 2000, 16, degenerate_tol=1e-6, fd_h=1e-5, plain_fd=False,
                               log=lambda m: print(m, flush=True))
    for k, v in enumerate(res["values"]):
        print(f"dir {k}: adjoint {v['adjoint']:.12e} tangent {v['tangent']:.12e} replay fd {v['replay_fd']:.12e} "
              f"rel {res['tangent_rel_errs'][k]:.1e} / {res['replay_fd_rel_errs'][k]:.1e}", flush=True)
    assert res["tangent_rel_err_max"] <= 1e-10
    assert res["replay_fd_rel_err_max"] <= 1e-6
