"""GPU parity of the MPI library's step on one GPU's subdomain
(ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/resistiveMHD2D.cu:147-221 with
sendrecv_U, mpi.cu:104-188, on a rank that is its own neighbour): the HIP
MHD_SCHEME_MULTI_SHARE step against the oracle's restatement of it.

Exact numerics must match the oracle bitwise; fast numerics to 1e-8
relative L2 per variable (the reference's HLLD region choice is
discontinuous, so rounding differences are not held to 1e-10).  At the full
8192^2 subdomain, where the oracle would take minutes per step, the checks
are size-independent: periodic translation equivariance (bitwise), mass and
energy conservation to round-off and the face divergence of B held at
round-off by the final CT.
"""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from oracle.oracle import SCHEME_MULTI_SHARE, OracleSim
from tests.helpers import NAMES, rel_l2

pytestmark = pytest.mark.gpu

TOL_FAST = 1e-8


def _halo(cfg, U0):
    o = OracleSim.from_config(cfg)
    o.set_scheme(SCHEME_MULTI_SHARE)
    o.set_state(U0)
    o.apply_halo()
    return o.get_state()


def _gpu(cfg, U0, steps, numerics="exact", sweep=None, planes=None):
    sim = mhd.ResistiveMHD2D(cfg)
    sim.set_numerics(numerics)
    sim.set_scheme("multi")
    if planes is not None:
        sim.setEtaStaggered(*planes)
    sim.setU(U0)
    if sweep:
        sim.set_sweep_length(sweep)
    sim.steps(steps)
    return sim


def _compare(cfg, U0, steps, numerics, sweep=None, tag="", planes=None):
    print(f"case {tag}: nx={cfg.nx} ny={cfg.ny} steps={steps} numerics={numerics} sweep={sweep}", flush=True)
    o = OracleSim.from_config(cfg)
    o.set_scheme(SCHEME_MULTI_SHARE)
    if planes is not None:
        o.set_eta_staggered(*planes)
    o.set_state(U0)
    o.step(steps)
    g = _gpu(cfg, U0, steps, numerics, sweep, planes)
    Uo, Ug = o.get_state(), g.getU()
    err = rel_l2(Ug, Uo)
    print(tag, numerics, " ".join(f"{n}={e:.2e}" for n, e in zip(NAMES, err)), flush=True)
    assert np.all(np.isfinite(Ug))
    assert not g.checkCalculationIsCrashed()
    assert not o.crashed()
    if numerics == "exact":
        assert np.array_equal(Ug, Uo), err
        assert g.dt == o.dt and g.totalTime == o.time
    else:
        assert np.all(err <= TOL_FAST), err
        assert abs(g.dt - o.dt) <= 1e-10 * abs(o.dt)
    return g, Ug


def _div_b(cfg, U):
    g = mhd.GHOST
    bx, by = U[..., 4], U[..., 5]
    return ((bx[g:-g, g:-g] - bx[g - 1:-g - 1, g:-g]) / cfg.dx
            + (by[g:-g, g:-g] - by[g:-g, g - 1:-g - 1]) / cfg.dy)


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_multi_share_orszag_tang_64(numerics):
    cfg = mhd.orszag_tang_config(64, eta=1e-3)
    U0 = _halo(cfg, mhd.fill_orszag_tang(cfg))
    _compare(cfg, U0, 100, numerics, tag="m-ot64")


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_multi_share_turbulence_ragged(numerics):
    """Turbulence on a non-square subdomain whose rows do not fill the last
    wavefront strip, split into short sweeps."""
    cfg = mhd.default_config(96, 70 + 61, 0.0, 2.0 * np.pi, 0.0, 3.0 * np.pi, eta=1e-4)
    U0 = _halo(cfg, mhd.fill_turbulence(cfg, seed=7, kmax=8, rms_v=0.5, rms_b=0.5))
    _compare(cfg, U0, 60, numerics, sweep=16, tag="m-turb")


def test_multi_share_column_walk_remainders():
    """The CT-side kernels walk columns in segments (32 for the corner EMF
    with the CT, 8 for the smoothed cell EMF) and take 63 rows per
    wavefront: 100 interior columns leave a short last segment of each, and
    130 rows a last wavefront of 4 rows."""
    cfg = mhd.default_config(100, 130, 0.0, 2.0 * np.pi, 0.0, 2.6 * np.pi, eta=1e-4)
    U0 = _halo(cfg, mhd.fill_turbulence(cfg, seed=11, kmax=6, rms_v=0.5, rms_b=0.5))
    _compare(cfg, U0, 40, "exact", tag="m-turb-remainders")


def test_multi_share_turbulence_512():
    """Bitwise at a size where every kernel of the step runs many work items:
    the stage kernels several strips and segments, the walking CT kernels
    16 (and 64) column segments and 9 row groups."""
    cfg = mhd.turbulence_config(512, eta=1e-4)
    U0 = _halo(cfg, mhd.fill_turbulence(cfg, seed=5, kmax=8, rms_v=0.5, rms_b=0.5))
    _compare(cfg, U0, 30, "exact", tag="m-turb512")


def _periodic_eta_planes(cfg, seed, base):
    """getEta of a smooth periodic resistivity (a random Fourier sum around
    `base`, positive) sampled at the MPI library's x-face, y-face and cell
    positions of every interior index; the ghost entries are copies of
    their periodic images (which a periodic getEta gives up to rounding)."""
    rng = np.random.default_rng(seed)
    lx = (cfg.nx - 6) * cfg.dx
    ly = (cfg.ny - 6) * cfg.dy
    modes = [(int(rng.integers(1, 4)), int(rng.integers(0, 4)), rng.uniform(0, 2 * np.pi), rng.uniform(0.05, 0.2))
             for _ in range(4)]

    def eta(x, y):
        v = np.ones(np.broadcast(x, y).shape)
        for kx, ky, ph, a in modes:
            v += a * np.sin(2 * np.pi * (kx * x / lx + ky * y / ly) + ph)
        return base * v

    i = np.arange(cfg.nx - 6)[:, None] + 0.0
    j = np.arange(cfg.ny - 6)[None, :] + 0.0
    planes = (eta((i + 0.5) * cfg.dx, j * cfg.dy), eta(i * cfg.dx, (j + 0.5) * cfg.dy), eta(i * cfg.dx, j * cfg.dy))
    return [np.pad(p, 3, mode="wrap") for p in planes]


@pytest.mark.parametrize("numerics", ["exact", "fast"])
def test_multi_share_eta_planes_orszag_tang(numerics):
    """Spatial resistivity at the MPI library's three sampling positions
    (M/flux_solver.cu:84,168, M/ct.cu:41): bitwise the oracle's step."""
    cfg = mhd.orszag_tang_config(64, eta=1e-3)
    U0 = _halo(cfg, mhd.fill_orszag_tang(cfg))
    _compare(cfg, U0, 60, numerics, tag="m-ot64-eta", planes=_periodic_eta_planes(cfg, 3, 5e-3))


def test_multi_share_eta_planes_turbulence_ragged():
    cfg = mhd.default_config(96, 70 + 61, 0.0, 2.0 * np.pi, 0.0, 3.0 * np.pi, eta=1e-4)
    U0 = _halo(cfg, mhd.fill_turbulence(cfg, seed=7, kmax=8, rms_v=0.5, rms_b=0.5))
    _compare(cfg, U0, 40, "exact", sweep=16, tag="m-turb-eta", planes=_periodic_eta_planes(cfg, 4, 2e-3))


def test_multi_share_eta_planes_refusals():
    """Planes only under the multi-GPU scheme, all three or none, no
    negative entries, ghost entries equal to their periodic images; switching
    back to the single-GPU scheme drops them."""
    cfg = mhd.orszag_tang_config(32, eta=1e-3)
    planes = _periodic_eta_planes(cfg, 5, 1e-3)
    sim = mhd.ResistiveMHD2D(cfg)
    with pytest.raises(mhd.MHDError):
        sim.setEtaStaggered(*planes)
    sim.set_scheme("multi")
    sim.setEtaStaggered(*planes)
    bad = [p.copy() for p in planes]
    bad[1][0, 10] += 1e-6                       # a ghost entry that is not its image
    with pytest.raises(mhd.MHDError):
        sim.setEtaStaggered(*bad)
    bad = [p.copy() for p in planes]
    bad[2][10, 10] = -1e-3
    with pytest.raises(mhd.MHDError):
        sim.setEtaStaggered(*bad)
    with pytest.raises(mhd.MHDError):
        sim._ok(sim._lib.mhd_set_eta_staggered(sim._h, mhd.mhd2d._dptr(planes[0]), None, None))
    # uniform planes of cfg.eta are the uniform step, bit for bit
    U0 = _halo(cfg, mhd.fill_orszag_tang(cfg))
    sim.setEtaStaggered(*[np.full((cfg.nx, cfg.ny), cfg.eta)] * 3)
    sim.setU(U0)
    sim.steps(5)
    ref = _gpu(cfg, U0, 5)
    assert np.array_equal(sim.getU(), ref.getU())
    sim.set_scheme("single")
    sim.set_scheme("multi")
    sim.setU(U0)
    sim.steps(5)
    assert np.array_equal(sim.getU(), ref.getU())


def test_multi_share_init_matches_host_fill():
    """mhd_init_turbulence = host fill + self-exchange halo."""
    cfg = mhd.turbulence_config(64)
    sim = mhd.ResistiveMHD2D(cfg)
    sim.set_scheme("multi")
    sim.initTurbulence(seed=11)
    assert np.array_equal(sim.getU(), _halo(cfg, mhd.fill_turbulence(cfg, seed=11)))


def test_multi_share_rejects_symmetric_and_eta_field():
    cfg = mhd.harris_config(64, 32)
    sim = mhd.ResistiveMHD2D(cfg)
    with pytest.raises(mhd.MHDError):
        sim.set_scheme("multi")
    cfg = mhd.orszag_tang_config(32)
    sim = mhd.ResistiveMHD2D(cfg)
    sim.setEtaField(np.full((cfg.nx, cfg.ny), 1e-3))
    with pytest.raises(mhd.MHDError):
        sim.set_scheme("multi")
    sim.setEtaField(None)
    sim.set_scheme("multi")
    with pytest.raises(mhd.MHDError):
        sim.setEtaField(np.full((cfg.nx, cfg.ny), 1e-3))
    with pytest.raises(mhd.MHDError):
        sim.setBoundary(mhd.MHD_BC_SYMMETRIC, mhd.MHD_BC_PERIODIC)


def _dt_pair(cfg, U0):
    """calculateDt of the HIP library and of the oracle, both schemes."""
    out = []
    for scheme in ("single", "multi"):
        sim = mhd.ResistiveMHD2D(cfg)
        sim.set_scheme(scheme)
        sim.setU(U0)
        o = OracleSim.from_config(cfg)
        if scheme == "multi":
            o.set_scheme(SCHEME_MULTI_SHARE)
        o.set_state(U0)
        out.append((sim.calculateDt(), o.calculate_dt()))
    return out


def test_multi_share_dt_interior_without_diffusion_limit():
    """calculateDt of the MPI library: interior cells only, no eta limit;
    the single-GPU library's: every cell, with the limit."""
    cfg = mhd.orszag_tang_config(32, eta=1e-3)
    U0 = _halo(cfg, mhd.fill_orszag_tang(cfg))
    U0[0, 5, 7] += 0.5 * 40.0 ** 2 / U0[0, 5, 0]   # a fast ghost cell with valid pressure
    U0[0, 5, 1] = 40.0
    (gs, os_), (gm, om) = _dt_pair(cfg, U0)
    assert gs == os_ and gm == om
    assert gs < 0.5 * gm            # only the all-cells dt sees the ghost cell
    cfg.eta = 10.0                  # diffusion limit far below the CFL limit
    (gs, os_), (gm, om) = _dt_pair(cfg, U0)
    assert gs == os_ and gm == om
    assert gs == min(cfg.dx ** 2 / 10.0, cfg.dy ** 2 / 10.0) * cfg.cfl
    assert gm > 5.0 * gs


def test_dt_nan_in_first_visited_cell():
    """A NaN per-cell dt is passed over unless it is the first cell the
    sequential minimum visits ((0, 0), or (3, 3) for the interior-only dt),
    where it sticks: the single scheme then falls back to the diffusion limit
    and the multi scheme's dt is NaN."""
    cfg = mhd.orszag_tang_config(32, eta=1e-3)
    base = _halo(cfg, mhd.fill_orszag_tang(cfg))
    U0 = base.copy()
    U0[0, 0, 1] = 1e6                 # negative pressure: NaN sound speed
    (gs, os_), (gm, om) = _dt_pair(cfg, U0)
    assert gs == os_ == min(cfg.dx ** 2, cfg.dy ** 2) / cfg.eta * cfg.cfl
    assert gm == om and np.isfinite(gm)
    U0 = base.copy()
    U0[3, 3, 1] = 1e6
    (gs, os_), (gm, om) = _dt_pair(cfg, U0)
    assert gs == os_ and np.isfinite(gs)
    assert np.isnan(gm) and np.isnan(om)
    U0 = base.copy()
    U0[7, 9, 1] = 1e6                 # elsewhere: passed over by both
    (gs, os_), (gm, om) = _dt_pair(cfg, U0)
    assert gs == os_ and gm == om and np.isfinite(gs) and np.isfinite(gm)


def test_multi_share_turbulence_8192_properties():
    """One GPU's share at full size: 8192^2 turbulence, eta 1e-4, 20 steps."""
    n, steps = 8192, 20
    cfg = mhd.turbulence_config(n, eta=1e-4)
    print(f"case m-turb8192: steps={steps}", flush=True)
    sim = mhd.ResistiveMHD2D(cfg)
    sim.set_scheme("multi")
    sim.initTurbulence(seed=0, kmax=8, rms_v=0.5, rms_b=0.5)
    U0 = sim.getU()
    g = mhd.GHOST
    mass0 = U0[g:-g, g:-g, 0].sum()
    energy0 = U0[g:-g, g:-g, 7].sum()
    div0 = np.abs(_div_b(cfg, U0)).max()
    sim.steps(steps)
    U1 = sim.getU()
    assert not sim.checkCalculationIsCrashed()
    mass1 = U1[g:-g, g:-g, 0].sum()
    energy1 = U1[g:-g, g:-g, 7].sum()
    div1 = np.abs(_div_b(cfg, U1)).max()
    bscale = np.abs(U1[..., 4:6]).max() / min(cfg.dx, cfg.dy)
    print(f"mass {abs(mass1 / mass0 - 1):.2e} energy {abs(energy1 / energy0 - 1):.2e} "
          f"divB {div0:.2e} -> {div1:.2e} (B/dx {bscale:.2e})", flush=True)
    assert abs(mass1 / mass0 - 1) < 1e-12
    assert abs(energy1 / energy0 - 1) < 1e-12
    assert div1 < 1e-12 * bscale
    del U1

    # translation equivariance: the shifted start gives the shifted result
    sx, sy = 1234, 4321
    Ui = U0[g:-g, g:-g]
    Us = np.empty_like(U0)
    Us[g:-g, g:-g] = np.roll(Ui, (sx, sy), axis=(0, 1))
    del U0
    sim2 = mhd.ResistiveMHD2D(cfg)
    sim2.set_scheme("multi")
    sim2.setU(Us)
    sim2.periodicBoundaryX2nd()
    sim2.periodicBoundaryY2nd()
    del Us
    sim2.steps(steps)
    a = sim.getU()[g:-g, g:-g]
    b = sim2.getU()[g:-g, g:-g]
    assert np.array_equal(np.roll(a, (sx, sy), axis=(0, 1)), b)
