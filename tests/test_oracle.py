"""Pins of the CPU oracle (oracle/mhd2d_oracle.c) against what the
mathematics fixes, independent of the oracle's own expressions.

The reference ships no tests, golden vectors or fixtures, so each oracle
function is checked against closed forms, invariants and special cases:
  minmod            -- truth table of the limiter including the EPS floor
  HLLD              -- consistency F(U,U) = ideal-MHD flux (textbook formula
                       written here), exact resolution of stationary contacts,
                       mirror symmetry under x -> -x
  resistive flux    -- linear B: constant current, closed-form eta*J flux
                       (also at the MPI library's face and cell sampling
                       positions of getEta)
  dt                -- closed form for a uniform state
  RK2 + CT + BC     -- conservation of mass, momentum and energy under
                       periodic boundaries, div B held to round-off, a uniform
                       state stays bitwise uniform, boundary copy/mirror
                       identities, div-free rebuilt ghost faces, the stencil
                       extent of one step.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import OracleSim, BC_PERIODIC, BC_SYMMETRIC

GAMMA = 5.0 / 3.0
EPS = 1e-20


# ----------------------------------------------------------------- minmod
@pytest.mark.parametrize("x,y,expect", [
    (1.0, 2.0, 1.0), (2.0, 1.0, 1.0), (-1.0, -3.0, -1.0), (-3.0, -1.0, -1.0),
    (1.0, -1.0, EPS), (-1.0, 1.0, -EPS), (3.0, 0.0, EPS), (-3.0, 0.0, -EPS),
    (0.0, 5.0, 0.0), (0.0, -5.0, 0.0), (1e-30, 1.0, EPS), (2.5, 2.5, 2.5),
])
def test_minmod_truth_table(x, y, expect):
    # limiter: sign(x) * max(min(|x|, sign(x) y), EPS); equal signs give the
    # smaller magnitude, opposite signs give sign(x)*EPS, x == 0 gives 0
    assert orc.minmod(x, y, EPS) == expect


# ------------------------------------------------------------------- HLLD
def ideal_mhd_flux(rho, u, v, w, bx, by, bz, p, gamma=GAMMA):
    """Textbook x-flux of ideal MHD in conservative variables."""
    b2 = bx * bx + by * by + bz * bz
    pt = p + 0.5 * b2
    e = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * b2
    return np.array([
        rho * u,
        rho * u * u + pt - bx * bx,
        rho * u * v - bx * by,
        rho * u * w - bx * bz,
        0.0,
        u * by - v * bx,
        u * bz - w * bx,
        (e + pt) * u - bx * (u * bx + v * by + w * bz),
    ])


def _states(seed, n):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        rho = rng.uniform(0.1, 3.0)
        u, v, w = rng.uniform(-2, 2, 3)
        bx, by, bz = rng.uniform(-2, 2, 3)
        p = rng.uniform(0.05, 3.0)
        yield np.array([rho, u, v, w, bx, by, bz, p])


def test_hlld_consistency():
    for q in _states(1, 200):
        f = orc.hlld_flux(q, q)
        ref = ideal_mhd_flux(*q)
        np.testing.assert_allclose(f, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("bx", [0.0, 0.7, -1.3])
def test_hlld_resolves_stationary_contact(bx):
    # density jump with equal velocity (at rest), pressure and field: the
    # exact solution is a stationary contact; HLLD resolves it exactly
    rng = np.random.default_rng(7)
    for _ in range(50):
        by, bz = rng.uniform(-1, 1, 2)
        p = rng.uniform(0.2, 2.0)
        L = np.array([rng.uniform(0.5, 2.0), 0.0, 0.0, 0.0, bx, by, bz, p])
        R = L.copy()
        R[0] = rng.uniform(0.05, 0.4)
        f = orc.hlld_flux(L, R)
        ref = ideal_mhd_flux(*L)  # zero velocity: the flux is the same on both sides
        np.testing.assert_allclose(f, ref, rtol=1e-12, atol=1e-12)


def test_hlld_mirror_symmetry():
    # x -> -x maps (u, Bx) -> (-u, -Bx) and the flux to
    # (-f0, f1, -f2, -f3, 0, -f5, -f6, -f7) with left and right swapped
    sgn = np.array([-1, 1, -1, -1, 1, -1, -1, -1.0])
    for L, R in zip(_states(11, 300), _states(12, 300)):
        R = R.copy()
        R[4] = L[4]                      # one normal field per interface
        m = lambda q: q * np.array([1, -1, 1, 1, -1, 1, 1, 1.0])
        f = orc.hlld_flux(L, R)
        g = orc.hlld_flux(m(R), m(L))
        np.testing.assert_allclose(g, sgn * f, rtol=1e-9, atol=1e-11 * (1 + np.abs(f).max()))


def _superfast_pairs(seed, n):
    """Pairs (L, R) sharing Bx whose flow is faster than the fast speed on
    both sides (u > max(cf_L, cf_R), so the unclamped HLLD speed S_L > 0)."""
    rng = np.random.default_rng(seed)

    def cf(q):
        rho, _, _, _, bx, by, bz, p = q
        cs2, ca2, va2 = GAMMA * p / rho, (bx * bx + by * by + bz * bz) / rho, bx * bx / rho
        return np.sqrt(0.5 * (cs2 + ca2 + np.sqrt((cs2 + ca2) ** 2 - 4.0 * cs2 * va2)))

    for _ in range(n):
        L = np.array([rng.uniform(0.5, 2.0), 0.0, *rng.uniform(-1, 1, 2), *rng.uniform(-1, 1, 3),
                      rng.uniform(0.2, 2.0)])
        R = np.array([rng.uniform(0.5, 2.0), 0.0, *rng.uniform(-1, 1, 2), *rng.uniform(-1, 1, 3),
                      rng.uniform(0.2, 2.0)])
        R[4] = L[4]
        c = max(cf(L), cf(R))
        L[1], R[1] = c + rng.uniform(0.05, 2.0), c + rng.uniform(0.05, 2.0)
        yield L, R


def test_hlld_outer_speed_clamp():
    """Q2, P/hlld.cu:315-318: the outer speeds are clamped, SL = min(SL, 0),
    SR = max(SR, 0).  For a flow faster than the fast speed towards +x the
    unclamped solver would select the outer-left region and return the
    left state's physical flux F(L) itself.  Clamped, SL = 0 and the flux is
    the physical flux of the left star state built with a zero wave speed:
    equal to F(L) only up to rounding (the star state satisfies the jump
    conditions across a wave of speed 0), and not bit for bit.  Towards -x
    the clamp gives SR = 0 exactly, which selects the outer-right region
    alone: the flux is F(R) bit for bit.  Mirroring x -> -x maps one case
    onto the other, and F of a mirrored state is the mirrored F bit for bit
    (every term only changes sign), so the mirrored leftward flux is the
    exact F(L) of the rightward case, which the rightward flux must differ
    from."""
    sgn = np.array([-1, 1, -1, -1, 1, -1, -1, -1.0])
    m = lambda q: q * np.array([1, -1, 1, 1, -1, 1, 1, 1.0])
    differs = 0
    n = 200
    for L, R in _superfast_pairs(21, n):
        f = orc.hlld_flux(L, R)                 # rightward: left star state at SL = 0
        g = orc.hlld_flux(m(R), m(L))           # leftward: outer-right region, F(m(L))
        ref = ideal_mhd_flux(*L)
        np.testing.assert_allclose(f, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())
        np.testing.assert_allclose(g, sgn * ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())
        differs += not np.array_equal(f, sgn * g)
    assert differs >= 0.9 * n, differs


# --------------------------------------------------------------------- dt
# HLLD region bits: outer-left, single-star left, double-star left,
# outer-right, single-star right, double-star right
R_OL, R_ML, R_IL, R_OR, R_MR, R_IR = 1, 2, 4, 8, 16, 32

# (normal velocity, expected bits) for a uniform state whose normal Alfven
# speed |Bn|/sqrt(rho) is ca and fast speed cf: with L = R the HLLD speeds
# are SM = u_n, S1L = u_n - ca, S1R = u_n + ca, SL = min(u_n - cf, 0),
# SR = max(u_n + cf, 0) (the clamp P/hlld.cu:315-318), so the region is set
# by where 0 falls among them.  The clamp makes a super-fast flow to the
# right single-star left (SL = 0 <= 0 < S1L) and one to the left outer-right
# only (SR = 0).
REGION_CASES = [
    (0.0, R_IR), ("+ca/2", R_IL), ("(ca+cf)/2", R_ML), ("2cf", R_ML),
    ("-ca/2", R_IR), ("-(ca+cf)/2", R_MR), ("-2cf", R_OR),
]


def _speed(tag, ca, cf):
    return {0.0: 0.0, "+ca/2": 0.5 * ca, "(ca+cf)/2": 0.5 * (ca + cf), "2cf": 2.0 * cf,
            "-ca/2": -0.5 * ca, "-(ca+cf)/2": -0.5 * (ca + cf), "-2cf": -2.0 * cf}[tag]


def region_case_states(nx=16, ny=14, rho=1.0, p=0.6, bx=0.8, by=0.5, bz=0.3):
    """Uniform states for each REGION_CASES row, with the expected x-face and
    y-face bits (shared by the GPU diagnostic's test)."""
    cs2, b2 = GAMMA * p / rho, (bx * bx + by * by + bz * bz) / rho

    def cf(bn):
        va2 = bn * bn / rho
        return np.sqrt(0.5 * (cs2 + b2 + np.sqrt((cs2 + b2) ** 2 - 4.0 * cs2 * va2)))

    cax, cay = abs(bx) / np.sqrt(rho), abs(by) / np.sqrt(rho)
    out = []
    for (tx, ex), (ty, ey) in zip(REGION_CASES, REGION_CASES[::-1]):
        u, v = _speed(tx, cax, cf(bx)), _speed(ty, cay, cf(by))
        out.append((_uniform_state(nx, ny, rho=rho, u=u, v=v, w=0.1, bx=bx, by=by, bz=bz, p=p), ex, ey))
    return out


def face_interiors(rx, ry):
    """The faces the update reads: x-faces (i+1/2, j) for i in [2, nx-4],
    j in [3, ny-4]; y-faces (i, j+1/2) for i in [3, nx-4], j in [2, ny-4]."""
    return rx[2:-3, 3:-3], ry[3:-3, 2:-3]


def test_face_regions_uniform_closed_forms():
    for U, ex, ey in region_case_states():
        o = OracleSim(U.shape[0], U.shape[1], 0.1, 0.1, GAMMA, 0.7, EPS, 1e-3)
        o.set_state(U)
        fx, fy = face_interiors(*o.face_regions())
        assert np.all(fx == ex), (ex, np.unique(fx))
        assert np.all(fy == ey), (ey, np.unique(fy))


def test_dt_uniform_closed_form():
    for q in _states(3, 50):
        rho, u, v, w, bx, by, bz, p = q
        e = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
        cell = np.array([rho, rho * u, rho * v, rho * w, bx, by, bz, e])
        dx, dy = 0.1, 0.05
        c = np.sqrt(GAMMA * p / rho + (bx * bx + by * by + bz * bz) / rho)
        expect = 1.0 / ((abs(u) + c) / dx + (abs(v) + c) / dy)
        assert orc.cell_dt(cell, GAMMA, dx, dy) == pytest.approx(expect, rel=1e-13)


def _uniform_state(nx, ny, rho=1.2, u=0.3, v=-0.2, w=0.1, bx=0.5, by=-0.4, bz=0.2, p=0.8):
    e = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
    U = np.empty((nx, ny, 8))
    U[...] = [rho, rho * u, rho * v, rho * w, bx, by, bz, e]
    return U


def test_calculate_dt_with_diffusion_limit():
    nx, ny, dx, dy = 20, 16, 0.1, 0.05
    U = _uniform_state(nx, ny)
    for eta, cfl in [(0.0, 0.7), (1e-3, 0.7), (0.5, 0.4)]:
        o = OracleSim(nx, ny, dx, dy, GAMMA, cfl, EPS, eta)
        o.set_state(U)
        c = orc.cell_dt(U[0, 0], GAMMA, dx, dy)
        lim = min(dx * dx / eta, dy * dy / eta) if eta > 0 else np.inf
        assert o.calculate_dt() == pytest.approx(cfl * min(c, lim), rel=1e-15)


# --------------------------------------------------------- resistive flux
def test_resistive_flux_linear_field():
    # rho, p uniform, v = 0, Bx = 0, By = a*x on the y-faces: J_z = a is
    # uniform, the HLLD part vanishes (MUSCL reconstructs a linear By
    # exactly so left and right states agree and u = v = 0), so
    # F.f5 = -eta*a, F.f6 = 0, F.f7 = -eta*a*(By_i + By_{i+1})/2 on every
    # interior x-face, and G.f4 = +eta*a on every interior y-face.
    nx, ny, dx, dy, eta, a = 20, 18, 0.1, 0.1, 3e-3, 0.8
    U = _uniform_state(nx, ny, u=0.0, v=0.0, w=0.0, bx=0.0, by=0.0, bz=0.0, p=1.0)
    x = (np.arange(nx) - 2.5) * dx
    U[:, :, 5] = a * x[:, None]
    U[:, :, 7] += 0.5 * U[:, :, 5] ** 2
    o = OracleSim(nx, ny, dx, dy, GAMMA, 0.7, EPS, eta)
    F, G = o.compute_fluxes(U)
    i, j = slice(3, nx - 4), slice(3, ny - 4)
    np.testing.assert_allclose(F[i, j, 5], -eta * a, rtol=1e-12)
    np.testing.assert_allclose(F[i, j, 6], 0.0, atol=1e-15)
    byc = U[:, :, 5]
    expect7 = -eta * a * 0.5 * (byc[3:nx - 4] + byc[4:nx - 3])
    np.testing.assert_allclose(F[i, j, 7], expect7[:, j] if expect7.ndim == 2 else expect7, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(G[i, j, 4], eta * a, rtol=1e-12)


# ----------------------------------------------- whole step: invariants
def _ot_state(n, seed=None):
    import resistivemhd_gpu_amd as mhd
    cfg = mhd.orszag_tang_config(n, eta=1e-3)
    U = mhd.fill_orszag_tang(cfg)
    o = OracleSim.from_config(cfg)
    o.set_state(U)
    o.apply_boundary()
    return cfg, o


def _divb(U, dx, dy):
    bx, by = U[:, :, 4], U[:, :, 5]
    return (bx[3:-3, 3:-3] - bx[2:-4, 3:-3]) / dx + (by[3:-3, 3:-3] - by[3:-3, 2:-4]) / dy


def test_conservation_periodic():
    cfg, o = _ot_state(48)
    U0 = o.get_state()
    o.step(25)
    U1 = o.get_state()
    for v in (0, 1, 2, 3, 7):
        s0, s1 = U0[3:-3, 3:-3, v].sum(), U1[3:-3, 3:-3, v].sum()
        scale = np.abs(U0[3:-3, 3:-3, v]).sum()
        assert abs(s1 - s0) <= 1e-12 * scale, (v, s0, s1)


def test_div_b_preserved():
    cfg, o = _ot_state(40)
    # a state whose div B is not zero stays at its own value: CT changes
    # face B by a discrete curl, whose divergence vanishes identically
    U = o.get_state()
    rng = np.random.default_rng(5)
    U[:, :, 4] += 1e-3 * rng.standard_normal(U.shape[:2])
    o.set_state(U)
    o.apply_boundary()
    U0 = o.get_state()
    d0 = _divb(U0, cfg.dx, cfg.dy)
    o.step(10)
    d1 = _divb(o.get_state(), cfg.dx, cfg.dy)
    assert np.abs(d1 - d0).max() <= 1e-10 * np.abs(d0).max()


def test_stage2_ct_reuses_stage1_emf():
    """Q7, P/resistiveMHD_2D.cu:167-168,190 and P/ct.cu:12-19: CT::setOldFlux2D
    is called once per step and CT::divBClean twice, so the second CT
    applies the stage-1 corner EMF to the same old face field.  The
    interior face B after the step must therefore equal the face B of the
    stage-1 state UBar right after its CT bit for bit."""
    cfg, o = _ot_state(40)
    U = o.get_state()
    rng = np.random.default_rng(8)
    U[..., 1:4] += 1e-2 * rng.standard_normal(U[..., 1:4].shape)
    o.set_state(U)
    o.apply_boundary()
    o.step(3)
    Un = o.get_state()
    bx, by = o.stage1_face_b()
    inner = (slice(3, -3), slice(3, -3))
    assert np.array_equal(Un[inner][..., 4], bx[inner])
    assert np.array_equal(Un[inner][..., 5], by[inner])
    # and the step did move the field (the check is not vacuous)
    o.step(1)
    assert not np.array_equal(o.get_state()[inner][..., 4], Un[inner][..., 4])


def _gs05_corner_emf(Ex, Ey, Ec, mx, my):
    """Corner EMF at (i+1/2, j+1/2) of Gardiner & Stone (2005), J. Comput.
    Phys. 205, 509, eqs. (41), (45)-(50), written from the paper: the
    average of the four face EMFs plus the upwinded derivatives of Ez from
    the face to the corner.  Ex[(a, b)]: Ez at x-face (i+1/2, j+b) for a = 0;
    Ey[(a, b)]: Ez at y-face (i+a, j+1/2); Ec[(a, b)]: Ez at cell (i+a, j+b);
    mx[b]: mass flux through x-face (i+1/2, j+b); my[a]: through y-face
    (i+a, j+1/2).  Grid spacings cancel: (dy/8) * 2/dy = 1/4."""
    def up(flux, lower, upper):
        return lower if flux > 0 else (upper if flux < 0 else 0.5 * (lower + upper))
    # (dE/dy) at (i+1/2, j+1/4) and (i+1/2, j+3/4): from the cell on the
    # upwind side of the x-face (i for positive mass flux, i+1 for negative)
    dy_lo = up(mx[0], Ey[(0, 0)] - Ec[(0, 0)], Ey[(1, 0)] - Ec[(1, 0)])
    dy_hi = up(mx[1], Ec[(0, 1)] - Ey[(0, 0)], Ec[(1, 1)] - Ey[(1, 0)])
    # (dE/dx) at (i+1/4, j+1/2) and (i+3/4, j+1/2): upwind along y
    dx_lo = up(my[0], Ex[(0, 0)] - Ec[(0, 0)], Ex[(0, 1)] - Ec[(0, 1)])
    dx_hi = up(my[1], Ec[(1, 0)] - Ex[(0, 0)], Ec[(1, 1)] - Ex[(0, 1)])
    avg = 0.25 * (Ex[(0, 0)] + Ex[(0, 1)] + Ey[(0, 0)] + Ey[(1, 0)])
    return avg + 0.25 * (dy_lo - dy_hi) + 0.25 * (dx_lo - dx_hi)


def test_ct_contact_emf_upwinding():
    """The MPI library's corner EMF (M/ct.cu:147-169) is the CT-contact EMF
    of Gardiner & Stone (2005): checked against the paper's formula on
    random patches, with every sign combination of the face mass fluxes
    (zero included).  Reversing the upwind direction changes the result by
    O(1) on these patches."""
    rng = np.random.default_rng(31)
    cells = [(0, 0), (0, 1), (1, 0), (1, 1)]     # the oracle's patch order
    for t in range(400):
        Ex = {(0, 0): rng.normal(), (0, 1): rng.normal()}
        Ey = {(0, 0): rng.normal(), (1, 0): rng.normal()}
        Ec = {c: rng.normal() for c in cells}
        signs = rng.integers(-1, 2, 4) if t % 4 == 0 else rng.choice([-1.0, 1.0], 4)
        mx = [signs[0] * rng.uniform(0.1, 2), signs[1] * rng.uniform(0.1, 2)]
        my = [signs[2] * rng.uniform(0.1, 2), signs[3] * rng.uniform(0.1, 2)]
        # the reference's arrays: x-face F.f5 = -Ez, y-face G.f4 = Ez, cell
        # terms F5 = -Ez_c and G4 = Ez_c (M/ct.cu:26-65), mass fluxes F.f0, G.f0
        NF5 = [-Ex[(0, 0)], -Ex[(0, 1)], 0.0, 0.0]
        NG4 = [Ey[(0, 0)], 0.0, Ey[(1, 0)], 0.0]
        F5 = [-Ec[c] for c in cells]
        G4 = [Ec[c] for c in cells]
        NF0 = [mx[0], mx[1], 0.0, 0.0]
        NG0 = [my[0], 0.0, my[1], 0.0]
        got = orc.ct_contact_emf(NF5, NG4, F5, G4, NF0, NG0)
        want = _gs05_corner_emf(Ex, Ey, Ec, mx, my)
        assert abs(got - want) <= 1e-14 * (1 + abs(want)), (t, got, want)


def test_uniform_state_stays_uniform():
    nx, ny = 24, 20
    U = _uniform_state(nx, ny)
    o = OracleSim(nx, ny, 0.1, 0.1, GAMMA, 0.7, EPS, 1e-3)
    o.set_state(U)
    o.apply_boundary()
    o.step(5)
    V = o.get_state()
    assert np.array_equal(V[3:-3, 3:-3], U[3:-3, 3:-3])


def test_stencil_extent_one_step():
    # a perturbation of one interior cell reaches at most 6 cells away in
    # one RK2 step (3 per stage: MUSCL + face flux + CT/resistive corner).
    # eta is large enough that dt is the diffusion limit, which does not
    # depend on the state (otherwise the perturbed cell could move the
    # global CFL minimum and change every cell).
    cfg, o = _ot_state(40)
    cfg.eta = 2.0
    U0 = o.get_state()
    o = OracleSim.from_config(cfg)
    o.set_state(U0)
    o.step(1)
    A = o.get_state()
    U1 = U0.copy()
    ic, jc = 22, 21
    U1[ic, jc, 0] *= 1.0 + 1e-6
    o2 = OracleSim.from_config(cfg)
    o2.set_state(U1)
    o2.step(1)
    B = o2.get_state()
    changed = np.argwhere(np.any(A[3:-3, 3:-3] != B[3:-3, 3:-3], axis=2)) + 3
    assert len(changed) > 0
    assert np.abs(changed[:, 0] - ic).max() <= 6
    assert np.abs(changed[:, 1] - jc).max() <= 6


# -------------------------------------------------------------- boundary
def test_periodic_boundary_copies():
    nx, ny = 18, 16
    rng = np.random.default_rng(2)
    U = rng.standard_normal((nx, ny, 8))
    o = OracleSim(nx, ny, 0.1, 0.1)
    o.set_state(U)
    o.apply_boundary()
    V = o.get_state()
    it = V[3:nx - 3, 3:ny - 3]
    # every ghost cell equals the interior cell one period away (x first,
    # then y over full columns, which fills the corners)
    for i in range(nx):
        for j in range(ny):
            ii = (i - 3) % (nx - 6)
            jj = (j - 3) % (ny - 6)
            assert np.array_equal(V[i, j], it[ii, jj])


def test_symmetric_boundary_mirror_and_divfree_ghosts():
    nx, ny, dx, dy = 20, 18, 0.1, 0.07
    rng = np.random.default_rng(4)
    U = rng.standard_normal((nx, ny, 8))
    o = OracleSim(nx, ny, dx, dy, bc_x=BC_SYMMETRIC, bc_y=BC_SYMMETRIC)
    o.set_state(U)
    o.apply_boundary_axis(0, BC_SYMMETRIC)
    V = o.get_state()
    # mirror of everything but Bx (normal momentum included, no sign flip)
    for v in (0, 1, 2, 3, 5, 6, 7):
        for g, s in ((0, 5), (1, 4), (2, 3), (nx - 1, nx - 6), (nx - 2, nx - 5), (nx - 3, nx - 4)):
            assert np.array_equal(V[g, :, v], V[s, :, v])
    # rebuilt ghost Bx makes the ghost cells divergence-free
    bx, by = V[:, :, 4], V[:, :, 5]
    for i in (1, 2, 3, nx - 3, nx - 2, nx - 1):
        div = (bx[i, 1:] - bx[i - 1, 1:]) / dx + (by[i, 1:] - by[i, :-1]) / dy
        assert np.abs(div).max() <= 1e-12 * (1 + np.abs(by).max() / dy)


# ------------------------------------- multi-GPU library step on one rank
from oracle.oracle import SCHEME_MULTI_SHARE


def _ot_multi(n, steps):
    cfg, o = _ot_state(n)
    o.set_scheme(SCHEME_MULTI_SHARE)
    return cfg, o


def test_multi_halo_equals_periodic_fill():
    # sendrecv_U to itself (M/mpi.cu:104-188) is the periodic fill
    nx, ny = 20, 17
    rng = np.random.default_rng(8)
    U = rng.standard_normal((nx, ny, 8))
    a = OracleSim(nx, ny, 0.1, 0.1)
    a.set_state(U)
    a.apply_boundary()
    b = OracleSim(nx, ny, 0.1, 0.1)
    b.set_state(U)
    b.apply_halo()
    assert np.array_equal(a.get_state(), b.get_state())


def test_multi_conservation_and_divb():
    cfg, o = _ot_multi(40, 0)
    U0 = o.get_state()
    d0 = _divb(U0, cfg.dx, cfg.dy)
    o.step(20)
    U1 = o.get_state()
    for v in (0, 1, 2, 3, 7):
        scale = np.abs(U0[3:-3, 3:-3, v]).sum()
        assert abs(U1[3:-3, 3:-3, v].sum() - U0[3:-3, 3:-3, v].sum()) <= 1e-12 * scale
    d1 = _divb(U1, cfg.dx, cfg.dy)
    assert np.abs(d1 - d0).max() <= 1e-11
    assert not o.crashed()


def test_multi_dt_interior_only_no_diffusion_limit():
    nx, ny, dx, dy = 20, 16, 0.1, 0.05
    U = _uniform_state(nx, ny)
    U[0, 0] = _uniform_state(1, 1, rho=0.01)[0, 0]    # a fast ghost cell is ignored
    o = OracleSim(nx, ny, dx, dy, GAMMA, 0.7, EPS, 10.0)
    o.set_scheme(SCHEME_MULTI_SHARE)
    o.set_state(U)
    assert o.calculate_dt() == pytest.approx(0.7 * orc.cell_dt(U[5, 5], GAMMA, dx, dy), rel=1e-15)


def test_multi_uniform_state_stays_uniform():
    nx, ny = 24, 20
    U = _uniform_state(nx, ny)
    o = OracleSim(nx, ny, 0.1, 0.1, GAMMA, 0.7, EPS, 1e-3)
    o.set_scheme(SCHEME_MULTI_SHARE)
    o.set_state(U)
    o.apply_halo()
    o.step(4)
    assert np.array_equal(o.get_state()[3:-3, 3:-3], U[3:-3, 3:-3])


def test_multi_scheme_rejects_symmetric_and_eta_field():
    o = OracleSim(20, 20, 0.1, 0.1, bc_y=BC_SYMMETRIC)
    with pytest.raises(ValueError):
        o.set_scheme(SCHEME_MULTI_SHARE)


# ------------------- the MPI library's getEta at its three sampling positions
def _periodic_planes(nx, ny, seed, g=3):
    """Three eta planes (x-face, y-face, cell) of smooth random values on the
    interior whose ghost entries are their periodic images, as a periodic
    getEta sampled at every stored index gives on one rank."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        I = 1e-3 * (1.0 + 0.5 * rng.uniform(-1, 1, (nx - 2 * g, ny - 2 * g)))
        out.append(np.pad(I, g, mode="wrap"))
    return out


def test_multi_eta_planes_of_the_uniform_value_are_the_uniform_step():
    cfg, o = _ot_multi(24, 0)
    U0 = o.get_state()
    o.step(3)
    A = o.get_state()
    b = OracleSim.from_config(cfg)
    b.set_scheme(SCHEME_MULTI_SHARE)
    b.set_state(U0)
    full = np.full((cfg.nx, cfg.ny), cfg.eta)
    b.set_eta_staggered(full, full, full)
    b.step(3)
    assert np.array_equal(b.get_state(), A)


def test_multi_resistive_flux_reads_eta_at_its_face():
    """By = a x (centred By linear), Bx = 0, at rest: the MPI library's
    face current is J_z = a exactly and the HLLD part of f5 and of G's f4
    vanishes, so F.f5 = -a eta_x(i, j) on x-face (i + 1/2, j) and G.f4 =
    a eta_y(i, j) on y-face (i, j + 1/2), each from its own plane at its own
    index (M/flux_solver.cu:84-96,168-181)."""
    nx, ny, dx, dy, a = 20, 18, 0.1, 0.1, 0.8
    U = _uniform_state(nx, ny, u=0.0, v=0.0, w=0.0, bx=0.0, by=0.0, bz=0.0, p=1.0)
    x = (np.arange(nx) - 2.5) * dx
    U[:, :, 5] = a * x[:, None]
    U[:, :, 7] += 0.5 * U[:, :, 5] ** 2
    o = OracleSim(nx, ny, dx, dy, GAMMA, 0.7, EPS, 0.0)
    o.set_scheme(SCHEME_MULTI_SHARE)
    rng = np.random.default_rng(4)
    ex, ey, ec = (1e-3 * (1.0 + rng.uniform(size=(nx, ny))) for _ in range(3))
    o.set_eta_staggered(ex, ey, ec)
    F, G = o.compute_fluxes(U)
    i, j = slice(3, nx - 4), slice(3, ny - 4)
    np.testing.assert_allclose(F[i, j, 5], -a * ex[i, j], rtol=1e-12)
    np.testing.assert_allclose(G[i, j, 4], a * ey[i, j], rtol=1e-12)


def test_multi_cell_eta_enters_the_ct_at_its_cell():
    """The cell plane enters only the cell EMF of the CT-contact average
    (M/ct.cu:26-65).  Where every face's mass flux is positive, the corner
    EMF (i + 1/2, j + 1/2) of M/ct.cu:147-169 reads the cell EMFs with
    weights 1/2 (cell i, j), 1/4 (i, j+1) and 1/4 (i+1, j), so raising eta at
    one cell (i0, j0) changes the corners (i0, j0), (i0, j0-1), (i0-1, j0) by
    q/2, q/4, q/4 (q the same for all three), and after one step exactly
    these faces, in these ratios (dx = dy): Bx(i0, j0-1), Bx(i0, j0),
    Bx(i0-1, j0) by -q/4, Bx(i0-1, j0+1) by q/4, Bx(i0, j0+1) by q/2 (times
    dt/dy), and By(i0-1, j0), By(i0, j0-1), By(i0, j0) by q/4, By(i0+1, j0-1)
    by -q/4, By(i0+1, j0) by -q/2 (times dt/dx).  The cell plane read one
    cell off, or at the transposed index, moves the pattern."""
    import resistivemhd_gpu_amd as mhd
    cfg = mhd.turbulence_config(24, eta=2e-3)
    U = mhd.fill_turbulence(cfg, seed=7, kmax=3, rms_v=0.05, rms_b=0.4)
    rho = U[..., 0]
    ke0 = 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2) / rho
    U[..., 1] += 0.8 * rho
    U[..., 2] += 0.6 * rho
    U[..., 7] += 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2) / rho - ke0
    o = OracleSim.from_config(cfg)
    o.set_state(U)
    o.apply_boundary()
    U = o.get_state()
    base = [np.full((cfg.nx, cfg.ny), cfg.eta) for _ in range(3)]
    runs = []
    i0, j0 = 11, 13
    for bump in (0.0, 2e-2):
        planes = [p.copy() for p in base]
        planes[2][i0, j0] += bump
        o = OracleSim.from_config(cfg)
        o.set_scheme(SCHEME_MULTI_SHARE)
        o.set_fixed_dt(0.02)
        o.set_eta_staggered(*planes)
        F, G = o.compute_fluxes(U)   # the stage-1 fluxes: mass flux positive on every face
        assert np.all(F[2:-3, 2:-3, 0] > 0) and np.all(G[2:-3, 2:-3, 0] > 0)
        o.set_state(U)
        o.step(1)
        runs.append(o.get_state())
    D = runs[1] - runs[0]
    assert np.array_equal(D[..., [0, 1, 2, 3, 6, 7]], np.zeros_like(D[..., :6]))
    dbx, dby = D[..., 4], D[..., 5]
    r = dbx[i0, j0]            # -q/4 dt/dy
    assert abs(r) > 1e-6
    want_bx = {(i0, j0 - 1): 1.0, (i0, j0): 1.0, (i0 - 1, j0): 1.0, (i0 - 1, j0 + 1): -1.0, (i0, j0 + 1): -2.0}
    want_by = {(i0 - 1, j0): -1.0, (i0, j0 - 1): -1.0, (i0, j0): -1.0, (i0 + 1, j0 - 1): 1.0, (i0 + 1, j0): 2.0}
    for plane, want in ((dbx, want_bx), (dby, want_by)):
        got = {(int(a), int(b)) for a, b in np.argwhere(plane != 0)}
        assert got == set(want), (sorted(got), sorted(want))
        for (i, j), k in want.items():
            assert abs(plane[i, j] - k * r) <= 1e-9 * abs(r), (i, j, plane[i, j], k * r)


# ------------------------------------- rotation by 90 degrees (both schemes)
def _rotate(U, g=3):
    """The state rotated by +90 degrees about the grid centre, (x, y) ->
    (-y, x), on the interior n x n cells of a periodic grid: cell (i', j')
    holds old cell (j', n-1-i'); velocities and B turn with the frame,
    (u, v) -> (-v, u); face fields move to the faces they land on (the new
    x-face Bx' is minus the old y-face By one row down, the new y-face By'
    is the old x-face Bx).  Ghosts are left to the boundary fill."""
    I = U[g:-g, g:-g]

    def rot(a):
        return a[:, ::-1].T

    out = np.zeros_like(U)
    R = out[g:-g, g:-g]
    R[..., 0] = rot(I[..., 0])
    R[..., 1] = -rot(I[..., 2])
    R[..., 2] = rot(I[..., 1])
    R[..., 3] = rot(I[..., 3])
    R[..., 4] = -rot(np.roll(I[..., 5], 1, axis=1))
    R[..., 5] = rot(I[..., 4])
    R[..., 6] = rot(I[..., 6])
    R[..., 7] = rot(I[..., 7])
    return out


def _generic_periodic_state(n, seed):
    import resistivemhd_gpu_amd as mhd
    cfg = mhd.turbulence_config(n, eta=2e-3)
    U = mhd.fill_turbulence(cfg, seed=seed, kmax=4, rms_v=0.6, rms_b=0.6)
    rng = np.random.default_rng(seed)
    U[..., 0] *= 1.0 + 0.2 * rng.uniform(-1, 1, U.shape[:2])
    U[..., 3] += 0.2 * rng.normal(size=U.shape[:2])
    U[..., 6] += 0.2 * rng.normal(size=U.shape[:2])
    U[..., 7] += 1.0 + 0.3 * rng.uniform(size=U.shape[:2])
    return cfg, U


def test_rotate_helper_is_a_quarter_turn():
    """Four quarter turns give the identity, and a quarter turn keeps the
    discrete divergence of B a rotated scalar field."""
    cfg, U = _generic_periodic_state(24, 3)
    o = OracleSim.from_config(cfg)
    o.set_state(U)
    o.apply_boundary()
    U = o.get_state()
    R = U
    for _ in range(4):
        o.set_state(_rotate(R))
        o.apply_boundary()
        R = o.get_state()
    assert np.array_equal(R[3:-3, 3:-3], U[3:-3, 3:-3])
    o.set_state(_rotate(U))
    o.apply_boundary()
    R1 = o.get_state()
    d0, d1 = _divb(U, cfg.dx, cfg.dy), _divb(R1, cfg.dx, cfg.dy)
    np.testing.assert_allclose(d1, d0[:, ::-1].T, atol=1e-12)


def _run_fixed_dt(cfg, state, scheme, nsteps, dt=0.02):
    o = OracleSim.from_config(cfg)
    if scheme:
        o.set_scheme(scheme)
    o.set_fixed_dt(dt)
    o.set_state(state)
    o.apply_boundary()
    o.step(nsteps)
    assert not o.crashed()
    return o.get_state()


def _assert_commutes(A, B, tol=1e-11):
    for v in range(8):
        a, b = A[3:-3, 3:-3, v], B[3:-3, 3:-3, v]
        scale = np.abs(a).max()
        assert np.abs(a - b).max() <= tol * scale, (v, np.abs(a - b).max(), scale)


def _transpose(U):
    """The state mirrored across the diagonal x = y: cell (i, j) -> (j, i),
    (u, v) -> (v, u), face Bx -> face By and back (with B -> -B, which the
    equations allow, the axial field keeps the same components)."""
    T = np.transpose(U, (1, 0, 2)).copy()
    T[..., 1], T[..., 2] = U[..., 2].T.copy(), U[..., 1].T.copy()
    T[..., 4], T[..., 5] = U[..., 5].T.copy(), U[..., 4].T.copy()
    return T


@pytest.mark.parametrize("scheme", [0, SCHEME_MULTI_SHARE])
def test_step_commutes_with_transpose(scheme):
    """step(transpose(U)) = transpose(step(U)) to round-off on a generic
    periodic state, dt held fixed (the reference's cell dt reads the stored
    face B as if it were centred, which no symmetry maps onto itself).
    Pins the y sweep against the x sweep (rotation into the x frame and
    back, the resistive stencils of both directions) and, for the MPI
    library's step, the CT-contact EMF: its S/N terms map onto its W/E terms,
    so an index, upwind-flag or sign slip in any one of them breaks the
    symmetry (checked by planting each in a copy of the oracle: 4e-5 to 3e-2
    against 1e-15)."""
    cfg, U = _generic_periodic_state(32, 5)
    o = OracleSim.from_config(cfg)
    o.set_state(U)
    o.apply_boundary()
    U = o.get_state()
    _assert_commutes(_transpose(_run_fixed_dt(cfg, U, scheme, 3)), _run_fixed_dt(cfg, _transpose(U), scheme, 3))


def test_multi_step_with_eta_planes_commutes_with_transpose():
    """The transpose maps an x-face ((i + 1/2) dx, j dy) onto the y-face
    (j dx, (i + 1/2) dy) and a cell onto a cell, so the step with eta
    planes (x-face, y-face, cell) commutes with it when the transposed run
    takes the planes (cell^T ... ): x-face' = y-face^T, y-face' = x-face^T,
    cell' = cell^T.  A plane read at the other face direction, or the cell
    plane read one cell off along one axis, breaks the symmetry."""
    cfg, U = _generic_periodic_state(32, 5)
    o = OracleSim.from_config(cfg)
    o.set_state(U)
    o.apply_boundary()
    U = o.get_state()
    ex, ey, ec = _periodic_planes(cfg.nx, cfg.ny, 6)

    def run(state, planes):
        r = OracleSim.from_config(cfg)
        r.set_scheme(SCHEME_MULTI_SHARE)
        r.set_fixed_dt(0.02)
        r.set_eta_staggered(*planes)
        r.set_state(state)
        r.apply_boundary()
        r.step(3)
        return r.get_state()

    A = _transpose(run(U, (ex, ey, ec)))
    _assert_commutes(A, run(_transpose(U), (ey.T, ex.T, ec.T)))
    # the planes matter: the same run with the face planes swapped differs
    B = _transpose(run(U, (ey, ex, ec)))
    assert np.abs(B[3:-3, 3:-3, 7] - A[3:-3, 3:-3, 7]).max() > 1e-9 * np.abs(A[3:-3, 3:-3, 7]).max()


def test_step_commutes_with_quarter_turn():
    """The single-GPU step also commutes with a quarter turn.  The MPI
    library's step does not, by construction: its stages advance the face
    fields by flux differences taken half a cell to the left (Bx) and below
    (By) the face (M/resistiveMHD2D.cu:64-100), offsets that a
    transpose keeps and a quarter turn does not."""
    cfg, U = _generic_periodic_state(32, 5)
    o = OracleSim.from_config(cfg)
    o.set_state(U)
    o.apply_boundary()
    U = o.get_state()
    A = _rotate(_run_fixed_dt(cfg, U, 0, 3))
    _assert_commutes(A, _run_fixed_dt(cfg, _rotate(U), 0, 3))
    # a wrong turn (velocity not turned) does not commute
    bad = _rotate(U)
    bad[..., 1], bad[..., 2] = bad[..., 2].copy(), bad[..., 1].copy()
    C = _run_fixed_dt(cfg, bad, 0, 3)
    assert np.abs(C[3:-3, 3:-3, 1] - A[3:-3, 3:-3, 1]).max() > 1e-3 * np.abs(A[3:-3, 3:-3, 1]).max()
