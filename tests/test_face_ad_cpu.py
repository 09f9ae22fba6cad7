"""CPU checks of the per-face flux pieces the adjoint is built from
(resistivemhd_gpu_amd/csrc/mhd2d_face_ad.h, generated from the *.ssa
listings, and the assembly in mhd2d_adjoint_face.h):

- the generated primal flux equals the oracle's F and G bitwise on every
  face the step reads (same operations in the same order);
- tangent-linear against central finite differences;
- adjoint against tangent-linear by the dot-product identity
  <J v, w> = <v, J^T w>, for the generated pieces and for the assembled
  face (gather, y rotation, resistive stencil, eta, centring).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from oracle.oracle import OracleSim
from tests.helpers import boundary_applied

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "resistivemhd_gpu_amd", "csrc")
D = ctypes.POINTER(ctypes.c_double)


def generate(out):
    subprocess.run(["python3", os.path.join(CSRC, "adjoint", "gen_face_ad.py"), out,
                    os.path.join(CSRC, "adjoint", "face_flux.ssa:face_flux"),
                    os.path.join(CSRC, "adjoint", "resistive_x.ssa:resist_x"),
                    os.path.join(CSRC, "adjoint", "resistive_y.ssa:resist_y")], check=True)


def test_generated_header_is_current(tmp_path):
    """mhd2d_face_ad.h is what the generator makes from the listings."""
    out = tmp_path / "face_ad.h"
    generate(str(out))
    assert out.read_text() == open(os.path.join(CSRC, "mhd2d_face_ad.h")).read()


@pytest.fixture(scope="module")
def ad(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("ad") / "libad.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I", CSRC,
                    os.path.join(ROOT, "tests", "ad_harness.cpp"), "-o", so], check=True)
    return ctypes.CDLL(so)


def p(a):
    return a.ctypes.data_as(D) if a is not None else None


def par_of(cfg, dtol=0.0):
    return np.array([cfg.gamma - 1.0, cfg.gamma, cfg.eps, dtol]), np.array([cfg.dx, cfg.dy])


def random_face_inputs(rng):
    """Centred states of 4 cells: positive density and pressure."""
    cells = []
    g = 5.0 / 3.0
    for _ in range(4):
        rho = rng.uniform(0.5, 2.0)
        u, v, w = rng.normal(0, 0.5, 3)
        bx, by, bz = rng.normal(0, 0.7, 3)
        pr = rng.uniform(0.3, 2.0)
        e = pr / (g - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
        cells += [rho, rho * u, rho * v, rho * w, bx, by, bz, e]
    return np.array(cells)


def noisy_ot(n, seed):
    cfg = mhd.orszag_tang_config(n, eta=1e-2)
    U = mhd.fill_orszag_tang(cfg)
    rng = np.random.default_rng(seed)
    U[..., 1:4] += rng.normal(0, 0.3, U[..., 1:4].shape)
    U[..., 4:7] += rng.normal(0, 0.3, U[..., 4:7].shape)
    U[..., 7] += 2.0
    return cfg, boundary_applied(cfg, U)


def face_ranges(cfg, d):
    """Faces the RK step reads (stage updates of interior cells and CT)."""
    if d == 0:
        return range(2, cfg.nx - 3), range(2, cfg.ny - 2)
    return range(2, cfg.nx - 2), range(2, cfg.ny - 3)


@pytest.mark.parametrize("case", ["ot", "harris"])
def test_primal_equals_oracle_fluxes(ad, case):
    if case == "ot":
        cfg, U = noisy_ot(20, 1)
        eta = None
    else:
        cfg = mhd.harris_config(24, 18)
        U = boundary_applied(cfg, mhd.fill_harris(cfg, seed=2, noise_amp=5e-2))
        eta = mhd.fill_harris_eta(cfg)
    o = OracleSim.from_config(cfg)
    if eta is not None:
        o.set_eta_field(eta)
    F, G = o.compute_fluxes(U)
    par, rpar = par_of(cfg)
    out = np.empty(8)
    for d, ref in ((0, F), (1, G)):
        ri, rj = face_ranges(cfg, d)
        for i in ri:
            for j in rj:
                ad.h_flux_full(p(U), p(eta), ctypes.c_double(cfg.eta), cfg.nx, cfg.ny, d, i, j, p(par), p(rpar),
                               p(out))
                assert np.array_equal(out, ref[i, j]), (d, i, j, out - ref[i, j])


def test_face_flux_tangent_matches_finite_differences(ad):
    rng = np.random.default_rng(11)
    par = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20, 0.0])
    errs = []
    for _ in range(200):
        x = random_face_inputs(rng)
        dx = rng.normal(size=32)
        out, out_d = np.empty(7), np.empty(7)
        ad.h_face_flux_tlm(p(x), p(dx), p(par), p(out), p(out_d))
        h = 1e-6
        fp, fm = np.empty(7), np.empty(7)
        ad.h_face_flux(p(x + h * dx), p(par), p(fp))
        ad.h_face_flux(p(x - h * dx), p(par), p(fm))
        fd = (fp - fm) / (2 * h)
        errs.append(np.abs(fd - out_d).max() / max(1.0, np.abs(out_d).max()))
    errs = np.array(errs)
    # a handful of draws straddle a region switch or a minmod kink within +-h
    assert np.median(errs) < 1e-7
    assert np.mean(errs < 1e-5) > 0.97


def test_face_flux_adjoint_dot_product(ad):
    rng = np.random.default_rng(12)
    par = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20, 0.0])
    for _ in range(200):
        x = random_face_inputs(rng)
        v = rng.normal(size=32)
        w = rng.normal(size=7)
        out, out_d = np.empty(7), np.empty(7)
        ad.h_face_flux_tlm(p(x), p(v), p(par), p(out), p(out_d))
        xb = np.empty(32)
        ad.h_face_flux_adj(p(x), p(par), p(w), p(xb))
        lhs, rhs = out_d @ w, v @ xb
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs)), (lhs, rhs)


@pytest.mark.parametrize("d", [0, 1])
def test_resistive_tangent_and_adjoint(ad, d):
    rng = np.random.default_rng(13 + d)
    rpar = np.array([0.05, 0.07])
    for _ in range(50):
        x = rng.normal(size=14)
        x[12:] = rng.uniform(1e-4, 1e-2, 2)
        v, w = rng.normal(size=14), rng.normal(size=3)
        out, out_d = np.empty(3), np.empty(3)
        ad.h_resist_tlm(d, p(x), p(v), p(rpar), p(out), p(out_d))
        h = 1e-6
        fp, fm, tmp = np.empty(3), np.empty(3), np.empty(3)
        ad.h_resist_tlm(d, p(x + h * v), p(v), p(rpar), p(fp), p(tmp))
        ad.h_resist_tlm(d, p(x - h * v), p(v), p(rpar), p(fm), p(tmp))
        np.testing.assert_allclose((fp - fm) / (2 * h), out_d, rtol=1e-6, atol=1e-9)
        xb = np.empty(14)
        ad.h_resist_adj(d, p(x), p(rpar), p(w), p(xb))
        assert abs(out_d @ w - v @ xb) <= 1e-12 * max(1.0, abs(out_d @ w))


def shift_transpose(cfg, lam_c):
    """Adjoint of the face-to-centre averaging of Bx (along i) and By (along j)."""
    lam = lam_c.copy()
    bx = np.zeros_like(lam_c[..., 4])
    bx[0] = lam_c[0, :, 4]
    bx[1:] += 0.5 * lam_c[1:, :, 4]
    bx[:-1] += 0.5 * lam_c[1:, :, 4]
    by = np.zeros_like(lam_c[..., 5])
    by[:, 0] = lam_c[:, 0, 5]
    by[:, 1:] += 0.5 * lam_c[:, 1:, 5]
    by[:, :-1] += 0.5 * lam_c[:, 1:, 5]
    lam[..., 4], lam[..., 5] = bx, by
    return lam


@pytest.mark.parametrize("d", [0, 1])
def test_assembled_face_adjoint_dot_product(ad, d):
    """Gather, rotation, resistive stencil, eta and centring: the assembled
    adjoint is the transpose of the assembled tangent."""
    cfg, U = noisy_ot(14, 5)
    rng = np.random.default_rng(20 + d)
    eta = rng.uniform(1e-3, 1e-2, (cfg.nx, cfg.ny))
    par, rpar = par_of(cfg)
    ri, rj = face_ranges(cfg, d)
    for i in list(ri)[::3]:
        for j in list(rj)[::3]:
            dU = rng.normal(size=U.shape)
            deta = rng.normal(size=eta.shape) * 1e-3
            lamF = rng.normal(size=8)
            F, dF = np.empty(8), np.empty(8)
            ad.h_flux_full_tlm(p(U), p(dU), p(eta), p(deta), ctypes.c_double(0.0), cfg.nx, cfg.ny, d, i, j, p(par),
                               p(rpar), p(F), p(dF))
            lam_c = np.zeros_like(U)
            geta = np.zeros_like(eta)
            ad.h_flux_full_adj(p(U), p(eta), ctypes.c_double(0.0), cfg.nx, cfg.ny, d, i, j, p(par), p(rpar),
                               p(lamF), p(lam_c), p(geta))
            lhs = dF @ lamF
            rhs = (shift_transpose(cfg, lam_c) * dU).sum() + (geta * deta).sum()
            assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs)), (i, j, lhs, rhs)


def test_adjoint_branches_read_ieee_values():
    """The adjoint's own primal pass uses refined reciprocal and square-root
    estimates only where no branch looks: every name an HLLD region
    indicator, a guard or a min/max/sign selection reads is computed,
    transitively, with IEEE division and sqrt.  (S1L = SM - sqrt(bXL^2 /
    rho1L) decides between the outer and inner star regions; an estimate
    there could let the adjoint differentiate a region the forward step did
    not take.)"""
    import re
    text = open(os.path.join(CSRC, "mhd2d_face_ad.h")).read()
    body = text[text.index("inline void face_flux_adj("):]
    body = body[:body.index("\n}\n")]
    primal = body[:body.index("_b = 0.0;")]
    defs = {}
    for m in re.finditer(r"^\s*(?:const double )?(\w+) = (.*);$", primal, re.M):
        if m.group(2) != "0.0":
            defs[m.group(1)] = m.group(2)
    ident = re.compile(r"[A-Za-z_]\w*")
    roots = set()
    for name, expr in defs.items():
        if "<" in expr or ">" in expr:          # ind, tmin, tmax, sgn
            roots.update(ident.findall(expr))
    for g in re.findall(r"^\s*if \((.*)\) \{$", primal, re.M):
        roots.update(ident.findall(g))
    assert {"S1L", "S1R", "SM", "SL", "SR"} <= roots
    seen, todo, bad = set(), list(roots), []
    while todo:
        n = todo.pop()
        if n in seen or n not in defs:
            continue
        seen.add(n)
        if "mhd_ad_" in defs[n]:
            bad.append((n, defs[n]))
        todo += ident.findall(defs[n])
    assert not bad, bad


def listing_flux(x, par, limit=False):
    """face_flux.ssa run as Python statements (IEEE double, the listing's own
    evaluation order), independently of the generated C.  limit=True
    evaluates dlim(x, limit, c) as its limit wherever c holds."""
    import math
    ns = {"gm1": float(par[0]), "gamma": float(par[1]), "eps": float(par[2]), "dtol": float(par[3]),
          "sqrt": math.sqrt, "fabs": abs}
    ns["tmin"] = lambda a, b: b if b < a else a
    ns["tmax"] = lambda a, b: b if a < b else a
    ns["sgn"] = lambda a: float((a > 0.0) - (a < 0.0))
    ns["ind"] = lambda c: float(c)
    ns["minmod"] = lambda a, b: ns["sgn"](a) * ns["tmax"](ns["tmin"](abs(a), ns["sgn"](a) * b), ns["eps"])
    ns["dlim"] = (lambda v, lim, c: lim if c != 0.0 else v) if limit else (lambda v, lim, c: v)
    outputs = []
    for line in open(os.path.join(CSRC, "adjoint", "face_flux.ssa")):
        line = line.split("#", 1)[0].strip()
        if line.startswith("inputs:"):
            for k, name in enumerate(line.split(":", 1)[1].split()):
                ns[name] = float(x[k])
        elif line.startswith("outputs:"):
            outputs = line.split(":", 1)[1].split()
        elif line and not line.split("=")[0].strip().endswith(("params:", "guard:", "segment:")) \
                and ":" not in line.split("=")[0]:
            exec(line, ns)
    return np.array([ns[o] for o in outputs])


def test_listing_run_as_python_equals_generated_primal(ad):
    rng = np.random.default_rng(31)
    par = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20, 0.0])
    out = np.empty(7)
    for _ in range(50):
        x = random_face_inputs(rng)
        ad.h_face_flux(p(x), p(par), p(out))
        assert np.array_equal(out, listing_flux(x, par))


def degenerate_face():
    """Four cells whose face has degenerate left and right star states, as
    the Harris sheet's far field does: the normal field dominates (Bx^2 /
    rho = 5 > c_s^2), the transverse field reconstructs to exactly 0 on both
    sides, and rho, u and p reconstruct to equal values on both sides, so
    rho (S - u)(S - SM) - Bx^2 is zero up to rounding.  Every limiter is
    off its kink (B_t = -4t, -t, t, 4t: the middle difference is the
    smallest, and t = 2^-10 makes the reconstruction exactly 0)."""
    g = 5.0 / 3.0
    t = 2.0 ** -10
    rho = [0.2, 0.202, 0.203, 0.205]
    u = [0.0, 2e-4, 3e-4, 5e-4]
    v = [1e-3, -2e-3, 5e-4, 1.5e-3]
    w = [2e-4, -1e-4, 3e-4, 0.0]
    by = [-4 * t, -t, t, 4 * t]
    bz = [-8 * t, -2 * t, 2 * t, 8 * t]
    pr = [0.1, 0.101, 0.1015, 0.1025]
    cells = []
    for k in range(4):
        e = pr[k] / (g - 1) + 0.5 * rho[k] * (u[k] ** 2 + v[k] ** 2 + w[k] ** 2) + 0.5 * (1.0 + by[k] ** 2 + bz[k] ** 2)
        cells += [rho[k], rho[k] * u[k], rho[k] * v[k], rho[k] * w[k], -1.0, by[k], bz[k], e]
    return np.array(cells)


def test_degenerate_rule_matches_differences_of_the_limit(ad):
    """At a degenerate face the reference's star-state quotients are rounding
    noise: their central differences change by O(1) between steps h = 1e-5
    and 1e-7, so no derivative rule can match them.  With the rule on, the
    tangent is the derivative of the flux whose star tangential components
    are the outer ones (Miyoshi & Kusano's degenerate limit): central
    differences of the listing run as Python with that substitution are
    steady in h and agree with it."""
    x = degenerate_face()
    rng = np.random.default_rng(41)
    on = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20, 1e-8])
    off = on.copy()
    off[3] = 0.0
    out, d_on = np.empty(7), np.empty(7)
    ad.h_face_flux(p(x), p(on), p(out))
    assert np.array_equal(out, listing_flux(x, off))      # the value is the reference's
    assert np.array_equal(out, listing_flux(x, on, limit=True))   # and equals the limit's here

    def fd(par, limit, dx, h):
        return (listing_flux(x + h * dx, par, limit) - listing_flux(x - h * dx, par, limit)) / (2 * h)
    worst_on, ref_spread = 0.0, 0.0
    for _ in range(6):
        dx = rng.normal(size=32) * np.abs(x) + rng.normal(size=32) * 1e-7
        ad.h_face_flux_tlm(p(x), p(dx), p(on), p(out), p(d_on))
        lim = fd(on, True, dx, 1e-7)
        scale = np.abs(lim).max()
        worst_on = max(worst_on, np.abs(d_on - lim).max() / scale,
                       np.abs(fd(on, True, dx, 1e-5) - lim).max() / scale)
        ref_spread = max(ref_spread, np.abs(fd(off, False, dx, 1e-5) - fd(off, False, dx, 1e-7)).max() / scale)
    assert worst_on < 1e-7, worst_on
    assert ref_spread > 0.1, ref_spread


def test_degenerate_rule_is_inactive_on_generic_faces(ad):
    """Random faces are far from the degenerate set: the tangent and adjoint
    with the rule on equal those with it off bit for bit."""
    rng = np.random.default_rng(43)
    on = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20, 1e-8])
    off = on.copy()
    off[3] = 0.0
    o1, o2, d1, d2, b1, b2 = (np.empty(n) for n in (7, 7, 7, 7, 32, 32))
    for _ in range(100):
        x = random_face_inputs(rng)
        v, w = rng.normal(size=32), rng.normal(size=7)
        ad.h_face_flux_tlm(p(x), p(v), p(on), p(o1), p(d1))
        ad.h_face_flux_tlm(p(x), p(v), p(off), p(o2), p(d2))
        ad.h_face_flux_adj(p(x), p(on), p(w), p(b1))
        ad.h_face_flux_adj(p(x), p(off), p(w), p(b2))
        assert np.array_equal(d1, d2) and np.array_equal(b1, b2)


def test_degenerate_rule_adjoint_is_transpose_of_tangent(ad):
    """Dot-product identity with the rule on, at the degenerate face and at
    faces of the Harris sheet start (where the far field is degenerate)."""
    rng = np.random.default_rng(47)
    on = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20, 1e-8])
    xs = [degenerate_face()]
    cfg = mhd.harris_config(64, 32)
    U = boundary_applied(cfg, mhd.fill_harris(cfg, seed=3, noise_amp=1e-3))
    inp = np.empty(32)
    for i, j in ((10, 3), (30, 28), (50, 5)):
        for c in range(4):
            inp[8 * c:8 * c + 8] = U[i - 1 + c, j]
        inp[4::8] = 0.5 * (U[i - 2:i + 2, j, 4] + U[i - 1:i + 3, j, 4])
        xs.append(inp.copy())
    out, out_d, xb = np.empty(7), np.empty(7), np.empty(32)
    for x in xs:
        for _ in range(5):
            v, w = rng.normal(size=32), rng.normal(size=7)
            ad.h_face_flux_tlm(p(x), p(v), p(on), p(out), p(out_d))
            ad.h_face_flux_adj(p(x), p(on), p(w), p(xb))
            lhs, rhs = out_d @ w, v @ xb
            assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs), np.abs(out_d).max() * np.abs(w).max()), (lhs, rhs)


def test_replay_equals_primal_at_the_base_inputs(ad):
    """The decision replay evaluated at the base inputs is the primal."""
    rng = np.random.default_rng(51)
    par = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20, 1e-8])
    out, out_p, ref = np.empty(7), np.empty(7), np.empty(7)
    for x in [degenerate_face()] + [random_face_inputs(rng) for _ in range(50)]:
        ad.h_face_flux_rep(p(x), p(x), p(par), p(out), p(out_p))
        ad.h_face_flux(p(x), p(par), p(ref))
        assert np.array_equal(out, ref) and np.array_equal(out_p, ref)


@pytest.mark.parametrize("dtol", [0.0, 1e-8])
def test_replay_differences_match_tangent_on_every_face(ad, dtol):
    """Central differences of the decision replay (branches frozen at the
    base inputs) agree with the tangent on every draw, with no exception for
    draws that straddle a region switch or a limiter kink (those are what
    break plain differences, test_face_flux_tangent_matches_finite_differences),
    and at the degenerate face with the limit rule on."""
    rng = np.random.default_rng(53)
    par = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20, dtol])
    xs = [random_face_inputs(rng) for _ in range(300)]
    if dtol > 0.0:
        xs.append(degenerate_face())
    out, out_d, fp, fm, tmp = (np.empty(7) for _ in range(5))
    worst = 0.0
    for x in xs:
        dx = rng.normal(size=32) * np.maximum(np.abs(x), 1e-3)
        ad.h_face_flux_tlm(p(x), p(dx), p(par), p(out), p(out_d))
        h = 1e-6
        ad.h_face_flux_rep(p(x), p(x + h * dx), p(par), p(tmp), p(fp))
        ad.h_face_flux_rep(p(x), p(x - h * dx), p(par), p(tmp), p(fm))
        fd = (fp - fm) / (2 * h)
        worst = max(worst, np.abs(fd - out_d).max() / max(1.0, np.abs(out_d).max()))
    assert worst < 1e-6, worst


@pytest.mark.parametrize("d", [0, 1])
def test_assembled_replay_differences_match_tangent(ad, d):
    """The same for the assembled face (gather, rotation, resistive stencil
    with a perturbed eta, centring)."""
    cfg, U = noisy_ot(14, 7)
    rng = np.random.default_rng(60 + d)
    eta = rng.uniform(1e-3, 1e-2, (cfg.nx, cfg.ny))
    par, rpar = par_of(cfg)
    ri, rj = face_ranges(cfg, d)
    F, dF, FP, FM, T = (np.empty(8) for _ in range(5))
    worst = 0.0
    for i in list(ri)[::2]:
        for j in list(rj)[::2]:
            dU = rng.normal(size=U.shape) * 0.1
            deta = rng.normal(size=eta.shape) * 1e-3
            ad.h_flux_full_tlm(p(U), p(dU), p(eta), p(deta), ctypes.c_double(0.0), cfg.nx, cfg.ny, d, i, j, p(par),
                               p(rpar), p(F), p(dF))
            h = 1e-6
            ad.h_flux_full_rep(p(U), p(U + h * dU), p(eta), p(eta + h * deta), ctypes.c_double(0.0), cfg.nx, cfg.ny,
                               d, i, j, p(par), p(rpar), p(T), p(FP))
            ad.h_flux_full_rep(p(U), p(U - h * dU), p(eta), p(eta - h * deta), ctypes.c_double(0.0), cfg.nx, cfg.ny,
                               d, i, j, p(par), p(rpar), p(T), p(FM))
            worst = max(worst, np.abs((FP - FM) / (2 * h) - dF).max() / max(1.0, np.abs(dF).max()))
    assert worst < 1e-6, worst
