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


def par_of(cfg):
    return np.array([cfg.gamma - 1.0, cfg.gamma, cfg.eps]), np.array([cfg.dx, cfg.dy])


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
    par = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20])
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
    par = np.array([2.0 / 3.0, 5.0 / 3.0, 1e-20])
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
