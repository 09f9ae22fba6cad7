"""CPU-side checks of the C ABI (no GPU calls): the library loads, exports
every entry point include/mhd2d.h declares, and its host-only setup
functions produce the documented initial states."""
import ctypes
import os
import re

import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from resistivemhd_gpu_amd import mhd2d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "mhd2d.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(mhd_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert declared_entry_points() == sorted(mhd2d.EXPORTS)


def test_library_exports_every_declared_symbol(mhdlib):
    for name in declared_entry_points():
        assert hasattr(mhdlib, name), name


def test_default_config():
    cfg = mhd.default_config(256, 128, 0.0, 2.0, -1.0, 1.0)
    assert (cfg.nx, cfg.ny) == (262, 134)
    assert cfg.dx == pytest.approx(2.0 / 256) and cfg.dy == pytest.approx(1.0 / 128)
    assert cfg.gamma == pytest.approx(5.0 / 3.0) and cfg.cfl == 0.7 and cfg.eps == 1e-20
    assert (cfg.bc_x, cfg.bc_y) == (mhd.MHD_BC_PERIODIC, mhd.MHD_BC_PERIODIC)
    bad = mhd2d.MhdConfig()
    rc = mhdlib_call("mhd_default_config", ctypes.byref(bad), 2, 2, 0.0, 1.0, 0.0, 1.0)
    assert rc != 0


def mhdlib_call(name, *args):
    return getattr(mhd.load_library(), name)(*args)


def _centres(cfg):
    x = cfg.xmin + (np.arange(cfg.nx) - 3 + 0.5) * cfg.dx
    y = cfg.ymin + (np.arange(cfg.ny) - 3 + 0.5) * cfg.dy
    return np.meshgrid(x, y, indexing="ij")


def test_orszag_tang_setup():
    cfg = mhd.orszag_tang_config(32)
    U = mhd.fill_orszag_tang(cfg)
    X, Y = _centres(cfg)
    g = 5.0 / 3.0
    rho, u, v = g * g, -np.sin(Y), np.sin(X)
    bx, by = -np.sin(Y), np.sin(2 * X)
    e = g / (g - 1) + 0.5 * rho * (u * u + v * v) + 0.5 * (bx * bx + by * by)
    np.testing.assert_allclose(U[..., 0], rho, rtol=1e-15)
    np.testing.assert_allclose(U[..., 1], rho * u, atol=1e-14)
    np.testing.assert_allclose(U[..., 2], rho * v, atol=1e-14)
    np.testing.assert_allclose(U[..., 4], bx, atol=1e-15)
    np.testing.assert_allclose(U[..., 5], by, atol=1e-15)
    np.testing.assert_allclose(U[..., 7], e, rtol=1e-14)
    # face fields are discretely divergence-free
    d = (U[1:, 1:, 4] - U[:-1, 1:, 4]) / cfg.dx + (U[1:, 1:, 5] - U[1:, :-1, 5]) / cfg.dy
    assert np.abs(d).max() < 1e-12


def test_harris_setup_and_eta():
    cfg = mhd.harris_config(64, 32)
    U = mhd.fill_harris(cfg, seed=9, noise_amp=1e-3)
    X, Y = _centres(cfg)
    rho = 1.0 / np.cosh(Y) ** 2 + 0.2
    np.testing.assert_allclose(U[..., 0], rho, rtol=1e-14)
    np.testing.assert_allclose(U[..., 4], np.tanh(Y), rtol=1e-14, atol=1e-16)
    vy = U[..., 2] / U[..., 0]
    assert np.abs(vy).max() <= 1e-3 and np.abs(vy).max() > 9e-4
    assert abs(vy.mean()) < 1e-4                       # white noise, zero mean
    e = 0.5 * rho / (2.0 / 3.0) + 0.5 * rho * vy ** 2 + 0.5 * np.tanh(Y) ** 2
    np.testing.assert_allclose(U[..., 7], e, rtol=1e-13)
    # reproducible and seed-dependent
    assert np.array_equal(U, mhd.fill_harris(cfg, seed=9, noise_amp=1e-3))
    assert not np.array_equal(U, mhd.fill_harris(cfg, seed=10, noise_amp=1e-3))
    eta = mhd.fill_harris_eta(cfg, 1e-3, 5e-3)
    np.testing.assert_allclose(eta, 1e-3 + 5e-3 * np.exp(-(X ** 2 + Y ** 2) / 4), rtol=1e-14)


def test_turbulence_setup():
    """rho = p = 1, interior rms |v| and |B| as requested, face divergence of
    B zero to round-off (B is the discrete curl of Az), modes |k| <= kmax,
    seed-reproducible."""
    cfg = mhd.default_config(64, 48, 0.0, 2.0 * np.pi, 0.0, 1.5 * np.pi, eta=1e-4)
    U = mhd.fill_turbulence(cfg, seed=5, kmax=8, rms_v=0.5, rms_b=0.25)
    g = mhd.GHOST
    I = U[g:-g, g:-g]
    assert np.all(U[..., 0] == 1.0) and np.all(U[..., 3] == 0.0) and np.all(U[..., 6] == 0.0)
    np.testing.assert_allclose(np.sqrt((I[..., 1] ** 2 + I[..., 2] ** 2).mean()), 0.5, rtol=1e-12)
    np.testing.assert_allclose(np.sqrt((I[..., 4] ** 2 + I[..., 5] ** 2).mean()), 0.25, rtol=1e-12)
    bx, by = U[..., 4], U[..., 5]
    div = ((bx[g:-g, g:-g] - bx[g - 1:-g - 1, g:-g]) / cfg.dx
           + (by[g:-g, g:-g] - by[g:-g, g - 1:-g - 1]) / cfg.dy)
    assert np.abs(div).max() < 1e-13 * np.abs(bx).max() / cfg.dx
    # energy from p = 1 and the cell-centred field
    bxc = 0.5 * (bx[1:, :] + bx[:-1, :])[:, 1:]
    byc = 0.5 * (by[:, 1:] + by[:, :-1])[1:, :]
    e = 1.0 / (2.0 / 3.0) + 0.5 * (U[1:, 1:, 1] ** 2 + U[1:, 1:, 2] ** 2) + 0.5 * (bxc ** 2 + byc ** 2)
    np.testing.assert_allclose(U[1:, 1:, 7], e, rtol=1e-13)
    # spectrum of the interior velocity: nothing beyond |n| = kmax
    spec = np.abs(np.fft.fft2(I[..., 1])) ** 2
    nx = np.fft.fftfreq(64, 1.0 / 64)[:, None]
    ny = np.fft.fftfreq(48, 1.0 / 48)[None, :]
    outside = (nx ** 2 + ny ** 2) > 64
    assert spec[outside].sum() < 1e-20 * spec.sum()
    assert spec[~outside].sum() > 0
    assert np.array_equal(U, mhd.fill_turbulence(cfg, seed=5, kmax=8, rms_v=0.5, rms_b=0.25))
    assert not np.array_equal(U, mhd.fill_turbulence(cfg, seed=6, kmax=8, rms_v=0.5, rms_b=0.25))


def test_fill_rejects_bad_config():
    cfg = mhd.orszag_tang_config(32)
    cfg.nx = 5
    with pytest.raises(mhd.MHDError):
        mhd.fill_orszag_tang(cfg)
    cfg = mhd.orszag_tang_config(16)
    with pytest.raises(mhd.MHDError):   # kmax = 8 is not resolved by 16 cells
        mhd.fill_turbulence(cfg, kmax=8)


CPP_MAIN = r"""
#include <cstdio>
#include <vector>
#include "mhd2d.hpp"
// one rank of the MPI library's example loop: its step on one subdomain and
// the per-rank interior-only save
void multi_rank(mhd2d::ResistiveMHD2D &sim, int rank) {
    sim.setScheme(MHD_SCHEME_MULTI_SHARE);
    mhd_init_turbulence(sim.handle(), 1, 8, 0.5, 0.5);
    sim.oneStepRK2();
    sim.saveSubdomain(".", "turb", 0, rank);
}
// the reference example's time loop (EX/bhattacharjee/main_bhattacharjee.cu:264-296)
int main() {
    mhd_config cfg;
    mhd_default_config(&cfg, 64, 64, 0.0, 1.0, 0.0, 1.0);
    cfg.eta = 2e-4;
    mhd2d::ResistiveMHD2D sim(cfg);
    std::vector<double> U(sim.size());
    mhd_fill_orszag_tang(&cfg, U.data());
    sim.initializeU(U);
    double totalTime = 0.0;
    for (int step = 0; step < 3; ++step) {
        if (step % 100 == 0) sim.save(".", "ot", step);
        sim.oneStepRK2();
        if (sim.checkCalculationIsCrashed()) return 1;
        totalTime += sim.lastDt();
    }
    std::printf("%g\n", totalTime);
    return 0;
}
"""


def test_cpp_wrapper_compiles_and_links(tmp_path, mhdlib):
    import subprocess
    src = tmp_path / "main.cpp"
    src.write_text(CPP_MAIN)
    libdir = os.path.join(ROOT, "resistivemhd_gpu_amd")
    res = subprocess.run(["g++", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), str(src),
                          "-L", libdir, "-lmhd2d", "-Wl,-rpath," + libdir, "-o", str(tmp_path / "main")],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
