"""GPU checks of the file formats: ResistiveMHD2D::save of the single-GPU
library ({dir}/{name}_{step}.bin, all cells, AoS, no header; read back by the
notebooks as np.fromfile(...).reshape(nx, ny, 8)) and the MPI library's
per-rank save ({dir}/{name}_{step}_{rank}.bin, interior cells only)."""
import os

import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd

pytestmark = pytest.mark.gpu


def _sim(steps=3):
    cfg = mhd.orszag_tang_config(40, eta=1e-3)
    cfg.ny = 40 + 6 + 9
    s = mhd.ResistiveMHD2D(cfg)
    s.initOrszagTang()
    s.steps(steps)
    return cfg, s


def test_save_layout_and_load_round_trip(tmp_path):
    cfg, s = _sim()
    U = s.getU()
    s.save(str(tmp_path), "ot", 3)
    path = tmp_path / "ot_3.bin"
    assert path.stat().st_size == cfg.nx * cfg.ny * 8 * 8
    assert np.array_equal(np.fromfile(path, dtype="<f8").reshape(cfg.nx, cfg.ny, 8), U)
    s2 = mhd.ResistiveMHD2D(cfg)
    s2.load(str(path))
    assert np.array_equal(s2.getU(), U)
    # same continuation from the loaded state
    s.steps(2)
    s2.steps(2)
    assert np.array_equal(s.getU(), s2.getU())


def test_subdomain_save_interior_only(tmp_path):
    cfg, s = _sim()
    U = s.getU()
    g = mhd.GHOST
    s.saveSubdomain(str(tmp_path), "turb", 7, rank=2)
    path = tmp_path / "turb_7_2.bin"
    assert path.stat().st_size == (cfg.nx - 2 * g) * (cfg.ny - 2 * g) * 8 * 8
    data = np.fromfile(path, dtype="<f8").reshape(cfg.nx - 2 * g, cfg.ny - 2 * g, 8)
    assert np.array_equal(data, U[g:-g, g:-g])
    s2 = mhd.ResistiveMHD2D(cfg)
    s2.loadSubdomain(str(path))
    s2.periodicBoundaryX2nd()
    s2.periodicBoundaryY2nd()
    assert np.array_equal(s2.getU()[g:-g, g:-g], U[g:-g, g:-g])


def test_io_errors(tmp_path):
    cfg, s = _sim(0)
    with pytest.raises(mhd.MHDError):
        s.load(str(tmp_path / "missing.bin"))
    bad = tmp_path / "short.bin"
    np.zeros(10).tofile(bad)
    with pytest.raises(mhd.MHDError):
        s.load(str(bad))
    with pytest.raises(mhd.MHDError):
        s.loadSubdomain(str(bad))
    with pytest.raises(mhd.MHDError):
        s.save(os.path.join(str(tmp_path), "no", "such", "dir"), "x", 0)


CPP_LOOP = r"""
#include <cstdio>
#include <vector>
#include "mhd2d.hpp"
// The reference example's time loop (examples2D/bhattacharjee/
// main_bhattacharjee.cu:264-296) against the library: initial state,
// save every recordStep steps, crash check, accumulated time.
int main() {
    mhd_config cfg;
    mhd_default_config(&cfg, 64, 64, 0.0, 6.283185307179586, 0.0, 6.283185307179586);
    cfg.eta = 1e-3;
    mhd2d::ResistiveMHD2D sim(cfg);
    std::vector<double> U(sim.size());
    mhd_fill_orszag_tang(&cfg, U.data());
    sim.initializeU(U);
    const int totalStep = 6, recordStep = 3;
    double totalTime = 0.0;
    for (int step = 0; step < totalStep + 1; step++) {
        if (step % recordStep == 0) sim.save(".", "ot", step);
        sim.oneStepRK2();
        if (sim.checkCalculationIsCrashed()) return 1;
        totalTime += sim.lastDt();
    }
    std::printf("%.17g\n", totalTime);
    return 0;
}
"""


def test_cpp_example_loop_runs_and_matches_binding(tmp_path):
    """The C++ drop-in (include/mhd2d.hpp over the C ABI) runs the
    reference example's loop on the GPU; its .bin snapshots and accumulated
    time equal the Python binding's on the same setup, bitwise."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "resistivemhd_gpu_amd")
    src = tmp_path / "main.cpp"
    src.write_text(CPP_LOOP)
    exe = tmp_path / "main"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(root, "include"), str(src), "-L", libdir,
                    "-lmhd2d", "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    print("case cpp-example-loop: OT 64^2, 7 steps", flush=True)
    res = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    total_cpp = float(res.stdout.strip())
    cfg = mhd.orszag_tang_config(64, eta=1e-3)
    s = mhd.ResistiveMHD2D(cfg)
    s.initOrszagTang()
    total = 0.0
    for step in range(7):
        if step % 3 == 0:
            got = np.fromfile(tmp_path / f"ot_{step}.bin", dtype="<f8").reshape(cfg.nx, cfg.ny, 8)
            assert np.array_equal(got, s.getU()), step
        s.oneStepRK2()
        total += s.dt
    assert total_cpp == total
