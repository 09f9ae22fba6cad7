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
