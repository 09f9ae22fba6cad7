"""GPU check of the allocation-free step (mhd_step_rk2_noalloc), called
through the C ABI: on two calls in a row, the second on a smaller grid (the
scratch keeps the first call's contents), the state, dt and time equal
those of mhd_step_rk2 on a handle bit for bit; sizes beyond the scratch and
bad layouts are refused with an error."""
import ctypes

import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from resistivemhd_gpu_amd import mhd2d

pytestmark = pytest.mark.gpu


def _planes(torch, U, pitch, plane):
    """AoS (nx, ny, 8) -> device planes (8, plane) with row pitch `pitch`."""
    nx, ny, _ = U.shape
    t = torch.zeros((8, plane), dtype=torch.float64, device="cuda")
    t[:, :nx * pitch].view(8, nx, pitch)[:, :, :ny] = torch.from_numpy(np.ascontiguousarray(U.transpose(2, 0, 1)))
    return t


def _aos(t, nx, ny, pitch):
    return t[:, :nx * pitch].view(8, nx, pitch)[:, :, :ny].cpu().numpy().transpose(1, 2, 0)


def _reference(cfg, U0, eta, steps):
    s = mhd.ResistiveMHD2D(cfg)
    s.initializeU(U0)
    if eta is not None:
        s.setEtaField(eta)
    Ub = s.getU()
    s.steps(steps)
    return Ub, s.getU(), s.dt, s.totalTime


def _noalloc(torch, lib, cfg, Ub, eta, steps, pitch):
    nx, ny = cfg.nx, cfg.ny
    plane = ((nx * pitch + 63) // 64) * 64 + 2
    st = _planes(torch, Ub, pitch, plane)
    et = None
    if eta is not None:
        et = torch.zeros(plane, dtype=torch.float64, device="cuda")
        et[:nx * pitch].view(nx, pitch)[:, :ny] = torch.from_numpy(eta)
    info = torch.zeros(3, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    rc = lib.mhd_step_rk2_noalloc(ctypes.byref(cfg), ctypes.c_void_p(st.data_ptr()), pitch, plane,
                                  ctypes.c_void_p(et.data_ptr()) if et is not None else None, steps,
                                  ctypes.c_void_p(stream.cuda_stream), ctypes.c_void_p(info.data_ptr()))
    assert rc == 0, lib.mhd_last_error(None)
    torch.cuda.synchronize()
    return _aos(st, nx, ny, pitch), info.cpu().numpy()


def test_noalloc_step_matches_handle_twice_then_smaller():
    import torch
    lib = mhd2d.load_library()
    c1 = mhd.harris_config(200, 120, eta_max=6e-3)      # periodic x, symmetric y, eta plane
    U1 = mhd.fill_harris(c1, seed=4, noise_amp=1e-2)
    e1 = mhd.fill_harris_eta(c1)
    c2 = mhd.orszag_tang_config(64, eta=1e-3)          # smaller, periodic, uniform eta
    U2 = mhd.fill_orszag_tang(c2)
    for cfg, U0, eta, steps, pitch in ((c1, U1, e1, 30, 132), (c2, U2, None, 25, 70)):
        print(f"case noalloc: nx={cfg.nx} ny={cfg.ny} pitch={pitch} steps={steps} eta_plane={eta is not None}",
              flush=True)
        Ub, Uref, dt, time = _reference(cfg, U0, eta, steps)
        U, info = _noalloc(torch, lib, cfg, Ub, eta, steps, pitch)
        assert np.array_equal(U, Uref)
        assert info[0] == dt and info[1] == time and info[2] == 0.0


def test_noalloc_refuses_sizes_beyond_the_scratch_and_bad_layouts():
    import torch
    lib = mhd2d.load_library()
    big = mhd.orszag_tang_config(1024)
    pitch = 1032
    plane = big.nx * pitch
    # buffers of the full layout size, so nothing could reach past them even
    # if a check let the call through (each call is refused before any launch)
    t = torch.zeros(8 * plane, dtype=torch.float64, device="cuda")
    rc = lib.mhd_step_rk2_noalloc(ctypes.byref(big), ctypes.c_void_p(t.data_ptr()), pitch, plane, None, 1, None, None)
    msg = lib.mhd_last_error(None).decode()
    print("refused:", msg, flush=True)
    assert rc == -1 and "scratch" in msg          # MHD_ERR_INVALID
    small = mhd.orszag_tang_config(32)
    for p, pl in ((small.ny + 1, small.nx * (small.ny + 1) + 1), (small.ny - 2, small.nx * small.ny)):
        rc = lib.mhd_step_rk2_noalloc(ctypes.byref(small), ctypes.c_void_p(t.data_ptr()), p, pl, None, 1, None, None)
        assert rc == -1, (p, pl)
    torch.cuda.synchronize()
    assert not torch.any(t != 0.0)
