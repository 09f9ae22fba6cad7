#!/usr/bin/env python3
"""Throughput of the plain FP64 resistive-MHD RK2 step on one MI355X.

Workload (default, --workload harris): Harris current sheet, 4096 x 4096
interior cells on [-100, 100]^2 (the cell size of the 4096 x 2048 case on
[-100, 100] x [-50, 50], which the default run also times), periodic x,
symmetric y, Bx = tanh y, rho = sech^2 y + 0.2, p = rho/2, per-cell
resistivity 1e-3 + 5e-3 exp(-(x^2 + y^2)/4), seeded white-noise vy of
amplitude 1e-3.
--workload turbulence: one GPU's subdomain of the MPI library's step,
8192 x 8192 decaying turbulence (rho = p = 1, eta 1e-4, random-phase modes
|k| <= 8 with rms v = rms B = 0.5), periodic halo filled on the device.
The state is resident in HBM before the timed region; K steps run with no
host synchronisation inside it.

Prints one JSON line: zone-cycles/s (interior cells x RK2 steps / s), the
roofline of the dominant stage kernel (HIP events on the library's stream)
and the CPU baseline: the Thrust twin of the reference's pass structure on
the host's cores, timed on a bounded sample of the same setup.
"""
import argparse
import json
import os
import sys
import time

METRIC = "zone-cycles/s FP64 4096² (%HBM peak); batched fwd+adjoint runs/s vs plain"
HBM_PEAK_GBS = 8000.0
STATE_B = 64   # 8 doubles per cell
ETA_B = 8      # per-cell resistivity plane


AUX_B = 32    # face fluxes F0, F5, G0, G4 the multi-GPU scheme's stages store


def stage_bytes(zones, eta_field, multi=False):
    """Algorithmic HBM bytes of one launch of each stage kernel.
    stage 1: read U^n, write U* ; stage 2: read U*, read U^n, write U^{n+1};
    each stage reads the resistivity plane once.  The multi-GPU scheme's
    stages also write the four face fluxes of the CT-contact EMF."""
    e = ETA_B if eta_field else 0
    a = AUX_B if multi else 0
    return zones * (2 * STATE_B + e + a), zones * (3 * STATE_B + e + a)


def multi_extra_bytes(zones):
    """Algorithmic HBM bytes per step of the multi-GPU scheme's EMF and CT
    kernels: two cell-EMF passes (read rho, rho u, rho v, Bx, By; write one
    plane), the corner EMF (read 2 x 4 flux planes and 2 EMF planes, write
    one) and the final CT (read U^n's Bx, By, E and the 6 other planes of U2,
    write Bx, By)."""
    return zones * (2 * (40 + 8) + (64 + 16 + 8) + (16 + 8 + 48 + 16))


def measured_traffic(root, nx, ny, numerics, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a
    1 GiB device copy), or None if there is none."""
    path = os.path.join(root, "profiles", f"rk_stage_kernel_harris{nx}x{ny}_{numerics}_hbm_counters.md")
    if not os.path.exists(path):
        return None
    want = f"traffic_bytes_per_launch {kernel} ="
    for line in open(path):
        if line.startswith(want):
            return float(line.split("=")[1])
    return None


def host_cpu():
    """CPU model, sockets and cores of this host (lscpu)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "cpus": info.get("CPU(s)")}


def _twin_rate(mhd, nx, ny, warm, steps, threads):
    from oracle.oracle import TwinSim
    cfg = mhd.harris_config(nx, ny, ly=harris_ly(nx, ny))
    TwinSim.set_threads(threads)
    t = TwinSim.from_config(cfg)
    t.set_state(mhd.fill_harris(cfg, seed=0, noise_amp=1e-3))
    t.set_eta_field(mhd.fill_harris_eta(cfg))
    t.apply_boundary()
    t.step(warm)   # first touch of the intermediates
    t0 = time.perf_counter()
    crashed = t.step(steps)
    sec = time.perf_counter() - t0
    del t
    return nx * ny * steps / sec, sec, crashed


def cpu_baseline(nx, ny, steps, threads=0, nx1=1024, ny1=1024, steps1=2):
    """The Thrust twin of the reference's pass structure (rocThrust's OpenMP
    backend, oracle/thrust_twin.cpp; its states equal the oracle's bit for
    bit) on the Harris-sheet workload: all host threads on the GPU line's
    grid for a short step count, and one thread on a smaller grid of the
    same cell size."""
    import resistivemhd_gpu_amd as mhd
    from oracle.oracle import TwinSim
    threads = threads or TwinSim.max_threads()
    rate, sec, crashed = _twin_rate(mhd, nx, ny, 1, steps, threads)
    rate1, sec1, _ = _twin_rate(mhd, nx1, ny1, 1, steps1, 1)
    return {
        "value": rate,
        "unit": "zone-cycles/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"Thrust twin of the reference's pass structure (rocThrust OpenMP backend, {threads} threads), "
                   f"Harris sheet {nx}x{ny}, {steps} RK2 steps after 1 warm-up step, {sec:.1f} s"),
        "one_thread": {"value": rate1, "unit": "zone-cycles/s",
                       "sample": f"same twin on 1 thread, Harris sheet {nx1}x{ny1}, {steps1} steps, {sec1:.1f} s"},
        "host": host_cpu(),
        "crashed": crashed,
    }


BATCH_SIZES = (64, 128, 192, 256, 384, 512)


def batch_workload(nruns):
    return (f"{nruns} runs: Orszag-Tang / Harris 50/50, N in {list(BATCH_SIZES)} squared, "
            "eta log-uniform [1e-4, 1e-2], seeded per run")


def batch_configs(mhd, nruns, seed0=0):
    """The batched workload: run s draws (setup, N, eta) from a generator
    seeded with seed0 + s: Orszag-Tang or Harris sheet (50/50) on N x N
    cells, N uniform over BATCH_SIZES, eta log-uniform in [1e-4, 1e-2]."""
    import numpy as np
    runs = []
    for s in range(nruns):
        rng = np.random.default_rng(seed0 + s)
        harris = bool(rng.integers(2))
        n = int(rng.choice(BATCH_SIZES))
        eta = float(10.0 ** rng.uniform(-4.0, -2.0))
        cfg = mhd.harris_config(n, n, eta_max=eta) if harris else mhd.orszag_tang_config(n, eta=eta)
        cfg.eta = eta
        runs.append((cfg, "harris" if harris else "orszag_tang", seed0 + s))
    return runs


def batch_bench(nruns, steps, plain_steps, run_steps, numerics):
    """Runs/s of the batched form against a loop of single handles advancing
    the same runs one after another, both with states resident on the GPU.
    runs/s = runs * (timed steps / steps per run) / seconds."""
    import torch
    import resistivemhd_gpu_amd as mhd
    runs = batch_configs(mhd, nruns)
    b = mhd.ResistiveMHD2DBatch([c for c, _, _ in runs])
    b.set_numerics(numerics)
    for r, (c, kind, seed) in enumerate(runs):
        if kind == "harris":
            b.initHarris(r, seed=seed, noise_amp=1e-3)
        else:
            b.initOrszagTang(r)
    b.applyBoundary()
    t0 = time.perf_counter()
    sweep = b.autotune(steps=2)
    waves = b.waves_per_cu()
    t_tune = time.perf_counter() - t0
    b.steps(4)
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.steps(steps)
    b.synchronize()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    crashed = any(b.checkCalculationIsCrashed(r) for r in range(nruns))
    del b

    singles = []
    for c, kind, seed in runs:
        s = mhd.ResistiveMHD2D(c)
        s.set_numerics(numerics)
        if kind == "harris":
            s.initializeU(mhd.fill_harris(c, seed=seed, noise_amp=1e-3))
        else:
            s.initOrszagTang()
        s.steps(2)
        singles.append(s)
    for s in singles:
        s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in singles:              # the plain per-instance loop: one run after another
        s.steps(plain_steps)
        s.synchronize()
    t_plain = time.perf_counter() - t0
    del singles
    cells = sum((c.nx - 6) * (c.ny - 6) for c, _, _ in runs)
    batched = nruns * steps / run_steps / t_batch
    plain = nruns * plain_steps / run_steps / t_plain
    return {
        "runs": nruns, "steps_per_run": run_steps, "numerics": numerics,
        "interior_cells_total": cells,
        "batched_runs_per_s": batched, "plain_runs_per_s": plain, "speedup_vs_plain": batched / plain,
        "batched_ms_per_step": t_batch / steps * 1e3, "plain_ms_per_step": t_plain / plain_steps * 1e3,
        "batched_zone_cycles_per_s": cells * steps / t_batch,
        "timed_steps": {"batched": steps, "plain": plain_steps},
        "autotuned_sweep_columns": sweep, "autotuned_waves_per_cu": waves, "autotune_s": t_tune,
        "crashed": crashed,
        "workload": batch_workload(nruns),
    }


def adjoint_bench(nruns, steps, run_steps, batched_only=False, stride=0):
    """Runs/s of batched forward+adjoint against a loop of single-run
    forward+adjoint objects, on the batched workload.  States stay on the
    GPU: dL/dU_N = U_N as the seed, moved with the device-pointer transfers
    (torch tensors) in both legs, as the torch binding does.  Both legs warm
    up with a forward+adjoint of the timed step count, so the timed call
    finds its checkpoint and history storage already laid out (steady
    state; a first call with a new step count also sizes the dt table and
    uploads the per-slot descriptors, which costs a single-run object about
    a millisecond of host round trips).  runs/s = runs * (steps / run_steps)
    / seconds.  batched_only: the batched leg alone (for profiling it;
    no plain loop, no ratio).  stride: the checkpoint stride of both legs
    (0: ceil(sqrt(steps)); -1: the largest that fits in memory)."""
    import numpy as np
    import torch
    import resistivemhd_gpu_amd as mhd
    runs = batch_configs(mhd, nruns)

    def init(obj, r, c, kind, seed):
        U0 = mhd.fill_harris(c, seed=seed, noise_amp=1e-3) if kind == "harris" else mhd.fill_orszag_tang(c)
        obj.setU0(r, U0)

    def fwd_adj(obj, buf):
        obj.forward(steps, stride=stride)
        obj.getU_all_device(buf)
        obj.set_seed_all_device(buf)
        obj.backward()

    a = mhd.ResistiveMHD2DAdjoint([c for c, _, _ in runs])
    for r, (c, kind, seed) in enumerate(runs):
        init(a, r, c, kind, seed)
    cols = a.autotune()   # as the batched-forward line does; single runs keep their default
    buf = torch.empty(a._total(), dtype=torch.float64, device="cuda")
    fwd_adj(a, buf)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd_adj(a, buf)
    t_batch = time.perf_counter() - t0
    finite = all(np.isfinite(a.grad_state(r)).all() for r in range(0, nruns, 17))
    del a, buf
    if batched_only:
        return {"runs": nruns, "steps_per_run": run_steps, "timed_steps": steps,
                "batched_fwd_adj_runs_per_s": nruns * steps / run_steps / t_batch, "batched_s": t_batch,
                "gradients_finite": bool(finite), "autotuned_sweep_columns": cols,
                "note": "batched leg only, after a warm-up of the same step count"}

    singles = []
    for c, kind, seed in runs:
        s = mhd.ResistiveMHD2DAdjoint([c])
        init(s, 0, c, kind, seed)
        b = torch.empty(s._total(), dtype=torch.float64, device="cuda")
        fwd_adj(s, b)
        singles.append((s, b))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, b in singles:
        fwd_adj(s, b)
    t_plain = time.perf_counter() - t0
    del singles
    return {
        "runs": nruns, "steps_per_run": run_steps, "timed_steps": steps,
        "batched_fwd_adj_runs_per_s": nruns * steps / run_steps / t_batch,
        "plain_fwd_adj_runs_per_s": nruns * steps / run_steps / t_plain,
        "speedup_vs_plain": t_plain / t_batch,
        "batched_s": t_batch, "plain_s": t_plain, "gradients_finite": bool(finite),
        "autotuned_sweep_columns": cols,
        "checkpoint_stride": "sqrt" if stride == 0 else ("fit" if stride < 0 else stride),
        "note": "forward, seed transfer on the GPU and backward timed in both legs, after a warm-up of the same "
                "step count (steady state)",
    }


def gradcheck_problem(nx, ny, nruns, seed=0):
    """The configured gradient check's inputs: the Harris sheet on nx x ny
    cells, its initial state with seeded vy noise (amplitude 1e-3) and vx
    noise (so that MUSCL's limiter is off its kink: vx = 0 exactly has
    directional derivatives but no gradient), and per run a seeded 'true'
    resistivity (sum of 6 Gaussians, amplitudes log-uniform in
    [1e-4, 5e-3], plus 1e-4)."""
    import numpy as np
    import resistivemhd_gpu_amd as mhd
    g = mhd.GHOST
    cfg = mhd.harris_config(nx, ny, eta_max=6e-3)
    rng = np.random.default_rng(seed)
    X = cfg.xmin + (np.arange(cfg.nx) - g + 0.5)[:, None] * cfg.dx
    Y = cfg.ymin + (np.arange(cfg.ny) - g + 0.5)[None, :] * cfg.dy

    def gaussians(r):
        eta = np.zeros((cfg.nx, cfg.ny))
        for _ in range(6):
            amp = 10 ** r.uniform(-4, np.log10(5e-3))
            x0, y0 = r.uniform(-80, 80), r.uniform(-5, 5)
            w = r.uniform(3, 15)
            eta += amp * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / w ** 2)
        return eta + 1e-4

    U0 = mhd.fill_harris(cfg, seed=seed, noise_amp=1e-3)
    vx = 1e-3 * (2 * rng.uniform(size=U0.shape[:2]) - 1)
    U0[..., 7] += 0.5 * U0[..., 0] * vx * vx
    U0[..., 1] = U0[..., 0] * vx
    etas = [gaussians(np.random.default_rng(seed + 100 + r)) for r in range(nruns)]
    return cfg, U0, etas


def gradient_check(nruns, nx, ny, steps, ndirs, seed=0, degenerate_tol=0.0, fd_h=1e-4, plain_fd=True,
                   log=None):
    """Harris sheet nx x ny, batch of nruns: 'observed' final states from the
    seeded resistivity of each run (gradcheck_problem); loss = |(rho, Bx,
    By)(T) - observed|^2 from eta = 1e-3; gradients with respect to eta(x, y),
    vx0 and vy0 from the adjoint of the whole batch.  Checked on run 0 along
    ndirs random directions against
      - the tangent-linear model (ndirs runs of one tangent rollout),
      - central differences of the decision replay (2 ndirs runs: the
        perturbed rollouts with every branch decision and dt of the recorded
        one, the map the adjoint differentiates),
      - and, reported only, central differences of the rollout itself with
        dt frozen (plain_fd), whose own spread between steps h and 10 h shows
        where branch switches within the stencil make the map non-smooth.
    degenerate_tol > 0 switches the adjoint, the tangent and the replay to
    the degenerate-limit derivative at HLLD star states whose denominator
    is below that fraction of rho (S - u)^2."""
    import numpy as np
    import resistivemhd_gpu_amd as mhd
    say = log or (lambda *a: None)
    g = mhd.GHOST
    cfg, U0, etas = gradcheck_problem(nx, ny, nruns, seed)
    rho = U0[..., 0]
    vx, vy = U0[..., 1] / rho, U0[..., 2] / rho
    cfgs = [cfg] * nruns
    t0 = time.perf_counter()
    obs = mhd.ResistiveMHD2DAdjoint(cfgs)
    for r in range(nruns):
        obs.setU0(r, U0)
        obs.setEtaField(r, etas[r])
    obs.forward(steps)
    target = [obs.getU(r) for r in range(nruns)]
    del obs
    say(f"observed states: {time.perf_counter() - t0:.1f} s")
    eta0 = np.full((cfg.nx, cfg.ny), 1e-3)
    mask = np.zeros((cfg.nx, cfg.ny, 8))
    mask[g:-g, g:-g, [0, 4, 5]] = 1.0

    a = mhd.ResistiveMHD2DAdjoint(cfgs)
    a.set_degenerate_tol(degenerate_tol)
    for r in range(nruns):
        a.setU0(r, U0)
        a.setEtaField(r, eta0)
    t1 = time.perf_counter()
    a.forward(steps)
    UN = [a.getU(r) for r in range(nruns)]
    for r in range(nruns):
        a.set_seed(r, 2.0 * mask * (UN[r] - target[r]))
    a.backward()
    t_grad = time.perf_counter() - t1
    say(f"batch forward + adjoint: {t_grad:.1f} s")
    dts = a.dts(0, steps)
    gU, geta = a.grad_state(0), a.grad_eta(0)
    gvx = gU[..., 1] * rho + gU[..., 7] * rho * vx
    gvy = gU[..., 2] * rho + gU[..., 7] * rho * vy
    del a

    def with_velocity(dvx, dvy, h):
        U = U0.copy()
        u, v = vx + h * dvx, vy + h * dvy
        U[..., 1], U[..., 2] = rho * u, rho * v
        U[..., 7] = U0[..., 7] - 0.5 * rho * (vx * vx + vy * vy) + 0.5 * rho * (u * u + v * v)
        return U

    drng = np.random.default_rng(seed + 7)
    dirs = [(drng.normal(size=eta0.shape) * 1e-3, drng.normal(size=rho.shape) * 1e-3,
             drng.normal(size=rho.shape) * 1e-3) for _ in range(ndirs)]
    ads = [float((geta * de).sum() + (gvx * dvx)[g:-g, g:-g].sum() + (gvy * dvy)[g:-g, g:-g].sum())
           for de, dvx, dvy in dirs]

    def fd_of(Up, Um, h):
        # L(Up) - L(Um) summed cell by cell as (Up - Um)(Up + Um - 2 T), so
        # two large sums do not cancel
        return float((mask * (Up - Um) * (Up + Um - 2.0 * target[0])).sum()) / (2 * h)

    t2 = time.perf_counter()
    tl = mhd.ResistiveMHD2DAdjoint([cfg] * ndirs)
    tl.set_degenerate_tol(degenerate_tol)
    for k, (de, dvx, dvy) in enumerate(dirs):
        dU0 = np.zeros_like(U0)
        dU0[..., 1] = rho * dvx
        dU0[..., 2] = rho * dvy
        dU0[..., 7] = rho * (vx * dvx + vy * dvy)
        tl.setU0(k, U0)
        tl.setEtaField(k, eta0)
        tl.set_dts(k, dts)
        tl.set_direction(k, dU0, de)
    tl.tangent(steps)
    tls = [float((2.0 * mask * (UN[0] - target[0]) * tl.get_tangent(k)).sum()) for k in range(ndirs)]
    del tl
    say(f"tangent of {ndirs} directions: {time.perf_counter() - t2:.1f} s")

    def replay_fd(h):
        rp = mhd.ResistiveMHD2DAdjoint([cfg] * (2 * ndirs))
        rp.set_degenerate_tol(degenerate_tol)
        for k, (de, dvx, dvy) in enumerate(dirs):
            for q, sg in ((0, 1.0), (1, -1.0)):
                r = 2 * k + q
                rp.setU0(r, U0)
                rp.setEtaField(r, eta0)
                rp.set_dts(r, dts)
                rp.set_replay_state(r, with_velocity(dvx, dvy, sg * h), eta0 + sg * h * de)
        rp.replay(steps)
        out = [fd_of(rp.get_tangent(2 * k), rp.get_tangent(2 * k + 1), h) for k in range(ndirs)]
        del rp
        return out

    t3 = time.perf_counter()
    rfd = replay_fd(fd_h)
    say(f"decision replay of {2 * ndirs} perturbed rollouts: {time.perf_counter() - t3:.1f} s")
    rel = lambda x, y: abs(x - y) / abs(y)
    res = {
        "workload": f"harris {nx}x{ny}, batch {nruns}, {steps} steps", "directions": ndirs,
        "degenerate_tol": degenerate_tol, "fd_h": fd_h,
        "tangent_rel_errs": [rel(x, y) for x, y in zip(ads, tls)],
        "replay_fd_rel_errs": [rel(x, y) for x, y in zip(ads, rfd)],
        "values": [{"adjoint": x, "tangent": y, "replay_fd": z} for x, y, z in zip(ads, tls, rfd)],
        "loss": float((mask * (UN[0] - target[0]) ** 2).sum()),
        "fwd_adj_seconds_batch": t_grad, "setup_seconds": t1 - t0,
    }
    res["tangent_rel_err_max"] = max(res["tangent_rel_errs"])
    res["replay_fd_rel_err_max"] = max(res["replay_fd_rel_errs"])
    if plain_fd:
        def plain(h):
            pf = mhd.ResistiveMHD2DAdjoint([cfg] * (2 * ndirs))
            for k, (de, dvx, dvy) in enumerate(dirs):
                for q, sg in ((0, 1.0), (1, -1.0)):
                    r = 2 * k + q
                    pf.setU0(r, with_velocity(dvx, dvy, sg * h))
                    pf.setEtaField(r, eta0 + sg * h * de)
                    pf.set_dts(r, dts)
            pf.forward(steps, frozen_dt=True)
            out = [fd_of(pf.getU(2 * k), pf.getU(2 * k + 1), h) for k in range(ndirs)]
            del pf
            return out
        fd, fd10 = plain(fd_h), plain(10 * fd_h)
        res["fd_rel_errs"] = [rel(x, y) for x, y in zip(ads, fd)]
        res["fd_rel_err_median"] = float(np.median(res["fd_rel_errs"]))
        res["fd_step_spread"] = [rel(y, x) for x, y in zip(fd, fd10)]
        res["fd_step_spread_median"] = float(np.median(res["fd_step_spread"]))
    return res


def harris_ly(nx, ny):
    """y extent of the Harris-sheet workloads: the cell size of the 4096 x
    2048 case on [-100, 100] x [-50, 50] (dx = dy = 200 / 4096), so the
    4096^2 grid covers [-100, 100]^2."""
    return 200.0 * ny / nx


def plain_bench(mhd, torch, root, workload, nx, ny, steps, warmup, numerics, sweep):
    """zone-cycles/s of the plain step on one workload, with the roofline of
    its dominant stage kernel (HIP events on the library's stream)."""
    multi = workload == "turbulence"
    if multi:
        cfg = mhd.default_config(nx, ny, 0.0, 2.0 * 3.141592653589793, 0.0, 2.0 * 3.141592653589793, eta=1e-4)
    else:
        cfg = mhd.harris_config(nx, ny, ly=harris_ly(nx, ny))
    sim = mhd.ResistiveMHD2D(cfg)
    sim.set_numerics(numerics)
    if sweep:
        sim.set_sweep_length(sweep)
    if multi:
        sim.set_scheme("multi")
        sim.initTurbulence(seed=0, kmax=8, rms_v=0.5, rms_b=0.5)
        name = f"turbulence_{nx}x{ny}_one_subdomain_of_multi_gpu_step_periodic_halo"
    else:
        sim.initHarris(seed=0, noise_amp=1e-3, eta0=1e-3, eta1=5e-3)
        name = f"harris_sheet_{nx}x{ny}_periodic_x_symmetric_y_eta_field"
    zones = nx * ny

    sim.steps(warmup)
    sim.synchronize()
    torch.cuda.synchronize()
    sim.profile_enable(True)
    t0 = time.perf_counter()
    sim.steps(steps)
    sim.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms1, ms2, nt = sim.profile_read()
    sim.profile_enable(False)
    crashed = sim.checkCalculationIsCrashed()
    sweep_cols = sim.sweep_length()
    del sim

    seconds = t1 - t0
    value = zones * steps / seconds
    b1, b2 = stage_bytes(zones, not multi, multi)
    avg1, avg2 = ms1 / nt / 1e3, ms2 / nt / 1e3
    # dominant kernel = the stage with the larger summed time
    exact = "true" if numerics == "exact" else "false"
    scheme = 1 if multi else 0
    if avg2 >= avg1:
        kname, kbytes, kavg = f"rk_stage_kernel<2, {exact}, {scheme}>", b2, avg2
    else:
        kname, kbytes, kavg = f"rk_stage_kernel<1, {exact}, {scheme}>", b1, avg1
    traffic = None if multi else measured_traffic(root, nx, ny, numerics, kname)
    achieved = kbytes / kavg / 1e9
    return {
        "value": value,
        "ms_per_step": seconds / steps * 1e3,
        "config": {
            "workload": name,
            "nx": nx, "ny": ny, "ghost": 3, "sweep_columns": sweep_cols,
            "numerics": numerics,
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": kbytes,
            "avg_launch_ms": kavg * 1e3,
            "stage_ms": [ms1 / nt, ms2 / nt],
        },
        "step_hbm_frac": value * (b1 + b2 + (multi_extra_bytes(zones) if multi else 0)) / zones
        / (HBM_PEAK_GBS * 1e9),
        "crashed": crashed,
    }


def fast_vs_exact(mhd, steps=100):
    """Fast arithmetic against exact (bitwise the oracle) after `steps` steps
    of the plain-method case (Orszag-Tang 256^2, eta 1e-3) and of the
    configured Harris case (4096 x 2048, eta field): relative L2 of the
    interior state per conserved variable, relative differences of dt and
    time, and the faces of the final exact state whose HLLD region differs
    under fast arithmetic (mhd_face_regions)."""
    import numpy as np
    names = ("rho", "rhoU", "rhoV", "rhoW", "bX", "bY", "bZ", "e")
    g = mhd.GHOST
    cases = [
        ("orszag_tang_256_eta1e-3", mhd.orszag_tang_config(256, eta=1e-3), lambda sim: sim.initOrszagTang()),
        ("harris_4096x2048_eta_field", mhd.harris_config(4096, 2048, ly=harris_ly(4096, 2048)),
         lambda sim: sim.initHarris(seed=0, noise_amp=1e-3, eta0=1e-3, eta1=5e-3)),
    ]
    out = {}
    for tag, cfg, init in cases:
        sims = {}
        for mode in ("exact", "fast"):
            sim = mhd.ResistiveMHD2D(cfg)
            sim.set_numerics(mode)
            init(sim)
            sim.steps(steps)
            sim.synchronize()
            sims[mode] = sim
        Ue = sims["exact"].getU()[g:-g, g:-g]
        Uf = sims["fast"].getU()[g:-g, g:-g]
        num = np.sqrt(((Uf - Ue) ** 2).reshape(-1, 8).sum(axis=0))
        den = np.sqrt((Ue ** 2).reshape(-1, 8).sum(axis=0))
        rel = num / np.where(den == 0.0, 1.0, den)
        rx, ry = sims["exact"].face_regions("exact")
        fx, fy = sims["exact"].face_regions("fast")
        ex, ey, qx, qy = rx[2:-3, 3:-3], ry[3:-3, 2:-3], fx[2:-3, 3:-3], fy[3:-3, 2:-3]
        de, df = sims["exact"].dt, sims["fast"].dt
        te, tf = sims["exact"].totalTime, sims["fast"].totalTime
        out[tag] = {
            "steps": steps,
            "rel_l2": {n: float(r) for n, r in zip(names, rel)},
            "rel_l2_max": float(rel.max()),
            "dt_rel_diff": abs(df - de) / de,
            "time_rel_diff": abs(tf - te) / te,
            "faces": int(ex.size + ey.size),
            "region_flips_x": int((ex != qx).sum()),
            "region_flips_y": int((ey != qy).sum()),
            "crashed": [bool(sims[m].checkCalculationIsCrashed()) for m in ("exact", "fast")],
        }
        del sims
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["harris", "turbulence", "batch", "adjoint", "gradcheck"],
                    default="harris")
    ap.add_argument("--adj-steps", type=int, default=20, help="timed forward+adjoint steps per run")
    ap.add_argument("--adj-stride", type=int, default=0,
                    help="adjoint workload: checkpoint stride of both legs (0: ceil(sqrt(steps)), -1: fit in memory)")
    ap.add_argument("--adj-batched-only", action="store_true",
                    help="adjoint workload: time the batched leg only (no plain loop, no ratio)")
    ap.add_argument("--check-nx", type=int, default=1024)
    ap.add_argument("--check-ny", type=int, default=512)
    ap.add_argument("--check-runs", type=int, default=32)
    ap.add_argument("--check-steps", type=int, default=2000)
    ap.add_argument("--check-dirs", type=int, default=16)
    ap.add_argument("--degenerate-tol", type=float, default=1e-6,
                    help="gradcheck: derivative rule at degenerate HLLD star states (0: the exact derivative, "
                         "which is rounding noise at the Harris far field's faces)")
    ap.add_argument("--fd-h", type=float, default=1e-5, help="gradcheck: central-difference step")
    ap.add_argument("--batch-runs", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=40, help="timed batched steps")
    ap.add_argument("--plain-steps", type=int, default=10, help="timed steps per run of the plain loop")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-form line in the default run")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the batched forward+adjoint line in the default run")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=0, help="interior cells along x (0: 4096 harris, 8192 turbulence)")
    ap.add_argument("--ny", type=int, default=0, help="interior cells along y (0: 2048 harris, 8192 turbulence)")
    ap.add_argument("--sweep", type=int, default=0, help="columns per wavefront sweep (0 = library default)")
    ap.add_argument("--numerics", choices=["exact", "fast"], default="exact",
                    help="exact: bitwise reference arithmetic (library default); fast: rounding-level differences")
    ap.add_argument("--cpu-steps", type=int, default=2, help="timed steps of the CPU baseline on the GPU line's grid")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip the 4096 x 2048 Harris line of the default run")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (reference point for HBM byte counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the step runs on a single MI355X "
              "(multi-GPU sharding is a design, not a built path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    root = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, root)
    import torch
    import resistivemhd_gpu_amd as mhd

    if args.workload == "adjoint":
        res = adjoint_bench(args.batch_runs, args.adj_steps, 500, args.adj_batched_only, args.adj_stride)
        print(json.dumps({"metric": "batched forward+adjoint runs/s vs plain per-instance loop",
                          "value": res["batched_fwd_adj_runs_per_s"], "unit": "runs/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": batch_workload(args.batch_runs)}, "adjoint": res}),
              flush=True)
        return
    if args.workload == "gradcheck":
        res = gradient_check(args.check_runs, args.check_nx, args.check_ny, args.check_steps, args.check_dirs,
                             degenerate_tol=args.degenerate_tol, fd_h=args.fd_h,
                             log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps({"metric": "adjoint gradient vs tangent-linear model and central differences",
                          "value": res["tangent_rel_err_max"], "unit": "relative", "higher_is_better": False,
                          "dtype": "f64", "data": "synthetic", "gradcheck": res}), flush=True)
        return
    if args.workload == "batch":
        res = batch_bench(args.batch_runs, args.batch_steps, args.plain_steps, 500, args.numerics)
        print(json.dumps({"metric": "batched runs/s vs plain per-instance loop", "value": res["batched_runs_per_s"],
                          "unit": "runs/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic", "config": {"workload": res["workload"]}, "batched": res}),
              flush=True)
        return

    multi = args.workload == "turbulence"
    if multi:
        args.nx = args.nx or 8192
        args.ny = args.ny or 8192
    else:
        args.nx = args.nx or 4096
        args.ny = args.ny or 4096
    res = plain_bench(mhd, torch, root, args.workload, args.nx, args.ny, args.steps, args.warmup, args.numerics,
                      args.sweep)
    line = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": res["config"],
        "roofline": res["roofline"],
        "step_hbm_frac": res["step_hbm_frac"],
        "crashed": res["crashed"],
    }
    if not multi and (args.nx, args.ny) == (4096, 4096) and not args.no_extra:
        # the headline run of the configured Harris case, 4096 x 2048 cells
        line["harris_4096x2048"] = plain_bench(mhd, torch, root, "harris", 4096, 2048, args.steps, args.warmup,
                                               args.numerics, args.sweep)
    if not multi and (args.nx, args.ny) == (4096, 4096) and not args.no_extra:
        # fast arithmetic: its throughput on the headline grid and how far it
        # is from exact (the oracle's arithmetic) on the two reference cases
        fast = plain_bench(mhd, torch, root, "harris", 4096, 4096, args.steps, args.warmup, "fast", args.sweep)
        line["fast_vs_exact"] = {"fast_4096x4096": {k: fast[k] for k in ("value", "ms_per_step", "roofline")},
                                 **fast_vs_exact(mhd)}
    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        line["calibration_copy_bytes"] = n * 8
        del src, dst
    if not multi and not args.no_batch and not args.no_cpu:
        line["batched"] = batch_bench(args.batch_runs, args.batch_steps, args.plain_steps, 500, args.numerics)
    if not multi and not args.no_adjoint and not args.no_cpu and args.numerics == "exact":
        line["batched_fwd_adjoint"] = adjoint_bench(args.batch_runs, args.adj_steps, 500)
    if not args.no_cpu and not multi:
        line["cpu_baseline"] = cpu_baseline(args.nx, args.ny, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
