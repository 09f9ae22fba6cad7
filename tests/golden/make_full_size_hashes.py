"""Writes full_size_hashes.json: the CPU oracle's state after 100 RK2 steps
of the full-size Harris-sheet workloads (4096 x 2048 on [-100, 100] x
[-50, 50] and 4096 x 4096 on [-100, 100]^2, periodic x, symmetric y,
per-cell resistivity 1e-3 + 5e-3 exp(-(x^2 + y^2)/4), vy noise 1e-3 from
seed 0), as a SHA-256 of every cell's 8 doubles (ghosts included, -0.0
written as +0.0) together with dt and the accumulated time after the last
step.  The GPU tests compare the HIP library's state of the same run with
these hashes, so that the bitwise check of the headline sizes needs no CPU
run on the GPU host.

Run from the repository root: python tests/golden/make_full_size_hashes.py
(about 15 GB of memory and some minutes per case on 8 cores).
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import resistivemhd_gpu_amd as mhd  # noqa: E402
from oracle.oracle import OracleSim  # noqa: E402

CASES = [("harris_4096x2048", 4096, 2048, 100.0), ("harris_4096x4096", 4096, 4096, 200.0)]
STEPS = 100


def state_hash(U):
    """SHA-256 of the state with negative zeros written as +0.0 (so that
    the hash compares values, as np.array_equal does)."""
    return hashlib.sha256(np.ascontiguousarray(U + 0.0).tobytes()).hexdigest()


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "full_size_hashes.json")
    res = json.load(open(out)) if os.path.exists(out) else {}
    for name, nx, ny, ly in CASES:
        if len(sys.argv) > 1 and name not in sys.argv[1:]:
            continue
        cfg = mhd.harris_config(nx, ny, ly=ly)
        t0 = time.time()
        o = OracleSim.from_config(cfg)
        o.set_state(mhd.fill_harris(cfg, seed=0, noise_amp=1e-3))
        o.set_eta_field(mhd.fill_harris_eta(cfg, 1e-3, 5e-3))
        o.apply_boundary()
        h0 = state_hash(o.get_state())
        o.step(STEPS)
        U = o.get_state()
        res[name] = {"nx": nx, "ny": ny, "ly": ly, "steps": STEPS, "seed": 0, "noise_amp": 1e-3,
                     "eta0": 1e-3, "eta1": 5e-3, "initial_state_sha256": h0, "state_sha256": state_hash(U),
                     "dt": o.dt, "time": o.time, "crashed": o.crashed(),
                     "rho_sum": float(U[3:-3, 3:-3, 0].sum())}
        print(name, res[name], f"{time.time() - t0:.0f} s", flush=True)
        del o, U
        json.dump(res, open(out, "w"), indent=1)


if __name__ == "__main__":
    main()
