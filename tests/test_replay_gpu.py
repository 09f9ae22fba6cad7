"""GPU checks of the decision replay (mhd_adjoint_replay): the rollout of a
perturbed initial state with every branch decision and dt of the recorded
rollout.  At the recorded initial state it is the forward rollout itself,
and central differences of a loss of its final state agree with the
adjoint to the truncation error of the differences, also on states where
differences of the plain rollout are broken by branch switches."""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from tests.test_adjoint_gpu import _eta_field, _generic, _harris, _ot, _weights

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["ot", "harris"])
def test_replay_at_base_is_the_rollout(case):
    cfg, U0 = _ot() if case == "ot" else _harris()
    eta = _eta_field(cfg, 2)
    print(f"case replay-at-base-{case}: nx={cfg.nx} ny={cfg.ny}", flush=True)
    a = mhd.ResistiveMHD2DAdjoint([cfg])
    a.set_degenerate_tol(1e-8)
    a.setU0(0, U0)
    a.setEtaField(0, eta)
    a.forward(15)
    UN = a.getU(0)
    a.set_replay_state(0, U0, eta)
    a.replay(15)
    assert np.array_equal(a.get_tangent(0), UN)


@pytest.mark.parametrize("case, noise, dtol", [("ot", 0.05, 1e-8), ("harris", 0.005, 1e-4), ("harris_exact", 0.0, 1e-4)])
def test_replay_differences_match_adjoint(case, noise, dtol):
    """L = sum(W * U_N): the adjoint's <dL/dU0, dU0> + <dL/deta, deta>
    against central differences of the replayed rollout along random
    directions (relative size 1e-2 h of every component).  harris_exact
    starts from the setup itself, whose exact symmetries put MUSCL's limiter
    on its kink, so plain differences of the rollout do not converge there.
    The Harris sheet has HLLD star states close to degenerate (transverse
    field of the size of the noise, rho (S - u)(S - SM) - Bx^2 down to
    ~1e-5 rho (S - u)^2): their quotients curve on the scale of their small
    denominators, far more than a step of the differences changes them
    unless the degenerate-limit rule covers them, here below 1e-4 rho (S -
    u)^2 (measured at h = 1e-4: 9e-10 with the rule at 1e-4, 4e-3 with it at
    1e-6)."""
    cfg, U0 = _ot() if case == "ot" else _harris()
    if noise:
        U0 = _generic(U0, 9, noise)
    nsteps = 20
    eta = _eta_field(cfg, 2)
    W = _weights(cfg, 3)
    print(f"case replay-fd-{case}: nx={cfg.nx} ny={cfg.ny} steps={nsteps} dtol={dtol}", flush=True)
    a = mhd.ResistiveMHD2DAdjoint([cfg])
    a.set_degenerate_tol(dtol)
    a.setU0(0, U0)
    a.setEtaField(0, eta)
    a.forward(nsteps, stride=6)
    a.set_seed(0, W)
    a.backward()
    gU, geta = a.grad_state(0), a.grad_eta(0)
    dts = a.dts(0)
    rng = np.random.default_rng(4)
    h = 1e-4
    rp = mhd.ResistiveMHD2DAdjoint([cfg] * 8)
    rp.set_degenerate_tol(dtol)
    dirs = []
    for k in range(4):
        dU0 = rng.normal(size=U0.shape) * np.maximum(np.abs(U0), 1e-2) * 1e-2 * (k != 1)
        deta = rng.normal(size=eta.shape) * eta * (k != 0)
        dirs.append((dU0, deta))
        for q, sg in ((0, 1.0), (1, -1.0)):
            r = 2 * k + q
            rp.setU0(r, U0)
            rp.setEtaField(r, eta)
            rp.set_dts(r, dts)
            rp.set_replay_state(r, U0 + sg * h * dU0, eta + sg * h * deta)
    rp.replay(nsteps)
    for k, (dU0, deta) in enumerate(dirs):
        fd = float((W * (rp.get_tangent(2 * k) - rp.get_tangent(2 * k + 1))).sum()) / (2 * h)
        ad = float((gU * dU0).sum() + (geta * deta).sum())
        print(f"dir {k}: adjoint {ad:.15e} replay fd {fd:.15e} rel {abs(ad - fd) / abs(fd):.2e}", flush=True)
        assert abs(ad - fd) <= 1e-8 * abs(fd)
