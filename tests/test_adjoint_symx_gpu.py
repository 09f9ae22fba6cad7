"""GPU checks of the discrete adjoint through the symmetric-x boundary
(ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_symmetricXY/
boundary.cu:4-104): mirrored ghost columns, the ghost Bx rebuilt from the
zero divergence of the mirrored By, which reads the bottom boundary face
By(i, 2) as the stage's CT left it.

- the adjoint object's rollout is the plain step bit for bit;
- the decision replay at the recorded start is that rollout bit for bit
  (the tangent's boundary and stage kernels advance a state as the forward
  step does);
- the adjoint is the transpose of the tangent-linear model (dot-product
  test along random directions);
- central differences of the decision replay agree with the adjoint;
- a batch that mixes periodic and symmetric x gives each run its own
  gradient.
"""
import numpy as np
import pytest

import resistivemhd_gpu_amd as mhd
from tests.test_adjoint_gpu import _eta_field, _generic, _harris, _ot, _weights, _rel_l2

pytestmark = pytest.mark.gpu


def _sym(case):
    """Orszag-Tang with the symmetric x and y fills of the S/ library, or
    the Harris sheet with symmetric x and y; both moved off the limiter kinks
    of their exact symmetries."""
    if case == "ot":
        cfg, U0 = _ot(24)
        cfg.bc_x = cfg.bc_y = mhd.MHD_BC_SYMMETRIC
        return cfg, _generic(U0, 9, 0.05)
    cfg, U0 = _harris(40, 24)
    cfg.bc_x = mhd.MHD_BC_SYMMETRIC
    return cfg, _generic(U0, 9, 0.005)


@pytest.mark.parametrize("case", ["ot", "harris"])
def test_symx_rollout_and_replay_are_the_plain_step(case):
    cfg, U0 = _sym(case)
    eta = _eta_field(cfg, 1)
    print(f"case symx-rollout-{case}: nx={cfg.nx} ny={cfg.ny}", flush=True)
    a = mhd.ResistiveMHD2DAdjoint([cfg])
    a.setU0(0, U0)
    a.setEtaField(0, eta)
    a.forward(15, stride=4)
    s = mhd.ResistiveMHD2D(cfg)
    s.initializeU(U0)
    s.setEtaField(eta)
    dts = []
    for _ in range(15):
        s.oneStepRK2()
        dts.append(s.dt)
    UN = s.getU()
    assert np.array_equal(a.getU(0), UN)
    assert np.array_equal(a.dts(0), np.array(dts))
    a.set_replay_state(0, U0, eta)
    a.replay(15)
    assert np.array_equal(a.get_tangent(0), UN)


@pytest.mark.parametrize("face", ["per_face", "sweep"])
@pytest.mark.parametrize("case, tol", [("ot", 1e-12), ("harris", 1e-8)])
def test_symx_adjoint_matches_tangent_linear(case, tol, face):
    """Tolerances as for the periodic runs: round-off on Orszag-Tang, and
    on the noisy Harris sheet the amplification of rounding through HLLD
    denominators near degeneracy."""
    cfg, U0 = _sym(case)
    nsteps = 20
    eta = _eta_field(cfg, 2)
    W = _weights(cfg, 3)
    print(f"case symx-adjoint-tlm-{case}-{face}: nx={cfg.nx} ny={cfg.ny} steps={nsteps}", flush=True)
    a = mhd.ResistiveMHD2DAdjoint([cfg])
    a.set_face_kernel(a.FACE_SWEEP if face == "sweep" else a.FACE_PER_FACE)
    a.setU0(0, U0)
    a.setEtaField(0, eta)
    a.forward(nsteps, stride=6)
    a.set_seed(0, W)
    a.backward()
    gU, geta = a.grad_state(0), a.grad_eta(0)
    assert np.all(np.isfinite(gU)) and np.all(np.isfinite(geta))
    rng = np.random.default_rng(4)
    for k in range(4):
        dU0 = rng.normal(size=U0.shape) * (k != 1)
        deta = rng.normal(size=eta.shape) * eta * (k != 0)
        a.set_direction(0, dU0, deta)
        a.tangent(nsteps)
        tl = float((W * a.get_tangent(0)).sum())
        ad = float((gU * dU0).sum() + (geta * deta).sum())
        print(f"dir {k}: adjoint {ad:.15e} tangent {tl:.15e} rel {abs(ad - tl) / abs(tl):.2e}", flush=True)
        assert abs(ad - tl) <= tol * abs(tl)


def test_symx_boundary_direction_reaches_the_gradient():
    """A seed on the ghost-rebuilt faces' neighbourhood: perturbing only By
    of the first interior column (which the ghost Bx rebuild reads, bottom
    boundary face included) changes the loss after three steps, and the adjoint
    gives that change (replay differences), so the rebuild's transpose is
    not a no-op."""
    cfg, U0 = _sym("ot")
    eta = _eta_field(cfg, 2)
    W = _weights(cfg, 3)
    a = mhd.ResistiveMHD2DAdjoint([cfg])
    a.setU0(0, U0)
    a.setEtaField(0, eta)
    a.forward(3)
    a.set_seed(0, W)
    a.backward()
    gU = a.grad_state(0)
    dts = a.dts(0)
    dU0 = np.zeros_like(U0)
    dU0[3, 3:-3, 5] = 1.0
    dU0[cfg.nx - 4, 3:-3, 5] = -1.0
    h = 1e-4
    rp = mhd.ResistiveMHD2DAdjoint([cfg] * 2)
    for r, sg in ((0, 1.0), (1, -1.0)):
        rp.setU0(r, U0)
        rp.setEtaField(r, eta)
        rp.set_dts(r, dts)
        rp.set_replay_state(r, U0 + sg * h * dU0, eta)
    rp.replay(3)
    fd = float((W * (rp.get_tangent(0) - rp.get_tangent(1))).sum()) / (2 * h)
    ad = float((gU * dU0).sum())
    print(f"boundary column By: adjoint {ad:.15e} replay fd {fd:.15e}", flush=True)
    assert abs(fd) > 0.0
    assert abs(ad - fd) <= 1e-8 * abs(fd)


@pytest.mark.parametrize("case", ["ot", "harris"])
def test_symx_replay_differences_match_adjoint(case):
    """Central differences of the decision replay along random directions
    against the adjoint (the degenerate-limit rule at 1e-4 on the Harris
    sheet, as for its periodic runs).  On the Harris sheet with symmetric x
    the differences' truncation error at h = 1e-4 is 1.2e-8 (measured, first
    direction), so its bound is 1e-7, ten times inside the 1e-6 of the
    gradient criterion; Orszag-Tang holds 1e-8."""
    cfg, U0 = _sym(case)
    dtol = 1e-8 if case == "ot" else 1e-4
    nsteps = 20
    eta = _eta_field(cfg, 2)
    W = _weights(cfg, 3)
    print(f"case symx-replay-fd-{case}: nx={cfg.nx} ny={cfg.ny} steps={nsteps}", flush=True)
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
    rp = mhd.ResistiveMHD2DAdjoint([cfg] * 6)
    rp.set_degenerate_tol(dtol)
    dirs = []
    for k in range(3):
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
        assert abs(ad - fd) <= (1e-8 if case == "ot" else 1e-7) * abs(fd)


def test_symx_mixed_batch_matches_single_runs():
    """Periodic and symmetric x side by side in one batch: each run's
    gradient is the one it gets alone (atomic sums in another order)."""
    c0, U0 = _sym("ot")
    c1, U1 = _ot(32)
    U1 = _generic(U1, 21, 0.05)
    c2, U2 = _sym("harris")
    cases = [(c0, U0), (c1, U1), (c2, U2)]
    nsteps = 9
    a = mhd.ResistiveMHD2DAdjoint([c for c, _ in cases])
    for r, (c, U) in enumerate(cases):
        a.setU0(r, U)
        a.setEtaField(r, _eta_field(c, r))
    a.forward(nsteps, stride=3)
    for r, (c, _) in enumerate(cases):
        a.set_seed(r, _weights(c, 30 + r))
    a.backward()
    for r, (c, U) in enumerate(cases):
        s = mhd.ResistiveMHD2DAdjoint([c])
        s.setU0(0, U)
        s.setEtaField(0, _eta_field(c, r))
        s.forward(nsteps, stride=3)
        s.set_seed(0, _weights(c, 30 + r))
        s.backward()
        assert np.array_equal(s.getU(0), a.getU(r))
        bound = 1e-8 if r == 2 else 1e-11
        assert _rel_l2(s.grad_state(0), a.grad_state(r)) <= bound, r
        assert _rel_l2(s.grad_eta(0), a.grad_eta(r)) <= bound, r
