"""Differentiable stepping (cc_fwd / cc_bwd) against central finite differences of the FP64
oracle.

The adjoint differentiates the velocity-level contact solve and the integrator with the contact
set of each step fixed (include/crashcat_amd.h, cc_bwd).  Central differences (h = 1e-4) of the
oracle's loss are its true derivative only where the loss is smooth at that scale, which is
checked here by requiring the differences at h = 1e-3, 1e-4 and 1e-5 to agree.  On such
trajectories the gradients must match: a projectile in flight, a box sliding to rest (friction
disc clamped, then sticking), a sphere striking a box face, and config 5's throw worlds before
the projectile arrives (the stack settling, friction gradient).  The config-5 loss after the
impact is not smooth at h = 1e-4 (the differences grow like 1/h: warm-start matches, contact
sets and manifold choices of the settling stack switch under perturbations of that size), so
there the check is the bench's reported comparison, not a test.
"""

import numpy as np
import pytest
import torch

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH
from crashcat_amd._lib import Context, CrashcatError
from crashcat_amd.autograd import WorldRun
from oracle import refcpu as O

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
W3 = np.array([1.0, 0.3, 0.1])  # loss = W3 . com of the tracked body


def _floor(sc):
    sc.add_body(SH.box((20, 0.5, 20)), (0, -0.5, 0), motion_type=_abi.CC_MOTION_STATIC, friction=0.5)


def _flight():
    sc = SC.Scene()
    sc.settings.allow_sleeping = 0
    _floor(sc)
    sc.add_body(SH.sphere(0.4), (0, 5, 0), linear_velocity=(3, 4, 1), friction=0.5)
    return sc, 1, 1


def _slide():
    sc = SC.Scene()
    sc.settings.allow_sleeping = 0
    _floor(sc)
    sc.add_body(SH.box((0.25, 0.25, 0.25)), (0, 0.251, 0), linear_velocity=(3, 0, 0.5), friction=0.5)
    return sc, 1, 1


def _strike():
    sc = SC.Scene()
    sc.settings.allow_sleeping = 0
    _floor(sc)
    sc.add_body(SH.box((0.5, 0.5, 0.5)), (0, 0.501, 0), friction=0.5)
    sc.add_body(SH.sphere(0.3, density=300.0), (-2.0, 0.55, 0.1), linear_velocity=(8, 0.5, 0.2), friction=0.5)
    return sc, 1, 2


def _oracle_loss(build, steps, mover, body, dv=np.zeros(3), dmu=0.0):
    sc, _, _ = build()
    b = sc.bodies()
    b["linear_velocity"][mover] += dv
    b["friction"][:] += dmu
    o = O.OracleWorld(sc)
    o.step(DT, steps)
    return float(o.bodies()["center_of_mass_position"][body] @ W3)


def _fd(build, steps, mover, body, h):
    g = [(_oracle_loss(build, steps, mover, body, e) - _oracle_loss(build, steps, mover, body, -e)) / (2 * h)
         for e in np.eye(3) * h]
    gm = (_oracle_loss(build, steps, mover, body, dmu=h) - _oracle_loss(build, steps, mover, body, dmu=-h)) / (2 * h)
    return np.array(g + [gm])


def _adjoint(build, steps, precision, k=7):
    sc, body, mover = build()
    g = Context(precision, island_groups=1)
    try:
        g.upload(sc)
        g.fwd(DT, steps, k)
        gc = np.zeros((sc.num_bodies, 3))
        gc[body] = W3
        out = g.bwd(gc)
    finally:
        g.close()
    return np.array(list(out["v0"][mover]) + [out["friction"].sum()])


# tolerances as a fraction of the gradient's largest component; the strike leaves the pose
# dependence of the contact normal and of the inverse inertias out (0.24 % measured)
@pytest.mark.parametrize("case,steps,tol64,tol32", [(_flight, 40, 2e-4, 2e-3), (_slide, 30, 2e-4, 2e-3),
                                                     (_slide, 60, 2e-4, 2e-3), (_strike, 45, 5e-3, 5e-3)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adjoint_matches_fd_on_smooth_runs(case, steps, tol64, tol32, precision):
    rtol = tol64 if precision == "fp64" else tol32
    _, body, mover = case()
    fds = [_fd(case, steps, mover, body, h) for h in (1e-3, 1e-4, 1e-5)]
    ref = fds[1]
    scale = np.abs(ref).max()
    # the loss is smooth at the FD scale: the three step sizes agree
    assert np.abs(fds[0] - ref).max() <= 1e-3 * scale and np.abs(fds[2] - ref).max() <= 1e-3 * scale, fds
    adj = _adjoint(case, steps, precision)
    np.testing.assert_allclose(adj, ref, rtol=0, atol=rtol * scale + 1e-12)


def test_throw_worlds_friction_gradient_before_impact():
    """Config 5 (4x4x4 stack + projectile), 4 worlds, 12 steps: the stack settles, the projectile
    is still in flight.  dL/dmu of every world within 3 % of FP64 central differences (h = 1e-4),
    dL/dv0 zero on both sides."""
    nw, steps = 4, 12
    sc = SC.throw_worlds(num_worlds=nw)
    th = sc.throw
    offs = sc.offsets()

    def loss(w, dmu=0.0):
        b = sc.bodies()
        keep = b.copy()
        b["friction"][offs[w]:offs[w + 1]] += dmu
        o = O.OracleWorld(sc, w)
        b[:] = keep
        o.step(DT, steps)
        st = th["stack"][w] - offs[w]
        c = o.bodies()["center_of_mass_position"][st].mean(axis=0)
        return float(np.sum((c - th["target"][w]) ** 2))

    for precision in ("fp64", "fp32"):
        g = Context(precision, island_groups=1)
        try:
            g.upload(sc)
            g.fwd(DT, steps, 5)
            bs = g.bodies()
            gc = np.zeros((sc.num_bodies, 3))
            for w in range(nw):
                st = th["stack"][w]
                c = bs["center_of_mass_position"][st].mean(axis=0)
                gc[st] = 2 * (c - th["target"][w]) / len(st)
            out = g.bwd(gc)
        finally:
            g.close()
        for w in range(nw):
            fd = (loss(w, 1e-4) - loss(w, -1e-4)) / 2e-4
            adj = out["friction"][offs[w]:offs[w + 1]].sum()
            assert abs(adj - fd) <= 0.03 * abs(fd), (precision, w, adj, fd)
            assert np.abs(out["v0"][th["projectile"][w]]).max() < 1e-12


def test_checkpoint_interval_changes_nothing():
    """Recomputation from checkpoints repeats the forward steps exactly: the gradients of a
    60-step throw are the same with a checkpoint every step, every 10 steps or only at the start
    (up to the summation order of atomically accumulated impulse adjoints)."""
    sc = SC.throw_worlds(num_worlds=2)
    th = sc.throw
    res = []
    for k in (1, 10, 60):
        g = Context("fp64", island_groups=1)
        try:
            g.upload(sc)
            g.fwd(DT, 60, k)
            after_fwd = g.bodies()
            gc = np.zeros((sc.num_bodies, 3))
            gc[th["stack"][0]] = 1.0
            out = g.bwd(gc)
            # cc_bwd leaves the context in the state cc_fwd left it in
            assert np.array_equal(g.bodies()["position"], after_fwd["position"])
            res.append(np.concatenate([out["v0"].ravel(), out["friction"]]))
        finally:
            g.close()
    for r in res[1:]:
        np.testing.assert_allclose(r, res[0], rtol=1e-9, atol=1e-12)


def test_autograd_function_matches_cc_bwd():
    sc, body, mover = _slide()
    n = sc.num_bodies
    b = sc.bodies()
    v0 = torch.tensor(b["linear_velocity"].copy(), requires_grad=True)
    mu = torch.tensor(b["friction"].copy(), requires_grad=True)
    g = Context("fp64", island_groups=1)
    try:
        com = WorldRun.apply(v0, mu, g, sc, DT, 30, 7)
        loss = (com[body] * torch.tensor(W3)).sum()
        loss.backward()
    finally:
        g.close()
    adj = _adjoint(_slide, 30, "fp64")
    np.testing.assert_allclose(v0.grad.numpy()[mover], adj[:3], rtol=1e-12)
    assert mu.grad.numpy().sum() == pytest.approx(adj[3], rel=1e-12)
    assert mu.grad.shape == (n,)


def test_bwd_state_rules():
    """cc_bwd needs a cc_fwd with nothing in between; differentiable stepping rejects split
    worlds, sleeping and point constraints."""
    sc, _, _ = _slide()
    g = Context("fp64", island_groups=1)
    try:
        g.upload(sc)
        g.fwd(DT, 5, 2)
        g.step(DT, 1)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_STATE}"):
            g.bwd(np.zeros((sc.num_bodies, 3)))
    finally:
        g.close()
    g = Context("fp64")  # a single world of few bodies is split into island groups by default
    try:
        big = SC.cube_heap(seed=1001, nx=5, ny=5, nz=5)
        big.settings.allow_sleeping = 0
        g.upload(big)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_UNSUPPORTED}"):
            g.fwd(DT, 2, 1)
    finally:
        g.close()
    g = Context("fp64", island_groups=1)
    try:
        sleepy = SC.cube_heap(seed=1001, nx=2, ny=2, nz=2)
        g.upload(sleepy)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_UNSUPPORTED}"):
            g.fwd(DT, 2, 1)
        chains = SC.batched_worlds(seed=1003, num_worlds=1, min_bodies=8, max_bodies=9)
        chains.settings.allow_sleeping = 0
        g.upload(chains)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_UNSUPPORTED}"):
            g.fwd(DT, 2, 1)
    finally:
        g.close()
