"""Step inputs of crashcat's updateWorld on the HIP path, FP64 bit for bit against the oracle:
external forces and torques (update.ts:245-319, cleared at :234-238), per-island iteration
overrides (islands.ts:344-386), moving kinematic bodies in worlds solved as one workgroup and
as several island groups, the warm-start match after moving a body between steps and a
changing dt; and the upload rejections (sensors, an unknown motion quality, overrides out of range, a failed upload
leaving no scene behind)."""
import ctypes as C

import numpy as np
import pytest

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH
from crashcat_amd._lib import Context, CrashcatError, lib
from oracle import refcpu as O
from tests.test_gpu_parity import _assert_bodies_equal, _assert_contacts_equal, _same_partition

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _pile(seed, n=120, kinematic=0, overrides=False):
    rng = np.random.default_rng(seed)
    sc = SC.Scene()
    sc.add_body(SH.box((20, 0.5, 20)), (0, -0.5, 0), motion_type=_abi.CC_MOTION_STATIC, friction=0.5)
    shapes = [SH.box((0.3, 0.3, 0.3)), SH.sphere(0.3), SH.capsule(0.2, 0.3), SH.box((0.5, 0.2, 0.3))]
    for i in range(n):
        pos = (rng.uniform(-3, 3), rng.uniform(0.5, 6), rng.uniform(-3, 3))
        kw = {}
        if overrides and i % 7 == 0:
            kw = dict(num_velocity_steps_override=int(rng.integers(1, 25)),
                      num_position_steps_override=int(rng.integers(0, 6)))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        sc.add_body(shapes[i % 4], pos, quaternion=tuple(q), friction=0.5, linear_velocity=tuple(rng.normal(0, 1, 3)),
                    **kw)
    for k in range(kinematic):
        z = -2.0 + 4.0 * k / max(kinematic - 1, 1)
        sc.add_body(SH.box((0.4, 0.6, 0.4)), (-6.0, 0.7, z), motion_type=_abi.CC_MOTION_KINEMATIC,
                    linear_velocity=(2.5, 0.0, 0.0), angular_velocity=(0.0, 1.5 * (k + 1), 0.0))
    return sc


def _run_lockstep(sc, steps, ctx, before_step=None, check_every=1):
    orc = O.OracleWorld(sc)
    g = Context("fp64", **ctx)
    g.upload(sc)
    try:
        for s in range(1, steps + 1):
            if before_step:
                before_step(s, orc, g)
            orc.step(DT)
            g.step(DT, 1)
            if s % check_every == 0 or s == steps:
                g.sync()
                ob, gb = orc.bodies(), g.bodies()
                _assert_bodies_equal(ob, gb, f"step {s}")
                assert np.array_equal(orc.pairs(), g.pairs()), f"step {s}: pairs"
                _assert_contacts_equal(orc.contacts(), g.contacts(), f"step {s}")
                assert _same_partition(ob["island"], gb["island"]), f"step {s}: islands"
        return orc, g
    finally:
        g.close()


def test_fp64_forces_and_torques_bitwise():
    """Random forces and torques on a third of the bodies every step (a thruster pile), plus
    initial accumulators in the descriptors; the GPU integrates and clears them as the oracle."""
    sc = _pile(11, n=100)
    b = sc.bodies()
    rng = np.random.default_rng(3)
    b["force"][1::5] = rng.normal(0, 200, (len(b["force"][1::5]), 3))
    b["torque"][1::5] = rng.normal(0, 20, (len(b["torque"][1::5]), 3))
    n = len(b)

    def push(s, orc, g):
        r = np.random.default_rng(1000 + s)
        f = np.zeros((n, 3))
        t = np.zeros((n, 3))
        f[s % 3::3] = r.normal(0, 300, f[s % 3::3].shape)
        t[s % 3::3] = r.normal(0, 30, t[s % 3::3].shape)
        f[0] = [1e6, 1e6, 1e6]  # the static floor: ignored by both
        orc.add_forces(f, t, wake=True)
        g.add_forces(f, t, wake=True)

    _run_lockstep(sc, 50, {}, push, check_every=5)


def test_fp64_iteration_overrides_bitwise():
    """Bodies with velocity/position overrides (1..24 / 0..5) among default bodies, and capsule
    chains whose point constraints carry overrides: every island solves its own counts."""
    sc = _pile(12, n=100, overrides=True)
    cap = SH.capsule(0.1, 0.3)
    prev = None
    for k in range(6):
        i = sc.add_body(cap, (4.0, 6.0 - 0.8 * k, 0.0), quaternion=(0.0, 0.0, 0.0, 1.0))
        if prev is not None:
            sc.add_point_constraint(prev, i, (0, -0.4, 0), (0, 0.4, 0), num_velocity_steps_override=3 * k,
                                    num_position_steps_override=k % 3)
        prev = i
    _run_lockstep(sc, 60, {}, check_every=6)


@pytest.mark.parametrize("groups", [0, 1])
def test_fp64_moving_kinematic_bodies_bitwise(groups):
    """Two kinematic boxes sweep through the pile while spinning.  Awake kinematic bodies join
    the islands of the dynamic bodies they touch (linkContactConstraints, islands.ts:184-204),
    so the workgroup that solves that island also integrates them; groups = 0 splits the world
    into island groups, 1 keeps it on one workgroup."""
    sc = _pile(13, n=150, kinematic=2)
    _run_lockstep(sc, 80, {"island_groups": groups}, check_every=4)


def test_fp64_warm_start_after_moving_a_body():
    """A settled box slid by 0.9 cm between steps keeps its impulses, by 1.1 cm loses them
    (contact-constraints.ts:651-686), on the GPU as in the oracle."""
    sc = SC.Scene()
    sc.add_body(SH.box((20, 0.5, 20)), (0, -0.5, 0), motion_type=_abi.CC_MOTION_STATIC, friction=0.5)
    sc.add_body(SH.box((0.5, 0.5, 0.5)), (0, 0.505, 0), friction=0.5)
    sc.add_body(SH.box((0.5, 0.5, 0.5)), (3, 0.505, 0), friction=0.5)
    sc.settings.allow_sleeping = 0

    def move(s, orc, g):
        if s in (60, 61):
            for w in (orc, g):
                b = w.bodies()
                b["position"][1, 0] += 0.009
                b["position"][2, 0] += 0.011
                w.set_bodies(b)

    _run_lockstep(sc, 70, {}, move)


def test_fp64_changing_dt_bitwise():
    """The warm-start ratio dt / prevDt (update.ts:123-124) over steps of changing length."""
    sc = _pile(14, n=60)
    orc = O.OracleWorld(sc)
    g = Context("fp64")
    g.upload(sc)
    try:
        for s, dt in enumerate([DT, DT, DT / 2, DT / 2, DT * 1.5, DT, DT / 3] * 5):
            orc.step(dt)
            g.step(dt, 1)
            g.sync()
            _assert_bodies_equal(orc.bodies(), g.bodies(), f"step {s}")
            _assert_contacts_equal(orc.contacts(), g.contacts(), f"step {s}")
    finally:
        g.close()


# ------------------------------------------------------------------ rejections
def _one_body(**kw):
    sc = SC.Scene()
    sc.add_body(SH.box((0.5, 0.5, 0.5)), (0, 1, 0), **kw)
    return sc


@pytest.mark.parametrize("kw,code", [
    (dict(sensor=True), _abi.CC_E_UNSUPPORTED),
    (dict(motion_quality=2), _abi.CC_E_ARG),
    (dict(num_velocity_steps_override=256), _abi.CC_E_ARG),
    (dict(num_position_steps_override=-1), _abi.CC_E_ARG),
])
def test_unsupported_body_features_are_rejected(kw, code):
    g = Context("fp32")
    try:
        with pytest.raises(CrashcatError, match=f"code {code}"):
            g.upload(_one_body(**kw))
    finally:
        g.close()


def test_failed_upload_leaves_no_scene():
    """An upload that fails after a good one leaves the context without a scene: stepping it
    reports CC_E_STATE instead of running on a half-built context."""
    g = Context("fp32")
    try:
        g.upload(_one_body())
        g.step(DT, 1)
        g.sync()
        with pytest.raises(CrashcatError):
            g.upload(_one_body(sensor=True))
        rc = lib().cc_step(g.h, C.c_double(DT), 1)
        assert rc == _abi.CC_E_STATE
        g.upload(_one_body())  # a later good upload works again
        g.step(DT, 2)
        g.sync()
        assert g.bodies()["position"][0, 1] < 1.0
    finally:
        g.close()
