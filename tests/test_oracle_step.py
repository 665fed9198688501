"""Whole-step checks of the FP64 oracle against what the mathematics fixes.

Closed forms (free fall with damping, constant-axis rotation), conservation and invariants
of the contact solve (resting load, restitution, Coulomb friction, lambda bounds), the pair
set against a brute-force restatement of the broadphase predicate, islands against connected
components, and the reference's behavioural tests (islands.test.ts, sleep.test.ts).
"""
import math

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH
from oracle import refcpu as O

DT = 1.0 / 60.0
ID = (0.0, 0.0, 0.0, 1.0)


def _quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def _axis_angle(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    return np.array([*(a * math.sin(angle / 2)), math.cos(angle / 2)])


def test_free_fall_closed_form():
    """accelerationIntegrationUpdate (src/update.ts:245-319) and the position integration of
    src/body/rigid-body-step.ts: v_k = (v_{k-1} + g dt)(1 - c dt), com_k = com_{k-1} + v_k dt."""
    sc = SC.Scene()
    v0 = np.array([1.5, 4.0, -2.0])
    sc.add_body(SH.sphere(0.5), (0, 100, 0), linear_velocity=tuple(v0), linear_damping=0.05)
    w = O.OracleWorld(sc)
    g = np.array([0, -9.81, 0])
    v = v0.copy()
    com = np.array([0, 100.0, 0])
    f = max(0.0, 1 - 0.05 * DT)
    for _ in range(120):
        w.step(DT)
        v = (v + g * DT) * f
        com = com + v * DT
    b = w.bodies()[0]
    np.testing.assert_allclose(b["linear_velocity"], v, rtol=0, atol=1e-12)
    np.testing.assert_allclose(b["center_of_mass_position"], com, rtol=0, atol=1e-10)


def test_velocity_clamp_and_locked_dofs():
    """Velocity clamped to max_linear_velocity; locked translation axes zeroed (update.ts:267-282)."""
    sc = SC.Scene()
    sc.add_body(SH.sphere(0.5), (0, 0, 0), linear_velocity=(300, 0, 400), max_linear_velocity=50.0,
                linear_damping=0.0)
    sc.add_body(SH.sphere(0.5), (10, 0, 0), linear_velocity=(3, 2, 1), allowed_dofs=0b111101, linear_damping=0.0)
    sc.settings.gravity_enabled = 0
    w = O.OracleWorld(sc)
    w.step(DT)
    b = w.bodies()
    np.testing.assert_allclose(b["linear_velocity"][0], [30, 0, 40], atol=1e-12)
    np.testing.assert_array_equal(b["linear_velocity"][1], [3, 0, 1])


def test_rotation_about_fixed_axis_closed_form():
    """With a sphere (isotropic inertia) and no contacts the axis of omega is fixed: after K
    steps the rotation is axis-angle(axis, |w0| dt sum_k f^k) applied to q0, f = 1 - c dt."""
    sc = SC.Scene()
    w0 = np.array([0.3, -1.2, 0.7])
    q0 = _axis_angle((1, 2, 3), 0.4)
    sc.add_body(SH.sphere(0.5), (0, 0, 0), quaternion=tuple(q0), angular_velocity=tuple(w0), angular_damping=0.05)
    sc.settings.gravity_enabled = 0
    sc.settings.allow_sleeping = 0
    w = O.OracleWorld(sc)
    K = 90
    w.step(DT, K)
    f = 1 - 0.05 * DT
    angle = np.linalg.norm(w0) * DT * sum(f ** k for k in range(1, K + 1))
    q = _quat_mul(_axis_angle(w0, angle), q0)
    b = w.bodies()[0]
    assert abs(abs(np.dot(b["quaternion"], q)) - 1) < 1e-12
    np.testing.assert_allclose(b["angular_velocity"], w0 * f ** K, atol=1e-13)


def _floor_scene(*, friction=0.5, restitution=0.0):
    sc = SC.Scene()
    sc.add_body(SH.box((20, 0.5, 20)), (0, -0.5, 0), motion_type=_abi.CC_MOTION_STATIC, friction=friction,
                restitution=restitution)
    return sc


def test_resting_box_carries_its_weight():
    """A box settled on the floor: v_y -> 0 and the normal impulses of its manifold add up
    to m g dt per step; four points (a face-face manifold); lambda_n >= 0."""
    sc = _floor_scene()
    box = SH.box((0.5, 0.5, 0.5))
    sc.add_body(box, (0, 0.52, 0), friction=0.5)
    sc.settings.allow_sleeping = 0
    w = O.OracleWorld(sc)
    w.step(DT, 120)
    b = w.bodies()[1]
    assert abs(b["linear_velocity"][1]) < 1e-3
    assert b["position"][1] == pytest.approx(0.5, abs=0.03)
    c = w.contacts()
    assert len(c) == 1 and c[0]["num_points"] == 4
    ln = c[0]["lambda_n"][:4]
    assert np.all(ln >= 0)
    # gravity impulse per step (the damping factor removes a sliver of it)
    assert ln.sum() == pytest.approx(box.mass * 9.81 * DT, rel=2e-3)
    n = c[0]["normal"]
    assert np.linalg.norm(n) == pytest.approx(1, abs=1e-12) and abs(n[1]) == pytest.approx(1, abs=1e-9)


def test_restitution_head_on():
    """Two touching spheres approaching at 4 m/s with restitution 0.5 (both bodies, so any
    combine mode gives 0.5): the normal relative velocity reverses to -e times its value
    (src/constraints/contact-constraints.ts:357-411); momentum is conserved."""
    sc = SC.Scene()
    sc.settings.gravity_enabled = 0
    sc.add_body(SH.sphere(0.5), (-0.499, 0, 0), linear_velocity=(2, 0, 0), restitution=0.5, linear_damping=0.0)
    sc.add_body(SH.sphere(0.5), (0.499, 0, 0), linear_velocity=(-2, 0, 0), restitution=0.5, linear_damping=0.0)
    w = O.OracleWorld(sc)
    w.step(DT)
    b = w.bodies()
    v = b["linear_velocity"][:, 0]
    assert v[1] - v[0] == pytest.approx(0.5 * 4.0, rel=1e-9)
    assert v[0] + v[1] == pytest.approx(0.0, abs=1e-12)


def test_coulomb_slide_to_stop():
    """A box sliding on the floor decelerates at about mu g (combined friction sqrt(0.5*0.5))
    and stops near v0^2 / (2 mu g)."""
    sc = _floor_scene(friction=0.5)
    sc.add_body(SH.box((0.5, 0.5, 0.5)), (0, 0.5, 0), linear_velocity=(3, 0, 0), friction=0.5, linear_damping=0.0)
    sc.settings.allow_sleeping = 0
    w = O.OracleWorld(sc)
    vs = []
    for _ in range(90):
        w.step(DT)
        vs.append(w.bodies()[1]["linear_velocity"][0])
    vs = np.array(vs)
    # steady deceleration during the slide
    dec = -(vs[30] - vs[10]) / (20 * DT)
    assert dec == pytest.approx(0.5 * 9.81, rel=0.03)
    assert abs(vs[-1]) < 1e-3
    x = w.bodies()[1]["position"][0]
    assert x == pytest.approx(3.0 ** 2 / (2 * 0.5 * 9.81), rel=0.05)
    c = w.contacts()[0]
    k = c["num_points"]
    lt = np.hypot(c["lambda_t1"][:k], c["lambda_t2"][:k])
    assert np.all(lt <= 0.5 * c["lambda_n"][:k] + 1e-12)


def _brute_pairs(mn, mx, motion, sleeping, layer, margin, table, nl):
    """The pair predicate of the broadphase, stated directly: a querying body (awake,
    non-static) pairs with every other body whose box overlaps its own box grown by the
    margin (inclusive), if the object-layer pair is enabled and one of the two is dynamic;
    two querying bodies pair once.
    src/broadphase/broadphase.ts:44-277, src/layers.ts, src/filter.ts:149-203."""
    n = len(motion)
    mn = mn.astype(np.float64)
    mx = mx.astype(np.float64)
    query = (motion != 0) & (sleeping == 0)
    out = set()
    for i in np.nonzero(query)[0]:
        lo = mn[i] - margin
        hi = mx[i] + margin
        ov = np.all((lo <= mx) & (mn <= hi), axis=1)
        ov[i] = False
        for j in np.nonzero(ov)[0]:
            if not table[layer[i] * 8 + layer[j]]:
                continue
            if query[j] and j < i:
                # the pair was (or will be) found from the lower-index querying body
                continue
            if motion[i] != 2 and motion[j] != 2:
                # kinematic-kinematic and kinematic-static need an opt-in (broadphase.ts:98-109)
                continue
            out.add((min(i, j), max(i, j)))
    return np.array(sorted(out), dtype=np.int32).reshape(-1, 2)


@pytest.mark.parametrize("seed,n", [(0, 1), (1, 2), (2, 40), (3, 400)])
def test_broadphase_pairs_brute_force(seed, n):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 10, (n, 3))
    h = rng.uniform(0.05, 1.0, (n, 3))
    mn = (c - h).astype(np.float32)
    mx = (c + h).astype(np.float32)
    motion = rng.integers(0, 3, n).astype(np.int32)
    sleeping = (rng.uniform(size=n) < 0.2).astype(np.int32) * (motion != 0)
    s, moving, fixed = SC.test_world_settings()
    layer = np.where(motion == 0, fixed, moving).astype(np.int32)
    got = O.broadphase_aabbs(mn, mx, motion, sleeping, layer, 0.02, s)
    got = np.array(sorted(map(tuple, np.sort(got, axis=1))), dtype=np.int32).reshape(-1, 2)
    want = _brute_pairs(mn, mx, motion, sleeping, layer, 0.02, list(s.object_layer_pairs), 8)
    np.testing.assert_array_equal(got, want)


def test_broadphase_touching_boxes_are_inclusive():
    """Boxes whose grown faces touch exactly overlap (inclusive test); 1 ulp further apart
    they do not."""
    s, moving, fixed = SC.test_world_settings()
    mn = np.array([[0, 0, 0], [1.5, 0, 0]], dtype=np.float32)
    mx = np.array([[1.0, 1, 1], [2.5, 1, 1]], dtype=np.float32)
    motion = np.array([2, 2], np.int32)
    zero = np.zeros(2, np.int32)
    lay = np.full(2, moving, np.int32)
    margin = float(np.float32(1.5) - np.float32(1.0))
    assert len(O.broadphase_aabbs(mn, mx, motion, zero, lay, margin, s)) == 1
    # the largest margin that leaves the grown box 1 ulp short of the other box
    short = float(np.nextafter(1.5, 0.0) - 1.0)
    assert len(O.broadphase_aabbs(mn, mx, motion, zero, lay, short, s)) == 0


def test_islands_group_connected_and_exclude_static():
    """islands.test.ts:6-58."""
    sc = SC.Scene()
    sph = SH.sphere(0.5)
    sc.add_body(sph, (0, 0.5, 0))
    sc.add_body(sph, (0, 0, 0))
    sc.add_body(sph, (5, 0, 0), motion_type=_abi.CC_MOTION_STATIC)
    sc.add_body(sph, (10, 0, 0))
    w = O.OracleWorld(sc)
    w.step(DT)
    isl = w.bodies()["island"]
    assert isl[2] == -1
    assert isl[0] == isl[1] and isl[0] >= 0
    assert isl[3] >= 0 and isl[3] != isl[0]


def test_islands_are_connected_components_of_the_contact_graph():
    """On a settling cube heap the island partition equals the connected components of the
    graph whose edges are the step's contact constraints between non-static bodies."""
    sc = SC.cube_heap(seed=1001, nx=4, ny=4, nz=4)
    w = O.OracleWorld(sc)
    w.step(DT, 80)
    bs = w.bodies()
    static = sc.bodies()["motion_type"] == 0
    c = w.contacts()
    e = [(a, b) for a, b in zip(c["body_a"], c["body_b"]) if not static[a] and not static[b]]
    n = len(bs)
    if e:
        e = np.array(e)
        g = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
    else:
        g = coo_matrix((n, n))
    _, comp = connected_components(g, directed=False)
    active = ~static & (bs["sleeping"] == 0)
    isl = bs["island"]
    assert np.all(isl[static] == -1)
    ids = np.nonzero(active)[0]
    for i in ids:
        for j in ids:
            if j <= i:
                continue
            assert (isl[i] == isl[j]) == (comp[i] == comp[j])


def test_sleep_after_rest_and_wake():
    """sleep.test.ts:346-386: a sphere at rest without gravity sleeps after
    ceil((0.5 + 0.1) * 60) steps, and a velocity change (impulse) wakes it."""
    sc = SC.Scene()
    sc.settings.gravity_enabled = 0
    sc.add_body(SH.sphere(1.0), (0, 0, 0))
    w = O.OracleWorld(sc)
    assert w.bodies()["sleeping"][0] == 0
    w.step(DT, math.ceil((0.5 + 0.1) * 60))
    st = w.bodies()
    assert st["sleeping"][0] == 1
    st["linear_velocity"][0] = (0, 5 * SH.sphere(1.0).mass ** -1, 0)
    st["sleeping"][0] = 0
    w.set_bodies(st)
    w.step(DT)
    assert w.bodies()["sleeping"][0] == 0


def test_sleep_timer_needs_half_a_second():
    """The body is still awake one step before time_before_sleep has elapsed."""
    sc = SC.Scene()
    sc.settings.gravity_enabled = 0
    sc.add_body(SH.sphere(1.0), (0, 0, 0))
    w = O.OracleWorld(sc)
    w.step(DT, 29)
    assert w.bodies()["sleeping"][0] == 0
    w.step(DT, 2)
    assert w.bodies()["sleeping"][0] == 1


def _chain_gaps(sc, bs):
    gaps = []
    for a, b, la, lb, *_ in sc.point_constraints:
        def world_pt(i, lp):
            x, y, z, ww = bs["quaternion"][i]
            r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * ww), 2 * (x * z + y * ww)],
                          [2 * (x * y + z * ww), 1 - 2 * (x * x + z * z), 2 * (y * z - x * ww)],
                          [2 * (x * z - y * ww), 2 * (y * z + x * ww), 1 - 2 * (x * x + y * y)]])
            return bs["position"][i] + r @ np.asarray(lp)
        gaps.append(np.linalg.norm(world_pt(a, la) - world_pt(b, lb)))
    return np.array(gaps)


def test_point_constraint_chain_in_free_space():
    """A chain of 8 capsules joined end to end (point-constraint.ts), kicked with random
    velocities and no gravity: the joints stay closed and the total linear momentum of the
    chain is conserved (the constraint impulses are internal)."""
    sc = SC.Scene()
    cap = SH.capsule(0.1, 0.3)
    rng = np.random.default_rng(2)
    q = (0.0, 0.0, math.sin(math.pi / 4), math.cos(math.pi / 4))
    prev = None
    for l in range(8):
        idx = sc.add_body(cap, ((l - 3.5) * 0.8, 0, 0), q, linear_velocity=tuple(rng.normal(0, 0.5, 3)),
                          angular_velocity=tuple(rng.normal(0, 0.5, 3)), linear_damping=0.0, angular_damping=0.0)
        if prev is not None:
            sc.add_point_constraint(prev, idx, (0.0, -0.4, 0.0), (0.0, 0.4, 0.0))
        prev = idx
    sc.settings.gravity_enabled = 0
    sc.settings.allow_sleeping = 0
    w = O.OracleWorld(sc)
    p0 = sc.bodies()["linear_velocity"].sum(axis=0)
    w.step(DT, 60)
    bs = w.bodies()
    assert _chain_gaps(sc, bs).max() < 5e-3
    np.testing.assert_allclose(bs["linear_velocity"].sum(axis=0), p0, atol=1e-9)


def test_point_constraint_chain_on_a_pile():
    """The capsule chains of the batched-worlds scene dropped onto their pile keep their joints
    within a few centimetres (Baumgarte position correction, two position iterations)."""
    sc = SC.batched_worlds(seed=1003, num_worlds=1, min_bodies=8, max_bodies=8, chains=2, chain_len=8)
    w = O.OracleWorld(sc)
    w.step(DT, 60)
    assert _chain_gaps(sc, w.bodies()).max() < 0.1


def test_worlds_step_on_threads_as_alone():
    """Separate oracle worlds stepped on separate threads at once (the CPU baseline's form)
    give exactly what each gives stepped alone: the oracle's scratch is per thread."""
    from concurrent.futures import ThreadPoolExecutor
    base = SC.mixed_pile(seed=1002, n=400, extent=(8.0, 6.0, 8.0), cylinders=True)
    batch = SC.replicate_world(base, 6, 1002, extent=(8.0, 6.0, 8.0))

    def run(w):
        o = O.OracleWorld(batch, w)
        o.step(DT, 40)
        return o.bodies()

    alone = [run(w) for w in range(6)]
    with ThreadPoolExecutor(6) as ex:
        threaded = list(ex.map(run, range(6)))
    for a, b in zip(alone, threaded):
        for f in ("position", "quaternion", "linear_velocity", "angular_velocity"):
            assert np.array_equal(a[f], b[f])
