"""Pins of the oracle's solver details that the closed-form tests leave open.

Each test is built so that a plausible mistake in the oracle fails it: the contact sort key
against literal values worked out with JavaScript's int32 semantics, the tangent basis on
both sides of its branch in closed form, the warm-start impulse match (both body-space points
within 1 cm, first match wins, slots already overwritten compared again), the dt / prevDt
warm-start ratio, the Baumgarte position step of one contact point in closed form, external
forces and torques, and the per-island iteration counts of the reference's islands test.
"""
import math

import numpy as np
import pytest

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH
from oracle import refcpu as O

DT = 1.0 / 60.0
EMPTY = 0xFFFFFFFF


# ------------------------------------------------------------------ sort key
def _js_sort_key(a, b, sa, sb):
    """computeContactSortKey (contact-constraints.ts:1222-1231) with JavaScript semantics:
    `(hash * 31 + x) | 0` is ToInt32 of the exact product-sum, `>>> 0` makes it unsigned."""
    h = 0
    for x in (a, b, sa, sb):
        v = (h * 31 + x) & 0xFFFFFFFF
        h = v - (1 << 32) if v >= (1 << 31) else v
    return h & 0xFFFFFFFF


@pytest.mark.parametrize("args,expected", [
    ((0, 1, EMPTY, EMPTY), 0x3A1),
    ((1, 2, EMPTY, EMPTY), 0x7BC1),
    ((7, 3, 5, EMPTY), 0x33A76),
    ((12345, 67890, EMPTY, EMPTY), 0x19CF3DB9),       # wraps past 2^31 more than once
    ((1000000, 1000001, EMPTY, EMPTY), 0x28F64BA1),
    ((8388607, 8388606, 0, 0), 0x0FFF841F),
    ((2, 1, 0x7FFFFFFF, 3), 0x8000EC63),               # high bit set: the unsigned reading matters
    ((0, 0, 0, 0), 0),
])
def test_contact_sort_key_literals(args, expected):
    assert _js_sort_key(*args) == expected  # the literal is what JavaScript computes
    assert O.sort_key(*args) == expected


def test_contact_sort_key_hand_worked():
    # (0, 1, EMPTY, EMPTY): h = 0 -> 1 -> (31 + 2^32 - 1) | 0 = 30 -> (930 + 2^32 - 1) | 0 = 929
    assert O.sort_key(0, 1, EMPTY, EMPTY) == 929


# ------------------------------------------------------------------ tangent basis
@pytest.mark.parametrize("n,t1,t2", [
    ((0.6, 0.0, 0.8), (0.8, 0.0, -0.6), (0.0, 1.0, 0.0)),          # |nx| > |ny|: x-z plane
    ((-0.6, 0.0, 0.8), (0.8, 0.0, 0.6), (0.0, 1.0, 0.0)),
    ((0.0, 0.6, 0.8), (0.0, 0.8, -0.6), (-1.0, 0.0, 0.0)),         # else branch: y-z plane
    ((0.0, 1.0, 0.0), (0.0, 0.0, -1.0), (-1.0, 0.0, 0.0)),         # floor normal
    ((1 / math.sqrt(2), 1 / math.sqrt(2), 0.0), (0.0, 0.0, -1.0),  # tie |nx| == |ny| takes the else branch
     (-1 / math.sqrt(2), 1 / math.sqrt(2), 0.0)),
])
def test_contact_tangents_closed_form(n, t1, t2):
    a, b = O.contact_tangents(n)
    np.testing.assert_allclose(a, t1, atol=1e-15)
    np.testing.assert_allclose(b, t2, atol=1e-15)


def test_contact_tangents_orthonormal_right_handed():
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        t1, t2 = O.contact_tangents(n)
        assert abs(np.dot(t1, n)) < 1e-15 and abs(np.dot(t2, n)) < 1e-15 and abs(np.dot(t1, t2)) < 1e-15
        assert np.linalg.norm(t1) == pytest.approx(1, abs=1e-15)
        np.testing.assert_allclose(np.cross(n, t1), t2, atol=0)


# ------------------------------------------------------------------ warm start
def _resting_box(he=0.5, friction=0.5):
    sc = SC.Scene()
    sc.add_body(SH.box((20, 0.5, 20)), (0, -0.5, 0), motion_type=_abi.CC_MOTION_STATIC, friction=friction)
    sc.add_body(SH.box((he, he, he), convex_radius=min(0.05, he / 2)), (0, he + 0.005, 0), friction=friction)
    sc.settings.allow_sleeping = 0
    return sc


def _settled(sc, steps=90):
    w = O.OracleWorld(sc)
    w.step(DT, steps)
    return w


def _no_velocity_iterations(w, sc):
    """From the next step on, the stored impulses are the warm-start impulses themselves
    (warmStart :500-537 scales them, no sweep changes them, storeAppliedImpulses copies them)."""
    s = _abi.Settings.from_buffer_copy(bytes(memoryview(sc.settings)))
    s.velocity_iterations = 0
    s.position_iterations = 0
    s.gravity_enabled = 0
    w.set_settings(s)


def _shift_box(w, dx):
    b = w.bodies()
    b["position"][1, 0] += dx
    b["linear_velocity"][1] = 0
    b["angular_velocity"][1] = 0
    w.set_bodies(b)


@pytest.mark.parametrize("dx,carried", [(0.009, True), (0.011, False)])
def test_warm_start_match_needs_both_points_within_1cm(dx, carried):
    """contact-constraints.ts:651-686: a cached point's impulses carry over only if BOTH its
    body-space points are within sqrt(1e-4) = 1 cm of the new ones.  Sliding the box by dx moves
    the floor-side (body A) points by dx and leaves the box-side points in place."""
    sc = _resting_box()
    w = _settled(sc)
    before = w.contacts()
    assert len(before) == 1 and before[0]["num_points"] == 4
    ln0 = before[0]["lambda_n"][:4].copy()
    assert np.all(ln0 > 0)
    _no_velocity_iterations(w, sc)
    _shift_box(w, dx)
    w.step(DT)  # same dt: ratio 1
    after = w.contacts()
    assert len(after) == 1 and after[0]["num_points"] == 4
    pa0 = before[0]["position_a"][:4]
    pa1 = after[0]["position_a"][:4]
    pb0, pb1 = before[0]["position_b"][:4], after[0]["position_b"][:4]
    # same corners, floor points moved by dx
    np.testing.assert_allclose(pb1 - pb0, np.tile([dx, 0, 0], (4, 1)), atol=1e-9)
    if carried:
        np.testing.assert_array_equal(after[0]["lambda_n"][:4], ln0)
        np.testing.assert_array_equal(after[0]["lambda_t1"][:4], before[0]["lambda_t1"][:4])
    else:  # an OR of the two tests, or a body-space test on one body only, would carry them
        np.testing.assert_array_equal(after[0]["lambda_n"][:4], 0)
        np.testing.assert_array_equal(after[0]["lambda_t1"][:4], 0)
    assert np.abs(pa1 - pa0).max() < 0.02


def test_warm_start_first_match_wins_in_place_slots():
    """A box of 8 mm side: its four corners are within 1 cm of their neighbours.  The reference
    compares each new point with the contact's slots in order and takes the FIRST match, and the
    slots below the new point's index already hold this step's points (written in place), so
    every point takes the impulses of cached slot 0.  A best-match or an own-slot rule would
    keep four different impulses."""
    sc = _resting_box(he=0.004, friction=0.5)
    w = _settled(sc, 120)
    before = w.contacts()
    assert len(before) == 1 and before[0]["num_points"] == 4
    ln0 = before[0]["lambda_n"][:4].copy()
    assert len(set(ln0.tolist())) == 4, "the cached impulses must differ for this test to decide"
    pts = before[0]["position_b"][:4]
    d01 = np.linalg.norm(pts[0] - pts[1])
    assert d01 < 0.01
    _no_velocity_iterations(w, sc)
    w.step(DT)
    after = w.contacts()
    assert after[0]["num_points"] == 4
    np.testing.assert_array_equal(after[0]["lambda_n"][:4], np.full(4, ln0[0]))


def test_warm_start_ratio_dt_over_prev_dt():
    """update.ts:123-124: ratio = dt / prevDt, 0 on the first step (prevDt = 0)."""
    sc = _resting_box()
    # first step: no warm start, so with no sweeps the stored impulses are 0
    first = O.OracleWorld(sc)
    _no_velocity_iterations(first, sc)
    first.step(DT)
    assert first.contacts()[0]["num_points"] == 4
    np.testing.assert_array_equal(first.contacts()[0]["lambda_n"][:4], 0)
    w = _settled(sc)
    ln0 = w.contacts()[0]["lambda_n"][:4].copy()
    _no_velocity_iterations(w, sc)
    w.step(DT / 2)
    np.testing.assert_array_equal(w.contacts()[0]["lambda_n"][:4], ln0 * 0.5)
    w.step(DT / 2)  # ratio 1 now
    np.testing.assert_array_equal(w.contacts()[0]["lambda_n"][:4], ln0 * 0.5)
    w.step(DT)      # ratio 2
    np.testing.assert_array_equal(w.contacts()[0]["lambda_n"][:4], ln0)


# ------------------------------------------------------------------ Baumgarte position step
def _quat_between(u, v):
    u = np.asarray(u, float) / np.linalg.norm(u)
    v = np.asarray(v, float) / np.linalg.norm(v)
    a = np.cross(u, v)
    q = np.array([*a, 1 + np.dot(u, v)])
    return q / np.linalg.norm(q)


def _rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("depth", [0.05, 0.35])
def test_baumgarte_step_of_one_point_closed_form(depth):
    """solvePositionConstraintWithMassOverride (axis-constraint-part.ts:790-847) for a box
    standing on one corner, sunk by `depth`: separation C = max(-(depth) + slop, -maxPen),
    lambda = -effMass * beta * C, the box moves by lambda * invM along n and turns by
    I^-1 (r x n) lambda, with r from the centre of mass to the midpoint of the two points
    and effMass = 1 / (invM + (r x n) . I^-1 (r x n)).  depth 0.35 checks the 0.2 m clamp."""
    he = 0.5
    box = SH.box((he, he, he))
    q = _quat_between((1, 1, 1), (0, -1, 0))  # the (+,+,+) corner points straight down
    R = _rot(q)
    corner = R @ np.array([he, he, he])
    com = np.array([0.0, -corner[1] - depth, 0.0])  # corner at y = -depth
    sc = SC.Scene()
    sc.add_body(SH.box((20, 0.5, 20)), (0, -0.5, 0), motion_type=_abi.CC_MOTION_STATIC)
    sc.add_body(box, tuple(com), quaternion=tuple(q))
    s = sc.settings
    s.gravity_enabled = 0
    s.velocity_iterations = 0
    s.position_iterations = 1
    s.warm_starting = 0
    s.allow_sleeping = 0
    w = O.OracleWorld(sc)
    w.step(DT)
    c = w.contacts()
    assert len(c) == 1 and c[0]["num_points"] == 1
    n = np.asarray(c[0]["normal"])
    np.testing.assert_allclose(n, [0, 1, 0], atol=1e-6)  # EPA stops within its tolerance
    pa, pb = c[0]["position_a"][0], c[0]["position_b"][0]
    slop, beta, max_pen = s.penetration_slop, s.baumgarte_factor, s.max_penetration_distance
    C = max(np.dot(pb - pa, n) + slop, -max_pen)
    assert C < 0
    mid = 0.5 * (pa + pb)
    r = mid - com
    inv_m = 1.0 / box.mass
    inv_i_world = R @ np.linalg.inv(box.inertia) @ R.T
    rxn = np.cross(r, n)
    eff = 1.0 / (inv_m + rxn @ inv_i_world @ rxn)
    lam = -eff * beta * C
    b = w.bodies()[1]
    np.testing.assert_allclose(b["center_of_mass_position"] - com, lam * inv_m * n, rtol=0, atol=1e-12)
    assert abs(lam * inv_m) > 1e-3  # the step moved it measurably
    # the turn: axis-angle of I^-1 (r x n) lambda applied to q
    rot = inv_i_world @ rxn * lam
    ang = np.linalg.norm(rot)
    ax = rot / ang
    dq = np.array([*(ax * math.sin(ang / 2)), math.cos(ang / 2)])
    x1, y1, z1, w1 = dq
    x2, y2, z2, w2 = q
    qn = np.array([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                   w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])
    qn /= np.linalg.norm(qn)
    assert abs(abs(np.dot(qn, b["quaternion"])) - 1) < 1e-12


# ------------------------------------------------------------------ forces and torques
def test_force_and_torque_integrate_then_clear():
    """update.ts:253-292: v += (F/m + g) dt then damping; w += I_w^-1 tau dt then damping;
    forces are cleared after the step (:234-238), so the next step sees gravity only."""
    sc = SC.Scene()
    box = SH.box((0.5, 0.25, 0.125))
    q = _quat_between((1, 0, 0), (1, 2, 3))
    sc.add_body(box, (0, 50, 0), quaternion=tuple(q), linear_damping=0.05, angular_damping=0.05)
    sc.settings.allow_sleeping = 0
    w = O.OracleWorld(sc)
    F = np.array([[30.0, -12.0, 7.0]])
    tau = np.array([[1.5, -2.0, 0.5]])
    w.add_forces(F, tau)
    w.step(DT)
    b = w.bodies()[0]
    m = box.mass
    g = np.array([0, -9.81, 0])
    f = 1 - 0.05 * DT
    v1 = (F[0] / m + g) * DT * f
    R = _rot(q)
    inv_i = R @ np.linalg.inv(box.inertia) @ R.T
    w1 = (inv_i @ tau[0]) * DT * f
    np.testing.assert_allclose(b["linear_velocity"], v1, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(b["angular_velocity"], w1, rtol=1e-12, atol=1e-15)
    w.step(DT)  # cleared: gravity only
    b2 = w.bodies()[0]
    np.testing.assert_allclose(b2["linear_velocity"], (v1 + g * DT) * f, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(b2["angular_velocity"], w1 * f, rtol=1e-12, atol=1e-15)


def test_force_on_static_body_is_ignored_and_wakes_sleeping():
    """addForce touches dynamic bodies only (rigid-body.ts:840-869); with wake it wakes a
    sleeping body (sleep.ts:188-203)."""
    sc = _floor_and_box()
    w = O.OracleWorld(sc)
    w.step(DT, 200)
    assert w.bodies()["sleeping"][1] == 1
    F = np.zeros((2, 3))
    F[0] = [0, 1e6, 0]   # the static floor: ignored
    F[1] = [0, 3000.0, 0]
    w.add_forces(F, None, wake=True)
    assert w.bodies()["sleeping"][1] == 0
    y0 = w.bodies()["position"][1, 1]
    w.step(DT)
    assert w.bodies()["linear_velocity"][1, 1] > 0 and w.bodies()["position"][1, 1] > y0
    np.testing.assert_array_equal(w.bodies()["position"][0], [0, -0.5, 0])


def _floor_and_box():
    sc = SC.Scene()
    sc.add_body(SH.box((20, 0.5, 20)), (0, -0.5, 0), motion_type=_abi.CC_MOTION_STATIC)
    sc.add_body(SH.box((0.5, 0.5, 0.5)), (0, 0.51, 0))
    return sc


# ------------------------------------------------------------------ iteration overrides
def test_island_iteration_counts_of_the_reference_test():
    """tst/islands.test.ts:71-137: world 10/5; an island of a default body (0) and a body with
    20/15 solves 20/15; a lone body with 8/3 solves 8/3 (no member uses 0, so no default)."""
    sc = SC.Scene()
    sc.settings.velocity_iterations = 10
    sc.settings.position_iterations = 5
    s = SH.sphere(0.5)
    a = sc.add_body(s, (0, 0.6, 0))
    b = sc.add_body(s, (0, 0, 0), num_velocity_steps_override=20, num_position_steps_override=15)
    c = sc.add_body(s, (10, 0, 0), num_velocity_steps_override=8, num_position_steps_override=3)
    w = O.OracleWorld(sc)
    w.step(DT)
    assert w.bodies()["island"][a] == w.bodies()["island"][b]
    assert w.island_iterations(a) == (20, 15)
    assert w.island_iterations(b) == (20, 15)
    assert w.island_iterations(c) == (8, 3)


def test_island_iteration_counts_point_constraints_and_below_default():
    """Point-constraint overrides count like bodies (islands.ts:366-374); an override below the
    world default wins when nothing in the island uses 0."""
    sc = SC.Scene()
    cap = SH.capsule(0.1, 0.3)
    p = sc.add_body(cap, (0, 5, 0), num_velocity_steps_override=4, num_position_steps_override=1)
    q = sc.add_body(cap, (0, 5.8, 0), num_velocity_steps_override=3, num_position_steps_override=2)
    sc.add_point_constraint(p, q, (0, 0.4, 0), (0, -0.4, 0), num_velocity_steps_override=6,
                            num_position_steps_override=0)
    w = O.OracleWorld(sc)
    w.step(DT)
    # velocity: max(4, 3, 6) = 6, nothing uses 0; position: max(1, 2, 0) and the constraint uses 0 -> 2 defaults
    assert w.island_iterations(p) == (6, max(2, sc.settings.position_iterations))
