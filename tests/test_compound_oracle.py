"""Compound shapes (src/shapes/compound.ts) in the oracle, pinned by what the geometry fixes:
a compound of one child at the identity transform is its child (bit for bit, sub-shape id
empty); each child of a compound touching its own neighbour makes that neighbour's contact
at the composed pose, with the child's index in its sub-shape id (pushIndex, src/body/
sub-shape.ts:43-64) and a compound against a compound pushes both; two cubes side by side have
the mass, centre of mass and inertia of the box they make; and a pile of compounds comes to rest.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from crashcat_amd import scene as SC
from crashcat_amd import shapes as S
from oracle import refcpu as R

DT = 1.0 / 60.0
Q_ID = (0.0, 0.0, 0.0, 1.0)


def _contacts_after_one_step(bodies, floor=True):
    """bodies: (shape, position, quaternion, motion_type); one oracle step, the step's contacts."""
    sc = SC.Scene()
    if floor:
        SC._floor(sc)
    for sh, p, q, mt in bodies:
        sc.add_body(sh, p, q, motion_type=mt, allow_sleeping=False)
    sc.end_world()
    o = R.OracleWorld(sc)
    o.step(DT)
    return o.contacts(), o.bodies()


def test_one_child_compound_is_its_child():
    for ch in (S.box((0.3, 0.2, 0.4)), S.sphere(0.3), S.capsule(0.2, 0.3), S.cylinder(0.3, 0.2)):
        q = Rotation.from_euler("xyz", [0.2, 0.5, -0.1]).as_quat()
        c1, b1 = _contacts_after_one_step([(S.compound([((0, 0, 0), Q_ID, ch)]), (0.1, 0.25, 0.0), q, 2)])
        c2, b2 = _contacts_after_one_step([(ch, (0.1, 0.25, 0.0), q, 2)])
        assert len(c1) == len(c2) == 1
        for f in ("num_points", "normal", "position_a", "position_b", "lambda_n", "sub_shape_a", "sub_shape_b"):
            assert np.array_equal(c1[f], c2[f]), f
        assert c1["sub_shape_b"][0] == 0xFFFFFFFF  # one child: pushIndex takes no bits


def test_children_make_their_own_contacts_with_pushed_ids():
    """A static compound of three boxes; a small dynamic box sits on each child.  Each contact is
    the one a standalone box at the child's composed pose makes (to rounding), with sub-shape id
    0xfffffffc | child on the compound's side."""
    r = Rotation.from_euler("xyz", [0.0, 0.3, 0.0])
    kids = [((-1.5, 0.0, 0.0), Q_ID, S.box((0.4, 0.2, 0.4))),
            ((0.0, 0.1, 0.2), Rotation.from_euler("y", 0.4).as_quat(), S.box((0.4, 0.2, 0.4))),
            ((1.5, -0.05, 0.0), Q_ID, S.box((0.5, 0.2, 0.3)))]
    comp = S.compound(kids)
    p0 = np.array([0.2, 0.0, -0.1])
    tops = []
    for pos, q, ch in kids:
        wp = p0 + r.apply(pos)
        wq = r * Rotation.from_quat(q)
        tops.append((wp, wq, ch))
    small = S.box((0.1, 0.1, 0.1))
    dyn = [(small, tuple(wp + wq.apply((0.0, 0.2 + 0.1 - 0.005, 0.0))), tuple(wq.as_quat()), 2) for wp, wq, _ in tops]
    c1, _ = _contacts_after_one_step([(comp, tuple(p0), tuple(r.as_quat()), 0)] + dyn, floor=False)
    assert len(c1) == 3
    for k in range(3):
        assert c1["body_a"][k] == 0 and c1["body_b"][k] == k + 1
        assert c1["sub_shape_a"][k] == (0xFFFFFFFC | k) and c1["sub_shape_b"][k] == 0xFFFFFFFF
        wp, wq, ch = tops[k]
        c2, _ = _contacts_after_one_step([(ch, tuple(wp), tuple(wq.as_quat()), 0), dyn[k]], floor=False)
        n = c2["num_points"][0]
        assert c1["num_points"][k] == n == 4
        assert np.allclose(c1["normal"][k], c2["normal"][0], atol=1e-12)
        pa1, pa2 = c1["position_a"][k][:n], c2["position_a"][0][:n]
        for p in pa1:
            assert np.min(np.linalg.norm(pa2 - p, axis=1)) < 1e-9


def test_compound_vs_compound_pushes_both_ids():
    """Two compounds of two spheres; child 1 of A touches child 0 of B only."""
    a = S.compound([((-1.0, 0, 0), Q_ID, S.sphere(0.3)), ((1.0, 0, 0), Q_ID, S.sphere(0.3))])
    b = S.compound([((0.0, 0, 0), Q_ID, S.sphere(0.3)), ((0.0, 2.0, 0), Q_ID, S.sphere(0.3))])
    c, _ = _contacts_after_one_step([(a, (0, 5, 0), Q_ID, 2), (b, (1.55, 5, 0), Q_ID, 2)], floor=False)
    assert len(c) == 1
    assert c["sub_shape_a"][0] == 0xFFFFFFFF & ~1 | 1 and c["sub_shape_b"][0] == 0xFFFFFFFE
    # the depth is the two radii less the centres' distance
    d = float(np.dot(c["position_a"][0][0] - c["position_b"][0][0], c["normal"][0]))
    assert d == pytest.approx(0.6 - 0.55, abs=1e-12)


def test_two_cubes_have_the_mass_of_their_box():
    """Two cubes side by side (compound.ts:169-227): the mass, the centre of mass and the inertia
    of the 1 x 0.5 x 0.5 box (half extents) they make."""
    cube = S.box((0.25, 0.25, 0.25))
    c = S.compound([((-0.25, 0, 0), Q_ID, cube), ((0.25, 0, 0), Q_ID, cube)])
    b = S.box((0.5, 0.25, 0.25))
    assert c.mass == pytest.approx(b.mass, rel=1e-15)
    assert np.allclose(c.center_of_mass, (0, 0, 0), atol=1e-16)
    assert np.allclose(c.inertia, b.inertia, rtol=1e-13, atol=1e-13)
    assert np.allclose(c.aabb_min, (-0.5, -0.25, -0.25)) and np.allclose(c.aabb_max, (0.5, 0.25, 0.25))
    # off centre: the centre of mass is the mass-weighted mean of the children's
    s1, s2 = S.sphere(0.2), S.sphere(0.4)
    c2 = S.compound([((1.0, 0, 0), Q_ID, s1), ((0, 2.0, 0), Q_ID, s2)])
    want = (np.array([1.0, 0, 0]) * s1.mass + np.array([0, 2.0, 0]) * s2.mass) / (s1.mass + s2.mass)
    assert np.allclose(c2.center_of_mass, want, atol=1e-15)


def test_compound_pile_steps_in_the_oracle():
    sc = SC.compound_pile(n=90, extent=(6.0, 4.0, 6.0))
    o = R.OracleWorld(sc)
    for _ in range(150):
        o.step(DT)
    b = o.bodies()
    assert np.all(np.isfinite(b["position"]))
    com = b["center_of_mass_position"][1:, 1]
    assert com.min() > -0.5 and com.max() < 6.0
    c = o.contacts()
    subs = c["sub_shape_a"][c["sub_shape_a"] != 0xFFFFFFFF].tolist() + c["sub_shape_b"][c["sub_shape_b"] != 0xFFFFFFFF].tolist()
    assert len(subs) > 20  # compound children are in contact
    v0 = np.linalg.norm(sc.bodies()["linear_velocity"][1:], axis=1)
    v1 = np.linalg.norm(b["linear_velocity"][1:], axis=1)
    assert np.median(v1) < 0.5 * np.median(v0)
