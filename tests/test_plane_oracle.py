"""The plane shape (src/shapes/plane.ts) in the oracle: collideConvexVsPlane (:474-590) with the
adaptive plane face (:631-725), pinned by closed forms -- a box's and a sphere's depth below a
plane (the DEFAULT support includes a box's convex radius and adds a sphere's back), the
reversed hit, a rotated and translated plane body against n_w . p + c_w, the bounded form's
local bounds (computePlaneLocalBounds :273-295) -- and by a step: a box resting on a plane
moves as it does on a box floor, with the same four corners.
"""
import math

import numpy as np
import pytest

from crashcat_amd import scene as SC
from crashcat_amd import shapes as S
from crashcat_amd._abi import CC_MOTION_STATIC
from oracle import refcpu as R


def test_plane_bounds_and_normalisation():
    p = S.plane((0.0, 2.0, 0.0), constant=-0.5, half_extent=100.0)
    assert p.half_extents == (0.0, 1.0, 0.0) and p.radius == -0.5
    # the quad at y = 0.5 (n . p + c = 0) and the same quad 100 m behind it
    assert np.allclose(p.aabb_min, (-100, 0.5 - 100, -100)) and np.allclose(p.aabb_max, (100, 0.5, 100))


def _scene():
    sc = SC.Scene()
    pi = sc.add_shape(S.plane((0.0, 1.0, 0.0)))
    return sc, pi


def test_box_and_sphere_depths():
    sc, pi = _scene()
    bi = sc.add_shape(S.box((0.5, 0.3, 0.4)))
    h = R.collide(sc, bi, (0.2, 0.27, -0.1), (0, 0, 0, 1), pi, (0, 0, 0), (0, 0, 0, 1))
    assert h["penetration"] == pytest.approx(0.03, abs=1e-15)  # hy - y; the radius is in the box
    assert np.allclose(h["axis"], (0, -1, 0), atol=1e-15)
    assert h["point_b"][1] == pytest.approx(0.0, abs=1e-15)  # on the plane
    assert h["point_a"][1] == pytest.approx(-0.03, abs=1e-15)  # the box's lowest corner
    assert len(h["face_a"]) == 4 and len(h["face_b"]) == 4  # the bottom face and the adaptive quad
    si = sc.add_shape(S.sphere(0.4))
    h = R.collide(sc, si, (0.0, 0.35, 0.0), (0, 0, 0, 1), pi, (0, 0, 0), (0, 0, 0, 1))
    assert h["penetration"] == pytest.approx(0.05, abs=1e-15)  # r - y: centre support plus the radius
    assert np.allclose(h["point_a"], (0, -0.05, 0), atol=1e-15) and np.allclose(h["point_b"], (0, 0, 0), atol=1e-15)
    assert len(h["face_a"]) == 0  # a sphere has no face: a single point
    assert R.collide(sc, si, (0.0, 0.43, 0.0), (0, 0, 0, 1), pi, (0, 0, 0), (0, 0, 0, 1)) is None
    assert R.collide(sc, si, (0.0, 0.41, 0.0), (0, 0, 0, 1), pi, (0, 0, 0), (0, 0, 0, 1), max_separation=0.02)
    # reversed: the plane as shape A
    r = R.collide(sc, pi, (0, 0, 0), (0, 0, 0, 1), si, (0.0, 0.35, 0.0), (0, 0, 0, 1))
    assert r["penetration"] == pytest.approx(0.05, abs=1e-15)
    assert np.allclose(r["axis"], (0, 1, 0)) and np.allclose(r["point_a"], (0, 0, 0), atol=1e-15)
    assert R.collide(sc, pi, (0, 0, 0), (0, 0, 0, 1), pi, (0, 1, 0), (0, 0, 0, 1)) is None


def test_rotated_translated_plane():
    sc, pi = _scene()
    si = sc.add_shape(S.sphere(0.5))
    a = math.radians(25)
    q = (0.0, 0.0, math.sin(a / 2), math.cos(a / 2))  # about z: the normal becomes (-sin a, cos a, 0)
    t = np.array([0.3, -0.2, 0.7])
    nw = np.array([-math.sin(a), math.cos(a), 0.0])
    for p in [(0.1, 0.3, 0.2), (1.0, 0.8, -0.4), (-0.6, 0.2, 0.9)]:
        d = nw @ (np.asarray(p) - t)  # n_w . p + c_w with c_w = -n_w . t
        h = R.collide(sc, si, p, (0, 0, 0, 1), pi, tuple(t), q, max_separation=0.1)
        if 0.5 - d > -0.1:
            assert h["penetration"] == pytest.approx(0.5 - d, abs=1e-12)
            assert np.allclose(h["axis"], -nw, atol=1e-12)
        else:
            assert h is None


def _box_world(plane_floor, y0):
    sc = SC.Scene()
    if plane_floor:
        sc.add_body(S.plane((0.0, 1.0, 0.0)), (0, 0, 0), motion_type=CC_MOTION_STATIC, friction=0.5)
    else:
        sc.add_body(S.box((50.0, 0.5, 50.0)), (0, -0.5, 0), motion_type=CC_MOTION_STATIC, friction=0.5)
    sc.add_body(S.box((0.5, 0.5, 0.5)), (0.23, y0, -0.31), friction=0.5)
    sc.end_world()
    return sc


def test_box_on_plane_matches_box_on_floor():
    y0 = 0.5 - 0.004
    op, of = R.OracleWorld(_box_world(True, y0)), R.OracleWorld(_box_world(False, y0))
    op.step(1 / 60)
    of.step(1 / 60)
    cp, cf = op.contacts(), of.contacts()
    assert len(cp) == 1 and len(cf) == 1 and cp[0]["num_points"] == 4 == cf[0]["num_points"]
    assert np.allclose(cp[0]["normal"], (0, 1, 0), atol=1e-15)  # the plane's normal, exactly
    assert np.allclose(cp[0]["normal"], cf[0]["normal"], atol=1e-9)  # the floor's comes from EPA, within its tolerance
    pp = np.sort(np.asarray(cp[0]["position_b"])[:4], axis=0)
    pf = np.sort(np.asarray(cf[0]["position_b"])[:4], axis=0)
    assert np.allclose(pp, pf, atol=1e-9)
    bp, bf = op.bodies(), of.bodies()
    assert np.allclose(bp["position"][1], bf["position"][1], atol=1e-9)
    vp, vf = bp["linear_velocity"][1], bf["linear_velocity"][1]
    assert vp[1] == pytest.approx(vf[1], abs=1e-9)
    assert math.hypot(vp[0], vp[2]) == pytest.approx(math.hypot(vf[0], vf[2]), abs=1e-9)
