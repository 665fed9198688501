"""The cylinder shape (src/shapes/cylinder.ts) in the oracle, pinned by the reference's own
collide cases (tst/collision/collide-shape-vs-shape.test.ts:2531-2666) made exact with their
closed-form depths, and by what geometry fixes: the mass properties of a solid cylinder, the
supporting faces (a vertical side edge at the radius; the cap octagon on the circle, turned
toward the direction's horizontal part), a standing cylinder's contact on a floor, and a
cylinder coming to rest on its cap.
"""
import math

import numpy as np
import pytest

from crashcat_amd import scene as SC
from crashcat_amd import shapes as S
from crashcat_amd._abi import CC_MOTION_STATIC
from oracle import refcpu as R


def test_mass_properties_of_a_solid_cylinder():
    r, h, rho = 0.4, 0.7, 800.0
    c = S.cylinder(r, h, density=rho)
    m = rho * math.pi * r * r * 2 * h
    assert c.mass == pytest.approx(m, rel=1e-15)
    # I_yy = m r^2 / 2, I_xx = I_zz = m (3 r^2 + (2h)^2) / 12
    assert c.inertia[1, 1] == pytest.approx(m * r * r / 2, rel=1e-14)
    assert c.inertia[0, 0] == pytest.approx(m * (3 * r * r + 4 * h * h) / 12, rel=1e-14)
    assert c.inertia[2, 2] == c.inertia[0, 0]
    assert np.count_nonzero(c.inertia - np.diag(np.diag(c.inertia))) == 0
    with pytest.raises(ValueError):
        S.cylinder(0.04, 1.0)  # radius below the convex radius (cylinder.ts:94-103)
    with pytest.raises(ValueError):
        S.cylinder(1.0, 0.04)


@pytest.mark.parametrize("a, pb, depth", [
    (lambda: S.sphere(1.0), (1.5, 0, 0), 0.5),
    (lambda: S.box((1, 1, 1)), (1.5, 0, 0), 0.5),
    (lambda: S.cylinder(0.5, 1.0), (0.8, 0, 0), 0.2),
])
def test_reference_cylinder_cases_with_exact_depths(a, pb, depth):
    sc = SC.Scene()
    sa = sc.add_shape(a())
    sb = sc.add_shape(S.cylinder(1.0, 2.0) if depth == 0.5 else S.cylinder(0.5, 1.0))
    hit = R.collide(sc, sa, (0, 0, 0), (0, 0, 0, 1), sb, pb, (0, 0, 0, 1))
    assert hit is not None and hit["penetration"] > 0  # the reference's assertion
    assert hit["penetration"] == pytest.approx(depth, abs=2e-4)  # the side of the cylinder is exact
    ax = hit["axis"] / np.linalg.norm(hit["axis"])
    assert np.allclose(ax, (1, 0, 0), atol=1e-3)
    # no collision between distant cylinders
    assert R.collide(sc, sb, (0, 0, 0), (0, 0, 0, 1), sb, (5, 0, 0), (0, 0, 0, 1)) is None


def test_supporting_faces():
    sc = SC.Scene()
    r, h = 0.6, 0.9
    ci = sc.add_shape(S.cylinder(r, h))
    # side: the vertical edge at the radius, on the side facing away from the direction
    d = np.array([0.6, 0.1, 0.8])
    f = R.supporting_face(sc, ci, d)
    assert len(f) == 2
    hd = np.array([d[0], 0, d[2]]) / math.hypot(d[0], d[2])
    assert np.allclose(f[0], -r * hd + (0, h, 0)) and np.allclose(f[1], -r * hd - (0, h, 0))
    # cap, straight down: the octagon on the top cap's circle, in the template's order
    f = R.supporting_face(sc, ci, (0, -1, 0))
    assert len(f) == 8
    assert np.allclose(f[:, 1], h)
    assert np.allclose(np.hypot(f[:, 0], f[:, 2]), r, rtol=1e-8)
    assert np.allclose(f[0], (0, h, r)) and np.allclose(f[2], (r, h, 0))
    # cap, tilted 30 degrees from the axis: turned so that a vertex points along the tilt
    t = math.radians(30)
    d = np.array([math.sin(t) * math.cos(0.4), -math.cos(t), math.sin(t) * math.sin(0.4)])
    f = R.supporting_face(sc, ci, d)
    hd = np.array([math.cos(0.4), 0, math.sin(0.4)])
    assert len(f) == 8 and np.allclose(f[:, 1], h)
    assert min(np.linalg.norm(f - (r * hd + (0, h, 0)), axis=1)) < 1e-8
    # within 5 degrees of the axis: not turned
    d = np.array([0.05, 1.0, 0.0])
    f = R.supporting_face(sc, ci, d)
    assert np.allclose(f[:, 1], -h) and np.allclose(f[0], (0, -h, r))


def _standing(y):
    sc = SC.Scene()
    sc.add_body(S.box((5.0, 0.5, 5.0)), (0, -0.5, 0), motion_type=CC_MOTION_STATIC, friction=0.5)
    sc.add_body(S.cylinder(0.4, 0.5), (0.1, y, -0.2), friction=0.5)
    sc.end_world()
    return sc


def test_standing_cylinder_contact_and_rest():
    o = R.OracleWorld(_standing(0.5 - 0.01))
    o.step(1 / 60)
    c = o.contacts()
    assert len(c) == 1 and c[0]["num_points"] == 4  # the cap octagon clipped and pruned to 4
    assert np.allclose(np.abs(c[0]["normal"]), (0, 1, 0), atol=1e-9)
    pts = np.asarray(c[0]["position_b"])[:4]
    assert np.allclose(np.hypot(pts[:, 0] - 0.1, pts[:, 2] + 0.2), 0.4, atol=1e-6)  # on the cap's rim
    for _ in range(180):
        o.step(1 / 60)
    b = o.bodies()
    assert np.linalg.norm(b["linear_velocity"][1]) < 1e-3
    assert 0.5 - 0.02 <= b["position"][1][1] <= 0.5 + 1e-3  # rests within the penetration slop
    q = b["quaternion"][1]
    assert abs(abs(q[3]) - 1) < 1e-6  # still upright
