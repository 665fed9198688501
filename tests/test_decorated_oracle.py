"""Decorated shapes (src/shapes/scaled.ts, transformed.ts, offset-center-of-mass.ts) in the oracle,
pinned by what the geometry fixes rather than by the oracle's own formulas:

* a shape scaled by powers of two collides exactly like the shape built at the scaled size (the
  scaled support functions and transformFace must then give the same bits), a mirrored box like
  the box (which needs the inside-out face reversal), a scaled sphere like a sphere of |scale.x|
  times the radius whatever the other components;
* a transformed shape collides like its leaf placed at the composed pose, computed here with
  scipy's rotations;
* an offset centre of mass changes no contact, only the centre of mass and the inertia
  (parallel-axis theorem);
* the scaled hull's shrunk points lie the scaled convex radius inside each of their scaled
  supporting planes;
* the mass properties match closed forms of the scaled or rotated solids.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from crashcat_amd import scene as SC
from crashcat_amd import shapes as S
from oracle import refcpu as R

Q_ID = (0.0, 0.0, 0.0, 1.0)


def _hit(a, pa, qa, b, pb, qb, max_sep=0.02):
    sc = SC.Scene()
    ia = sc.add_shape(a)
    ib = sc.add_shape(b)
    return R.collide(sc, ia, pa, qa, ib, pb, qb, max_separation=max_sep)


def _same_hit(h1, h2, exact=True, tol=1e-12):
    assert (h1 is None) == (h2 is None)
    if h1 is None:
        return
    for k in ("point_a", "point_b", "axis"):
        if exact:
            assert np.array_equal(h1[k], h2[k]), k
        else:
            assert np.allclose(h1[k], h2[k], rtol=0, atol=tol), (k, h1[k], h2[k])
    if exact:
        assert h1["penetration"] == h2["penetration"]
    else:
        assert h1["penetration"] == pytest.approx(h2["penetration"], abs=tol)


def _manifold(h):
    nrm, pa, pb = R.manifold_from_hit(h, 0.021)
    return nrm, pa, pb


def _same_point_set(a, b, tol):
    assert len(a) == len(b)
    used = set()
    for p in a:
        d = np.linalg.norm(b - p, axis=1)
        j = int(np.argmin(d))
        assert d[j] <= tol and j not in used, (p, b)
        used.add(j)


POSES = [
    ((0.3, 0.9, -0.2), Rotation.from_euler("xyz", [0.3, -0.5, 0.8]).as_quat()),
    ((0.05, 1.1, 0.1), Rotation.from_euler("xyz", [0.01, 0.02, 0.0]).as_quat()),
    ((-0.6, 0.7, 0.4), Rotation.from_euler("xyz", [1.2, 0.4, -0.3]).as_quat()),
]


@pytest.mark.parametrize("pb, qb", POSES)
def test_power_of_two_scale_equals_the_scaled_build(pb, qb):
    """scaled(box(he), 2) is box(2 he) (the convex radius min(0.05 * 2, 0.05) is the box's), and
    scaled(box(he), (2, 0.5, 1)) is box(he * s) with the scaled radius 0.025, bit for bit, hits and
    faces; the same for capsules, cylinders and spheres under a uniform power-of-two scale."""
    floor = S.box((2.0, 0.5, 2.0))
    he = (0.4, 0.3, 0.5)
    cases = [
        (S.scaled(S.box(he), (2.0, 2.0, 2.0)), S.box((0.8, 0.6, 1.0))),
        (S.scaled(S.box(he), (2.0, 0.5, 1.0)), S.box((0.8, 0.15, 0.5), convex_radius=0.025)),
        (S.scaled(S.capsule(0.2, 0.3), (2.0, 2.0, 2.0)), S.capsule(0.4, 0.6)),
        (S.scaled(S.cylinder(0.3, 0.25), (0.5, 0.5, 0.5)), S.cylinder(0.15, 0.125, convex_radius=0.025)),
        (S.scaled(S.sphere(0.3), (2.0, 2.0, 2.0)), S.sphere(0.6)),
    ]
    for deco, built in cases:
        h1 = _hit(floor, (0, 0, 0), Q_ID, deco, pb, qb)
        h2 = _hit(floor, (0, 0, 0), Q_ID, built, pb, qb)
        _same_hit(h1, h2)
        if h1 is not None:
            for x, y in zip(_manifold(h1), _manifold(h2)):
                assert np.array_equal(x, y)


def test_sphere_scale_is_uniform_from_the_x_component():
    """sphere.ts:195-241 takes |scale[0]| whatever the other components (uniform scale only)."""
    floor = S.box((2.0, 0.5, 2.0))
    h1 = _hit(floor, (0, 0, 0), Q_ID, S.scaled(S.sphere(0.25), (-1.5, 7.0, 0.2)), (0.1, 0.8, 0.0), Q_ID)
    h2 = _hit(floor, (0, 0, 0), Q_ID, S.sphere(0.375), (0.1, 0.8, 0.0), Q_ID)
    _same_hit(h1, h2)
    assert h1["penetration"] == pytest.approx(0.375 - 0.3, abs=1e-9)
    sc = SC.Scene()
    a = sc.add_shape(S.scaled(S.sphere(0.3), (2.0, 1.0, 1.0)))
    b = sc.add_shape(S.sphere(0.5))
    h = R.collide(sc, a, (0, 0, 0), Q_ID, b, (1.0, 0, 0), Q_ID)
    assert h["penetration"] == pytest.approx(0.6 + 0.5 - 1.0, abs=1e-12)  # the closed form, radius 0.6


FLAT_POSES = [
    ((0.05, 0.75, 0.1), Rotation.from_euler("xyz", [0.01, 0.02, 0.0]).as_quat()),
    ((-0.3, 0.7, 0.2), Rotation.from_euler("xyz", [0.15, 0.9, -0.1]).as_quat()),
    ((0.2, 0.78, -0.1), Rotation.from_euler("xyz", [-0.2, -0.4, 0.12]).as_quat()),
]


@pytest.mark.parametrize("mirror", [(-1, 1, 1), (1, 1, -1), (-1, 1, -1)])
def test_mirrored_box_collides_like_the_box(mirror):
    """A box mirrored along x or z is the same solid; against a floor its supporting face is chosen
    along y (the box rests nearly flat), comes in the reverse order when the scale is inside out
    (box.ts:136-330, isScaleInsideOut face.ts:87-89), and the manifold is the same set of points.
    Without the reversal the clipping against the inverted face would lose the points.  (When the
    face's own axis is mirrored, transformFace's signed scale puts the face chosen in the unscaled
    frame on the far side of the box; the oracle keeps that behaviour of the reference.)"""
    floor = S.box((2.0, 0.5, 2.0))
    found = 0
    for pb, qb in FLAT_POSES:
        h1 = _hit(S.scaled(S.box((0.4, 0.3, 0.5)), mirror), pb, qb, floor, (0, 0, 0), Q_ID)
        h2 = _hit(S.box((0.4, 0.3, 0.5)), pb, qb, floor, (0, 0, 0), Q_ID)
        _same_hit(h1, h2, exact=False, tol=1e-12)
        if h1 is None:
            continue
        n1, a1, b1 = _manifold(h1)
        n2, a2, b2 = _manifold(h2)
        assert np.allclose(n1, n2, atol=1e-12)
        _same_point_set(a1, a2, 1e-9)
        _same_point_set(b1, b2, 1e-9)
        found += len(a1) > 1
    assert found >= 2
    # the face itself: the box's face mirrored, its vertex order reversed when inside out
    sc = SC.Scene()
    i1 = sc.add_shape(S.scaled(S.box((0.4, 0.3, 0.5)), mirror))
    i2 = sc.add_shape(S.box((0.4, 0.3, 0.5)))
    d = np.array([0.2, -1.0, 0.1])
    f1 = R.supporting_face(sc, i1, d)
    f2 = R.supporting_face(sc, i2, d)
    sign = np.array(mirror, dtype=np.float64)
    if np.prod(sign) < 0:
        assert np.array_equal(f1, (f2 * sign)[::-1])
    else:
        assert np.array_equal(f1, f2 * sign)
    # the mirrored face is the same quad as the box's
    _same_point_set(f1, f2, 1e-15)


@pytest.mark.parametrize("leaf", ["box", "capsule", "cylinder", "hull", "sphere"])
def test_transformed_shape_collides_like_its_leaf_at_the_composed_pose(leaf):
    """collideTransformedVsShape (transformed.ts:370-420): the leaf at p + R(q) t with rotation
    q * tq; the composed pose is computed here with scipy's rotations."""
    rng = np.random.default_rng(7)
    pts = rng.normal(size=(16, 3))
    pts = 0.4 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
    inner = {"box": S.box((0.4, 0.3, 0.2)), "capsule": S.capsule(0.2, 0.3), "cylinder": S.cylinder(0.3, 0.25),
             "hull": S.convex_hull(pts), "sphere": S.sphere(0.3)}[leaf]
    t = np.array([0.25, -0.1, 0.15])
    tq = Rotation.from_euler("xyz", [0.7, -0.2, 1.1])
    deco = S.transformed(inner, t, tq.as_quat())
    other = S.box((0.6, 0.3, 0.5))
    for pb, qb in POSES:
        p = np.array([0.0, 0.45, 0.05])
        r = Rotation.from_euler("xyz", [0.2, 0.9, -0.4])
        h1 = _hit(deco, p, r.as_quat(), other, pb, qb)
        pc = p + r.apply(t)
        qc = (r * tq).as_quat()
        h2 = _hit(inner, pc, qc, other, pb, qb)
        _same_hit(h1, h2, exact=False, tol=1e-12)
        if h1 is not None:
            n1, a1, _ = _manifold(h1)
            n2, a2, _ = _manifold(h2)
            assert np.allclose(n1, n2, atol=1e-12)
            _same_point_set(a1, a2, 1e-9)


def test_offset_center_of_mass_moves_only_the_mass():
    """offset-center-of-mass.ts: collision is the inner shape's (bit for bit), the centre of mass
    moves by the offset, and the inertia about the new centre obeys the parallel-axis theorem
    (a sphere: 2/5 m r^2 + m |o|^2 about axes at right angles to o)."""
    r, o = 0.3, np.array([0.0, -0.12, 0.0])
    sp = S.sphere(r)
    oc = S.offset_center_of_mass(sp, o)
    assert oc.center_of_mass == pytest.approx(tuple(o))
    m = sp.mass
    i0 = 0.4 * m * r * r
    assert oc.inertia[0, 0] == pytest.approx(i0 + m * 0.12 ** 2, rel=1e-14)
    assert oc.inertia[2, 2] == pytest.approx(i0 + m * 0.12 ** 2, rel=1e-14)
    assert oc.inertia[1, 1] == pytest.approx(i0, rel=1e-14)
    floor = S.box((2.0, 0.5, 2.0))
    for inner in (S.sphere(0.3), S.box((0.3, 0.2, 0.4))):
        for pb, qb in POSES:
            h1 = _hit(floor, (0, 0, 0), Q_ID, S.offset_center_of_mass(inner, o), pb, qb)
            h2 = _hit(floor, (0, 0, 0), Q_ID, inner, pb, qb)
            _same_hit(h1, h2)


def test_scaled_mass_properties_match_the_scaled_solid():
    """massProperties.scale (src/body/mass-properties.ts:59-114): a box scaled by (2, 3, 0.5) has
    the mass and inertia of the box of those half extents; a mirror changes neither.  A rotated
    box's inertia is R I R^T (massProperties.rotate :121-161), computed here with scipy."""
    he = np.array([0.4, 0.3, 0.2])
    s = np.array([2.0, 3.0, 0.5])
    a = S.scaled(S.box(tuple(he)), tuple(s))
    b = S.box(tuple(he * s))
    assert a.mass == pytest.approx(b.mass, rel=1e-14)
    assert np.allclose(a.inertia, b.inertia, rtol=1e-13, atol=0)
    m = S.scaled(S.box(tuple(he)), (-1.0, 1.0, 1.0))
    assert m.mass == pytest.approx(S.box(tuple(he)).mass, rel=1e-15)
    assert np.allclose(m.inertia, S.box(tuple(he)).inertia, rtol=1e-14)
    assert a.center_of_mass == (0.0, 0.0, 0.0)  # computeScaledCenterOfMass scales the shape's own zero
    rot = Rotation.from_euler("xyz", [0.3, 1.0, -0.6])
    t = S.transformed(S.box(tuple(he)), (0.1, 0.2, 0.3), rot.as_quat())
    R3 = rot.as_matrix()
    assert np.allclose(t.inertia, R3 @ S.box(tuple(he)).inertia @ R3.T, rtol=0, atol=1e-12)
    assert np.allclose(t.center_of_mass, (0.1, 0.2, 0.3), atol=1e-15)
    # transformed bounds hold the rotated box's corners, and are tight on them
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]) * he
    w = rot.apply(corners) + (0.1, 0.2, 0.3)
    assert np.allclose(t.aabb_min, w.min(axis=0), atol=1e-12) and np.allclose(t.aabb_max, w.max(axis=0), atol=1e-12)


@pytest.mark.parametrize("scale", [(1.5, 1.5, 1.5), (2.0, 0.7, 1.2), (-1.0, 0.8, 1.4)])
def test_scaled_hull_shrunk_points_sit_inside_their_scaled_planes(scale):
    """setConvexHullNoConvexSupportScaled (convex-hull.ts:842-935): each shrunk point lies the
    scaled convex radius (radius times the smallest |scale|) inside each supporting plane of its
    point, the planes being the hull's face planes taken through the scaled point with normals
    n / scale normalised; a point with two planes also lies on the third plane through the
    scaled point at right angles to both.  Under a uniform scale the shrunk points are the
    unscaled ones times the scale."""
    rng = np.random.default_rng(11)
    pts = rng.normal(size=(16, 3))
    pts = 0.5 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
    h = S.convex_hull(pts)
    sc = np.array(scale)
    d = S.scaled(h, scale)
    r = h.convex_radius * np.min(np.abs(sc))
    assert r > 0
    sp = np.asarray(h.points) * sc
    for i, pf in enumerate(h.point_faces):
        q = d.shrunk_points[i]
        ns = [np.asarray(h.plane_normals[f]) / sc for f in pf]
        ns = [n / np.linalg.norm(n) for n in ns]
        if len(pf) == 1:
            assert np.allclose(q, sp[i] - ns[0] * r, atol=1e-12)
            continue
        for n in ns:
            assert np.dot(n, q - sp[i]) == pytest.approx(-r, abs=1e-12)
        if len(pf) == 2:
            perp = np.cross(ns[0], ns[1])
            perp /= np.linalg.norm(perp)
            assert np.dot(perp, q - sp[i]) == pytest.approx(0.0, abs=1e-12)
    if sc[0] == sc[1] == sc[2]:
        assert np.allclose(d.shrunk_points, np.asarray(h.shrunk_points) * sc[0], atol=1e-12)


def test_scaled_hull_collides_like_the_hull_built_scaled():
    """A hull of zero convex radius scaled by 2 collides like the hull built from the doubled
    points (a hull built independently: Qhull on the scaled points, its own centre of mass)."""
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(16, 3))
    pts = 0.4 * pts / np.linalg.norm(pts, axis=1, keepdims=True)
    h = S.convex_hull(pts, convex_radius=0.0)
    h2 = S.convex_hull(pts * 2.0, convex_radius=0.0)
    com1 = np.asarray(h.center_of_mass) * 2.0
    com2 = np.asarray(h2.center_of_mass)
    floor = S.box((2.0, 0.5, 2.0))
    for pb, qb in POSES:
        # the hull's geometry is relative to its centre of mass: place both so the solids coincide
        rot = Rotation.from_quat(qb)
        p1 = np.asarray(pb)
        p2 = p1 + rot.apply(com2 - com1)
        h1 = _hit(floor, (0, 0, 0), Q_ID, S.scaled(h, (2.0, 2.0, 2.0)), p1, qb)
        hb = _hit(floor, (0, 0, 0), Q_ID, h2, p2, qb)
        _same_hit(h1, hb, exact=False, tol=1e-9)


def test_decorated_pile_steps_in_the_oracle():
    """The generator's pile of every decorator runs in the oracle and comes to rest on its floor:
    every body stays in the pile, no contact is deeper than 0.6 m (a dense pile presses some bodies
    deep), and the bodies have slowed."""
    sc = SC.decorated_pile(n=120, extent=(6.0, 4.0, 6.0))
    o = R.OracleWorld(sc)
    for _ in range(150):
        o.step(1.0 / 60.0)
    b = o.bodies()
    assert np.all(np.isfinite(b["position"]))
    com = b["center_of_mass_position"][1:, 1]
    assert com.min() > -0.5 and com.max() < 6.0
    c = o.contacts()
    depth = [float(np.dot(c["position_a"][k][j] - c["position_b"][k][j], c["normal"][k]))
             for k in range(len(c)) for j in range(c["num_points"][k])]
    assert len(c) > 100 and max(depth) < 0.6, (len(c), max(depth))
    # kinetic energy has fallen well below the throw's
    v0 = np.linalg.norm(sc.bodies()["linear_velocity"][1:], axis=1)
    v1 = np.linalg.norm(b["linear_velocity"][1:], axis=1)
    assert np.median(v1) < 0.5 * np.median(v0)
    assert math.isfinite(float(np.abs(b["linear_velocity"]).max()))
