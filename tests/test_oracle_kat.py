"""Known-answer tests of the FP64 oracle, restated from the reference's own unit tests.

Each test names the reference test it restates (tst/..., file:line).  The inputs and
expected values are copied from those tests; where the reference only checks a count or a
range, the exact value that the geometry fixes is checked as well (tighter than the
reference), so that a dropped term or a swapped index fails here.
"""
import math

import numpy as np
import pytest

from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH
from oracle import refcpu as O


def _rows(vals):
    return np.asarray(vals, dtype=np.float64).reshape(-1, 3)


def _same_polygon(a, b, tol=1e-12):
    """Same vertex cycle, up to the starting vertex."""
    a, b = _rows(a), _rows(b)
    if len(a) != len(b):
        return False
    for s in range(len(a)):
        if np.allclose(np.roll(a, -s, axis=0), b, atol=tol):
            return True
    return False


# ----------------------------------------------------------------- clipping (tst/manifold/clip.test.ts)

def test_clip_plane_trapezoid():
    """clip.test.ts:21-43 — triangle cut at y=1 gives (0.5,1),(0,0),(2,0),(1.5,1) in that order."""
    out = O.clip_poly_vs_plane([0, 0, 0, 2, 0, 0, 1, 2, 0], [0, 1, 0], [0, -1, 0])
    assert len(out) == 4
    np.testing.assert_allclose(out, _rows([0.5, 1, 0, 0, 0, 0, 2, 0, 0, 1.5, 1, 0]), atol=1e-15)


def test_clip_plane_inside_outside_parallel_unnormalized():
    """clip.test.ts:45-111."""
    tri = [0, 0, 0, 1, 0, 0, 0.5, 1, 0]
    np.testing.assert_allclose(O.clip_poly_vs_plane(tri, [0, 2, 0], [0, -1, 0]), _rows(tri))
    assert len(O.clip_poly_vs_plane(tri, [0, -1, 0], [0, -1, 0])) == 0
    sq = [0, -0.1, 0, 1, -0.1, 0, 1, 1, 0, 0, 1, 0]
    out = O.clip_poly_vs_plane(sq, [0, 0, 0], [0, -1, 0])
    assert len(out) == 4
    # the kept part is the strip y in [-0.1, 0]
    assert np.all(out[:, 1] <= 1e-15) and np.isclose(out[:, 1].min(), -0.1)
    out = O.clip_poly_vs_plane([0, 0, 0, 2, 0, 0, 1, 2, 0], [0, 1, 0], [0, -5, 0])
    np.testing.assert_allclose(out, _rows([0.5, 1, 0, 0, 0, 0, 2, 0, 0, 1.5, 1, 0]), atol=1e-15)


def test_clip_poly_octagon():
    """clip.test.ts:114-132 — unit square vs 45-degree square of circumradius sqrt(2)/2: an
    octagon.  Its vertices are where the diamond |x|+|y|=sqrt(2)/2 crosses the square's edges."""
    s = math.sqrt(2) / 2
    out = O.clip_poly_vs_poly([-0.5, -0.5, 0, 0.5, -0.5, 0, 0.5, 0.5, 0, -0.5, 0.5, 0],
                              [-s, 0, 0, 0, -s, 0, s, 0, 0, 0, s, 0], [0, 0, 1])
    assert len(out) == 8
    np.testing.assert_allclose(np.abs(out[:, 0]) + np.abs(out[:, 1]), s, atol=1e-12)
    assert np.all(np.isclose(np.abs(out[:, 0]), 0.5) | np.isclose(np.abs(out[:, 1]), 0.5))
    # area of the octagon = square minus four corner triangles of legs (1 - s)
    x, y = out[:, 0], out[:, 1]
    area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
    assert area == pytest.approx(1 - 2 * (1 - s) ** 2, abs=1e-12)


def test_clip_poly_contained_disjoint_partial_degenerate():
    """clip.test.ts:134-205."""
    small = [0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0]
    out = O.clip_poly_vs_poly(small, [-1, -1, 0, 2, -1, 0, 2, 2, 0, -1, 2, 0], [0, 0, 1])
    assert _same_polygon(out, small)
    assert len(O.clip_poly_vs_poly(small, [5, 5, 0, 6, 5, 0, 6, 6, 0, 5, 6, 0], [0, 0, 1])) == 0
    out = O.clip_poly_vs_poly([0, 0, 0, 2, 0, 0, 2, 2, 0, 0, 2, 0], [1, 1, 0, 3, 1, 0, 3, 3, 0, 1, 3, 0], [0, 0, 1])
    assert 3 <= len(out) <= 8
    # the overlap is exactly the square [1,2]^2
    assert np.allclose(out.min(axis=0)[:2], [1, 1]) and np.allclose(out.max(axis=0)[:2], [2, 2])
    assert len(O.clip_poly_vs_poly(small, [0.5, 0, 0, 0.5, 1, 0], [0, 0, 1])) == 0


def test_clip_edge_cases():
    """clip.test.ts:208-319."""
    sq2 = [0, 0, 0, 2, 0, 0, 2, 2, 0, 0, 2, 0]
    out = O.clip_poly_vs_edge(sq2, [1, 0, 5], [1, 2, 5], [0, 0, 1])
    assert len(out) == 2
    out = O.clip_poly_vs_edge(sq2, [0.5, 0, 5], [0.5, 2, 5], [0, 0, 1])
    assert len(out) == 2
    out = O.clip_poly_vs_edge([0, 0, 0, 4, 0, 0, 4, 4, 0, 0, 4, 0], [2, 1.5, 5], [2, 2.5, 5], [0, 0, 1])
    assert len(out) >= 1
    assert np.all(out[:, 1] >= 1.5 - 1e-12) and np.all(out[:, 1] <= 2.5 + 1e-12)
    assert len(O.clip_poly_vs_edge([0, 0, 0, 1, 0, 0], [0.5, -1, 0], [0.5, 1, 0], [-1, 0, 0])) == 0
    out = O.clip_poly_vs_edge([0, 0, 0, 2, 0, 0, 1, 2, 0], [1, 0, 5], [1, 2, 5], [0, 0, 1])
    assert len(out) >= 1
    np.testing.assert_allclose(out[:, 2], 0.0, atol=1e-12)


# ----------------------------------------------------------------- manifold (tst/manifold/manifold.test.ts)

SQ0 = [-0.5, -0.5, 0, 0.5, -0.5, 0, 0.5, 0.5, 0, -0.5, 0.5, 0]
SQ1 = [-0.5, -0.5, 0.1, 0.5, -0.5, 0.1, 0.5, 0.5, 0.1, -0.5, 0.5, 0.1]


@pytest.mark.parametrize("face1,face2,axis,maxd,count,ref", [
    (SQ0, SQ1, [0, 0, 0.1], 0.2, 4, "manifold.test.ts:14-33"),
    (SQ0, [0, -0.5, 0.1, 0.5, 0, 0.1, 0, 0.5, 0.1, -0.5, 0, 0.1], [0, 0, 0.1], 0.2, 8, "manifold.test.ts:35-56"),
    ([0, -0.5, 0, 0, 0.5, 0], SQ1, [0, 0, 0.1], 0.2, 2, "manifold.test.ts:58-77"),
    (SQ0, [10, 10, 0.1, 11, 10, 0.1, 11, 11, 0.1, 10, 11, 0.1], [0, 0, 0.1], 0.2, 1, "manifold.test.ts:79-98"),
    (SQ0, SQ1, [0, 0, 0.1], 0.05, 1, "manifold.test.ts:100-118"),
    ([0, 0, 0], SQ1, [0, 0, 0.1], 0.2, 1, "manifold.test.ts:120-138"),
    (SQ0, [0, 0, 0.1, 1, 0, 0.1], [0, 0, 0.1], 0.2, 1, "manifold.test.ts:140-158"),
    (SQ0, SQ1, [1, 0, 0], 0.2, 1, "manifold.test.ts:199-218"),
])
def test_manifold_between_two_faces_counts(face1, face2, axis, maxd, count, ref):
    a, b = O.manifold_between_two_faces([0, 0, 0], [0, 0, 0.1], axis, maxd, face1, face2)
    assert len(a) == count, ref
    if count == 1:  # the fallback keeps the input closest points
        np.testing.assert_array_equal(a[0], [0, 0, 0])
        np.testing.assert_array_equal(b[0], [0, 0, 0.1])


def test_manifold_aligned_points_are_corners():
    """Aligned faces 0.1 apart: the four corners of face 2 paired with their projections on face 1."""
    a, b = O.manifold_between_two_faces([0, 0, 0], [0, 0, 0.1], [0, 0, 0.1], 0.2, SQ0, SQ1)
    assert _same_polygon(b, SQ1) or _same_polygon(b[::-1], SQ1)
    np.testing.assert_allclose(a[:, :2], b[:, :2], atol=1e-15)
    np.testing.assert_allclose(a[:, 2], 0.0, atol=1e-15)


def _pts(points, depth):
    p1 = np.asarray(points, dtype=np.float64)
    p2 = p1 + np.asarray([[0, 0, d] for d in depth])
    return p1, p2


def test_prune_8_to_4_corners():
    """manifold.test.ts:221-254, 332-366 — a 3x3 grid rim (4 corners + 4 edge midpoints) at
    uniform depth prunes to 4 points; the metric picks the corners, in P1, P3, P2, P4 order
    (P2 the corner opposite P1, P3/P4 the two other corners on either side of P1-P2)."""
    pts = [(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0), (0, -1, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0)]
    a, b = _pts(pts, [0.02] * 8)
    ra, rb = O.prune(a, b, [0, 0, 1])
    assert len(ra) == 4
    corners = {(-1, -1), (1, -1), (1, 1), (-1, 1)}
    assert {(int(round(p[0])), int(round(p[1]))) for p in ra} == corners
    assert np.allclose(ra[0][:2], -ra[2][:2])  # P1 and P2 diagonal
    np.testing.assert_allclose(rb - ra, [[0, 0, 0.02]] * 4, atol=1e-15)


def test_prune_hexagon_and_collinear():
    """manifold.test.ts:289-330 (6 points on a circle) and :368-399 (5 collinear points keep 2,
    the two ends)."""
    ang = np.arange(6) / 6 * 2 * np.pi
    a, b = _pts(np.stack([np.cos(ang) * 0.5, np.sin(ang) * 0.5, np.zeros(6)], 1), [0.01] * 6)
    ra, rb = O.prune(a, b, [0, 0, 1])
    assert len(ra) == 4
    a, b = _pts([(-1, 0, 0), (-0.5, 0, 0), (0, 0, 0), (0.5, 0, 0), (1, 0, 0)], [0.01] * 5)
    ra, rb = O.prune(a, b, [0, 0, 1])
    assert len(ra) == 2
    assert sorted(p[0] for p in ra) == [-1.0, 1.0]


def test_prune_weights_distance_and_depth():
    """manifold.test.ts:401-457 — depth grows with distance; the far points win."""
    pts = [(0.5, 0, 0), (2, 0, 0), (0.5, 0.5, 0), (2, 2, 0), (-0.5, 0, 0), (-2, 0, 0), (-0.5, -0.5, 0), (-2, -2, 0)]
    a, b = _pts(pts, [np.linalg.norm(p) * 0.1 for p in pts])
    ra, _ = O.prune(a, b, [0, 0, 1])
    assert len(ra) == 4
    assert all(np.linalg.norm(p) >= 2 - 1e-12 for p in ra)


# ----------------------------------------------------------------- barycentrics (tst/collision/closest-points.test.ts)

def test_barycentric_kats():
    """closest-points.test.ts:8-144, with the exact values the geometry fixes."""
    u, v, ok = O.barycentric2([1, 0, 0], [0, 1, 0])
    assert ok and u == pytest.approx(0.5, abs=1e-15) and v == pytest.approx(0.5, abs=1e-15)
    u, v, ok = O.barycentric2([0, 0, 0], [1, 0, 0])
    assert ok and u == 1.0 and v == 0.0
    u, v, ok = O.barycentric2([0, 0, 0], [2, 0, 0])
    assert ok and u + v == pytest.approx(1.0)
    assert not O.barycentric2([1, 1, 1], [1, 1, 1])[2]
    u, v, w, ok = O.barycentric3([0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert ok and (u, v, w) == pytest.approx((1, 0, 0), abs=1e-15)
    u, v, w, ok = O.barycentric3([1, 0, 0], [0, 1, 0], [0, 0, 1])
    assert ok and (u, v, w) == pytest.approx((1 / 3, 1 / 3, 1 / 3), abs=1e-15)
    assert not O.barycentric3([0, 0, 0], [1, 0, 0], [2, 0, 0])[3]
    # origin projected on a generic line / triangle: u*a + v*b (+ w*c) is the closest point
    rng = np.random.default_rng(5)
    for _ in range(50):
        a, b, c = rng.normal(size=(3, 3))
        u, v, ok = O.barycentric2(a, b)
        p = u * a + v * b
        assert ok and u + v == pytest.approx(1.0) and abs(np.dot(p, b - a)) < 1e-12
        u, v, w, ok = O.barycentric3(a, b, c)
        p = u * a + v * b + w * c
        n = np.cross(b - a, c - a)
        assert ok and u + v + w == pytest.approx(1.0)
        np.testing.assert_allclose(np.cross(p, n), 0.0, atol=1e-10)  # p is the foot of the normal


# ----------------------------------------------------------------- inverse inertia (tst/body/motion-properties.test.ts)

def _axis_angle(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    s = math.sin(angle / 2)
    return np.array([a[0] * s, a[1] * s, a[2] * s, math.cos(angle / 2)])


def _rotmat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


ID = (0, 0, 0, 1)


def test_inverse_inertia_kats():
    """motion-properties.test.ts:66-237."""
    np.testing.assert_allclose(O.inverse_inertia((2.5, 2.5, 2.5), ID, ID), np.eye(3) * 2.5, atol=1e-15)
    np.testing.assert_allclose(O.inverse_inertia((1, 2, 3), ID, ID), np.diag([1, 2, 3]), atol=1e-15)
    rz = _axis_angle((0, 0, 1), math.pi / 2)
    np.testing.assert_allclose(np.diag(O.inverse_inertia((1, 2, 3), ID, rz)), [2, 1, 3], atol=1e-12)
    np.testing.assert_allclose(np.diag(O.inverse_inertia((1, 2, 3), rz, ID)), [2, 1, 3], atol=1e-12)
    m = O.inverse_inertia((1, 2, 3), ID, ID, dofs=0b110111)  # rotation X locked
    assert np.all(m[0, :] == 0) and np.all(m[:, 0] == 0) and m[1, 1] == 2 and m[2, 2] == 3
    c = s = math.cos(math.pi / 4)
    m = O.inverse_inertia((1, 2, 3), ID, _axis_angle((0, 1, 0), math.pi / 4))
    assert m[0, 0] == pytest.approx(c * c + 3 * s * s, abs=1e-12)
    assert m[1, 1] == pytest.approx(2, abs=1e-12)
    assert m[2, 2] == pytest.approx(s * s + 3 * c * c, abs=1e-12)
    assert m[0, 2] == pytest.approx(2 * c * s, abs=1e-12) and m[2, 0] == pytest.approx(2 * c * s, abs=1e-12)


def test_inverse_inertia_matches_r_d_rt():
    """motion-properties.test.ts:6-64, 190-211: (R_body R_inertia) diag(D) (R_body R_inertia)^T,
    symmetric, for arbitrary rotations."""
    rng = np.random.default_rng(7)
    for _ in range(30):
        d = rng.uniform(0.1, 5, 3)
        qi = _axis_angle(rng.normal(size=3), rng.uniform(0, 6))
        qb = _axis_angle(rng.normal(size=3), rng.uniform(0, 6))
        r = _rotmat(qb) @ _rotmat(qi)
        m = O.inverse_inertia(d, qi, qb)
        np.testing.assert_allclose(m, r @ np.diag(d) @ r.T, atol=1e-12)
        np.testing.assert_allclose(m, m.T, atol=1e-13)
        # DOF masking zeroes the rows/columns of locked rotation axes
        dof = int(rng.integers(0, 8)) << 3 | 0b111
        mm = O.inverse_inertia(d, qi, qb, dofs=dof)
        keep = np.array([(dof >> (3 + k)) & 1 for k in range(3)], dtype=float)
        np.testing.assert_allclose(mm, (r @ np.diag(d) @ r.T) * np.outer(keep, keep), atol=1e-12)


# ----------------------------------------------------------------- shapes (tst/shapes/*.test.ts)

def test_box_mass_and_aabb():
    """box.test.ts:6-47."""
    b = SH.box((1, 2, 3))
    assert b.aabb_min == (-1, -2, -3) and b.aabb_max == (1, 2, 3) and b.center_of_mass == (0, 0, 0)
    assert b.density == 1000
    assert SH.box((1, 1, 1)).mass == pytest.approx(8000, abs=1e-9)
    # solid-box inertia: m (b^2 + c^2) / 12 with full side lengths
    np.testing.assert_allclose(np.diag(b.inertia), b.mass / 12 * np.array([4 * 4 + 6 * 6, 2 * 2 + 6 * 6, 2 * 2 + 4 * 4]))


def test_capsule_properties():
    """capsule.test.ts:37-170."""
    c = SH.capsule(1.0, 1.0)
    assert c.volume == pytest.approx(2 * math.pi + 4 / 3 * math.pi, abs=1e-12)
    c2 = SH.capsule(0.5, 2.0)
    assert c2.volume == pytest.approx(math.pi * 0.25 * 4 + 4 / 3 * math.pi * 0.125, abs=1e-12)
    assert c2.aabb_min == (-0.5, -2.5, -0.5) and c2.aabb_max == (0.5, 2.5, 0.5)
    assert c.mass == pytest.approx(1000 * c.volume)
    ixx, iyy, izz = np.diag(c.inertia)
    assert ixx == izz and iyy < ixx and iyy > 0
    assert np.count_nonzero(c.inertia - np.diag(np.diag(c.inertia))) == 0
    c500 = SH.capsule(1.0, 1.0, density=2000)
    np.testing.assert_allclose(c500.inertia, 2 * c.inertia, rtol=1e-14)
    long_ = SH.capsule(0.5, 10.0)
    assert long_.inertia[1, 1] < 0.1 * long_.inertia[0, 0]
    wide = SH.capsule(2.0, 0.1)
    assert wide.inertia[1, 1] / wide.inertia[0, 0] > 0.5
    with pytest.raises(ValueError):
        SH.capsule(0.0, 1.0)
    with pytest.raises(ValueError):
        SH.capsule(1.0, -1.0)


def test_sphere_mass():
    s = SH.sphere(0.5)
    assert s.mass == pytest.approx(4 / 3 * math.pi * 0.125 * 1000)
    np.testing.assert_allclose(np.diag(s.inertia), 0.4 * s.mass * 0.25)


def test_unit_cube_hull():
    """A hull of the 8 corners of the unit cube is the unit cube: volume 1, mass 1000,
    inertia m/6 on the diagonal, six quad faces, and (with convex radius 0.05) inner
    points at +-0.45."""
    corners = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)])
    h = SH.convex_hull(corners)
    assert h.volume == pytest.approx(1.0, abs=1e-12)
    assert h.mass == pytest.approx(1000.0, abs=1e-9)
    np.testing.assert_allclose(h.inertia, np.eye(3) * 1000 / 6, atol=1e-9)
    assert len(h.faces) == 6 and all(len(f) == 4 for f in h.faces)
    np.testing.assert_allclose(np.abs(h.shrunk_points), 0.45, atol=1e-12)


def _one_shape_scene(*shapes):
    sc = SC.Scene()
    for s in shapes:
        sc.add_shape(s)
    return sc


@pytest.mark.parametrize("he,d,axis,val", [
    ((2, 1, 0.5), (1, 0, 0), 0, -2), ((1.5, 2, 1), (-1, 0, 0), 0, 1.5), ((1, 2, 0.5), (0, 1, 0), 1, -2),
    ((1, 1.5, 1), (0, -1, 0), 1, 1.5), ((1, 1, 3), (0, 0, 1), 2, -3), ((1, 1, 2), (0, 0, -1), 2, 2),
    ((3, 1, 1), (10, 1, 1), 0, -3),
])
def test_box_supporting_face(he, d, axis, val):
    """supporting-face.test.ts:51-210 — the face of a box most opposed to the direction."""
    sc = _one_shape_scene(SH.box(he))
    f = O.supporting_face(sc, 0, d)
    assert len(f) == 4
    np.testing.assert_allclose(f[:, axis], val, atol=1e-15)


def test_sphere_supporting_face_empty():
    """supporting-face.test.ts:5-48."""
    sc = _one_shape_scene(SH.sphere(2.0))
    for d in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (-1, -1, -1)]:
        assert len(O.supporting_face(sc, 0, d)) == 0


# ----------------------------------------------------------------- narrowphase (tst/collision/collide-shape-vs-shape.test.ts)

def test_sphere_sphere_closed_form():
    """collide-shape-vs-shape.test.ts:22-160 and src/shapes/sphere.ts:362-395: penetration
    r1 + r2 - d, axis along the centre line, points on each surface."""
    sc = _one_shape_scene(SH.sphere(1.0), SH.sphere(0.7))
    assert O.collide(sc, 0, (0, 0, 0), ID, 0, (10, 0, 0), ID) is None
    rng = np.random.default_rng(3)
    for _ in range(20):
        pa = rng.normal(size=3)
        dirn = rng.normal(size=3)
        dirn /= np.linalg.norm(dirn)
        dist = rng.uniform(0.2, 1.6)
        pb = pa + dirn * dist
        h = O.collide(sc, 0, pa, ID, 1, pb, ID)
        assert h is not None
        assert h["penetration"] == pytest.approx(1.7 - dist, abs=1e-12)
        np.testing.assert_allclose(h["axis"] / np.linalg.norm(h["axis"]), dirn, atol=1e-12)
        np.testing.assert_allclose(h["point_a"], pa + dirn * 1.0, atol=1e-12)
        np.testing.assert_allclose(h["point_b"], pb - dirn * 0.7, atol=1e-12)


def test_convex_path_coincident_and_deep_spheres():
    """collide-shape-vs-shape.test.ts:166-263: through GJK/EPA, coincident r=2 spheres give a
    depth in (3.5, 4.5); r=2 and r=1.8 spheres 0.1 apart give 3.7 to one decimal."""
    sc = _one_shape_scene(SH.sphere(2.0), SH.sphere(1.8))
    h = O.collide(sc, 0, (1, 2, 3), ID, 0, (1, 2, 3), ID, convex_path=True)
    assert h is not None and 3.5 < h["penetration"] < 4.5
    h = O.collide(sc, 0, (1, 2, 3), ID, 1, (1.1, 2, 3), ID, convex_path=True)
    assert h is not None and round(h["penetration"], 1) == round(2.0 + 1.8 - 0.1, 1)
    assert abs(h["penetration"] - 3.7) < 0.05


def test_box_box_depth_is_sat_minimum():
    """Axis-aligned box-box overlaps (collide-shape-vs-shape.test.ts:267-494 checks only that
    a hit with positive depth exists).  The separating-axis theorem fixes the true depth: the
    smallest per-axis overlap.  EPA's polytope lies inside the Minkowski difference, so its
    depth never exceeds that; when EPA runs to convergence it equals it with the normal on that
    axis.  EPA may also stop early: when the next queued face is farther than the best support
    distance seen (penetration.ts:430) the result is taken from the last expanded face
    (:519-525), whose plane lies below the true depth.  So the depth is bounded above by the
    SAT value and most cases must hit it exactly."""
    rng = np.random.default_rng(11)
    sc = _one_shape_scene(SH.box((0.5, 0.5, 0.5)), SH.box((0.6, 0.3, 0.4)))
    exact = total = 0
    for _ in range(200):
        off = rng.uniform(-0.9, 0.9, 3) * np.array([1.1, 0.8, 0.9])
        ov = np.array([1.1, 0.8, 0.9]) - np.abs(off)
        if np.any(ov <= 0.01) or np.sort(ov)[1] - np.sort(ov)[0] < 0.01:
            continue
        h = O.collide(sc, 0, (0, 0, 0), ID, 1, off, ID)
        assert h is not None
        k = int(np.argmin(ov))
        total += 1
        assert 0 < h["penetration"] <= ov[k] + 1e-9
        n = h["axis"] / np.linalg.norm(h["axis"])
        if abs(h["penetration"] - ov[k]) < 1e-4 * ov[k] + 1e-12 and abs(abs(n[k]) - 1) < 1e-6:
            exact += 1
    assert total > 50 and exact >= 0.6 * total


def test_separated_convex_distance_within_speculative_margin():
    """Separated box and capsule within the speculative distance: a hit with negative
    penetration equal to minus the gap (analytic distance, tolerance 1e-4)."""
    sc = _one_shape_scene(SH.box((0.5, 0.5, 0.5)), SH.capsule(0.3, 0.5))
    gap = 0.01
    h = O.collide(sc, 0, (0, 0, 0), ID, 1, (0.5 + 0.3 + gap, 0, 0), ID, max_separation=0.02)
    assert h is not None
    assert h["penetration"] == pytest.approx(-gap, abs=1e-4)
    assert O.collide(sc, 0, (0, 0, 0), ID, 1, (0.5 + 0.3 + 0.05, 0, 0), ID, max_separation=0.02) is None
