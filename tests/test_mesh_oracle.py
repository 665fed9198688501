"""Triangle meshes on the CPU: the host mesh builder against the fixtures of the reference's
builder tests (tst/shapes/triangle-mesh-builder.test.ts) and the BVH's structural invariants,
and the oracle's convex-vs-mesh narrowphase against the reference's collide tests
(tst/collision/collide-shape-vs-shape.test.ts:994-1177) and closed forms."""
import math

import numpy as np
import pytest

from crashcat_amd import scene as SC
from crashcat_amd import shapes as S
from crashcat_amd._abi import CC_MOTION_STATIC
from crashcat_amd._lib import build_triangle_mesh
from oracle import refcpu as R

COS5 = 0.996195


def build(positions, indices, tol=1e-6, leaf=2):
    return build_triangle_mesh(np.asarray(positions, dtype=np.float64).reshape(-1, 3),
                               np.asarray(indices, dtype=np.int32).reshape(-1, 3), tol, COS5, leaf)


# ---- builder: triangle-mesh-builder.test.ts fixtures (triangle counts, active edges)
@pytest.mark.parametrize("positions,indices,tol,count", [
    ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 1, 2], 1e-6, 1),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 1], [0, 1, 2, 1, 2, 3], 1e-6, 2),
    ([0, 0, 0, 1, 0, 0, 2, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 1, 2, 3, 4, 5], 1e-6, 1),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, 1], 1e-6, 0),
    ([0, 0, 0, 1, 0, 0, 2, 0, 0, 3, 0, 0, 4, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 1, 2, 2, 3, 4, 5, 6, 7], 1e-6, 1),
    ([0, 0, 0, 0.00001, 0, 0, 0, 0.00001, 0], [0, 1, 2], 1e-10, 1),
    ([0, 0, 0, 0.00001, 0, 0, 0, 0.00001, 0], [0, 1, 2], 1e-2, 0),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 1, 2, 0, 1, 2], 1e-6, 1),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 1, 2, 1, 2, 0], 1e-6, 1),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0], [0, 1, 2, 1, 2, 3], 1e-6, 2),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 1, 2, 0, 1, 2, 1, 2, 0], 1e-6, 1),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0, 0.000000001, 0.000000001, 0.000000001, 1.000000001, 0, 0, 0, 1.000000001, 0],
     [0, 1, 2, 3, 4, 5], 1e-6, 1),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0, 2, 0, 0, 1, 1, 1], [0, 1, 2, 0, 1, 3, 1, 2, 4], 1e-6, 2),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0, 2, 0, 0], [0, 1, 2, 0, 1, 3, 0, 1, 2], 1e-6, 1),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0, 2, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 1], [0, 1, 2, 0, 1, 3, 4, 5, 6, 1, 2, 7],
     1e-6, 2),
    ([0, 0, 0, 1, 0, 0, 0, 1, 0, 2, 0, 0, 1, 1, 1, 2, 1, 1], [0, 1, 2, 0, 1, 3, 1, 2, 4], 1e-6, 2),
    ([], [], 1e-6, 0),
    ([0, 0, 0], [0, 0, 0], 1e-6, 0),
])
def test_builder_triangle_counts(positions, indices, tol, count):
    pos, tri, nodes = build(positions, indices, tol)
    assert len(tri) == count
    assert (len(nodes) > 0) == (count > 0)


def test_builder_large_soup_and_active_edges():
    # 'handles very large mesh': 100 disjoint triangles, more than 50 node values
    positions, indices = [], []
    for x in range(10):
        for y in range(10):
            positions += [x, y, 0, x + 1, y, 0, x, y + 1, 0]
            b = len(positions) // 3 - 3
            indices += [b, b + 1, b + 2]
    pos, tri, nodes = build(positions, indices)
    assert len(tri) == 100 and nodes.size > 50
    # 'single triangle has all edges active'
    _, tri, _ = build([0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 1, 2])
    assert tri[0, 3] == 0b111
    # 'smooth adjacent triangles have shared edge inactive'
    _, tri, _ = build([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0], [0, 1, 2, 1, 3, 2])
    assert tri[0, 3] & 0b010 == 0 and tri[1, 3] & 0b100 == 0
    assert tri[0, 3] & 0b101 == 0b101 and tri[1, 3] & 0b011 == 0b011
    # a 90 degree crease stays active on both sides
    _, tri, _ = build([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 1, 1], [0, 1, 2, 1, 3, 2])
    assert tri[0, 3] == 0b111 and tri[1, 3] == 0b111


def test_builder_flat_grid_active_edges_are_the_boundary():
    nx, nz = 9, 7
    pos, idx = S.height_field_mesh(nx, nz, 0.5, lambda x, z: 0.0 * x)
    _, tri, _ = build_triangle_mesh(pos, idx)
    bits = sum(bin(int(a)).count("1") for a in tri[:, 3])
    assert bits == 2 * (nx - 1) + 2 * (nz - 1)


def _leaves(nodes, o=0, out=None):
    out = [] if out is None else out
    n = nodes[o // 8]
    if n[7] < 0:
        out.append((int(n[6]), -int(n[7]) - 1, o // 8))
    else:
        for c in (o + 8, int(n[6])):
            ch = nodes[c // 8]
            assert np.all(ch[0:3] >= n[0:3]) and np.all(ch[3:6] <= n[3:6])
            _leaves(nodes, c, out)
    return out


def test_bvh_invariants_on_a_terrain():
    rng = np.random.default_rng(5)
    amp = rng.uniform(0.5, 3.0, 4)
    lam = rng.uniform(8, 64, 4)
    ph = rng.uniform(0, 2 * math.pi, 4)
    dirs = rng.normal(size=(4, 2))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)

    def h(x, z):
        return sum(amp[i] * np.sin(2 * np.pi * (dirs[i, 0] * x + dirs[i, 1] * z) / lam[i] + ph[i]) for i in range(4))
    pos, idx = S.height_field_mesh(40, 33, 0.5, h)
    P, T, N = build_triangle_mesh(pos, idx)
    assert len(T) == 2 * 39 * 32 and len(P) == 40 * 33
    leaves = sorted(_leaves(N))
    covered = 0
    for st, cnt, ni in leaves:
        assert st == covered and 1 <= cnt <= 8
        v = P[T[st:st + cnt, :3]]
        # node bounds are centre -/+ half extent (triangle-mesh-bvh.ts:445-490): within rounding of the vertices
        assert np.all(v.min(axis=(0, 1)) >= N[ni, 0:3] - 1e-12) and np.all(v.max(axis=(0, 1)) <= N[ni, 3:6] + 1e-12)
        covered += cnt
    assert covered == len(T)
    # the welded triangles are the input ones, reordered
    key = lambda t: tuple(sorted(t))
    assert sorted(map(key, T[:, :3].tolist())) == sorted(map(key, idx.tolist()))


# ---- convex vs mesh: collide-shape-vs-shape.test.ts:1033-1177 and closed forms
def _scene():
    sc = SC.Scene()
    mesh = S.triangle_mesh([0, 0, 0, 2, 0, 0, 0, 2, 0], [0, 1, 2])  # normal +z
    return sc, sc.add_shape(mesh)


def test_reference_box_and_capsule_vs_mesh_cases():
    sc, mi = _scene()
    box5 = sc.add_shape(S.box((0.5, 0.5, 0.5)))
    hits = R.collide_mesh(sc, box5, (0.5, 0.5, 0.2), (0, 0, 0, 1), mi)
    assert len(hits) > 0 and hits[0][0].penetration > 0
    hits = R.collide_mesh(sc, sc.add_shape(S.box((0.3, 0.3, 0.3))), (2.2, 0, 0), (0, 0, 0, 1), mi)
    assert len(hits) > 0
    assert R.collide_mesh(sc, box5, (10, 10, 10), (0, 0, 0, 1), mi) == []
    cap = sc.add_shape(S.capsule(0.3, 0.5))
    hits = R.collide_mesh(sc, cap, (0.5, 0.5, 0.15), (0, 0, 0, 1), mi)
    assert len(hits) == 1
    h, sub = hits[0]
    # closed form: the capsule's side (radius 0.3, axis along y) sinks 0.15 below the plane z = 0
    assert h.penetration == pytest.approx(0.15, abs=1e-9)
    assert np.allclose(np.asarray(h.axis) / np.linalg.norm(h.axis), (0, 0, -1), atol=1e-9)
    assert sub == 0xFFFFFFFF  # one triangle: pushIndex adds no bits (sub-shape.ts:60-64)


def test_back_faces_and_sub_shape_ids():
    sc = SC.Scene()
    pos, idx = S.height_field_mesh(3, 3, 1.0, lambda x, z: 0.0 * x)  # 8 triangles, normals +y
    mi = sc.add_shape(S.triangle_mesh(pos, idx))
    bi = sc.add_shape(S.box((0.2, 0.2, 0.2)))
    below = (0.7, -0.15, 0.6)
    assert R.collide_mesh(sc, bi, below, (0, 0, 0, 1), mi, backfaces=False) == []
    assert len(R.collide_mesh(sc, bi, below, (0, 0, 0, 1), mi, backfaces=True)) > 0
    hits = R.collide_mesh(sc, bi, (0.7, 0.15, 0.6), (0, 0, 0, 1), mi, backfaces=False)
    assert hits
    for h, sub in hits:  # 8 triangles: 3 index bits under the empty id
        assert sub & ~0x7 == 0xFFFFFFF8 and (sub & 7) < 8


def _flat_world(floor_mesh, y0, nx=11):
    sc = SC.Scene()
    if floor_mesh:
        pos, idx = S.height_field_mesh(nx, nx, 1.0, lambda x, z: 0.0 * x, origin=(-(nx - 1) / 2, -(nx - 1) / 2))
        sc.add_body(S.triangle_mesh(pos, idx), (0, 0, 0), motion_type=CC_MOTION_STATIC, friction=0.5)
    else:
        sc.add_body(S.box((5.0, 0.5, 5.0)), (0, -0.5, 0), motion_type=CC_MOTION_STATIC, friction=0.5)
    sc.add_body(S.box((0.5, 0.5, 0.5)), (0.23, y0, -0.31), friction=0.5)
    sc.end_world()
    return sc


def test_box_on_flat_mesh_matches_box_on_floor():
    # a resting box on a flat grid of 200 triangles: every triangle hit merges into one manifold
    # with the box's four bottom corners, and the step equals the one on a flat box floor
    y0 = 0.5 - 0.004
    om = R.OracleWorld(_flat_world(True, y0))
    of = R.OracleWorld(_flat_world(False, y0))
    om.step(1 / 60)
    of.step(1 / 60)
    cm, cf = om.contacts(), of.contacts()
    assert len(cm) == 1 and len(cf) == 1
    assert cm[0]["num_points"] == 4 and cf[0]["num_points"] == 4
    assert np.allclose(cm[0]["normal"], (0, 1, 0), atol=1e-12)
    pm = np.sort(np.asarray(cm[0]["position_b"])[:4].reshape(4, 3), axis=0)
    pf = np.sort(np.asarray(cf[0]["position_b"])[:4].reshape(4, 3), axis=0)
    assert np.allclose(pm, pf, atol=1e-9)
    assert cm[0]["sub_shape_a"] != 0xFFFFFFFF and cm[0]["sub_shape_b"] == 0xFFFFFFFF
    bm, bf = om.bodies(), of.bodies()
    assert np.allclose(bm["position"][1], bf["position"][1], atol=1e-9)
    # the two clippings list the corners from different starts, so the sequential friction
    # solve turns the sliding direction about the normal; its speed and the normal velocity agree
    vm, vf = bm["linear_velocity"][1], bf["linear_velocity"][1]
    assert vm[1] == pytest.approx(vf[1], abs=1e-9)
    assert math.hypot(vm[0], vm[2]) == pytest.approx(math.hypot(vf[0], vf[2]), abs=1e-9)


def test_inactive_edges_keep_the_face_normal_on_a_flat_grid():
    # a tilted box whose lowest corner sits on a grid vertex, no movement hint: every hit normal
    # is the face normal (the vertex's edges are all inactive); with a hint straight down the
    # computed normal impedes the movement less and is kept (active-edges.ts:98-108)
    sc = SC.Scene()
    pos, idx = S.height_field_mesh(5, 5, 1.0, lambda x, z: 0.0 * x, origin=(-2, -2))
    mi = sc.add_shape(S.triangle_mesh(pos, idx))
    bi = sc.add_shape(S.box((0.4, 0.4, 0.4)))
    a = 0.3
    q = (math.sin(a / 2) / math.sqrt(2), 0.0, math.sin(a / 2) / math.sqrt(2), math.cos(a / 2))
    hits = R.collide_mesh(sc, bi, (0.0, 0.45, 0.0), q, mi, max_separation=0.02)
    assert hits
    for h, _ in hits:
        n = np.asarray(h.axis) / np.linalg.norm(h.axis)
        assert np.allclose(n, (0, -1, 0), atol=1e-9)
    hits = R.collide_mesh(sc, bi, (0.0, 0.45, 0.0), q, mi, max_separation=0.02, move_dir=(0.0, -1.0, 0.0))
    assert any(abs(np.asarray(h.axis)[1]) / np.linalg.norm(h.axis) < 0.999 for h, _ in hits)
