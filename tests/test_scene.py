"""Host-side scene building: the seeded generators, shape tables and batch layout."""
import math

import numpy as np

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH


def test_splitmix64_published_vector():
    """Vigna's splitmix64 reference output for seed 1234567."""
    x = 1234567
    out = []
    for _ in range(5):
        x, z = SC._splitmix64(x)
        out.append(z)
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]


def test_xoshiro256ss_published_vector():
    """xoshiro256** from state {1, 2, 3, 4}."""
    r = SC.Rng(0)
    r.s = [1, 2, 3, 4]
    assert [r.next_u64() for _ in range(4)] == [11520, 0, 1509978240, 1215971899390074240]


def test_generators_are_deterministic():
    a = SC.cube_heap(seed=1001, nx=3, ny=3, nz=3).bodies()
    b = SC.cube_heap(seed=1001, nx=3, ny=3, nz=3).bodies()
    assert a.tobytes() == b.tobytes()
    c = SC.cube_heap(seed=1002, nx=3, ny=3, nz=3).bodies()
    assert a.tobytes() != c.tobytes()


def test_cube_heap_layout():
    """1 static floor + nx*ny*nz boxes of half-extent 0.5 on a 1.1 m grid, jittered by at most
    0.05 m, with yaw-only rotations and zero velocity."""
    sc = SC.cube_heap(seed=1001)
    b = sc.bodies()
    assert len(b) == 1001
    assert b["motion_type"][0] == _abi.CC_MOTION_STATIC
    dyn = b[1:]
    assert np.all(dyn["motion_type"] == _abi.CC_MOTION_DYNAMIC)
    q = dyn["quaternion"]
    np.testing.assert_array_equal(q[:, [0, 2]], 0.0)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-15)
    assert np.all(dyn["linear_velocity"] == 0)
    assert dyn["position"][:, 1].min() > 0.9
    assert np.all(dyn["inv_mass"] == 1.0 / 1000.0)


def test_mixed_pile_layout():
    sc = SC.mixed_pile(seed=1002, n=400, extent=(12, 8, 12))
    b = sc.bodies()
    assert len(b) == 401 and b["motion_type"][0] == _abi.CC_MOTION_STATIC
    types = np.array([sc.shapes[s].type for s in b["shape"][1:]])
    assert set(types) == {SH.SPHERE, SH.BOX, SH.CAPSULE, SH.CONVEX_HULL}
    pos = b["position"][1:]
    assert pos[:, 1].min() > 0 and np.abs(pos[:, 0]).max() <= 6 and np.abs(pos[:, 2]).max() <= 6
    assert np.allclose(np.linalg.norm(b["quaternion"], axis=1), 1.0)


def test_batched_worlds_layout():
    sc = SC.batched_worlds(seed=1003, num_worlds=16)
    offs = sc.offsets()
    assert len(offs) == 17
    counts = np.diff(offs)
    # floor + n in [8, 512] + 2 chains of 8
    assert np.all(counts >= 1 + 8 + 16) and np.all(counts <= 1 + 512 + 16)
    assert len(sc.point_constraints) == 16 * 2 * 7
    for a, b, *_ in sc.point_constraints:
        w = np.searchsorted(offs, a, side="right") - 1
        assert offs[w] <= b < offs[w + 1]


def test_replicate_world_redraws_dynamic_bodies_only():
    base = SC.mixed_pile(seed=1002, n=200, extent=(10, 8, 10))
    batch = SC.replicate_world(base, 3, 1002)
    b = batch.bodies()
    offs = batch.offsets()
    assert list(offs) == [0, 201, 402, 603]
    w0, w1 = b[0:201], b[201:402]
    assert w0.tobytes() == base.bodies().tobytes()
    assert np.array_equal(w0["shape"], w1["shape"])
    assert w1[0].tobytes() == w0[0].tobytes()  # the static floor is copied unchanged
    assert not np.allclose(w0["position"][1:], w1["position"][1:])


def test_motion_mass_diagonalises_inertia():
    """inverse inertia diagonal + rotation reproduce the inverse of the inertia tensor
    (src/body/motion-properties.ts setMassProperties)."""
    rng = np.random.default_rng(0)
    pts = rng.normal(size=(16, 3))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    h = SH.convex_hull(pts * 0.7)
    inv_m, inv_d, q = SH.motion_mass(h.mass, h.inertia)
    x, y, z, w = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    np.testing.assert_allclose(r @ np.diag(inv_d) @ r.T, np.linalg.inv(h.inertia), rtol=1e-9,
                               atol=1e-12 * np.abs(np.linalg.inv(h.inertia)).max())
    assert inv_m == 1.0 / h.mass


def test_hull_of_sphere_points_is_convex_and_closed():
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(16, 3))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    h = SH.convex_hull(pts)
    # every hull point lies on or behind every face plane
    P = np.asarray(h.points)
    for n, c in zip(h.plane_normals, h.plane_constants):
        assert np.all(P @ np.asarray(n) + c <= 1e-9)
    # Euler characteristic of a closed convex polyhedron: V - E + F = 2
    edges = set()
    for f in h.faces:
        for i in range(len(f)):
            edges.add(tuple(sorted((f[i], f[(i + 1) % len(f)]))))
    used = {v for f in h.faces for v in f}
    assert len(used) - len(edges) + len(h.faces) == 2
    assert h.volume < 4 / 3 * math.pi
