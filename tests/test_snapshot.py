"""World snapshots through the C ABI (cc_snapshot / cc_restore): the state one step hands to the
next, in place of the reference's JSON-serialisable world (README.md:2296-2298) with its warm
start in world.contacts (src/contacts.ts:9-58) and world.previousTimeStep (src/world.ts:38).

The property that pins it whatever the scene: a run that is snapshotted and continued, and the
same snapshot restored into a fresh context (or back into the first, after it ran on) and
continued, give bit-identical bodies, contacts (points, normals, impulses, persisted flags) and
removed contacts, in both precisions, on a pile that starts to sleep, on a terrain with
spheres (runs of mesh manifolds with triangle sub-shape ids in the cache) and on batched
worlds with point-constraint chains (their impulses warm-start too).
"""
import numpy as np
import pytest

from crashcat_amd import scene as SC
from crashcat_amd._lib import Context, CrashcatError

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
SCENES = {
    "mixed_pile": lambda: SC.mixed_pile(seed=1002, n=400, extent=(12.0, 8.0, 12.0)),
    "terrain_spheres": lambda: SC.terrain_drop(vertices=24, n=48, drop=(0.5, 3.0), spheres=True),
    "batched_chains": lambda: SC.batched_worlds(seed=1003, num_worlds=8, min_bodies=8, max_bodies=64),
    "ccd_gallery": lambda: SC.ccd_gallery(seed=1013),
}
FIELDS = ["position", "quaternion", "center_of_mass_position", "linear_velocity", "angular_velocity", "aabb_min",
          "aabb_max", "sleeping"]


def _same_partition(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if not np.array_equal(a < 0, b < 0):
        return False
    m = a >= 0
    pairs = set(zip(a[m].tolist(), b[m].tolist()))
    return len(pairs) == len(set(a[m].tolist())) == len(set(b[m].tolist()))


def _state(ctx):
    ctx.sync()
    return ctx.bodies(), ctx.contacts(), ctx.removed_contacts()


def _assert_same(s1, s2, ctx):
    (b1, c1, r1), (b2, c2, r2) = s1, s2
    for f in FIELDS:
        assert np.array_equal(b1[f], b2[f]), f"{ctx}: bodies {f}"
    assert _same_partition(b1["island"], b2["island"]), f"{ctx}: islands"
    assert c1.tobytes() == c2.tobytes(), f"{ctx}: contacts"
    assert r1.tobytes() == r2.tobytes(), f"{ctx}: removed contacts"


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("which", list(SCENES))
def test_restore_continues_bit_for_bit(prec, which):
    sc = SCENES[which]()
    a = Context(prec, max_pairs_per_body=16)
    b = Context(prec, max_pairs_per_body=16)
    try:
        a.upload(sc)
        a.step(DT, 40)
        a.sync()
        blob = a.snapshot()
        assert len(blob) == len(a.snapshot())
        a.step(DT, 30)
        ref = _state(a)
        assert len(ref[1]) > 0
        b.upload(sc)
        b.restore(blob)
        assert len(b.contacts()) == 0 and len(b.pairs()) == 0  # the taps wait for a step
        b.step(DT, 30)
        _assert_same(ref, _state(b), f"{which} {prec}: fresh context")
        a.step(DT, 25)  # run on, then rewind
        a.restore(blob)
        a.step(DT, 30)
        _assert_same(ref, _state(a), f"{which} {prec}: rewound context")
        # same state, same size; the bytes may differ in record slots past a contact's point
        # count, which keep older values and are never read
        assert len(b.snapshot()) == len(a.snapshot())
    finally:
        a.close()
        b.close()


def test_restore_rejects_other_scenes_and_damaged_blobs():
    sc = SCENES["mixed_pile"]()
    a = Context("fp32", max_pairs_per_body=16)
    other = Context("fp32", max_pairs_per_body=16)
    wide = Context("fp64", max_pairs_per_body=16)
    try:
        a.upload(sc)
        a.step(DT, 5)
        blob = a.snapshot()
        other.upload(SC.mixed_pile(seed=1002, n=300, extent=(12.0, 8.0, 12.0)))
        wide.upload(sc)
        for ctx, data in [(other, blob), (wide, blob), (a, blob[:-1]), (a, blob + b"\0"), (a, b"x" * 8 + blob[8:]),
                          (a, blob[:20])]:
            with pytest.raises(CrashcatError):
                ctx.restore(data)
        a.restore(blob)  # a failed restore leaves the context usable
        a.step(DT, 1)
    finally:
        a.close()
        other.close()
        wide.close()
