"""The state one step hands to the next, brought from the FP64 oracle into a GPU context through
plain records: bodies (cc_set_bodies), the contact cache with the previous time step
(cc_set_contact_cache; world.contacts and world.previousTimeStep, src/contacts.ts:9-58,
src/world.ts:38), the sleep tests (cc_set_sleep_state; motionProperties.sleepTestSpheres /
sleepTestTimer) and the point constraints' impulses (cc_set_point_impulses).  From there the GPU
must continue bit for bit with the oracle, its own cache must equal the oracle's record for
record, and a snapshot of it must restore and continue the same way.  Without the cache the run
differs at once, so the comparison is sensitive to it."""
import numpy as np
import pytest

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH
from crashcat_amd._lib import Context, CrashcatError
from oracle import refcpu as O
from tests.test_gpu_parity import _assert_bodies_equal, _assert_contacts_equal, _same_partition

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
FIELDS = ("position", "quaternion", "center_of_mass_position", "linear_velocity", "angular_velocity")


def _chains_pile():
    sc = SC.mixed_pile(seed=1002, n=120, extent=(6.0, 6.0, 6.0), cylinders=True)
    sc.world_offsets = [0]  # the chain joins the pile's world
    cap = SH.capsule(0.1, 0.3)
    prev = None
    for k in range(6):
        i = sc.add_body(cap, (4.0, 6.0 - 0.8 * k, 4.0), quaternion=(0.0, 0.0, 0.0, 1.0))
        if prev is not None:
            sc.add_point_constraint(prev, i, (0, -0.4, 0), (0, 0.4, 0))
        prev = i
    assert sc.num_worlds == 1
    return sc


def _from_oracle(sc, orc, precision="fp64", cache=True):
    g = Context(precision)
    g.upload(sc)
    g.set_bodies(orc.bodies())
    g.set_sleep_state(orc.sleep_state())
    if orc.npc:
        g.set_point_impulses(orc.point_impulses())
    recs, pdt = orc.contact_cache()
    g.set_contact_cache(recs if cache else recs[:0], pdt)
    return g


def _assert_cache_equal(a, b, what):
    ra, pa = a
    rb, pb = b
    assert pa == pb, what
    assert len(ra) == len(rb), what
    for f in ("body_a", "body_b", "sub_shape_a", "sub_shape_b", "num_points"):
        assert np.array_equal(ra[f], rb[f]), (what, f)
    for k in range(len(ra)):
        n = int(ra["num_points"][k])
        for f in ("local_a", "local_b", "lambda_n", "lambda_t1", "lambda_t2"):
            assert np.array_equal(ra[f][k][:n], rb[f][k][:n]), (what, f, k)


@pytest.mark.parametrize("which,start,steps", [("cube_heap", 45, 100), ("chains", 30, 50)])
def test_oracle_state_continues_bit_for_bit(which, start, steps):
    sc = SC.cube_heap(seed=1001, nx=4, ny=4, nz=4) if which == "cube_heap" else _chains_pile()
    orc = O.OracleWorld(sc)
    orc.step(DT, start)
    g = _from_oracle(sc, orc)
    try:
        slept = 0
        for s in range(steps):
            orc.step(DT)
            g.step(DT, 1)
            g.sync()
            ob, gb = orc.bodies(), g.bodies()
            _assert_bodies_equal(ob, gb, f"step {start + s + 1}")
            _assert_contacts_equal(orc.contacts(), g.contacts(), f"step {start + s + 1}")
            assert np.array_equal(orc.pairs(), g.pairs())
            assert _same_partition(ob["island"], gb["island"])
            if s % 10 == 9:
                _assert_cache_equal(orc.contact_cache(), g.contact_cache(), f"cache at step {start + s + 1}")
            slept = max(slept, int(gb["sleeping"].sum()))
        if which == "cube_heap":
            assert slept > 0  # the run crossed the heap falling asleep, driven by the imported sleep tests
    finally:
        g.close()


def test_without_the_cache_the_run_differs():
    sc = SC.cube_heap(seed=1001, nx=4, ny=4, nz=4)
    orc = O.OracleWorld(sc)
    orc.step(DT, 45)
    g = _from_oracle(sc, orc, cache=False)
    try:
        orc.step(DT)
        g.step(DT, 1)
        g.sync()
        ob, gb = orc.bodies(), g.bodies()
        assert any(not np.array_equal(ob[f], gb[f]) for f in FIELDS)
    finally:
        g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_snapshot_of_an_imported_state_restores(precision):
    """A snapshot taken right after the import restores into a fresh context; both continue
    identically (and, in FP64, identically to the oracle)."""
    sc = _chains_pile()
    orc = O.OracleWorld(sc)
    orc.step(DT, 30)
    a = _from_oracle(sc, orc, precision)
    b = Context(precision)
    try:
        blob = a.snapshot()
        b.upload(sc)
        b.restore(blob)
        for s in range(25):
            orc.step(DT)
            a.step(DT, 1)
            b.step(DT, 1)
        a.sync()
        b.sync()
        ga, gb = a.bodies(), b.bodies()
        for f in FIELDS:
            assert np.array_equal(ga[f], gb[f]), f
        if precision == "fp64":
            _assert_bodies_equal(orc.bodies(), ga, "after 25 steps")
    finally:
        a.close()
        b.close()


def test_warm_state_records_are_validated():
    sc = SC.cube_heap(seed=1001, nx=2, ny=2, nz=2)
    orc = O.OracleWorld(sc)
    orc.step(DT, 30)
    recs, pdt = orc.contact_cache()
    assert len(recs) > 0
    g = Context("fp64")
    try:
        g.upload(sc)
        bad = recs.copy()
        bad["body_a"][0], bad["body_b"][0] = bad["body_b"][0], bad["body_a"][0]
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_ARG}"):
            g.set_contact_cache(bad, pdt)
        bad = recs.copy()
        bad["num_points"][0] = 0
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_ARG}"):
            g.set_contact_cache(bad, pdt)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_ARG}"):
            g.set_contact_cache(np.concatenate([recs[:1], recs[:1]]), pdt)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_CAPACITY}"):
            g.set_contact_cache(np.concatenate([recs] * 200), pdt)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_ARG}"):
            g.set_sleep_state(orc.sleep_state()[:-1])
    finally:
        g.close()
