"""Continuous collision detection (MotionQuality.LINEAR_CAST, src/update.ts:896-1677) on the GPU
path against the FP64 oracle: bit for bit in FP64 over the gallery scene (bullets of every convex
shape at 30-150 m/s into thin walls, a tilted slab, resting and sleeping targets and each other,
a kinematic LINEAR_CAST paddle, slow LINEAR_CAST bodies below the threshold), every step through
the first second and then every few; batched worlds with and without LINEAR_CAST bodies in one
context; FP32 within 1e-4 m of the oracle after one step from one state, at the start and in
mid-flight; and the scenes the GPU path does not take refused at upload."""
import numpy as np
import pytest

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd import shapes as S
from crashcat_amd._lib import Context, CrashcatError
from oracle import refcpu as O

from test_gpu_parity import DT, _assert_bodies_equal, _fp32_one_step, _lockstep_fp64, _step_error

pytestmark = pytest.mark.gpu


def test_fp64_ccd_gallery_bitwise():
    _lockstep_fp64(SC.ccd_gallery(seed=1013), 60, check_every=1)


def test_fp64_ccd_records_bitwise():
    """The CCD records of every step (world.ccd.ccdBodies: time-of-impact order, hit body,
    fraction, fraction plus slop, normal, contact point, material) are the oracle's, bit for bit."""
    sc = SC.ccd_gallery(seed=1017, bullets=32, targets=80)
    orc = O.OracleWorld(sc)
    g = Context("fp64", max_pairs_per_body=16)
    g.upload(sc)
    hits = 0
    try:
        for s in range(1, 41):
            orc.step(DT)
            g.step(DT, 1)
            oc, gc = orc.ccd_bodies(), g.ccd_bodies()
            assert len(oc) == len(gc), s
            for f in oc.dtype.names:
                assert np.array_equal(oc[f], gc[f]), (s, f)
            hits += int((oc["hit_body_index"] >= 0).sum())
        assert hits > 20
    finally:
        g.close()


def test_fp64_ccd_gallery_on_a_plane_bitwise():
    _lockstep_fp64(SC.ccd_gallery(seed=1018, plane_floor=True), 60, check_every=2)


def test_fp64_ccd_gallery_with_compound_targets_bitwise():
    sc = SC.ccd_gallery(seed=1019, compound_targets=True)
    orc = O.OracleWorld(sc)
    g = Context("fp64", max_pairs_per_body=16)
    g.upload(sc)
    try:
        for s in range(1, 41):
            orc.step(DT)
            g.step(DT, 1)
            oc, gc = orc.ccd_bodies(), g.ccd_bodies()
            for f in oc.dtype.names:
                assert np.array_equal(oc[f], gc[f]), (s, f)
            if s % 4 == 0:
                _assert_bodies_equal(orc.bodies(), g.bodies(), f"step {s}")
    finally:
        g.close()


def test_fp64_ccd_gallery_with_decorated_targets_bitwise():
    """Targets in scaled, transformed and offset-centre-of-mass decorators and compounds with scaled
    children (castShapeVsScaled / Transformed / OffsetCenterOfMass / Compound), and scaled and
    offset-centre-of-mass bullets (castScaledVsShape / castOffsetCenterOfMassVsShape)."""
    sc = SC.ccd_gallery(seed=1020, decorated_targets=True, decorated_bullets=True)
    orc = O.OracleWorld(sc)
    g = Context("fp64", max_pairs_per_body=16)
    g.upload(sc)
    try:
        for s in range(1, 41):
            orc.step(DT)
            g.step(DT, 1)
            oc, gc = orc.ccd_bodies(), g.ccd_bodies()
            for f in oc.dtype.names:
                assert np.array_equal(oc[f], gc[f]), (s, f)
            if s % 4 == 0:
                _assert_bodies_equal(orc.bodies(), g.bodies(), f"step {s}")
    finally:
        g.close()


def test_fp64_ccd_gallery_long_bitwise():
    _lockstep_fp64(SC.ccd_gallery(seed=1016, bullets=40, targets=120), 180, check_every=6)


def test_fp64_ccd_batched_worlds_bitwise():
    """Worlds with LINEAR_CAST bodies (one workgroup each) beside worlds without them."""
    sc = SC.Scene()
    for w in range(6):
        if w % 3 == 2:
            SC.mixed_pile(seed=4000 + w, n=40, extent=(5.0, 4.0, 5.0), scene=sc)
        else:
            SC.ccd_gallery(seed=4000 + w, bullets=8 + 4 * w, targets=20 + 5 * w, scene=sc)
    g = Context("fp64", max_pairs_per_body=16)
    g.upload(sc)
    orcs = [O.OracleWorld(sc, w) for w in range(sc.num_worlds)]
    offs = sc.offsets()
    try:
        for s in range(1, 61):
            g.step(DT, 1)
            for o in orcs:
                o.step(DT)
            if s % 5 == 0 or s < 5:
                gb = g.bodies()
                for w, o in enumerate(orcs):
                    ob = o.bodies()
                    for f in ("position", "quaternion", "linear_velocity", "angular_velocity", "sleeping"):
                        assert np.array_equal(ob[f], gb[f][offs[w]:offs[w + 1]]), (s, w, f)
    finally:
        g.close()


@pytest.mark.parametrize("warm", [0, 3])
def test_fp32_ccd_gallery_step_tolerance(warm):
    """One FP32 step from the oracle's state after `warm` steps (rounded to FP32 for both): every
    body within 1e-4 m."""
    sc = SC.ccd_gallery(seed=1013)
    run = O.OracleWorld(sc)
    for _ in range(warm):
        run.step(DT)
    ob, gb, oc, gc = _fp32_one_step(sc, states=SC.as_fp32(run.bodies()), pairs=16, same_contacts=False)
    err, dyn = _step_error(sc, ob, gb)
    print(f"fp32 ccd gallery step after {warm}: p99 {np.quantile(err[dyn], 0.99):.3g} max {err.max():.3g}")
    assert err.max() <= 1e-4


def _ccd_body(sc):
    sc.add_body(S.sphere(0.2), (-2.0, 1.0, 0.0), linear_velocity=(200.0, 0.0, 0.0),
                motion_quality=_abi.CC_MOTION_QUALITY_LINEAR_CAST)


def test_ccd_uploads_refuse_what_the_path_does_not_take():
    g = Context("fp64")
    try:
        bad = []
        sc = SC.Scene()
        pos, idx = S.height_field_mesh(3, 3, 1.0, lambda x, z: 0.0 * x)
        sc.add_body(S.triangle_mesh(pos, idx), (0, 0, 0), motion_type=0)
        _ccd_body(sc)
        sc.end_world()
        bad.append(sc)
        for sc in bad:
            with pytest.raises(CrashcatError, match="LINEAR_CAST"):
                g.upload(sc)
        # a kinematic LINEAR_CAST body is never cast, so any world may hold one
        sc = SC.Scene()
        sc.add_body(S.plane((0.0, 1.0, 0.0)), (0, 0, 0), motion_type=0)
        sc.add_body(S.box((0.3, 0.3, 0.3)), (0, 1, 0), motion_type=_abi.CC_MOTION_KINEMATIC,
                    motion_quality=_abi.CC_MOTION_QUALITY_LINEAR_CAST, linear_velocity=(1.0, 0.0, 0.0))
        sc.end_world()
        g.upload(sc)
        g.step(DT, 1)
    finally:
        g.close()


def test_differentiable_stepping_refuses_ccd_worlds():
    g = Context("fp64", island_groups=1)
    try:
        sc = SC.Scene()
        sc.settings.allow_sleeping = 0
        sc.add_body(S.box((5.0, 0.5, 5.0)), (0, -0.5, 0), motion_type=0)
        _ccd_body(sc)
        sc.end_world()
        g.upload(sc)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_UNSUPPORTED}"):
            g.fwd(DT, 2, 1)
    finally:
        g.close()


def test_fp32_narrowphase_mode_ccd_keeps_bullets_in_the_range():
    """With the narrowphase in FP32 (cc_config.narrowphase_fp32) the shape casts run in FP32 too:
    one step from the oracle's state stays within 1e-3 m at the 99th percentile, and over 40 steps
    no LINEAR_CAST body leaves the walled range below the walls' top."""
    sc = SC.ccd_gallery(seed=1013)
    run = O.OracleWorld(sc)
    for _ in range(3):
        run.step(DT)
    ob, gb, oc, gc = _fp32_one_step(sc, states=SC.as_fp32(run.bodies()), pairs=16, narrowphase="fp32",
                                    same_contacts=False)
    err, dyn = _step_error(sc, ob, gb)
    print(f"fp32-narrowphase ccd gallery step: p99 {np.quantile(err[dyn], 0.99):.3g} max {err.max():.3g}")
    assert np.quantile(err[dyn], 0.99) <= 1e-3
    bd = sc.bodies()
    lc = (bd["motion_quality"] == _abi.CC_MOTION_QUALITY_LINEAR_CAST) & (bd["motion_type"] == _abi.CC_MOTION_DYNAMIC)
    g = Context("fp32", max_pairs_per_body=16, narrowphase="fp32")
    g.upload(sc)
    try:
        for _ in range(40):
            g.step(DT, 1)
            p = g.bodies()["position"]
            out = lc & ((np.abs(p[:, 0]) > 8.05) | (np.abs(p[:, 2]) > 8.05)) & (p[:, 1] < 5.5)
            assert not out.any(), np.where(out)
    finally:
        g.close()
