"""Compound shapes (src/shapes/compound.ts) on the GPU path against the FP64 oracle: bit for bit
in FP64 over a pile of compounds of 2-4 children (boxes, spheres, capsules, cylinders, hulls,
some scaled) with plain bodies, including every contact's sub-shape ids; batched worlds of
them; FP32 within 1e-4 m after one step from one state with the oracle's contact points; and
compounds in scenes the GPU path does not take refused at upload."""
import numpy as np
import pytest

from crashcat_amd import scene as SC
from crashcat_amd import shapes as S
from crashcat_amd._lib import Context, CrashcatError
from oracle import refcpu as O

from test_gpu_parity import DT, _fp32_one_step, _lockstep_fp64, _point_set_differs, _step_error

pytestmark = pytest.mark.gpu


def test_fp64_compound_pile_bitwise():
    _lockstep_fp64(SC.compound_pile(seed=1012, n=200), 120, check_every=3)


def test_fp64_compound_batched_worlds_bitwise():
    sc = SC.Scene()
    for w in range(6):
        SC.compound_pile(seed=3000 + w, n=30 + 15 * w, extent=(5.0, 4.0, 5.0), scene=sc)
    g = Context("fp64", max_pairs_per_body=16)
    g.upload(sc)
    orcs = [O.OracleWorld(sc, w) for w in range(sc.num_worlds)]
    offs = sc.offsets()
    try:
        for s in range(1, 61):
            g.step(DT, 1)
            for o in orcs:
                o.step(DT)
            if s % 10 == 0:
                gb, gc = g.bodies(), g.contacts()
                for w, o in enumerate(orcs):
                    ob = o.bodies()
                    for f in ("position", "quaternion", "linear_velocity", "angular_velocity"):
                        assert np.array_equal(ob[f], gb[f][offs[w]:offs[w + 1]]), (s, w, f)
                    oc = o.contacts()
                    sel = (gc["body_a"] >= offs[w]) & (gc["body_a"] < offs[w + 1])
                    assert np.array_equal(oc["sub_shape_a"], gc["sub_shape_a"][sel]), (s, w)
                    assert np.array_equal(oc["sub_shape_b"], gc["sub_shape_b"][sel]), (s, w)
    finally:
        g.close()


def test_fp32_compound_pile_step_tolerance():
    sc = SC.compound_pile(seed=1013, n=200)
    run = O.OracleWorld(sc)
    for _ in range(40):
        run.step(DT)
    ob, gb, oc, gc = _fp32_one_step(sc, states=SC.as_fp32(run.bodies()), pairs=16)
    err, dyn = _step_error(sc, ob, gb)
    print(f"fp32 compound pile step: p99 {np.quantile(err[dyn], 0.99):.3g} max {err.max():.3g}")
    assert np.array_equal(oc["sub_shape_a"], gc["sub_shape_a"]) and np.array_equal(oc["sub_shape_b"], gc["sub_shape_b"])
    assert not [k for k in range(len(oc)) if _point_set_differs(oc, gc, k)]
    assert err.max() <= 1e-4


def test_compound_uploads_refuse_what_the_path_does_not_take():
    g = Context("fp64")
    try:
        sc = SC.Scene()
        sc.add_body(S.plane((0.0, 1.0, 0.0)), (0, 0, 0), motion_type=0)
        sc.add_body(S.compound([((0, 0, 0), (0, 0, 0, 1), S.box((0.3, 0.3, 0.3)))]), (0, 1, 0))
        sc.end_world()
        with pytest.raises(CrashcatError):
            g.upload(sc)
    finally:
        g.close()
