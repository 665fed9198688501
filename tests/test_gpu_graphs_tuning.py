"""The step's launch machinery: each phase replayed as a captured hipGraph, and the block sizes
picked by the autotuner.  Neither may change a result: FP64 stays bit-identical to the oracle
with graphs on and off and at every block size the autotuner can pick, and FP32 results are
bit-identical between graphs on and off, between block sizes and between an autotuned and a
fixed-size context.  Every graph replay is checked after the host has synchronised and read the
state back (eager work between replays), over steps whose dt changes (a recapture)."""
import numpy as np
import pytest

from crashcat_amd import scene as SC
from crashcat_amd import shapes as SH
from crashcat_amd._lib import PHASES, Context
from oracle import refcpu as O
from tests.test_gpu_parity import _assert_bodies_equal, _assert_contacts_equal
from tests.test_gpu_step_inputs import _pile

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
DTS = [DT] * 6 + [DT / 2] * 3 + [DT] * 6


@pytest.mark.parametrize("block_solver,block_bodies,graphs", [(128, 128, True), (512, 512, True), (256, 256, False),
                                                              (0, 0, True), (64, 512, False)])
def test_fp64_bitwise_at_every_block_size_with_and_without_graphs(block_solver, block_bodies, graphs):
    sc = SC.cube_heap(seed=1001, nx=6, ny=6, nz=6)
    orc = O.OracleWorld(sc)
    g = Context("fp64", block_solver=block_solver, block_bodies=block_bodies, graphs=graphs)
    g.upload(sc)
    try:
        for s, dt in enumerate(DTS * 3):
            orc.step(dt)
            g.step(dt, 1)
            g.sync()
            _assert_bodies_equal(orc.bodies(), g.bodies(), f"step {s}")
            _assert_contacts_equal(orc.contacts(), g.contacts(), f"step {s}")
    finally:
        g.close()


def _fp32_run(sc, steps, **kw):
    g = Context("fp32", **kw)
    g.upload(sc)
    try:
        out = []
        for s in range(steps):
            g.step(DTS[s % len(DTS)], 1)
            if s % 5 == 4:
                g.sync()
                out.append(g.bodies().copy())
        ph = g.phase_times()
        return out, ph
    finally:
        g.close()


def test_fp32_graphs_and_block_sizes_change_nothing():
    sc = SC.batched_worlds(seed=1003, num_worlds=96, min_bodies=8, max_bodies=200)
    ref, ph = _fp32_run(sc, 40, block_solver=256, block_bodies=256, graphs=False)
    assert list(ph) == PHASES and all(v >= 0 for v in ph.values())
    for kw in (dict(block_solver=256, block_bodies=256, graphs=True), dict(block_solver=128, block_bodies=512),
               dict(block_solver=512, block_bodies=128), dict()):
        got, _ = _fp32_run(sc, 40, **kw)
        for k, (a, b) in enumerate(zip(ref, got)):
            for f in ("position", "quaternion", "linear_velocity", "angular_velocity"):
                assert np.array_equal(a[f], b[f]), (kw, k, f)


def test_single_large_world_fp32_at_1024_threads():
    """The largest solver block (1024 threads) on one world of a few thousand bodies equals 256."""
    sc = SC.cube_heap(seed=1001, nx=14, ny=14, nz=14)
    ref, _ = _fp32_run(sc, 20, block_solver=256, island_groups=1)
    got, _ = _fp32_run(sc, 20, block_solver=1024, island_groups=1)
    for a, b in zip(ref, got):
        assert np.array_equal(a["position"], b["position"])
        assert np.array_equal(a["linear_velocity"], b["linear_velocity"])


@pytest.mark.parametrize("groups,block_solver", [(0, 0), (1, 0), (1, 128), (0, 512)])
def test_fp64_chains_and_kinematics_at_tuned_and_fixed_block_sizes(groups, block_solver):
    """A pile with a capsule chain on point constraints and a moving kinematic body, solved as
    island groups and as one workgroup, at autotuned and fixed solver block sizes, FP64 bit for
    bit against the oracle."""
    sc = _pile(21, n=120, kinematic=1)
    cap = SH.capsule(0.1, 0.3)
    prev = None
    for k in range(8):
        i = sc.add_body(cap, (2.0, 6.0 - 0.8 * k, 1.0), quaternion=(0.0, 0.0, 0.0, 1.0))
        if prev is not None:
            sc.add_point_constraint(prev, i, (0, -0.4, 0), (0, 0.4, 0))
        prev = i
    assert sc.num_worlds == 1
    orc = O.OracleWorld(sc)
    g = Context("fp64", island_groups=groups, block_solver=block_solver)
    g.upload(sc)
    try:
        for s_ in range(60):
            orc.step(DT)
            g.step(DT, 1)
            if s_ % 10 == 9:
                g.sync()
                _assert_bodies_equal(orc.bodies(), g.bodies(), f"step {s_}")
                _assert_contacts_equal(orc.contacts(), g.contacts(), f"step {s_}")
    finally:
        g.close()
