"""The graph-coloured solve order (cc_config.solver_order = CC_ORDER_COLORED): constraints are
solved colour by colour, where no two constraints of a colour share a dynamic body, instead of
in crashcat's order (contact-constraints.ts:963-1053, the islands' key order).  It is a
different Gauss-Seidel order, so it is judged by its velocity residual against the reference
order (cc_get_solver_residual: the impulse change one more sweep would make), and checked for
what any order must keep: runs repeat bit for bit (a colour that shared a body would race), the
block size changes nothing, heaps stay stacked, and the colouring is much shallower than the
reference order's level schedule."""
import numpy as np
import pytest

from crashcat_amd import _abi
from crashcat_amd import scene as SC
from crashcat_amd._lib import Context, CrashcatError

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
FIELDS = ("position", "quaternion", "linear_velocity", "angular_velocity")


def _run(sc, steps, precision="fp64", **kw):
    g = Context(precision, **kw)
    g.upload(sc)
    res, lev = [], []
    try:
        for _ in range(steps):
            g.step(DT, 1)
            if kw.get("measure_residual"):
                r = g.solver_residual()
                res.append(r["sum"] / max(r["lambda_n_sum"], 1e-300))
            lev.append(g.counters()["max_levels"])
        g.sync()
        return g.bodies(), np.array(res), np.array(lev)
    finally:
        g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_colored_runs_repeat_and_block_size_changes_nothing(precision):
    sc = SC.cube_heap(seed=1001, nx=8, ny=8, nz=8)
    a, _, _ = _run(sc, 90, precision, solver_order="colored", block_solver=256)
    b, _, _ = _run(sc, 90, precision, solver_order="colored", block_solver=256)
    c, _, _ = _run(sc, 90, precision, solver_order="colored", block_solver=128, island_groups=1)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
        assert np.array_equal(a[f], c[f]), f


def test_colored_residual_against_reference_order():
    """Config 1 (1,000 boxes), FP64, 150 steps from the same start: per step, the relative
    residual (sum of |dlambda| over the sum of normal impulses) after the 10 velocity sweeps.
    Over the settled steps 60-150 the coloured order's median stays within 2x of crashcat's
    order's; both are small (the sweeps have converged to within a few percent of the impulses)."""
    sc = SC.cube_heap(seed=1001)
    br, rr, lr = _run(sc, 150, "fp64", measure_residual=True)
    bc, rc, lc = _run(sc, 150, "fp64", measure_residual=True, solver_order="colored")
    mr, mc = np.median(rr[60:]), np.median(rc[60:])
    assert mc <= 2.0 * mr, (mc, mr)
    assert mr < 0.05 and mc < 0.05, (mr, mc)
    # the colouring is shallower than the reference schedule (about 112 levels here)
    assert np.max(lc[60:]) * 3 <= np.min(lr[60:]), (lc[60:].max(), lr[60:].min())
    # the heap still stands as in crashcat's order: every box above the floor, the top layer as
    # high (within half a box), nearly every box at rest
    y, yr = bc["position"][1:, 1], br["position"][1:, 1]
    assert y.min() > 0.3 and y.max() > yr.max() - 0.5, (y.min(), y.max(), yr.max())
    speed = np.linalg.norm(bc["linear_velocity"][1:], axis=1)
    speed_r = np.linalg.norm(br["linear_velocity"][1:], axis=1)
    assert np.median(speed) < 0.05 and np.median(speed) <= 3 * np.median(speed_r) + 0.01, (np.median(speed),
                                                                                         np.median(speed_r))


def test_residual_shrinks_with_more_sweeps():
    """The residual measures convergence: 30 velocity sweeps leave a smaller residual than 5,
    in both orders (settled cube heap, 100 steps)."""
    out = {}
    for order in ("reference", "colored"):
        for it in (5, 30):
            sc = SC.cube_heap(seed=1001, nx=6, ny=6, nz=6)
            sc.settings.velocity_iterations = it
            _, r, _ = _run(sc, 100, "fp64", measure_residual=True, solver_order=order)
            out[(order, it)] = np.median(r[50:])
    for order in ("reference", "colored"):
        assert out[(order, 30)] < out[(order, 5)], out


def test_colored_batched_chains_fp32():
    """Batched worlds with capsule chains on point constraints, FP32, coloured: runs repeat, and
    the chains stay joined as well as in crashcat's order (largest joint gap within 2x + 1 mm)."""
    sc = SC.batched_worlds(seed=1003, num_worlds=96, min_bodies=8, max_bodies=200)
    a, _, _ = _run(sc, 60, "fp32", solver_order="colored")
    b, _, _ = _run(sc, 60, "fp32", solver_order="colored", graphs=False)
    r, _, _ = _run(sc, 60, "fp32")
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f

    def rot(qq, v):
        x, y, z, w = qq
        u = np.array([x, y, z])
        return v + 2.0 * np.cross(u, np.cross(u, v) + w * v)

    def max_gap(st):
        pos, q = st["center_of_mass_position"], st["quaternion"]
        return max(np.linalg.norm(pos[i] + rot(q[i], np.array(la)) - pos[j] - rot(q[j], np.array(lb)))
                   for (i, j, la, lb, _, _) in sc.point_constraints)

    assert max_gap(a) <= 2.0 * max_gap(r) + 1e-3, (max_gap(a), max_gap(r))


def test_residual_needs_the_config_and_diff_needs_reference_order():
    sc = SC.cube_heap(seed=1001, nx=2, ny=2, nz=2)
    g = Context("fp64")
    try:
        g.upload(sc)
        g.step(DT, 1)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_STATE}"):
            g.solver_residual()
    finally:
        g.close()
    sc.settings.allow_sleeping = 0
    g = Context("fp64", island_groups=1, solver_order="colored")
    try:
        g.upload(sc)
        with pytest.raises(CrashcatError, match=f"code {_abi.CC_E_UNSUPPORTED}"):
            g.fwd(DT, 2, 1)
    finally:
        g.close()
