"""Benchmark of the crashcat step hot path on one MI355X.

Headline (`value`): the mixed convex pile -- 20,000 bodies drawn from spheres, boxes, capsules
and 16-point convex hulls in a 60 x 40 x 60 m box above a static floor (seed 1002), dt 1/60,
a 200-step trajectory -- stepped in FP32 as a batch of 256 independent worlds (world 0 is the
seeded scene, worlds 1.. redraw positions, orientations and velocities over the same shapes:
5.12 M bodies).  The timed window is the LAST `--steps` steps of the trajectory, after
`--warmup` untimed steps; the steps before the warmup are advanced untimed too, so that the
window is the same steps of the same trajectory whatever the flags (with --steps 20 --warmup 5:
steps 176-180 warm up, 181-200 are timed).  `value` = bodies x timed steps / timed seconds,
with the scene resident in device memory.  Bulk state, broadphase boxes and the solver are FP32;
GJK, EPA and the convex manifolds run in FP64 from the FP32 poses (the context's default).

Beside it on the same JSON line:
  trajectory        body-steps/s over all steps of the trajectory (the autotuner's one-time
                    trial steps included)
  fp64_plain        the plain single-world FP64 path on world 0, the same window
  phases            every phase's ms per step in the window and its byte-model GB/s
  roofline          the phase with the largest time in the window (all twelve phases); its
                    byte model over its HIP-event time; `traffic` (HBM bytes per launch from
                    calibrated FETCH_SIZE/WRITE_SIZE counters) only when the committed counter
                    summary covers this exact window and the figure stays under peak
  accuracy          FP32 against the FP64 oracle from one state (the FP32 context's rounding of
                    it, given to both): one step from the initial state, one step from the
                    settled pile, and the drift over 60 steps, against 1e-4 m and 1e-3 m; the
                    oracle's own sensitivity to that rounding; the same steps with GJK, EPA and
                    the manifolds in FP32
  narrowphase_fp32_variant
                    the headline window with GJK, EPA and the manifolds in FP32
  fp64_plain_colored_order_body_steps_per_s
                    the same single world in the graph-coloured solve order
  config1           cube heap, 1,000 boxes, 300 steps: FP64 plain path (bit-for-bit check
                    against the oracle at step 300), FP32, and the oracle's own time; both again
                    in the graph-coloured solve order, with both orders' velocity residuals
  config3           4,096 uneven worlds with capsule chains, 60 steps: FP32 and FP64, and FP32
                    with direct launches at the rule-based block sizes (before graphs and tuning)
  config4           50,000 convex bodies on a 522k-triangle terrain, 300 steps, FP32 and FP64
  config5           1,024 throw worlds, 120 steps: forward + adjoint throughput; gradients
                    against the FP64 oracle: central differences before the impact (12 steps,
                    where the loss is smooth), and over the 120 steps (where it is not) the raw
                    central differences by step size and a line fit of 41 samples with the
                    loss's scatter about it
  cpu_baseline      the FP64 oracle on the host, one world per thread, over a bounded sample of
                    the headline workload (worlds of the batch from the GPU state at the start of
                    the timed window)

Profiling helpers (for the committed summaries): --profile runs only the headline trajectory
and exits after its window; --summarize-trace and --summarize-counters reduce rocprofv3 CSV
output to that window; --calibrate copies a buffer of known size for counter calibration.
"""
import argparse
import ctypes as C
import csv
import glob
import json
import os
import re
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
DT = 1.0 / 60.0
TRAJ = 200  # config 2's trajectory length

PHASE_UNITS = {
    # algorithmic bytes per unit of each phase (FP32; 4-byte words, inputs read once, outputs
    # written once); units are the device's running totals (cc_get_totals)
    "accel": "144 B/awake body",
    "grid": "40 B/body",
    "cell_pairs": "28 B/body + 8 B/pair",
    "pair_layout": "8 B/body + 12 B/pair",
    "filter": "16 B/pair + 8 B/GJK pair",
    "gjk": "88 B/GJK pair",
    "epa": "96 B/EPA job",
    "manifold": "876 B/manifold + 228 B/point (both manifold phases; attributed to the first)",
    "islands": "16 B/body + 16 B/pair",
    "solver": "396 B/awake body + (I_v + 1)(160 B/manifold + 192 B/point) + I_p (240 B/manifold + 24 B/point)",
    "cache": "24 B/manifold",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def phase_bytes(t):
    """Algorithmic bytes of each phase from the difference of two cc_get_totals readings."""
    return {
        "accel": 144 * t["awake_body_steps"],
        "grid": 40 * t["body_steps"],
        "cell_pairs": 28 * t["body_steps"] + 8 * t["pairs"],
        "pair_layout": 8 * t["body_steps"] + 12 * t["pairs"],
        "filter": 16 * t["pairs"] + 8 * t["gjk"],
        "gjk": 88 * t["gjk"],
        "epa": 96 * t["epa"],
        "manifold": 876 * t["contacts"] + 228 * t["points"],
        "islands": 16 * t["body_steps"] + 16 * t["pairs"],
        "solver": 396 * t["awake_body_steps"] + t["solver_bytes"],
        "cache": 24 * t["contacts"],
    }


def totals_delta(a, b):
    return {k: b[k] - a[k] for k in b}


def headline_window(steps, warmup):
    traj = max(TRAJ, steps + warmup)
    return traj, traj - steps - warmup


def window_label(worlds, bodies, steps, warmup):
    traj, _ = headline_window(steps, warmup)
    return (f"mixed convex pile, {worlds} worlds x {bodies} bodies, seed 1002, FP32, timed steps "
            f"{traj - steps + 1}-{traj} of {traj} (warmup {traj - steps - warmup + 1}-{traj - steps})")


def read_traffic(kernel, window):
    """HBM bytes per launch of `kernel` from profiles/<kernel>_traffic.md, or None when that
    summary is missing or covers another window."""
    path = os.path.join(HERE, "profiles", f"{kernel}_traffic.md")
    if not os.path.exists(path):
        return None, "no counter summary"
    txt = open(path).read()
    w = re.search(r"^window:\s*(.+)$", txt, re.M)
    if not w or w.group(1).strip() != window:
        return None, f"counter summary covers another window ({w.group(1).strip() if w else 'none'})"
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    return (float(m.group(1)) if m else None), window


def calibrate_copy(nbytes):
    """A device-to-device copy of known size, for counter calibration runs."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copied {nbytes} bytes device to device 3 times")


# ------------------------------------------------------------------ rocprofv3 summaries
def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _short(name):
    """`void cc::k_world<float, 1024, false>(cc::Dev<float>, float)` -> `k_world<float, 1024, false>`"""
    name = re.sub(r"^void\s+", "", name.strip())
    return re.sub(r"\(.*$", "", name.replace("cc::", "")).strip()


def summarize_trace(d, steps):
    """Per-kernel launch count, total and average duration over the window (the last `steps`
    steps of the run: every dispatch from the `steps`-th last k_accel on) of a
    `rocprofv3 --kernel-trace` directory."""
    rows = _rows(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)[0])
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    acc = [r for r in rows if _short(r["Kernel_Name"]).startswith("k_accel")]
    t0 = int(acc[-steps]["Start_Timestamp"])
    win = [r for r in rows if int(r["Start_Timestamp"]) >= t0]
    per = {}
    for r in win:
        k = _short(r["Kernel_Name"])
        per.setdefault(k, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6)
    # the window holds whole steps only; kernels launched fewer times are the autotuner's trials
    per = {k: v for k, v in per.items() if len(v) >= steps or not k.startswith("k_world")}
    span = (max(int(r["End_Timestamp"]) for r in win) - t0) * 1e-6
    print(f"window: last {steps} steps, {len(win)} dispatches, {span:.3f} ms from the first k_accel to the last end")
    print("| kernel | launches | total ms | ms/launch |")
    print("|---|---|---|---|")
    for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        print(f"| `{k}` | {len(v)} | {sum(v):.3f} | {sum(v) / len(v):.4f} |")


def summarize_counters(fdir, wdir, cfdir, cwdir, tdir, steps, window):
    """Calibrated HBM bytes per launch over the window (each kernel's last `steps` dispatches)
    from FETCH_SIZE and WRITE_SIZE runs, with a 256 MiB copy's counters as the scale and the
    window's kernel-trace averages as the time."""
    def per_kernel(d, counter):
        rows = _rows(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)[0])
        by = {}
        for r in rows:
            if r["Counter_Name"] != counter:
                continue
            by.setdefault(_short(r["Kernel_Name"]), []).append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
        # kernels with fewer dispatches than the window are the autotuner's trial block sizes
        return {k: sum(v for _, v in sorted(x)[-steps:]) / steps for k, x in by.items() if len(x) >= steps}

    def copy_kib(d, counter):
        rows = _rows(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)[0])
        return max(float(r["Counter_Value"]) for r in rows if r["Counter_Name"] == counter)

    fetch = per_kernel(fdir, "FETCH_SIZE")
    write = per_kernel(wdir, "WRITE_SIZE")
    nbytes_kib = 256 * 1024  # the calibration copy: 256 MiB read and 256 MiB written
    sf = nbytes_kib / copy_kib(cfdir, "FETCH_SIZE")
    sw = nbytes_kib / copy_kib(cwdir, "WRITE_SIZE")
    rows = _rows(glob.glob(os.path.join(tdir, "**", "*kernel_trace.csv"), recursive=True)[0])
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    times = {}
    for r in rows:
        times.setdefault(_short(r["Kernel_Name"]), []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6)
    print(f"window: {window}")
    print(f"calibration: FETCH_SIZE x {sf:.4f}, WRITE_SIZE x {sw:.4f} (256 MiB device-to-device copy)")
    print("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | GB read | GB written | GB total | ms/launch | GB/s |")
    print("|---|---|---|---|---|---|---|---|")
    for k in sorted(fetch, key=lambda k: -(fetch[k] * sf + write.get(k, 0) * sw)):
        gr = fetch[k] * sf * 1024 / 1e9
        gw = write.get(k, 0.0) * sw * 1024 / 1e9
        t = times.get(k, [])[-steps:]
        ms = sum(t) / len(t) if t else float("nan")
        print(f"| `{k}` | {fetch[k]:.4g} | {write.get(k, 0):.4g} | {gr:.3f} | {gw:.3f} | {gr + gw:.3f} | {ms:.3f} | "
              f"{(gr + gw) / (ms * 1e-3):.0f} |")


# ------------------------------------------------------------------ legs
def pos_err(ob, gb, dyn, bound):
    e = np.linalg.norm(ob["position"] - gb["position"], axis=1)[dyn]
    return {"max_m": float(e.max()), "p99_m": float(np.quantile(e, 0.99)), "median_m": float(np.median(e)),
            "bound_m": bound, "frac_within_bound": float((e <= bound).mean()), "bodies": int(dyn.sum()),
            "bodies_over_bound": int((e > bound).sum())}


def advance(ctx, n, label, chunk=10):
    """Step n steps in chunks, synchronising after each chunk; returns the wall seconds."""
    t = time.perf_counter()
    done = 0
    while done < n:
        k = min(chunk, n - done)
        ctx.step(DT, k)
        ctx.sync()
        done += k
        if done % 50 == 0 or done == n:
            log(f"  {label}: {done}/{n} steps")
    return time.perf_counter() - t


def headline(args, S, Context, batch, nbody, profile=False, narrowphase=None):
    traj, pre = headline_window(args.steps, args.warmup)
    ctx = Context("fp32", max_pairs_per_body=args.pairs_per_body, block_solver=args.block,
                  max_contacts_per_body=args.contacts_per_body, narrowphase=narrowphase or args.narrowphase)
    ctx.upload(batch)
    t_pre = advance(ctx, pre, "untimed advance") if pre else 0.0
    t_warm = advance(ctx, args.warmup, "warmup", chunk=max(1, args.warmup))
    start_state = None
    if not profile and not args.skip_cpu:
        b = ctx.bodies()
        offs = batch.offsets()
        start_state = [b[offs[w]:offs[w + 1]].copy() for w in range(min(args.cpu_worlds, batch.num_worlds))]
    tot0 = ctx.totals()
    t = time.perf_counter()
    ctx.step(DT, args.steps)
    ctx.sync()
    el = time.perf_counter() - t
    tot1 = ctx.totals()
    phases = ctx.phase_times()
    cnt = ctx.counters()
    ctx.close()
    log(f"fp32 batch: steps {traj - args.steps + 1}-{traj} in {el:.3f}s -> {nbody * args.steps / el:.4g} body-steps/s; "
        f"phases {phases}; {cnt}")
    return {"el": el, "traj_s": t_pre + t_warm + el, "traj": traj, "pre": pre, "phases": phases, "counters": cnt,
            "totals": totals_delta(tot0, tot1), "start_state": start_state}


def fp64_plain(args, S, Context, base, order="reference"):
    traj, pre = headline_window(args.steps, args.warmup)
    one = S.replicate_world(base, 1, 1002)
    c = Context("fp64", max_pairs_per_body=args.pairs_per_body, max_contacts_per_body=args.contacts_per_body,
                solver_order=order)
    c.upload(one)
    if pre:
        advance(c, pre, "fp64 untimed advance")
    advance(c, args.warmup, "fp64 warmup", chunk=max(1, args.warmup))
    t = time.perf_counter()
    c.step(DT, args.steps)
    c.sync()
    e = time.perf_counter() - t
    ph = c.phase_times()
    c.close()
    v = args.bodies * args.steps / e
    log(f"fp64 plain ({order} order): {args.steps} steps in {e:.3f}s -> {v:.4g} body-steps/s; phases {ph}")
    return v, {k: x / args.steps for k, x in ph.items()}


def accuracy(args, S, Context, base):
    """FP32 against the FP64 oracle from one state: the scene's (or the settled pile's) poses and
    velocities as the FP32 context holds them, given to both.  The oracle's own sensitivity to that
    rounding (the oracle from the exact FP64 scene against the oracle from its FP32 rounding, after
    one step and after 60) is reported beside it: no program that stores the state in FP32 can sit
    closer to the unrounded run than that, and the pile is chaotic enough that two FP64 runs one
    FP32 rounding apart drift as far in 60 steps as the FP32 path does."""
    from oracle.refcpu import OracleWorld
    one = S.replicate_world(base, 1, 1002)
    dyn = one.bodies()["motion_type"] == 2
    start = S.as_fp32(one.bodies())
    acc = {"state": "the scene's FP32 rounding (positions, orientations, velocities) given to both sides"}
    exact = OracleWorld(one)
    exact.step(DT)
    orc = OracleWorld(one, bodies=start)
    orc.step(DT)
    ob1 = orc.bodies()
    acc["oracle_input_rounding_floor"] = pos_err(exact.bodies(), ob1, dyn, 1e-4)
    g = Context("fp32")
    g.upload(one)
    g.step(DT, 1)
    acc["step_from_initial_state"] = pos_err(ob1, g.bodies(), dyn, 1e-4)
    for s in range(2, 61):
        orc.step(DT)
        exact.step(DT)
        g.step(DT, 1)
        if s % 20 == 0:
            log(f"  oracle/fp32 lockstep {s} steps")
    acc["drift_60_steps"] = pos_err(orc.bodies(), g.bodies(), dyn, 1e-3)
    acc["oracle_input_rounding_floor_60_steps"] = pos_err(exact.bodies(), orc.bodies(), dyn, 1e-3)
    del exact
    g.close()
    g = Context("fp32", narrowphase="fp32")
    g.upload(one)
    g.step(DT, 1)
    acc["narrowphase_fp32_step_from_initial_state"] = pos_err(ob1, g.bodies(), dyn, 1e-4)
    g.close()
    # the settled pile: FP32 state after 100 steps handed to the oracle and to a fresh FP32
    # context; one step of each from that identical state
    g = Context("fp32")
    g.upload(one)
    g.step(DT, 100)
    st = g.bodies()
    g.close()
    orc = OracleWorld(one)
    orc.set_bodies(st)
    orc.step(DT)
    ob = orc.bodies()
    for mode in ("fp64", "fp32"):
        g = Context("fp32", narrowphase=mode)
        g.upload(one)
        g.set_bodies(st)
        g.step(DT, 1)
        key = "step_from_settled_pile" if mode == "fp64" else "narrowphase_fp32_step_from_settled_pile"
        acc[key] = pos_err(ob, g.bodies(), dyn, 1e-4)
        g.close()
    log(f"accuracy {acc}")
    return acc


def cpu_baseline(args, batch, start_state, first_step):
    """The FP64 oracle stepping worlds of the headline batch from the GPU state at the start of
    the timed window, one world per host thread (the oracle's C calls release the GIL)."""
    from oracle.refcpu import OracleWorld
    threads = max(1, min(args.cpu_threads, os.cpu_count() or 1, len(start_state)))
    worlds = []
    for w, st in enumerate(start_state[:threads]):
        o = OracleWorld(batch, w)
        o.set_bodies(st)
        worlds.append(o)
    n = args.cpu_steps

    def run(o):
        for _ in range(n):
            o.step(DT)

    t = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(run, worlds))
    el = time.perf_counter() - t
    v = args.bodies * threads * n / el
    sample = (f"FP64 oracle (C restatement of crashcat's step), worlds 0-{threads - 1} of the headline batch "
              f"({args.bodies} bodies each) from the FP32 GPU state at step {first_step} (contact cache cold), "
              f"{n} steps each, one world per thread on {threads} threads, {el:.1f} s wall")
    log(f"cpu baseline {v:.4g} body-steps/s: {sample}")
    return {"value": v, "unit": "body-steps/s", "cores": threads, "kind": "port", "sample": sample}


def config1(args, S, Context):
    from oracle.refcpu import OracleWorld
    sc = S.cube_heap(seed=1001)
    n = sc.num_bodies - 1
    out = {"workload": "cube heap: 1,000 boxes (he 0.5, density 1000, friction 0.5) on a 10x10x10 grid at 1.1 m "
                       "over a static floor, jitter 0.05 m and random yaw (seed 1001), dt 1/60, 10 velocity + 2 "
                       "position iterations, all 300 steps timed", "bodies": n}
    res = {}
    for prec in ("fp64", "fp32"):
        c = Context(prec)
        c.upload(sc)
        t = time.perf_counter()
        c.step(DT, 300)
        c.sync()
        e = time.perf_counter() - t
        res[prec] = c.bodies()
        out[f"{prec}_body_steps_per_s"] = n * 300 / e
        out[f"{prec}_ms_per_step"] = 1000 * e / 300
        c.close()
    # the graph-coloured solve order, timed the same way, then both orders' velocity residuals
    # (sum |dlambda| of one more sweep over the sum of normal impulses, median of steps 61-300)
    for prec in ("fp64", "fp32"):
        c = Context(prec, solver_order="colored")
        c.upload(sc)
        t = time.perf_counter()
        c.step(DT, 300)
        c.sync()
        e = time.perf_counter() - t
        out[f"{prec}_colored_body_steps_per_s"] = n * 300 / e
        out[f"{prec}_colored_ms_per_step"] = 1000 * e / 300
        out[f"{prec}_colored_levels_last_step"] = c.counters()["max_levels"]
        c.close()
    for order in ("reference", "colored"):
        c = Context("fp64", solver_order=order, measure_residual=True)
        c.upload(sc)
        r = []
        for s_ in range(300):
            c.step(DT, 1)
            q = c.solver_residual()
            r.append(q["sum"] / max(q["lambda_n_sum"], 1e-300))
        out[f"fp64_{order}_relative_residual_median"] = float(np.median(r[60:]))
        out[f"fp64_{order}_levels_last_step"] = c.counters()["max_levels"]
        c.close()
    orc = OracleWorld(sc)
    t = time.perf_counter()
    orc.step(DT, 300)
    e = time.perf_counter() - t
    ob = orc.bodies()
    out["oracle_body_steps_per_s_1_thread"] = n * 300 / e
    out["fp64_bitwise_equal_to_oracle_at_step_300"] = bool(all(np.array_equal(ob[f], res["fp64"][f]) for f in (
        "position", "quaternion", "linear_velocity", "angular_velocity")))
    dyn = sc.bodies()["motion_type"] == 2
    out["fp32_position_diff_at_step_300"] = pos_err(ob, res["fp32"], dyn, 1e-3)
    out["fp64_over_oracle"] = out["fp64_body_steps_per_s"] / out["oracle_body_steps_per_s_1_thread"]
    log(f"config1 {out}")
    return out


def config3(args, S, Context):
    t0 = time.time()
    bw = S.batched_worlds(seed=1003, num_worlds=args.batched_worlds)
    log(f"batched worlds: {bw.num_bodies} bodies in {bw.num_worlds} worlds, {len(bw.point_constraints)} point "
        f"constraints, built in {time.time() - t0:.1f}s")
    out = {"workload": "batched uneven worlds: body count per world log-uniform in [8,512], boxes and spheres on a "
                       "1.1 m grid over a floor plus 2 chains of 8 capsules joined by point constraints, seed 1003, "
                       "dt 1/60, 5 untimed + 60 timed steps",
           "worlds": bw.num_worlds, "bodies": bw.num_bodies, "point_constraints": len(bw.point_constraints)}
    runs = [("fp32", dict()), ("fp64", dict()),
            ("fp32_direct_launches_rule_sizes", dict(block_solver=256, block_bodies=256, graphs=False))]
    for name, kw in runs:
        c = Context(name[:4], max_pairs_per_body=16, max_contacts_per_body=8, **kw)
        c.upload(bw)
        c.step(DT, 5)
        c.sync()
        t = time.perf_counter()
        c.step(DT, 60)
        c.sync()
        e = time.perf_counter() - t
        ph = c.phase_times()
        out[f"{name}_body_steps_per_s"] = bw.num_bodies * 60 / e
        out[f"{name}_phase_ms_per_step"] = {k: v / 60 for k, v in ph.items()}
        log(f"batched worlds {name}: 60 steps in {e:.3f}s -> {bw.num_bodies * 60 / e:.4g} body-steps/s")
        c.close()
    out["graphs_and_tuning_speedup_fp32"] = out["fp32_body_steps_per_s"] / out["fp32_direct_launches_rule_sizes_body_steps_per_s"]
    return out


def config4(args, S, Context):
    t0 = time.time()
    tw = S.terrain_drop(seed=1004, n=args.terrain_bodies)
    log(f"terrain: {tw.num_bodies} bodies, built in {time.time() - t0:.1f}s")
    out = {"workload": "terrain drop: static mesh on a 512 x 512 vertex grid at 0.5 m (522,242 triangles, heights "
                       "from 4 seeded sinusoids, amplitudes 0.5-3 m, wavelengths 8-64 m), boxes (he 0.5), capsules "
                       "(r 0.25, hh 0.25) and 16-point hulls (r 0.5) dropped from 5-30 m, seed 1004, dt 1/60, all "
                       "300 steps timed (in chunks of 20 with a host synchronisation after each)",
           "bodies": args.terrain_bodies}
    for prec in ("fp32", "fp64"):
        c = Context(prec, max_pairs_per_body=16, max_contacts_per_body=8)
        c.upload(tw)
        e = advance(c, 300, f"terrain {prec}", chunk=20)
        cn = c.counters()
        out[f"{prec}_body_steps_per_s"] = args.terrain_bodies * 300 / e
        out[f"{prec}_contacts_last_step"] = cn["num_contacts"]
        log(f"terrain {prec}: 300 steps in {e:.3f}s -> {args.terrain_bodies * 300 / e:.4g} body-steps/s; {cn}")
        c.close()
    return out


def ccd_batch(args, S, Context):
    """CCD (MotionQuality.LINEAR_CAST): a batch of gallery worlds (bullets of every convex shape at
    30-150 m/s into thin walls, resting and sleeping targets and each other) through the first
    second, where the CCD records are; FP32 and FP64, and the same worlds with the bullets
    discrete for the cost of the CCD pass."""
    sc = S.Scene()
    for w in range(args.ccd_worlds):
        S.ccd_gallery(seed=5000 + w, scene=sc)
    nb = sc.num_bodies
    out = {"workload": f"{args.ccd_worlds} CCD gallery worlds ({nb // args.ccd_worlds} bodies each: 24 LINEAR_CAST "
                       "bullets, 6 head-on, a kinematic paddle, 60 targets, 4 slow LINEAR_CAST bodies, walls), seeds "
                       "5000+, dt 1/60, steps 1-60 timed in chunks of 10", "bodies": nb}
    for prec in ("fp32", "fp64"):
        c = Context(prec, max_pairs_per_body=16)
        c.upload(sc)
        c.step(1.0 / 60.0, 1)  # kernels loaded, autotuned
        c.sync()
        c.upload(sc)
        e = advance(c, 60, f"ccd {prec}", chunk=10)
        recs = c.ccd_bodies()
        out[f"{prec}_body_steps_per_s"] = nb * 60 / e
        out[f"{prec}_ms_per_step"] = e / 60 * 1e3
        out[f"{prec}_ccd_records_last_step"] = int(len(recs))
        log(f"ccd {prec}: 60 steps in {e:.3f}s -> {nb * 60 / e:.4g} body-steps/s, {len(recs)} CCD records in the last step")
        c.close()
    d = S.Scene()
    for w in range(args.ccd_worlds):
        S.ccd_gallery(seed=5000 + w, scene=d)
    bodies = d.bodies().copy()  # the same worlds with every body discrete
    bodies["motion_quality"] = 0
    d._bodies = bodies
    c = Context("fp32", max_pairs_per_body=16)
    c.upload(d)
    c.step(1.0 / 60.0, 1)
    c.sync()
    c.upload(d)
    e = advance(c, 60, "ccd fp32 discrete", chunk=10)
    out["fp32_discrete_ms_per_step"] = e / 60 * 1e3
    c.close()
    return out


def config5(args, S, Context):
    from oracle.refcpu import OracleWorld
    nw, steps, K = args.throw_worlds, 120, 10
    sc = S.throw_worlds(num_worlds=nw)
    th = sc.throw
    offs = sc.offsets()
    nb = sc.num_bodies
    out = {"workload": f"differentiable throw: {nw} worlds, 4x4x4 stack of 0.25 m boxes + a 0.4 m 10 kg sphere at "
                       "v0 ~ N((12,2,0),1) m/s, friction 0.5, sleeping off, dt 1/60, 120 steps, checkpoint every "
                       "10 steps; loss = squared distance of the stack centroid to a seeded target",
           "worlds": nw, "bodies": nb}

    def seeds(bs):
        gc = np.zeros((nb, 3))
        for w in range(nw):
            st = th["stack"][w]
            c = bs["center_of_mass_position"][st].mean(axis=0)
            gc[st] = 2 * (c - th["target"][w]) / len(st)
        return gc

    def run_adjoint(prec, nsteps):
        g = Context(prec, island_groups=1)
        try:
            g.upload(sc)
            t = time.perf_counter()
            g.fwd(DT, nsteps, K)
            tf = time.perf_counter() - t
            gc = seeds(g.bodies())
            t = time.perf_counter()
            o = g.bwd(gc)
            tb = time.perf_counter() - t
        finally:
            g.close()
        return o, tf, tb

    grads = {}
    for prec in ("fp32", "fp64"):
        o, tf, tb = run_adjoint(prec, steps)
        grads[prec] = o
        out[f"{prec}_forward_s"] = tf
        out[f"{prec}_backward_s"] = tb
        out[f"{prec}_fwd_bwd_body_steps_per_s"] = nb * steps / (tf + tb)
        log(f"config5 {prec}: fwd {tf:.3f}s bwd {tb:.3f}s")

    def loss(w, nsteps, dv=np.zeros(3), dmu=0.0):
        b = sc.bodies()[offs[w]:offs[w + 1]].copy()  # a private copy: the samples run on threads
        b["linear_velocity"][th["projectile"][w] - offs[w]] += dv
        b["friction"] += dmu
        o = OracleWorld(sc, w, bodies=b)
        o.step(DT, nsteps)
        st = th["stack"][w] - offs[w]
        c = o.bodies()["center_of_mass_position"][st].mean(axis=0)
        return float(np.sum((c - th["target"][w]) ** 2))

    def perturbed(w, nsteps, k, x):
        return loss(w, nsteps, dv=np.eye(3)[k] * x) if k < 3 else loss(w, nsteps, dmu=x)

    def adj(o, w):
        return np.array(list(o["v0"][th["projectile"][w]]) + [o["friction"][offs[w]:offs[w + 1]].sum()])

    threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
    pool = ThreadPoolExecutor(threads)  # oracle worlds step on separate threads
    # (1) before the impact (12 steps: the stack settles, the ball is in flight) the loss is
    # smooth: central differences h = 1e-4
    early = {p: run_adjoint(p, 12)[0] for p in ("fp32", "fp64")}
    smooth = []
    for w in range(args.fd_worlds):
        fdm = list(pool.map(lambda x: loss(w, 12, dmu=x), (1e-4, -1e-4)))
        ref = (fdm[0] - fdm[1]) / 2e-4
        smooth.append({"world": w, "fd_dL_dmu": ref, "adjoint_fp32": float(adj(early["fp32"], w)[3]),
                       "adjoint_fp64": float(adj(early["fp64"], w)[3]),
                       "rel_err_fp32": abs(adj(early["fp32"], w)[3] - ref) / max(abs(ref), 1e-12),
                       "rel_err_fp64": abs(adj(early["fp64"], w)[3] - ref) / max(abs(ref), 1e-12)})
    out["gradients_12_steps_vs_central_fd"] = smooth
    # (2) 120 steps: after the impact the loss of the oracle is not smooth at finite-difference
    # scales (central differences grow like 1/h: contact sets, manifold choices and warm-start
    # matches of the settling stack switch under perturbations of that size).  Reported: the
    # central differences by step size, and for each parameter of two worlds a least-squares line
    # through 41 samples over [-2e-3, 2e-3] with the loss's scatter about it, beside the
    # adjoint's slope times the same range (what a smooth loss would move by)
    sweep = {}
    for h in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
        a, b = pool.map(lambda x: loss(0, steps, dmu=x), (h, -h))
        sweep[f"{h:g}"] = (a - b) / (2 * h)
    out["central_fd_world0_dL_dmu_by_h"] = sweep
    pw = []
    xs = np.linspace(-2e-3, 2e-3, 41)
    for w in range(min(args.fd_worlds, 2)):
        rows = []
        for k in range(4):
            ys = np.array(list(pool.map(lambda x: perturbed(w, steps, k, x), xs)))
            slope, icpt = np.polyfit(xs, ys, 1)
            rows.append({"param": ["v0x", "v0y", "v0z", "mu"][k], "fit_slope": float(slope),
                         "scatter_std": float(np.std(ys - (slope * xs + icpt))),
                         "adjoint_fp64": float(adj(grads["fp64"], w)[k]), "adjoint_fp32": float(adj(grads["fp32"], w)[k]),
                         "adjoint_change_over_range": float(abs(adj(grads["fp64"], w)[k]) * 4e-3)})
        pw.append({"world": w, "params": rows})
    pool.shutdown()
    out["loss_120_steps_line_fit"] = pw
    log(f"config5 gradients: 12 steps {smooth}; 120 steps {pw}; central FD by h {sweep}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps: the last steps of the trajectory")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps just before the timed window")
    ap.add_argument("--worlds", type=int, default=256, help="worlds in the FP32 batch")
    ap.add_argument("--bodies", type=int, default=20000)
    ap.add_argument("--block", type=int, default=0, help="solver workgroup size (0: autotuned)")
    ap.add_argument("--pairs-per-body", type=int, default=24, help="pair buffer capacity per body")
    ap.add_argument("--contacts-per-body", type=int, default=6, help="contact constraint capacity per body")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-accuracy", action="store_true")
    ap.add_argument("--skip-np32-variant", action="store_true", help="skip the FP32-narrowphase headline run")
    ap.add_argument("--cpu-steps", type=int, default=6, help="oracle steps per world in the CPU baseline")
    ap.add_argument("--cpu-worlds", type=int, default=16, help="worlds (one per thread) in the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--batched-worlds", type=int, default=4096, help="worlds of config 3 (0 skips it)")
    ap.add_argument("--terrain-bodies", type=int, default=50000, help="bodies of config 4 (0 skips it)")
    ap.add_argument("--throw-worlds", type=int, default=1024, help="worlds of config 5 (0 skips it)")
    ap.add_argument("--fd-worlds", type=int, default=4, help="config-5 worlds compared with finite differences")
    ap.add_argument("--skip-config1", action="store_true")
    ap.add_argument("--ccd-worlds", type=int, default=512, help="CCD gallery worlds (0 skips them)")
    ap.add_argument("--profile", action="store_true", help="only the headline trajectory (for rocprofv3)")
    ap.add_argument("--narrowphase", default="fp64", choices=("fp64", "fp32"),
                    help="arithmetic of GJK and EPA in the FP32 batch")
    ap.add_argument("--calibrate", type=int, default=0, help="only copy this many bytes device to device")
    ap.add_argument("--summarize-trace", default=None, help="rocprofv3 --kernel-trace output directory")
    ap.add_argument("--summarize-counters", nargs=5, default=None,
                    metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR", "TRACE_DIR"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); worlds would shard across GPUs by "
              "contiguous world ranges, which this program does not launch.", file=sys.stderr)
        sys.exit(2)
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.steps)
        return
    if args.summarize_counters:
        summarize_counters(*args.summarize_counters, args.steps,
                           window_label(args.worlds, args.bodies, args.steps, args.warmup))
        return
    if args.calibrate:
        calibrate_copy(args.calibrate)
        return

    from crashcat_amd import scene as S
    from crashcat_amd._lib import PHASE_KERNEL, Context

    t0 = time.time()
    base = S.mixed_pile(seed=1002, n=args.bodies)
    log(f"scene: {base.num_bodies} bodies, {len(base.shapes)} shapes, built in {time.time() - t0:.1f}s")
    batch = S.replicate_world(base, args.worlds, 1002)
    nbody = args.bodies * args.worlds
    log(f"batch: {args.worlds} worlds, {nbody} dynamic bodies, built in {time.time() - t0:.1f}s")
    window = window_label(args.worlds, args.bodies, args.steps, args.warmup)

    H = headline(args, S, Context, batch, nbody, profile=args.profile)
    if args.profile:
        log(f"profile run done: {window}")
        return
    el, phases, K = H["el"], H["phases"], args.steps
    value = nbody * K / el
    pb = phase_bytes(H["totals"])
    per_phase = {}
    for k, ms in phases.items():
        msps = ms / K
        gbps = pb[k] / K / (msps * 1e-3) / 1e9 if k in pb and msps > 0 else None
        per_phase[k] = {"ms_per_step": msps, "model_GBps": gbps, "kernel": PHASE_KERNEL.get(k),
                        "model": PHASE_UNITS.get(k)}
    dom = max(phases, key=lambda k: phases[k])
    dom_ms = phases[dom] / K
    achieved = pb[dom] / K / (dom_ms * 1e-3) / 1e9 if dom in pb else None
    kernel = PHASE_KERNEL.get(dom, f"{dom} (several kernels)")
    traffic, traffic_note = read_traffic(kernel, window)
    if traffic is not None and traffic / (dom_ms * 1e-3) / 1e9 > PEAK_HBM_GBS:
        traffic_note = f"counter figure {traffic:.4g} B over {dom_ms:.3f} ms exceeds peak: not used"
        traffic = None
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS if achieved else None, "traffic": traffic,
                "kernel": kernel, "phase": dom, "algorithmic_bytes_per_launch": pb.get(dom, 0) / K,
                "launch_ms": dom_ms, "traffic_source": traffic_note}
    if traffic is not None:
        roofline["traffic_GBps"] = traffic / (dom_ms * 1e-3) / 1e9
        roofline["traffic_over_algorithmic"] = traffic / (pb[dom] / K)

    np32 = None
    if not args.skip_np32_variant:
        V = headline(args, S, Context, batch, nbody, profile=True, narrowphase="fp32")
        np32 = {"body_steps_per_s": nbody * K / V["el"], "ms_per_step": 1000.0 * V["el"] / K,
                "phase_ms_per_step": {k: x / K for k, x in V["phases"].items()},
                "note": "the same window with GJK, EPA and the manifolds in FP32 (EPA faces FP64)"}
    fp64, fp64_phases = (None, None) if args.skip_fp64 else fp64_plain(args, S, Context, base)
    fp64c = None if args.skip_fp64 else fp64_plain(args, S, Context, base, "colored")[0]
    acc = None if args.skip_accuracy else accuracy(args, S, Context, base)
    c1 = None if args.skip_config1 else config1(args, S, Context)
    c3 = config3(args, S, Context) if args.batched_worlds else None
    c4 = config4(args, S, Context) if args.terrain_bodies else None
    c5 = config5(args, S, Context) if args.throw_worlds else None
    ccd = ccd_batch(args, S, Context) if args.ccd_worlds else None
    cpu = None if args.skip_cpu else cpu_baseline(args, batch, H["start_state"], H["traj"] - K + 1)

    out = {
        "metric": "body-steps/s batched FP32 vs plain FP64 GPU at <=1e-4 m err; % HBM peak",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * el / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generators, seeds 1001-1005)",
        "config": {
            "workload": f"{window}: 20k spheres/boxes/capsules/16-point hulls in 60x40x60 m over a static floor per "
                        f"world, dt 1/60, 10 velocity + 2 position iterations, steps before the warmup advanced "
                        f"untimed",
            "worlds": args.worlds,
            "bodies_per_world": args.bodies,
        },
        "trajectory": {"steps": H["traj"], "body_steps_per_s": nbody * H["traj"] / H["traj_s"],
                       "note": "every step of the trajectory, host synchronisation every 10 steps, the autotuner's "
                               "one-time trial steps included"},
        "awake_fraction": H["counters"]["num_awake"] / max(1, H["counters"]["num_bodies"]),
        "counters_last_step": H["counters"],
        "fp64_plain_body_steps_per_s": fp64,
        "fp64_plain_phase_ms_per_step": fp64_phases,
        "fp64_plain_colored_order_body_steps_per_s": fp64c,
        "batched_fp32_over_plain_fp64": (value / fp64) if fp64 else None,
        "accuracy_vs_fp64_oracle": acc,
        "narrowphase_fp32_variant": np32,
        "phases": per_phase,
        "roofline": roofline,
        "model_bytes_per_body_step": sum(pb.values()) / max(1, nbody * K),
        "whole_step_model_GBps": sum(pb.values()) / el / 1e9,
        "config1": c1,
        "config3": c3,
        "config4": c4,
        "config5": c5,
        "ccd": ccd,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
