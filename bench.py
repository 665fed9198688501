"""Benchmark of the crashcat step hot path on one MI355X.

Workload: the mixed convex pile -- 20,000 bodies drawn from spheres, boxes, capsules and
16-point convex hulls above a static floor (seed 1002), 200 steps of 1/60 s -- stepped in FP32
as a batch of independent worlds (world 0 is the seeded scene, worlds 1.. redraw positions,
orientations and velocities over the same shapes).  Warmup steps are the first steps of the
same trajectory, the timed steps the rest of it (20 + 180 = the 200 steps by default).

Beside it: the plain single-world FP64 path on the same GPU, the FP32 path's error against
the FP64 oracle (one step from identical state, at the start and in the settled pile; drift
over 60 steps), and the oracle timed on one host core over a bounded sample.

Prints one JSON line.  `value` is the whole-batch body-steps/s of the FP32 path with the
scene resident in device memory (upload excluded).
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
DT = 1.0 / 60.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def phase_bytes(t):
    """Algorithmic bytes per phase of the v1 byte model (4-byte words, each phase reads its
    inputs once and writes its outputs once) from the device's running totals:
    accel 144 B/awake body; manifold kernel (narrowphase output, contact cache, constraint
    setup) 876 B/manifold + 228 B/point; narrowphase (pair read) 16 B/pair; solver kernel (velocity integration, final
    pose, sleep 396 B/awake body, plus the sweeps accumulated on the device:
    (I_v + 1) (160 B/manifold + 192 B/point) + I_p (240 B/manifold + 24 B/point))."""
    return {
        "accel": 144 * t["awake_body_steps"],
        "narrowphase": 16 * t["pairs"],
        "manifold": 876 * t["contacts"] + 228 * t["points"],
        "solver": 396 * t["awake_body_steps"] + t["solver_bytes"],
    }


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary (FETCH_SIZE +
    WRITE_SIZE runs, calibrated against a copy of known size), or None."""
    path = os.path.join(HERE, "profiles", f"{kernel}_traffic.md")
    if not os.path.exists(path):
        return None, None
    txt = open(path).read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    w = re.search(r"workload:\s*(.+)", txt)
    return (float(m.group(1)) if m else None), (w.group(1).strip() if w else None)


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


def pos_err(ob, gb, dyn):
    e = np.linalg.norm(ob["position"] - gb["position"], axis=1)[dyn]
    return {"max_m": float(e.max()), "p99_m": float(np.quantile(e, 0.99)), "median_m": float(np.median(e)),
            "frac_le_1e-4": float((e <= 1e-4).mean())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=180)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--worlds", type=int, default=256, help="worlds in the FP32 batch")
    ap.add_argument("--bodies", type=int, default=20000)
    ap.add_argument("--block", type=int, default=0, help="solver workgroup size (0: by world size)")
    ap.add_argument("--pairs-per-body", type=int, default=24, help="pair buffer capacity per body")
    ap.add_argument("--contacts-per-body", type=int, default=6, help="contact constraint capacity per body")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-accuracy", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=60, help="oracle steps timed from the settled pile")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run the FP32 batch for warmup + this many steps (for rocprofv3)")
    ap.add_argument("--profile-workload", choices=["pile", "terrain"], default="pile",
                    help="workload of --profile-steps (terrain: the 50,000-body drop, 60 warmup steps)")
    ap.add_argument("--calibrate", type=int, default=0, help="only copy this many bytes device to device")
    ap.add_argument("--batched-worlds", type=int, default=4096,
                    help="worlds in the batched uneven-worlds measurement (0 skips it)")
    ap.add_argument("--terrain-bodies", type=int, default=50000,
                    help="bodies dropped on the 512 x 512 terrain mesh in its measurement (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); worlds would shard across GPUs by "
              "contiguous world ranges, which this program does not launch.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate:
        calibrate_copy(args.calibrate)
        return

    from crashcat_amd import scene as S
    from crashcat_amd._lib import Context

    def ctx_for(prec):
        return Context(prec, max_pairs_per_body=args.pairs_per_body, block_solver=args.block,
                       max_contacts_per_body=args.contacts_per_body)

    if args.profile_steps and args.profile_workload == "terrain":
        tw = S.terrain_drop(seed=1004, n=args.terrain_bodies or 50000)
        ctx = Context("fp32", max_pairs_per_body=16, max_contacts_per_body=8, block_solver=args.block)
        ctx.upload(tw)
        for _ in range(3):
            ctx.step(DT, 20)
            ctx.sync()
        ctx.step(DT, args.profile_steps)
        ctx.sync()
        log("terrain profile run done", ctx.counters())
        return

    t0 = time.time()
    base = S.mixed_pile(seed=1002, n=args.bodies)
    log(f"scene: {base.num_bodies} bodies, {len(base.shapes)} shapes, built in {time.time() - t0:.1f}s")
    batch = S.replicate_world(base, args.worlds, 1002)
    nbody = args.bodies * args.worlds
    log(f"batch: {args.worlds} worlds, {nbody} dynamic bodies, built in {time.time() - t0:.1f}s")

    if args.profile_steps:
        ctx = ctx_for("fp32")
        ctx.upload(batch)
        ctx.step(DT, args.warmup)
        ctx.sync()
        ctx.step(DT, args.profile_steps)
        ctx.sync()
        log("profile run done", ctx.counters())
        return

    # ---- FP32 batch: the timed region
    ctx = ctx_for("fp32")
    ctx.upload(batch)
    for w0 in range(0, args.warmup, 10):
        ctx.step(DT, min(10, args.warmup - w0))
        ctx.sync()
        log(f"  warmup {w0 + min(10, args.warmup - w0)}/{args.warmup}")
    tot0 = ctx.totals()
    t = time.perf_counter()
    ctx.step(DT, args.steps)
    ctx.sync()
    el = time.perf_counter() - t
    tot1 = ctx.totals()
    phases = ctx.phase_times()
    cnt = ctx.counters()
    value = nbody * args.steps / el
    log(f"fp32 batch: {args.steps} steps in {el:.3f}s -> {value:.4g} body-steps/s; phases {phases}; {cnt}")
    dtot = {k: tot1[k] - tot0[k] for k in tot1}
    pb = phase_bytes(dtot)
    ctx.close()

    # dominant kernel: the phase with the largest measured time among the single-kernel phases
    kernels = {"solver": "k_world", "manifold": "k_np_manifold", "narrowphase": "k_np_gjk"}
    dom = max(kernels, key=lambda k: phases.get(k, 0.0))
    dom_ms = phases[dom] / args.steps
    achieved = pb[dom] / args.steps / (dom_ms * 1e-3) / 1e9
    traffic, traffic_workload = read_traffic(kernels[dom])
    per_phase = {k: {"ms_per_step": phases[k] / args.steps,
                     "model_GBps": (pb[k] / args.steps / (phases[k] / args.steps * 1e-3) / 1e9) if k in pb and phases[k] > 0
                     else None} for k in phases}

    # ---- FP64 plain single world
    fp64 = None
    if not args.skip_fp64:
        one = S.replicate_world(base, 1, 1002)
        c64 = ctx_for("fp64")
        c64.upload(one)
        c64.step(DT, args.warmup)
        c64.sync()
        k64 = args.steps  # the same steps of the trajectory as the timed FP32 region
        t = time.perf_counter()
        c64.step(DT, k64)
        c64.sync()
        e64 = time.perf_counter() - t
        fp64 = args.bodies * k64 / e64
        log(f"fp64 plain: {k64} steps in {e64:.3f}s -> {fp64:.4g} body-steps/s; phases {c64.phase_times()}")
        c64.close()

    # ---- batched uneven worlds: thousands of piles of 8-512 bodies with capsule chains, 60 steps
    batched = None
    if args.batched_worlds:
        t0 = time.time()
        bw = S.batched_worlds(seed=1003, num_worlds=args.batched_worlds)
        log(f"batched worlds: {bw.num_bodies} bodies in {bw.num_worlds} worlds, "
            f"{len(bw.point_constraints)} point constraints, built in {time.time() - t0:.1f}s")
        batched = {"workload": "batched uneven worlds: body count per world log-uniform in [8,512], boxes and "
                               "spheres on a 1.1 m grid over a floor plus 2 chains of 8 capsules joined by point "
                               "constraints, seed 1003, dt 1/60, 5 warmup + 60 timed steps",
                   "worlds": bw.num_worlds, "bodies": bw.num_bodies,
                   "point_constraints": len(bw.point_constraints)}
        for prec in ("fp32", "fp64"):
            c = Context(prec, max_pairs_per_body=16, max_contacts_per_body=8, block_solver=args.block)
            c.upload(bw)
            c.step(DT, 5)
            c.sync()
            t = time.perf_counter()
            c.step(DT, 60)
            c.sync()
            e = time.perf_counter() - t
            ph = c.phase_times()
            batched[f"{prec}_body_steps_per_s"] = bw.num_bodies * 60 / e
            batched[f"{prec}_phase_ms_per_step"] = {k: v / 60 for k, v in ph.items()}
            log(f"batched worlds {prec}: 60 steps in {e:.3f}s -> {bw.num_bodies * 60 / e:.4g} body-steps/s; "
                f"phases {ph}")
            c.close()
        del bw

    # ---- convex bodies dropped on a triangle-mesh terrain (522,242 triangles)
    terrain = None
    if args.terrain_bodies:
        t0 = time.time()
        tw = S.terrain_drop(seed=1004, n=args.terrain_bodies)
        log(f"terrain: {tw.num_bodies} bodies, built in {time.time() - t0:.1f}s")
        terrain = {"workload": "terrain drop: static mesh on a 512 x 512 vertex grid at 0.5 m (522,242 triangles, "
                               "heights from 4 seeded sinusoids, amplitudes 0.5-3 m, wavelengths 8-64 m), boxes "
                               "(he 0.5), capsules (r 0.25, hh 0.25) and 16-point hulls (r 0.5) dropped from 5-30 m, "
                               "seed 1004, dt 1/60; 60 steps after a 60-step warmup (bodies reach the terrain)",
                   "bodies": args.terrain_bodies}
        for prec in ("fp32", "fp64"):
            c = Context(prec, max_pairs_per_body=16, max_contacts_per_body=8, block_solver=args.block)
            c.upload(tw)
            for w0 in range(0, 60, 20):
                c.step(DT, 20)
                c.sync()
            t = time.perf_counter()
            c.step(DT, 60)
            c.sync()
            e = time.perf_counter() - t
            ph = c.phase_times()
            cn = c.counters()
            terrain[f"{prec}_body_steps_per_s"] = args.terrain_bodies * 60 / e
            terrain[f"{prec}_phase_ms_per_step"] = {k: v / 60 for k, v in ph.items()}
            terrain[f"{prec}_contacts_last_step"] = cn["num_contacts"]
            log(f"terrain {prec}: 60 steps in {e:.3f}s -> {args.terrain_bodies * 60 / e:.4g} body-steps/s; "
                f"phases {ph}; {cn}")
            c.close()
        del tw

    # ---- accuracy vs the FP64 oracle, and the oracle's own speed on the host
    acc = {}
    cpu = None
    cpu_sample = None
    if not (args.skip_accuracy and args.skip_cpu):
        from oracle.refcpu import OracleWorld
        one = S.replicate_world(base, 1, 1002)
        dyn = one.bodies()["motion_type"] == 2
        if not args.skip_accuracy:
            orc = OracleWorld(one)
            g = ctx_for("fp32")
            g.upload(one)
            orc.step(DT)
            g.step(DT, 1)
            acc["step_from_initial_state"] = pos_err(orc.bodies(), g.bodies(), dyn)
            for s in range(2, 61):
                orc.step(DT)
                g.step(DT, 1)
                if s % 20 == 0:
                    log(f"  oracle/fp32 lockstep {s} steps")
            e = pos_err(orc.bodies(), g.bodies(), dyn)
            acc["drift_60_steps"] = e
            g.close()
        # the settled pile: FP32 state after 100 steps, handed to the oracle and to a fresh
        # FP32 context; one step of each from that identical state, then the oracle is timed
        g = ctx_for("fp32")
        g.upload(one)
        g.step(DT, 100)
        st = g.bodies()
        g.close()
        orc = OracleWorld(one)
        orc.set_bodies(st)
        g = ctx_for("fp32")
        g.upload(one)
        g.set_bodies(st)
        t = time.perf_counter()
        orc.step(DT)
        t1 = time.perf_counter() - t
        g.step(DT, 1)
        if not args.skip_accuracy:
            acc["step_from_settled_pile"] = pos_err(orc.bodies(), g.bodies(), dyn)
        g.close()
        if not args.skip_cpu:
            n = max(1, args.cpu_steps)
            tc = t1
            for _ in range(n - 1):
                t = time.perf_counter()
                orc.step(DT)
                tc += time.perf_counter() - t
            cpu = args.bodies * n / tc
            cpu_sample = (f"FP64 oracle (C restatement of crashcat's step), {n} steps of one {args.bodies}-body world "
                          f"from the pile state at step 100 (contacts cold), single thread, {tc:.1f} s")
        log(f"accuracy {acc}; cpu {cpu}")

    out = {
        "metric": "body-steps/s batched FP32 vs plain FP64 GPU at <=1e-4 m err; % HBM peak",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, seed 1002)",
        "config": {
            "workload": "mixed convex pile: 20k spheres/boxes/capsules/16-point hulls in 60x40x60 m over a static "
                        "floor, dt 1/60, 10 velocity + 2 position iterations, 200 steps (warmup + timed), batched "
                        "as independent worlds",
            "worlds": args.worlds,
            "bodies_per_world": args.bodies,
        },
        "awake_fraction": cnt["num_awake"] / max(1, cnt["num_bodies"]),
        "fp64_plain_body_steps_per_s": fp64,
        "batched_fp32_over_plain_fp64": (value / fp64) if fp64 else None,
        "accuracy_vs_fp64_oracle": acc,
        "batched_uneven_worlds": batched,
        "terrain_mesh": terrain,
        "phases": per_phase,
        "roofline": {
            "bound": "hbm",
            "kernel": kernels[dom],
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "traffic_workload": traffic_workload,
            # the same launch by the calibrated FETCH_SIZE + WRITE_SIZE counters (bytes moved)
            "traffic_GBps": (traffic / (dom_ms * 1e-3) / 1e9) if traffic else None,
            "traffic_frac": (traffic / (dom_ms * 1e-3) / 1e9 / PEAK_HBM_GBS) if traffic else None,
        },
        "model_bytes_per_body_step": sum(pb.values()) / max(1, nbody * args.steps),
        "cpu_baseline": None if cpu is None else {
            "value": cpu,
            "unit": "body-steps/s",
            "cores": 1,
            "kind": "port",
            "sample": cpu_sample,
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
