"""Benchmark of the crashcat step hot path on one MI355X.

Workload (default): the mixed convex pile — 20,000 bodies drawn from spheres, boxes,
capsules and 16-point convex hulls above a static floor (seed 1002) — stepped in FP32 as a
batch of independent worlds (world 0 is the seeded scene, worlds 1.. redraw positions,
orientations and velocities over the same shapes), next to the plain single-world FP64
path on the same GPU and the FP64 CPU oracle on the host.

Prints one JSON line.  `value` is the whole-batch body-steps/s of the FP32 path with the
scene resident in device memory (upload excluded).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
DT = 1.0 / 60.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def byte_model(tot, body_bytes=540):
    """Algorithmic bytes of the v1 byte model, split by phase.  Per item: body 540 B,
    pair 16 B, manifold 876 B, point 228 B; the velocity/position sweeps are accumulated on
    the device (160 B/manifold + 192 B/point per velocity sweep, 240 B/manifold + 24 B/point
    per executed position sweep)."""
    return {
        "body": body_bytes * tot["awake_body_steps"],
        "pairs": 16 * tot["pairs"],
        "setup": 876 * tot["contacts"] + 228 * tot["points"],
        "solve": tot["solver_bytes"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--worlds", type=int, default=32, help="worlds in the FP32 batch")
    ap.add_argument("--bodies", type=int, default=20000)
    ap.add_argument("--block", type=int, default=0, help="solver workgroup size (0: by world size)")
    ap.add_argument("--pairs-per-body", type=int, default=24, help="pair buffer capacity per body")
    ap.add_argument("--contacts-per-body", type=int, default=6, help="contact constraint capacity per body")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-accuracy", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run the FP32 batch for this many steps (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); worlds shard across GPUs by "
              "contiguous world ranges, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    from crashcat_amd import scene as S
    from crashcat_amd._lib import Context

    t0 = time.time()
    base = S.mixed_pile(seed=1002, n=args.bodies)
    log(f"scene: {base.num_bodies} bodies, {len(base.shapes)} shapes, built in {time.time() - t0:.1f}s")
    batch = S.replicate_world(base, args.worlds, 1002)
    nbody = args.bodies * args.worlds

    if args.profile_steps:
        ctx = Context("fp32", max_pairs_per_body=args.pairs_per_body, block_solver=args.block,
                      max_contacts_per_body=args.contacts_per_body)
        ctx.upload(batch)
        ctx.step(DT, args.profile_steps)
        ctx.sync()
        log("profile run done", ctx.counters())
        return

    # ---- FP32 batched
    ctx = Context("fp32", max_pairs_per_body=args.pairs_per_body, block_solver=args.block,
                      max_contacts_per_body=args.contacts_per_body)
    ctx.upload(batch)
    log(f"fp32 batch uploaded: {args.worlds} worlds, {nbody} dynamic bodies")
    for w0 in range(0, args.warmup, 10):
        ctx.step(DT, min(10, args.warmup - w0))
        ctx.sync()
        log(f"  warmup {w0 + min(10, args.warmup - w0)}/{args.warmup}")
    tot0 = ctx.totals()
    ctx.sync()
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
    bm = byte_model(dtot)
    solver_ms = phases.get("solver", 0.0)
    achieved = bm["solve"] / (solver_ms * 1e-3) / 1e9 if solver_ms > 0 else 0.0
    ctx.close()

    # ---- FP64 plain single world
    fp64 = None
    if not args.skip_fp64:
        one = S.replicate_world(base, 1, 1002)
        c64 = Context("fp64", max_pairs_per_body=args.pairs_per_body, block_solver=args.block,
                      max_contacts_per_body=args.contacts_per_body)
        c64.upload(one)
        w64 = max(1, min(args.warmup, 20))
        c64.step(DT, w64)
        c64.sync()
        k64 = max(1, min(args.steps, 40))
        t = time.perf_counter()
        c64.step(DT, k64)
        c64.sync()
        e64 = time.perf_counter() - t
        fp64 = args.bodies * k64 / e64
        log(f"fp64 plain: {k64} steps in {e64:.3f}s -> {fp64:.4g} body-steps/s; phases {c64.phase_times()}")
        c64.close()

    # ---- accuracy vs the FP64 oracle and CPU baseline
    acc = {}
    cpu = None
    if not (args.skip_accuracy and args.skip_cpu):
        from oracle.refcpu import OracleWorld
        one = S.replicate_world(base, 1, 1002)
        orc = OracleWorld(one)
        g = Context("fp32", max_pairs_per_body=args.pairs_per_body, block_solver=args.block,
                      max_contacts_per_body=args.contacts_per_body)
        g.upload(one)
        t = time.perf_counter()
        orc.step(DT)
        t_first = time.perf_counter() - t
        g.step(DT, 1)
        ob, gb = orc.bodies(), g.bodies()
        dyn = one.bodies()["motion_type"] == 2
        acc["step1_max_pos_err_m"] = float(np.abs(ob["position"] - gb["position"])[dyn].max())
        nsteps = 1
        cpu_steps = max(1, args.cpu_steps)
        drift_steps = 0 if args.skip_accuracy else 60
        tcpu = t_first
        while nsteps < max(drift_steps, cpu_steps):
            t = time.perf_counter()
            orc.step(DT)
            dtc = time.perf_counter() - t
            if nsteps < cpu_steps:
                tcpu += dtc
            g.step(DT, 1)
            nsteps += 1
            if nsteps % 10 == 0:
                log(f"  oracle/fp32 lockstep {nsteps} steps")
        if drift_steps:
            ob, gb = orc.bodies(), g.bodies()
            err = np.linalg.norm(ob["position"] - gb["position"], axis=1)[dyn]
            acc["drift_60_max_m"] = float(err.max())
            acc["drift_60_median_m"] = float(np.median(err))
            acc["drift_60_frac_le_1mm"] = float((err <= 1e-3).mean())
        cpu = args.bodies * cpu_steps / tcpu
        log(f"oracle: {cpu_steps} steps in {tcpu:.2f}s -> {cpu:.4g} body-steps/s; accuracy {acc}")
        g.close()

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
        "data": "synthetic (seeded scene generator)",
        "config": {
            "workload": "mixed convex pile: 20k spheres/boxes/capsules/16-pt hulls over a static floor, dt 1/60, "
                        "10 velocity + 2 position iterations, batched as independent worlds",
            "worlds": args.worlds,
            "bodies_per_world": args.bodies,
            "warmup_steps_before_timing": args.warmup,
        },
        "fp64_plain_body_steps_per_s": fp64,
        "batched_fp32_over_plain_fp64": (value / fp64) if fp64 else None,
        "accuracy_vs_fp64_oracle": acc,
        "phase_ms_total": phases,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_world (sort, levels, warm start, velocity/position sweeps, integration, sleep)",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": None,
        },
        "model_bytes_per_body_step": sum(bm.values()) / max(1, nbody * args.steps),
        "cpu_baseline": None if cpu is None else {
            "value": cpu,
            "unit": "body-steps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"FP64 oracle (C restatement of crashcat's step), first {args.cpu_steps} steps of one "
                      f"{args.bodies}-body world of the same workload, single thread",
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
