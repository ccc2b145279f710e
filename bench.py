"""Benchmark: body-steps/s of the device-resident rigid-body step at 100k bodies.

Workload (one GPU): the single large pile — 100,000 mixed convex bodies (40%
spheres, 30% boxes, 15% capsules, 15% 16-32 point hulls) dropped into a walled
100x100 m bin, 20 solver iterations, dt = 1/60.  A step is one
stepSimulation(dt) of the whole world on the device (AABB update, broadphase,
narrowphase, coloured solver, integration); inputs are resident in HBM.
`--warmup` steps run untimed (bodies fall and start piling), then `--steps`
steps are timed; value = bodies * steps / seconds.

Also reported: the roofline of the dominant kernel (algorithmic bytes per launch
divided by its average launch time, measured with HIP events on the library's
stream) and the CPU baseline (the single-threaded oracle stepping the same pile
from the same state with the reference's island/Random(0) solver order).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "libgdx-jbullet_amd"))

METRIC = "body-steps/s @100k bodies; solver iters to 1e-4 residual, learned vs plain"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per unit of work (compulsory traffic, each array touched once)
ROW_BYTES = 180          # solver row per sweep: 80 B row + 4 B impulse + 2 x 48 B body velocities
PAIR_BYTES = 1056        # narrowphase pair: ids, transforms, shapes, manifold read + write
PROXY_BYTES = 128        # broadphase proxy: new AABB + stored state read + write
PAIR_OUT_BYTES = 8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--bodies", type=int, default=100_000)
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import jbgpu
    from jbgpu import scenes

    t = time.time()
    sc = scenes.mixed_pile(seed=args.seed, n=args.bodies)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    log(f"[bench] scene {sc.name}: {sc.n_bodies} bodies ({args.bodies} dynamic) built in {time.time() - t:.1f}s")
    ctx = jbgpu.Context(num_iterations=args.iterations)
    ctx.world_create(shapes, verts, bodies)
    dt = 1.0 / 60.0
    t = time.time()
    for k in range(args.warmup):
        ctx.step(dt)
        if (k + 1) % 50 == 0:
            st = ctx.stats()
            log(f"[bench] warmup {k + 1}/{args.warmup}: pairs {st['pairs']} manifolds {st['manifolds']} "
                f"colors {st['colors']} epa {st['epa_calls']} ({time.time() - t:.1f}s)")
    ctx.timing(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.step(dt)
        if (k + 1) % 50 == 0:
            log(f"[bench] timed {k + 1}/{args.steps} ({time.perf_counter() - t0:.1f}s)")
    t1 = time.perf_counter()
    secs = t1 - t0
    st = ctx.stats()
    regions, nsteps = ctx.timing_get()
    value = args.bodies * args.steps / secs

    # dominant kernel region and its roofline (bytes per launch / average launch time)
    tim = {k: v for k, v in regions.items() if v[1] > 0}
    dom = max(("solver_iter", "gjk", "manifold", "traverse", "setaabb", "epa"),
              key=lambda k: tim.get(k, (0.0, 0))[0])
    roof = roofline(dom, st, ctx, args)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * secs / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"mixed_pile: {args.bodies} convex bodies (spheres/boxes/capsules/hulls) in a walled "
                               f"100x100 m bin, {args.iterations} solver iterations, dt 1/60, steps "
                               f"{args.warmup + 1}-{args.warmup + args.steps} of the drop",
                   "bodies": args.bodies, "solver_iterations": args.iterations, "broadphase": "dbvt",
                   "solver_order": "colour batches"},
        "phase_ms_per_step": {k: round(v[0] / max(nsteps, 1), 3) for k, v in tim.items()},
        "last_step": {k: st[k] for k in ("pairs", "contacts", "manifolds", "colors", "epa_calls", "epa_fails")},
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ctx, shapes, verts, bodies, args)
    ctx.close()
    print(json.dumps(out))


def roofline(dom, st, ctx, args):
    """Achieved GB/s of the dominant kernel region from its algorithmic bytes per launch."""
    regions, nsteps = ctx.timing_get()
    ms_total, launches = regions[dom]
    if dom == "solver_iter":
        # bytes per step of the iteration kernels: each sweep touches every contact's rows once
        k, man = ctx.pairs()
        npts = man[:, 0].view(np.int32)
        dist = man[:, 4:].reshape(len(man), 4, 32)[:, :, 15]
        valid = (np.arange(4)[None, :] < npts[:, None]) & (dist <= 0)
        rows_normal = int(valid.sum())
        per_iter_bytes = ROW_BYTES * rows_normal * 3  # normal sweep + two friction rows per contact
        launches_per_step = launches / max(nsteps, 1)
        bytes_per_launch = per_iter_bytes * args.iterations / max(launches_per_step, 1)
        unit_note = f"{ROW_BYTES} B per solver row per sweep x {rows_normal} contacts x 3 rows"
    elif dom in ("gjk", "manifold"):
        bytes_per_launch = PAIR_BYTES * st["pairs"]
        unit_note = f"{PAIR_BYTES} B per pair"
    else:
        bytes_per_launch = PROXY_BYTES * args.bodies + PAIR_OUT_BYTES * st["pairs"]
        unit_note = f"{PROXY_BYTES} B per proxy + {PAIR_OUT_BYTES} B per pair"
    avg_s = (ms_total / max(launches, 1)) / 1000.0
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    return {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "avg_launch_us": round(avg_s * 1e6, 3),
            "bytes_per_launch": int(bytes_per_launch), "launches": int(launches), "unit_bytes": unit_note}


def cpu_baseline(ctx, shapes, verts, bodies, args):
    """The oracle (C++ restatement of the Java path, 1 core) stepping the same pile from the GPU's current state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.lib()
    orc = pyoracle.World(shapes, verts, bodies, iterations=args.iterations, order=0)
    orc.set_bodies(ctx.bodies())
    prox, s0 = ctx.bp_proxies()
    orc.set_proxies(prox, s0)
    k, m = ctx.pairs()
    orc.set_pairs(k, m)
    n = 0
    secs = 0.0
    while secs < args.cpu_seconds and n < 50:
        secs += orc.time_steps(1)
        n += 1
        log(f"[bench] cpu baseline step {n}: {secs:.1f}s")
    v = args.bodies * n / secs
    return {"value": round(v, 1), "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} steps of the same {args.bodies}-body pile continued from the GPU state after step "
                      f"{args.warmup + args.steps}, single-threaded C++ restatement of the Java path "
                      f"(reference island order, Random(0) shuffle), host {os.cpu_count()} cores, 1 used"}


if __name__ == "__main__":
    main()
