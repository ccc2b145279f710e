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
DROP_STEPS = 600         # the pile's drop: the timed window ends at its last step when the steps fit
TUNE_CANDIDATES = ((2048, 0), (512, 0), (0, 0), (0, 256), (0, 512), (0, 1024))
TUNE_STEPS = 8
TRACE_MARK_BYTES = 48 << 10  # marker copy around the timed window (a kCopyProbe of this size)
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bodies", type=int, default=100_000)
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-every", type=int, default=25,
                    help="time each solver sweep launch with its own events on every k-th timed step")
    ap.add_argument("--batched-worlds", type=int, default=4096)
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-broadphase", action="store_true")
    ap.add_argument("--no-rays", action="store_true")
    ap.add_argument("--rays", type=int, default=1 << 20)
    ap.add_argument("--bp-frames", type=int, default=20)
    ap.add_argument("--phase-steps", type=int, default=30, help="instrumented steps after the timed region")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--sweep-shape", default=None, metavar="BLOCKS,TAIL",
                    help="fix the sweep launch shape instead of autotuning it (e.g. 0,256)")
    ap.add_argument("--trace-window", action="store_true",
                    help="launch a marker copy right before and after the timed steps (for --summarize-trace)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="summarise the kernels between the two window markers of a rocprofv3 kernel trace")
    ap.add_argument("--world-blocks", default=None, metavar="BODIES,THREADS,PIPELINED",
                    help="fix the batched solve's world -> workgroup mapping instead of autotuning it (e.g. 512,256,0)")
    ap.add_argument("--section", choices=("batched", "learned", "broadphase_1m"),
                    help="run only this secondary section and print its JSON (for profiling it alone)")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-kind", default="friction", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-blocks", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-tail", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-warmup", type=int, default=300, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import jbgpu
    from jbgpu import scenes

    if args.summarize_trace:
        return summarize_trace(args.summarize_trace, args.steps)
    if args.pmc_child:
        return pmc_child(args)
    if args.section:
        fn = {"batched": batched, "learned": learned_warm_start, "broadphase_1m": broadphase_1m}[args.section]
        print(json.dumps({args.section: fn(args)}))
        return
    t = time.time()
    sc = scenes.mixed_pile(seed=args.seed, n=args.bodies)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    log(f"[bench] scene {sc.name}: {sc.n_bodies} bodies ({args.bodies} dynamic) built in {time.time() - t:.1f}s")
    ctx = jbgpu.Context(num_iterations=args.iterations)
    ctx.world_create(shapes, verts, bodies)
    dt = 1.0 / 60.0
    # The timed window is fixed by the scene, not by --warmup: the untimed settle phase runs the drop so that
    # settle + autotune + warmup + timed steps end at step DROP_STEPS (the last steps of the 600-step drop, the
    # settled pile) whenever they fit.
    tune_steps = 0 if args.sweep_shape else len(TUNE_CANDIDATES) * (TUNE_STEPS + 1)
    settle = max(0, DROP_STEPS - tune_steps - args.warmup - args.steps)
    t = time.time()
    for k in range(settle):
        ctx.step(dt)
        if (k + 1) % 100 == 0:
            st = ctx.stats()
            log(f"[bench] settle {k + 1}/{settle}: pairs {st['pairs']} manifolds {st['manifolds']} "
                f"colors {st['colors']} epa {st['epa_calls']} ({time.time() - t:.1f}s)")
    if args.sweep_shape:
        blocks, tail = (int(v) for v in args.sweep_shape.split(","))
        ctx.solver_set_blocks(blocks)
        ctx.solver_set_tail(tail)
        tune = {"block_size": blocks, "tail_threshold": tail, "ms_per_step": None, "steps": 0}
    else:
        tune = autotune(ctx, dt)
    ctx.step(dt, n=args.warmup)
    first_timed = settle + tune["steps"] + args.warmup + 1  # the pile's step number of the first timed step
    window = f"steps {first_timed}-{first_timed + args.steps - 1} of the drop"
    # the state at the start of the window: the CPU baseline and the residual comparison start from it
    snap = dict(bodies=ctx.bodies(), proxies=ctx.bp_proxies(), pairs=ctx.pairs())
    st0 = ctx.stats()
    log(f"[bench] window {window}: pairs {st0['pairs']} manifolds {st0['manifolds']} colors {st0['colors']}")
    # timed region: only the sampled per-launch sweep events (the roofline's kernel durations) are recorded
    ctx.timing_kernels(args.kernel_every)
    if args.trace_window:
        ctx.measure_copy(TRACE_MARK_BYTES, 1)  # kCopyProbe of a marker size: the window's first dispatch
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.step(dt)
        if (k + 1) % 50 == 0:
            log(f"[bench] timed {k + 1}/{args.steps} ({time.perf_counter() - t0:.1f}s)")
    t1 = time.perf_counter()
    if args.trace_window:
        ctx.measure_copy(TRACE_MARK_BYTES, 1)  # and its last
    secs = t1 - t0
    st = ctx.stats()
    value = args.bodies * args.steps / secs
    # phase breakdown from an instrumented pass after the timed region (events between phases)
    kstats = ctx.timing_kernels_get()
    ctx.timing_kernels(0)
    regions, nsteps = {}, 0
    if args.phase_steps > 0:
        ctx.timing(True)
        ctx.step(dt, n=args.phase_steps)
        regions, nsteps = ctx.timing_get()
        ctx.timing(False)

    tim = {k: v for k, v in regions.items() if v[1] > 0}
    roof = roofline(kstats, args)
    if roof is not None:
        roof["copy_ceiling_gbs"] = round(ctx.measure_copy(1 << 30, 10), 1)
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
        "config": {"workload": f"mixed_pile: {args.bodies} convex bodies (spheres/boxes/capsules/hulls) dropped in a "
                               f"walled 100x100 m bin, {args.iterations} solver iterations, dt 1/60; timed: {window} "
                               f"(untimed before it: {settle} settle, {tune['steps']} autotune, {args.warmup} warmup "
                               f"steps)",
                   "bodies": args.bodies, "solver_iterations": args.iterations, "broadphase": "dbvt",
                   "solver_order": "colour batches", "first_timed_step": first_timed,
                   "window_start_state": {k: st0[k] for k in ("pairs", "contacts", "manifolds", "colors")}},
        "phase_ms_per_step": {k: round(v[0] / max(nsteps, 1), 3) for k, v in tim.items()},
        "phase_roofline": phase_roofline(tim, nsteps, ctx.stats(), args),
        "phase_note": f"event-instrumented pass of {nsteps} steps after the timed region",
        "last_step": {k: st[k] for k in ("pairs", "contacts", "manifolds", "colors", "epa_calls", "epa_fails")},
        "roofline": roof,
        "autotune": tune,
    }
    # the first window step again from the snapshot, on the device (colour order): its final residual beside the
    # reference order's from the same state (the CPU baseline's first step)
    ctx.set_bodies(snap["bodies"])
    ctx.bp_set_proxies(*snap["proxies"])
    ctx.set_pairs(*snap["pairs"])
    ctx.step(dt)
    rst = ctx.stats()
    out["residual"] = {"step": first_timed, "colour_order": float(rst["residual"]), "iterations": args.iterations,
                       "definition": "||l_k - l_(k-1)|| / max(1, ||l_k||) over every normal and friction row after "
                                     "the last iteration, both from the state before the step"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(shapes, verts, bodies, snap, args, first_timed, out["residual"])
    if not args.no_rays:
        try:
            out["ray_test"] = ray_section(ctx, shapes, verts, bodies, args)
        except Exception as e:  # an optional section must not lose the headline line
            log(f"[bench] ray_test failed: {e!r}")
            out["ray_test"] = {"error": repr(e)}
    ctx.close()
    if roof is not None and not args.no_traffic:
        try:
            roof["traffic"], roof["traffic_detail"] = traffic(args, roof["kind"], tune["block_size"],
                                                              tune["tail_threshold"], first_timed - 1)
        except Exception as e:
            roof["traffic"], roof["traffic_detail"] = None, f"counter passes failed: {e!r}"
    for key, flag, fn in (("batched", args.no_batched, batched), ("learned", args.no_learned, learned_warm_start),
                          ("broadphase_1m", args.no_broadphase, broadphase_1m)):
        if flag:
            continue
        try:
            out[key] = fn(args)
        except Exception as e:  # an optional section must not lose the headline line
            log(f"[bench] {key} failed: {e!r}")
            out[key] = {"error": repr(e)}
    print(json.dumps(out))


PMC_STEPS, PMC_COPY_BYTES = 10, 64 << 20


def pmc_child(args):
    """Workload of one counter pass: the same pile up to the bench's first timed step (--pmc-warmup steps), then
    PMC_STEPS steps whose sweep launches are tallied (launch count and rows, for the algorithmic bytes), then one
    copy of PMC_COPY_BYTES as the counter's scale reference."""
    import jbgpu
    from jbgpu import scenes
    sc = scenes.mixed_pile(seed=args.seed, n=args.bodies)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=args.iterations)
    ctx.world_create(shapes, verts, bodies)
    ctx.solver_set_blocks(args.pmc_blocks)
    ctx.solver_set_tail(args.pmc_tail)
    # every sweep launch is counted; the launches (and rows) of the last PMC_STEPS steps are tallied here and
    # the parent keeps that many of the last counter records
    ctx.step(n=args.pmc_warmup)
    ctx.timing_kernels(1, stride=1)
    ctx.step(n=PMC_STEPS)
    ks = ctx.timing_kernels_get()[args.pmc_kind]
    ctx.measure_copy(PMC_COPY_BYTES, 1)
    ctx.close()
    print(json.dumps({"rows": ks["rows"], "launches": ks["launches"]}))


KERNEL_MATCH = {"friction": (r"kFriction\(", "%kFriction(%"), "normal": (r"kNormal\(", "%kNormal(%"),
                "friction_blocks": (r"kSweepBlocks<true>", "%kSweepBlocks<true>%"),
                "normal_blocks": (r"kSweepBlocks<false>", "%kSweepBlocks<false>%")}


def traffic(args, kind, blocks, tail, warmup):
    """HBM bytes per launch of the dominant sweep kernel from two rocprofv3 counter passes (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each scaled by the ratio of known bytes to what the same counter reads for
    the plain copy probe (kCopyProbe, PMC_COPY_BYTES read and written) in the same pass."""
    regex, like = KERNEL_MATCH[kind]
    import shutil
    import sqlite3
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="jbgpu_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", "150", exe, "--pmc", counter, "--kernel-include-regex", regex + "|kCopyProbe",
               "-d", d, "-o", "run", "--", sys.executable, os.path.abspath(__file__), "--pmc-child", "--bodies",
               str(args.bodies), "--iterations", str(args.iterations), "--seed", str(args.seed), "--pmc-kind", kind,
               "--pmc-blocks", str(blocks), "--pmc-tail", str(tail), "--pmc-warmup", str(warmup)]
        t = time.time()
        p = subprocess.run(cmd, capture_output=True, text=True)
        log(f"[bench] counter pass {counter}: rc {p.returncode} in {time.time() - t:.1f}s")
        dbs = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("results.db")]
        lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
        if p.returncode != 0 or not dbs or not lines:
            return None, f"counter pass {counter} failed (rc {p.returncode})"
        child = json.loads(lines[-1])
        cur = sqlite3.connect(dbs[0]).cursor()
        cols = [r[1] for r in cur.execute("pragma table_info(counters_collection)")]
        order = "dispatch_id" if "dispatch_id" in cols else "rowid"
        # the sweep launches of the counted steps are the last `launches` dispatches of the kernel
        q = (f"select avg(value), count(*) from (select value from counters_collection where counter_name = ? and "
             f"kernel_name like ? order by {order} desc limit ?)")
        kf, nkf = cur.execute(q, (counter, like, int(child["launches"]))).fetchone()
        cp, ncp = cur.execute(q, (counter, "%kCopyProbe%", 1)).fetchone()
        log(f"[bench] counter pass {counter}: ordered by {order}, {nkf} sweep dispatches of the counted steps")
        if not nkf or not ncp or not cp or nkf != child["launches"]:
            return None, f"counter pass {counter}: no dispatches recorded"
        scale = PMC_COPY_BYTES / (cp * 1024.0)
        res[counter] = dict(kib_per_dispatch=kf, dispatches=nkf, copy_kib=cp, scale=scale, child=child)
        shutil.rmtree(d, ignore_errors=True)
    fetch = res["FETCH_SIZE"]["kib_per_dispatch"] * 1024.0 * res["FETCH_SIZE"]["scale"]
    write = res["WRITE_SIZE"]["kib_per_dispatch"] * 1024.0 * res["WRITE_SIZE"]["scale"]
    child = res["FETCH_SIZE"]["child"]
    algo = ROW_BYTES * child["rows"] / max(child["launches"], 1)
    detail = {"workload": f"same pile, block size {blocks}, tail threshold {tail}, every {regex} launch of steps "
                          f"{warmup + 1}-{warmup + PMC_STEPS} (the first timed steps; the last dispatches of the "
                          f"counter pass)",
              "fetch_bytes": int(fetch),
              "write_bytes": int(write), "algorithmic_bytes": int(algo), "traffic_over_algorithmic":
              round((fetch + write) / max(algo, 1.0), 3), "scale_fetch": round(res["FETCH_SIZE"]["scale"], 4),
              "scale_write": round(res["WRITE_SIZE"]["scale"], 4), "dispatches": res["FETCH_SIZE"]["dispatches"]}
    return int(fetch + write), detail


def phase_roofline(tim, nsteps, st, args):
    """Fraction of the 8 TB/s HBM roofline per phase: compulsory bytes per unit of work (predict 208 B/body, AABB
    88 B/body, broadphase 128 B/proxy + 8 B/pair, narrowphase 1056 B/pair, solver setup 552 B/contact, solver
    180 B per row per sweep, integrate 156 B/body) x this step's counts, over the phase's event-timed duration."""
    if not tim or nsteps <= 0:
        return None
    ms = lambda *names: sum(tim[n][0] for n in names if n in tim) / nsteps
    nb, pairs, contacts, rows = args.bodies + 5, st["pairs"], st["contacts"], st["rows"]
    sweeps = args.iterations
    phases = {
        "predict": (208 * args.bodies, ms("predict")),
        "aabb": (88 * nb, ms("setaabb")),
        "broadphase": (128 * nb + 8 * pairs, ms("bvh", "traverse", "keysort", "remap")),
        "narrowphase": (1056 * pairs, ms("bin", "gjk", "epa", "manifold")),
        "solver_setup": (552 * contacts, ms("color", "solver_setup")),
        "solver_iterations": (180 * 3 * rows * sweeps, ms("solver_iter")),
        "integrate": (156 * args.bodies, ms("integrate")),
    }
    out = {}
    for k, (b, t) in phases.items():
        if t <= 0:
            continue
        gbs = b / (t * 1e-3) / 1e9
        out[k] = {"bytes": int(b), "ms": round(t, 4), "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["units"] = {"bodies": args.bodies, "pairs": pairs, "contacts": contacts, "normal_rows": rows,
                    "rows_per_sweep_pair": 3 * rows, "iterations": sweeps}
    return out


def autotune(ctx, dt, steps=TUNE_STEPS):
    """Launch shaping of the sweeps (results are bit-identical for every choice): the block size of the
    block-partitioned sweeps (0: one launch per colour) and, without blocks, the largest colour swept by the
    single-workgroup tail launch.  Each candidate runs a few untimed warmup steps."""
    cands = TUNE_CANDIDATES
    ms = {}
    for blocks, tail in cands:
        ctx.solver_set_blocks(blocks)
        ctx.solver_set_tail(tail)
        ctx.step(dt)
        t = time.perf_counter()
        ctx.step(dt, n=steps)
        ms[(blocks, tail)] = 1e3 * (time.perf_counter() - t) / steps
    best = min(ms, key=ms.get)
    ctx.solver_set_blocks(best[0])
    ctx.solver_set_tail(best[1])
    log(f"[bench] autotune (block size, tail threshold): {ms} -> {best}")
    return {"block_size": best[0], "tail_threshold": best[1],
            "ms_per_step": {f"blocks={b},tail={t}": round(v, 3) for (b, t), v in ms.items()},
            "steps": len(cands) * (steps + 1)}


def roofline(ks, args):
    """Dominant kernel: the solver's colour sweeps.  The sampled launches (the sweeps of every fifth iteration on
    a sampled step of the timed region) are each timed by their own start/stop events
    on the library stream (kernel execution only); achieved = algorithmic bytes per launch / average duration,
    with ROW_BYTES per row per sweep (80 B row + 4 B impulse + 2 x 48 B body velocity read and write)."""
    name = max(ks, key=lambda k: ks[k]["ms"])
    k = ks[name]
    if k["launches"] == 0:
        return None
    avg_s = k["ms"] / k["launches"] / 1000.0
    bytes_per_launch = ROW_BYTES * k["rows"] / k["launches"]
    achieved = bytes_per_launch / avg_s / 1e9
    kernel = {"normal": "kNormal (normal-row colour sweep)", "friction": "kFriction (friction-row colour sweep)",
              "normal_blocks": "kSweepBlocks<false> (normal rows, block-partitioned sweep)",
              "friction_blocks": "kSweepBlocks<true> (friction rows, block-partitioned sweep)"}[name]
    others = {n: {"avg_launch_us": round(v["ms"] * 1e3 / v["launches"], 3), "launches_timed": v["launches"],
                  "achieved_gbs": round(ROW_BYTES * v["rows"] / max(v["ms"], 1e-9) / 1e6, 2)}
              for n, v in ks.items() if n != name and v["launches"] > 0}
    return {"bound": "hbm", "kernel": kernel, "kind": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "avg_launch_us": round(avg_s * 1e6, 3),
            "bytes_per_launch": int(bytes_per_launch), "launches_timed": k["launches"],
            "unit_bytes": f"{ROW_BYTES} B per row per sweep; {k['rows']} rows over the timed launches",
            "other_sweeps": others}


def batched(args):
    """Batched independent worlds: towers of spheres and boxes (sizes as in the mixed pile), own floor plane per
    world, 10 iterations.  The world -> workgroup mapping of the solve is autotuned first (untimed steps), then
    --batched-steps steps run with phase timing, as many as ordinary launches, and as many timed as replayed
    hipGraphs (the headline).  CPU baseline: the oracle stepping the first worlds of the same state (as 4096
    separate worlds would be stepped one after another), about --cpu-seconds of work."""
    import jbgpu
    from jbgpu import autotune, scenes
    t = time.time()
    sc = scenes.batched_towers(seed=args.seed, n_worlds=args.batched_worlds)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    nb = int((bodies[:, 18] != 0).sum())
    log(f"[bench] batched: {args.batched_worlds} worlds, {nb} dynamic bodies, built in {time.time() - t:.1f}s")
    ctx = jbgpu.Context(num_iterations=10)
    ctx.world_create(shapes, verts, bodies, world_ids=sc.world)
    warm = max(args.batched_steps // 2, 1)
    ctx.step(n=warm)
    if args.world_blocks:  # fixed mapping (profiling: the profiler distorts the autotuner's timings)
        cand = tuple(int(x) for x in args.world_blocks.split(","))
        ctx.solver_set_world_blocks(*cand)
        ctx.step()
        tune = {"bodies_per_block": cand[0], "threads": cand[1], "pipelined": bool(cand[2]), "fixed": True, "steps": 1}
    else:
        tune = autotune.world_blocks(ctx, log=log)
    first = warm + tune["steps"] + 1
    snap = dict(bodies=ctx.bodies(), proxies=ctx.bp_proxies(), pairs=ctx.pairs())
    n = args.batched_steps

    def restore():  # the window's first state: bodies, proxies (with the Dbvt stage), pairs and manifolds
        ctx.set_bodies(snap["bodies"])
        ctx.bp_set_proxies(*snap["proxies"])
        ctx.set_pairs(*snap["pairs"])

    # 1. ordinary launches with event timing between the phases (the phase breakdown)
    ctx.timing(not args.trace_window)
    ctx.step(n=n)
    st_phases = ctx.stats()
    regions, nsteps = ctx.timing_get()
    ctx.timing(False)
    # 2. ordinary launches, no events, the same window
    restore()
    t0 = time.perf_counter()
    ctx.step(n=n)
    ordinary_secs = time.perf_counter() - t0
    ordinary_end = ctx.bodies()
    # 3. the same window as replayed hipGraphs (the configuration's launch form), after two untimed steps that
    # capture both graphs.  Under a kernel-trace profiler (--trace-window) the window keeps ordinary launches: the
    # profiler's host side has crashed while a step was being captured, and the kernels are the same either way
    graph = not args.trace_window
    if graph:
        restore()
        ctx.set_step_graph(True)
        ctx.step(n=2)
    restore()
    if args.trace_window:
        ctx.measure_copy(TRACE_MARK_BYTES, 1)  # window markers for --summarize-trace
    t0 = time.perf_counter()
    ctx.step(n=n)
    secs = time.perf_counter() - t0
    if args.trace_window:
        ctx.measure_copy(TRACE_MARK_BYTES, 1)
    same = bool(np.array_equal(ctx.bodies().view(np.uint32), ordinary_end.view(np.uint32)))
    st = ctx.stats()
    # the rest of the configuration's run (its step count), graph replays
    rest_out = None
    rest = int(sc.meta.get("steps", 0)) - (first + n - 1)
    if graph and rest > 0:
        t0 = time.perf_counter()
        ctx.step(n=rest)
        rest_secs = time.perf_counter() - t0
        rest_out = {"timed_window": f"steps {first + n}-{first + n + rest - 1}",
                    "ms_per_step": round(1e3 * rest_secs / rest, 3),
                    "body_steps_per_s": round(nb * rest / rest_secs, 1), "last_step_pairs": ctx.stats()["pairs"],
                    "graph_redos": ctx.stats()["graph_redos"]}
    ctx.close()
    out = {"worlds": args.batched_worlds, "dynamic_bodies": nb, "steps": n,
           "launch_form": "hipGraph replay" if graph else "ordinary launches (profiled window)",
           "timed_window": f"steps {first}-{first + n - 1}",
           "ms_per_step": round(1e3 * secs / n, 3),
           "body_steps_per_s": round(nb * n / secs, 1), "worlds_steps_per_s":
           round(args.batched_worlds * n / secs, 1),
           "ordinary_launches": {"ms_per_step": round(1e3 * ordinary_secs / n, 3), "timed_window": "the same",
                                 "end_state_bit_identical": same},
           "last_step_pairs": st["pairs"], "last_step_colors": st["colors"], "last_step_manifolds": st["manifolds"],
           "last_step_epa_calls": st["epa_calls"], "graph_redos": st["graph_redos"], "autotune": tune,
           "rest_of_run": rest_out,
           "phase_ms_per_step": {k: round(v[0] / max(nsteps, 1), 3) for k, v in regions.items() if v[1] > 0},
           "phase_note": f"ordinary launches with event timing between the phases, the same window "
                         f"({st_phases['pairs']} pairs at the last)"}
    if not args.no_cpu:
        out["cpu_baseline"] = batched_cpu_baseline(shapes, verts, bodies, sc.world, snap, args, first)
    return out


def batched_cpu_baseline(shapes, verts, bodies, world, snap, args, first):
    """The oracle (1 core) stepping whole worlds of the batched scene from the device state at the start of the
    timed window: worlds 0, 1, 2, ... (one DiscreteDynamicsWorld each, stepped one after another), as many as
    fit in about --cpu-seconds, each for two steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    world = np.asarray(world)
    secs, body_steps, worlds = 0.0, 0, 0
    keys, man = snap["pairs"]
    lo = (keys >> np.uint64(32)).astype(np.int64) - 1
    prox, s0 = snap["proxies"]
    for w in range(int(world.max()) + 1):
        if secs >= args.cpu_seconds:
            break
        idx = np.nonzero(world == w)[0]
        remap = np.full(len(world), -1, np.int64)
        remap[idx] = np.arange(len(idx))
        sel = np.nonzero(world[lo] == w)[0]
        k = keys[sel]
        a = remap[(k >> np.uint64(32)).astype(np.int64) - 1] + 1
        b = remap[(k & np.uint64(0xffffffff)).astype(np.int64) - 1] + 1
        m = man[sel].copy()
        hv = m[:, 1:3].view(np.int32)
        hv[:] = remap[hv]
        orc = pyoracle.World(shapes, verts, bodies[idx], iterations=10, order=0)
        orc.set_bodies(snap["bodies"][idx])
        orc.set_proxies(prox[idx], s0)
        orc.set_pairs((a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64), m)
        secs += orc.time_steps(2)
        body_steps += 2 * int((bodies[idx, 18] != 0).sum())
        worlds += 1
    log(f"[bench] batched cpu baseline: {worlds} worlds, {body_steps} body-steps in {secs:.1f}s")
    return {"value": round(body_steps / secs, 1), "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": f"worlds 0-{worlds - 1} of the same scene, each stepped 2 steps (steps {first}-{first + 1}) "
                      f"as its own world from the device state at the timed window's start; single-threaded C++ "
                      f"restatement of the Java path (reference island order, Random(0) shuffle per world)"}


def broadphase_1m(args):
    """Broadphase stress: 1M AABBs (centres U[0,1000)^3, lognormal half extents, constant velocities); each
    frame's AABBs are already resident in HBM; a step = setAabb of every proxy + calculateOverlappingPairs
    (sorted pair keys on the device)."""
    import jbgpu
    from jbgpu import scenes
    n = 1_000_000
    c, h, v = scenes.aabb_cloud(seed=args.seed, n=n)
    first = np.concatenate([c - h, c + h], 1).astype(np.float32)
    frames = np.stack([scenes.aabb_cloud_step(c, h, v) for _ in range(args.bp_frames)])
    out = {"aabbs": n, "frames": args.bp_frames}
    for name, mode in (("dbvt", jbgpu.BP_DBVT), ("sap32", jbgpu.BP_SAP32)):
        ctx = jbgpu.Context(bp_mode=mode, world_min=(-100.0,) * 3, world_max=(1100.0,) * 3)
        dev = ctx.buffer_alloc(frames.nbytes)
        ctx.buffer_upload(dev, frames)
        frame_bytes = frames[0].nbytes
        ctx.bp_create_proxies(first)
        for k in range(3):  # untimed: first frames build the stored state
            ctx.bp_set_aabbs_device(dev + k * frame_bytes, n)
            ctx.bp_calculate_overlapping_pairs()
        per_frame = []
        mark = args.trace_window and name == "dbvt"  # window markers for --summarize-trace (DBVT frames)
        if mark:
            ctx.measure_copy(TRACE_MARK_BYTES, 1)
        t0 = time.perf_counter()
        for k in range(3, args.bp_frames):
            ctx.bp_set_aabbs_device(dev + k * frame_bytes, n)
            per_frame.append(ctx.bp_calculate_overlapping_pairs())
        secs = time.perf_counter() - t0
        pairs = sum(per_frame)
        if mark:
            ctx.measure_copy(TRACE_MARK_BYTES, 1)
        steps = args.bp_frames - 3
        ctx.buffer_free(dev)
        ctx.close()
        out[name] = {"ms_per_step": round(1e3 * secs / steps, 3), "pairs_per_step": pairs // steps,
                     "proxies_per_s": round(n * steps / secs, 1), "pairs_per_s": round(pairs / secs, 1)}
        log(f"[bench] broadphase 1M {name}: {out[name]}")
        if name == "dbvt":
            dbvt_counts = per_frame
    if not args.no_cpu and not args.trace_window:
        out["cpu_baseline"] = broadphase_cpu_baseline(first, frames, dbvt_counts, args)
    return out


def broadphase_cpu_baseline(first, frames, gpu_counts, args):
    """The oracle's DBVT pair search (1 core) on the same frames after the same 3 untimed ones, as many timed
    frames as fit in about --cpu-seconds; its pair counts are compared with the device's for those frames."""
    import jbgpu
    import pyoracle
    orc = pyoracle.Broadphase(jbgpu.BP_DBVT, np.array((-100.0,) * 3, np.float32), np.array((1100.0,) * 3, np.float32))
    orc.create(first)
    cap = 4 * max(gpu_counts) + 4096
    for k in range(3):
        orc.set_all(frames[k])
        orc.step(cap=cap)
    secs, counts = 0.0, []
    for k in range(3, len(frames)):
        if secs >= args.cpu_seconds:
            break
        t0 = time.perf_counter()
        orc.set_all(frames[k])
        counts.append(len(orc.step(cap=cap)))
        secs += time.perf_counter() - t0
    n = len(first)
    log(f"[bench] broadphase cpu baseline: {len(counts)} frames in {secs:.1f}s")
    return {"value": round(n * len(counts) / secs, 1), "unit": "proxies/s", "cores": 1, "kind": "port",
            "sample": f"frames 4-{3 + len(counts)} of the same 1M-AABB sequence (DBVT semantics) after the same 3 "
                      f"untimed frames: setAabb of every proxy + the pair search; single-threaded C++ restatement "
                      f"(not the Java reference: no JDK on the box)",
            "ms_per_frame": round(1e3 * secs / max(len(counts), 1), 1),
            "pair_counts_agree": counts == list(gpu_counts[:len(counts)])}


def learned_warm_start(args):
    """The learned warm-start configuration: 64k boxes (half extents U(0.2, 1)) in 32 pyramids and walls plus 2,000
    falling spheres.  Training contacts from converged solves of seeds 1000-1002 (disjoint from the evaluation
    seed), 200k of them drawn for training on the device; evaluation on seed 1 from step 150 (spheres landing),
    every solve capped at 100 iterations: iterations until the residual first drops to each threshold for the
    plain warm start (0.85 x cached impulses), the model for every contact, the model for contacts without cached
    impulses only, and factor 1.0."""
    import jbgpu
    from jbgpu import learned
    kw = dict(n_boxes=64_000, n_pyramids=32, n_spheres=2000, area=200.0)
    settle = 150
    t = time.time()
    x, y = learned.collect(range(1000, 1003), kw, settle=settle, samples=4, every=8, log=log)
    t_collect = time.time() - t
    pick = np.random.default_rng(0).choice(len(x), min(len(x), 200_000), replace=False)
    x, y = x[pick], y[pick]
    ctx = jbgpu.Context()
    t = time.time()
    losses = ctx.mlp_train(x, y, epochs=60, batch=4096, lr=2e-3, seed=0)
    t_train = time.time() - t
    params, norm = ctx.mlp_get_params()
    ctx.close()
    log(f"[bench] learned: {len(x)} training contacts ({t_collect:.1f}s), loss {losses[0]:.4f} -> {losses[-1]:.4f} "
        f"({t_train:.1f}s)")
    ev = learned.evaluate(params, norm, seed=1, scene_kw=kw, settle=settle, steps=20)
    ev["mfma"] = mlp_throughput(params, norm, ev, len(x) * len(losses), t_train)
    ev.update({"training_contacts": int(len(x)), "epochs": len(losses), "loss_first": round(float(losses[0]), 5),
               "loss_last": round(float(losses[-1]), 5), "train_s": round(t_train, 2),
               "scene": "pyramids: 64,000 boxes in 32 pyramids and walls, 2,000 spheres; eval seed 1, steps "
                        f"{settle + 1}-{settle + 20}, each solve run 100 iterations; training seeds 1000-1002",
               "residual": "r_k = ||l_k - l_(k-1)|| / max(1, ||l_k||) over normal and friction rows"})
    return ev


MLP_FWD_FLOP = 2 * (16 * 64 + 64 * 64 + 64 * 64 + 64 * 3)  # 18,816 per contact
BF16_DENSE_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA


def mlp_throughput(params, norm, ev, train_samples, train_s):
    """MFMA use of the 16-64-64-64-3 bf16 MLP: inference inside the solve (4 slots per active manifold, kernel
    time from the solve's timing region), a 2M-contact inference batch (kernel time only), and training
    (forward + backward = 3 x forward flops per sample, wall time of jbgpu_mlp_train)."""
    import jbgpu
    ctx = jbgpu.Context()
    ctx.mlp_set_params(params, norm)
    x = np.random.default_rng(0).normal(size=(1 << 21, 16)).astype(np.float32)
    ctx.mlp_infer(x[:4096])
    ctx.timing(True)
    reps = 3
    for _ in range(reps):
        ctx.mlp_infer(x)
    reg, _ = ctx.timing_get()
    ctx.close()
    big_ms = reg["mlp"][0] / reps
    big = MLP_FWD_FLOP * len(x) / (big_ms * 1e-3) / 1e12
    le = ev["learned"]
    slots = 4 * le["manifolds"]
    solve = MLP_FWD_FLOP * slots / (le["mlp_ms"] * 1e-3) / 1e12 if le["mlp_ms"] > 0 else None
    train = 3 * MLP_FWD_FLOP * train_samples / train_s / 1e12
    return {"peak_tflops": BF16_DENSE_PEAK_TFLOPS, "flop_per_contact_forward": MLP_FWD_FLOP,
            "infer_2m_contacts_ms": round(big_ms, 4), "infer_2m_tflops": round(big, 2),
            "infer_2m_frac": round(big / BF16_DENSE_PEAK_TFLOPS, 5),
            "in_solve_slots": slots, "in_solve_ms": le["mlp_ms"],
            "in_solve_tflops": None if solve is None else round(solve, 3),
            "train_tflops_wall": round(train, 3), "train_frac": round(train / BF16_DENSE_PEAK_TFLOPS, 6)}


def ray_section(ctx, shapes, verts, bodies, args):
    """World queries on the pile as the timed region left it: CollisionWorld.rayTest with a closest-hit callback
    for args.rays segments between random points of the bin's volume (half of them vertical), rays and results
    resident in device buffers; each call builds the query tree over the current AABBs.  The oracle is timed on
    a sample of the same rays (1 core) and its results checked against the device's."""
    rng = np.random.default_rng(args.seed)
    n = args.rays
    lo, hi = np.float32([-52, -1, -52]), np.float32([52, 40, 52])
    a = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    b = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    v = rng.random(n) < 0.5
    b[v, 0], b[v, 2] = a[v, 0], a[v, 2]
    a[v, 1], b[v, 1] = hi[1], lo[1]
    dev = {k: ctx.buffer_alloc(nb) for k, nb in (("f", a.nbytes), ("t", b.nbytes), ("o", 4 * n), ("fr", 4 * n),
                                                  ("n", 12 * n), ("p", 12 * n))}
    ctx.buffer_upload(dev["f"], a)
    ctx.buffer_upload(dev["t"], b)
    call = lambda: ctx.ray_test_device(n, dev["f"], dev["t"], dev["o"], dev["fr"], dev["n"], dev["p"])
    call()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    secs = (time.perf_counter() - t0) / reps
    for k in dev:
        ctx.buffer_free(dev[k])
    sample = 64
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    orc = pyoracle.World(shapes, verts, bodies, iterations=args.iterations, order=0)
    orc.set_bodies(ctx.bodies())
    t0 = time.perf_counter()
    want = orc.ray_test(a[:sample], b[:sample])
    cpu = time.perf_counter() - t0
    got = ctx.ray_test(a[:sample], b[:sample])
    same = bool(np.array_equal(got[0], want[0]) and all(np.array_equal(x.view(np.uint32), y.view(np.uint32))
                                                        for x, y in zip(got[1:], want[1:])))
    log(f"[bench] ray_test: {n} rays in {1e3 * secs:.2f} ms; oracle {sample} rays in {cpu:.2f}s; agree {same}")
    return {"rays": n, "ms_per_call": round(1e3 * secs, 3), "rays_per_s": round(n / secs, 1),
            "hits_sample": int((want[0] >= 0).sum()), "sample": sample, "sample_matches_oracle": same,
            "cpu_rays_per_s": round(sample / cpu, 1), "cpu_cores": 1,
            "workload": "closest-hit rays between random points of the 104 x 41 x 104 m volume over the 100k pile "
                        "(half vertical), after the timed region; query tree rebuilt in every call"}


def cpu_baseline(shapes, verts, bodies, snap, args, first_timed, residual):
    """The oracle (C++ restatement of the Java path, 1 core) stepping the same pile from the state at the start of
    the timed window, in the reference's island order with the Random(0) shuffle: the first steps of the same
    window, as many as fit in --cpu-seconds.  The first step's final residual is the reference order's residual
    for the window's first step (the device's colour-order residual from the same state is beside it)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.lib()
    orc = pyoracle.World(shapes, verts, bodies, iterations=args.iterations, order=0)
    orc.set_bodies(snap["bodies"])
    orc.set_proxies(*snap["proxies"])
    orc.set_pairs(*snap["pairs"])
    n = 0
    secs = 0.0
    res = []
    while secs < args.cpu_seconds and n < args.steps:
        secs += orc.time_steps(1)
        res.append(orc.stats()["step_residual"])
        n += 1
        log(f"[bench] cpu baseline step {first_timed + n - 1}: {secs:.1f}s, residual {res[-1]:.3e}")
    residual["reference_order"] = float(res[0])
    residual["reference_order_per_step"] = [float(r) for r in res]
    v = args.bodies * n / secs
    return {"value": round(v, 1), "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": f"steps {first_timed}-{first_timed + n - 1} (the first {n} of the timed window) of the same "
                      f"{args.bodies}-body pile from the device state at the window's start; single-threaded C++ "
                      f"restatement of the Java path (reference island order, Random(0) shuffle), host "
                      f"{os.cpu_count()} cores, 1 used"}


def summarize_trace(d, steps):
    """Per-kernel summary of the dispatches between the two window markers (kCopyProbe of TRACE_MARK_BYTES,
    launched by --trace-window right before and after the timed steps) in a rocprofv3 kernel trace (csv)."""
    import csv
    files = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("kernel_trace.csv")]
    if not files:
        raise SystemExit(f"no kernel_trace.csv under {d}")
    rows = []
    for f in files:
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    # each marker is one jbgpu_measure_copy call of TRACE_MARK_BYTES: a warm-up and a timed kCopyProbe dispatch;
    # the window lies between the second dispatch of the first marker and the first of the second (the
    # copy-ceiling probe after the window copies 1 GiB, far longer than a marker)
    marks = [i for i, r in enumerate(rows) if "kCopyProbe" in r["Kernel_Name"]
             and int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) < 200_000]
    if len(marks) < 4:
        raise SystemExit("window markers not found")
    a, b = marks[1], marks[2]
    win = [r for r in rows[a + 1:b] if not r["Kernel_Name"].startswith("__amd_rocclr")]
    span_ms = (int(rows[b]["Start_Timestamp"]) - int(rows[a]["End_Timestamp"])) / 1e6
    agg = {}
    import re
    for r in win:
        name = re.sub(r"^void ", "", r["Kernel_Name"].replace("(anonymous namespace)::", "")).split("(")[0]
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3  # us
        e = agg.setdefault(name, [0, 0.0])
        e[0] += 1
        e[1] += dur
    total = sum(v[1] for v in agg.values())
    print(f"window: {len(win)} dispatches over {steps} steps, {span_ms:.3f} ms between the markers "
          f"({span_ms / steps:.3f} ms/step wall), kernel time {total / 1e3:.3f} ms ({total / 1e3 / steps:.3f} ms/step)")
    print(f"{'kernel':64s} {'calls':>7s} {'calls/step':>10s} {'total_ms':>9s} {'ms/step':>8s} {'avg_us':>8s} {'%':>6s}")
    for name, (n, us) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"{name[:64]:64s} {n:7d} {n / steps:10.1f} {us / 1e3:9.3f} {us / 1e3 / steps:8.3f} {us / n:8.2f} "
              f"{100 * us / total:6.2f}")


if __name__ == "__main__":
    main()
