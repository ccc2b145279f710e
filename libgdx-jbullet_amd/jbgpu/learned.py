"""Learned warm start on the device: training data from converged solves, training, evaluation.

Training pairs come from the solver itself: for each contact slot of a solve run to a tight residual, the
16 features the solver extracts before the solve (normal, relative velocities, depth, inverse masses, cached
impulses, friction) and the converged (normal, lateral1, lateral2) impulses it ends with.  The model is then
trained on the device (``Context.mlp_train``) and evaluated by running the same scene to a residual with the
plain warm start (0.85 x cached impulses, as ``SequentialImpulseConstraintSolver``) and with the model.
"""
import time

import numpy as np

import jbgpu
from jbgpu import scenes


def valid_slots(ctx, order):
    """Slots (4 per active manifold, colour order) that carried a solver contact in the last solve."""
    _, man = ctx.pairs()
    mm = man[np.asarray(order, np.int64)]
    npts = mm[:, 0].view(np.int32)
    dist = mm[:, 4:].reshape(len(mm), 4, jbgpu.POINT_WORDS)[:, :, 15]
    return ((np.arange(4)[None, :] < npts[:, None]) & (dist <= 0)).reshape(-1)


def _world(sc, iterations):
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=iterations)
    ctx.world_create(shapes, verts, bodies, world_ids=sc.world)
    return ctx


def collect(seeds, scene_kw, settle=90, samples=6, every=5, tol=1e-6, cap=300, log=None):
    """Features and converged impulses of solves run to ``tol`` (cap ``cap`` iterations)."""
    xs, ys = [], []
    for seed in seeds:
        ctx = _world(scenes.pyramids(seed=seed, **scene_kw), 10)
        ctx.solver_set_params(tolerance=tol, max_iterations=cap, num_iterations=0)
        ctx.step(n=settle)
        ctx.solver_enable_features(True)
        for _ in range(samples):
            ctx.step(n=every)
            order, _ = ctx.solver_order()
            if len(order) == 0:
                continue
            v = valid_slots(ctx, order)
            xs.append(ctx.solver_features(4 * len(order))[v])
            ys.append(ctx.solver_impulses(4 * len(order))[v])
        ctx.close()
        if log:
            log(f"seed {seed}: {sum(len(x) for x in xs)} contacts")
    return np.concatenate(xs), np.concatenate(ys)


def evaluate(params, norm, seed, scene_kw, settle=90, steps=30, tol=1e-4, cap=100):
    """Solver iterations and time per step to ``tol``: plain warm start vs the model, from one settled state."""
    out = {}
    for mode in ("plain", "learned"):
        ctx = _world(scenes.pyramids(seed=seed, **scene_kw), 10)
        ctx.solver_set_params(tolerance=tol, max_iterations=cap, num_iterations=0)
        ctx.step(n=settle)
        if mode == "learned":
            ctx.mlp_set_params(params, norm)
            ctx.solver_use_mlp(True)
        ctx.timing(True)
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.step()
            its.append(ctx.stats()["iterations"])
        wall = time.perf_counter() - t0
        reg, nsteps = ctx.timing_get()
        solve_ms = sum(reg[k][0] for k in ("solver_setup", "solver_iter", "mlp")) / max(nsteps, 1)
        out[mode] = dict(iterations=float(np.mean(its)), solve_ms=solve_ms, step_ms=1e3 * wall / steps,
                         contacts=int(ctx.stats()["contacts"]))
        ctx.close()
    out["iteration_ratio"] = out["learned"]["iterations"] / max(out["plain"]["iterations"], 1e-9)
    out["time_ratio"] = out["learned"]["solve_ms"] / max(out["plain"]["solve_ms"], 1e-9)
    return out


def main():
    """Full protocol: pyramids of 64k boxes + 2000 spheres; training contacts from seeds 1000..; evaluation on
    seeds 1-3 to a 1e-4 residual (cap 100 iterations).  Prints one JSON line."""
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--boxes", type=int, default=64_000)
    ap.add_argument("--pyramids", type=int, default=32)
    ap.add_argument("--spheres", type=int, default=2000)
    ap.add_argument("--area", type=float, default=200.0)
    ap.add_argument("--train-seeds", type=int, default=8)
    ap.add_argument("--train-contacts", type=int, default=200_000)
    ap.add_argument("--settle", type=int, default=90)
    ap.add_argument("--eval-steps", type=int, default=30)
    ap.add_argument("--epochs", type=int, default=40)
    args = ap.parse_args()
    kw = dict(n_boxes=args.boxes, n_pyramids=args.pyramids, n_spheres=args.spheres, area=args.area)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t = time.time()
    x, y = collect(range(1000, 1000 + args.train_seeds), kw, settle=args.settle, samples=3, every=10, log=log)
    if len(x) > args.train_contacts:
        pick = np.random.default_rng(0).choice(len(x), args.train_contacts, replace=False)
        x, y = x[pick], y[pick]
    t_collect = time.time() - t
    ctx = jbgpu.Context()
    t = time.time()
    losses = ctx.mlp_train(x, y, epochs=args.epochs, batch=4096, lr=2e-3, seed=0)
    t_train = time.time() - t
    params, norm = ctx.mlp_get_params()
    ctx.close()
    log(f"trained on {len(x)} contacts: loss {losses[0]:.4f} -> {losses[-1]:.4f} in {t_train:.1f}s")
    evals = {s: evaluate(params, norm, s, kw, settle=args.settle, steps=args.eval_steps) for s in (1, 2, 3)}
    out = {"scene": kw, "training_contacts": int(len(x)), "collect_s": round(t_collect, 1),
           "train_s": round(t_train, 2), "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
           "eval": evals,
           "iteration_ratio_mean": float(np.mean([e["iteration_ratio"] for e in evals.values()])),
           "time_ratio_mean": float(np.mean([e["time_ratio"] for e in evals.values()]))}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
