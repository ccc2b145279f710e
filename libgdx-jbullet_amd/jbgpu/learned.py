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


def collect(seeds, scene_kw, settle=90, samples=6, every=5, tol=1e-12, cap=100, new_only=False, log=None):
    """Features and final impulses of solves run to ``tol`` with at most ``cap`` iterations (by default the
    evaluation's regime: every step runs the full ``cap`` iterations).  new_only: only the contacts that enter the
    solve without cached impulses (the ones the model warm-starts in mode 2)."""
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
            f = ctx.solver_features(4 * len(order))
            if new_only:
                v &= (f[:, 12] == 0) & (f[:, 13] == 0) & (f[:, 14] == 0)
            xs.append(f[v])
            ys.append(ctx.solver_impulses(4 * len(order))[v])
        ctx.close()
        if log:
            log(f"seed {seed}: {sum(len(x) for x in xs)} contacts")
    return np.concatenate(xs), np.concatenate(ys)


THRESHOLDS = (1e-2, 5e-3, 2e-3, 1e-3, 1e-4)


def _iterations_to(hist, tol, cap):
    below = np.nonzero(np.asarray(hist) <= tol)[0]
    return int(below[0]) + 1 if len(below) else cap


def evaluate(params, norm, seed, scene_kw, settle=90, steps=30, cap=100,
             modes=("plain", "learned", "learned_new", "factor_1")):
    """Every solve of `steps` steps runs `cap` iterations with the residual recorded after each one, from the same
    settled state for every mode: plain (0.85 x cached impulses, as the reference), learned (the model for every
    contact), learned_new (the model for the contacts without cached impulses, 0.85 x cached for the others),
    factor_1 (1.0 x cached impulses).  Reported per mode: mean iterations until the residual first drops to each
    threshold (cap when it never does), the residual after 20 iterations, and the solve time per iteration."""
    out = {}
    for mode in modes:
        ctx = _world(scenes.pyramids(seed=seed, **scene_kw), 10)
        ctx.solver_set_params(tolerance=1e-12, max_iterations=cap, num_iterations=0)
        ctx.step(n=settle)
        if mode in ("learned", "learned_new"):
            ctx.mlp_set_params(params, norm)
            ctx.solver_use_mlp(1 if mode == "learned" else 2)
        if mode == "factor_1":
            ctx.solver_set_warmstart_factor(1.0)
        ctx.timing(True)
        to = {t: [] for t in THRESHOLDS}
        r20 = []
        for _ in range(steps):
            ctx.step()
            h = ctx.residual_history()
            for t in THRESHOLDS:
                to[t].append(_iterations_to(h, t, cap))
            r20.append(float(h[min(19, len(h) - 1)]) if len(h) else 0.0)
        reg, nsteps = ctx.timing_get()
        it_ms = reg["solver_iter"][0] / max(nsteps, 1) / cap
        setup_ms = (reg["solver_setup"][0] + reg["mlp"][0]) / max(nsteps, 1)
        st = ctx.stats()
        out[mode] = dict(iterations_to={f"{t:g}": float(np.mean(v)) for t, v in to.items()},
                         residual_after_20=float(np.mean(r20)), ms_per_iteration=round(it_ms, 4),
                         setup_ms=round(setup_ms, 4), mlp_ms=round(reg["mlp"][0] / max(nsteps, 1), 4),
                         contacts=int(st["contacts"]), manifolds=int(st["manifolds"]))
        ctx.close()
    pl = out["plain"]["iterations_to"]
    for mode in modes:
        if mode.startswith("learned"):
            le = out[mode]["iterations_to"]
            out[f"iteration_ratio_{mode}"] = {t: le[t] / max(pl[t], 1e-9) for t in pl}
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
           }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
