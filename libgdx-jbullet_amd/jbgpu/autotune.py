"""Launch-shape autotuning on the device (results are bit-identical for every candidate; only time changes).

``world_blocks`` picks the world -> workgroup mapping of the batched solve: how many dynamic bodies of whole
worlds one workgroup takes (worlds packed in world order), the workgroup size and whether the sweeps load the next
colour's rows ahead (more registers, fewer resident waves), against the per-colour-launch solve
(``bodies_per_block`` 0).  Each candidate runs one untimed step (first launch, block tables) and then
``steps`` timed steps of the caller's scene, which advance the scene as any other steps would.
"""
import time

WORLD_BLOCK_CANDIDATES = ((0, 256, 0), (128, 128, 0), (256, 128, 0), (256, 256, 0), (512, 128, 0), (512, 256, 0),
                          (1024, 256, 0), (1024, 512, 0), (128, 128, 1), (256, 256, 1), (512, 256, 1), (1024, 512, 1))


def world_blocks(ctx, dt=1.0 / 60.0, steps=4, candidates=WORLD_BLOCK_CANDIDATES, log=None):
    """Time every (bodies_per_block, threads, pipelined) candidate on the context's scene, keep the fastest."""
    ms = {}
    for cand in candidates:
        ctx.solver_set_world_blocks(*cand)
        ctx.step(dt)
        t = time.perf_counter()
        ctx.step(dt, n=steps)
        ms[cand] = 1e3 * (time.perf_counter() - t) / steps
        if log:
            log(f"[autotune] world blocks {cand[0]} bodies / {cand[1]} threads / pipelined {cand[2]}: "
                f"{ms[cand]:.3f} ms/step")
    best = min(ms, key=ms.get)
    ctx.solver_set_world_blocks(*best)
    return {"bodies_per_block": best[0], "threads": best[1], "pipelined": bool(best[2]),
            "ms_per_step": {f"{b}/{t}/{'p' if p else 'n'}": round(v, 3) for (b, t, p), v in ms.items()},
            "steps": len(candidates) * (steps + 1)}
