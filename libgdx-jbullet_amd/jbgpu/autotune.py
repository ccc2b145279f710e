"""Launch-shape autotuning on the device (results are bit-identical for every candidate; only time changes).

``world_blocks`` picks the world -> workgroup mapping of the batched solve: how many dynamic bodies of whole
worlds one workgroup takes (worlds packed in world order) and the workgroup size, against the per-colour-launch
solve (``bodies_per_block`` 0).  Each candidate runs one untimed step (first launch, block tables) and then
``steps`` timed steps of the caller's scene, which advance the scene as any other steps would.
"""
import time

WORLD_BLOCK_CANDIDATES = ((0, 256), (64, 64), (128, 128), (256, 128), (256, 256), (512, 256), (512, 512),
                          (1024, 512))


def world_blocks(ctx, dt=1.0 / 60.0, steps=4, candidates=WORLD_BLOCK_CANDIDATES, log=None):
    """Time every (bodies_per_block, threads) candidate on the context's scene and keep the fastest."""
    ms = {}
    for bodies, threads in candidates:
        ctx.solver_set_world_blocks(bodies, threads)
        ctx.step(dt)
        t = time.perf_counter()
        ctx.step(dt, n=steps)
        ms[(bodies, threads)] = 1e3 * (time.perf_counter() - t) / steps
        if log:
            log(f"[autotune] world blocks {bodies} bodies / {threads} threads: {ms[(bodies, threads)]:.3f} ms/step")
    best = min(ms, key=ms.get)
    ctx.solver_set_world_blocks(*best)
    return {"bodies_per_block": best[0], "threads": best[1],
            "ms_per_step": {f"{b}/{t}": round(v, 3) for (b, t), v in ms.items()},
            "steps": len(candidates) * (steps + 1)}
