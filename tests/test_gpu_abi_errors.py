"""Error paths of the C ABI on the device: a rejected call leaves the context as it was and the next valid
call succeeds.  Every condition is produced from the host (bad arguments); each kernel involved checks its own
bounds and exits normally."""
import numpy as np
import pytest

import jbgpu
import pyoracle
from jbgpu import scenes

pytestmark = pytest.mark.gpu


def _world(sc, iterations=10):
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=iterations)
    ctx.world_create(shapes, verts, bodies, world_ids=sc.world)
    return ctx, shapes, verts, bodies


def test_error_does_not_stick_to_later_steps():
    """A failed call (bad step arguments, malformed pair records) returns its status; the following valid
    steps return JBGPU_OK and match the oracle."""
    sc = scenes.boxes_grid(seed=4, nx=3, ny=3, nz=3, base=0.05)
    ctx, shapes, verts, bodies = _world(sc)
    ctx.step(n=20)
    with pytest.raises(jbgpu.JbgpuError, match=r"\(-1\)"):
        ctx.step(dt=-1.0)
    k, m = ctx.pairs()
    bad = m.copy()
    bad[0, 0] = np.int32(9).view(np.float32)
    with pytest.raises(jbgpu.JbgpuError, match="point count"):
        ctx.set_pairs(k, bad)
    k2, m2 = ctx.pairs()  # the rejected records never reached the device
    assert np.array_equal(k, k2) and np.array_equal(m.view(np.uint32), m2.view(np.uint32))
    ctx.set_pairs(k, m)  # well-formed records are taken
    orc = pyoracle.World(shapes, verts, bodies, iterations=10, order=jbgpu.ORDER_EXPLICIT)
    for step in range(5):
        b0 = ctx.bodies()
        prox, s0 = ctx.bp_proxies()
        k0, m0 = ctx.pairs()
        ctx.step()
        order, _ = ctx.solver_order()
        orc.set_bodies(b0)
        orc.set_proxies(prox, s0)
        orc.set_pairs(k0, m0)
        orc.set_order(order)
        orc.step()
        assert np.array_equal(ctx.bodies().view(np.uint32), orc.bodies().view(np.uint32)), f"step {step}"
    ctx.close()


def test_device_ray_world_ids_are_bounds_checked():
    """Rays resident in device memory with world ids outside the proxies' worlds: the call reports an argument
    error (the candidate kernel gives such rays no candidates and raises the flag); in-range ids work after."""
    sc = scenes.batched_towers(seed=5, n_worlds=6, lo=16, hi=40)
    ctx, shapes, verts, bodies = _world(sc)
    ctx.step(n=5)
    n = 64
    rng = np.random.default_rng(3)
    frm = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    to = frm.copy()
    frm[:, 1], to[:, 1] = 30.0, -1.0
    dev = {k: ctx.buffer_alloc(b) for k, b in (("f", frm.nbytes), ("t", to.nbytes), ("w", 4 * n), ("o", 4 * n),
                                                ("fr", 4 * n), ("n", 12 * n), ("p", 12 * n))}
    ctx.buffer_upload(dev["f"], frm)
    ctx.buffer_upload(dev["t"], to)
    call = lambda: ctx.ray_test_device(n, dev["f"], dev["t"], dev["o"], dev["fr"], dev["n"], dev["p"],
                                       world=dev["w"])
    maxw = int(sc.world.max())
    for bad in (maxw + 1, -1, 1 << 30):
        w = np.zeros(n, np.int32)
        w[n // 2] = bad
        ctx.buffer_upload(dev["w"], w)
        with pytest.raises(jbgpu.JbgpuError, match="world id out of range"):
            call()
    w = (np.arange(n) % (maxw + 1)).astype(np.int32)
    ctx.buffer_upload(dev["w"], w)
    call()
    got = ctx.ray_test(frm, to, world=w)
    assert (got[0] >= 0).sum() > n // 4
    for v in dev.values():
        ctx.buffer_free(v)
    ctx.close()


def test_failed_destroy_changes_nothing():
    """destroyProxy with one bad uid (out of range, or repeated in the call) is rejected as a whole: no proxy of
    the call is destroyed, so the same call without the bad uid succeeds and the pair set matches the oracle."""
    c, h, v = scenes.aabb_cloud(seed=9, n=4000, extent=200.0)
    boxes = np.concatenate([c - h, c + h], 1).astype(np.float32)
    ctx = jbgpu.Context()
    ctx.bp_create_proxies(boxes)
    ctx.bp_calculate_overlapping_pairs()
    before = ctx.bp_pairs()
    good = np.array([5, 17, 400], np.int32)
    for bad in (np.array([5, 17, 400, 99999], np.int32), np.array([5, 17, 5, 400], np.int32)):
        with pytest.raises(jbgpu.JbgpuError, match=r"\(-1\)"):
            ctx.bp_destroy_proxies(bad)
        assert np.array_equal(ctx.bp_pairs(), before)
    ctx.bp_destroy_proxies(good)
    orc = pyoracle.Broadphase(jbgpu.BP_DBVT)
    orc.create(boxes)
    orc.step()
    orc.destroy(good)
    boxes = scenes.aabb_cloud_step(c, h, v)
    ctx.bp_set_aabbs(boxes)
    orc.set_all(boxes)
    ctx.bp_calculate_overlapping_pairs()
    assert np.array_equal(ctx.bp_pairs(), orc.step())
    # the first pair query came with no setAabb: the stored boxes stayed and the stages rotated as in the oracle
    pg, sg = ctx.bp_proxies()
    po_ = orc.proxies()
    assert sg == pyoracle.lib().orc_bp_stage_current(orc.h)
    assert np.array_equal(pg[:, :13].view(np.uint32), po_[:, :13].view(np.uint32))
    ctx.close()


def test_aabb_overflow_disables_simulation_on_device():
    """A body whose AABB overflows (extent squared >= 1e12) is set to DISABLE_SIMULATION by the device's AABB
    kernel as by CollisionWorld.updateSingleAabb (CollisionWorld.java:211-227); every step bit-exact against
    the oracle fed the same order, and the disabled body never moves."""
    sc = scenes.aabb_overflow()
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=10)
    ctx.world_create(shapes, verts, bodies)
    orc = pyoracle.World(shapes, verts, bodies, iterations=10, order=jbgpu.ORDER_EXPLICIT)
    for step in range(30):
        b0 = ctx.bodies()
        prox, s0 = ctx.bp_proxies()
        k0, m0 = ctx.pairs()
        ctx.step()
        order, _ = ctx.solver_order()
        orc.set_bodies(b0)
        orc.set_proxies(prox, s0)
        orc.set_pairs(k0, m0)
        orc.set_order(order)
        orc.step()
        bg = ctx.bodies()
        assert np.array_equal(bg.view(np.uint32), orc.bodies().view(np.uint32)), f"step {step}"
        assert bg[2, 29].view(np.int32) == 5 and bg[1, 29].view(np.int32) == 1
    # the transform never changes (its velocity took one step of gravity before the AABB update disabled it)
    assert np.array_equal(bg[2, :12].view(np.uint32), bodies[2, :12].view(np.uint32))
    ctx.close()


def test_sweep_timing_stride():
    """Sampled sweep timing times the launches of every stride-th iteration: from the same state, 20 iterations
    with a stride of 5 time a fifth of the launches and rows a stride of 1 does (every colour its own launch);
    the step itself is the same; a stride below 1 is refused."""
    sc = scenes.pyramids(seed=7, n_boxes=800, n_pyramids=6, n_spheres=0, area=60.0)
    ctx, shapes, verts, bodies = _world(sc, iterations=20)
    ctx.solver_set_tail(0)
    ctx.step(n=5)
    snap = (ctx.bodies(), ctx.bp_proxies(), ctx.pairs())
    got, ends = {}, {}
    for stride in (1, 5):
        ctx.set_bodies(snap[0])
        ctx.bp_set_proxies(*snap[1])
        ctx.set_pairs(*snap[2])
        ctx.timing_kernels(1, stride=stride)
        ctx.step()
        got[stride] = ctx.timing_kernels_get()
        ctx.timing_kernels(0)
        ends[stride] = ctx.bodies()
    for kind in ("normal", "friction"):
        a, b = got[1][kind], got[5][kind]
        assert b["launches"] > 0 and a["launches"] == 5 * b["launches"], (kind, a, b)
        assert a["rows"] == 5 * b["rows"], (kind, a, b)
    assert np.array_equal(ends[1].view(np.uint32), ends[5].view(np.uint32))
    with pytest.raises(jbgpu.JbgpuError):
        ctx.timing_kernels(1, stride=0)
    ctx.close()
