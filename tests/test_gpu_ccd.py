"""GPU parity of the convex sweeps and of continuous collision detection, through the C ABI.

convexSweepTest (closest hit per query) against the oracle from the same body state: hit body, fraction,
normal and point bit for bit, for every convex cast shape type.  CCD: worlds with motion thresholds set are
stepped one step at a time from identical state (as the other world parity tests do) and must agree bit for
bit after every step; the motion clamp runs inside integrateTransforms, so any difference in the sweep, the
filters or the body order shows up in the body records.
"""
import numpy as np
import pytest

import jbgpu
import pyoracle
from jbgpu import scenes

pytestmark = pytest.mark.gpu

SPHERE, BOX, CAPSULE, HULL, PLANE = 0, 1, 2, 3, 4


def _world_pair(sc, iterations=10):
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    if "velocity" in sc.meta:
        bodies[:, 12:15] = sc.meta["velocity"]
    if "angular" in sc.meta:
        bodies[:, 15:18] = sc.meta["angular"]
    ctx = jbgpu.Context(num_iterations=iterations)
    ctx.world_create(shapes, verts, bodies, world_ids=sc.world)
    orc = pyoracle.World(shapes, verts, bodies, iterations=iterations, order=jbgpu.ORDER_EXPLICIT)
    if sc.world is not None:
        orc.set_worlds(sc.world)
    return ctx, orc, shapes


def _sync_and_step(ctx, orc):
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


def _compare_bodies(ctx, orc, tag):
    bg, bo = ctx.bodies(), orc.bodies()
    if not np.array_equal(bg.view(np.uint32), bo.view(np.uint32)):
        bad = np.nonzero((bg.view(np.uint32) != bo.view(np.uint32)).any(1))[0]
        d = np.abs(bg.astype(np.float64) - bo) / np.maximum(1.0, np.abs(bo))
        raise AssertionError(f"{tag}: {len(bad)} bodies differ (first {bad[0]}), max rel {d.max():.3e}")
    kg, _ = ctx.pairs()
    ko, _ = orc.pairs()
    assert np.array_equal(kg, ko), f"{tag}: pair keys differ"


def _rot(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    return np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                     2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                     2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1).astype(np.float32)


def _queries(rng, n, lo, hi):
    """From/to transforms: random points of the box [lo, hi], random orientations (different at both ends for
    two thirds of the queries, so the cast AABB carries an angular part), a third of them vertical drops."""
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    a = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    b = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    v = rng.random(n) < 0.33
    b[v, 0], b[v, 2] = a[v, 0], a[v, 2]
    a[v, 1], b[v, 1] = hi[1], lo[1]
    ra = _rot(rng, n)
    rb = np.where((rng.random(n) < 0.66)[:, None], _rot(rng, n), ra)
    return np.concatenate([ra, a], 1), np.concatenate([rb, b], 1)


def _one_of_each(shapes):
    """Index of the first shape of every convex type in the world's shape table."""
    types = shapes[:, 0].view(np.int32)
    return {t: int(np.nonzero(types == t)[0][0]) for t in (SPHERE, BOX, CAPSULE, HULL) if (types == t).any()}


def _check_sweeps(ctx, orc, shapes, k, frm, to, what, world=None, group=1, mask=-1):
    got = ctx.convex_sweep(shapes[k], frm, to, group=group, mask=mask, world=world)
    want = orc.convex_sweep(k, frm, to, group=group, mask=mask, world=world)
    hits = int((want[0] >= 0).sum())
    print(f"[sweep {what}] {len(frm)} queries, {hits} hits", flush=True)
    assert np.array_equal(got[0], want[0]), f"{what}: hit objects differ"
    for g, w_, name in zip(got[1:], want[1:], ("fraction", "normal", "point")):
        assert np.array_equal(g.view(np.uint32), w_.view(np.uint32)), f"{what}: {name} bits differ"
    return hits


@pytest.mark.parametrize("scene", ["mixed", "batched"])
def test_convex_sweep_matches_oracle(scene):
    """CollisionWorld.convexSweepTest (closest hit) of a sphere, a box, a capsule and a hull from the scene's
    shape table (the batched towers have spheres and boxes only), from the same body state as the oracle: every
    query's result bit for bit."""
    rng = np.random.default_rng(29)
    world = None
    if scene == "mixed":
        sc = scenes.mixed_pile(seed=4, n=3000, bin_half=15.0)
        lo, hi = (-17, -1, -17), (17, 30, 17)
    else:
        sc = scenes.batched_towers(seed=5, n_worlds=24)
        lo, hi = (-15, -1, -15), (15, 25, 15)
    ctx, orc, shapes = _world_pair(sc)
    ctx.step(n=40)
    orc.set_bodies(ctx.bodies())
    cast = _one_of_each(shapes)
    assert len(cast) == (4 if scene == "mixed" else 2), cast  # the towers are boxes and spheres
    for t, k in cast.items():
        frm, to = _queries(rng, 1500, lo, hi)
        if scene == "batched":
            world = rng.integers(0, int(sc.world.max()) + 1, len(frm)).astype(np.int32)
        hits = _check_sweeps(ctx, orc, shapes, k, frm, to, f"{scene} type {t}", world=world)
        assert hits > len(frm) // (4 if scene == "mixed" else 10)  # the towers fill less of the query box
    if scene == "mixed":  # the callback filter: static bodies (group 2) excluded
        frm, to = _queries(rng, 500, lo, hi)
        _check_sweeps(ctx, orc, shapes, cast[BOX], frm, to, "mixed, static excluded", mask=np.int16(-1 ^ 2))
    ctx.close()


def test_convex_sweep_100k_pile_matches_oracle():
    """convexSweepTest in the 100k-body pile after 120 steps: 200 box sweeps (the oracle tests every body)."""
    rng = np.random.default_rng(31)
    sc = scenes.mixed_pile(seed=1, n=100_000)
    ctx, orc, shapes = _world_pair(sc)
    ctx.step(n=120)
    orc.set_bodies(ctx.bodies())
    frm, to = _queries(rng, 200, (-52, -1, -52), (52, 40, 52))
    hits = _check_sweeps(ctx, orc, shapes, _one_of_each(shapes)[BOX], frm, to, "pile 100k")
    assert hits > 50
    ctx.close()


def test_ccd_projectiles_step_parity():
    """Projectiles at 60-300 m/s against thin walls, planks and stacks with CCD (threshold 0.05, swept radius
    0.08): 90 steps, bit-exact after each; and the clamp does its work: without CCD most projectiles in a wall's
    span pass through it, with CCD the spheres never do."""
    sc = scenes.projectiles(seed=1)
    ids, thr, rad = sc.meta["ccd"]
    zs = (np.arange(6) - 2.5) * 2.5
    runs = {}
    for ccd in (False, True):
        ctx, orc, shapes = _world_pair(sc)
        if ccd:
            ctx.set_ccd(ids, thr, rad)
            orc.set_ccd(ids, thr, rad)
        tunnelled = set()
        for step in range(90):
            prev = ctx.bodies()[ids, 9:12].copy()
            _sync_and_step(ctx, orc)
            _compare_bodies(ctx, orc, f"projectiles ccd={ccd} step {step}")
            cur = ctx.bodies()[ids, 9:12]
            for k in range(len(ids)):
                if prev[k, 0] < 7.95 and cur[k, 0] > 8.05:
                    t = (8.0 - prev[k, 0]) / (cur[k, 0] - prev[k, 0])
                    p = prev[k] + t * (cur[k] - prev[k])
                    if np.min(np.abs(p[2] - zs)) < 1.0 and p[1] < 4.0:
                        tunnelled.add(int(ids[k]))
        types = shapes[ctx.bodies()[ids, 27].view(np.int32), 0].view(np.int32)
        spheres = {int(i) for i, t in zip(ids, types) if t == SPHERE}
        runs[ccd] = (len(tunnelled), len(tunnelled & spheres))
        print(f"[ccd projectiles] ccd={ccd}: {runs[ccd][0]} through a static wall ({runs[ccd][1]} spheres)", flush=True)
        ctx.close()
    assert runs[False][0] >= 20 and runs[False][1] > 0
    assert runs[True][1] == 0 and runs[True][0] < runs[False][0] // 2


def test_ccd_thin_wall():
    """The oracle's thin-wall case on the device: a sphere of radius 0.1 at 300 m/s stops in front of a static
    wall 0.1 m thick instead of passing it, bit-exact with the oracle over 30 steps."""
    b = scenes._Builder("wall")
    wall = b.shape(BOX, [0.05, 2.0, 2.0])
    b.body(wall, scenes.IDENT, np.array([2.5, 0, 0], np.float32), 0.0)
    s = b.shape(SPHERE, [0.1])
    b.body(s, scenes.IDENT, np.array([0, 0, 0], np.float32), 1.0)
    sc = b.done(velocity=np.array([[0, 0, 0], [300, 0, 0]], np.float32))
    ctx, orc, _ = _world_pair(sc)
    ctx.set_ccd([1], 0.1, 0.1)
    orc.set_ccd([1], 0.1, 0.1)
    for step in range(30):
        _sync_and_step(ctx, orc)
        _compare_bodies(ctx, orc, f"thin wall step {step}")
        assert ctx.bodies()[1, 9] < 2.45
    ctx.close()


@pytest.mark.parametrize("scene", ["mixed", "batched"])
def test_ccd_every_body_step_parity(scene):
    """CCD on every dynamic body with a low threshold (a falling pile: hundreds of bodies clamp-tested per
    step, each sweeping against the bodies before it at their new transforms), bit-exact for 60 steps; in the
    batched worlds each body sweeps only its own world."""
    if scene == "mixed":
        sc = scenes.mixed_pile(seed=6, n=1500, bin_half=10.0)
    else:
        sc = scenes.batched_towers(seed=7, n_worlds=12, lo=16, hi=80)
    ctx, orc, shapes = _world_pair(sc)
    dyn = np.nonzero(sc.mass > 0)[0].astype(np.int32)
    ctx.set_ccd(dyn, 0.02, 0.05)
    orc.set_ccd(dyn, 0.02, 0.05)
    for step in range(60):
        _sync_and_step(ctx, orc)
        _compare_bodies(ctx, orc, f"ccd {scene} step {step}")
        if step % 20 == 0:
            print(f"[ccd {scene}] step {step}: pairs {ctx.stats()['pairs']}", flush=True)
    ctx.close()


def test_ccd_and_sweep_argument_errors():
    """Bad CCD settings are refused whole (nothing changes); the cast shape must be convex with its hull points in
    the world's table."""
    sc = scenes.projectiles(seed=2, lanes=2, shots=5)
    ctx, orc, shapes = _world_pair(sc)
    ids = sc.meta["ccd"][0]
    for bodies, thr, rad in (([0, ctx.nb], 0.1, 0.1), ([1], np.nan, 0.1), ([1], 0.1, -0.5), ([1], np.inf, 0.1)):
        with pytest.raises(jbgpu.JbgpuError):
            ctx.set_ccd(bodies, thr, rad)
    ctx.set_ccd(ids, 0.05, 0.08)
    orc.set_ccd(ids, 0.05, 0.08)
    for step in range(20):
        _sync_and_step(ctx, orc)
        _compare_bodies(ctx, orc, f"errors scene step {step}")
    frm = np.concatenate([scenes.IDENT, [0, 1, 0]]).astype(np.float32)[None]
    to = np.concatenate([scenes.IDENT, [20, 1, 0]]).astype(np.float32)[None]
    with pytest.raises(jbgpu.JbgpuError):
        ctx.convex_sweep(shapes[0], frm, to)  # the floor plane: not convex
    bad = shapes[_one_of_each(shapes)[HULL]].copy()
    bad[5:7] = np.array([0, 10_000], np.int32).view(np.float32)
    with pytest.raises(jbgpu.JbgpuError):
        ctx.convex_sweep(bad, frm, to)
    with pytest.raises(jbgpu.JbgpuError):
        ctx.convex_sweep(shapes[_one_of_each(shapes)[SPHERE]], frm, to, world=[5])
    ctx.close()
