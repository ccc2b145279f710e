"""GPU tests of graph-captured steps (jbgpu_world_set_step_graph): every step replayed from a captured hipGraph must
leave the world bit-identical to the ordinary step, over many replays with the program's own reads and
synchronisations between them, across a pair-capacity overflow (the step is finished on the ordinary path) and
against the oracle."""
import numpy as np
import pytest

import jbgpu
import pyoracle
from jbgpu import scenes

pytestmark = pytest.mark.gpu

SPHERE, BOX, CAPSULE, HULL, PLANE = 0, 1, 2, 3, 4


def _contexts(sc, blocks=(512, 256, 0), iterations=10):
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    if "velocity" in sc.meta:
        bodies[:, 12:15] = sc.meta["velocity"]
    out = []
    for graph in (False, True):
        ctx = jbgpu.Context(num_iterations=iterations)
        ctx.world_create(shapes, verts, bodies, world_ids=sc.world)
        ctx.solver_set_world_blocks(*blocks)
        if graph:
            ctx.set_step_graph(True)
        out.append(ctx)
    return out[0], out[1], (shapes, verts, bodies)


def _live_manifolds(m):
    """Manifold records with the words no reader uses cleared: points past the count, friction directions of
    points whose directions were never set."""
    m = m.copy()
    if len(m) == 0:
        return m
    pts = m[:, 4:].reshape(len(m), 4, 32)
    npts = m[:, 0].view(np.int32)
    for j in range(4):
        pts[j >= npts, j, :] = 0
        uninit = pts[:, j, 27].view(np.int32) == 0
        pts[uninit, j, 21:27] = 0
    m[:, 4:] = pts.reshape(len(m), -1)
    return m


def _same(a, b, tag):
    ba, bb = a.bodies(), b.bodies()
    if not np.array_equal(ba.view(np.uint32), bb.view(np.uint32)):
        bad = np.nonzero((ba.view(np.uint32) != bb.view(np.uint32)).any(1))[0]
        raise AssertionError(f"{tag}: {len(bad)} bodies differ (first {bad[0]})")
    ka, ma = a.pairs()
    kb, mb = b.pairs()
    assert np.array_equal(ka, kb), f"{tag}: pair keys differ ({len(ka)} vs {len(kb)})"
    la, lb = _live_manifolds(ma).view(np.uint32), _live_manifolds(mb).view(np.uint32)
    if not np.array_equal(la, lb):
        rows = np.nonzero((la != lb).any(1))[0]
        for r in rows[:4]:
            w = np.nonzero(la[r] != lb[r])[0]
            print(f"[{tag}] pair {r} key {int(ka[r]):#x} points {la[r, 0]}/{lb[r, 0]} words {w.tolist()[:12]} "
                  f"a {ma[r, w[:6]].tolist()} b {mb[r, w[:6]].tolist()}", flush=True)
        raise AssertionError(f"{tag}: {len(rows)} of {len(la)} manifolds differ")
    sa, sb = a.stats(), b.stats()
    for k in ("pairs", "manifolds", "iterations", "colors", "huge_proxies", "rows", "epa_calls"):
        assert sa[k] == sb[k], (tag, k, sa[k], sb[k])


@pytest.mark.parametrize("blocks", [(512, 256, 0), (16, 64, 1)])
def test_graph_steps_match_ordinary_steps(blocks):
    """24 batched worlds (towers of spheres and boxes): 60 steps in calls of 1, 5, 10, 20 and 24 steps (each call
    replays its graphs back to back), bodies, pairs, manifolds and statistics compared after every call."""
    sc = scenes.batched_towers(seed=3, n_worlds=24, lo=16, hi=120)
    a, b, _ = _contexts(sc, blocks)
    for n in (1, 5, 10, 20, 24):
        a.step(n=n)
        b.step(n=n)
        _same(a, b, f"blocks {blocks}, after a call of {n}")
        print(f"[graph {blocks}] call of {n}: pairs {b.stats()['pairs']} manifolds {b.stats()['manifolds']}",
              flush=True)
    assert b.stats()["graph_redos"] == 0
    a.close()
    b.close()


def _sphere_grid(n_worlds=64, side=10, spacing=1.5):
    """Per world a floor plane and side x side spheres of radius 0.5 resting on it, `spacing` apart."""
    bl = scenes._Builder("sphere grid")
    plane = bl.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    sph = bl.shape(SPHERE, [0.5])
    for w in range(n_worlds):
        bl.body(plane, scenes.IDENT, np.zeros(3, np.float32), 0.0, world=w)
        for i in range(side):
            for k in range(side):
                bl.body(sph, scenes.IDENT, np.array([i * spacing, 0.5, k * spacing], np.float32), 1.0, world=w)
    return bl.done()


def test_graph_step_pair_overflow_is_finished_on_the_ordinary_path():
    """Spheres 1.5 m apart on their floors (one pair per sphere, with the floor), then moved 0.9 m apart: the
    next graph-captured step finds five times the pairs it was captured for, gives up after its AABB update, is
    finished on the ordinary path and the graphs are captured again at the larger capacity.  Bit-identical to the
    ordinary steps throughout."""
    sc = _sphere_grid()
    a, b, (shapes, verts, bodies) = _contexts(sc)
    a.step(n=3)
    b.step(n=3)
    _same(a, b, "spread out")
    packed = a.bodies().copy()
    dyn = sc.mass > 0
    packed[dyn, 9] *= np.float32(0.6)
    packed[dyn, 11] *= np.float32(0.6)
    a.set_bodies(packed)
    b.set_bodies(packed)
    a.step(n=6)
    b.step(n=6)
    _same(a, b, "packed")
    print(f"[graph overflow] pairs {b.stats()['pairs']}, redone steps {b.stats()['graph_redos']}", flush=True)
    assert b.stats()["graph_redos"] >= 1 and a.stats()["graph_redos"] == 0
    b.step(n=10)
    a.step(n=10)
    _same(a, b, "after the redo")
    a.close()
    b.close()


def test_graph_steps_match_oracle():
    """Single graph-captured steps from the device state, checked against the oracle fed the block solve's
    manifold order (the captured solve reports its order like the ordinary one)."""
    sc = scenes.batched_towers(seed=4, n_worlds=12, lo=16, hi=100)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=10)
    ctx.world_create(shapes, verts, bodies, world_ids=sc.world)
    ctx.solver_set_world_blocks(512, 256, 0)
    ctx.set_step_graph(True)
    orc = pyoracle.World(shapes, verts, bodies, iterations=10, order=jbgpu.ORDER_EXPLICIT)
    orc.set_worlds(sc.world)
    for step in range(40):
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
        bg, bo = ctx.bodies(), orc.bodies()
        assert np.array_equal(bg.view(np.uint32), bo.view(np.uint32)), f"step {step}: bodies differ"
        kg, _ = ctx.pairs()
        ko, _ = orc.pairs()
        assert np.array_equal(kg, ko), f"step {step}: pair keys differ"
    ctx.close()


def test_graph_steps_with_ccd():
    """Projectiles with continuous collision detection through graph-captured steps: bit-identical to the
    ordinary steps (the clamp pass reads the pair count of the step from the device)."""
    sc = scenes.projectiles(seed=3)
    a, b, _ = _contexts(sc)
    ids, thr, rad = sc.meta["ccd"]
    for ctx in (a, b):
        ctx.set_ccd(ids, thr, rad)
    for n in (1, 4, 15, 40):
        a.step(n=n)
        b.step(n=n)
        _same(a, b, f"ccd, after a call of {n}")
    a.close()
    b.close()


def test_graph_step_requirements():
    """Graph-captured steps need the world-block solve and no constraints; what is missing is named."""
    sc = scenes.batched_towers(seed=5, n_worlds=4, lo=16, hi=40)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=10)
    ctx.world_create(shapes, verts, bodies, world_ids=sc.world)
    with pytest.raises(jbgpu.JbgpuError, match="world-block"):
        ctx.set_step_graph(True)
    ctx.solver_set_world_blocks(512, 256, 0)
    ctx.set_step_graph(True)
    ctx.step(n=3)
    dyn = np.nonzero(sc.mass > 0)[0]
    ctx.add_point2point([int(dyn[0])], [int(dyn[1])], [[0, 0.5, 0]], [[0, -0.5, 0]])
    with pytest.raises(jbgpu.JbgpuError, match="constraints"):
        ctx.step()
    ctx.close()


def test_graph_4096_worlds_match_ordinary_steps():
    """The batched-worlds configuration at its full size (4,096 worlds, ~588k bodies): 40 steps in calls of 1, 9
    and 30, graph-captured against ordinary, bit-identical."""
    sc = scenes.batched_towers(seed=1, n_worlds=4096)
    print(f"[graph 4096] {sc.n_bodies} bodies", flush=True)
    a, b, _ = _contexts(sc)
    for n in (1, 9, 30):
        a.step(n=n)
        b.step(n=n)
        _same(a, b, f"4096 worlds, after a call of {n}")
        print(f"[graph 4096] call of {n}: pairs {b.stats()['pairs']}", flush=True)
    a.close()
    b.close()


def test_graph_steps_without_pairs():
    """A world with no pair at all (spheres 3 m apart, no floor, falling together): every count of the captured
    step is zero (no tree leaves overlap, no manifolds, no rows); bit-identical to the ordinary steps."""
    bl = scenes._Builder("apart")
    sph = bl.shape(SPHERE, [0.5])
    for w in range(4):
        for i in range(25):
            bl.body(sph, scenes.IDENT, np.array([3.0 * (i % 5), 3.0 * (i // 5), 0.0], np.float32), 1.0, world=w)
    sc = bl.done()
    a, b, _ = _contexts(sc)
    for n in (1, 3, 10):
        a.step(n=n)
        b.step(n=n)
        _same(a, b, f"no pairs, after a call of {n}")
        assert b.stats()["pairs"] == 0
    a.close()
    b.close()
