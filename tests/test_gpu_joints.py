"""GPU parity of Point2Point constraints (dynamics/constraintsolver/Point2PointConstraint.java) through the C ABI.

Every step starts the oracle from the device's state (bodies, proxies, pair cache) and feeds it the solver's
manifold order and the constraint order the device used (jbgpu_world_joint_order); bodies, pair records and every
constraint's appliedImpulse after the step must agree bit for bit.
"""
import numpy as np
import pytest

import jbgpu
import pyoracle
from jbgpu import scenes

pytestmark = pytest.mark.gpu


def _records(sc, vel_scale):
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    bodies = np.asarray(bodies, np.float32).reshape(-1, jbgpu.BODY_WORDS).copy()
    bodies[:, 12:15] = sc.meta["velocity"] * np.float32(vel_scale)
    return shapes, verts, bodies


def _pair(sc, vel_scale=1.0, deactivation=False, iterations=10):
    shapes, verts, bodies = _records(sc, vel_scale)
    ctx = jbgpu.Context(num_iterations=iterations, deactivation_disabled=int(not deactivation))
    ctx.world_create(shapes, verts, bodies)
    orc = pyoracle.World(shapes, verts, bodies, iterations=iterations, order=jbgpu.ORDER_EXPLICIT,
                         deactivation_disabled=not deactivation)
    a, b, pa, pb = sc.meta["joints"]
    ctx.add_point2point(a, b, pa, pb)
    orc.add_point2point(a, b, pa, pb)
    return ctx, orc


def _step_and_compare(ctx, orc, tag):
    b0 = ctx.bodies()
    prox, s0 = ctx.bp_proxies()
    k0, m0 = ctx.pairs()
    ctx.step()
    order, _ = ctx.solver_order()
    jorder = ctx.joint_order()
    orc.set_bodies(b0)
    orc.set_proxies(prox, s0)
    orc.set_pairs(k0, m0)
    orc.set_order(order)
    orc.set_joint_order(jorder)
    orc.step()
    bg, bo = ctx.bodies(), orc.bodies()
    if not np.array_equal(bg.view(np.uint32), bo.view(np.uint32)):
        bad = np.nonzero((bg.view(np.uint32) != bo.view(np.uint32)).any(1))[0]
        d = np.abs(bg.astype(np.float64) - bo) / np.maximum(1.0, np.abs(bo))
        raise AssertionError(f"{tag}: {len(bad)} bodies differ (first {bad[0]}), max rel {d.max():.3e}")
    kg, mg = ctx.pairs()
    ko, mo = orc.pairs()
    assert np.array_equal(kg, ko), f"{tag}: pair keys differ"
    ng = mg[:, 0].view(np.int32)
    for m in (mg, mo):  # dead points and uninitialised friction directions carry no meaning
        pts = m[:, 4:].reshape(len(m), 4, 32)
        for j in range(4):
            pts[j >= ng, j, :] = 0
            pts[pts[:, j, 27].view(np.int32) == 0, j, 21:27] = 0
    assert np.array_equal(mg.view(np.uint32), mo.view(np.uint32)), f"{tag}: manifolds differ"
    ig, io = ctx.joint_impulses(), orc.joint_impulses()
    if not np.array_equal(ig.view(np.uint32), io.view(np.uint32)):
        bad = np.nonzero(ig.view(np.uint32) != io.view(np.uint32))[0]
        raise AssertionError(f"{tag}: {len(bad)} constraint impulses differ (first {bad[0]}: {ig[bad[0]]} vs "
                             f"{io[bad[0]]})")
    return jorder


@pytest.mark.parametrize("mode", ["colours", "blocks", "world_blocks", "tolerance"])
def test_chains_step_parity(mode):
    """64 chains of 8 links (every other one on the floor), 200 boxes dropped onto them; 150 steps.  'blocks':
    block-partitioned contact sweeps; 'world_blocks': the one-launch batched-worlds solve is asked for, and the
    step falls back to the colour sweeps because the world has constraints; 'tolerance': iterate to a residual
    of 1e-3 (at most 40 iterations), the constraints stopping with the contact rows."""
    sc = scenes.chains(seed=3)
    ctx, orc = _pair(sc)
    if mode == "blocks":
        ctx.solver_set_blocks(128)
    elif mode == "world_blocks":
        ctx.solver_set_world_blocks(256)
    elif mode == "tolerance":
        ctx.solver_set_params(tolerance=1e-3, max_iterations=40, num_iterations=10)
        orc.set_solver(jbgpu.ORDER_EXPLICIT, 1e-3, 40, 10)
    nj = ctx.num_joints()
    assert nj == 512
    for step in range(150):
        jorder = _step_and_compare(ctx, orc, f"{mode} step {step}")
        assert len(jorder) == nj and np.array_equal(np.sort(jorder), np.arange(nj))
        if step % 25 == 0:
            st = ctx.stats()
            print(f"[chains {mode}] step {step}: manifolds {st['manifolds']} colours {st['colors']} "
                  f"max |joint impulse| {np.abs(ctx.joint_impulses()).max():.3f}", flush=True)
    ctx.close()


def test_chains_islands_sleep_parity():
    """Deactivation on: constraints join islands, so a chain sleeps only as a whole and its constraints stop
    being solved while it sleeps (their appliedImpulse keeps the last solve's value); slow swings let the free
    chains fall asleep within the run, the dropped boxes wake some."""
    sc = scenes.chains(seed=4, nx=6, nz=6, links=6, drops=12)
    ctx, orc = _pair(sc, vel_scale=0.3, deactivation=True)
    nj = ctx.num_joints()
    skipped = asleep = 0
    for step in range(300):
        jorder = _step_and_compare(ctx, orc, f"sleep step {step}")
        skipped = max(skipped, nj - len(jorder))
        state = ctx.bodies()[:, 29].view(np.int32)
        asleep = max(asleep, int((state == 2).sum()))
        if step % 50 == 0:
            print(f"[chains sleep] step {step}: solved constraints {len(jorder)}/{nj}, asleep {int((state == 2).sum())}",
                  flush=True)
    assert asleep > 0 and skipped > 0, (asleep, skipped)
    ctx.close()


def test_constraints_without_contacts():
    """A world without pairs (no floor): the solve runs the constraints alone (solveGroup without manifolds);
    a free chain of 4 links hung from a fixed point, and a chain joined to a static anchor body."""
    b = scenes._Builder("free_chains")
    link = b.shape(scenes.BOX, [0.15, 0.2, 0.15])
    vel = []
    for m in range(4):  # links 0.7 m apart: 0.3 m gaps, no AABB overlap at the start
        b.body(link, scenes.IDENT, np.array([0.0, 10.0 - 0.7 * m, 0.0], np.float32), 1.0)
        vel.append([0.5 * (m + 1), 0.0, 0.2])
    anchor = b.shape(scenes.BOX, [0.1, 0.1, 0.1])
    b.body(anchor, scenes.IDENT, np.array([5.0, 10.0, 0.0], np.float32), 0.0)
    b.body(link, scenes.IDENT, np.array([5.0, 9.5, 0.0], np.float32), 2.0)
    vel += [[0, 0, 0], [0.0, 0.0, -1.0]]
    joints = (np.array([0, 0, 1, 2, 4], np.int32), np.array([-1, 1, 2, 3, 5], np.int32),
              np.array([[0, 0.35, 0], [0, -0.35, 0], [0, -0.35, 0], [0, -0.35, 0], [0, -0.2, 0]], np.float32),
              np.array([[0, 10.35, 0], [0, 0.35, 0], [0, 0.35, 0], [0, 0.35, 0], [0, 0.3, 0]], np.float32))
    sc = b.done(joints=joints, velocity=np.array(vel, np.float32))
    ctx, orc = _pair(sc)
    for step in range(120):
        jorder = _step_and_compare(ctx, orc, f"free step {step}")
        assert len(jorder) == 5
        if step == 0:
            assert ctx.stats()["pairs"] == 0  # no broadphase pair at all on the first step
    ctx.close()


def test_add_point2point_argument_errors():
    """Bodies out of range, body A equal to body B, or a short pivot array are refused and change nothing."""
    sc = scenes.chains(seed=5, nx=2, nz=2, links=3, drops=0)
    ctx, _ = _pair(sc)
    nj = ctx.num_joints()
    nb = ctx.nb
    for a, b in ((nb, -1), (-1, 1), (1, nb), (1, -2), (2, 2)):
        with pytest.raises(jbgpu.JbgpuError):
            ctx.add_point2point([1, a], [-1, b], np.zeros((2, 3)), np.zeros((2, 3)))
        assert ctx.num_joints() == nj
    with pytest.raises(ValueError):
        ctx.add_point2point([1], [2], np.zeros((1, 3)), np.zeros((2, 3)))
    ctx.step()
    ctx.add_point2point([1], [2], [[0, 0.1, 0]], [[0, -0.1, 0]])  # added after steps: appliedImpulse starts at 0
    assert ctx.num_joints() == nj + 1 and ctx.joint_impulses()[-1] == 0.0
    ctx.step()
    assert len(ctx.joint_order()) == nj + 1
    ctx.close()


def _hinge_rows(sc, bodies):
    """The scene's hinges as (body_a, body_b, frame_a, frame_b, params) arrays; frames from the constructors'
    host helper on body A's initial transform."""
    A, B, FA, FB, P = [], [], [], [], []
    free = [1e30, -1e30, 0.9, 0.3, 1.0, 0.0, 0.0, 0.0, 0.0]
    for (a, b, pa, xa, pb, xb, F, params) in sc.meta["hinges"]:
        tr = bodies[a, :12]
        if F is not None:
            fa, fb = jbgpu.hinge_frames(tr, frame_a=F)
        elif pb is None:
            fa, fb = jbgpu.hinge_frames(tr, pa, xa)
        else:
            fa, fb = jbgpu.hinge_frames(tr, pa, xa, pb, xb)
        A.append(a), B.append(b), FA.append(fa), FB.append(fb), P.append(free if params is None else params)
    return (np.array(A, np.int32), np.array(B, np.int32), np.array(FA), np.array(FB), np.array(P, np.float32))


@pytest.mark.parametrize("deactivation", [False, True])
def test_hinges_step_parity(deactivation):
    """Hinged chains, doors (limits, motors, both constructor forms), motor wheels on static posts (some
    angular-only) and falling boxes, with Point2Point constraints joining some of the falling boxes in pairs
    (both constraint types in one colouring); 180 steps bit-exact, constraint impulses included."""
    sc = scenes.hinges(seed=2)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    bodies = np.asarray(bodies, np.float32).reshape(-1, jbgpu.BODY_WORDS).copy()
    bodies[:, 12:15] = sc.meta["velocity"]
    bodies[:, 15:18] = sc.meta["spin"]
    ctx = jbgpu.Context(num_iterations=10, deactivation_disabled=int(not deactivation))
    ctx.world_create(shapes, verts, bodies)
    orc = pyoracle.World(shapes, verts, bodies, iterations=10, order=jbgpu.ORDER_EXPLICIT,
                         deactivation_disabled=not deactivation)
    a, b, fa, fb, params = _hinge_rows(sc, bodies)
    ctx.add_hinge(a, b, fa, fb, params)
    orc.add_hinge(a, b, fa, fb, params)
    drops = np.arange(len(bodies) - 40, len(bodies), 2, dtype=np.int32)
    pa = np.tile(np.float32([0.3, 0.0, 0.0]), (len(drops), 1))
    ctx.add_point2point(drops, drops + 1, pa, -pa)
    orc.add_point2point(drops, drops + 1, pa, -pa)
    nj = ctx.num_joints()
    assert nj == len(a) + len(drops)
    for step in range(180):
        jorder = _step_and_compare(ctx, orc, f"hinges step {step}")
        if not deactivation:
            assert len(jorder) == nj
        if step % 30 == 0:
            st = ctx.stats()
            print(f"[hinges deact={deactivation}] step {step}: manifolds {st['manifolds']} solved constraints "
                  f"{len(jorder)}/{nj} max |impulse| {np.abs(ctx.joint_impulses()).max():.3f}", flush=True)
    ctx.close()
