"""GPU parity: libjbgpu (through the C ABI) against the oracle on the same inputs.

Broadphase pair sets must be bit-exact.  World steps are compared one step at a
time from identical state: the device state before step k is loaded into the
oracle, which is fed the solver's colour-batch manifold order and stepped; body
records, pair keys and manifold records after the step must agree bit for bit
(the declared tolerance, 1e-5 relative, is only used to report how far apart a
mismatch is).
"""
import numpy as np
import pytest

import jbgpu
import pyoracle
from jbgpu import scenes

pytestmark = pytest.mark.gpu

MODES = {"dbvt": jbgpu.BP_DBVT, "sap16": jbgpu.BP_SAP16, "sap32": jbgpu.BP_SAP32, "simple": jbgpu.BP_SIMPLE}
WMIN, WMAX = (-100.0, -100.0, -100.0), (1100.0, 1100.0, 1100.0)


def _cloud(n, seed, extent=1000.0 * 0.2):
    c, h, v = scenes.aabb_cloud(seed=seed, n=n, extent=extent)
    return c, h, v


@pytest.mark.parametrize("mode", list(MODES))
def test_broadphase_pairs_match_oracle(mode):
    n = 20000
    c, h, v = _cloud(n, seed=11)
    boxes = np.concatenate([c - h, c + h], 1).astype(np.float32)
    ctx = jbgpu.Context(bp_mode=MODES[mode], world_min=WMIN, world_max=WMAX)
    ctx.bp_create_proxies(boxes)
    orc = pyoracle.Broadphase(MODES[mode], np.array(WMIN, np.float32), np.array(WMAX, np.float32))
    orc.create(boxes)
    for step in range(6):
        boxes = scenes.aabb_cloud_step(c, h, v)
        ctx.bp_set_aabbs(boxes)
        orc.set_all(boxes)
        ctx.bp_calculate_overlapping_pairs()
        got = ctx.bp_pairs()
        want = orc.step()
        print(f"[bp {mode}] step {step}: {len(got)} pairs (oracle {len(want)})", flush=True)
        assert np.array_equal(got, want), f"{mode} step {step}"
        if mode == "dbvt":
            pg, sg = ctx.bp_proxies()
            po = orc.proxies()
            assert np.array_equal(pg[:, :13].view(np.uint32), po[:, :13].view(np.uint32))
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", list(MODES))
def test_broadphase_destroy_append_and_deltas(mode):
    """destroyProxy / later createProxy calls and the pair deltas: pair sets bit-exact against the oracle
    doing the same operations; deltas equal the set differences of consecutive pair sets."""
    rng = np.random.default_rng(5)
    n = 8000
    c, h, v = _cloud(n, seed=21)
    boxes = np.concatenate([c - h, c + h], 1).astype(np.float32)
    ctx = jbgpu.Context(bp_mode=MODES[mode], world_min=WMIN, world_max=WMAX)
    uids = ctx.bp_create_proxies(boxes)
    assert np.array_equal(uids, np.arange(1, n + 1))
    orc = pyoracle.Broadphase(MODES[mode], np.array(WMIN, np.float32), np.array(WMAX, np.float32))
    orc.create(boxes)
    prev = np.zeros(0, np.uint64)
    dead = np.zeros(0, np.int64)
    for step in range(6):
        if step == 2:  # destroy a tenth of the proxies: their pairs leave the current set at once
            dead = rng.choice(n, n // 10, replace=False) + 1
            ctx.bp_destroy_proxies(dead)
            orc.destroy(dead)
            cur = ctx.bp_pairs()
            lo, hi = (cur >> np.uint64(32)).astype(np.int64), (cur & np.uint64(0xffffffff)).astype(np.int64)
            assert not (np.isin(lo, dead) | np.isin(hi, dead)).any()
            keep = ~(np.isin((prev >> np.uint64(32)).astype(np.int64), dead) |
                     np.isin((prev & np.uint64(0xffffffff)).astype(np.int64), dead))
            assert np.array_equal(cur, prev[keep])
            prev = cur
        if step == 4:  # createProxy after the first batch: uids continue
            c2, h2, v2 = _cloud(1000, seed=22)
            more = np.concatenate([c2 - h2, c2 + h2], 1).astype(np.float32)
            u2 = ctx.bp_create_proxies(more)
            assert np.array_equal(u2, np.arange(n + 1, n + 1001))
            orc.create(more)
            c, h, v = np.concatenate([c, c2]), np.concatenate([h, h2]), np.concatenate([v, v2])
        boxes = scenes.aabb_cloud_step(c, h, v)
        ctx.bp_set_aabbs(boxes)
        orc.set_all(boxes)
        ctx.bp_calculate_overlapping_pairs()
        got, want = ctx.bp_pairs(), orc.step()
        added, removed = ctx.bp_pair_deltas()
        print(f"[bp-destroy {mode}] step {step}: {len(got)} pairs (oracle {len(want)}), +{len(added)} -{len(removed)}",
              flush=True)
        assert np.array_equal(got, want), f"{mode} step {step}"
        assert np.array_equal(added, np.setdiff1d(got, prev))
        assert np.array_equal(removed, np.setdiff1d(prev, got))
        prev = got
    ctx.close()


def _world_pair(sc, iterations=10, **kw):
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=iterations, **kw)
    ctx.world_create(shapes, verts, bodies, world_ids=sc.world)
    orc = pyoracle.World(shapes, verts, bodies, iterations=iterations, order=jbgpu.ORDER_EXPLICIT)
    if sc.world is not None:
        orc.set_worlds(sc.world)
    return ctx, orc


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


def _compare_world(ctx, orc, tag):
    bg, bo = ctx.bodies(), orc.bodies()
    if not np.array_equal(bg.view(np.uint32), bo.view(np.uint32)):
        d = np.abs(bg.astype(np.float64) - bo) / np.maximum(1.0, np.abs(bo))
        raise AssertionError(f"{tag}: bodies differ, max rel {d.max():.3e}")
    kg, mg = ctx.pairs()
    ko, mo = orc.pairs()
    assert np.array_equal(kg, ko), f"{tag}: pair keys differ ({len(kg)} vs {len(ko)})"
    # manifold records: friction directions are meaningful only once initialised
    mg = mg.copy()
    mo = mo.copy()
    for m in (mg, mo):
        pts = m[:, 4:].reshape(len(m), 4, 32)
        npts = m[:, 0].view(np.int32)
        for j in range(4):
            dead = j >= npts
            pts[dead, j, :] = 0
            uninit = pts[:, j, 27].view(np.int32) == 0
            pts[uninit, j, 21:27] = 0
        m[:, 4:] = pts.reshape(len(m), -1)
    if not np.array_equal(mg.view(np.uint32), mo.view(np.uint32)):
        bad = np.nonzero((mg.view(np.uint32) != mo.view(np.uint32)).any(1))[0]
        raise AssertionError(f"{tag}: {len(bad)} manifolds differ, first pair index {bad[0]}")


@pytest.mark.parametrize("blocks,tail", [(2048, 0), (256, 0), (0, 1024), (0, 0)])
def test_world_boxes_grid_step_parity(blocks, tail):
    """The 1000-box configuration for its 300 steps under each launch shape of the sweeps: blocks = bodies per
    block of the block-partitioned sweep (0: none); tail: trailing colours of at most this many manifolds swept in
    one workgroup (0: one launch each)."""
    sc = scenes.boxes_grid(seed=1)
    ctx, orc = _world_pair(sc)
    ctx.solver_set_blocks(blocks)
    ctx.solver_set_tail(tail)
    for step in range(300):  # the 1000-box configuration's 300 steps, under every sweep shape
        _sync_and_step(ctx, orc)
        st = ctx.stats()
        if step % 15 == 0:
            print(f"[boxes] step {step}: pairs {st['pairs']} contacts {st['contacts']} manifolds {st['manifolds']} "
                  f"colors {st['colors']} epa {st['epa_calls']}", flush=True)
        _compare_world(ctx, orc, f"boxes step {step}")
    ctx.close()


def test_tolerance_mode_step_parity():
    """Tolerance mode (iterate until ||dlambda|| / max(1, ||lambda||) <= 1e-3, at most 60 iterations): the host
    stops launching sweeps once the device's residual check has passed; the iteration count and the state are
    those of the oracle run with the same tolerance in the device's order."""
    sc = scenes.boxes_grid(seed=2)
    ctx, orc = _world_pair(sc)
    ctx.solver_set_params(tolerance=1e-3, max_iterations=60, num_iterations=10)
    orc.set_solver(jbgpu.ORDER_EXPLICIT, 1e-3, 60, 10)
    early = 0
    for step in range(120):
        _sync_and_step(ctx, orc)
        _compare_world(ctx, orc, f"tolerance step {step}")
        it_d, it_o = ctx.stats()["iterations"], orc.stats()["iterations"]
        assert it_d == it_o, (step, it_d, it_o)
        early += 0 < it_d < 60
    assert early > 0  # some solves converged before the cap
    ctx.close()


@pytest.mark.parametrize("degree_first,recolour", [(1, 0), (0, 0), (1, 2)])
def test_world_mixed_shapes_step_parity(degree_first, recolour):
    """Mixed pile (all shape pairs, EPA), under each colouring: largest-degree-first or hashed Jones-Plassmann
    priorities, with and without iterated-greedy recolouring passes (each gives its own valid order)."""
    sc = scenes.mixed_pile(seed=5, n=3000, bin_half=12.0)
    ctx, orc = _world_pair(sc, iterations=20)
    ctx.solver_set_colouring(degree_first, recolour)
    for step in range(120):
        _sync_and_step(ctx, orc)
        st = ctx.stats()
        if step % 20 == 0:
            print(f"[mixed] step {step}: pairs {st['pairs']} contacts {st['contacts']} manifolds {st['manifolds']} "
                  f"colors {st['colors']} epa {st['epa_calls']}", flush=True)
        _compare_world(ctx, orc, f"mixed step {step}")
    ctx.close()


@pytest.mark.parametrize("tail", [1 << 30, 0])
def test_tail_sweep_paths_step_parity(tail):
    """Resting box pyramids (four-point manifolds from the first step, ~6,000 manifolds): every colour in the
    single-workgroup tail sweep (colours of up to ~1,000 manifolds, several passes each), and no tail at all."""
    sc = scenes.pyramids(seed=7, n_boxes=3000, n_pyramids=6, n_spheres=0, area=60.0)
    ctx, orc = _world_pair(sc, iterations=10)
    ctx.solver_set_tail(tail)
    for step in range(40):
        _sync_and_step(ctx, orc)
        st = ctx.stats()
        if step % 10 == 0:
            print(f"[tail {tail}] step {step}: manifolds {st['manifolds']} rows {st['rows']} colors {st['colors']}",
                  flush=True)
        _compare_world(ctx, orc, f"tail {tail} step {step}")
    ctx.close()


def test_batched_worlds_step_parity():
    """Independent worlds batched in one context (own floor plane each, pairs never cross worlds)."""
    sc = scenes.batched_towers(seed=3, n_worlds=24, lo=16, hi=120)
    ctx, orc = _world_pair(sc, iterations=10)
    for step in range(80):
        _sync_and_step(ctx, orc)
        st = ctx.stats()
        if step % 20 == 0:
            print(f"[batched] step {step}: pairs {st['pairs']} contacts {st['contacts']} manifolds {st['manifolds']} "
                  f"colors {st['colors']}", flush=True)
        _compare_world(ctx, orc, f"batched step {step}")
    kg, _ = ctx.pairs()
    lo = (kg >> np.uint64(32)).astype(np.int64) - 1
    hi = (kg & np.uint64(0xffffffff)).astype(np.int64) - 1
    assert np.array_equal(sc.world[lo], sc.world[hi])
    ctx.close()


def test_sparse_world_ids_step_parity():
    """The same batched towers with world ids spread up to ~1M (20 bits): the packed Morton sort key (world id, 12
    Morton bits, huge marker) no longer fits 32 bits, so the tree is sorted and built on 64-bit keys. Bit-exact
    bodies, pairs and manifolds every step, and pairs never cross worlds."""
    sc = scenes.batched_towers(seed=3, n_worlds=24, lo=16, hi=120)
    sc.world = (sc.world.astype(np.int64) * 45_000 + 11).astype(np.int32)
    assert int(sc.world.max()) >= 1 << 19
    ctx, orc = _world_pair(sc, iterations=10)
    for step in range(40):
        _sync_and_step(ctx, orc)
        if step % 20 == 0:
            st = ctx.stats()
            print(f"[sparse worlds] step {step}: pairs {st['pairs']} manifolds {st['manifolds']}", flush=True)
        _compare_world(ctx, orc, f"sparse worlds step {step}")
    kg, _ = ctx.pairs()
    lo = (kg >> np.uint64(32)).astype(np.int64) - 1
    hi = (kg & np.uint64(0xffffffff)).astype(np.int64) - 1
    assert len(kg) > 0 and np.array_equal(sc.world[lo], sc.world[hi])
    ctx.close()


@pytest.mark.parametrize("mode", ["dbvt", "sap32"])
def test_broadphase_1m_aabbs_match_oracle(mode):
    """Full-size broadphase stress: 1M moving AABBs, pair keys bit-exact every step (count + SHA-256)."""
    import hashlib
    c, h, v = scenes.aabb_cloud(seed=3, n=1_000_000)
    boxes = np.concatenate([c - h, c + h], 1).astype(np.float32)
    ctx = jbgpu.Context(bp_mode=MODES[mode], world_min=WMIN, world_max=WMAX)
    ctx.bp_create_proxies(boxes)
    orc = pyoracle.Broadphase(MODES[mode], np.array(WMIN, np.float32), np.array(WMAX, np.float32))
    orc.create(boxes)
    for step in range(4):
        boxes = scenes.aabb_cloud_step(c, h, v)
        ctx.bp_set_aabbs(boxes)
        orc.set_all(boxes)
        ctx.bp_calculate_overlapping_pairs()
        got, want = ctx.bp_pairs(), orc.step()
        dg, do = hashlib.sha256(got.tobytes()).hexdigest()[:16], hashlib.sha256(want.tobytes()).hexdigest()[:16]
        print(f"[bp1m {mode}] step {step}: {len(got)} pairs sha {dg} (oracle {len(want)} {do})", flush=True)
        assert len(got) == len(want) and dg == do
    ctx.close()


def test_sleeping_islands_step_parity():
    """Deactivation enabled: island sleep/wake decisions on the device match the oracle every step."""
    sc = scenes.sleep_and_wake()
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=10, deactivation_disabled=0)
    ctx.world_create(shapes, verts, bodies)
    orc = pyoracle.World(shapes, verts, bodies, iterations=10, order=jbgpu.ORDER_EXPLICIT, deactivation_disabled=False)
    seen = set()
    for step in range(400):
        _sync_and_step(ctx, orc)
        _compare_world(ctx, orc, f"sleep step {step}")
        seen |= set(ctx.bodies()[:, 29].view(np.int32).tolist())
        if step % 50 == 0:
            print(f"[sleep] step {step}: states {sorted(seen)}", flush=True)
    assert {1, 2, 3} <= seen
    ctx.close()


def test_world_mixed_pile_100k_parity():
    """Full-size pile (100,000 bodies): 300 device-only steps, then 4 steps checked bit for bit."""
    sc = scenes.mixed_pile(seed=1, n=100_000)
    ctx, orc = _world_pair(sc, iterations=20)
    ctx.step(n=300)
    st = ctx.stats()
    print(f"[100k] after 300 steps: pairs {st['pairs']} manifolds {st['manifolds']} colors {st['colors']}", flush=True)
    for step in range(4):
        _sync_and_step(ctx, orc)
        _compare_world(ctx, orc, f"100k step {301 + step}")
        print(f"[100k] step {301 + step} bit-exact", flush=True)
    ctx.close()


def test_colour_batches_share_no_dynamic_body():
    """The 100k pile's colour batches (Jones-Plassmann rounds, the first batch's length adapting from step to
    step): within a batch no two manifolds share a dynamic body, and every active manifold is in exactly one
    batch, on every checked step."""
    sc = scenes.mixed_pile(seed=1, n=100_000)
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=20)
    ctx.world_create(shapes, verts, bodies)
    dyn = bodies[:, 18] != 0  # inverse mass
    ctx.step(n=300)
    for step in range(6):
        ctx.step()
        order, offs = ctx.solver_order()
        _, m = ctx.pairs()
        st = ctx.stats()
        assert len(order) == st["manifolds"] and len(np.unique(order)) == len(order), f"step {301 + step}"
        b0 = m[order, 1].view(np.int32)
        b1 = m[order, 2].view(np.int32)
        for c in range(len(offs) - 1):
            x, y = b0[offs[c]:offs[c + 1]], b1[offs[c]:offs[c + 1]]
            bs = np.concatenate([x[dyn[x]], y[dyn[y]]])
            assert len(np.unique(bs)) == len(bs), f"step {301 + step}: colour {c} repeats a dynamic body"
        print(f"[colours] step {301 + step}: {len(offs) - 1} batches over {len(order)} manifolds", flush=True)
    ctx.close()


def test_epa_workspace_retry_step_parity():
    """EPA jobs run first with a small LDS workspace; the few that outgrow it are rerun with the large one. The
    100k pile has such jobs from step ~130 on (a few per ten steps): steps whose EPA pass had retries are
    replayed by the oracle from the device state before them and must agree bit for bit."""
    sc = scenes.mixed_pile(seed=1, n=100_000)
    ctx, orc = _world_pair(sc, iterations=20)
    ctx.step(n=140)
    checked = 0
    for step in range(141, 241):
        b0 = ctx.bodies()
        prox, s0 = ctx.bp_proxies()
        k0, m0 = ctx.pairs()
        ctx.step()
        st = ctx.stats()
        if st["epa_retries"] == 0:
            continue
        order, _ = ctx.solver_order()
        orc.set_bodies(b0)
        orc.set_proxies(prox, s0)
        orc.set_pairs(k0, m0)
        orc.set_order(order)
        orc.step()
        _compare_world(ctx, orc, f"retry step {step}")
        print(f"[epa retry] step {step}: {st['epa_calls']} EPA jobs, {st['epa_retries']} rerun; bit-exact",
              flush=True)
        checked += 1
        if checked == 2:
            break
    assert checked > 0, "no EPA job outgrew the first-pass workspace in steps 141-240"
    ctx.close()


@pytest.mark.parametrize("small,large", [(16, 16), (16, 1024), (24, 48)])
def test_epa_tiers_step_parity(small, large):
    """EPA jobs run in a small LDS workspace, then a large LDS one, then a global-memory one, each job moving on
    only when it outgrows the last.  With the LDS tiers' face capacities lowered (16 faces: nearly every job moves
    on), the mixed pile's EPA jobs go through all three tiers; the steps stay bit-exact against the oracle (the
    capacities move work, they never change a result)."""
    sc = scenes.mixed_pile(seed=5, n=3000, bin_half=12.0)
    ctx, orc = _world_pair(sc, iterations=20)
    ctx.np_set_epa_capacity(small, large)
    retried = huge = 0
    for step in range(90):
        _sync_and_step(ctx, orc)
        _compare_world(ctx, orc, f"EPA tiers {small}/{large} step {step}")
        st = ctx.stats()
        retried += st["epa_retries"]
        huge += st["epa_huge"]
    print(f"[epa tiers {small}/{large}] {retried} jobs rerun in the large tier, {huge} in the global one", flush=True)
    assert retried > 0
    if large == 16:
        assert huge > 0
    ctx.close()


def _rays(rng, n, lo, hi, vertical_share=0.3):
    """Ray segments between random points of the box [lo, hi]; a share of them vertical (axis-parallel)."""
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    a = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    b = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    v = rng.random(n) < vertical_share
    b[v, 0], b[v, 2] = a[v, 0], a[v, 2]
    a[v, 1], b[v, 1] = hi[1], lo[1]
    return a, b


def _check_rays(ctx, orc, frm, to, what, world=None, group=1, mask=-1):
    got = ctx.ray_test(frm, to, group=group, mask=mask, world=world)
    want = orc.ray_test(frm, to, group=group, mask=mask, world=world)
    hits = int((want[0] >= 0).sum())
    print(f"[rays {what}] {len(frm)} rays, {hits} hits", flush=True)
    assert np.array_equal(got[0], want[0]), f"{what}: hit objects differ"
    for g, w_, name in zip(got[1:], want[1:], ("fraction", "normal", "point")):
        assert np.array_equal(g.view(np.uint32), w_.view(np.uint32)), f"{what}: {name} bits differ"
    return hits


@pytest.mark.parametrize("scene", ["boxes", "mixed", "batched"])
def test_ray_test_matches_oracle(scene):
    """CollisionWorld.rayTest (closest hit) on the device vs the oracle, from the same body state: hit body,
    fraction, normal and point bit for bit; vertical rays included (axis-parallel segments)."""
    rng = np.random.default_rng(17)
    world = None
    if scene == "boxes":
        sc = scenes.boxes_grid(seed=2)
        ctx, orc = _world_pair(sc, 10)
        lo, hi = (-10, -1, -10), (10, 20, 10)
    elif scene == "mixed":
        sc = scenes.mixed_pile(seed=4, n=3000, bin_half=15.0)
        ctx, orc = _world_pair(sc, 10)
        lo, hi = (-17, -1, -17), (17, 30, 17)
    else:
        sc = scenes.batched_towers(seed=5, n_worlds=24)
        ctx, orc = _world_pair(sc, 10)
        lo, hi = (-15, -1, -15), (15, 25, 15)
    ctx.step(n=40)
    orc.set_bodies(ctx.bodies())
    frm, to = _rays(rng, 3000, lo, hi)
    if scene == "batched":
        world = rng.integers(0, int(sc.world.max()) + 1, len(frm)).astype(np.int32)
    hits = _check_rays(ctx, orc, frm, to, scene, world=world)
    assert hits > len(frm) // 4
    if scene == "mixed":  # the callback filter: static bodies (group 2) excluded
        _check_rays(ctx, orc, frm[:500], to[:500], "mixed, static excluded", mask=np.int16(-1 ^ 2))
    ctx.close()


def test_ray_test_100k_pile_matches_oracle():
    """rayTest in the 100k-body pile after 120 steps: 300 rays (the oracle tests every body per ray)."""
    rng = np.random.default_rng(23)
    sc = scenes.mixed_pile(seed=1, n=100_000)
    ctx, orc = _world_pair(sc, 10)
    ctx.step(n=120)
    orc.set_bodies(ctx.bodies())
    frm, to = _rays(rng, 300, (-52, -1, -52), (52, 40, 52), vertical_share=0.5)
    hits = _check_rays(ctx, orc, frm, to, "pile 100k")
    assert hits > 100
    ctx.close()


@pytest.mark.parametrize("bodies,threads,pipelined", [(16, 64, 0), (256, 256, 0), (1024, 512, 0), (16, 64, 1),
                                                      (256, 256, 1), (1024, 512, 1)])
def test_batched_world_blocks_step_parity(bodies, threads, pipelined):
    """The batched solve with one workgroup per block of whole worlds (colouring, warm start and all iterations
    in one launch): bit-exact against the oracle fed the block-by-block manifold order, for blocks of one small
    world (16 bodies: every world is a block), of several worlds and of many."""
    sc = scenes.batched_towers(seed=3, n_worlds=24, lo=16, hi=120)
    ctx, orc = _world_pair(sc, iterations=10)
    ctx.solver_set_world_blocks(bodies, threads, pipelined)
    for step in range(60):
        _sync_and_step(ctx, orc)
        st = ctx.stats()
        if step % 20 == 0:
            print(f"[world blocks {bodies}/{threads}/{pipelined}] step {step}: manifolds {st['manifolds']} rows {st['rows']} "
                  f"colours {st['colors']} residual {st['residual']:.3e}", flush=True)
        _compare_world(ctx, orc, f"world blocks {bodies} step {step}")
    ctx.close()


@pytest.mark.parametrize("world_blocks", [0, 512])
def test_batched_4096_worlds_step_parity(world_blocks):
    """The batched-worlds configuration at its full size: 4,096 worlds of log-uniform 16-512 bodies (~588k
    bodies, sizes as in the mixed pile), towers on their own floors, 10 iterations.  50 device-only steps, then
    two steps checked bit for bit against the oracle stepping all worlds from the same state."""
    sc = scenes.batched_towers(seed=1, n_worlds=4096)
    print(f"[batched 4096] {sc.n_bodies} bodies, world blocks {world_blocks}", flush=True)
    ctx, orc = _world_pair(sc, iterations=10)
    ctx.solver_set_world_blocks(world_blocks, 256, 1)
    ctx.step(n=50)
    st = ctx.stats()
    print(f"[batched 4096] after 50 steps: pairs {st['pairs']} manifolds {st['manifolds']} colors {st['colors']}",
          flush=True)
    for step in range(2):
        _sync_and_step(ctx, orc)
        _compare_world(ctx, orc, f"batched 4096 step {51 + step}")
        print(f"[batched 4096] step {51 + step} bit-exact", flush=True)
    kg, _ = ctx.pairs()
    lo = (kg >> np.uint64(32)).astype(np.int64) - 1
    hi = (kg & np.uint64(0xffffffff)).astype(np.int64) - 1
    assert np.array_equal(sc.world[lo], sc.world[hi])
    ctx.close()
