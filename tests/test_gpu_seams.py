"""Each plugin seam alone on the device, the rest of the step on the CPU (the oracle), through the C ABI:

* Dispatcher: the oracle runs predict, AABBs and the broadphase; the device runs
  dispatchAllCollisionPairs (jbgpu_np_dispatch) on the oracle's pair cache; the
  manifolds must equal the oracle's own narrowphase bit for bit.
* ConstraintSolver: the oracle runs everything up to the narrowphase; the device
  solves the islands recorded through prepareSolve / solveGroup / allSolved
  (jbgpu_solver_prepare / _add_island / _all_solved); velocities and impulses must
  equal the oracle's solveGroup fed the device's manifold order bit for bit.
"""
import numpy as np
import pytest

import jbgpu
import pyoracle
from jbgpu import scenes

pytestmark = pytest.mark.gpu


def _pair(sc, iterations, order=jbgpu.ORDER_EXPLICIT):
    shapes, verts, bodies = scenes.build_records(sc, jbgpu)
    ctx = jbgpu.Context(num_iterations=iterations)
    ctx.world_create(shapes, verts, bodies)
    orc = pyoracle.World(shapes, verts, bodies, iterations=iterations, order=order)
    return ctx, orc, bodies


def _masked(m):
    """Manifold records with the words that carry no state cleared: points past the count, and friction
    directions of points whose directions were never set."""
    m = m.copy()
    pts = m[:, 4:].reshape(len(m), 4, 32)
    npts = m[:, 0].view(np.int32)
    for j in range(4):
        pts[j >= npts, j, :] = 0
        uninit = pts[:, j, 27].view(np.int32) == 0
        pts[uninit, j, 21:27] = 0
    m[:, 4:] = pts.reshape(len(m), -1)
    return m


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _islands(keys, man, bodies):
    """The manifolds buildAndProcessIslands hands to solveGroup (deactivation disabled: every manifold with
    points and a dynamic body), grouped by connected component of the dynamic bodies."""
    static = bodies[:, 28].view(np.int32) & 1
    parent = np.arange(len(bodies))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    lo = (keys >> np.uint64(32)).astype(np.int64) - 1
    hi = (keys & np.uint64(0xffffffff)).astype(np.int64) - 1
    for a, b in zip(lo, hi):
        if not static[a] and not static[b]:
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    groups = {}
    npts = man[:, 0].view(np.int32)
    for i, (a, b) in enumerate(zip(lo, hi)):
        if npts[i] == 0 or (static[a] and static[b]):
            continue
        groups.setdefault(find(b if static[a] else a), []).append(i)
    return [np.array(v, np.int32) for _, v in sorted(groups.items())]


def test_dispatcher_seam_matches_oracle():
    sc = scenes.mixed_pile(seed=6, n=3000, bin_half=12.0)
    ctx, orc, _ = _pair(sc, 20, order=jbgpu.ORDER_REFERENCE)  # the CPU solves in the reference's own order
    checked = 0
    for step in range(100):
        orc.phase(0)
        orc.phase(1)
        b = orc.bodies()
        k, m = orc.pairs()
        ctx.set_bodies(b)
        got = ctx.np_dispatch(k, m)
        orc.phase(2)
        ko, mo = orc.pairs()
        assert np.array_equal(k, ko)
        assert _same(_masked(got), _masked(mo)), f"step {step}: manifolds differ"
        checked += int(got[:, 0].view(np.int32).sum())
        if step % 25 == 0:
            st = ctx.stats()
            print(f"[dispatcher seam] step {step}: {len(k)} pairs, {checked} points so far, epa {st['epa_calls']}",
                  flush=True)
        orc.phase(3)
        orc.phase(4)
    assert checked > 10000
    ctx.close()


def test_solver_seam_matches_oracle():
    sc = scenes.mixed_pile(seed=7, n=3000, bin_half=12.0)
    ctx, orc, bodies = _pair(sc, 20)
    solved = 0
    for step in range(80):
        orc.phase(0)
        orc.phase(1)
        orc.phase(2)
        b = orc.bodies()
        k, m = orc.pairs()
        ctx.set_bodies(b)
        ctx.set_pairs(k, m)
        islands = _islands(k, m, b)
        ctx.solver_prepare(len(b), len(k))
        for isl in islands:
            ctx.solver_add_island(isl)
        ctx.solver_all_solved()
        order, offs = ctx.solver_order()
        assert set(order.tolist()) <= set(np.concatenate(islands).tolist()) if islands else len(order) == 0
        orc.set_order(order)
        orc.phase(3)
        assert _same(ctx.bodies(), orc.bodies()), f"step {step}: velocities differ"
        kg, mg = ctx.pairs()
        ko, mo = orc.pairs()
        assert np.array_equal(kg, ko) and _same(_masked(mg), _masked(mo)), f"step {step}: impulses differ"
        solved += len(order)
        if step % 20 == 0:
            print(f"[solver seam] step {step}: {len(islands)} islands, {len(order)} manifolds, "
                  f"{len(offs) - 1} colours", flush=True)
        orc.phase(4)
    assert solved > 1000
    ctx.close()


def test_seam_argument_errors():
    """Seam misuse is refused with a status and changes nothing: allSolved without prepareSolve, an island index
    outside the pair cache, one manifold in two islands, a manifold header that is not the dispatcher's."""
    sc = scenes.mixed_pile(seed=8, n=400, bin_half=6.0)
    ctx, orc, _ = _pair(sc, 10)
    ctx.step(n=40)
    k, m = ctx.pairs()
    with pytest.raises(jbgpu.JbgpuError, match=r"\(-3\)"):
        ctx.solver_all_solved()
    with pytest.raises(jbgpu.JbgpuError, match=r"\(-1\)"):
        ctx.solver_prepare(ctx.nb, len(k) + 1)
    ctx.solver_prepare(ctx.nb, len(k))
    with pytest.raises(jbgpu.JbgpuError, match="out of range"):
        ctx.solver_add_island([len(k)])
    ctx.solver_add_island([0, 1])
    ctx.solver_add_island([1])
    with pytest.raises(jbgpu.JbgpuError, match="two islands"):
        ctx.solver_all_solved()
    full = np.nonzero(m[:, 0].view(np.int32) > 0)[0]
    bad = m.copy()
    i = int(full[0])
    bad[i, 1:3] = bad[i, 2:0:-1].copy()  # body order swapped on a record with points
    with pytest.raises(jbgpu.JbgpuError, match="header"):
        ctx.np_dispatch(k, bad)
    k2, m2 = ctx.pairs()
    assert np.array_equal(k, k2) and _same(m, m2)
    ctx.step()  # the world carries on
    ctx.close()
