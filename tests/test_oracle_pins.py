"""CPU pins of the oracle (no GPU).

The reference ships no tests or golden vectors, so each part of the oracle is
pinned by something independent of its own code: published known answers
(java.util.Random), closed forms (sphere/plane contacts, inertia of a cube,
rotations of known axes, free fall), brute force (O(N^2) pair sets, polytope
distances in float64) and invariants (pair-set symmetry and filtering,
friction-cone complementarity, manifold capacity and deepest-point retention).
"""
import itertools

import numpy as np
import pytest

import pyoracle as po
from jbgpu import scenes

SPHERE, BOX, CAPSULE, HULL, PLANE = 0, 1, 2, 3, 4


def tr(basis=None, origin=(0, 0, 0)):
    b = np.eye(3, dtype=np.float32) if basis is None else np.asarray(basis, np.float32).reshape(3, 3)
    return np.concatenate([b.reshape(-1), np.asarray(origin, np.float32)]).astype(np.float32)


def rot(axis, ang):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def random_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ------------------------------------------------------------------ java.util.Random
def test_java_random_known_answers():
    # values published for the Java SE LCG: new Random(0).nextInt(), new Random(42).nextInt()
    assert po.random_ints(0, 1)[0] == -1155484576
    assert po.random_ints(42, 1)[0] == -1170105035
    # the solver's shuffle draws nextInt(j + 1), j = 0..9, from one Random(0) stream
    assert list(po.random_bounded(0, list(range(1, 11)))) == [0, 1, 1, 2, 0, 5, 2, 0, 6, 4]


def test_java_random_bounds():
    b = np.arange(1, 2000)
    v = po.random_bounded(7, b)
    assert (v >= 0).all() and (v < b).all()
    assert (po.random_bounded(7, [1] * 50) == 0).all()


# ------------------------------------------------------------------ libGDX rotation semantics
def test_quaternion_matrix_round_trip():
    rng = np.random.default_rng(0)
    for _ in range(200):
        q = rng.normal(size=4).astype(np.float32)
        q /= np.linalg.norm(q)
        m = np.zeros(9, np.float32)
        po.lib().orc_matrix_from_quat(q, m)
        q2 = np.zeros(4, np.float32)
        po.lib().orc_quat_from_matrix(m, q2)
        assert min(np.abs(q2 - q).max(), np.abs(q2 + q).max()) < 2e-6
        M = m.reshape(3, 3).astype(np.float64)
        assert np.abs(M @ M.T - np.eye(3)).max() < 1e-6


def test_matrix_of_known_rotation():
    # quarter turn about z: (x, y, z, w) = (0, 0, sin 45, cos 45) maps x to y
    s = np.float32(np.sqrt(0.5))
    m = np.zeros(9, np.float32)
    po.lib().orc_matrix_from_quat(np.array([0, 0, s, s], np.float32), m)
    assert np.abs(m.reshape(3, 3) @ np.array([1, 0, 0]) - np.array([0, 1, 0])).max() < 1e-6


# ------------------------------------------------------------------ shapes
def test_box_record_support_and_inertia():
    rec = po.make_shape(BOX, [0.5, 0.5, 0.5])
    assert rec[1] == np.float32(0.04) and np.allclose(rec[2:5], 0.46)
    # support with margin is the full box; -0.0 picks the positive side (fsel: a >= 0)
    assert np.array_equal(po.support(rec, [1, -1, -0.0], True), np.float32([0.5, -0.5, 0.5]))
    assert np.array_equal(po.support(rec, [-1, 1, 2], False), np.float32([-0.46, 0.46, 0.46]))
    # a cube of side 1 and mass 6 has I = m s^2 / 6 = 1 on every axis
    assert np.allclose(po.shape_inertia(rec, 6.0), 1.0, atol=1e-6)


def test_sphere_record_aabb_inertia():
    rec = po.make_shape(SPHERE, [0.7])
    assert rec[1] == np.float32(0.7)
    a = po.aabb(rec, tr(origin=(1, 2, 3)))
    assert np.allclose(a, [0.3, 1.3, 2.3, 1.7, 2.7, 3.7])
    assert np.allclose(po.shape_inertia(rec, 2.0), 0.4 * 2 * 0.49, rtol=1e-6)  # 2/5 m r^2


def test_rotated_box_aabb_closed_form():
    rec = po.make_shape(BOX, [1.0, 0.25, 0.5])
    ang = 0.6
    a = po.aabb(rec, tr(rot([0, 1, 0], ang), (0, 0, 0)))
    ex = 1.0 * np.cos(ang) + 0.5 * np.sin(ang)
    ez = 1.0 * np.sin(ang) + 0.5 * np.cos(ang)
    assert np.allclose(a[3:], [ex, 0.25, ez], atol=1e-6) and np.allclose(a[:3], -a[3:])


def test_hull_support_is_first_maximum():
    rng = np.random.default_rng(3)
    pts = rng.normal(size=(24, 3)).astype(np.float32)
    rec = po.make_shape(HULL, [0, 24], pts)
    for _ in range(100):
        d = rng.normal(size=3).astype(np.float32)
        dn = d / np.float32(np.sqrt(np.float32(d @ d)))
        dots = pts.astype(np.float64) @ dn.astype(np.float64)
        want = pts[int(np.argmax(dots))]
        got = po.support(rec, d, False, pts)
        assert np.array_equal(got, want)
    # local AABB = support along the axes +- margin
    assert np.allclose(rec[10:13], pts.max(0) + 0.04) and np.allclose(rec[7:10], pts.min(0) - 0.04)


def test_capsule_support_closed_form():
    rec = po.make_shape(CAPSULE, [0.3, 1.0])   # radius 0.3, cylinder height 1 (half 0.5)
    for d in ([0, 1, 0], [0, -1, 0], [1, 0.2, 0], [0.3, -0.8, 0.5]):
        d = np.asarray(d, np.float64)
        dn = d / np.linalg.norm(d)
        cap = np.array([0, 0.5 if d[1] >= 0 else -0.5, 0])
        want = cap + dn * 0.3  # full radius once the margin is added back
        got = po.support(rec, d, True)
        assert np.abs(got - want).max() < 1e-6


# ------------------------------------------------------------------ narrowphase closed forms
def test_sphere_sphere_contact():
    a = po.make_shape(SPHERE, [0.5])
    b = po.make_shape(SPHERE, [0.4])
    out = po.sphere_sphere(a, tr(origin=(0.8, 0, 0)), b, tr(origin=(0, 0, 0)))
    assert out[0] == 1 and np.allclose(out[1:4], [1, 0, 0])
    assert np.allclose(out[4:7], [0.4, 0, 0]) and np.isclose(out[7], 0.8 - 0.9)
    far = po.sphere_sphere(a, tr(origin=(2, 0, 0)), b, tr())
    assert far[0] == 0


def test_box_on_plane_contact():
    box = po.make_shape(BOX, [0.5, 0.5, 0.5])
    plane = po.make_shape(PLANE, [0, 1, 0, 0])
    out = po.convex_plane(box, tr(origin=(1, 0.49, 2)), plane, tr())
    assert out[0] == 1 and np.allclose(out[1:4], [0, 1, 0])
    assert np.isclose(out[7], -0.01, atol=1e-6)
    assert np.isclose(out[5], 0.0, atol=1e-6)  # projected onto the plane
    assert convex_plane_far(box, plane)


def convex_plane_far(box, plane):
    return po.convex_plane(box, tr(origin=(0, 0.6, 0)), plane, tr())[0] == 0


def _box_vertices(h):
    return np.array(list(itertools.product(*[(-x, x) for x in h])), np.float64)


def _seg_seg(p1, q1, p2, q2):
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-15 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    return np.linalg.norm((p1 + d1 * s) - (p2 + d2 * t))


def box_box_distance(ha, Ra, ta, hb, Rb, tb):
    """Exact distance between two separated boxes (vertex-box and edge-edge features), float64."""
    va = _box_vertices(ha) @ Ra.T + ta
    vb = _box_vertices(hb) @ Rb.T + tb
    best = np.inf
    for (verts, h, R, t) in ((va, hb, Rb, tb), (vb, ha, Ra, ta)):
        loc = (verts - t) @ R
        clamped = np.clip(loc, -h, h)
        best = min(best, np.linalg.norm(loc - clamped, axis=1).min())
    ea = [(i, j) for i in range(8) for j in range(i + 1, 8) if bin(i ^ j).count("1") == 1]
    for i, j in ea:
        for k, l in ea:
            best = min(best, _seg_seg(va[i], va[j], vb[k], vb[l]))
    return best


def test_gjk_distance_against_brute_force():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(60):
        ha = rng.uniform(0.2, 1.0, 3)
        hb = rng.uniform(0.2, 1.0, 3)
        Ra, Rb = random_rot(rng), random_rot(rng)
        ta = np.zeros(3)
        d = rng.normal(size=3)
        tb = d / np.linalg.norm(d) * (np.linalg.norm(ha) + np.linalg.norm(hb) + rng.uniform(0.05, 1.0))
        recA = po.make_shape(BOX, ha)
        recB = po.make_shape(BOX, hb)
        core_a, core_b = recA[2:5].astype(np.float64), recB[2:5].astype(np.float64)
        TA = tr(Ra, ta)
        TB = tr(Rb, tb)
        out = po.gjk(recA, TA, recB, TB, 1e30)
        want = box_box_distance(core_a, TA[:9].reshape(3, 3).astype(np.float64), TA[9:].astype(np.float64),
                                core_b, TB[:9].reshape(3, 3).astype(np.float64), TB[9:].astype(np.float64)) - 0.08
        assert out[0] == 1
        # GJK stops when the squared distance improves by less than REL_ERROR2 = 1e-6 of itself
        # (GjkPairDetector.java:44,187-196); the float32 restatement lands within ~2e-7 of the float64 distance
        assert abs(out[7] - want) < 1e-5 * max(1.0, want), (out[7], want)
        assert abs(np.linalg.norm(out[1:4]) - 1) < 1e-5
        checked += 1
    assert checked == 60


def test_gjk_sphere_box_distance():
    rng = np.random.default_rng(5)
    box = po.make_shape(BOX, [0.7, 0.4, 0.5])
    sph = po.make_shape(SPHERE, [0.3])
    for _ in range(40):
        R = random_rot(rng)
        c = rng.normal(size=3)
        c = c / np.linalg.norm(c) * rng.uniform(1.5, 3.0)
        TB = tr(R, (0, 0, 0))
        Rf = TB[:9].reshape(3, 3).astype(np.float64)
        loc = Rf.T @ c
        core = box[2:5].astype(np.float64)
        want = np.linalg.norm(loc - np.clip(loc, -core, core)) - 0.04 - 0.3
        out = po.gjk(sph, tr(origin=c), box, TB, 1e30)
        assert out[0] == 1 and abs(out[7] - want) < 1e-5 * max(1.0, abs(want)), (out[7], want)


def test_epa_depth_axis_aligned_boxes():
    for (ha, hb, off, axis) in [((1, 1, 1), (0.5, 0.5, 0.5), (1.3, 0.1, 0.05), 0),
                                ((0.6, 0.3, 0.8), (0.6, 0.3, 0.8), (0.2, 0.5, 0.1), 1),
                                ((2, 1, 0.4), (0.5, 0.5, 0.5), (0.3, 0.2, 0.75), 2)]:
        recA = po.make_shape(BOX, ha)
        recB = po.make_shape(BOX, hb)
        out = po.epa(recA, tr(), recB, tr(origin=off))
        overlap = np.array(ha) + np.array(hb) - np.abs(np.array(off))
        assert out[0] == 1
        depth = np.linalg.norm(out[1:4] - out[4:7])
        assert abs(depth - overlap.min()) < 2e-3, (depth, overlap)
        n = (out[4:7] - out[1:4]) / depth
        assert abs(abs(n[axis]) - 1) < 1e-3


def test_epa_sphere_depth():
    a = po.make_shape(SPHERE, [0.5])
    b = po.make_shape(SPHERE, [0.4])
    out = po.epa(a, tr(), b, tr(origin=(0.6, 0.2, -0.1)))
    d = np.linalg.norm([0.6, 0.2, -0.1])
    assert out[0] == 1 and abs(np.linalg.norm(out[1:4] - out[4:7]) - (0.9 - d)) < 2e-3


# ------------------------------------------------------------------ persistent manifold
def _cand(n, p, depth):
    return np.array([1, *n, *p, depth], np.float32)


def _manifold():
    m = np.zeros(po.MANIFOLD_WORDS, np.float32)
    m[1:3] = np.array([0, 1], np.int32).view(np.float32)
    return m


def test_manifold_replace_keeps_impulse_and_lifetime():
    I = tr()
    fr = [0.5, 0.5, 0.0, 0.0]
    m = po.manifold_add(_manifold(), I, I, fr, _cand([0, 1, 0], [0, 0, 0], -0.01), refresh=False)
    assert m[0].view(np.int32) == 1
    m[4 + 18] = 3.5  # cached impulse
    m = po.manifold_add(m, I, I, fr, _cand([0, 1, 0], [0.005, 0, 0], -0.012), refresh=True)
    assert m[0].view(np.int32) == 1 and m[4 + 18] == np.float32(3.5)
    assert np.isclose(m[4 + 15], -0.012) and m[4 + 28].view(np.int32) == 1
    assert np.isclose(m[4 + 16], 0.25)  # combined friction 0.5 * 0.5


def test_manifold_capacity_and_deepest_point():
    I = tr()
    fr = [0.5, 0.5, 0.0, 0.0]
    m = _manifold()
    corners = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)]
    depths = [-0.01, -0.015, -0.005, -0.018]
    for p, d in zip(corners, depths):
        m = po.manifold_add(m, I, I, fr, _cand([0, 1, 0], p, d), refresh=False)
    assert m[0].view(np.int32) == 4
    m = po.manifold_add(m, I, I, fr, _cand([0, 1, 0], (0.5, 0, 0.5), -0.001), refresh=False)
    assert m[0].view(np.int32) == 4
    pts = m[4:].reshape(4, 32)
    assert np.isclose(pts[:, 15].min(), -0.018)  # the deepest point survives the eviction


def test_manifold_refresh_drops_separated_and_drifted():
    I = tr()
    fr = [0.5, 0.5, 0.0, 0.0]
    m = po.manifold_add(_manifold(), I, I, fr, _cand([0, 1, 0], (0, 0, 0), -0.01), refresh=False)
    m = po.manifold_add(m, I, I, fr, _cand([0, 1, 0], (1, 0, 0), -0.01), refresh=False)
    assert m[0].view(np.int32) == 2
    # body A moves up 5 cm: both points separate beyond the 2 cm breaking threshold
    out = po.manifold_add(m, tr(origin=(0, 0.05, 0)), I, fr, np.zeros(8, np.float32), refresh=True)
    assert out[0].view(np.int32) == 0
    # body A slides 3 cm sideways: tangential drift beyond 2 cm
    out = po.manifold_add(m, tr(origin=(0.03, 0, 0)), I, fr, np.zeros(8, np.float32), refresh=True)
    assert out[0].view(np.int32) == 0
    out = po.manifold_add(m, tr(origin=(0.005, 0.004, 0)), I, fr, np.zeros(8, np.float32), refresh=True)
    assert out[0].view(np.int32) == 2


# ------------------------------------------------------------------ broadphase
def _boxes(rng, n, extent=20.0, size=1.0):
    c = rng.uniform(0, extent, (n, 3))
    h = rng.uniform(0.05, size, (n, 3))
    return np.concatenate([c - h, c + h], 1).astype(np.float32)


def _brute_pairs(boxes):
    n = len(boxes)
    mn, mx = boxes[:, :3], boxes[:, 3:]
    out = []
    for i in range(n):
        ok = (mn[i] <= mx[i + 1:]).all(1) & (mx[i] >= mn[i + 1:]).all(1)
        out += [((i + 1) << 32) | (j + i + 2) for j in np.nonzero(ok)[0]]
    return np.array(sorted(out), np.uint64)


@pytest.mark.parametrize("mode", [0, 3])
def test_pair_set_equals_brute_force(mode):
    rng = np.random.default_rng(4)
    boxes = _boxes(rng, 1500)
    boxes = np.concatenate([boxes, np.array([[-1e30] * 3 + [1e30] * 3], np.float32)])  # a plane's AABB
    bp = po.Broadphase(mode)
    bp.create(boxes)
    bp.set_all(boxes)
    got = bp.step()
    if mode == 0:  # the stored (possibly fattened) boxes are what the set is computed on
        stored = bp.proxies()[:, 7:13]
        want = _brute_pairs(stored)
    else:
        want = _brute_pairs(boxes)
    assert np.array_equal(got, want)
    assert np.array_equal(got, bp.step(brute=True))


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_destroy_and_append_proxies(mode):
    """destroyProxy removes a proxy from every pair; later createProxy calls append (uids of the remaining
    proxies unchanged): the set equals brute force over the live proxies."""
    rng = np.random.default_rng(11)
    boxes = _boxes(rng, 900)
    bp = po.Broadphase(mode) if mode != 2 else po.Broadphase(2, (-50.0,) * 3, (150.0,) * 3)
    bp.create(boxes)
    bp.set_all(boxes)
    bp.step()
    dead = rng.choice(900, 120, replace=False) + 1
    bp.destroy(dead)
    more = _boxes(rng, 200)
    bp.create(more)
    allb = np.concatenate([boxes, more])
    bp.set_all(allb)
    got = bp.step()
    lo = (got >> np.uint64(32)).astype(np.int64)
    hi = (got & np.uint64(0xffffffff)).astype(np.int64)
    assert not np.isin(lo, dead).any() and not np.isin(hi, dead).any()
    if mode == 0:
        allb = bp.proxies()[:, 7:13]
    if mode != 2:
        want = _brute_pairs(allb)
        wl = (want >> np.uint64(32)).astype(np.int64)
        wh = (want & np.uint64(0xffffffff)).astype(np.int64)
        want = want[~(np.isin(wl, dead) | np.isin(wh, dead))]
        assert np.array_equal(got, want)
    assert np.array_equal(got, bp.step(brute=True))
    assert hi.max() > 900  # appended proxies take uids after the first batch


def test_pair_filter_groups():
    rng = np.random.default_rng(9)
    boxes = _boxes(rng, 400, extent=5.0)
    group = np.where(np.arange(400) % 3 == 0, 2, 1).astype(np.int16)    # every third proxy static
    mask = np.where(group == 2, np.int16(-1 ^ 2), np.int16(-1)).astype(np.int16)
    bp = po.Broadphase(3)
    bp.create(boxes, group, mask)
    bp.set_all(boxes)
    got = bp.step()
    lo = (got >> np.uint64(32)).astype(np.int64) - 1
    hi = (got & np.uint64(0xffffffff)).astype(np.int64) - 1
    assert (lo < hi).all()
    assert not ((group[lo] == 2) & (group[hi] == 2)).any()  # static-static never pairs
    allp = _brute_pairs(boxes)
    a = (allp >> np.uint64(32)).astype(np.int64) - 1
    b = (allp & np.uint64(0xffffffff)).astype(np.int64) - 1
    keep = ~((group[a] == 2) & (group[b] == 2))
    assert np.array_equal(got, allp[keep])


def test_dbvt_fattening_state_machine():
    box = np.array([[0, 0, 0, 1, 1, 1]], np.float32)
    bp = po.Broadphase(0)
    bp.create(box)
    moved = box + np.float32([0.1, 0, -0.02, 0.1, 0, -0.02])
    bp.set_all(moved)
    p = bp.proxies()[0]
    # not contained in the old leaf: expand by 0.05, then by 2*delta on the side of motion
    want = moved[0].copy()
    want[:3] -= np.float32(0.05)
    want[3:] += np.float32(0.05)
    delta = np.float32(2.0) * ((moved[0, :3] + moved[0, 3:]) * np.float32(0.5) - (box[0, :3] + box[0, 3:]) *
                               np.float32(0.5))
    for k in range(3):
        if delta[k] > 0:
            want[3 + k] += delta[k]
        else:
            want[k] += delta[k]
    assert np.array_equal(p[7:13], want) and np.array_equal(p[1:7], want)
    # a small move inside the fattened leaf keeps the leaf and stores the raw box
    small = moved + np.float32([0.01, 0, 0, 0.01, 0, 0])
    bp.set_all(small)
    p2 = bp.proxies()[0]
    assert np.array_equal(p2[1:7], want) and np.array_equal(p2[7:13], small[0])


@pytest.mark.parametrize("mode", [1, 2])
def test_axis_sweep_quantised_overlap(mode):
    rng = np.random.default_rng(6)
    wmin, wmax = np.float32([-10, -10, -10]), np.float32([30, 30, 30])
    boxes = _boxes(rng, 800)
    bp = po.Broadphase(mode, wmin, wmax)
    bp.create(boxes)
    bp.set_all(boxes)
    got = bp.step()
    q = bp.quantized().astype(np.int64)
    mn, mx = q[:, :3], q[:, 3:]
    assert (mn % 2 == 0).all() and (mx % 2 == 1).all()  # min edges even, max edges odd
    want = []
    for i in range(len(q)):
        ok = (mn[i] < mx[i + 1:]).all(1) & (mx[i] > mn[i + 1:]).all(1)
        want += [((i + 1) << 32) | (j + i + 2) for j in np.nonzero(ok)[0]]
    assert np.array_equal(got, np.array(sorted(want), np.uint64))
    # the quantised set contains the float set
    fl = _brute_pairs(boxes)
    assert np.isin(fl, got).all()
    # the world corners quantise to the edge values
    lim = 0xfffe if mode == 1 else 0x7ffffffe
    assert q[:, :3].min() >= 0 and q[:, 3:].max() <= lim + 1


# ------------------------------------------------------------------ solver and integrator
def _single_world(bodies_spec, iterations=10, order=0, **kw):
    b = scenes._Builder("t")
    pl = b.shape(PLANE, [0, 1, 0, 0])
    b.body(pl, scenes.IDENT, np.zeros(3, np.float32), 0.0)
    for (stype, params, pos, mass) in bodies_spec:
        s = b.shape(stype, params)
        b.body(s, scenes.IDENT, np.asarray(pos, np.float32), mass, **kw)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    return po.World(shapes, verts, bodies, iterations=iterations, order=order)


def test_free_fall_is_semi_implicit_euler():
    w = _single_world([(SPHERE, [0.5], (0, 100, 0), 2.0)])
    dt = np.float32(1 / 60)
    v = np.float32(0)
    y = np.float32(100)
    g = np.float32(-9.81)
    for k in range(60):
        w.step()
        v = v + np.float32(dt * g)  # (invMass * dt) * (g / invMass) == dt * g for mass 2
        y = y + dt * v
        b = w.bodies()[1]
        assert b[13] == v and b[10] == y, k


def test_constant_spin_matches_analytic_rotation():
    w = _single_world([(BOX, [0.5, 0.2, 0.3], (0, 50, 0), 1.0)])
    b = w.bodies()
    b[1, 15:18] = [0.0, 1.0, 0.0]
    b[1, 22:25] = 0.0  # no gravity
    w.set_bodies(b)
    for k in range(1, 121):
        w.step()
    R = w.bodies()[1, :9].reshape(3, 3).astype(np.float64)
    assert np.abs(R - rot([0, 1, 0], 120 / 60.0)).max() < 2e-5


def test_resting_box_normal_impulse_balances_gravity():
    w = _single_world([(BOX, [0.5, 0.5, 0.5], (0, 0.5, 0), 3.0)])
    w.step(n=240)
    b = w.bodies()[1]
    assert np.abs(b[12:18]).max() < 2e-2
    keys, man = w.pairs()
    m = man[0]
    n = m[0].view(np.int32)
    total = m[4:].reshape(4, 32)[:n, 18].sum()
    assert abs(total - 3.0 * 9.81 / 60) < 0.05 * 3.0 * 9.81 / 60
    pts = m[4:].reshape(4, 32)[:n]
    assert (np.abs(pts[:, 19]) <= pts[:, 16] * pts[:, 18] + 1e-6).all()  # friction cone


def test_head_on_collision_without_restitution():
    b = scenes._Builder("t")
    s = b.shape(SPHERE, [0.5])
    b.body(s, scenes.IDENT, np.array([-1.0, 5, 0], np.float32), 1.0)
    b.body(s, scenes.IDENT, np.array([1.0, 5, 0], np.float32), 1.0)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    bodies[:, 22:25] = 0.0
    bodies[0, 12] = 3.0
    bodies[1, 12] = -3.0
    w = po.World(shapes, verts, bodies, iterations=10)
    w.step(n=40)
    bb = w.bodies()
    assert abs(bb[0, 12] - bb[1, 12]) < 1e-3  # relative normal velocity removed, no bounce
    assert abs(bb[0, 12] + bb[1, 12]) < 1e-5  # momentum conserved


def test_reference_order_residual_decreases():
    sc = scenes.boxes_grid(seed=2, nx=4, ny=4, nz=4, base=0.02, spacing=1.0001)
    shapes, verts, bodies = scenes.build_records(sc, po)
    w = po.World(shapes, verts, bodies, iterations=60, order=0)
    w.step(n=30)
    h = w.residual_history()
    assert len(h) > 0 and h[-1] < h[0]


def test_oracle_world_runs_clean():
    sc = scenes.mixed_pile(seed=2, n=400, bin_half=6.0)
    shapes, verts, bodies = scenes.build_records(sc, po)
    w = po.World(shapes, verts, bodies, iterations=10)
    w.step(n=60)
    b = w.bodies()
    assert np.isfinite(b).all()


def _colour_order(keys, man, bodies, reverse_within=False):
    """Greedy colouring of the manifolds on their dynamic bodies; manifolds listed colour by colour."""
    static = (bodies[:, 28].view(np.int32) & 1) != 0
    colours = []
    used = []
    for i in range(len(keys)):
        if man[i, 0].view(np.int32) == 0:
            continue
        bs = [b for b in man[i, 1:3].view(np.int32) if not static[b]]
        c = 0
        while c < len(used) and any(b in used[c] for b in bs):
            c += 1
        if c == len(used):
            used.append(set())
            colours.append([])
        used[c].update(bs)
        colours[c].append(i)
    if reverse_within:
        colours = [c[::-1] for c in colours]
    return [i for c in colours for i in c], len(colours)


def test_permutation_within_a_colour_is_exact():
    """Rows of one colour share no dynamic body, so their order cannot change a single bit."""
    sc = scenes.boxes_grid(seed=5, nx=3, ny=4, nz=3, base=0.5, spacing=1.0)
    shapes, verts, bodies = scenes.build_records(sc, po)
    worlds = [po.World(shapes, verts, bodies, iterations=10, order=2) for _ in range(2)]
    probe = po.World(shapes, verts, bodies, iterations=10, order=2)
    ncol_max = 0
    for step in range(40):
        b, (k, m) = worlds[0].bodies(), worlds[0].pairs()
        prox, stage = worlds[0].proxies()
        probe.set_bodies(b)
        probe.set_pairs(k, m)
        probe.set_proxies(prox, stage)
        probe.step_until(2)
        keys, man = probe.pairs()
        for w, rev in zip(worlds, (False, True)):
            order, ncol = _colour_order(keys, man, b, rev)
            w.set_order(order)
            w.step()
        ncol_max = max(ncol_max, ncol)
        assert np.array_equal(worlds[0].bodies().view(np.uint32), worlds[1].bodies().view(np.uint32)), step
    assert ncol_max >= 2


# ------------------------------------------------------------------ learned warm-start model reference
def test_mlp_reference_backprop_against_finite_differences():
    import mlp_reference as mr
    rng = np.random.default_rng(21)
    p = mr.init_params(rng, last_scale=1.0).astype(np.float64)
    xn = rng.normal(size=(40, 16))
    tn = rng.normal(size=(40, 3))
    g = mr.exact_grad(p, xn, tn)
    assert mr.NPARAMS == 9603 and g.shape == (9603,)
    idx = np.concatenate([rng.choice(9603, 60, replace=False), [0, 1087, 1088, 5247, 9407, 9600, 9602]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (mr.exact_loss(p + e, xn, tn) - mr.exact_loss(p - e, xn, tn)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_mlp_reference_rounding_points():
    import mlp_reference as mr
    # bf16 round to nearest even: 1 + 2^-8 is a tie and rounds to 1; 1 + 3 * 2^-8 rounds up
    assert mr.bf16(np.float32(1 + 2 ** -8)) == 1.0
    assert mr.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6
    rng = np.random.default_rng(2)
    p = mr.init_params(rng)
    norm = mr.random_norm(rng)
    x = rng.normal(size=(64, 16)).astype(np.float32)
    t = rng.normal(size=(64, 3)).astype(np.float32)
    g, loss = mr.grad(p, norm, x, t)
    n = norm.astype(np.float64)
    xn = (x.astype(np.float64) - n[:16]) * n[16:32]
    tn = (t - n[35:38]) / n[32:35]
    ge = mr.exact_grad(p.astype(np.float64), xn, tn)
    cos = g @ ge / np.linalg.norm(g) / np.linalg.norm(ge)
    assert cos > 0.999 and abs(loss - mr.exact_loss(p.astype(np.float64), xn, tn)) < 1e-2 * loss


def test_batched_worlds_never_pair_across_worlds():
    rng = np.random.default_rng(12)
    one = _boxes(rng, 300, extent=6.0)
    boxes = np.concatenate([one, one, one])            # three identical worlds, same place
    world = np.repeat(np.arange(3), 300).astype(np.int32)
    bp = po.Broadphase(3)
    bp.create(boxes)
    bp.set_worlds(world)
    bp.set_all(boxes)
    got = bp.step()
    single = _brute_pairs(one)
    want = np.sort(np.concatenate([single + np.uint64(((300 * w) << 32) + 300 * w) for w in range(3)]))
    assert np.array_equal(got, want)


def test_batched_world_steps_like_independent_worlds():
    """A batched world fed a manifold order equals each world alone fed the same order restricted to it."""
    sc = scenes.batched_towers(seed=2, n_worlds=3, lo=16, hi=40)
    shapes, verts, bodies = scenes.build_records(sc, po)
    wid = sc.world.astype(np.int32)
    comb = po.World(shapes, verts, bodies, iterations=10, order=2)
    comb.set_worlds(wid)
    probe = po.World(shapes, verts, bodies, iterations=10, order=2)
    probe.set_worlds(wid)
    sel = [np.nonzero(wid == w)[0] for w in range(3)]
    subs = [po.World(shapes, verts, bodies[s_], iterations=10, order=2) for s_ in sel]
    for step in range(60):
        b = comb.bodies()
        k, m = comb.pairs()
        prox, stage = comb.proxies()
        probe.set_bodies(b)
        probe.set_pairs(k, m)
        probe.set_proxies(prox, stage)
        probe.step_until(2)
        keys, man = probe.pairs()
        order, _ = _colour_order(keys, man, b)
        comb.set_order(order)
        comb.step()
        pw = wid[man[:, 1].view(np.int32)]
        for w, sub in enumerate(subs):
            mine = np.nonzero(pw == w)[0]
            rank = {int(i): r for r, i in enumerate(mine)}
            sub.set_order([rank[i] for i in order if i in rank])
            sub.step()
            assert np.array_equal(sub.bodies().view(np.uint32), comb.bodies()[sel[w]].view(np.uint32)), (step, w)


def test_islands_sleep_and_wake():
    """Deactivation enabled: bodies at rest for more than 2 s sleep with zero velocity and stay put;
    an active body touching a sleeping island wakes it (DiscreteDynamicsWorld.updateActivationState,
    SimulationIslandManager.buildIslands)."""
    sc = scenes.sleep_and_wake()
    shapes, verts, bodies = scenes.build_records(sc, po)
    w = po.World(shapes, verts, bodies, iterations=10, deactivation_disabled=False)
    states, pos = [], []
    for _ in range(400):
        w.step()
        b = w.bodies()
        states.append(b[:, 29].view(np.int32).copy())
        pos.append(b[:, 9:12].copy())
        asleep = states[-1] == 2
        assert np.all(b[asleep, 12:18] == 0.0)
    states = np.array(states)
    boxes = slice(1, 19)
    assert (states[:100, boxes] == 1).all()                       # moving at first (< 2 s of rest)
    first_sleep = np.argmax((states[:, boxes] == 2).all(axis=1))
    assert 100 < first_sleep < 141                                  # whole pile asleep after > 2 s
    # the sphere's island reaches the pile: sleeping boxes are woken (ISLAND_SLEEPING -> WANTS_DEACTIVATION)
    woken = (states[:-1, boxes] == 2) & (states[1:, boxes] == 3)
    assert woken[first_sleep:].any()
    for k in range(1, 400):                                         # a body asleep on both steps does not move
        still = (states[k - 1] == 2) & (states[k] == 2)
        assert np.array_equal(pos[k][still], pos[k - 1][still])
    assert (states[-1, boxes] == 2).all()                           # and the pile sleeps again


# ------------------------------------------------------------------ rayTest (CollisionWorld.java:552-591)
def _ray_world(spec):
    b = scenes._Builder("rays")
    for (stype, params, pos, mass) in spec:
        s = b.shape(stype, params)
        b.body(s, scenes.IDENT, np.asarray(pos, np.float32), mass)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    return po.World(shapes, verts, bodies, iterations=10)


def test_ray_closed_forms():
    """Sphere, box face and plane hits against their closed forms (the subsimplex cast stops within 0.01 of the
    surface: |v|^2 <= 1e-4), a miss, and the nearer of two objects winning whatever their order."""
    w = _ray_world([(PLANE, [0, 1, 0, -2], (0, 0, 0), 0.0),       # y = -2
                    (SPHERE, [1.0], (0, 0, 0), 1.0),
                    (BOX, [1.0, 1.0, 1.0], (5, 0, 0), 1.0),      # implicit 0.96 + margin 0.04: faces at 4, 6
                    (SPHERE, [0.5], (0, 0, 30), 1.0),            # farther along the last ray, lower index
                    (SPHERE, [0.5], (0, 0, 20), 1.0)])
    frm = np.array([[0, 5, 0], [10, 0, 0], [20, 5, 3], [50, 50, 50], [0, 0, 10]], np.float32)
    to = np.array([[0, -5, 0], [0, 0, 0], [20, -5, 3], [60, 60, 60], [0, 0, 40]], np.float32)
    obj, fr, nrm, pt = w.ray_test(frm, to)
    assert obj.tolist() == [1, 2, 0, -1, 4]
    assert abs(fr[0] - 0.4) < 1e-3 and np.allclose(nrm[0], [0, 1, 0], atol=1e-3)
    assert abs(fr[1] - (10 - 6.0) / 10) < 1e-3 and np.allclose(nrm[1], [1, 0, 0], atol=1e-3)
    # plane: the triangle hit is exact up to rounding; the reported normal is the unnormalised triangle normal
    assert abs(fr[2] - 0.7) < 1e-6 and np.allclose(pt[2], [20, -2, 3], atol=1e-4)
    assert nrm[2][1] > 0 and np.allclose(nrm[2] / np.linalg.norm(nrm[2]), [0, 1, 0], atol=1e-6)
    assert fr[3] == 1.0
    assert abs(fr[4] - (19.5 - 10) / 30) < 1e-3
    assert np.allclose(pt[4], [0, 0, 19.5], atol=0.02)


def test_ray_ties_and_filters():
    """Equal fractions keep the first object in world order (strict <); the callback's group/mask filter
    (needsCollision) removes static bodies (group 2) when the mask excludes them."""
    w = _ray_world([(SPHERE, [1.0], (0, 0, 0), 1.0), (SPHERE, [1.0], (0, 0, 0), 1.0), (BOX, [4, 0.1, 4], (0, -3, 0), 0.0)])
    frm = np.array([[0, 5, 0], [3, 5, 0]], np.float32)
    to = np.array([[0, -5, 0], [3, -5, 0]], np.float32)
    obj, fr, _, _ = w.ray_test(frm, to)
    assert obj.tolist() == [0, 2]
    obj2, fr2, _, _ = w.ray_test(frm, to, group=1, mask=np.int16(-1 ^ 2))
    assert obj2.tolist() == [0, -1] and fr2[0] == fr[0] and fr2[1] == 1.0


def test_ray_plane_triangle_radius_quirk():
    """StaticPlaneShape.processAllTriangles takes the radius from aabbMin only (halfExtents.set(max).set(min)):
    a ray whose local box has its min corner at the origin yields zero-size triangles and misses the plane."""
    w = _ray_world([(PLANE, [0, 1, 0, 0], (0, 0, 0), 0.0)])
    obj, fr, _, _ = w.ray_test(np.array([[0, 1, 0]], np.float32), np.array([[0, -1, 0]], np.float32))
    assert obj.tolist() == [0] and abs(fr[0] - 0.5) < 1e-6      # box min (0,-1,0): radius 0.5, hit
    obj, fr, _, _ = w.ray_test(np.array([[0, 1, 0]], np.float32), np.array([[0, 0, 0]], np.float32))
    assert obj.tolist() == [-1]                                  # ends on the plane: no sign change
    # crossing at x = 9.09, triangles of radius |(0,-1,0)|/2 around the box centre x = 5: the reference misses
    # the plane (with the box's half extents as radius it would hit)
    obj, _, _, _ = w.ray_test(np.array([[0, 10, 0]], np.float32), np.array([[10, -1, 0]], np.float32))
    assert obj.tolist() == [-1]


# ------------------------------------------------------------------ hand-computed pins of the remaining quirks
def test_shuffle_permutation_known_answer():
    """One row-order shuffle (SequentialImpulseConstraintSolver.java:805-821) of 0..9 from new Random(0): the
    draws nextInt(j + 1), j = 0..9, are 0 1 1 2 0 5 2 0 6 4 (published-value KAT above); swapping order[j] with
    order[draw] by hand gives j=2: 0 2 1 ..; j=3: 0 2 3 1 ..; j=4: 4 2 3 1 0 ..; j=6: 4 2 6 1 0 5 3 ..;
    j=7: 7 2 6 1 0 5 3 4 ..; j=8: 7 2 6 1 0 5 8 4 3 9; j=9: 7 2 6 1 9 5 8 4 3 0."""
    assert po.shuffle(0, 10).tolist() == [7, 2, 6, 1, 9, 5, 8, 4, 3, 0]
    assert po.shuffle(0, 1).tolist() == [0]
    assert po.shuffle(0, 0).tolist() == []


def _quad_manifold(points, depths):
    """A full manifold on the y = 0 plane: candidate pB = (x, -depth, z) with normal (0, 1, 0), so the stored
    localPointA is exactly (x, 0, z) (pA = pB + n * depth) and every area product below is exact in float."""
    I = tr()
    fr = [0.5, 0.5, 0.0, 0.0]
    m = _manifold()
    for (x, z), d in zip(points, depths):
        m = po.manifold_add(m, I, I, fr, _cand([0, 1, 0], [x, -d, z], d), refresh=False)
    return m


SQUARE = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]


@pytest.mark.parametrize("case", ["A", "B", "C", "E"])
def test_sort_cached_points_hand_computed(case):
    """The fifth point of a full manifold replaces the point sortCachedPoints picks (PersistentManifold.java:
    83-156): with a = new - p_i and the two remaining points' difference b, res_i = |a x b|^2 for every i but
    the deepest cached point, and closestAxis4 takes the first largest |res| (VectorUtil.java:60-90).  For
    points in the y = 0 plane, a x b = (0, a.z b.x - a.x b.z, 0).  Each case is worked out by hand:

    A  square, p0 deepest, new (2, 2):  res = (skip, 4, 16, 4)            -> 2
    B  square, p2 deepest, new (2, 2):  res = (4, 4, skip, 4), first max  -> 0   (a >= tie rule gives 3,
       not skipping the deepest point gives 2)
    C  square, new deepest, new (-1, 0.5): res = (0.25, 0.25, 0.25, 1)    -> 3   (p3 - p2 in res3 instead of
       p2 - p1 would give a four-way tie -> 0)
    E  p = (0,0) (2,1) (3,3) (0,2), new deepest at (0.625, 0.125):
       res0 = ((z-1)(-3) - (x-2)(-1))^2 = 1.25^2 = 1.5625, res1 = (x - 3z)^2 = 0.0625,
       res2 = (-2z - x)^2 = 0.765625, res3 = (z - 2x)^2 = 1.265625           -> 0   (new - p0 in res0 instead of
       new - p1 would give res0 = 0.0625 -> 3)
    """
    pts, depths, new, nd, want = {
        "A": (SQUARE, [-0.010, -0.005, -0.005, -0.005], (2.0, 2.0), -0.001, 2),
        "B": (SQUARE, [-0.005, -0.005, -0.010, -0.005], (2.0, 2.0), -0.001, 0),
        "C": (SQUARE, [-0.005] * 4, (-1.0, 0.5), -0.010, 3),
        "E": ([(0.0, 0.0), (2.0, 1.0), (3.0, 3.0), (0.0, 2.0)], [-0.005] * 4, (0.625, 0.125), -0.010, 0),
    }[case]
    m = _quad_manifold(pts, depths)
    assert m[0].view(np.int32) == 4
    before = m[4:].reshape(4, 32).copy()
    assert np.array_equal(before[:, [0, 1, 2]], np.float32([[x, 0, z] for x, z in pts]))
    x, z = new
    out = po.manifold_add(m, tr(), tr(), [0.5, 0.5, 0.0, 0.0], _cand([0, 1, 0], [x, -nd, z], nd), refresh=False)
    assert out[0].view(np.int32) == 4
    after = out[4:].reshape(4, 32)
    assert np.array_equal(after[want, [0, 1, 2]], np.float32([x, 0, z])) and after[want, 15] == np.float32(nd)
    keep = [j for j in range(4) if j != want]
    assert np.array_equal(after[keep].view(np.uint32), before[keep].view(np.uint32))


# solver records for the warm-start pins: a sphere (r 0.5, inverse mass 0.5, inverse inertia 5 on every axis,
# identity basis, centre (0, 0.5, 0)) on a static plane body; one contact point at the origin, normal (0, 1, 0)
def _sphere_on_plane(lv=(0, 0, 0), av=(0, 0, 0)):
    b = np.zeros((2, 32), np.float32)
    b[:, [0, 4, 8]] = 1.0
    b[0, 9:12] = (0.0, 0.5, 0.0)
    b[0, 12:15], b[0, 15:18] = lv, av
    b[0, 18], b[0, 19:22] = 0.5, 5.0
    b[:, 25] = 0.5
    iv = b.view(np.int32)
    iv[1, 28] = 1  # static
    iv[:, 29] = 1
    return b


def _contact(applied=0.0, lat1=0.0, lat2=0.0, dirs=None, distance=-0.001):
    m = np.zeros((1, po.MANIFOLD_WORDS), np.float32)
    m[0, :4] = np.array([1, 0, 1, 1], np.int32).view(np.float32)
    P = m[0, 4:36]
    P[12:15] = (0, 1, 0)
    P[15], P[16] = distance, 0.25
    P[18:21] = (applied, lat1, lat2)
    if dirs is not None:
        P[21:24], P[24:27] = dirs
        P[27:28] = np.array([1], np.int32).view(np.float32)
    P[28:29] = np.array([3], np.int32).view(np.float32)
    return m


F85 = np.float32(0.85)


def test_warm_start_factor_and_lateral_writeback_by_hand():
    """0 iterations: only the warm start acts (SequentialImpulseConstraintSolver.java:659-747).  Cached
    (2, 0.5, 0.25) along n = (0,1,0), d1 = (1,0,0), d2 = (0,0,1) are scaled by 0.85 and applied: r = (0,-0.5,0),
    r x n = 0, r x d1 = (0,0,0.5), r x d2 = (-0.5,0,0), so with 1/m = 0.5 and I^-1 = 5:
    v = 0.5 * 0.85 * (0.5, 2, 0.25), w = 5 * 0.85 * (-0.5 * 0.25, 0, 0.5 * 0.5).  Writeback (:892-924):
    appliedImpulse = 0.85 * 2; appliedImpulseLateral1 is stored twice, the second time with the second
    friction row, so it ends at 0.85 * 0.25; appliedImpulseLateral2 is never stored and keeps 0.25."""
    b, m, it, _ = po.solve_group(_sphere_on_plane(), _contact(2.0, 0.5, 0.25, ((1, 0, 0), (0, 0, 1))),
                                 iterations=0)
    h = np.float32(0.5)
    assert np.array_equal(b[0, 12:15], np.float32([h * (F85 * h), h * (F85 * 2), h * (F85 * np.float32(0.25))]))
    assert np.array_equal(b[0, 15:18], np.float32([np.float32(-2.5) * (F85 * np.float32(0.25)), 0,
                                                   np.float32(2.5) * (F85 * h)]))
    P = m[0, 4:36]
    assert P[18] == F85 * 2 and P[19] == F85 * np.float32(0.25) and P[20] == np.float32(0.25)
    assert np.array_equal(P[21:27], np.float32([1, 0, 0, 0, 0, 1]))  # directions kept
    assert np.array_equal(b[1], _sphere_on_plane()[1])  # the static body is never written
    # factor 1 (warm start without the 0.85 scale) would change every velocity component
    b1, _, _, _ = po.solve_group(_sphere_on_plane(), _contact(2.0, 0.5, 0.25, ((1, 0, 0), (0, 0, 1))),
                                 iterations=0, warm_factor=1.0)
    assert b1[0, 13] == np.float32(1.0)


def test_friction_directions_kept_for_the_point_lifetime():
    """Directions are computed once per point (:678-708).  A sphere sliding along +z with a point whose
    directions were set earlier to (1,0,0) and (0,0,1) keeps them: the cached lateral impulses 0.5 and 0.25
    act along x and z.  A new point (flag 0) takes d1 = normalised lateral velocity = (0,0,1) and
    d2 = normalise(d1 x n) = normalise((0,0,1) x (0,1,0)) = (-1,0,0), and stores them with the flag set;
    without lateral velocity planeSpace1((0,1,0)) gives d1 = (-1,0,0), d2 = (0,0,1) (TransformUtil.java:45-61)."""
    h, q = np.float32(0.5), np.float32(0.25)
    b, m, _, _ = po.solve_group(_sphere_on_plane(lv=(0, 0, 3)), _contact(0.0, 0.5, 0.25, ((1, 0, 0), (0, 0, 1))),
                                iterations=0)
    assert b[0, 12] == h * (F85 * h) and b[0, 14] == np.float32(3) + h * (F85 * q)
    b, m, _, _ = po.solve_group(_sphere_on_plane(lv=(0, 0, 3)), _contact(0.0, 0.5, 0.25), iterations=0)
    P = m[0, 4:36]
    assert np.array_equal(P[21:27], np.float32([0, 0, 1, -1, 0, 0])) and P[27:28].view(np.int32)[0] == 1
    assert b[0, 14] == np.float32(3) + h * (F85 * h) and b[0, 12] == -(h * (F85 * q))
    _, m, _, _ = po.solve_group(_sphere_on_plane(), _contact(), iterations=0)
    assert np.array_equal(m[0, 4 + 21:4 + 27], np.float32([-1, 0, 0, 0, 0, 1]))


def test_lateral2_never_stored_across_two_solves():
    """Solve 1 (one iteration, no cached impulses) of a sphere sliding with v = (1, 0, 0.005) on a penetrating
    contact (distance -0.001, erp 0.2, dt 1/60, friction 0.25 combined):
      normal row: jacDiagABInv = 1/(1/m) = 2, impulse = 0.001 * 0.2 * 60 * 2 = 0.024;
      friction rows: 1/(1/m + 1.25) = 1/1.75 each (r x d = 0.5, I^-1 = 5); limit 0.25 * 0.024 = 0.006;
      row 1 (x): -1 / 1.75 clamps to -0.006; row 2 (z): -0.005 / 1.75 = -0.0028571 (inside the cone).
    Writeback: lateral 1 := row 2 = -0.0028571 (not row 1's -0.006), lateral 2 stays 0.
    Solve 2 (0 iterations) warm-starts 0.85 * (0.024, -0.0028571, 0): along x v changes by
    0.5 * 0.85 * -0.0028571, along z not at all."""
    b, m, it, _ = po.solve_group(_sphere_on_plane(lv=(1, 0, 0.005)), _contact(dirs=((1, 0, 0), (0, 0, 1))),
                                 iterations=1, warm=False)
    assert it == 1
    P = m[0, 4:36]
    row2 = -0.005 / 1.75
    assert np.isclose(P[18], 0.024, rtol=1e-5) and np.isclose(P[19], row2, rtol=1e-5) and P[20] == 0.0
    assert np.isclose(b[0, 12], 1 - 0.5 * 0.006, rtol=1e-6) and np.isclose(b[0, 14], 0.005 + 0.5 * row2, rtol=1e-5)
    b2, m2, _, _ = po.solve_group(b, m, iterations=0)
    assert np.isclose(b2[0, 12] - b[0, 12], 0.5 * 0.85 * row2, rtol=1e-5)
    assert b2[0, 14] == b[0, 14]
    assert np.isclose(b2[0, 13] - b[0, 13], 0.5 * 0.85 * 0.024, rtol=1e-5)


def test_aabb_overflow_disables_simulation():
    """CollisionWorld.updateSingleAabb (CollisionWorld.java:211-227): a dynamic body whose AABB extent squared
    is not below 1e12 (here 1.6e6 m wide; also a NaN origin, for which the comparison is false) gets
    DISABLE_SIMULATION and is skipped by predictUnconstraintMotion and integrateTransforms from then on; the
    other bodies are unaffected (the 1 m box falls freely: v_y = -9.81 * dt * k)."""
    sc = scenes.aabb_overflow(nan_body=True)
    shapes, verts, bodies = scenes.build_records(sc, po)
    w = po.World(shapes, verts, bodies, iterations=10, order=0)
    w.step()
    b1 = w.bodies()
    act = b1[:, 29].view(np.int32)
    assert act[2] == 5 and act[3] == 5 and act[1] == 1
    for _ in range(9):
        w.step()
    b10 = w.bodies()
    # not moved after the first step (transform and velocities; its deactivation clock keeps running, as
    # RigidBody.updateDeactivation does for every state but ISLAND_SLEEPING and DISABLE_DEACTIVATION)
    assert np.array_equal(b10[2, :18].view(np.uint32), b1[2, :18].view(np.uint32))
    dt, g = np.float32(1 / 60), np.float32(-9.81)
    v = np.float32(0)
    for _ in range(10):
        v = v + np.float32(1.0) * dt * g  # invMass * dt * (gravity force = g / invMass)
    assert abs(b10[1, 13] - v) < 1e-5


# ------------------------------------------------------------------ Point2PointConstraint
def _pendulum(theta0=0.1, L=1.0, r=0.1, pivot=(0.0, 10.0, 0.0), iterations=10):
    """A sphere (radius r, mass 1) hanging from a world point by a Point2PointConstraint whose pivot lies on
    the sphere's body at distance L from its centre (a physical pendulum), released at theta0 from rest; the
    floor is 50 m below, out of reach."""
    b = scenes._Builder("pendulum")
    pl = b.shape(PLANE, [0, 1, 0, -50.0])
    b.body(pl, scenes.IDENT, np.zeros(3, np.float32), 0.0)
    s = b.shape(SPHERE, [r])
    pos = np.array([pivot[0] + L * np.sin(theta0), pivot[1] - L * np.cos(theta0), pivot[2]], np.float32)
    b.body(s, scenes.IDENT, pos, 1.0)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    w = po.World(shapes, verts, bodies, iterations=iterations, order=0)
    w.add_point2point([1], [-1], np.float32(pivot) - pos, np.float32(pivot))
    return w, np.float32(pivot) - pos


def test_point2point_pendulum_period_closed_form():
    """Physical pendulum: period 2 pi sqrt(I_p / (m g L)) with I_p = 2/5 m r^2 + m L^2 (parallel axes),
    times (1 + theta0^2 / 16) for the amplitude; the simulated period (zero crossings over about five periods at
    dt 1/60, 10 iterations, Baumgarte tau 0.3) lies within 2% of it, and the pivot stays within 1% of L of its
    world point (the constraint holds)."""
    L, r, th = 1.0, 0.1, 0.1
    w, pa = _pendulum(th, L, r)
    xs, drift = [], 0.0
    for _ in range(660):
        w.step()
        b = w.bodies()[1]
        R = b[:9].reshape(3, 3).astype(np.float64)
        pw = R @ pa.astype(np.float64) + b[9:12]
        drift = max(drift, np.linalg.norm(pw - np.array([0.0, 10.0, 0.0])))
        xs.append(float(b[9]))
    xs = np.array(xs)
    t = (np.arange(len(xs)) + 1) / 60.0
    cross = [t[i] + (t[i + 1] - t[i]) * xs[i] / (xs[i] - xs[i + 1]) for i in range(len(xs) - 1)
             if xs[i] > 0 >= xs[i + 1] or xs[i] < 0 <= xs[i + 1]]
    period = 2.0 * np.mean(np.diff(cross))
    want = 2 * np.pi * np.sqrt((0.4 * r * r + L * L) / (9.81 * L)) * (1 + th * th / 16)
    assert len(cross) >= 9
    assert abs(period / want - 1) < 0.02, (period, want)
    assert drift < 0.01 * L, drift


def test_point2point_joined_bodies_fall_together():
    """Two boxes joined at a shared point in free fall: gravity alone keeps the joint closed, so the solver's
    impulses stay near zero and the pivots stay together while the pair falls freely (v_y = -9.81 t)."""
    b = scenes._Builder("pair")
    pl = b.shape(PLANE, [0, 1, 0, -500.0])
    b.body(pl, scenes.IDENT, np.zeros(3, np.float32), 0.0)
    bx = b.shape(BOX, [0.5, 0.5, 0.5])
    b.body(bx, scenes.IDENT, np.array([0, 10, 0], np.float32), 1.0)
    b.body(bx, scenes.IDENT, np.array([1.0, 10, 0], np.float32), 2.0)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    w = po.World(shapes, verts, bodies, iterations=10, order=0)
    w.add_point2point([1], [2], [[0.5, 0, 0]], [[-0.5, 0, 0]])
    for k in range(60):
        w.step()
    bb = w.bodies()
    assert abs(float(bb[1, 13]) + 9.81) < 1e-3 and abs(float(bb[2, 13]) + 9.81) < 1e-3
    assert abs(float(bb[2, 9] - bb[1, 9]) - 1.0) < 1e-4
    assert np.abs(w.joint_impulses()).max() < 1e-4


def test_point2point_hanging_chain_carries_the_weights():
    """world <- box 1 (1 kg) <- box 2 (2 kg), both joints vertical: at rest each joint's impulse per step is the
    weight it carries times dt, with the sign of the push on its body A (box 1 in both): joint 1 lifts box 1 by
    (1 + 2) 9.81 / 60 = 0.4905, joint 2 pulls it down by 2 * 9.81 / 60 = 0.327; within 2%, and the chain hangs
    still (both boxes within 1 cm of where they started)."""
    b = scenes._Builder("chain")
    pl = b.shape(PLANE, [0, 1, 0, -50.0])
    b.body(pl, scenes.IDENT, np.zeros(3, np.float32), 0.0)
    bx = b.shape(BOX, [0.4, 0.4, 0.4])
    b.body(bx, scenes.IDENT, np.array([0, 9, 0], np.float32), 1.0)
    b.body(bx, scenes.IDENT, np.array([0, 8, 0], np.float32), 2.0)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    w = po.World(shapes, verts, bodies, iterations=10, order=0)
    w.add_point2point([1, 1], [-1, 2], [[0, 0.5, 0], [0, -0.5, 0]], [[0, 9.5, 0], [0, 0.5, 0]])
    for _ in range(120):
        w.step()
    bb = w.bodies()
    assert np.abs(bb[1, 9:12] - [0, 9, 0]).max() < 0.01 and np.abs(bb[2, 9:12] - [0, 8, 0]).max() < 0.01
    imp = w.joint_impulses()
    assert abs(imp[0] / (3 * 9.81 / 60) - 1) < 0.02 and abs(imp[1] / (-2 * 9.81 / 60) - 1) < 0.02, imp


def test_point2point_holds_a_body_in_contact():
    """A box resting on the ground, pinned at its centre to that world point and given 1 m/s sideways: the
    joint runs on the rigid-body velocities between the contact rows' solver bodies' write-back and read-back
    (SolverBody.writebackVelocity / readVelocity), so it holds the box while it touches the ground; friction
    alone would let it slide v^2 / (2 mu g) = 0.1 m."""
    b = scenes._Builder("pinned")
    pl = b.shape(PLANE, [0, 1, 0, 0.0])
    b.body(pl, scenes.IDENT, np.zeros(3, np.float32), 0.0)
    bx = b.shape(BOX, [0.5, 0.5, 0.5])
    b.body(bx, scenes.IDENT, np.array([0, 0.5, 0], np.float32), 1.0)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    bodies = np.asarray(bodies, np.float32).reshape(-1, po.BODY_WORDS).copy()
    bodies[1, 12] = 1.0
    w = po.World(shapes, verts, bodies, iterations=10, order=0)
    w.add_point2point([1], [-1], [[0, 0, 0]], [[0, 0.5, 0]])
    for _ in range(60):
        w.step()
        assert abs(float(w.bodies()[1, 9])) < 0.01
    assert len(w.pairs()[0]) == 1                        # still touching the ground


def test_point2point_joins_islands_for_sleep():
    """A heavy box at rest on a ledge with a small box hanging from it as a pendulum, swinging beside the ledge:
    the joint puts both in one simulation island (DiscreteDynamicsWorld.calculateSimulationIslands), so the
    resting box comes to want deactivation after 2 s but the island never sleeps while the pendulum swings."""
    b = scenes._Builder("ledge")
    lb = b.shape(BOX, [2.0, 1.0, 2.0])
    b.body(lb, scenes.IDENT, np.array([0, 1, 0], np.float32), 0.0)
    big = b.shape(BOX, [0.5, 0.5, 0.5])
    b.body(big, scenes.IDENT, np.array([1.5, 2.5, 0], np.float32), 100.0)
    small = b.shape(BOX, [0.1, 0.1, 0.1])
    b.body(small, scenes.IDENT, np.array([2.3, 1.5, 0], np.float32), 0.1)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    bodies = np.asarray(bodies, np.float32).reshape(-1, po.BODY_WORDS).copy()
    bodies[2, 14] = 2.0
    w = po.World(shapes, verts, bodies, iterations=10, order=0, deactivation_disabled=False)
    w.add_point2point([1], [2], [[0.8, 0, 0]], [[0, 1.0, 0]])
    states = []
    for _ in range(300):
        w.step()
        states.append(w.bodies()[:, 29].view(np.int32).copy())
    states = np.array(states)
    assert (states[:, 1:] != 2).all()                    # nobody sleeps
    assert (states[150:, 1] == 3).any()                  # the resting box wants to, after 2 s at rest
    assert (states[:, 2] != 3).all()                     # the swinging box never does


def test_point2point_pair_on_the_ground_balances_momentum():
    """Two boxes (1 and 2 kg) on the ground, joined between their facing sides: the joint's impulses are internal,
    so over steps a+1..b the ground's normal impulses (on the points solved, distance <= 0; the others keep an
    older step's value) equal the weight's impulse plus the pair's momentum change,
    sum = 3 * 9.81 / 60 * (b - a) + p_y(b) - p_y(a), to 1e-5.  The joint rows run between the contact rows on the
    same velocities; a joint that ran on stale rigid-body velocities would wipe the contacts' earlier iterations
    and their stored impulses would pile up about one weight per iteration."""
    b = scenes._Builder("pair_on_ground")
    pl = b.shape(PLANE, [0, 1, 0, 0.0])
    b.body(pl, scenes.IDENT, np.zeros(3, np.float32), 0.0)
    bx = b.shape(BOX, [0.5, 0.5, 0.5])
    b.body(bx, scenes.IDENT, np.array([0, 0.5, 0], np.float32), 1.0)
    b.body(bx, scenes.IDENT, np.array([1.1, 0.5, 0], np.float32), 2.0)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    w = po.World(shapes, verts, bodies, iterations=10, order=0)
    w.add_point2point([1], [2], [[0.55, 0, 0]], [[-0.55, 0, 0]])
    totals, p = [], []
    for k in range(300):
        w.step()
        keys, man = w.pairs()
        hdr = man[:, :4].view(np.int32)                  # points, body 0, body 1, pair type; 32 words a point
        assert hdr[(hdr[:, 1] != 0) & (hdr[:, 2] != 0), 0].sum() == 0    # the boxes do not touch
        totals.append(sum(float(m[4 + 32 * j + 18]) for m, h in zip(man, hdr) for j in range(h[0])
                          if m[4 + 32 * j + 15] <= 0.0))
        bb = w.bodies()
        p.append(1.0 * float(bb[1, 13]) + 2.0 * float(bb[2, 13]))
    a, b_ = 20, 299
    want = 3 * 9.81 / 60 * (b_ - a) + p[b_] - p[a]
    assert abs(sum(totals[a + 1:b_ + 1]) / want - 1) < 1e-5, (sum(totals[a + 1:b_ + 1]), want)


# ------------------------------------------------------------------ HingeConstraint
def test_atan2_fast_error_bound_and_exact_octants():
    """ScalarUtil.atan2Fast is the first-order octant approximation: exact at the eight multiples of 45 degrees and
    within 0.072 rad of atan2 everywhere (the approximation's known worst case is ~0.0711 rad, near 22.5 degrees
    from an axis); the sign follows y and the quadrant x."""
    q = np.float32(np.pi / 4)
    for (y, x), want in {(1, 1): q, (1, 0): 2 * q, (1, -1): 3 * q, (0, -1): 4 * q, (-1, -1): -3 * q, (-1, 0): -2 * q,
                         (-1, 1): -q, (0, 1): 0.0}.items():
        assert abs(po.atan2_fast(y, x) - want) <= 1e-6, (y, x, po.atan2_fast(y, x), want)
    rng = np.random.default_rng(3)
    th = rng.uniform(-np.pi, np.pi, 2000)
    got = np.array([po.atan2_fast(np.sin(t), np.cos(t)) for t in th])
    err = np.abs(np.angle(np.exp(1j * (got - th))))
    assert err.max() < 0.072 and err.max() > 0.06, err.max()


def _rot(axis, ang):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def test_hinge_frames_of_the_constructors():
    """The constructors' frames (HingeConstraint.java:84-238) for a body turned 30 degrees about y: A's frame has
    the hinge axis (in A) as its third column and the pivot as its origin; its first two columns are the world x
    axis of body A made orthogonal to the axis by two cross products and, as in the reference, not renormalised:
    both have length s = sqrt(1 - (axis . x_A)^2), so the basis is diag(s^2, s^2, 1)-orthogonal with determinant
    s^2.  B's third column is -axisInB (flipped), and for aligned axes the hinge angle at construction is 0.  The
    fixed-body forms put B's frame in world space: origin = A's transform of the pivot, third column =
    -(A's basis * axis)."""
    R = _rot([0, 1, 0], np.pi / 6).astype(np.float32)
    tr = np.concatenate([R.reshape(-1), [1.0, 2.0, 3.0]]).astype(np.float32)
    axis = np.array([0.0, 0.6, 0.8], np.float32)
    fa, fb = po.hinge_frames(tr, [0.5, 0, 0], axis, [-0.5, 0, 0], axis)
    Ba, Bb = fa[:9].reshape(3, 3).astype(np.float64), fb[:9].reshape(3, 3).astype(np.float64)
    s2 = 1.0 - float(axis @ R[:, 0]) ** 2
    assert np.allclose(Ba.T @ Ba, np.diag([s2, s2, 1.0]), atol=1e-6) and abs(np.linalg.det(Ba) - s2) < 1e-6
    assert np.allclose(Ba[:, 2], axis, atol=1e-7) and np.allclose(fa[9:], [0.5, 0, 0])
    assert np.allclose(Bb[:, 2], -axis, atol=1e-7) and np.allclose(fb[9:], [-0.5, 0, 0])
    # B's first two columns are A's carried along the (identity) shortest arc, the second as axis x first
    assert np.allclose(Bb[:, 0], Ba[:, 0], atol=1e-6) and np.allclose(Bb[:, 1], np.cross(axis, Ba[:, 0]), atol=1e-6)
    fa2, fb2 = po.hinge_frames(tr, [0.0, 0.5, 0.0], axis)
    assert np.allclose(fb2[9:], R @ np.float32([0, 0.5, 0]) + tr[9:], atol=1e-6)
    assert np.allclose(fb2[:9].reshape(3, 3)[:, 2], -(R @ axis), atol=1e-6)
    # frame constructors: B's third column flipped; fixed body: B = A's frame, origin moved to world space
    F = np.concatenate([_rot([1, 2, 3], 0.7).reshape(-1), [0, 0.5, 0]]).astype(np.float32)
    ga, gb = po.hinge_frames_from_frames(tr, F, F)
    assert np.array_equal(ga, F) and np.array_equal(gb[[0, 1, 3, 4, 6, 7]], F[[0, 1, 3, 4, 6, 7]])
    assert np.array_equal(gb[[2, 5, 8]], -F[[2, 5, 8]])
    ga, gb = po.hinge_frames_from_frames(tr, F)
    assert np.allclose(gb[9:], R @ F[9:] + tr[9:], atol=1e-6)


def _hinge_world(L=1.0, half=(0.1, 0.1, 0.1), axis=(0.0, 0.0, 1.0), w0=(0.0, 0.0, 0.0), params=None, tilt=0.0,
                 iterations=10, frames=False):
    """A box of 1 kg hinged to the world (the fixed body) at a pivot L above its centre, about `axis` (in the
    box's frame, which starts aligned with the world), released at `tilt` radians about the axis with angular
    velocity w0 and the matching centre velocity w0 x (centre - pivot); the floor is far below.  frames: built with
    the (rbA, rbAFrame) constructor (frame = identity basis at the pivot; axis z), else (rbA, pivotInA, axisInA)."""
    b = scenes._Builder("hinge")
    pl = b.shape(PLANE, [0, 1, 0, -50.0])
    b.body(pl, scenes.IDENT, np.zeros(3, np.float32), 0.0)
    bx = b.shape(BOX, list(half))
    pivot = np.array([0.0, 10.0, 0.0])
    Rt = _rot(axis, tilt)
    pos = pivot + Rt @ np.array([0.0, -L, 0.0])
    b.body(bx, Rt.astype(np.float32).reshape(-1), pos.astype(np.float32), 1.0)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    bodies = np.asarray(bodies, np.float32).reshape(-1, po.BODY_WORDS).copy()
    bodies[1, 15:18] = w0
    bodies[1, 12:15] = np.cross(np.float64(w0), pos - pivot)
    w = po.World(shapes, verts, bodies, iterations=iterations, order=0)
    if frames:
        F = np.concatenate([np.eye(3).reshape(-1), [0.0, L, 0.0]]).astype(np.float32)
        fa, fb = po.hinge_frames_from_frames(bodies[1, :12], F)
    else:
        fa, fb = po.hinge_frames(bodies[1, :12], [0.0, L, 0.0], np.float32(axis))
    w.add_hinge([1], [-1], fa[None], fb[None], None if params is None else np.float32(params)[None])
    return w


def test_hinge_pendulum_period_and_plane():
    """A box pendulum on a hinge about z: period 2 pi sqrt(I_p / (m g L)) (I_p = box inertia about z + m L^2) times
    (1 + theta0^2 / 16) within 2%; the motion stays in the hinge plane (|v_z|, |w_x|, |w_y| below 1e-3) and the
    pivot within 1% of L of its world point."""
    L, h, th = 1.0, 0.1, 0.15
    w = _hinge_world(L=L, half=(h, h, h), tilt=th)
    xs, off = [], 0.0
    for _ in range(660):
        w.step()
        b = w.bodies()[1]
        R = b[:9].reshape(3, 3).astype(np.float64)
        off = max(off, np.linalg.norm(R @ [0, L, 0] + b[9:12] - [0, 10, 0]))
        assert abs(b[14]) < 1e-3 and abs(b[15]) < 1e-3 and abs(b[16]) < 1e-3, b[12:18]
        xs.append(float(b[9]))
    xs = np.array(xs)
    t = (np.arange(len(xs)) + 1) / 60.0
    cross = [t[i] + (t[i + 1] - t[i]) * xs[i] / (xs[i] - xs[i + 1]) for i in range(len(xs) - 1)
             if xs[i] * xs[i + 1] < 0]
    period = 2.0 * np.mean(np.diff(cross))
    Ip = (2 * h) ** 2 * 2 / 12.0 + L * L
    want = 2 * np.pi * np.sqrt(Ip / (9.81 * L)) * (1 + th * th / 16)
    assert len(cross) >= 9 and abs(period / want - 1) < 0.02, (period, want)
    assert off < 0.01 * L, off


def test_hinge_removes_off_axis_spin_and_keeps_axis_spin():
    """A wheel hinged at its centre about z, spun at 5 rad/s about z and 3 rad/s about x: gravity exerts no torque
    about the pivot, the hinge takes out the x spin within the first step and leaves the z spin at 5 rad/s (to
    1e-4) for two seconds."""
    w = _hinge_world(L=0.0, half=(0.5, 0.5, 0.1), w0=(3.0, 0.0, 5.0))
    for k in range(120):
        w.step()
        b = w.bodies()[1]
        assert abs(b[15]) < 1e-3 and abs(b[16]) < 1e-3, (k, b[15:18])
        assert abs(b[17] - 5.0) < 1e-4, (k, b[17])


def test_hinge_limit_holds_the_angle():
    """A pendulum on a hinge (frame constructor, hinge angle 0 at rest) with limits [-0.3, 0.3] rad (softness 0.9),
    kicked at 4 rad/s: unlimited it swings past 1 rad; with the limit its orientation about the axis stays below
    0.3 rad and reaches 0.2.  The limit acts on atan2Fast's angle, which reads 0.3 at a true 0.232 rad
    (r = (1 - 0.3 / (pi / 4)) solves (cos - sin) / (cos + sin) = r), and the bias lets it overshoot a little."""
    w = _hinge_world(L=1.0, w0=(0.0, 0.0, 4.0), params=(-0.3, 0.3, 0.9, 0.3, 1.0, 0, 0, 0, 0), frames=True)
    assert w.hinge_angle(0) == 0.0
    free = _hinge_world(L=1.0, w0=(0.0, 0.0, 4.0), frames=True)
    amax = fmax = 0.0
    for _ in range(120):
        w.step()
        free.step()
        R = w.bodies()[1, :9].reshape(3, 3)
        amax = max(amax, abs(float(np.arctan2(R[1, 0], R[0, 0]))))
        R = free.bodies()[1, :9].reshape(3, 3)
        fmax = max(fmax, abs(float(np.arctan2(R[1, 0], R[0, 0]))))
    assert fmax > 1.0, fmax
    assert 0.2 < amax < 0.3, amax


def test_hinge_motor_reaches_its_target():
    """A wheel hinged at its centre with the angular motor on (target 2 rad/s): with a large impulse cap it spins
    at 2 rad/s (to 1e-4) after the first step; with a cap of 0.001 per iteration its spin after one step of 10
    iterations is at most 10 * 0.001 / I_zz (the motor impulse is clipped each iteration)."""
    w = _hinge_world(L=0.0, half=(0.5, 0.5, 0.1), params=(1e30, -1e30, 0.9, 0.3, 1.0, 0, 1, 2.0, 100.0))
    for _ in range(3):
        w.step()
    assert abs(w.bodies()[1, 17] - 2.0) < 1e-4, w.bodies()[1, 15:18]
    w = _hinge_world(L=0.0, half=(0.5, 0.5, 0.1), params=(1e30, -1e30, 0.9, 0.3, 1.0, 0, 1, 2.0, 0.001))
    w.step()
    Izz = (1.0 + 1.0) / 12.0
    assert 0.0 < w.bodies()[1, 17] <= 10 * 0.001 / Izz * 1.0001, w.bodies()[1, 17]


def test_hinge_angle_of_the_fixed_body_axis_constructor_is_pi():
    """(rbA, pivotInA, axisInA) with A unrotated and the axis along z: B's axis is -z, so the shortest arc from z
    to -z takes the "pick any vector" branch, the quaternion (0, 1, 0, 0), which turns A's reference axis y into
    itself; B's swing axis is then (-z) x y = +x, A's second reference axis z x y = -x, and the hinge angle at
    rest is atan2Fast(0, -1) = pi (as in the reference, so limits about 0 act at once)."""
    w = _hinge_world(L=1.0)
    assert w.hinge_angle(0) == np.float32(np.pi)


def test_hinge_realigns_a_tilted_axis():
    """A wheel hinged at its centre about z whose body is then tilted 0.2 rad about x: the angular error term
    (-(axisA x axisB) / dt, scaled by the inverse angular mass) turns the axle back; after one second the body's
    z axis is within 0.01 rad of the world's, and it never tilts further than at the start."""
    w = _hinge_world(L=0.0, half=(0.5, 0.5, 0.1))
    b = w.bodies()
    b[1, :9] = _rot([1, 0, 0], 0.2).astype(np.float32).reshape(-1)
    w.set_bodies(b)
    for _ in range(60):
        w.step()
        tilt = np.arccos(min(1.0, float(w.bodies()[1, :9].reshape(3, 3)[2, 2])))
        assert tilt < 0.2 + 1e-4, tilt
    assert tilt < 0.01, tilt


# ------------------------------------------------------------------ convex sweeps and CCD
def _tr12(o, R=None):
    return np.concatenate([(np.eye(3) if R is None else np.asarray(R)).reshape(-1), o]).astype(np.float32)


def test_calculate_velocity_closed_form():
    """TransformUtil.calculateVelocity: from A to R(axis, angle)·A over dt the angular velocity w is a rotation
    vector of R: rot(w/|w|, |w| dt) = R with |w| dt in [0, 2pi]; while the trace of R is not negative the difference
    quaternion comes out with w >= 0 and the vector is axis·angle/dt itself (beyond that it may be the equivalent
    turn 2pi - angle about -axis, as the reference gives no sign to the quaternion). The linear velocity is the
    origin difference over dt."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        ang = rng.uniform(0.1, 2.5)
        A = random_rot(rng)
        o0, o1 = rng.uniform(-5, 5, 3), rng.uniform(-5, 5, 3)
        dt = 0.25
        lv, av = po.calculate_velocity(_tr12(o0, A), _tr12(o1, rot(a, ang) @ A), dt)
        assert np.allclose(lv, (np.float32(o1) - np.float32(o0)) / dt, rtol=1e-6, atol=1e-5)
        th = float(np.linalg.norm(av)) * dt
        assert 0.0 <= th <= 2 * np.pi + 1e-5
        assert np.abs(rot(av / np.linalg.norm(av), th) - rot(a, ang)).max() < 5e-5
        if 1 + 2 * np.cos(ang) >= 1e-3:
            assert np.abs(av * dt - a * ang).max() < 5e-5, (av * dt, a * ang)
    lv, av = po.calculate_velocity(_tr12([0, 0, 0]), _tr12([1, 0, 0]), 1.0)
    assert np.array_equal(av, [0, 0, 0])  # no rotation: the axis falls back to x with angle 0


def test_temporal_aabb_closed_form():
    """CollisionShape.calculateTemporalAabb: the shape's AABB, its max side grown by positive and its min side by
    negative linear motion, then both by |w| dt times the angular motion disc (half the diagonal of the AABB at the
    identity plus the distance of its centre)."""
    box = po.make_shape(BOX, [1.0, 2.0, 3.0])
    mn, mx = po.temporal_aabb(box, _tr12([1, 1, 1]), [2.0, -4.0, 0.0], [0.0, 3.0, 4.0], 0.5)
    am = 5.0 * np.sqrt(1 + 4 + 9) * 0.5
    assert np.allclose(mn, np.array([0, -1, -2]) + [0, -2, 0] - am, atol=1e-5)
    assert np.allclose(mx, np.array([2, 3, 4]) + [1, 0, 0] + am, atol=1e-5)
    # a hull off its origin: the disc adds the distance of the AABB centre to the half diagonal
    pts = np.array([[x, y, z] for x in (2, 4) for y in (-1, 1) for z in (-1, 1)], np.float32)
    hull = po.make_shape(HULL, [0, 8], pts)
    a = po.aabb(hull, _tr12([0, 0, 0])).astype(np.float64)
    lo, hi = a[:3], a[3:]
    disc = np.linalg.norm(hi - lo) * 0.5 + np.linalg.norm((hi + lo) * 0.5)
    mn, mx = po.temporal_aabb(hull, _tr12([0, 0, 0]), [0.0, 0.0, 0.0], [0.0, 0.0, 2.0], 1.0)
    assert np.allclose(mn, lo - 2 * disc, atol=1e-5) and np.allclose(mx, hi + 2 * disc, atol=1e-5)


def test_gjk_convex_cast_closed_forms():
    """GjkConvexCast: a sphere (r 0.5) swept along x from 0 to 10 meets a box of half-extent 1 centred at x = 5 when
    its centre reaches 3.5: fraction 0.35, normal -x, contact point on the box face (4, 0, 0); swept away, or past
    the box with a lateral offset larger than the sum of the half sizes, it reports nothing; against a sphere of
    radius 1 at 5 the fraction is (5 - 1.5) / 10."""
    sph = po.make_shape(SPHERE, [0.5])
    box = po.make_shape(BOX, [1.0, 1.0, 1.0])
    hit, f, n, p = po.gjk_convex_cast(sph, _tr12([0, 0, 0]), _tr12([10, 0, 0]), box, _tr12([5, 0, 0]))
    assert hit and abs(f - 0.35) < 1e-4 and np.allclose(n, [-1, 0, 0], atol=1e-5) and np.allclose(p, [4, 0, 0],
                                                                                                    atol=2e-3)
    assert not po.gjk_convex_cast(sph, _tr12([0, 0, 0]), _tr12([-10, 0, 0]), box, _tr12([5, 0, 0]))[0]
    assert not po.gjk_convex_cast(sph, _tr12([0, 1.6, 0]), _tr12([10, 1.6, 0]), box, _tr12([5, 0, 0]))[0]
    # starting within the cast's 0.001 tolerance of the face: moving in hits at fraction 0, moving away (or
    # along the face) does not, as the motion along the normal is not below -allowedPenetration
    hit, f, n, p = po.gjk_convex_cast(sph, _tr12([3.5 - 5e-4, 0, 0]), _tr12([5.5, 0, 0]), box, _tr12([5, 0, 0]))
    assert hit and f == 0.0 and np.allclose(n, [-1, 0, 0], atol=1e-5)
    assert not po.gjk_convex_cast(sph, _tr12([3.5 - 5e-4, 0, 0]), _tr12([1.5, 0, 0]), box, _tr12([5, 0, 0]))[0]
    assert not po.gjk_convex_cast(sph, _tr12([3.5 - 5e-4, 0, 0]), _tr12([3.5 - 5e-4, 0.5, 0]), box,
                                  _tr12([5, 0, 0]))[0]
    big = po.make_shape(SPHERE, [1.0])
    hit, f, n, p = po.gjk_convex_cast(sph, _tr12([0, 0, 0]), _tr12([10, 0, 0]), big, _tr12([5, 0, 0]))
    assert hit and abs(f - 0.35) < 1e-4


def _sweep_world():
    """Static boxes of half-extent 1 at x = 5, 8 and 11 (the middle one in filter group 2 only as every static
    body), a dynamic box above the line, no floor plane."""
    b = scenes._Builder("sweep")
    bx = b.shape(BOX, [1.0, 1.0, 1.0])
    for x in (11.0, 5.0, 8.0):
        b.body(bx, scenes.IDENT, np.array([x, 0, 0], np.float32), 0.0)
    b.body(bx, scenes.IDENT, np.array([6.5, 5, 0], np.float32), 1.0)
    sph = b.shape(SPHERE, [0.5])
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    return po.World(shapes, verts, bodies), sph


def test_convex_sweep_closest_hit_and_filters():
    """convexSweepTest with a closest-hit callback: the sphere swept from x = 0 to 20 stops at the nearest box (list
    order does not matter: the box at 5 is the second object), fraction 0.175; a callback whose mask excludes the
    static group sees nothing; a sweep above the boxes passes below the dynamic box only when it clears it."""
    w, sph = _sweep_world()
    obj, fr, n, p = w.convex_sweep(sph, _tr12([0, 0, 0])[None], _tr12([20, 0, 0])[None])
    assert obj[0] == 1 and abs(fr[0] - 0.175) < 1e-4 and np.allclose(n[0], [-1, 0, 0], atol=1e-5)
    obj, fr, n, p = w.convex_sweep(sph, _tr12([0, 0, 0])[None], _tr12([20, 0, 0])[None], group=1, mask=1)
    assert obj[0] == -1 and fr[0] == 1.0
    obj, fr, n, p = w.convex_sweep(sph, _tr12([0, 5, 0])[None], _tr12([20, 5, 0])[None])
    assert obj[0] == 3 and abs(fr[0] - (6.5 - 1 - 0.5) / 20) < 1e-4
    obj, fr, n, p = w.convex_sweep(sph, _tr12([0, 6.6, 0])[None], _tr12([20, 6.6, 0])[None])
    assert obj[0] == -1


def test_ccd_stops_a_fast_sphere_at_a_thin_wall():
    """A sphere (r 0.1) at 300 m/s towards a static wall 0.1 m thick 2.5 m away moves 5 m per step: without CCD it
    passes through the wall in the first step; with a motion threshold of 0.1 and a swept sphere of radius 0.1 the
    step is clamped at the time of impact (the sphere's centre just short of the wall face at 2.45 - 0.1) and the
    sphere never gets past the wall."""
    def world(ccd):
        b = scenes._Builder("wall")
        wall = b.shape(BOX, [0.05, 2.0, 2.0])
        b.body(wall, scenes.IDENT, np.array([2.5, 0, 0], np.float32), 0.0)
        s = b.shape(SPHERE, [0.1])
        b.body(s, scenes.IDENT, np.array([0, 0, 0], np.float32), 1.0)
        sc = b.done()
        shapes, verts, bodies = scenes.build_records(sc, po)
        bodies = np.asarray(bodies, np.float32).reshape(-1, po.BODY_WORDS).copy()
        bodies[1, 12] = 300.0
        w = po.World(shapes, verts, bodies)
        if ccd:
            w.set_ccd([1], 0.1, 0.1)
        return w
    w = world(False)
    w.step()
    assert w.bodies()[1, 9] > 4.0
    w = world(True)
    w.step()
    x = float(w.bodies()[1, 9])
    assert 2.3 < x < 2.35 + 1e-3, x
    for _ in range(30):
        w.step()
        assert w.bodies()[1, 9] < 2.45, w.bodies()[1, 9]


def _ccd_world(bodies_spec, ccd_bodies, radius):
    """bodies_spec: (shape type, params, x, vx, mass) per body in list order, no floor."""
    b = scenes._Builder("ccd")
    for st, par, x, vx, mass in bodies_spec:
        s = b.shape(st, par)
        b.body(s, scenes.IDENT, np.array([x, 0, 0], np.float32), mass)
    sc = b.done()
    shapes, verts, bodies = scenes.build_records(sc, po)
    bodies = np.asarray(bodies, np.float32).reshape(-1, po.BODY_WORDS).copy()
    for i, spec in enumerate(bodies_spec):
        bodies[i, 12] = spec[3]
    w = po.World(shapes, verts, bodies)
    w.set_ccd(ccd_bodies, 0.1, radius)
    return w


def test_ccd_skips_bodies_it_already_touches():
    """Two boxes (half 0.5) 0.01 apart, both at 300 m/s along x: the pair has a contact point (within the 0.02
    breaking threshold) and no impulse, so the rear box's sweep skips the front box and both move 5 m; the same
    boxes 0.03 apart have no contact point, the sweep meets the front box at its old place and clamps the rear box
    to about 0.03 m of travel."""
    for gap, far in ((0.01, True), (0.03, False)):
        w = _ccd_world([(BOX, [0.5, 0.5, 0.5], 0.0, 300.0, 1.0), (BOX, [0.5, 0.5, 0.5], 1.0 + gap, 300.0, 1.0)],
                       [0], 0.5)
        w.step()
        x = float(w.bodies()[0, 9])
        assert (x > 4.9) if far else (0.0 < x < 0.05), (gap, x)


def test_ccd_sweeps_earlier_bodies_at_their_new_place():
    """integrateTransforms moves the bodies one after another: a CCD sphere behind a sphere that moves away at the
    same speed passes freely when the front sphere comes first in the body list (already moved), and stops at
    the front sphere's old place (centres 1 m apart: a fifth of its 5 m step) when it comes later."""
    sp = (SPHERE, [0.5])
    w = _ccd_world([sp + (2.0, 300.0, 1.0), sp + (0.0, 300.0, 1.0)], [1], 0.5)
    w.step()
    assert float(w.bodies()[1, 9]) > 4.9
    w = _ccd_world([sp + (0.0, 300.0, 1.0), sp + (2.0, 300.0, 1.0)], [0], 0.5)
    w.step()
    x = float(w.bodies()[0, 9])
    assert abs(x - 1.0) < 0.01, x
