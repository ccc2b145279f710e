"""CPU checks of the C ABI (no GPU): the library loads, exports exactly what
include/jbgpu.h declares, and its host-side record builders agree bit for bit
with the oracle's.  No compute entry point is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import jbgpu
import pyoracle as po
from jbgpu import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "jbgpu.h")).read()
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(jbgpu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_set():
    assert declared_symbols() == sorted(jbgpu.SYMBOLS)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(jbgpu.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    out = os.popen(f"nm -D --defined-only {jbgpu.LIB_PATH}").read()
    exported = set(re.findall(r"\bT (jbgpu_\w+)", out))
    assert exported == set(declared_symbols())


def test_version_and_default_config():
    L = jbgpu.load()
    assert L.jbgpu_version() >= 1
    cfg = jbgpu.default_config()
    assert cfg.bp_mode == jbgpu.BP_DBVT and cfg.num_iterations == 10
    assert cfg.solver_order == jbgpu.ORDER_COLOR and cfg.deactivation_disabled == 1
    assert np.isclose(cfg.erp, 0.2) and np.isclose(cfg.warmstart_factor, 0.85)


def test_argument_errors_are_status_codes():
    rec = np.zeros(jbgpu.SHAPE_WORDS, np.float32)
    L = jbgpu.load()
    assert L.jbgpu_make_shape(99, np.zeros(4, np.float32), None, rec) == -1          # JBGPU_ERR_ARG
    assert L.jbgpu_make_shape(3, np.float32([0, 0, 0, 0]), None, rec) == -1         # hull without points
    buf = C.create_string_buffer(64)
    assert L.jbgpu_last_error(None, buf, 64) >= 0


def test_reference_solver_order_is_rejected():
    """The device solves in colour batches only: asking for the island order with the Random(0) shuffle is an
    argument error (checked before any device is touched)."""
    with pytest.raises(jbgpu.JbgpuError, match=r"\(-1\).*graph-colour"):
        jbgpu.Context(solver_order=jbgpu.ORDER_REFERENCE)


def _shape_cases():
    rng = np.random.default_rng(1)
    pts = rng.normal(size=(20, 3)).astype(np.float32)
    return [(0, [0.37], None), (1, [0.5, 0.25, 1.5], None), (2, [0.3, 1.2], None),
            (3, [0, 20], pts), (4, [0.0, 1.0, 0.0, -2.0], None), (1, [0.1, 0.1, 0.1], None)]


@pytest.mark.parametrize("case", range(6))
def test_make_shape_matches_oracle(case):
    t, params, verts = _shape_cases()[case]
    a = jbgpu.make_shape(t, params, verts)
    b = po.make_shape(t, params, verts)
    assert a.view(np.uint32).tolist() == b.view(np.uint32).tolist()
    if t != 4:
        ia = jbgpu.shape_inertia(a, 2.5, verts)
        ib = po.shape_inertia(b, 2.5, verts)
        assert ia.view(np.uint32).tolist() == ib.view(np.uint32).tolist()


@pytest.mark.parametrize("gen", ["boxes_grid", "mixed_pile", "batched_towers", "pyramids"])
def test_scene_records_match_oracle(gen):
    kw = dict(boxes_grid=dict(nx=4, ny=3, nz=2), mixed_pile=dict(n=500, bin_half=8.0),
              batched_towers=dict(n_worlds=8, lo=16, hi=40), pyramids=dict(n_boxes=300, n_pyramids=2, n_spheres=20,
                                                                           area=30.0))[gen]
    sc = getattr(scenes, gen)(seed=3, **kw)
    s1, v1, b1 = scenes.build_records(sc, jbgpu)
    s2, v2, b2 = scenes.build_records(sc, po)
    for x, y in ((s1, s2), (v1, v2), (b1, b2)):
        assert x.shape == y.shape and np.array_equal(x.view(np.uint32), y.view(np.uint32))


def _pair_records():
    """Three well-formed pairs of a 5-body world as the device keeps them: keys (uid_lo << 32 | uid_hi),
    manifold headers (count, body0, body1, type); the last is a convex-plane pair listing the convex body first."""
    keys = np.array([(1 << 32) | 2, (1 << 32) | 4, (3 << 32) | 5], np.uint64)
    man = np.zeros((3, jbgpu.MANIFOLD_WORDS), np.float32)
    hdr = np.array([[2, 0, 1, 2], [0, 0, 3, 0], [1, 4, 2, 1]], np.int32)
    man[:, :4] = hdr.view(np.float32)
    return keys, man


def test_pair_record_check_accepts_well_formed_records():
    keys, man = _pair_records()
    jbgpu.check_pair_records(5, keys, man)
    jbgpu.check_pair_records(5, keys[:0], man[:0])


@pytest.mark.parametrize("case,reason", [
    ("count5", "point count"), ("count_neg", "point count"), ("type", "pair type"), ("body", "bodies do not match"),
    ("unsorted", "strictly increasing"), ("dup", "strictly increasing"), ("uid_range", "out of range"),
    ("uid_zero", "out of range"), ("self", "uid_lo < uid_hi")])
def test_pair_record_check_rejects_each_bad_record(case, reason):
    """Every malformed record a caller could hand to jbgpu_world_set_pairs is refused on the host, before any
    of it reaches a kernel (a point count above 4 or a foreign body index would index past device buffers)."""
    keys, man = _pair_records()
    h = man[:, :4].view(np.int32)
    if case == "count5":
        h[0, 0] = 5
    elif case == "count_neg":
        h[1, 0] = -1
    elif case == "type":
        h[2, 3] = 7
    elif case == "body":
        h[1, 2] = 2
    elif case == "unsorted":
        keys = keys[[1, 0, 2]].copy()
        man = man[[1, 0, 2]].copy()
    elif case == "dup":
        keys[1] = keys[0]
        man[1] = man[0]
    elif case == "uid_range":
        keys[2] = (3 << 32) | 6
    elif case == "uid_zero":
        keys[0] = 2
        h[0, 1:3] = [-1, 1]
    elif case == "self":
        keys[2] = (5 << 32) | 5
    with pytest.raises(jbgpu.JbgpuError, match=reason):
        jbgpu.check_pair_records(5, keys, man)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]], np.float32)


def test_hinge_frames_match_oracle():
    """jbgpu_hinge_frames (host side, no device) builds the same frames as the oracle's restatement of the
    HingeConstraint constructors, bit for bit: random transforms and axes, the axis equal to / opposite to A's x
    axis (the constructor's two special cases), antiparallel A and B axes (the shortest arc's fallback), and the
    frame forms."""
    rng = np.random.default_rng(7)
    for k in range(200):
        R = _random_rotation(rng) if k % 4 else np.eye(3, dtype=np.float32)
        tr = np.concatenate([R.reshape(-1), rng.uniform(-5, 5, 3)]).astype(np.float32)
        ax = rng.normal(size=3).astype(np.float32)
        ax /= np.linalg.norm(ax)
        if k % 10 == 1:
            ax = R[:, 0].copy()
        elif k % 10 == 2:
            ax = -R[:, 0]
        bx = -ax if k % 10 == 3 else ax
        pa, pb = rng.uniform(-1, 1, 3).astype(np.float32), rng.uniform(-1, 1, 3).astype(np.float32)
        got = jbgpu.hinge_frames(tr, pa, ax, pb, bx)
        want = po.hinge_frames(tr, pa, ax, pb, bx)
        assert all(np.array_equal(g.view(np.uint32), w.view(np.uint32)) for g, w in zip(got, want)), k
        got = jbgpu.hinge_frames(tr, pa, ax)
        want = po.hinge_frames(tr, pa, ax)
        assert all(np.array_equal(g.view(np.uint32), w.view(np.uint32)) for g, w in zip(got, want)), k
        F = np.concatenate([_random_rotation(rng).reshape(-1), pa]).astype(np.float32)
        G = np.concatenate([_random_rotation(rng).reshape(-1), pb]).astype(np.float32)
        for fb in (G, None):
            got = jbgpu.hinge_frames(tr, frame_a=F, frame_b=fb)
            want = po.hinge_frames_from_frames(tr, F, fb)
            assert all(np.array_equal(g.view(np.uint32), w.view(np.uint32)) for g, w in zip(got, want)), k
