"""jbgpu — Python host binding of libjbgpu (the gfx950 rigid-body hot path).

The product is the C ABI in ``include/jbgpu.h``; this module is a thin ctypes
mirror of it with numpy buffers, arranged after the reference's plugin seams:
:class:`Context` exposes the BroadphaseInterface entry points (``bp_*``), the
fused DynamicsWorld step (``world_*``) and the ConstraintSolver hooks
(``solver_*``).  It raises :class:`JbgpuError` on any non-zero status; there is
no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(PKG_DIR), "libjbgpu.so")

SHAPE_WORDS, BODY_WORDS, PROXY_WORDS, POINT_WORDS, MANIFOLD_WORDS = 20, 32, 16, 32, 132
BP_DBVT, BP_SAP16, BP_SAP32, BP_SIMPLE = 0, 1, 2, 3
ORDER_REFERENCE, ORDER_COLOR, ORDER_EXPLICIT = 0, 1, 2
TIMER_REGIONS = 16
TIMER_NAMES = ["predict", "setaabb", "traverse", "gjk", "manifold", "solver_setup", "solver_iter", "integrate",
               "step", "epa", "remap", "mlp", "bvh", "keysort", "bin", "color"]
MLP_PARAMS, MLP_NORM_WORDS = 9603, 38
MLP_STATIC_MAX = 65536

# every exported symbol of include/jbgpu.h: (restype, argtypes)
_f = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class Config(C.Structure):
    _fields_ = [("bp_mode", C.c_int32), ("world_min", C.c_float * 3), ("world_max", C.c_float * 3),
                ("num_iterations", C.c_int32), ("solver_order", C.c_int32), ("deactivation_disabled", C.c_int32),
                ("tolerance", C.c_float), ("max_iterations", C.c_int32), ("warmstart", C.c_int32),
                ("erp", C.c_float), ("warmstart_factor", C.c_float)]


class Stats(C.Structure):
    _fields_ = [("pairs", C.c_int64), ("contacts", C.c_int64), ("manifolds", C.c_int64), ("colors", C.c_int64),
                ("epa_calls", C.c_int64), ("epa_fails", C.c_int64), ("iterations", C.c_int64),
                ("residual", C.c_float), ("huge_proxies", C.c_int64), ("rows", C.c_int64),
                ("epa_retries", C.c_int64), ("epa_huge", C.c_int64), ("graph_redos", C.c_int64)]


SYMBOLS = {
    "jbgpu_version": (C.c_int, []),
    "jbgpu_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "jbgpu_config_default": (None, [C.POINTER(Config)]),
    "jbgpu_ctx_create": (C.c_int, [C.POINTER(Config), C.POINTER(_vp)]),
    "jbgpu_ctx_destroy": (C.c_int, [_vp]),
    "jbgpu_last_error": (C.c_int, [_vp, C.c_char_p, C.c_int32]),
    "jbgpu_make_shape": (C.c_int, [C.c_int32, _f, _vp, _f]),
    "jbgpu_shape_local_inertia": (C.c_int, [_f, _vp, C.c_float, _f]),
    "jbgpu_bp_create_proxies": (C.c_int, [_vp, C.c_int32, _f, _vp, _vp, _vp, _vp]),
    "jbgpu_bp_set_aabbs": (C.c_int, [_vp, C.c_int32, _vp, C.c_int32]),
    "jbgpu_bp_calculate_overlapping_pairs": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "jbgpu_bp_get_pairs": (C.c_int, [_vp, _vp, C.c_int64, C.POINTER(C.c_int64)]),
    "jbgpu_bp_destroy_proxies": (C.c_int, [_vp, C.c_int32, _vp]),
    "jbgpu_world_ray_test": (C.c_int, [_vp, C.c_int32, _vp, _vp, C.c_int16, C.c_int16, _vp, _vp, _vp, _vp, _vp,
                                       C.c_int32]),
    "jbgpu_world_convex_sweep_test": (C.c_int, [_vp, _f, C.c_int32, _vp, _vp, C.c_int16, C.c_int16, _vp, _vp, _vp,
                                                _vp, _vp, C.c_int32]),
    "jbgpu_world_set_ccd": (C.c_int, [_vp, C.c_int32, _vp, _vp, _vp]),
    "jbgpu_bp_get_pair_deltas": (C.c_int, [_vp, _vp, C.c_int64, C.POINTER(C.c_int64), _vp, C.c_int64,
                                           C.POINTER(C.c_int64)]),
    "jbgpu_bp_get_proxies": (C.c_int, [_vp, _f, C.POINTER(C.c_int32)]),
    "jbgpu_bp_set_proxies": (C.c_int, [_vp, _f, C.c_int32]),
    "jbgpu_bp_get_quantized": (C.c_int, [_vp, _i32]),
    "jbgpu_world_create": (C.c_int, [_vp, C.c_int32, _f, C.c_int32, _f, C.c_int32, _f, _vp]),
    "jbgpu_world_step": (C.c_int, [_vp, C.c_float, C.c_int32]),
    "jbgpu_world_set_step_graph": (C.c_int, [_vp, C.c_int32]),
    "jbgpu_world_step_until": (C.c_int, [_vp, C.c_float, C.c_int32]),
    "jbgpu_world_get_bodies": (C.c_int, [_vp, _f]),
    "jbgpu_world_set_bodies": (C.c_int, [_vp, _f]),
    "jbgpu_world_num_pairs": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "jbgpu_world_get_pairs": (C.c_int, [_vp, _vp, _vp]),
    "jbgpu_world_set_pairs": (C.c_int, [_vp, C.c_int64, _vp, _vp]),
    "jbgpu_world_stats": (C.c_int, [_vp, C.POINTER(Stats)]),
    "jbgpu_world_add_point2point": (C.c_int, [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp]),
    "jbgpu_world_add_hinge": (C.c_int, [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp]),
    "jbgpu_hinge_frames": (C.c_int, [C.c_int32, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jbgpu_world_num_joints": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "jbgpu_np_set_epa_capacity": (C.c_int, [_vp, C.c_int32, C.c_int32]),
    "jbgpu_prim_sort": (C.c_int, [_vp, _vp, C.c_int32, _vp, C.c_int64, C.c_int32, C.c_int32]),
    "jbgpu_world_joint_impulses": (C.c_int, [_vp, _vp]),
    "jbgpu_world_joint_order": (C.c_int, [_vp, _vp, C.c_int32, C.POINTER(C.c_int32)]),
    "jbgpu_check_pair_records": (C.c_int, [C.c_int32, C.c_int64, _vp, _vp]),
    "jbgpu_np_dispatch": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp]),
    "jbgpu_solver_prepare": (C.c_int, [_vp, C.c_int32, C.c_int64]),
    "jbgpu_solver_add_island": (C.c_int, [_vp, C.c_int64, _vp]),
    "jbgpu_solver_all_solved": (C.c_int, [_vp, C.c_float]),
    "jbgpu_solver_get_order": (C.c_int, [_vp, _vp, C.c_int64, C.POINTER(C.c_int64), _vp, C.POINTER(C.c_int32)]),
    "jbgpu_solver_set_warmstart": (C.c_int, [_vp, _vp, C.c_int64, C.c_int32]),
    "jbgpu_solver_enable_features": (C.c_int, [_vp, C.c_int32]),
    "jbgpu_solver_get_features": (C.c_int, [_vp, _f, C.c_int64]),
    "jbgpu_solver_get_impulses": (C.c_int, [_vp, _f, C.c_int64]),
    "jbgpu_solver_set_params": (C.c_int, [_vp, C.c_float, C.c_int32, C.c_int32]),
    "jbgpu_solver_set_tail": (C.c_int, [_vp, C.c_int32]),
    "jbgpu_solver_set_blocks": (C.c_int, [_vp, C.c_int32]),
    "jbgpu_solver_set_world_blocks": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32]),
    "jbgpu_solver_set_colouring": (C.c_int, [_vp, C.c_int32, C.c_int32]),
    "jbgpu_solver_set_warmstart_factor": (C.c_int, [_vp, C.c_float]),
    "jbgpu_solver_residual_history": (C.c_int, [_vp, _f, C.c_int32, C.POINTER(C.c_int32)]),
    "jbgpu_mlp_set_params": (C.c_int, [_vp, _f, _f]),
    "jbgpu_mlp_get_params": (C.c_int, [_vp, _f, _f]),
    "jbgpu_mlp_infer": (C.c_int, [_vp, _f, C.c_int64, _f]),
    "jbgpu_mlp_infer_static": (C.c_int, [_vp, _f, C.c_int64, _f]),
    "jbgpu_mlp_grad": (C.c_int, [_vp, _f, _f, C.c_int64, _f, C.POINTER(C.c_float)]),
    "jbgpu_mlp_train": (C.c_int, [_vp, _f, _f, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _vp]),
    "jbgpu_solver_use_mlp": (C.c_int, [_vp, C.c_int32]),
    "jbgpu_timing_enable": (C.c_int, [_vp, C.c_int32]),
    "jbgpu_timing_get": (C.c_int, [_vp, _vp, _vp, C.POINTER(C.c_int64)]),
    "jbgpu_timing_kernels": (C.c_int, [_vp, C.c_int32]),
    "jbgpu_timing_kernels_stride": (C.c_int, [_vp, C.c_int32]),
    "jbgpu_measure_copy": (C.c_int, [_vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "jbgpu_buffer_alloc": (C.c_int, [_vp, C.c_int64, C.POINTER(_vp)]),
    "jbgpu_buffer_upload": (C.c_int, [_vp, _vp, _vp, C.c_int64]),
    "jbgpu_buffer_free": (C.c_int, [_vp, _vp]),
    "jbgpu_timing_kernels_get": (C.c_int, [_vp, _vp, _vp, _vp]),
}


class JbgpuError(RuntimeError):
    pass


_lib = None


def load():
    """Load libjbgpu.so (built in-tree by __graft_entry__.build()); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise JbgpuError(f"{LIB_PATH} not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def device_count():
    n = C.c_int32(0)
    load().jbgpu_device_count(C.byref(n))
    return n.value


def default_config(**kw):
    cfg = Config()
    load().jbgpu_config_default(C.byref(cfg))
    for k, v in kw.items():
        if k in ("world_min", "world_max"):
            getattr(cfg, k)[:] = [float(x) for x in v]
        else:
            setattr(cfg, k, v)
    return cfg


# ---- shape helpers (host; the record builder for jbgpu.scenes.build_records) ----
def make_shape(stype, params, verts=None):
    rec = np.zeros(SHAPE_WORDS, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    rc = load().jbgpu_make_shape(int(stype), f32(np.asarray(params, np.float32).reshape(-1)), _p(vv), rec)
    if rc:
        raise JbgpuError(f"jbgpu_make_shape failed ({rc})")
    return rec


def shape_inertia(rec, mass, verts=None):
    out = np.zeros(3, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    rc = load().jbgpu_shape_local_inertia(f32(rec), _p(vv), float(mass), out)
    if rc:
        raise JbgpuError(f"jbgpu_shape_local_inertia failed ({rc})")
    return out


def hinge_frames(tr_a, pivot_a=None, axis_a=None, pivot_b=None, axis_b=None, frame_a=None, frame_b=None):
    """The solver frames of a HingeConstraint built from body A's transform tr_a (12 floats) as its constructors
    build them: pivot_a / axis_a (and pivot_b / axis_b for two bodies; None: the fixed body), or frame_a (and
    frame_b; None: the fixed body).  Host-only.  Returns (frame_a, frame_b), 12 floats each."""
    fa, fb = np.zeros(12, np.float32), np.zeros(12, np.float32)
    tr = f32(tr_a).reshape(-1)
    if frame_a is not None:
        ia = f32(frame_a).reshape(-1)
        ib = None if frame_b is None else f32(frame_b).reshape(-1)
        rc = load().jbgpu_hinge_frames(1, int(ib is None), _p(tr), None, None, None, None, _p(ia), _p(ib), _p(fa),
                                       _p(fb))
    else:
        pa, xa = f32(pivot_a).reshape(-1), f32(axis_a).reshape(-1)
        pb = None if pivot_b is None else f32(pivot_b).reshape(-1)
        xb = None if axis_b is None else f32(axis_b).reshape(-1)
        rc = load().jbgpu_hinge_frames(0, int(pb is None), _p(tr), _p(pa), _p(pb), _p(xa), _p(xb), None, None,
                                       _p(fa), _p(fb))
    if rc:
        buf = C.create_string_buffer(256)
        load().jbgpu_last_error(None, buf, 256)
        raise JbgpuError(f"hinge_frames failed ({rc}): {buf.value.decode()}")
    return fa, fb


def check_pair_records(nbodies, keys, manifolds):
    """jbgpu_check_pair_records: raises JbgpuError with the library's reason if a record is malformed."""
    k = np.ascontiguousarray(keys, np.uint64)
    m = f32(manifolds).reshape(-1)
    if m.size != len(k) * MANIFOLD_WORDS:
        raise ValueError("one manifold record per key")
    L = load()
    rc = L.jbgpu_check_pair_records(int(nbodies), len(k), _p(k), _p(m))
    if rc:
        buf = C.create_string_buffer(256)
        L.jbgpu_last_error(None, buf, 256)
        raise JbgpuError(f"jbgpu_check_pair_records failed ({rc}): {buf.value.decode()}")


class Context:
    """One jbgpu context (one HIP stream) with a broadphase, dispatcher and solver."""

    def __init__(self, **cfg):
        L = load()
        self._cfg = default_config(**cfg)
        h = C.c_void_p()
        rc = L.jbgpu_ctx_create(C.byref(self._cfg), C.byref(h))
        if rc:
            buf = C.create_string_buffer(256)
            L.jbgpu_last_error(None, buf, 256)
            raise JbgpuError(f"jbgpu_ctx_create failed ({rc}): {buf.value.decode()}")
        self.h = h
        self.nb = 0
        self.np = 0

    def close(self):
        if getattr(self, "h", None):
            load().jbgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ck(self, rc, what):
        if rc:
            buf = C.create_string_buffer(512)
            load().jbgpu_last_error(self.h, buf, 512)
            raise JbgpuError(f"{what} failed ({rc}): {buf.value.decode()}")

    # ---- BroadphaseInterface
    def bp_create_proxies(self, aabb6, group=None, mask=None, world=None):
        a = f32(aabb6).reshape(-1)
        n = a.size // 6
        g = None if group is None else np.ascontiguousarray(group, np.int16)
        m = None if mask is None else np.ascontiguousarray(mask, np.int16)
        w = None if world is None else np.ascontiguousarray(world, np.int32)
        uid = np.zeros(n, np.int32)
        self._ck(load().jbgpu_bp_create_proxies(self.h, n, a, _p(g), _p(m), _p(w), _p(uid)), "bp_create_proxies")
        self.np += n
        return uid

    def bp_destroy_proxies(self, uids):
        u = np.ascontiguousarray(uids, np.int32)
        self._ck(load().jbgpu_bp_destroy_proxies(self.h, int(u.size), _p(u)), "bp_destroy_proxies")

    def bp_pair_deltas(self):
        """(added, removed) sorted keys of the last pair computation."""
        na, nr = C.c_int64(0), C.c_int64(0)
        self._ck(load().jbgpu_bp_get_pair_deltas(self.h, None, 0, C.byref(na), None, 0, C.byref(nr)),
                 "bp_get_pair_deltas")
        a = np.zeros(max(na.value, 1), np.uint64)
        r = np.zeros(max(nr.value, 1), np.uint64)
        self._ck(load().jbgpu_bp_get_pair_deltas(self.h, _p(a), a.size, C.byref(na), _p(r), r.size, C.byref(nr)),
                 "bp_get_pair_deltas")
        return a[:na.value], r[:nr.value]

    def bp_set_aabbs(self, aabb6):
        a = f32(aabb6).reshape(-1)
        self._ck(load().jbgpu_bp_set_aabbs(self.h, a.size // 6, _p(a), 0), "bp_set_aabbs")

    def bp_set_aabbs_device(self, ptr, n):
        self._ck(load().jbgpu_bp_set_aabbs(self.h, int(n), C.c_void_p(int(ptr)), 1), "bp_set_aabbs")

    def bp_calculate_overlapping_pairs(self):
        n = C.c_int64(0)
        self._ck(load().jbgpu_bp_calculate_overlapping_pairs(self.h, C.byref(n)), "calculate_overlapping_pairs")
        return n.value

    def bp_pairs(self):
        n = C.c_int64(0)
        self._ck(load().jbgpu_bp_get_pairs(self.h, None, 0, C.byref(n)), "bp_get_pairs")
        keys = np.zeros(max(n.value, 1), np.uint64)
        self._ck(load().jbgpu_bp_get_pairs(self.h, _p(keys), keys.size, C.byref(n)), "bp_get_pairs")
        return keys[:n.value]

    def bp_proxies(self):
        out = np.zeros(self.np * PROXY_WORDS, np.float32)
        sc = C.c_int32(0)
        self._ck(load().jbgpu_bp_get_proxies(self.h, out, C.byref(sc)), "bp_get_proxies")
        return out.reshape(self.np, PROXY_WORDS), sc.value

    def bp_set_proxies(self, rec, stage_current):
        self._ck(load().jbgpu_bp_set_proxies(self.h, f32(rec).reshape(-1), int(stage_current)), "bp_set_proxies")

    def bp_quantized(self):
        out = np.zeros(self.np * 6, np.int32)
        self._ck(load().jbgpu_bp_get_quantized(self.h, out), "bp_get_quantized")
        return out.reshape(self.np, 6)

    # ---- DynamicsWorld
    def world_create(self, shapes, verts, bodies, world_ids=None):
        s = f32(shapes).reshape(-1)
        v = f32(verts if verts is not None and len(verts) else np.zeros((1, 3))).reshape(-1)
        b = f32(bodies).reshape(-1)
        w = None if world_ids is None else np.ascontiguousarray(world_ids, np.int32)
        self.nb = b.size // BODY_WORDS
        self._ck(load().jbgpu_world_create(self.h, s.size // SHAPE_WORDS, s, v.size // 3, v, self.nb, b, _p(w)),
                 "world_create")
        self.np = self.nb

    def step(self, dt=1.0 / 60.0, n=1):
        self._ck(load().jbgpu_world_step(self.h, dt, int(n)), "world_step")

    def set_step_graph(self, enable=True):
        """Run jbgpu_world_step as replayed hipGraphs (batched worlds with the world-block solve)."""
        self._ck(load().jbgpu_world_set_step_graph(self.h, 1 if enable else 0), "world_set_step_graph")

    def ray_test(self, frm, to, group=1, mask=-1, world=None):
        """rayTest with a closest-hit callback per ray: (body or -1, fraction, normal, point)."""
        frm = f32(frm).reshape(-1, 3)
        to = f32(to).reshape(-1, 3)
        n = len(frm)
        obj = np.zeros(n, np.int32)
        fr = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        pt = np.zeros((n, 3), np.float32)
        w = None if world is None else np.ascontiguousarray(world, np.int32)
        self._ck(load().jbgpu_world_ray_test(self.h, n, _p(frm), _p(to), int(np.int16(group)), int(np.int16(mask)),
                                             _p(w), _p(obj), _p(fr), _p(nrm), _p(pt), 0), "world_ray_test")
        return obj, fr, nrm, pt

    def ray_test_device(self, n, frm, to, obj, fr, nrm, pt, group=1, mask=-1, world=None):
        """rayTest on rays and results resident in device buffers (jbgpu_buffer_alloc addresses); world: device
        address of n int32 world ids or None."""
        v = lambda a: C.c_void_p(int(a))
        self._ck(load().jbgpu_world_ray_test(self.h, int(n), v(frm), v(to), int(np.int16(group)), int(np.int16(mask)),
                                             None if world is None else v(world), v(obj), v(fr), v(nrm), v(pt), 1),
                 "world_ray_test")

    def convex_sweep(self, shape, frm, to, group=1, mask=-1, world=None):
        """convexSweepTest of a convex shape record with a closest-hit callback per query (from/to: 12-float
        transforms, basis rows then origin): (body or -1, fraction, normal, point)."""
        frm = f32(frm).reshape(-1, 12)
        to = f32(to).reshape(-1, 12)
        n = len(frm)
        if len(to) != n:
            raise ValueError("from and to differ in length")
        obj = np.zeros(n, np.int32)
        fr = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        pt = np.zeros((n, 3), np.float32)
        w = None if world is None else np.ascontiguousarray(world, np.int32)
        self._ck(load().jbgpu_world_convex_sweep_test(self.h, f32(shape).reshape(-1), n, _p(frm), _p(to),
                                                      int(np.int16(group)), int(np.int16(mask)), _p(w), _p(obj),
                                                      _p(fr), _p(nrm), _p(pt), 0), "world_convex_sweep_test")
        return obj, fr, nrm, pt

    def set_ccd(self, bodies, threshold, radius):
        """setCcdMotionThreshold / setCcdSweptSphereRadius of the given bodies (scalars or one value per body)."""
        b = np.ascontiguousarray(bodies, np.int32).reshape(-1)
        t = np.broadcast_to(f32(threshold), b.shape).copy()
        r = np.broadcast_to(f32(radius), b.shape).copy()
        self._ck(load().jbgpu_world_set_ccd(self.h, len(b), _p(b), _p(t), _p(r)), "world_set_ccd")

    def step_until(self, phase, dt=1.0 / 60.0):
        self._ck(load().jbgpu_world_step_until(self.h, dt, int(phase)), "world_step_until")

    def bodies(self):
        out = np.zeros(self.nb * BODY_WORDS, np.float32)
        self._ck(load().jbgpu_world_get_bodies(self.h, out), "world_get_bodies")
        return out.reshape(self.nb, BODY_WORDS)

    def set_bodies(self, b):
        self._ck(load().jbgpu_world_set_bodies(self.h, f32(b).reshape(-1)), "world_set_bodies")

    def pairs(self):
        n = C.c_int64(0)
        self._ck(load().jbgpu_world_num_pairs(self.h, C.byref(n)), "world_num_pairs")
        keys = np.zeros(max(n.value, 1), np.uint64)
        man = np.zeros(max(n.value, 1) * MANIFOLD_WORDS, np.float32)
        self._ck(load().jbgpu_world_get_pairs(self.h, _p(keys), _p(man)), "world_get_pairs")
        return keys[:n.value], man.reshape(-1, MANIFOLD_WORDS)[:n.value]

    def set_pairs(self, keys, man):
        k = np.ascontiguousarray(keys, np.uint64)
        m = f32(man).reshape(-1)
        self._ck(load().jbgpu_world_set_pairs(self.h, len(k), _p(k), _p(m)), "world_set_pairs")

    def add_point2point(self, body_a, body_b, pivot_a, pivot_b, settings=None):
        """addConstraint of Point2PointConstraints; body_b -1 is the fixed body (pivot_b a world point);
        settings rows (tau, damping, impulse clamp) or None for (0.3, 1, 0)."""
        a = np.ascontiguousarray(body_a, np.int32).reshape(-1)
        b = np.ascontiguousarray(body_b, np.int32).reshape(-1)
        pa = f32(pivot_a).reshape(-1)
        pb = f32(pivot_b).reshape(-1)
        if not (len(b) == len(a) and pa.size == 3 * len(a) and pb.size == 3 * len(a)):
            raise ValueError("body_a, body_b, pivot_a (n x 3), pivot_b (n x 3) disagree")
        st = None if settings is None else f32(settings).reshape(-1)
        if st is not None and st.size != 3 * len(a):
            raise ValueError("settings must be n x 3")
        self._ck(load().jbgpu_world_add_point2point(self.h, len(a), _p(a), _p(b), _p(pa), _p(pb), _p(st)),
                 "world_add_point2point")

    def add_hinge(self, body_a, body_b, frame_a, frame_b, params=None):
        """addConstraint of HingeConstraints from their solver frames (n x 12, see hinge_frames); params rows
        (lower, upper, softness, bias, relaxation, angular_only, motor_on, motor_target, max_motor) or None."""
        a = np.ascontiguousarray(body_a, np.int32).reshape(-1)
        b = np.ascontiguousarray(body_b, np.int32).reshape(-1)
        fa = f32(frame_a).reshape(-1)
        fb = f32(frame_b).reshape(-1)
        if not (len(b) == len(a) and fa.size == 12 * len(a) and fb.size == 12 * len(a)):
            raise ValueError("body_a, body_b, frame_a (n x 12), frame_b (n x 12) disagree")
        pr = None if params is None else f32(params).reshape(-1)
        if pr is not None and pr.size != 9 * len(a):
            raise ValueError("params must be n x 9")
        self._ck(load().jbgpu_world_add_hinge(self.h, len(a), _p(a), _p(b), _p(fa), _p(fb), _p(pr)),
                 "world_add_hinge")

    def np_set_epa_capacity(self, small_faces=192, large_faces=1024):
        """Face capacities of the two LDS EPA tiers (work moves between tiers; results never change)."""
        self._ck(load().jbgpu_np_set_epa_capacity(self.h, int(small_faces), int(large_faces)), "np_set_epa_capacity")

    def prim_sort(self, keys, values=None, bit_lo=0, bit_hi=None):
        """The device radix sort on copies of `keys` (uint32 or uint64) and `values` (int32): stable by bits
        [bit_lo, bit_hi) of the keys as unsigned words.  Returns (keys, values)."""
        k = np.ascontiguousarray(keys).copy()
        if k.dtype not in (np.uint32, np.uint64):
            raise ValueError("keys must be uint32 or uint64")
        v = None if values is None else np.ascontiguousarray(values, np.int32).copy()
        if v is not None and v.size != k.size:
            raise ValueError("values must match the keys")
        hi = 8 * k.itemsize if bit_hi is None else int(bit_hi)
        self._ck(load().jbgpu_prim_sort(self.h, _p(k), k.itemsize, None if v is None else _p(v), k.size,
                                        int(bit_lo), hi), "prim_sort")
        return k, v

    def num_joints(self):
        n = C.c_int32(0)
        self._ck(load().jbgpu_world_num_joints(self.h, C.byref(n)), "world_num_joints")
        return n.value

    def joint_impulses(self):
        out = np.zeros(max(self.num_joints(), 1), np.float32)
        self._ck(load().jbgpu_world_joint_impulses(self.h, _p(out)), "world_joint_impulses")
        return out[:self.num_joints()]

    def joint_order(self):
        n = C.c_int32(0)
        self._ck(load().jbgpu_world_joint_order(self.h, None, 0, C.byref(n)), "world_joint_order")
        out = np.zeros(max(n.value, 1), np.int32)
        self._ck(load().jbgpu_world_joint_order(self.h, _p(out), len(out), C.byref(n)), "world_joint_order")
        return out[:n.value]

    def stats(self):
        s = Stats()
        self._ck(load().jbgpu_world_stats(self.h, C.byref(s)), "world_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    # ---- Dispatcher
    def np_dispatch(self, keys, manifolds):
        """dispatchAllCollisionPairs on the given pair cache (sorted keys + manifold records): updated manifolds."""
        k = np.ascontiguousarray(keys, np.uint64)
        m = f32(manifolds).reshape(-1)
        if m.size != len(k) * MANIFOLD_WORDS:
            raise ValueError("one manifold record per key")
        out = np.zeros(max(len(k), 1) * MANIFOLD_WORDS, np.float32)
        self._ck(load().jbgpu_np_dispatch(self.h, len(k), _p(k), _p(m), _p(out)), "np_dispatch")
        return out.reshape(-1, MANIFOLD_WORDS)[:len(k)]

    # ---- ConstraintSolver
    def solver_prepare(self, num_bodies, num_manifolds):
        self._ck(load().jbgpu_solver_prepare(self.h, int(num_bodies), int(num_manifolds)), "solver_prepare")

    def solver_add_island(self, pair_index):
        ix = np.ascontiguousarray(pair_index, np.int32)
        self._ck(load().jbgpu_solver_add_island(self.h, len(ix), _p(ix)), "solver_add_island")

    def solver_all_solved(self, dt=1.0 / 60.0):
        self._ck(load().jbgpu_solver_all_solved(self.h, float(dt)), "solver_all_solved")

    def solver_set_world_blocks(self, bodies_per_block, threads=256, pipelined=False):
        """One workgroup per block of whole worlds (at most bodies_per_block dynamic bodies; 0 = off);
        pipelined: rows of the next colour loaded while a colour is swept."""
        self._ck(load().jbgpu_solver_set_world_blocks(self.h, int(bodies_per_block), int(threads), int(pipelined)),
                 "solver_set_world_blocks")

    def solver_set_colouring(self, degree_first=False, recolour_passes=0):
        """Colouring priorities (largest degree first or hashed) and iterated-greedy recolouring passes."""
        self._ck(load().jbgpu_solver_set_colouring(self.h, int(bool(degree_first)), int(recolour_passes)),
                 "solver_set_colouring")

    def solver_order(self):
        n = C.c_int64(0)
        nc = C.c_int32(0)
        self._ck(load().jbgpu_solver_get_order(self.h, None, 0, C.byref(n), None, C.byref(nc)), "solver_get_order")
        order = np.zeros(max(n.value, 1), np.int32)
        offs = np.zeros(nc.value + 1, np.int64)
        self._ck(load().jbgpu_solver_get_order(self.h, _p(order), order.size, C.byref(n), _p(offs), C.byref(nc)),
                 "solver_get_order")
        return order[:n.value], offs

    def solver_set_warmstart(self, init):
        if init is None:
            self._ck(load().jbgpu_solver_set_warmstart(self.h, None, 0, 0), "solver_set_warmstart")
            return
        a = f32(init).reshape(-1)
        self._keep = a
        self._ck(load().jbgpu_solver_set_warmstart(self.h, _p(a), a.size // 3, 0), "solver_set_warmstart")

    def solver_enable_features(self, on=True):
        self._ck(load().jbgpu_solver_enable_features(self.h, int(on)), "solver_enable_features")

    def solver_features(self, nslots):
        out = np.zeros(max(nslots, 1) * 16, np.float32)
        self._ck(load().jbgpu_solver_get_features(self.h, out, int(nslots)), "solver_get_features")
        return out.reshape(-1, 16)[:nslots]

    def solver_impulses(self, nslots):
        """Final (normal, lateral1, lateral2) impulses of the last solve, by contact slot."""
        out = np.zeros(max(nslots, 1) * 3, np.float32)
        self._ck(load().jbgpu_solver_get_impulses(self.h, out, int(nslots)), "solver_get_impulses")
        return out.reshape(-1, 3)[:nslots]

    def solver_use_mlp(self, on=True):
        self._ck(load().jbgpu_solver_use_mlp(self.h, int(on)), "solver_use_mlp")

    # ---- learned warm start model
    def mlp_set_params(self, params, norm):
        self._ck(load().jbgpu_mlp_set_params(self.h, f32(params).reshape(-1), f32(norm).reshape(-1)),
                 "mlp_set_params")

    def mlp_get_params(self):
        w = np.zeros(MLP_PARAMS, np.float32)
        nm = np.zeros(MLP_NORM_WORDS, np.float32)
        self._ck(load().jbgpu_mlp_get_params(self.h, w, nm), "mlp_get_params")
        return w, nm

    def mlp_infer(self, x):
        x = f32(x).reshape(-1, 16)
        y = np.zeros((max(len(x), 1), 3), np.float32)
        self._ck(load().jbgpu_mlp_infer(self.h, x.reshape(-1) if len(x) else np.zeros(16, np.float32), len(x),
                                        y.reshape(-1)), "mlp_infer")
        return y[:len(x)]

    def mlp_infer_static(self, x):
        """mlp_infer through the library's fixed per-device staging buffer (no device allocation; at most
        MLP_STATIC_MAX contacts)."""
        x = f32(x).reshape(-1, 16)
        y = np.zeros((max(len(x), 1), 3), np.float32)
        self._ck(load().jbgpu_mlp_infer_static(self.h, x.reshape(-1) if len(x) else np.zeros(16, np.float32),
                                               len(x), y.reshape(-1)), "mlp_infer_static")
        return y[:len(x)]

    def mlp_grad(self, x, t):
        x = f32(x).reshape(-1, 16)
        t = f32(t).reshape(-1, 3)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float(0)
        self._ck(load().jbgpu_mlp_grad(self.h, x.reshape(-1), t.reshape(-1), len(x), g, C.byref(loss)), "mlp_grad")
        return g, loss.value

    def mlp_train(self, x, t, epochs=10, batch=4096, lr=1e-3, seed=0):
        x = f32(x).reshape(-1, 16)
        t = f32(t).reshape(-1, 3)
        losses = np.zeros(epochs, np.float32)
        self._ck(load().jbgpu_mlp_train(self.h, x.reshape(-1), t.reshape(-1), len(x), int(epochs), int(batch),
                                        float(lr), int(seed), _p(losses)), "mlp_train")
        return losses

    def solver_set_params(self, tolerance=0.0, max_iterations=100, num_iterations=10):
        self._ck(load().jbgpu_solver_set_params(self.h, float(tolerance), int(max_iterations), int(num_iterations)),
                 "solver_set_params")

    def solver_set_warmstart_factor(self, factor):
        self._ck(load().jbgpu_solver_set_warmstart_factor(self.h, float(factor)), "solver_set_warmstart_factor")

    def solver_set_blocks(self, block_size):
        self._ck(load().jbgpu_solver_set_blocks(self.h, int(block_size)), "solver_set_blocks")

    def solver_set_tail(self, threshold):
        self._ck(load().jbgpu_solver_set_tail(self.h, int(threshold)), "solver_set_tail")

    def residual_history(self):
        out = np.zeros(1024, np.float32)
        n = C.c_int32(0)
        self._ck(load().jbgpu_solver_residual_history(self.h, out, 1024, C.byref(n)), "residual_history")
        return out[:min(n.value, 1024)].copy()

    # ---- measurement
    def timing(self, on=True):
        self._ck(load().jbgpu_timing_enable(self.h, int(on)), "timing_enable")

    def buffer_alloc(self, nbytes):
        """Device buffer owned by the caller (free with buffer_free); returns the device address."""
        p = C.c_void_p()
        self._ck(load().jbgpu_buffer_alloc(self.h, int(nbytes), C.byref(p)), "buffer_alloc")
        return p.value

    def buffer_upload(self, dev, host):
        host = np.ascontiguousarray(host)
        self._ck(load().jbgpu_buffer_upload(self.h, C.c_void_p(dev), host.ctypes.data_as(C.c_void_p), host.nbytes),
                 "buffer_upload")

    def buffer_free(self, dev):
        self._ck(load().jbgpu_buffer_free(self.h, C.c_void_p(dev)), "buffer_free")

    def measure_copy(self, nbytes=1 << 30, reps=10):
        """Device copy ceiling (GB/s, read + write) of `nbytes` copied `reps` times."""
        g = C.c_double(0)
        self._ck(load().jbgpu_measure_copy(self.h, int(nbytes), int(reps), C.byref(g)), "measure_copy")
        return g.value

    def timing_kernels(self, every, stride=None):
        """Time the solver sweep launches by their own events on every `every`-th step (0 = off): those of every
        `stride`-th iteration (default 5; 1 = all), each launch its own events."""
        if stride is not None:
            self._ck(load().jbgpu_timing_kernels_stride(self.h, int(stride)), "timing_kernels_stride")
        self._ck(load().jbgpu_timing_kernels(self.h, int(every)), "timing_kernels")

    def timing_kernels_get(self):
        ms = np.zeros(4, np.float64)
        ln = np.zeros(4, np.int64)
        rows = np.zeros(4, np.int64)
        self._ck(load().jbgpu_timing_kernels_get(self.h, _p(ms), _p(ln), _p(rows)), "timing_kernels_get")
        return {name: dict(ms=float(ms[i]), launches=int(ln[i]), rows=int(rows[i]))
                for i, name in enumerate(("normal", "friction", "normal_blocks", "friction_blocks"))}

    def timing_get(self):
        ms = np.zeros(TIMER_REGIONS, np.float64)
        ln = np.zeros(TIMER_REGIONS, np.int64)
        st = C.c_int64(0)
        self._ck(load().jbgpu_timing_get(self.h, _p(ms), _p(ln), C.byref(st)), "timing_get")
        return {TIMER_NAMES[i]: (float(ms[i]), int(ln[i])) for i in range(TIMER_REGIONS)}, st.value
