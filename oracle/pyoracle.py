"""Oracle (test infrastructure only): ctypes binding of oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.  Record layouts follow
include/jbgpu_format.h.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SHAPE_WORDS, BODY_WORDS, PROXY_WORDS, POINT_WORDS, MANIFOLD_WORDS = 20, 32, 16, 32, 132

_f = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i16 = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
_u64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class WorldCfg(C.Structure):
    _fields_ = [("bpMode", C.c_int32), ("numIterations", C.c_int32), ("solverOrder", C.c_int32),
                ("deactivationDisabled", C.c_int32), ("wmin", C.c_float * 3), ("wmax", C.c_float * 3),
                ("tolerance", C.c_float), ("maxIterations", C.c_int32), ("solverMode", C.c_int32)]


class SolverInfo(C.Structure):
    _fields_ = [("timeStep", C.c_float), ("erp", C.c_float), ("warmstartingFactor", C.c_float),
                ("tolerance", C.c_float), ("numIterations", C.c_int32), ("maxIterations", C.c_int32),
                ("solverMode", C.c_int32), ("shuffle", C.c_int32), ("seed", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "orc_random_next_int": (C.c_int32, [C.c_int64, C.c_int32, _i32]),
            "orc_random_next_bounded": (C.c_int32, [C.c_int64, C.c_int32, _i32, _i32]),
            "orc_quat_from_matrix": (None, [_f, _f]),
            "orc_matrix_from_quat": (None, [_f, _f]),
            "orc_make_shape": (None, [C.c_int, _f, vp, _f]),
            "orc_support": (None, [_f, vp, _f, C.c_int, _f]),
            "orc_aabb": (None, [_f, vp, _f, _f]),
            "orc_inertia": (None, [_f, vp, C.c_float, _f]),
            "orc_gjk": (None, [_f, _f, _f, _f, vp, C.c_float, _f]),
            "orc_epa": (None, [_f, _f, _f, _f, vp, _f]),
            "orc_sphere_sphere": (None, [_f, _f, _f, _f, _f]),
            "orc_convex_plane": (None, [_f, _f, _f, _f, vp, _f]),
            "orc_manifold_add": (None, [_f, _f, _f, _f, _f, C.c_int]),
            "orc_bp_new": (vp, [C.c_int, vp, vp]),
            "orc_bp_free": (None, [vp]),
            "orc_bp_create": (None, [vp, C.c_int, _f, vp, vp]),
            "orc_bp_set_all": (None, [vp, _f]),
            "orc_bp_set_worlds": (None, [vp, _i32]),
            "orc_bp_destroy": (None, [vp, C.c_int, _i32]),
            "orc_world_ray_test": (None, [vp, C.c_int, _f, _f, C.c_int16, C.c_int16, vp, _i32, _f, _f, _f]),
            "orc_bp_step": (C.c_int64, [vp, _u64, C.c_int64, C.c_int]),
            "orc_bp_get_proxies": (None, [vp, _f]),
            "orc_bp_get_quantized": (None, [vp, _i32]),
            "orc_bp_stage_current": (C.c_int, [vp]),
            "orc_world_new": (vp, [C.POINTER(WorldCfg), C.c_int, _f, C.c_int, _f, C.c_int, _f]),
            "orc_world_free": (None, [vp]),
            "orc_world_set_order": (None, [vp, C.c_int, _i32]),
            "orc_world_set_worlds": (None, [vp, _i32]),
            "orc_world_set_solver": (None, [vp, C.c_int, C.c_float, C.c_int, C.c_int]),
            "orc_world_step": (C.c_int, [vp, C.c_float, C.c_int]),
            "orc_world_step_until": (C.c_int, [vp, C.c_float, C.c_int]),
            "orc_world_get_bodies": (None, [vp, _f]),
            "orc_world_set_bodies": (None, [vp, _f]),
            "orc_world_num_pairs": (C.c_int64, [vp]),
            "orc_world_get_pairs": (None, [vp, _u64, _f]),
            "orc_world_set_pairs": (None, [vp, C.c_int64, _u64, _f]),
            "orc_world_get_proxies": (None, [vp, _f]),
            "orc_world_set_proxies": (None, [vp, _f, C.c_int]),
            "orc_world_stage_current": (C.c_int, [vp]),
            "orc_world_stats": (None, [vp, _f]),
            "orc_world_residual_history": (C.c_int, [vp, _f, C.c_int]),
            "orc_world_set_learned_init": (None, [vp, vp]),
            "orc_world_enable_features": (None, [vp, C.c_int]),
            "orc_world_get_features": (C.c_int64, [_f, C.c_int64]),
            "orc_world_clear_features": (None, []),
            "orc_world_time_steps": (C.c_double, [vp, C.c_float, C.c_int]),
            "orc_shuffle": (None, [C.c_int64, C.c_int, C.c_int, _i32]),
            "orc_world_phase": (C.c_int, [vp, C.c_float, C.c_int]),
            "orc_world_add_p2p": (None, [vp, C.c_int, _i32, _i32, _f, _f, vp]),
            "orc_world_add_hinge": (None, [vp, C.c_int, _i32, _i32, _f, _f, vp]),
            "orc_world_convex_sweep": (None, [vp, C.c_int, C.c_int, _f, _f, C.c_int16, C.c_int16, vp, _i32, _f, _f,
                                              _f]),
            "orc_world_set_ccd": (None, [vp, C.c_int, _i32, _f, _f]),
            "orc_calculate_velocity": (None, [_f, _f, C.c_float, _f, _f]),
            "orc_temporal_aabb": (None, [_f, _f, _f, _f, C.c_float, _f, _f]),
            "orc_gjk_convex_cast": (C.c_int, [_f, _f, _f, _f, _f, vp, C.c_float, _f]),
            "orc_hinge_frames": (None, [C.c_int, C.c_int, _f, _f, _f, _f, _f, _f, _f, _f, _f]),
            "orc_atan2_fast": (C.c_float, [C.c_float, C.c_float]),
            "orc_world_hinge_angle": (C.c_float, [vp, C.c_int]),
            "orc_world_set_joint_order": (None, [vp, C.c_int, _i32]),
            "orc_world_joint_impulses": (None, [vp, _f]),
            "orc_solve_group": (None, [C.c_int, _f, C.c_int, _f, C.POINTER(SolverInfo), _f]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None or len(a) == 0 else a.ctypes.data_as(C.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ----------------------------------------------------------------- unit helpers
def make_shape(stype, params, verts=None):
    rec = np.zeros(SHAPE_WORDS, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    lib().orc_make_shape(int(stype), f32(np.asarray(params, np.float32).reshape(-1)), _ptr(vv), rec)
    return rec


def shape_inertia(rec, mass, verts=None):
    out = np.zeros(3, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    lib().orc_inertia(f32(rec), _ptr(vv), float(mass), out)
    return out


def support(rec, d, margin, verts=None):
    out = np.zeros(3, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    lib().orc_support(f32(rec), _ptr(vv), f32(d), int(margin), out)
    return out


def aabb(rec, tr12, verts=None):
    out = np.zeros(6, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    lib().orc_aabb(f32(rec), _ptr(vv), f32(tr12), out)
    return out


def gjk(recA, trA, recB, trB, max_dist2, verts=None):
    out = np.zeros(8, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    lib().orc_gjk(f32(recA), f32(trA), f32(recB), f32(trB), _ptr(vv), float(max_dist2), out)
    return out


def epa(recA, trA, recB, trB, verts=None):
    out = np.zeros(8, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    lib().orc_epa(f32(recA), f32(trA), f32(recB), f32(trB), _ptr(vv), out)
    return out


def sphere_sphere(recA, trA, recB, trB):
    out = np.zeros(8, np.float32)
    lib().orc_sphere_sphere(f32(recA), f32(trA), f32(recB), f32(trB), out)
    return out


def convex_plane(recC, trC, recP, trP, verts=None):
    out = np.zeros(8, np.float32)
    vv = None if verts is None else f32(verts).reshape(-1)
    lib().orc_convex_plane(f32(recC), f32(trC), f32(recP), f32(trP), _ptr(vv), out)
    return out


def manifold_add(mrec, trA, trB, fr, cand, refresh=True):
    m = f32(mrec).copy()
    lib().orc_manifold_add(m, f32(trA), f32(trB), f32(fr), f32(cand), int(refresh))
    return m


def random_ints(seed, n):
    out = np.zeros(n, np.int32)
    lib().orc_random_next_int(seed, n, out)
    return out


def random_bounded(seed, bounds):
    b = np.ascontiguousarray(bounds, np.int32)
    out = np.zeros(len(b), np.int32)
    lib().orc_random_next_bounded(seed, len(b), b, out)
    return out


def shuffle(seed, n, rounds=1):
    """Permutation of 0..n-1 after `rounds` row-order shuffles drawn from new Random(seed)."""
    out = np.zeros(max(n, 1), np.int32)
    lib().orc_shuffle(int(seed), int(n), int(rounds), out)
    return out[:n]


def solve_group(bodies, manifolds, dt=1.0 / 60.0, iterations=10, erp=0.2, warm_factor=0.85, warm=True,
                shuffle=False, seed=0, tolerance=0.0, max_iterations=100):
    """solveGroup on caller records (manifolds solved in the given order); returns (bodies, manifolds,
    iterations, residual) after the solve."""
    b = f32(bodies).reshape(-1, BODY_WORDS).copy()
    m = f32(manifolds).reshape(-1, MANIFOLD_WORDS).copy()
    si = SolverInfo(np.float32(dt), erp, warm_factor, tolerance, iterations, max_iterations,
                    (1 if shuffle else 0) | 8 | (4 if warm else 0), int(shuffle), seed)
    out = np.zeros(2, np.float32)
    lib().orc_solve_group(len(b), b.reshape(-1), len(m), m.reshape(-1) if len(m) else np.zeros(1, np.float32),
                          C.byref(si), out)
    return b, m, int(out[0]), float(out[1])


# ----------------------------------------------------------------- broadphase
class Broadphase:
    def __init__(self, mode, wmin=None, wmax=None):
        self.h = lib().orc_bp_new(mode, _ptr(None if wmin is None else f32(wmin)),
                                  _ptr(None if wmax is None else f32(wmax)))
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bp_free(self.h)
            self.h = None

    def create(self, aabb6, group=None, mask=None):
        a = f32(aabb6).reshape(-1)
        n = a.size // 6
        g = None if group is None else np.ascontiguousarray(group, np.int16)
        m = None if mask is None else np.ascontiguousarray(mask, np.int16)
        lib().orc_bp_create(self.h, n, a, _ptr(g), _ptr(m))
        self.n += n

    def set_worlds(self, world):
        lib().orc_bp_set_worlds(self.h, np.ascontiguousarray(world, np.int32))

    def destroy(self, uids):
        u = np.ascontiguousarray(uids, np.int32)
        lib().orc_bp_destroy(self.h, int(u.size), u)

    def set_all(self, aabb6):
        lib().orc_bp_set_all(self.h, f32(aabb6).reshape(-1))

    def step(self, brute=False, cap=None):
        cap = cap or max(1024, 64 * self.n)
        keys = np.zeros(cap, np.uint64)
        n = lib().orc_bp_step(self.h, keys, cap, int(brute))
        if n > cap:
            keys = np.zeros(n, np.uint64)
            raise RuntimeError("pair capacity")
        return keys[:n].copy()

    def proxies(self):
        out = np.zeros(self.n * PROXY_WORDS, np.float32)
        lib().orc_bp_get_proxies(self.h, out)
        return out.reshape(self.n, PROXY_WORDS)

    def quantized(self):
        out = np.zeros(self.n * 6, np.int32)
        lib().orc_bp_get_quantized(self.h, out)
        return out.reshape(self.n, 6)


# ----------------------------------------------------------------- world
def hinge_frames(tr_a, pivot_a, axis_a, pivot_b=None, axis_b=None):
    """rbAFrame / rbBFrame of HingeConstraint(rbA, rbB, pivotInA, pivotInB, axisInA, axisInB), or of
    HingeConstraint(rbA, pivotInA, axisInA) when pivot_b is None (B the fixed body); tr_a = body A's 12 words."""
    fa, fb = np.zeros(12, np.float32), np.zeros(12, np.float32)
    fixed = pivot_b is None
    z = np.zeros(3, np.float32)
    lib().orc_hinge_frames(0, int(fixed), f32(tr_a), f32(pivot_a), z if fixed else f32(pivot_b), f32(axis_a),
                           z if fixed else f32(axis_b), np.zeros(12, np.float32), np.zeros(12, np.float32), fa, fb)
    return fa, fb


def hinge_frames_from_frames(tr_a, frame_a, frame_b=None):
    """The frame constructors: HingeConstraint(rbA, rbB, rbAFrame, rbBFrame), or (rbA, rbAFrame) when frame_b is
    None."""
    fa, fb = np.zeros(12, np.float32), np.zeros(12, np.float32)
    fixed = frame_b is None
    z = np.zeros(3, np.float32)
    lib().orc_hinge_frames(1, int(fixed), f32(tr_a), z, z, z, z, f32(frame_a),
                           f32(frame_a) if fixed else f32(frame_b), fa, fb)
    return fa, fb


def calculate_velocity(t0, t1, dt):
    """TransformUtil.calculateVelocity: (linear, angular) velocity taking transform t0 to t1 (12 floats) in dt."""
    lv, av = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().orc_calculate_velocity(f32(t0), f32(t1), np.float32(dt), lv, av)
    return lv, av


def temporal_aabb(rec, tr, lin, ang, dt):
    """CollisionShape.calculateTemporalAabb of a shape record at tr (12 floats) moving with (lin, ang) for dt."""
    mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().orc_temporal_aabb(f32(rec), f32(tr), f32(lin), f32(ang), np.float32(dt), mn, mx)
    return mn, mx


def gjk_convex_cast(rec_a, from_a, to_a, rec_b, tr_b, hull=None, allowed_penetration=0.04):
    """GjkConvexCast.calcTimeOfImpact of shape record a moving from_a -> to_a against b at rest at tr_b:
    (hit, fraction, normal, point)."""
    out = np.zeros(7, np.float32)
    hv = None if hull is None else f32(hull).reshape(-1)
    hit = lib().orc_gjk_convex_cast(f32(rec_a), f32(from_a), f32(to_a), f32(rec_b), f32(tr_b),
                                    None if hv is None else hv.ctypes.data_as(C.c_void_p),
                                    np.float32(allowed_penetration), out)
    return bool(hit), float(out[0]), out[1:4].copy(), out[4:7].copy()


def atan2_fast(y, x):
    return float(lib().orc_atan2_fast(np.float32(y), np.float32(x)))


class World:
    def __init__(self, shapes, verts, bodies, bp_mode=0, iterations=10, order=0, deactivation_disabled=True,
                 wmin=(-1000, -1000, -1000), wmax=(1000, 1000, 1000), tolerance=0.0, max_iterations=100,
                 solver_mode=0):
        cfg = WorldCfg()
        cfg.bpMode = bp_mode
        cfg.numIterations = iterations
        cfg.solverOrder = order
        cfg.deactivationDisabled = int(deactivation_disabled)
        cfg.wmin[:] = list(map(float, wmin))
        cfg.wmax[:] = list(map(float, wmax))
        cfg.tolerance = tolerance
        cfg.maxIterations = max_iterations
        cfg.solverMode = solver_mode
        self.shapes = f32(shapes).reshape(-1, SHAPE_WORDS)
        self.verts = f32(verts if verts is not None and len(verts) else np.zeros((1, 3))).reshape(-1, 3)
        self.bodies0 = f32(bodies).reshape(-1, BODY_WORDS)
        self.nb = self.bodies0.shape[0]
        self.h = lib().orc_world_new(C.byref(cfg), self.shapes.shape[0], self.shapes.reshape(-1),
                                     self.verts.shape[0], self.verts.reshape(-1), self.nb, self.bodies0.reshape(-1))
        self._init = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_world_free(self.h)
            self.h = None

    def step(self, dt=1.0 / 60.0, n=1):
        lib().orc_world_step(self.h, dt, n)

    def add_point2point(self, body_a, body_b, pivot_a, pivot_b, settings=None):
        """Point2PointConstraint per row (body_b -1: the fixed world body, pivot_b a world point)."""
        a = np.ascontiguousarray(body_a, np.int32)
        b = np.ascontiguousarray(body_b, np.int32)
        self._settings = None if settings is None else f32(settings).reshape(-1)
        lib().orc_world_add_p2p(self.h, len(a), a, b, f32(pivot_a).reshape(-1), f32(pivot_b).reshape(-1),
                                _ptr(self._settings))
        self.nj = getattr(self, "nj", 0) + len(a)

    def add_hinge(self, body_a, body_b, frame_a, frame_b, params=None):
        """HingeConstraints from the solver frames (n x 12: basis row-major, origin); params rows
        (lower, upper, softness, bias, relaxation, angular_only, motor_on, motor_target, max_motor) or None."""
        a = np.ascontiguousarray(body_a, np.int32)
        b = np.ascontiguousarray(body_b, np.int32)
        fa = f32(frame_a).reshape(-1)
        fb = f32(frame_b).reshape(-1)
        self._hinge_params = None if params is None else f32(params).reshape(-1)
        lib().orc_world_add_hinge(self.h, len(a), a, b, fa, fb, _ptr(self._hinge_params))
        self.nj = getattr(self, "nj", 0) + len(a)

    def hinge_angle(self, j):
        return float(lib().orc_world_hinge_angle(self.h, int(j)))

    def set_joint_order(self, order):
        o = np.ascontiguousarray(order, np.int32)
        lib().orc_world_set_joint_order(self.h, len(o), o if len(o) else np.zeros(1, np.int32))

    def joint_impulses(self):
        out = np.zeros(max(getattr(self, "nj", 0), 1), np.float32)
        lib().orc_world_joint_impulses(self.h, out)
        return out[:getattr(self, "nj", 0)]

    def phase(self, p, dt=1.0 / 60.0):
        """One phase of the step alone (0 predict, 1 AABBs + pairs, 2 narrowphase, 3 islands + solve,
        4 integrate + activation)."""
        lib().orc_world_phase(self.h, dt, int(p))

    def step_until(self, phase, dt=1.0 / 60.0):
        lib().orc_world_step_until(self.h, dt, phase)

    def set_worlds(self, world):
        """World id per body: bodies of different worlds never pair (batched independent worlds)."""
        lib().orc_world_set_worlds(self.h, np.ascontiguousarray(world, np.int32))

    def ray_test(self, frm, to, group=1, mask=-1, world=None):
        """CollisionWorld.rayTest with a ClosestRayResultCallback per ray: (object, fraction, normal, point)."""
        frm = f32(frm).reshape(-1, 3)
        to = f32(to).reshape(-1, 3)
        n = len(frm)
        obj = np.zeros(n, np.int32)
        fr = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        pt = np.zeros((n, 3), np.float32)
        w = None if world is None else np.ascontiguousarray(world, np.int32)
        lib().orc_world_ray_test(self.h, n, frm.reshape(-1), to.reshape(-1), int(group), int(mask),
                                 None if w is None else w.ctypes.data_as(C.c_void_p), obj, fr, nrm.reshape(-1),
                                 pt.reshape(-1))
        return obj, fr, nrm, pt

    def convex_sweep(self, shape, frm, to, group=1, mask=-1, world=None):
        """convexSweepTest of the world's shape `shape` from each transform of frm to to (n x 12) with a
        ClosestConvexResultCallback: (object, fraction, normal, point)."""
        frm = f32(frm).reshape(-1, 12)
        to = f32(to).reshape(-1, 12)
        n = len(frm)
        obj = np.zeros(n, np.int32)
        fr = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        pt = np.zeros((n, 3), np.float32)
        w = None if world is None else np.ascontiguousarray(world, np.int32)
        lib().orc_world_convex_sweep(self.h, n, int(shape), frm.reshape(-1), to.reshape(-1), int(group), int(mask),
                                     None if w is None else w.ctypes.data_as(C.c_void_p), obj, fr, nrm.reshape(-1),
                                     pt.reshape(-1))
        return obj, fr, nrm, pt

    def set_ccd(self, bodies, threshold, radius):
        """setCcdMotionThreshold / setCcdSweptSphereRadius of the given bodies."""
        b = np.ascontiguousarray(bodies, np.int32).reshape(-1)
        t = np.broadcast_to(f32(threshold), b.shape).copy()
        r = np.broadcast_to(f32(radius), b.shape).copy()
        lib().orc_world_set_ccd(self.h, len(b), b, t, r)

    def set_order(self, order):
        o = np.ascontiguousarray(order, np.int32)
        lib().orc_world_set_order(self.h, len(o), o)

    def set_solver(self, order, tol=0.0, max_it=100, num_it=10):
        lib().orc_world_set_solver(self.h, order, tol, max_it, num_it)

    def bodies(self):
        out = np.zeros(self.nb * BODY_WORDS, np.float32)
        lib().orc_world_get_bodies(self.h, out)
        return out.reshape(self.nb, BODY_WORDS)

    def set_bodies(self, b):
        lib().orc_world_set_bodies(self.h, f32(b).reshape(-1))

    def pairs(self):
        n = lib().orc_world_num_pairs(self.h)
        keys = np.zeros(max(n, 1), np.uint64)
        man = np.zeros(max(n, 1) * MANIFOLD_WORDS, np.float32)
        lib().orc_world_get_pairs(self.h, keys, man)
        return keys[:n], man.reshape(-1, MANIFOLD_WORDS)[:n]

    def set_pairs(self, keys, man):
        k = np.ascontiguousarray(keys, np.uint64)
        lib().orc_world_set_pairs(self.h, len(k), k if len(k) else np.zeros(1, np.uint64),
                                  f32(man).reshape(-1) if len(k) else np.zeros(MANIFOLD_WORDS, np.float32))

    def proxies(self):
        out = np.zeros(self.nb * PROXY_WORDS, np.float32)
        lib().orc_world_get_proxies(self.h, out)
        return out.reshape(self.nb, PROXY_WORDS), lib().orc_world_stage_current(self.h)

    def set_proxies(self, rec, stage_current):
        lib().orc_world_set_proxies(self.h, f32(rec).reshape(-1), int(stage_current))

    def stats(self):
        out = np.zeros(8, np.float32)
        lib().orc_world_stats(self.h, out)
        return dict(iterations=int(out[0]), residual=float(out[1]), rows=int(out[2]), friction_rows=int(out[3]),
                    gjk_iters=int(out[4]), epa_calls=int(out[5]), epa_fails=int(out[6]),
                    step_residual=float(out[7]))

    def residual_history(self):
        out = np.zeros(1024, np.float32)
        n = lib().orc_world_residual_history(self.h, out, 1024)
        return out[:min(n, 1024)].copy()

    def set_learned_init(self, init):
        self._init = None if init is None else f32(init).reshape(-1)
        lib().orc_world_set_learned_init(self.h, _ptr(self._init))

    def enable_features(self, on=True):
        lib().orc_world_enable_features(self.h, int(on))

    def features(self):
        n = lib().orc_world_get_features(np.zeros(1, np.float32), 0)
        out = np.zeros(max(n, 1), np.float32)
        lib().orc_world_get_features(out, n)
        lib().orc_world_clear_features()
        return out[:n].reshape(-1, 16)

    def time_steps(self, n, dt=1.0 / 60.0):
        return lib().orc_world_time_steps(self.h, dt, n)
