"""Seeded synthetic scenes for the rigid-body hot path.

Each generator returns a :class:`Scene` of raw shape parameters and initial
body state in float32.  :func:`build_records` turns a scene into the shape and
body records of ``include/jbgpu_format.h`` with Java float semantics, using a
record builder (the jbgpu library's host helpers, or any object exposing the
same two functions), so that every consumer starts from identical bits.

Scenes (sizes as in the benchmark configurations):
  * ``boxes_grid``      1,000 boxes (half-extent 0.5 m, 1 kg) in a 10x10x10 grid
                        spaced 1.2 m, 5 m above a plane; jitter +-0.05 m, random yaw.
  * ``mixed_pile``      100,000 bodies: 40% spheres (r log-uniform 0.1-1 m), 30% boxes
                        (half-extents U(0.1, 1)), 15% capsules (r log-uniform 0.1-0.5,
                        height U(0.2, 2)), 15% hulls of 16-32 points on an ellipsoid
                        (semi-axes U(0.1, 1)); random orientation; density U(500, 2000);
                        dropped on a jittered 2.2 m lattice into a 100x100 m bin
                        (floor plane + 4 static box walls).
  * ``aabb_cloud``      1,000,000 AABBs, centres U[0, 1000)^3, half-extents
                        lognormal(0, 0.7), velocities N(0, 5 m/s): broadphase only.
  * ``batched_towers``  4,096 worlds of floor(exp(U(ln 16, ln 513))) bodies, spheres and
                        boxes 4:3, towers of height 4-20 on a 3 m grid, own floor each.
  * ``pyramids``        boxes (half-extents U(0.2, 1)) in pyramids and walls plus
                        falling spheres (learned warm-start workload).
  * ``chains``          hanging chains of boxes joined by Point2Point constraints, with
                        boxes dropped onto them.
  * ``hinges``          hinged chains, doors (limits, motors) and motor-driven wheels on
                        static posts, with boxes dropped onto them.
"""
from dataclasses import dataclass, field

import numpy as np

SPHERE, BOX, CAPSULE, HULL, PLANE = 0, 1, 2, 3, 4
SHAPE_WORDS, BODY_WORDS = 20, 32
GRAVITY = np.array([0.0, -9.81, 0.0], np.float32)
ACTIVE_TAG = 1


@dataclass
class Scene:
    name: str
    shape_type: np.ndarray            # (S,) int32
    shape_param: np.ndarray           # (S, 4) float32; hull: (first, count)
    verts: np.ndarray                 # (V, 3) float32 hull points
    body_shape: np.ndarray            # (N,) int32
    basis: np.ndarray                 # (N, 9) float32 row-major
    origin: np.ndarray                # (N, 3) float32
    mass: np.ndarray                  # (N,) float32, 0 = static
    friction: np.ndarray              # (N,) float32
    restitution: np.ndarray           # (N,) float32
    world: np.ndarray = None          # (N,) int32 world id (batched scenes)
    meta: dict = field(default_factory=dict)

    @property
    def n_bodies(self):
        return int(self.body_shape.shape[0])


def _yaw_basis(theta):
    c, s = np.cos(theta), np.sin(theta)
    n = theta.shape[0]
    b = np.zeros((n, 9), np.float64)
    b[:, 0], b[:, 2], b[:, 4], b[:, 6], b[:, 8] = c, s, 1.0, -s, c
    return b.astype(np.float32)


def _quat_basis(q):
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    b = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=1)
    return b.astype(np.float32)


def _random_quats(rng, n):
    # Shoemake's uniform random rotation
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)


class _Builder:
    def __init__(self, name):
        self.name = name
        self.st, self.sp, self.verts = [], [], []
        self.bs, self.basis, self.origin, self.mass, self.fr, self.rest, self.world = [], [], [], [], [], [], []
        self.nverts = 0

    def shape(self, t, params):
        self.st.append(t)
        p = list(params) + [0.0] * (4 - len(params))
        self.sp.append(p)
        return len(self.st) - 1

    def hull(self, pts):
        first = self.nverts
        self.verts.append(np.asarray(pts, np.float32))
        self.nverts += len(pts)
        return self.shape(HULL, [first, len(pts)])

    def body(self, shape, basis, origin, mass, friction=0.5, restitution=0.0, world=0):
        self.bs.append(shape)
        self.basis.append(basis)
        self.origin.append(origin)
        self.mass.append(mass)
        self.fr.append(friction)
        self.rest.append(restitution)
        self.world.append(world)

    def bodies(self, shapes, basis, origin, mass, friction=0.5, restitution=0.0, world=0):
        n = len(shapes)
        self.bs.extend(list(shapes))
        self.basis.extend(list(np.asarray(basis, np.float32)))
        self.origin.extend(list(np.asarray(origin, np.float32)))
        self.mass.extend(list(np.asarray(mass, np.float32)))
        self.fr.extend([friction] * n)
        self.rest.extend([restitution] * n)
        self.world.extend(list(np.broadcast_to(np.asarray(world, np.int32), (n,))))

    def done(self, **meta):
        v = np.concatenate(self.verts).astype(np.float32) if self.verts else np.zeros((0, 3), np.float32)
        return Scene(self.name, np.asarray(self.st, np.int32), np.asarray(self.sp, np.float32), v,
                     np.asarray(self.bs, np.int32), np.asarray(self.basis, np.float32).reshape(-1, 9),
                     np.asarray(self.origin, np.float32).reshape(-1, 3), np.asarray(self.mass, np.float32),
                     np.asarray(self.fr, np.float32), np.asarray(self.rest, np.float32),
                     np.asarray(self.world, np.int32), meta)


IDENT = np.eye(3, dtype=np.float32).reshape(9)


def boxes_grid(seed=1, nx=10, ny=10, nz=10, spacing=1.2, half=0.5, base=5.0):
    """1,000 unit boxes in a grid above a plane (the CPU-runnable configuration)."""
    rng = np.random.default_rng(seed)
    b = _Builder("boxes_grid")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    b.body(plane, IDENT, np.zeros(3, np.float32), 0.0)
    box = b.shape(BOX, [half, half, half])
    n = nx * ny * nz
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x = (ii.reshape(-1) - (nx - 1) / 2.0) * spacing
    y = base + half + jj.reshape(-1) * spacing
    z = (kk.reshape(-1) - (nz - 1) / 2.0) * spacing
    pos = np.stack([x, y, z], 1) + rng.uniform(-0.05, 0.05, (n, 3))
    b.bodies([box] * n, _yaw_basis(rng.uniform(0, 2 * np.pi, n)), pos.astype(np.float32), np.ones(n, np.float32))
    return b.done(dt=1 / 60, iterations=10, steps=300)


def _ellipsoid_volume(a):
    return 4.0 / 3.0 * np.pi * a[0] * a[1] * a[2]


def mixed_pile(seed=1, n=100_000, bin_half=50.0, spacing=2.2, wall_height=30.0):
    """Single large pile of mixed convex bodies dropped into a walled bin."""
    rng = np.random.default_rng(seed)
    b = _Builder("mixed_pile")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    b.body(plane, IDENT, np.zeros(3, np.float32), 0.0)
    t = 1.0
    hw = wall_height / 2
    for (hx, hz, px, pz) in [(t, bin_half + t, bin_half + t, 0.0), (t, bin_half + t, -bin_half - t, 0.0),
                             (bin_half + t, t, 0.0, bin_half + t), (bin_half + t, t, 0.0, -bin_half - t)]:
        w = b.shape(BOX, [hx, hw, hz])
        b.body(w, IDENT, np.array([px, hw, pz], np.float32), 0.0)
    kinds = rng.choice(4, size=n, p=[0.40, 0.30, 0.15, 0.15])
    density = rng.uniform(500, 2000, n)
    q = _random_quats(rng, n)
    per = int((2 * bin_half - spacing) // spacing)
    layer = per * per
    idx = np.arange(n)
    lx, lz, ly = idx % per, (idx // per) % per, idx // layer
    pos = np.stack([(lx - (per - 1) / 2.0) * spacing, 2.0 + ly * spacing, (lz - (per - 1) / 2.0) * spacing], 1)
    pos = pos + rng.uniform(-0.1, 0.1, (n, 3))
    shapes, masses = [], []
    for i in range(n):
        k = kinds[i]
        if k == 0:
            r = float(np.exp(rng.uniform(np.log(0.1), np.log(1.0))))
            s = b.shape(SPHERE, [r])
            vol = 4.0 / 3.0 * np.pi * r ** 3
        elif k == 1:
            h = rng.uniform(0.1, 1.0, 3)
            s = b.shape(BOX, list(h))
            vol = 8.0 * h[0] * h[1] * h[2]
        elif k == 2:
            r = float(np.exp(rng.uniform(np.log(0.1), np.log(0.5))))
            hgt = float(rng.uniform(0.2, 2.0))
            s = b.shape(CAPSULE, [r, hgt])
            vol = np.pi * r * r * hgt + 4.0 / 3.0 * np.pi * r ** 3
        else:
            axes = rng.uniform(0.1, 1.0, 3)
            m = int(rng.integers(16, 33))
            d = rng.normal(size=(m, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            s = b.hull((d * axes).astype(np.float32))
            vol = _ellipsoid_volume(axes)
        shapes.append(s)
        masses.append(density[i] * vol)
    b.bodies(shapes, _quat_basis(q), pos.astype(np.float32), np.asarray(masses, np.float32))
    return b.done(dt=1 / 60, iterations=20, steps=600)


def aabb_cloud(seed=1, n=1_000_000, extent=1000.0, sigma=0.7, vsig=5.0):
    """Broadphase-only stress: returns (centres, half extents, velocities) in float32."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, extent, (n, 3)).astype(np.float32)
    h = rng.lognormal(0.0, sigma, (n, 3)).astype(np.float32)
    v = rng.normal(0.0, vsig, (n, 3)).astype(np.float32)
    return c, h, v


def aabb_cloud_step(c, h, v, dt=np.float32(1.0 / 60.0)):
    """One step of the stress workload: c += v*dt (float32), AABB = c -+ h."""
    c += v * np.float32(dt)
    return np.concatenate([c - h, c + h], axis=1).astype(np.float32)


def batched_towers(seed=1, n_worlds=4096, lo=16, hi=512, spacing=3.0, world_gap=0.0):
    """Many independent worlds, each a few towers of spheres and boxes on its own floor.  Sizes as in the mixed
    pile: sphere radius log-uniform 0.1-1 m, box half extents U(0.1, 1) m per axis, density U(500, 2000);
    spheres and boxes 4:3; towers of height U{4..20} on a `spacing` grid, eight towers per row."""
    rng = np.random.default_rng(seed)
    b = _Builder("batched_towers")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    counts = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), n_worlds))).astype(np.int64)
    for w in range(n_worlds):
        b.body(plane, IDENT, np.zeros(3, np.float32), 0.0, world=w)
        nb = int(counts[w])
        placed = 0
        tower = 0
        while placed < nb:
            height = int(rng.integers(4, 21))
            tx = (tower % 8) * spacing
            tz = (tower // 8) * spacing
            y = 0.0
            for _ in range(min(height, nb - placed)):
                if rng.random() < 4.0 / 7.0:
                    r = float(np.exp(rng.uniform(np.log(0.1), np.log(1.0))))
                    s = b.shape(SPHERE, [r])
                    ext = r
                    m = rng.uniform(500, 2000) * 4.0 / 3.0 * np.pi * r ** 3
                else:
                    h = rng.uniform(0.1, 1.0, 3)
                    s = b.shape(BOX, list(h))
                    ext = h[1]
                    m = rng.uniform(500, 2000) * 8 * h[0] * h[1] * h[2]
                y += ext + 0.01
                b.body(s, IDENT, np.array([tx, y, tz], np.float32), np.float32(m), world=w)
                y += ext
                placed += 1
            tower += 1
    return b.done(dt=1 / 60, iterations=10, steps=1000, n_worlds=n_worlds)


def pyramids(seed=1, n_boxes=64_000, n_pyramids=32, n_spheres=2000, area=200.0):
    """Box structures plus falling spheres (learned warm-start workload): even structures are square-base
    pyramids (level L of a base-k pyramid holds (k-L)^2 boxes), odd ones are walls eight times wider than tall;
    each structure uses one box size, half extents U(0.2, 1) per axis."""
    rng = np.random.default_rng(seed)
    b = _Builder("pyramids")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    b.body(plane, IDENT, np.zeros(3, np.float32), 0.0)
    per = n_boxes // n_pyramids
    for p in range(n_pyramids):
        cx, cz = rng.uniform(-area / 2, area / 2, 2)
        h = rng.uniform(0.2, 1.0, 3).astype(np.float32)
        box = b.shape(BOX, list(h))
        mass = np.float32(1000.0 * 8 * h[0] * h[1] * h[2])
        gx, gz = 2 * h[0] + 0.01, 2 * h[2] + 0.01
        placed, level = 0, 0
        if p % 2 == 0:
            k = 1
            while k * (k + 1) * (2 * k + 1) // 6 < per:
                k += 1
            while placed < per and level < k:
                side = k - level
                for a in range(side):
                    for c in range(side):
                        if placed >= per:
                            break
                        x = cx + (a - (side - 1) / 2.0) * gx
                        z = cz + (c - (side - 1) / 2.0) * gz
                        y = h[1] + level * (2 * h[1] + 0.001)
                        b.body(box, IDENT, np.array([x, y, z], np.float32), mass)
                        placed += 1
                level += 1
        else:
            height = max(int(np.sqrt(per / 8.0)), 1)
            width = int(np.ceil(per / height))
            while placed < per:
                for a in range(width):
                    if placed >= per:
                        break
                    x = cx + (a - (width - 1) / 2.0) * gx
                    y = h[1] + level * (2 * h[1] + 0.001)
                    b.body(box, IDENT, np.array([x, y, cz], np.float32), mass)
                    placed += 1
                level += 1
    for _ in range(n_spheres):
        r = float(rng.uniform(0.2, 0.6))
        s = b.shape(SPHERE, [r])
        pos = np.array([rng.uniform(-area / 2, area / 2), rng.uniform(20, 60), rng.uniform(-area / 2, area / 2)],
                       np.float32)
        b.body(s, IDENT, pos, np.float32(1000.0 * 4.0 / 3.0 * np.pi * r ** 3))
    return b.done(dt=1 / 60, iterations=100, steps=200)


def aabb_overflow(nan_body=False):
    """AABB overflow (CollisionWorld.updateSingleAabb, CollisionWorld.java:211-227): a dynamic box whose AABB
    extent squared exceeds 1e12 (half extents 8e5 x 0.5 x 0.5) is taken out of the simulation
    (DISABLE_SIMULATION), beside a 1 m box falling onto the floor; optionally a box with a NaN origin, whose
    AABB test fails the same way."""
    b = _Builder("aabb_overflow")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    b.body(plane, IDENT, np.zeros(3, np.float32), 0.0)
    box = b.shape(BOX, [0.5, 0.5, 0.5])
    b.body(box, IDENT, np.array([0.0, 3.0, 0.0], np.float32), 1.0)
    big = b.shape(BOX, [8e5, 0.5, 0.5])
    b.body(big, IDENT, np.array([0.0, 20.0, 0.0], np.float32), 1.0)
    if nan_body:
        b.body(box, IDENT, np.array([np.nan, 5.0, 0.0], np.float32), 1.0)
    return b.done(dt=1 / 60, iterations=10, steps=30)


def sleep_and_wake(n=3, layers=2, drop_height=30.0):
    """Resting boxes that fall asleep (deactivation enabled) and a sphere that lands on them later."""
    b = _Builder("sleep_and_wake")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    b.body(plane, IDENT, np.zeros(3, np.float32), 0.0)
    box = b.shape(BOX, [0.5, 0.5, 0.5])
    for i in range(n):
        for j in range(n):
            for k in range(layers):
                b.body(box, IDENT, np.array([i * 1.5, 0.5 + k * 1.0, j * 1.5], np.float32), 1.0)
    sphere = b.shape(SPHERE, [0.4])
    b.body(sphere, IDENT, np.array([(n - 1) * 0.75, drop_height, (n - 1) * 0.75], np.float32), 1.0)
    return b.done(dt=1 / 60, iterations=10, steps=400)


def chains(seed=1, nx=8, nz=8, links=8, spacing=0.9, drops=200):
    """Chains of boxes joined by Point2Point constraints: nx x nz chains of `links` links (half extents
    0.15 x 0.2 x 0.15, 0.5 m apart), each hung from a fixed world point; every other chain reaches the floor, the
    rest hang free; `drops` boxes (half-extent 0.2-0.4) fall onto them from 2-8 m above.  The links start with
    random horizontal velocities (meta["velocity"]), so chains swing into their neighbours.  meta["joints"] =
    (body_a, body_b, pivot_a, pivot_b), body_b -1 for the fixed body."""
    rng = np.random.default_rng(seed)
    b = _Builder("chains")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    b.body(plane, IDENT, np.zeros(3, np.float32), 0.0)
    link = b.shape(BOX, [0.15, 0.2, 0.15])
    ja, jb, pa, pb = [], [], [], []
    vel = [np.zeros(3, np.float32)]
    for i in range(nx):
        for k in range(nz):
            x, z = (i - (nx - 1) / 2.0) * spacing, (k - (nz - 1) / 2.0) * spacing
            top = links * 0.5 + (0.0 if (i + k) % 2 == 0 else 0.6)  # the even chains' last link touches the floor
            for m in range(links):
                idx = len(b.bs)
                b.body(link, IDENT, np.array([x, top - 0.25 - 0.5 * m, z], np.float32), 1.0)
                vel.append(np.array([rng.uniform(-1, 1), 0.0, rng.uniform(-1, 1)], np.float32))
                if m == 0:
                    ja.append(idx), jb.append(-1), pa.append([0, 0.25, 0]), pb.append([x, top, z])
                else:
                    ja.append(idx - 1), jb.append(idx), pa.append([0, -0.25, 0]), pb.append([0, 0.25, 0])
    ext = (nx * spacing / 2.0, nz * spacing / 2.0)
    for _ in range(drops):
        h = rng.uniform(0.2, 0.4, 3).astype(np.float32)
        bx = b.shape(BOX, list(h))
        pos = np.array([rng.uniform(-ext[0], ext[0]), links * 0.5 + rng.uniform(2.0, 8.0),
                        rng.uniform(-ext[1], ext[1])], np.float32)
        b.body(bx, _yaw_basis(rng.uniform(0, 2 * np.pi, 1))[0], pos, np.float32(8.0 * h.prod() * 200.0))
        vel.append(np.zeros(3, np.float32))
    joints = (np.array(ja, np.int32), np.array(jb, np.int32), np.array(pa, np.float32), np.array(pb, np.float32))
    return b.done(dt=1 / 60, iterations=10, steps=300, joints=joints, velocity=np.array(vel, np.float32))


def hinges(seed=1, n_chains=12, links=5, n_doors=24, n_wheels=12, drops=120):
    """Hinged mechanisms on a floor: chains of boxes hinged end to end about z (the top link to the world),
    doors hinged to the world about a vertical edge (resting on the floor, half of them limited to +-1 rad, a
    quarter driven by a motor), wheels hinged at their centre to static posts (motors on, some angular-only), and
    `drops` boxes falling onto all of it; random initial spins.  meta["hinges"]: one tuple per hinge,
    (body_a, body_b, pivot_a, axis_a, pivot_b, axis_b, frame_a, params): pivot/axis constructor when frame_a is
    None (pivot_b None: the fixed body), else the frame constructor (rbA, rbAFrame) with B the fixed body;
    params (9 floats) or None.  meta["velocity"], meta["spin"]: initial linear / angular velocities."""
    rng = np.random.default_rng(seed)
    b = _Builder("hinges")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    b.body(plane, IDENT, np.zeros(3, np.float32), 0.0)
    vel, spin, hs = [np.zeros(3)], [np.zeros(3)], []
    z = np.array([0.0, 0.0, 1.0], np.float32)
    link = b.shape(BOX, [0.15, 0.2, 0.1])
    for c in range(n_chains):
        x = (c - (n_chains - 1) / 2.0) * 1.1
        top = links * 0.5 + 0.3 * (c % 2)
        for m in range(links):
            idx = len(b.bs)
            b.body(link, IDENT, np.array([x, top - 0.25 - 0.5 * m, -6.0], np.float32), 1.0)
            vel.append(np.zeros(3))
            spin.append(np.array([0.0, 0.0, rng.uniform(-2, 2)]))
            if m == 0:
                hs.append((idx, -1, [0, 0.25, 0], z, None, None, None, None))
            else:
                hs.append((idx - 1, idx, [0, -0.25, 0], z, [0, 0.25, 0], z, None, None))
    door = b.shape(BOX, [0.4, 0.6, 0.05])
    y = np.array([0.0, 1.0, 0.0], np.float32)
    for d in range(n_doors):
        idx = len(b.bs)
        pos = np.array([(d % 6) * 1.6 - 4.0, 0.6, (d // 6) * 1.6 - 1.0], np.float32)
        b.body(door, IDENT, pos, 4.0)
        vel.append(np.zeros(3))
        spin.append(np.array([0.0, rng.uniform(-3, 3), 0.0]))
        params = None
        if d % 2 == 0:
            params = [-1.0, 1.0, 0.9, 0.3, 1.0, 0.0, 0.0, 0.0, 0.0]
        if d % 4 == 1:
            params = [1e30, -1e30, 0.9, 0.3, 1.0, 0.0, 1.0, rng.uniform(-2, 2), 0.5]
        if d % 3 == 0:  # the frame constructor: the hinge angle starts at 0, so the limit is about the rest pose
            F = np.concatenate([np.array([1, 0, 0, 0, 0, -1, 0, 1, 0], np.float32), [-0.4, 0.0, 0.0]])
            hs.append((idx, -1, None, None, None, None, F.astype(np.float32), params))
        else:
            hs.append((idx, -1, [-0.4, 0.0, 0.0], y, None, None, None, params))
    wheel = b.shape(BOX, [0.35, 0.35, 0.08])
    post = b.shape(BOX, [0.05, 0.05, 0.05])
    for w in range(n_wheels):
        pos = np.array([(w % 6) * 1.6 - 4.0, 0.45, 5.0 + (w // 6) * 1.2], np.float32)
        pidx = len(b.bs)
        b.body(post, IDENT, pos + np.array([0, 0, -0.3], np.float32), 0.0)
        vel.append(np.zeros(3))
        spin.append(np.zeros(3))
        idx = len(b.bs)
        b.body(wheel, IDENT, pos, 2.0)
        vel.append(np.zeros(3))
        spin.append(np.array([0.0, 0.0, rng.uniform(-4, 4)]))
        params = [1e30, -1e30, 0.9, 0.3, 1.0, float(w % 3 == 0), 1.0, rng.uniform(-5, 5), 0.3]
        hs.append((idx, pidx, [0, 0, 0], z, [0, 0, 0.3], z, None, params))
    ext = 6.0
    for _ in range(drops):
        h = rng.uniform(0.15, 0.3, 3).astype(np.float32)
        bx = b.shape(BOX, list(h))
        p = np.array([rng.uniform(-ext, ext), rng.uniform(3.0, 8.0), rng.uniform(-7.0, 7.0)], np.float32)
        b.body(bx, _yaw_basis(rng.uniform(0, 2 * np.pi, 1))[0], p, np.float32(8.0 * h.prod() * 300.0))
        vel.append(np.zeros(3))
        spin.append(np.zeros(3))
    return b.done(dt=1 / 60, iterations=10, steps=300, hinges=hs, velocity=np.array(vel, np.float32),
                  spin=np.array(spin, np.float32))


def projectiles(seed=1, lanes=6, shots=40, speed=(60.0, 300.0)):
    """Fast small bodies fired at thin walls, for continuous collision detection.  Per lane (2.5 m apart along
    z): a static wall 0.1 m thick at x = 8, a dynamic plank 0.1 m thick standing at x = 14 and a stack of three
    boxes at x = 20, on a floor.  `shots` projectiles (spheres of radius 0.1-0.15, boxes of half-extent 0.1,
    tetrahedral hulls of 0.15) start at x = -2 to 0, 0.4-2 m up, aimed along +x at 60-300 m/s (2-5 m per 1/60 s
    step, far more than the wall or plank is thick) with a little spread, some spinning; a few are fired
    downwards at the floor.  meta["velocity"], meta["angular"]: initial velocities; meta["ccd"] = (bodies,
    motion threshold, swept sphere radius) for the projectiles."""
    rng = np.random.default_rng(seed)
    b = _Builder("projectiles")
    plane = b.shape(PLANE, [0.0, 1.0, 0.0, 0.0])
    b.body(plane, IDENT, np.zeros(3, np.float32), 0.0)
    wall = b.shape(BOX, [0.05, 2.0, 1.0])
    plank = b.shape(BOX, [0.05, 1.0, 0.8])
    block = b.shape(BOX, [0.4, 0.4, 0.4])
    for k in range(lanes):
        z = (k - (lanes - 1) / 2.0) * 2.5
        b.body(wall, IDENT, np.array([8.0, 2.0, z], np.float32), 0.0)
        b.body(plank, IDENT, np.array([14.0, 1.0, z], np.float32), 2.0)
        for m in range(3):
            b.body(block, IDENT, np.array([20.0, 0.4 + 0.8 * m, z], np.float32), 5.0)
    nstat = len(b.bs)
    vel = [np.zeros(3, np.float32)] * nstat
    ang = [np.zeros(3, np.float32)] * nstat
    sph = [b.shape(SPHERE, [r]) for r in (0.1, 0.12, 0.15)]
    cube = b.shape(BOX, [0.1, 0.1, 0.1])
    tet = b.hull(np.array([[0.15, 0.15, 0.15], [-0.15, -0.15, 0.15], [-0.15, 0.15, -0.15], [0.15, -0.15, -0.15]],
                          np.float32))
    ccd_bodies, thr, rad = [], [], []
    span = (lanes - 1) / 2.0 * 2.5 + 0.5
    for i in range(shots):
        kind = i % 5
        shape = sph[i % 3] if kind < 3 else (cube if kind == 3 else tet)
        pos = np.array([rng.uniform(-2.0, 0.0), rng.uniform(0.4, 2.0), rng.uniform(-span, span)], np.float32)
        v = np.float32(rng.uniform(*speed))
        if i % 8 == 7:  # fired down at the floor
            d = np.array([rng.uniform(0.2, 0.6), -1.0, rng.uniform(-0.2, 0.2)], np.float32)
            pos[1] = rng.uniform(3.0, 6.0)
        else:
            d = np.array([1.0, rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05)], np.float32)
        d = d / np.float32(np.linalg.norm(d))
        ccd_bodies.append(len(b.bs))
        b.body(shape, _quat_basis(_random_quats(rng, 1))[0] if kind >= 3 else IDENT, pos, 0.2)
        vel.append((d * v).astype(np.float32))
        ang.append((rng.normal(size=3) * 20.0).astype(np.float32) if i % 3 == 0 else np.zeros(3, np.float32))
        thr.append(0.05)
        rad.append(0.08)
    ccd = (np.array(ccd_bodies, np.int32), np.array(thr, np.float32), np.array(rad, np.float32))
    return b.done(dt=1 / 60, iterations=10, steps=90, velocity=np.array(vel, np.float32),
                  angular=np.array(ang, np.float32), ccd=ccd)


def build_records(scene, builder, gravity=GRAVITY):
    """Shape and body records for ``scene``.

    ``builder`` provides ``make_shape(type, params, verts) -> (20,) float32`` and
    ``shape_inertia(record, mass, verts) -> (3,) float32`` with the semantics of the
    reference's shape constructors and ``calculateLocalInertia``.
    Body setup follows RigidBody.setMassProps / setGravity (dynamics/RigidBody.java:
    207-217, 301-313): invMass = 1/mass, inverse inertia 1/I (0 where I == 0),
    gravity force = g * (1/invMass) for dynamic bodies.
    """
    S = scene.shape_type.shape[0]
    verts = scene.verts if scene.verts.size else np.zeros((1, 3), np.float32)
    shapes = np.zeros((S, SHAPE_WORDS), np.float32)
    for i in range(S):
        shapes[i] = builder.make_shape(int(scene.shape_type[i]), scene.shape_param[i], verts)
    N = scene.n_bodies
    bodies = np.zeros((N, BODY_WORDS), np.float32)
    bodies[:, 0:9] = scene.basis
    bodies[:, 9:12] = scene.origin
    one = np.float32(1.0)
    iview = bodies.view(np.int32)
    for i in range(N):
        m = np.float32(scene.mass[i])
        rec = shapes[scene.body_shape[i]]
        if m != 0:
            inertia = builder.shape_inertia(rec, float(m), verts)
            inv_mass = one / m
            bodies[i, 18] = inv_mass
            bodies[i, 19:22] = [one / x if x != 0 else np.float32(0) for x in np.asarray(inertia, np.float32)]
            g = np.asarray(gravity, np.float32)
            bodies[i, 22:25] = g * (one / inv_mass)
        iview[i, 28] = 0 if m != 0 else 1
    bodies[:, 25] = scene.friction
    bodies[:, 26] = scene.restitution
    iview[:, 27] = scene.body_shape
    iview[:, 29] = ACTIVE_TAG
    return shapes, verts.astype(np.float32), bodies
