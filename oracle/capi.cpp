// Oracle (test infrastructure only): extern "C" entry points used by tests/
// (through ctypes) and by bench.py's cpu_baseline leg.  Record layouts are the
// ones documented in include/jbgpu_format.h.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "oracle.h"

using namespace jb;

namespace {

V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
void st3(float* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
int32_t ldi(const float* p) { int32_t i; std::memcpy(&i, p, 4); return i; }
void sti(float* p, int32_t i) { std::memcpy(p, &i, 4); }

Shape shapeFromRecord(const float* r) {
  Shape s{};
  s.type = ldi(r + 0);
  s.margin = r[1];
  s.dims = ld3(r + 2);
  s.vfirst = ldi(r + 5);
  s.vcount = ldi(r + 6);
  s.lmin = ld3(r + 7);
  s.lmax = ld3(r + 10);
  s.pn = ld3(r + 13);
  s.pc = r[16];
  return s;
}
void shapeToRecord(const Shape& s, float* r) {
  std::memset(r, 0, sizeof(float) * JBGPU_SHAPE_WORDS);
  sti(r + 0, s.type);
  r[1] = s.margin;
  st3(r + 2, s.dims);
  sti(r + 5, s.vfirst);
  sti(r + 6, s.vcount);
  st3(r + 7, s.lmin);
  st3(r + 10, s.lmax);
  st3(r + 13, s.pn);
  r[16] = s.pc;
}
Tr trFrom(const float* p) {
  Tr t;
  for (int i = 0; i < 9; i++) t.b.m[i] = p[i];
  t.o = ld3(p + 9);
  return t;
}
void trTo(const Tr& t, float* p) {
  for (int i = 0; i < 9; i++) p[i] = t.b.m[i];
  st3(p + 9, t.o);
}
Body bodyFromRecord(const float* r) {
  Body b{};
  b.tr = trFrom(r);
  b.lv = ld3(r + 12);
  b.av = ld3(r + 15);
  b.invMass = r[18];
  b.invInertiaLocal = ld3(r + 19);
  b.gravity = ld3(r + 22);
  b.friction = r[25];
  b.restitution = r[26];
  b.shape = ldi(r + 27);
  b.flags = ldi(r + 28);
  b.activation = ldi(r + 29);
  b.deactivationTime = r[30];
  return b;
}
void bodyToRecord(const Body& b, float* r) {
  std::memset(r, 0, sizeof(float) * JBGPU_BODY_WORDS);
  trTo(b.tr, r);
  st3(r + 12, b.lv);
  st3(r + 15, b.av);
  r[18] = b.invMass;
  st3(r + 19, b.invInertiaLocal);
  st3(r + 22, b.gravity);
  r[25] = b.friction;
  r[26] = b.restitution;
  sti(r + 27, b.shape);
  sti(r + 28, b.flags);
  sti(r + 29, b.activation);
  r[30] = b.deactivationTime;
}
void pointTo(const MPoint& p, float* r) {
  std::memset(r, 0, sizeof(float) * JBGPU_POINT_WORDS);
  st3(r + 0, p.localA);
  st3(r + 3, p.localB);
  st3(r + 6, p.worldA);
  st3(r + 9, p.worldB);
  st3(r + 12, p.normal);
  r[15] = p.distance;
  r[16] = p.friction;
  r[17] = p.restitution;
  r[18] = p.applied;
  r[19] = p.lat1;
  r[20] = p.lat2;
  if (p.latInit) {
    st3(r + 21, p.latDir1);
    st3(r + 24, p.latDir2);
  }
  sti(r + 27, p.latInit ? 1 : 0);
  sti(r + 28, p.lifetime);
}
MPoint pointFrom(const float* r) {
  MPoint p{};
  p.localA = ld3(r + 0);
  p.localB = ld3(r + 3);
  p.worldA = ld3(r + 6);
  p.worldB = ld3(r + 9);
  p.normal = ld3(r + 12);
  p.distance = r[15];
  p.friction = r[16];
  p.restitution = r[17];
  p.applied = r[18];
  p.lat1 = r[19];
  p.lat2 = r[20];
  p.latDir1 = ld3(r + 21);
  p.latDir2 = ld3(r + 24);
  p.latInit = ldi(r + 27) != 0;
  p.lifetime = ldi(r + 28);
  return p;
}
void manifoldTo(const Manifold& m, int type, float* r) {
  std::memset(r, 0, sizeof(float) * JBGPU_MANIFOLD_WORDS);
  sti(r + 0, m.n);
  sti(r + 1, m.body0);
  sti(r + 2, m.body1);
  sti(r + 3, type);
  for (int i = 0; i < m.n; i++) pointTo(m.p[i], r + JBGPU_MANIFOLD_HDR + i * JBGPU_POINT_WORDS);
}
Manifold manifoldFrom(const float* r) {
  Manifold m{};
  m.n = ldi(r + 0);
  m.body0 = ldi(r + 1);
  m.body1 = ldi(r + 2);
  for (int i = 0; i < m.n; i++) m.p[i] = pointFrom(r + JBGPU_MANIFOLD_HDR + i * JBGPU_POINT_WORDS);
  return m;
}
int pairTypeOf(const World& w, uint64_t key) {
  int a = (int)(key >> 32) - 1, b = (int)(key & 0xffffffffu) - 1;
  int ta = w.shapes[w.bodies[a].shape].type, tb = w.shapes[w.bodies[b].shape].type;
  if (ta == JBGPU_SHAPE_SPHERE && tb == JBGPU_SHAPE_SPHERE) return JBGPU_PAIR_SPHERE_SPHERE;
  if (ta == JBGPU_SHAPE_PLANE && tb == JBGPU_SHAPE_PLANE) return JBGPU_PAIR_NONE;
  if (ta == JBGPU_SHAPE_PLANE || tb == JBGPU_SHAPE_PLANE) return JBGPU_PAIR_CONVEX_PLANE;
  return JBGPU_PAIR_CONVEX_CONVEX;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- math / rng
int32_t orc_random_next_int(int64_t seed, int32_t n, int32_t* out) {
  JavaRandom r(seed);
  for (int i = 0; i < n; i++) out[i] = r.nextInt();
  return 0;
}
// consecutive nextInt(bounds[i]) draws from one new Random(seed)
int32_t orc_random_next_bounded(int64_t seed, int32_t n, const int32_t* bounds, int32_t* out) {
  JavaRandom r(seed);
  for (int i = 0; i < n; i++) out[i] = r.nextInt(bounds[i]);
  return 0;
}
void orc_quat_from_matrix(const float* m9, float* q4out) {
  M3 m;
  for (int i = 0; i < 9; i++) m.m[i] = m9[i];
  Q q = qFromMatrix(m);
  q4out[0] = q.x; q4out[1] = q.y; q4out[2] = q.z; q4out[3] = q.w;
}
void orc_matrix_from_quat(const float* q4in, float* m9) {
  M3 m = m3FromQ(q4(q4in[0], q4in[1], q4in[2], q4in[3]));
  for (int i = 0; i < 9; i++) m9[i] = m.m[i];
}

// ---------------------------------------------------------------- shapes
// type, params: sphere (r); box (hx,hy,hz); capsule (r, height); hull (first, count);
// plane (nx,ny,nz,c). Hull points are read from verts.
void orc_make_shape(int type, const float* params, const float* verts, float* rec) {
  Shape s{};
  switch (type) {
    case JBGPU_SHAPE_SPHERE: s = makeSphere(params[0]); break;
    case JBGPU_SHAPE_BOX: s = makeBox(v3(params[0], params[1], params[2])); break;
    case JBGPU_SHAPE_CAPSULE: s = makeCapsule(params[0], params[1]); break;
    case JBGPU_SHAPE_HULL: s = makeHull((const V3*)verts, (int)params[0], (int)params[1]); break;
    case JBGPU_SHAPE_PLANE: s = makePlane(v3(params[0], params[1], params[2]), params[3]); break;
  }
  shapeToRecord(s, rec);
}
void orc_support(const float* rec, const float* verts, const float* dir, int withMargin, float* out) {
  Shape s = shapeFromRecord(rec);
  V3 v = withMargin ? supportMargin(s, (const V3*)verts, ld3(dir)) : supportNoMargin(s, (const V3*)verts, ld3(dir));
  st3(out, v);
}
void orc_aabb(const float* rec, const float* verts, const float* tr12, float* out6) {
  (void)verts;
  Shape s = shapeFromRecord(rec);
  V3 mn, mx;
  shapeAabb(s, trFrom(tr12), mn, mx);
  st3(out6, mn);
  st3(out6 + 3, mx);
}
void orc_inertia(const float* rec, const float* verts, float mass, float* out3) {
  st3(out3, shapeLocalInertia(shapeFromRecord(rec), (const V3*)verts, mass));
}
// narrowphase unit entry points; out = [valid, nx,ny,nz, px,py,pz, depth]
void orc_gjk(const float* recA, const float* trA, const float* recB, const float* trB, const float* verts,
             float maxDist2, float* out) {
  ContactCandidate c = gjkPairDetector(shapeFromRecord(recA), trFrom(trA), shapeFromRecord(recB), trFrom(trB),
                                       (const V3*)verts, maxDist2, nullptr);
  out[0] = c.valid ? 1.0f : 0.0f;
  st3(out + 1, c.normalOnB);
  st3(out + 4, c.pointOnB);
  out[7] = c.depth;
}
// GjkEpaSolver.collide with radialmargin 0: out = [ok, wA xyz, wB xyz, epaIters]
void orc_epa(const float* recA, const float* trA, const float* recB, const float* trB, const float* verts, float* out) {
  V3 a, b;
  int it = 0;
  bool ok = gjkEpaPenDepth(shapeFromRecord(recA), trFrom(trA), shapeFromRecord(recB), trFrom(trB), (const V3*)verts, a,
                           b, &it);
  out[0] = ok ? 1.0f : 0.0f;
  st3(out + 1, a);
  st3(out + 4, b);
  out[7] = (float)it;
}
void orc_sphere_sphere(const float* recA, const float* trA, const float* recB, const float* trB, float* out) {
  bool touching;
  ContactCandidate c = sphereSphere(shapeFromRecord(recA), trFrom(trA), shapeFromRecord(recB), trFrom(trB), touching);
  out[0] = touching ? 1.0f : 0.0f;
  st3(out + 1, c.normalOnB);
  st3(out + 4, c.pointOnB);
  out[7] = c.depth;
}
void orc_convex_plane(const float* recC, const float* trC, const float* recP, const float* trP, const float* verts,
                      float* out) {
  ContactCandidate c = convexPlane(shapeFromRecord(recC), (const V3*)verts, trFrom(trC), shapeFromRecord(recP), trFrom(trP));
  out[0] = c.valid ? 1.0f : 0.0f;
  st3(out + 1, c.normalOnB);
  st3(out + 4, c.pointOnB);
  out[7] = c.depth;
}
// manifold: apply a sequence of candidates to a manifold record (body transforms given)
void orc_manifold_add(float* mrec, const float* trA, const float* trB, const float* fr2, const float* cand,
                      int refresh) {
  Manifold m = manifoldFrom(mrec);
  Body b0{}, b1{};
  b0.tr = trFrom(trA);
  b1.tr = trFrom(trB);
  b0.friction = fr2[0];
  b1.friction = fr2[1];
  b0.restitution = fr2[2];
  b1.restitution = fr2[3];
  m.body0 = 0;
  m.body1 = 1;
  if (cand[0] != 0.0f) {
    ContactCandidate c{true, ld3(cand + 1), ld3(cand + 4), cand[7]};
    addContactPoint(m, b0, b1, 0, c);
  }
  if (refresh && m.n) manifoldRefresh(m, b0.tr, b1.tr);
  manifoldTo(m, ldi(mrec + 3), mrec);
}

// ---------------------------------------------------------------- broadphase
void* orc_bp_new(int mode, const float* wmin, const float* wmax) {
  Broadphase* bp = new Broadphase();
  V3 a = wmin ? ld3(wmin) : v3(-1000, -1000, -1000), b = wmax ? ld3(wmax) : v3(1000, 1000, 1000);
  bp->init(mode, a, b);
  return bp;
}
void orc_bp_free(void* h) { delete (Broadphase*)h; }
void orc_bp_create(void* h, int n, const float* aabb6, const int16_t* group, const int16_t* mask) {
  Broadphase* bp = (Broadphase*)h;
  for (int i = 0; i < n; i++)
    bp->createProxy(Aabb{ld3(aabb6 + 6 * i), ld3(aabb6 + 6 * i + 3)}, group ? group[i] : 1, mask ? mask[i] : -1);
}
void orc_bp_destroy(void* h, int n, const int32_t* uids) {
  Broadphase* bp = (Broadphase*)h;
  for (int i = 0; i < n; i++) bp->destroyProxy(uids[i] - 1);
}
void orc_bp_set_worlds(void* h, const int32_t* world) {
  Broadphase* bp = (Broadphase*)h;
  for (size_t i = 0; i < bp->proxies.size(); i++) bp->proxies[i].world = world[i];
}
void orc_bp_set_all(void* h, const float* aabb6) {
  Broadphase* bp = (Broadphase*)h;
  for (size_t i = 0; i < bp->proxies.size(); i++) bp->setAabb((int)i, Aabb{ld3(aabb6 + 6 * i), ld3(aabb6 + 6 * i + 3)});
}
// calculateOverlappingPairs: collide() then the pair set; returns the count, writes min(count, cap) keys
int64_t orc_bp_step(void* h, uint64_t* keys, int64_t cap, int brute) {
  Broadphase* bp = (Broadphase*)h;
  bp->collide();
  std::vector<uint64_t> out;
  if (brute)
    bp->pairsBrute(out);
  else
    bp->pairs(out);
  for (int64_t i = 0; i < (int64_t)out.size() && i < cap; i++) keys[i] = out[i];
  return (int64_t)out.size();
}
void orc_bp_get_proxies(void* h, float* rec) {
  Broadphase* bp = (Broadphase*)h;
  for (size_t i = 0; i < bp->proxies.size(); i++) {
    const Proxy& p = bp->proxies[i];
    float* r = rec + i * JBGPU_PROXY_WORDS;
    std::memset(r, 0, sizeof(float) * JBGPU_PROXY_WORDS);
    sti(r + 0, p.stage);
    st3(r + 1, p.leaf.mn);
    st3(r + 4, p.leaf.mx);
    st3(r + 7, p.aabb.mn);
    st3(r + 10, p.aabb.mx);
    sti(r + 13, p.group);
    sti(r + 14, p.mask);
  }
}
void orc_bp_get_quantized(void* h, int32_t* q6) {
  Broadphase* bp = (Broadphase*)h;
  for (size_t i = 0; i < bp->proxies.size(); i++)
    for (int k = 0; k < 6; k++) q6[i * 6 + k] = bp->proxies[i].q[k];
}
int orc_bp_stage_current(void* h) { return ((Broadphase*)h)->stageCurrent; }

// ---------------------------------------------------------------- world
struct OrcWorldCfg {
  int32_t bpMode;
  int32_t numIterations;
  int32_t solverOrder;
  int32_t deactivationDisabled;
  float wmin[3], wmax[3];
  float tolerance;
  int32_t maxIterations;
  int32_t solverMode;
};
void* orc_world_new(const OrcWorldCfg* cfg, int nshapes, const float* shapes, int nverts, const float* verts,
                    int nbodies, const float* bodies) {
  World* w = new World();
  w->cfg.bpMode = cfg->bpMode;
  w->cfg.worldMin = ld3(cfg->wmin);
  w->cfg.worldMax = ld3(cfg->wmax);
  w->cfg.solver.numIterations = cfg->numIterations;
  w->cfg.solver.tolerance = cfg->tolerance;
  w->cfg.solver.maxIterations = cfg->maxIterations;
  if (cfg->solverMode) w->cfg.solver.solverMode = cfg->solverMode;
  w->cfg.solverOrder = cfg->solverOrder;
  w->cfg.deactivationDisabled = cfg->deactivationDisabled != 0;
  w->bp.init(cfg->bpMode, w->cfg.worldMin, w->cfg.worldMax);
  for (int i = 0; i < nshapes; i++) w->shapes.push_back(shapeFromRecord(shapes + i * JBGPU_SHAPE_WORDS));
  for (int i = 0; i < nverts; i++) w->hullv.push_back(ld3(verts + 3 * i));
  for (int i = 0; i < nbodies; i++) w->addBody(bodyFromRecord(bodies + i * JBGPU_BODY_WORDS));
  return w;
}
void orc_world_free(void* h) { delete (World*)h; }
void orc_world_set_order(void* h, int n, const int32_t* order) {
  World* w = (World*)h;
  w->explicitOrder.assign(order, order + n);
}
void orc_world_set_solver(void* h, int order, float tol, int maxIt, int numIt) {
  World* w = (World*)h;
  w->cfg.solverOrder = order;
  w->cfg.solver.tolerance = tol;
  w->cfg.solver.maxIterations = maxIt;
  w->cfg.solver.numIterations = numIt;
}
// world id per body (batched independent worlds; bodies of different worlds never pair)
// rayTest with a ClosestRayResultCallback per ray: object index (-1: none), fraction, normal, point
void orc_world_ray_test(void* h, int n, const float* from3, const float* to3, int16_t group, int16_t mask,
                        const int32_t* world, int32_t* obj, float* frac, float* normal3, float* point3) {
  World* w = (World*)h;
  for (int i = 0; i < n; i++) {
    RayHit r = rayTestClosest(*w, ld3(from3 + 3 * i), ld3(to3 + 3 * i), group, mask, world ? world[i] : 0);
    obj[i] = r.object;
    frac[i] = r.fraction;
    st3(normal3 + 3 * i, r.normal);
    st3(point3 + 3 * i, r.point);
  }
}
// convexSweepTest with a ClosestConvexResultCallback per query: cast shape = shape index of the world, from / to
// 12 floats each (basis row-major, origin)
void orc_world_convex_sweep(void* h, int n, int shape, const float* from12, const float* to12, int16_t group,
                            int16_t mask, const int32_t* world, int32_t* obj, float* frac, float* normal3,
                            float* point3) {
  World* w = (World*)h;
  for (int i = 0; i < n; i++) {
    RayHit r = convexSweepClosest(*w, w->shapes[shape], trFrom(from12 + 12 * i), trFrom(to12 + 12 * i), group, mask,
                                  world ? world[i] : 0);
    obj[i] = r.object;
    frac[i] = r.fraction;
    st3(normal3 + 3 * i, r.normal);
    st3(point3 + 3 * i, r.point);
  }
}
// CollisionObject.setCcdMotionThreshold / setCcdSweptSphereRadius of bodies
void orc_world_set_ccd(void* h, int n, const int32_t* body, const float* threshold, const float* radius) {
  World* w = (World*)h;
  for (int i = 0; i < n; i++) {
    w->bodies[body[i]].ccdMotionThreshold = threshold[i];
    w->bodies[body[i]].ccdSweptSphereRadius = radius[i];
  }
}
// TransformUtil.calculateVelocity and CollisionShape.calculateTemporalAabb (of a shape record) for the pins
void orc_calculate_velocity(const float* t0, const float* t1, float dt, float* lin3, float* ang3) {
  V3 l, a;
  calculateVelocity(trFrom(t0), trFrom(t1), dt, l, a);
  st3(lin3, l);
  st3(ang3, a);
}
void orc_temporal_aabb(const float* rec, const float* tr12, const float* lin3, const float* ang3, float dt, float* mn3,
                       float* mx3) {
  V3 mn, mx;
  temporalAabb(shapeFromRecord(rec), trFrom(tr12), ld3(lin3), ld3(ang3), dt, mn, mx);
  st3(mn3, mn);
  st3(mx3, mx);
}
// GjkConvexCast of shape record a from fromA to toA against shape record b at tb (hull points hv, may be null)
int orc_gjk_convex_cast(const float* recA, const float* fromA, const float* toA, const float* recB, const float* tb,
                        const float* hv, float allowedPenetration, float* out) {
  V3 n = v3(0, 0, 0), p = v3(0, 0, 0);
  float f = 1.0f;
  const bool hit = gjkConvexCast(shapeFromRecord(recA), trFrom(fromA), trFrom(toA), shapeFromRecord(recB), trFrom(tb),
                                 (const V3*)hv, allowedPenetration, f, n, p);
  out[0] = f;
  st3(out + 1, n);
  st3(out + 4, p);
  return hit ? 1 : 0;
}
void orc_world_set_worlds(void* h, const int32_t* world) {
  World* w = (World*)h;
  for (size_t i = 0; i < w->bp.proxies.size(); i++) w->bp.proxies[i].world = world[i];
}
int orc_world_step(void* h, float dt, int nsteps) {
  World* w = (World*)h;
  for (int i = 0; i < nsteps; i++) w->step(dt);
  return 0;
}
// run the step phase by phase: stop after phase p (0 predict, 1 aabb+bp, 2 narrowphase, 3 islands+solve, 4 full)
int orc_world_step_until(void* h, float dt, int phase) {
  World* w = (World*)h;
  w->predictUnconstraintMotion(dt);
  if (phase == 0) return 0;
  w->updateAabbs();
  w->calculateOverlappingPairs();
  if (phase == 1) return 0;
  w->dispatchAllCollisionPairs();
  if (phase == 2) return 0;
  w->calculateSimulationIslands();
  w->solveConstraints();
  if (phase == 3) return 0;
  w->integrateTransforms(dt);
  w->updateActivationState(dt);
  w->stepCount++;
  return 0;
}
// Point2PointConstraint(rbA, rbB, pivotInA, pivotInB) added with DiscreteDynamicsWorld.addConstraint; bodyB -1:
// TypedConstraint's fixed body (pivotB a world point); settings per joint (tau, damping, impulseClamp) or null
void orc_world_add_p2p(void* h, int n, const int32_t* bodyA, const int32_t* bodyB, const float* pivotA,
                       const float* pivotB, const float* settings) {
  World* w = (World*)h;
  for (int i = 0; i < n; i++) {
    Joint j{};
    j.bodyA = bodyA[i];
    j.bodyB = bodyB[i];
    j.pivotA = ld3(pivotA + 3 * i);
    j.pivotB = ld3(pivotB + 3 * i);
    j.tau = settings ? settings[3 * i] : 0.3f;
    j.damping = settings ? settings[3 * i + 1] : 1.0f;
    j.impulseClamp = settings ? settings[3 * i + 2] : 0.0f;
    w->joints.push_back(j);
  }
}
// HingeConstraints from the solver frames (rbAFrame, rbBFrame as the constructors leave them, 12 floats each:
// basis row-major, origin); bodyB -1 for the fixed body; params per hinge (lowerLimit, upperLimit, limitSoftness,
// biasFactor, relaxationFactor, angularOnly, enableAngularMotor, motorTargetVelocity, maxMotorImpulse) or null
void orc_world_add_hinge(void* h, int n, const int32_t* bodyA, const int32_t* bodyB, const float* frameA,
                         const float* frameB, const float* params) {
  World* w = (World*)h;
  for (int i = 0; i < n; i++) {
    Joint j{};
    j.type = 1;
    j.bodyA = bodyA[i];
    j.bodyB = bodyB[i];
    j.frameA = trFrom(frameA + 12 * i);
    j.frameB = trFrom(frameB + 12 * i);
    if (params) {
      const float* p = params + 9 * i;
      j.lowerLimit = p[0];
      j.upperLimit = p[1];
      j.limitSoftness = p[2];
      j.biasFactor = p[3];
      j.relaxationFactor = p[4];
      j.angularOnly = p[5] != 0.0f;
      j.enableMotor = p[6] != 0.0f;
      j.motorTarget = p[7];
      j.maxMotorImpulse = p[8];
    }
    w->joints.push_back(j);
  }
}
// the constructors' frames (HingeConstraint.java:84-238) from body A's transform trA (12 floats):
// fromFrames = 0: pivot/axis constructors (fixedB: (rbA, pivotInA, axisInA)); 1: frame constructors
void orc_hinge_frames(int fromFrames, int fixedB, const float* trA, const float* pivotA, const float* pivotB,
                      const float* axisA, const float* axisB, const float* inA, const float* inB, float* outA,
                      float* outB) {
  Tr fa, fb;
  if (fromFrames)
    hingeFramesFromFrames(trFrom(trA), trFrom(inA), fixedB ? trFrom(inA) : trFrom(inB), fixedB != 0, fa, fb);
  else
    hingeFramesFromAxes(trFrom(trA), ld3(pivotA), fixedB ? v3(0, 0, 0) : ld3(pivotB), ld3(axisA),
                        fixedB ? v3(0, 0, 0) : ld3(axisB), fixedB != 0, fa, fb);
  trTo(fa, outA);
  trTo(fb, outB);
}
float orc_atan2_fast(float y, float x) { return atan2Fast(y, x); }
float orc_world_hinge_angle(void* h, int j) {
  World* w = (World*)h;
  return hingeAngle(w->joints[j], w->bodies);
}
void orc_world_set_joint_order(void* h, int n, const int32_t* order) {
  ((World*)h)->explicitJointOrder.assign(order, order + n);
}
// appliedImpulse of every joint after the last solve (the sum over its three axes and the iterations)
void orc_world_joint_impulses(void* h, float* out) {
  World* w = (World*)h;
  for (size_t i = 0; i < w->joints.size(); i++) out[i] = w->joints[i].applied;
}
// one phase of the step alone: 0 predictUnconstraintMotion, 1 updateAabbs + calculateOverlappingPairs,
// 2 dispatchAllCollisionPairs, 3 calculateSimulationIslands + solveConstraints, 4 integrateTransforms +
// updateActivationState (DiscreteDynamicsWorld.java:356-400)
int orc_world_phase(void* h, float dt, int phase) {
  World* w = (World*)h;
  switch (phase) {
    case 0: w->predictUnconstraintMotion(dt); break;
    case 1: w->updateAabbs(); w->calculateOverlappingPairs(); break;
    case 2: w->dispatchAllCollisionPairs(); break;
    case 3: w->calculateSimulationIslands(); w->solveConstraints(); break;
    case 4: w->integrateTransforms(dt); w->updateActivationState(dt); w->stepCount++; break;
    default: return -1;
  }
  return 0;
}
void orc_world_get_bodies(void* h, float* out) {
  World* w = (World*)h;
  for (size_t i = 0; i < w->bodies.size(); i++) bodyToRecord(w->bodies[i], out + i * JBGPU_BODY_WORDS);
}
void orc_world_set_bodies(void* h, const float* in) {
  World* w = (World*)h;
  for (size_t i = 0; i < w->bodies.size(); i++) {
    Body b = bodyFromRecord(in + i * JBGPU_BODY_WORDS);
    b.islandTag = w->bodies[i].islandTag;
    b.companionId = w->bodies[i].companionId;
    b.ccdMotionThreshold = w->bodies[i].ccdMotionThreshold;  // not part of the record
    b.ccdSweptSphereRadius = w->bodies[i].ccdSweptSphereRadius;
    updateInertiaTensor(b);
    w->bodies[i] = b;
  }
}
int64_t orc_world_num_pairs(void* h) { return (int64_t)((World*)h)->pairs.size(); }
void orc_world_get_pairs(void* h, uint64_t* keys, float* manifolds) {
  World* w = (World*)h;
  for (size_t i = 0; i < w->pairs.size(); i++) {
    if (keys) keys[i] = w->pairs[i].key;
    if (manifolds) manifoldTo(w->pairs[i].m, pairTypeOf(*w, w->pairs[i].key), manifolds + i * JBGPU_MANIFOLD_WORDS);
  }
}
void orc_world_set_pairs(void* h, int64_t n, const uint64_t* keys, const float* manifolds) {
  World* w = (World*)h;
  w->pairs.resize(n);
  for (int64_t i = 0; i < n; i++) {
    w->pairs[i].key = keys[i];
    w->pairs[i].m = manifoldFrom(manifolds + i * JBGPU_MANIFOLD_WORDS);
  }
}
void orc_world_get_proxies(void* h, float* rec) { orc_bp_get_proxies(&((World*)h)->bp, rec); }
void orc_world_set_proxies(void* h, const float* rec, int stageCurrent) {
  World* w = (World*)h;
  for (size_t i = 0; i < w->bp.proxies.size(); i++) {
    const float* r = rec + i * JBGPU_PROXY_WORDS;
    Proxy& p = w->bp.proxies[i];
    p.stage = ldi(r + 0);
    p.leaf = Aabb{ld3(r + 1), ld3(r + 4)};
    p.aabb = Aabb{ld3(r + 7), ld3(r + 10)};
  }
  w->bp.stageCurrent = stageCurrent;
}
int orc_world_stage_current(void* h) { return ((World*)h)->bp.stageCurrent; }
// stats: [iterations, residual, rows, frictionRows, gjkIters, epaCalls, epaFails, step residual]; the step
// residual is ||dl|| / max(1, ||l||) of the last iteration over the rows of every island solved in the step
void orc_world_stats(void* h, float* out) {
  World* w = (World*)h;
  out[0] = (float)w->lastSolve.iterations;
  out[1] = w->lastSolve.residual;
  out[2] = (float)w->lastSolve.rows;
  out[3] = (float)w->lastSolve.frictionRows;
  out[4] = (float)w->gjk.gjkIters;
  out[5] = (float)w->gjk.epaCalls;
  out[6] = (float)w->gjk.epaFails;
  const double nrm = std::sqrt(w->lastSolve.finalNN);
  out[7] = (float)(std::sqrt(w->lastSolve.finalDD) / (nrm > 1.0 ? nrm : 1.0));
}
// the permutation of 0..n-1 after `rounds` shuffles of one pool drawn from new Random(seed)
void orc_shuffle(int64_t seed, int n, int rounds, int32_t* out) {
  JavaRandom rng(seed);
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  for (int r = 0; r < rounds; r++) shuffleOrder(order, rng);
  for (int i = 0; i < n; i++) out[i] = order[i];
}
struct OrcSolverInfo {
  float timeStep, erp, warmstartingFactor, tolerance;
  int32_t numIterations, maxIterations, solverMode, shuffle;
  int64_t seed;
};
// SequentialImpulseConstraintSolver.solveGroup alone on caller records: nb bodies (world inertia recomputed from
// each transform; islandTag 0 for dynamic bodies, -1 for static ones, as calculateSimulationIslands leaves them),
// nm manifolds solved in the given order; bodies and manifolds written back.  out: [iterations, residual].
void orc_solve_group(int nb, float* bodies, int nm, float* manifolds, const OrcSolverInfo* si, float* out) {
  std::vector<Body> bs(nb);
  for (int i = 0; i < nb; i++) {
    bs[i] = bodyFromRecord(bodies + (size_t)i * JBGPU_BODY_WORDS);
    bs[i].islandTag = bs[i].isStatic() ? -1 : 0;
    bs[i].companionId = -1;
    updateInertiaTensor(bs[i]);
  }
  std::vector<Manifold> ms(nm);
  std::vector<int> types(nm);
  std::vector<Manifold*> mp(nm);
  for (int k = 0; k < nm; k++) {
    ms[k] = manifoldFrom(manifolds + (size_t)k * JBGPU_MANIFOLD_WORDS);
    types[k] = ldi(manifolds + (size_t)k * JBGPU_MANIFOLD_WORDS + 3);
    mp[k] = &ms[k];
  }
  SolverInfo info;
  info.timeStep = si->timeStep;
  info.erp = si->erp;
  info.warmstartingFactor = si->warmstartingFactor;
  info.tolerance = si->tolerance;
  info.numIterations = si->numIterations;
  info.maxIterations = si->maxIterations;
  info.solverMode = si->solverMode;
  JavaRandom rng(si->seed);
  SolverStats st;
  solveGroup(bs, mp, info, si->shuffle != 0, rng, &st, nullptr, nullptr);
  for (int i = 0; i < nb; i++) bodyToRecord(bs[i], bodies + (size_t)i * JBGPU_BODY_WORDS);
  for (int k = 0; k < nm; k++) manifoldTo(ms[k], types[k], manifolds + (size_t)k * JBGPU_MANIFOLD_WORDS);
  out[0] = (float)st.iterations;
  out[1] = st.residual;
}
int orc_world_residual_history(void* h, float* out, int cap) {
  World* w = (World*)h;
  int n = (int)w->lastSolve.residualHistory.size();
  for (int i = 0; i < n && i < cap; i++) out[i] = w->lastSolve.residualHistory[i];
  return n;
}
// learned warm start hooks: per-row init impulses (3 per normal row) and feature dump
void orc_world_set_learned_init(void* h, const float* init) { ((World*)h)->learnedInit = init; }
static std::vector<float> g_features;
void orc_world_enable_features(void* h, int on) {
  World* w = (World*)h;
  g_features.clear();
  w->features = on ? &g_features : nullptr;
}
int64_t orc_world_get_features(float* out, int64_t cap) {
  int64_t n = (int64_t)g_features.size();
  for (int64_t i = 0; i < n && i < cap; i++) out[i] = g_features[i];
  return n;
}
void orc_world_clear_features() { g_features.clear(); }
// wall-clock seconds of nsteps world steps (CPU baseline timing)
double orc_world_time_steps(void* h, float dt, int nsteps) {
  World* w = (World*)h;
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < nsteps; i++) w->step(dt);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
