// Oracle (test infrastructure only): DiscreteDynamicsWorld.stepSimulation with
// one substep (dynamics/DiscreteDynamicsWorld.java:290-400), restated with the
// CollisionWorld AABB update (collision/dispatch/CollisionWorld.java:195-245),
// the dispatcher (CollisionDispatcher.java, DefaultNearCallback.java), islands
// (SimulationIslandManager.java, UnionFind.java, MiscUtil.quickSort), the
// integrator (RigidBody.java, TransformUtil.java) and activation.
//
// Canonical orders: the reference's pair-array and manifold-array orders depend
// on identity hash codes (bp/Dbvt.java:804) and are not reproducible; this
// oracle iterates pairs and manifolds in ascending (uid_lo, uid_hi) order.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>

#include "oracle.h"

namespace jb {

static const float ANGULAR_MOTION_THRESHOLD = 0.5f * SIMD_HALF_PI;  // TransformUtil.java:39
static const float MAX_ANGVEL = SIMD_HALF_PI;                        // RigidBody.java:56

void updateInertiaTensor(Body& b) {  // RigidBody.java:449-459
  M3 m1 = mscale(b.tr.b, b.invInertiaLocal);
  M3 m2 = transpose(b.tr.b);
  b.invInertiaWorld = mulMM(m1, m2);
}

// TransformUtil.integrateTransform (TransformUtil.java:64-103)
Tr integrateTransform(const Tr& cur, V3 lv, V3 av, float dt) {
  Tr p;
  p.o.x = cur.o.x + dt * lv.x;
  p.o.y = cur.o.y + dt * lv.y;
  p.o.z = cur.o.z + dt * lv.z;
  float fAngle = len(av);
  if (fAngle * dt > ANGULAR_MOTION_THRESHOLD) fAngle = ANGULAR_MOTION_THRESHOLD / dt;
  V3 axis;
  if (fAngle < 0.001f) {
    axis = scl(av, 0.5f * dt - (dt * dt * dt) * (0.020833333333f) * fAngle * fAngle);
  } else {
    axis = scl(av, (float)std::sin((double)(0.5f * fAngle * dt)) / fAngle);
  }
  Q dorn = q4(axis.x, axis.y, axis.z, (float)std::cos((double)(fAngle * dt * 0.5f)));
  Q orn0 = qFromMatrix(cur.b);
  Q pr = qnor(qmul(dorn, orn0));
  p.b = m3FromQ(pr);
  return p;
}

void World::addBody(const Body& b0) {
  Body b = b0;
  b.islandTag = -1;
  b.companionId = -1;
  updateInertiaTensor(b);
  bodies.push_back(b);
  // CollisionWorld.addCollisionObject (:102-121): AABB without the contact threshold,
  // DiscreteDynamicsWorld.addRigidBody (:426-440): filter groups by dynamic/static
  V3 mn, mx;
  shapeAabb(shapes[b.shape], b.tr, mn, mx);
  bool dyn = !b.isStatic();
  int16_t group = dyn ? 1 : 2;
  int16_t mask = dyn ? (int16_t)-1 : (int16_t)(-1 ^ 2);
  bp.createProxy(Aabb{mn, mx}, group, mask);
}

bool World::needsCollision(int a, int b) const {  // CollisionDispatcher.java:198-223
  return bodies[a].isActive() || bodies[b].isActive();
}
bool World::needsResponse(int a, int b) const {  // CollisionDispatcher.java:225-231
  return (!bodies[a].isStatic()) || (!bodies[b].isStatic());
}

void World::predictUnconstraintMotion(float dt) {  // DiscreteDynamicsWorld.java:741-767
  for (auto& b : bodies) {
    if (b.isStatic() || !b.isActive()) continue;
    // RigidBody.integrateVelocities (:323-342); totalForce = gravity (applyGravity), torque 0
    V3 tf = v3(0.0f + b.gravity.x * 1.0f, 0.0f + b.gravity.y * 1.0f, 0.0f + b.gravity.z * 1.0f);
    b.lv.x += b.invMass * dt * tf.x;
    b.lv.y += b.invMass * dt * tf.y;
    b.lv.z += b.invMass * dt * tf.z;
    V3 tmp = mulMV(b.invInertiaWorld, v3(0, 0, 0));
    tmp = scl(tmp, dt);
    b.av = add(b.av, tmp);
    float angvel = len(b.av);
    if (angvel * dt > MAX_ANGVEL) b.av = scl(b.av, (MAX_ANGVEL / dt) / angvel);
    // applyDamping (:246-299) with zero damping: scale by (float)Math.pow(1, dt)
    float ld = (float)std::pow(1.0 - 0.0, (double)dt);
    b.lv = scl(b.lv, ld);
    b.av = scl(b.av, ld);
  }
}

void World::updateAabbs() {  // CollisionWorld.java:195-245
  for (size_t i = 0; i < bodies.size(); i++) {
    Body& b = bodies[i];
    if (!b.isActive()) continue;
    V3 mn, mx;
    shapeAabb(shapes[b.shape], b.tr, mn, mx);
    V3 ct = v3(CONTACT_BREAKING_THRESHOLD, CONTACT_BREAKING_THRESHOLD, CONTACT_BREAKING_THRESHOLD);
    mn = sub(mn, ct);
    mx = add(mx, ct);
    if (b.isStatic() || len2(sub(mx, mn)) < 1e12f) {
      bp.setAabb((int)i, Aabb{mn, mx});
    } else {
      if (b.activation != JBGPU_DISABLE_DEACTIVATION && b.activation != JBGPU_DISABLE_SIMULATION)
        b.activation = JBGPU_DISABLE_SIMULATION;
    }
  }
}

void World::calculateOverlappingPairs() {
  bp.collide();
  std::vector<uint64_t> keys;
  bp.pairs(keys);
  std::vector<PairState> np(keys.size());
  size_t o = 0;
  for (size_t i = 0; i < keys.size(); i++) {
    np[i].key = keys[i];
    while (o < pairs.size() && pairs[o].key < keys[i]) o++;
    if (o < pairs.size() && pairs[o].key == keys[i]) {
      np[i].m = pairs[o].m;  // surviving pair keeps its algorithm and manifold
    } else {
      int a = (int)(keys[i] >> 32) - 1, b = (int)(keys[i] & 0xffffffffu) - 1;
      Manifold& m = np[i].m;
      m.n = 0;
      int ta = shapes[bodies[a].shape].type, tb = shapes[bodies[b].shape].type;
      if (tb == JBGPU_SHAPE_PLANE && ta != JBGPU_SHAPE_PLANE) { m.body0 = a; m.body1 = b; }
      else if (ta == JBGPU_SHAPE_PLANE && tb != JBGPU_SHAPE_PLANE) { m.body0 = b; m.body1 = a; }  // convex is body0
      else { m.body0 = a; m.body1 = b; }
    }
  }
  pairs.swap(np);
}

void World::dispatchAllCollisionPairs() {  // CollisionDispatcher.java:252-257, DefaultNearCallback.java:39-66
  const V3* hv = hullv.empty() ? nullptr : hullv.data();
  for (auto& ps : pairs) {
    int a = (int)(ps.key >> 32) - 1, b = (int)(ps.key & 0xffffffffu) - 1;
    if (!needsCollision(a, b)) continue;
    const Shape& sa = shapes[bodies[a].shape];
    const Shape& sb = shapes[bodies[b].shape];
    Manifold& m = ps.m;
    if (sa.type == JBGPU_SHAPE_SPHERE && sb.type == JBGPU_SHAPE_SPHERE) {
      bool touching;
      ContactCandidate c = sphereSphere(sa, bodies[a].tr, sb, bodies[b].tr, touching);
      if (touching) addContactPoint(m, bodies[a], bodies[b], a, c);
      refreshContactPoints(m, bodies.data());
    } else if (sa.type == JBGPU_SHAPE_PLANE && sb.type == JBGPU_SHAPE_PLANE) {
      continue;  // EmptyAlgorithm
    } else if (sa.type == JBGPU_SHAPE_PLANE || sb.type == JBGPU_SHAPE_PLANE) {
      int cv = (sb.type == JBGPU_SHAPE_PLANE) ? a : b;
      int pl = (cv == a) ? b : a;
      ContactCandidate c = convexPlane(shapes[bodies[cv].shape], hv, bodies[cv].tr, shapes[bodies[pl].shape], bodies[pl].tr);
      if (c.valid) addContactPoint(m, bodies[a], bodies[b], a, c);
      if (m.n != 0) refreshContactPoints(m, bodies.data());
    } else {
      float md = sa.margin + sb.margin + CONTACT_BREAKING_THRESHOLD;  // ConvexConvexAlgorithm.java:122-123
      md *= md;
      ContactCandidate c = gjkPairDetector(sa, bodies[a].tr, sb, bodies[b].tr, hv, md, &gjk);
      if (c.valid) addContactPoint(m, bodies[a], bodies[b], a, c);
      refreshContactPoints(m, bodies.data());
    }
  }
}

// ---- islands -----------------------------------------------------------------
void UnionFind::reset(int n) {  // UnionFind.java:73-79
  id.resize(n);
  sz.resize(n);
  for (int i = 0; i < n; i++) { id[i] = i; sz[i] = 1; }
}
int UnionFind::find(int x) {  // :127-143, path halving
  while (x != id[x]) {
    id[x] = id[id[x]];
    x = id[x];
  }
  return x;
}
void UnionFind::unite(int p, int q) {  // :106-125
  int i = find(p), j = find(q);
  if (i == j) return;
  id[i] = j;
  sz[j] += sz[i];
}
// MiscUtil.quickSortInternal (MiscUtil.java:184-210) with a comparator that never returns 0
template <class T, class Less>
static void jquick(std::vector<T>& a, Less lt, int lo, int hi) {
  int i = lo, j = hi;
  T x = a[(lo + hi) / 2];
  do {
    while (lt(a[i], x)) i++;
    while (lt(x, a[j])) j--;
    if (i <= j) {
      std::swap(a[i], a[j]);
      i++;
      j--;
    }
  } while (i <= j);
  if (lo < j) jquick(a, lt, lo, j);
  if (i < hi) jquick(a, lt, i, hi);
}
template <class T, class Less>
static void jquickSort(std::vector<T>& a, Less lt) {
  if (a.size() > 1) jquick(a, lt, 0, (int)a.size() - 1);
}
void UnionFind::sortIslands() {  // :55-71
  int n = (int)id.size();
  std::vector<std::pair<int, int>> el(n);
  for (int i = 0; i < n; i++) el[i] = {find(i), i};
  jquickSort(el, [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.first < b.first; });
  for (int i = 0; i < n; i++) { id[i] = el[i].first; sz[i] = el[i].second; }
}

void World::calculateSimulationIslands() {  // DiscreteDynamicsWorld.java:649-677
  int n = (int)bodies.size();
  uf.reset(n);
  for (int i = 0; i < n; i++) {  // SimulationIslandManager.updateActivationState (:72-90)
    bodies[i].islandTag = i;
    bodies[i].companionId = -1;
  }
  for (auto& ps : pairs) {  // findUnions (:57-70); mergesSimulationIslands = not static
    int a = (int)(ps.key >> 32) - 1, b = (int)(ps.key & 0xffffffffu) - 1;
    if (!bodies[a].isStatic() && !bodies[b].isStatic()) uf.unite(bodies[a].islandTag, bodies[b].islandTag);
  }
  for (const Joint& j : joints) {  // constraints join islands (DiscreteDynamicsWorld.java:655-669)
    if (j.bodyB < 0) continue;       // TypedConstraint's fixed body is static
    const Body& a = bodies[j.bodyA];
    const Body& b = bodies[j.bodyB];
    if (!a.isStatic() && !b.isStatic() && (a.isActive() || b.isActive())) uf.unite(a.islandTag, b.islandTag);
  }
  for (int i = 0; i < n; i++) {  // storeIslandActivationState (:92-110)
    if (!bodies[i].isStatic()) {
      bodies[i].islandTag = uf.find(i);
      bodies[i].companionId = -1;
    } else {
      bodies[i].islandTag = -1;
      bodies[i].companionId = -2;
    }
  }
}

static inline void setActivation(Body& b, int s) {  // CollisionObject.java:149-153
  if (b.activation != JBGPU_DISABLE_DEACTIVATION && b.activation != JBGPU_DISABLE_SIMULATION) b.activation = s;
}

void World::solveConstraints() {  // DiscreteDynamicsWorld.java:620-647 -> buildAndProcessIslands
  SolverInfo info = cfg.solver;
  lastSolve = SolverStats();
  // buildIslands (SimulationIslandManager.java:120-240)
  uf.sortIslands();
  int numElem = (int)uf.id.size();
  for (int s = 0, e = 1; s < numElem; s = e) {
    int islandId = uf.id[s];
    for (e = s + 1; e < numElem && uf.id[e] == islandId; e++) {
    }
    bool allSleeping = true;
    for (int k = s; k < e; k++) {
      Body& b = bodies[uf.sz[k]];
      if (b.islandTag == islandId) {
        if (b.activation == JBGPU_ACTIVE_TAG) allSleeping = false;
        if (b.activation == JBGPU_DISABLE_DEACTIVATION) allSleeping = false;
      }
    }
    for (int k = s; k < e; k++) {
      Body& b = bodies[uf.sz[k]];
      if (b.islandTag != islandId) continue;
      if (allSleeping)
        setActivation(b, JBGPU_ISLAND_SLEEPING);
      else if (b.activation == JBGPU_ISLAND_SLEEPING)
        setActivation(b, JBGPU_WANTS_DEACTIVATION);
    }
  }
  if (cfg.solverOrder == JBGPU_ORDER_EXPLICIT) {  // island sleep states above, then the caller's manifold order
    std::vector<Manifold*> ms;
    for (int idx : explicitOrder) ms.push_back(&pairs[idx].m);
    std::vector<Joint*> js;
    for (int idx : explicitJointOrder) js.push_back(&joints[idx]);
    if (!ms.empty() || !js.empty()) solveGroup(bodies, ms, info, false, rng, &lastSolve, learnedInit, features, &js);
    return;
  }
  // sortedConstraints: the constraints quicksorted by island (DiscreteDynamicsWorld.java:621-629)
  std::vector<Joint*> sj;
  for (Joint& j : joints) sj.push_back(&j);
  auto jointIsland = [&](const Joint* j) {  // getConstraintIslandId (:550-556)
    const int a = bodies[j->bodyA].islandTag;
    return a >= 0 ? a : (j->bodyB >= 0 ? bodies[j->bodyB].islandTag : -1);
  };
  jquickSort(sj, [&](const Joint* l, const Joint* r) { return jointIsland(l) < jointIsland(r); });
  const int numJoints = (int)sj.size();
  std::vector<Manifold*> im;
  for (auto& ps : pairs) {
    Manifold& m = ps.m;
    const Body& b0 = bodies[m.body0];
    const Body& b1 = bodies[m.body1];
    if (b0.activation != JBGPU_ISLAND_SLEEPING || b1.activation != JBGPU_ISLAND_SLEEPING)
      if (needsResponse(m.body0, m.body1)) im.push_back(&m);
  }
  auto islandOf = [&](const Manifold* m) {
    return bodies[m->body0].islandTag >= 0 ? bodies[m->body0].islandTag : bodies[m->body1].islandTag;
  };
  jquickSort(im, [&](const Manifold* l, const Manifold* r) { return islandOf(l) < islandOf(r); });
  // buildAndProcessIslands (:242-333)
  int numManifolds = (int)im.size();
  int sm = 0, em = 1;
  for (int s = 0, e = 0; s < numElem; s = e) {
    int islandId = uf.id[s];
    bool islandSleeping = false;
    for (e = s; e < numElem && uf.id[e] == islandId; e++)
      if (!bodies[uf.sz[e]].isActive()) islandSleeping = true;
    int numIslandManifolds = 0;
    int startIdx = -1;
    if (sm < numManifolds) {
      if (islandOf(im[sm]) == islandId) {
        startIdx = sm;
        for (em = sm + 1; em < numManifolds && islandId == islandOf(im[em]); em++) {
        }
        numIslandManifolds = em - sm;
      }
    }
    // the island's constraints (InplaceSolverIslandCallback.processIsland, DiscreteDynamicsWorld.java:586-612)
    int startC = -1, numC = 0, ci = 0;
    for (; ci < numJoints; ci++)
      if (jointIsland(sj[ci]) == islandId) {
        startC = ci;
        break;
      }
    for (; ci < numJoints; ci++)
      if (jointIsland(sj[ci]) == islandId) numC++;
    if (!islandSleeping && numIslandManifolds + numC > 0) {
      std::vector<Manifold*> ms;
      if (numIslandManifolds > 0) ms.assign(im.begin() + startIdx, im.begin() + startIdx + numIslandManifolds);
      std::vector<Joint*> js;
      if (numC > 0) js.assign(sj.begin() + startC, sj.begin() + startC + numC);
      solveGroup(bodies, ms, info, cfg.solverOrder == JBGPU_ORDER_REFERENCE, rng, &lastSolve, learnedInit, features,
                 &js);
    }
    if (numIslandManifolds != 0) sm = em;
  }
}

void World::integrateTransforms(float dt) {  // DiscreteDynamicsWorld.java:679-739
  for (size_t i = 0; i < bodies.size(); i++) {
    Body& b = bodies[i];
    if (b.isActive() && !b.isStatic()) {
      Tr pred = integrateTransform(b.tr, b.lv, b.av, dt);
      // CCD motion clamp (:693-725): a convex body moving more than its threshold sweeps a sphere against the
      // world (bodies earlier in the list already at their new transforms) and stops at the first hit
      const float sq = b.ccdMotionThreshold * b.ccdMotionThreshold;  // getCcdSquareMotionThreshold
      const float squareMotion = len2(sub(pred.o, b.tr.o));
      if (sq != 0.0f && sq < squareMotion && shapes[b.shape].type != JBGPU_SHAPE_PLANE) pred = ccdClamp(*this, (int)i, pred, dt);
      b.tr = pred;
      updateInertiaTensor(b);  // proceedToTransform -> setCenterOfMassTransform (RigidBody.java:344-354)
    }
  }
}

void World::updateActivationState(float dt) {  // DiscreteDynamicsWorld.java:475-512
  for (auto& b : bodies) {
    // RigidBody.updateDeactivation (:543-557)
    if (!(b.activation == JBGPU_ISLAND_SLEEPING || b.activation == JBGPU_DISABLE_DEACTIVATION)) {
      if ((len2(b.lv) < 0.8f * 0.8f) && (len2(b.av) < 1.0f * 1.0f)) {
        b.deactivationTime += dt;
      } else {
        b.deactivationTime = 0.0f;
        setActivation(b, 0);
      }
    }
    // RigidBody.wantsSleeping (:559-576)
    bool wants;
    if (b.activation == JBGPU_DISABLE_DEACTIVATION) wants = false;
    else if (cfg.deactivationDisabled) wants = false;
    else if (b.activation == JBGPU_ISLAND_SLEEPING || b.activation == JBGPU_WANTS_DEACTIVATION) wants = true;
    else wants = b.deactivationTime > 2.0f;
    if (wants) {
      if (b.isStatic()) {
        setActivation(b, JBGPU_ISLAND_SLEEPING);
      } else {
        if (b.activation == JBGPU_ACTIVE_TAG) setActivation(b, JBGPU_WANTS_DEACTIVATION);
        if (b.activation == JBGPU_ISLAND_SLEEPING) {
          b.av = v3(0, 0, 0);
          b.lv = v3(0, 0, 0);
        }
      }
    } else {
      if (b.activation != JBGPU_DISABLE_DEACTIVATION) setActivation(b, JBGPU_ACTIVE_TAG);
    }
  }
}

void World::step(float dt) {
  // stepSimulation(dt, 0): exactly one substep (DiscreteDynamicsWorld.java:301-318)
  predictUnconstraintMotion(dt);
  updateAabbs();
  calculateOverlappingPairs();
  dispatchAllCollisionPairs();
  calculateSimulationIslands();
  solveConstraints();
  integrateTransforms(dt);
  updateActivationState(dt);
  stepCount++;
}

}  // namespace jb
