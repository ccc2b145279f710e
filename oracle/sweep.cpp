// Oracle (test infrastructure only): convex sweeps and continuous collision detection.
//
// CollisionWorld.convexSweepTest (collision/dispatch/CollisionWorld.java:596-646) with objectQuerySingle's convex
// branch (:392-424), GjkConvexCast.calcTimeOfImpact (collision/narrowphase/GjkConvexCast.java:67-193),
// ClosestConvexResultCallback (:766-810), DiscreteDynamicsWorld.integrateTransforms' motion clamp
// (dynamics/DiscreteDynamicsWorld.java:679-739) with ClosestNotMeConvexResultCallback (:1129-1198).
#include <cmath>

#include "oracle.h"

namespace jb {

namespace {
// MatrixUtil.invert (MatrixUtil.java:362-394): cofactors over the determinant
float cofac(const M3& m, int r1, int c1, int r2, int c2) {
  return m.m[3 * r1 + c1] * m.m[3 * r2 + c2] - m.m[3 * r1 + c2] * m.m[3 * r2 + c1];
}
M3 invert(const M3& a) {
  const float co_x = cofac(a, 1, 1, 2, 2);
  const float co_y = cofac(a, 1, 2, 2, 0);
  const float co_z = cofac(a, 1, 0, 2, 1);
  const float det = a.m[0] * co_x + a.m[1] * co_y + a.m[2] * co_z;
  const float s = 1.0f / det;
  M3 r;
  r.m[0] = co_x * s;
  r.m[1] = cofac(a, 0, 2, 2, 1) * s;
  r.m[2] = cofac(a, 0, 1, 1, 2) * s;
  r.m[3] = co_y * s;
  r.m[4] = cofac(a, 0, 0, 2, 2) * s;
  r.m[5] = cofac(a, 0, 2, 1, 0) * s;
  r.m[6] = co_z * s;
  r.m[7] = cofac(a, 0, 1, 2, 0) * s;
  r.m[8] = cofac(a, 0, 0, 1, 1) * s;
  return r;
}

// PointCollector (PointCollector.java): keeps the smallest distance reported to it
struct PointCollector {
  V3 normal = v3(0, 0, 0), point = v3(0, 0, 0);
  float distance = 1e30f;
  bool hasResult = false;
  void add(const ContactCandidate& c) {
    if (!c.valid) return;
    if (c.depth < distance) {
      hasResult = true;
      normal = c.normalOnB;
      point = c.pointOnB;
      distance = c.depth;
    }
  }
};
}  // namespace

void calculateVelocity(const Tr& t0, const Tr& t1, float timeStep, V3& linVel, V3& angVel) {
  linVel = scl(sub(t1.o, t0.o), 1.0f / timeStep);
  // calculateDiffAxisAngle: dmat = basis1 * basis0^-1 as a quaternion (Quaternion.setFromMatrix), normalised
  const M3 dmat = mulMM(t1.b, invert(t0.b));
  Q dorn = qnor(qFromMatrix(dmat));
  const float angle = 2.0f * (float)std::acos((double)dorn.w);  // QuaternionUtil.getAngle
  V3 axis = v3(dorn.x, dorn.y, dorn.z);
  const float l = len2(axis);
  if (l < FLT_EPS * FLT_EPS)
    axis = v3(1.0f, 0.0f, 0.0f);
  else
    axis = scl(axis, 1.0f / (float)std::sqrt((double)l));
  angVel = scl(axis, angle / timeStep);
}

void temporalAabb(const Shape& s, const Tr& cur, V3 linvel, V3 angvel, float timeStep, V3& mn, V3& mx) {
  shapeAabb(s, cur, mn, mx);
  const V3 lm = scl(linvel, timeStep);
  if (lm.x > 0.0f) mx.x += lm.x; else mn.x += lm.x;
  if (lm.y > 0.0f) mx.y += lm.y; else mn.y += lm.y;
  if (lm.z > 0.0f) mx.z += lm.z; else mn.z += lm.z;
  // getAngularMotionDisc: the bounding sphere of the AABB at the identity, plus its centre's distance
  Tr id;
  id.b = m3ident();
  id.o = v3(0, 0, 0);
  V3 a0, a1;
  shapeAabb(s, id, a0, a1);
  const float radius = len(sub(a1, a0)) * 0.5f;
  const V3 center = scl(add(a0, a1), 0.5f);
  const float disc = radius + len(center);
  const float am = len(angvel) * disc * timeStep;
  mn = sub(mn, v3(am, am, am));
  mx = add(mx, v3(am, am, am));
}

bool gjkConvexCast(const Shape& sa, const Tr& fromA, const Tr& toA, const Shape& sb, const Tr& tb, const V3* hv,
                   float allowedPenetration, float& fraction, V3& normal, V3& hitPoint) {
  const V3 linVelA = sub(toA.o, fromA.o);
  const V3 linVelB = sub(tb.o, tb.o);
  const float radius = 0.001f;
  float lambda = 0.0f;
  V3 n = v3(0, 0, 0);
  const V3 r = sub(linVelA, linVelB);
  float lastLambda = lambda;
  int numIter = 0;
  const int maxIter = 32;
  PointCollector pc;
  Tr ia = fromA, ib = tb;
  // GjkPairDetector without a penetration solver, maximumDistanceSquared = Float.MAX_VALUE
  pc.add(gjkPairDetector(sa, ia, sb, ib, hv, 3.4028235e38f, nullptr, false));
  bool hasResult = pc.hasResult;
  V3 c = pc.point;
  if (!hasResult) return false;
  float dist = pc.distance;
  n = pc.normal;
  while (dist > radius) {
    numIter++;
    if (numIter > maxIter) return false;
    const float projectedLinearVelocity = dot(r, n);
    const float dLambda = dist / projectedLinearVelocity;
    lambda = lambda - dLambda;
    if (lambda > 1.0f) return false;
    if (lambda < 0.0f) return false;
    if (lambda <= lastLambda) return false;
    lastLambda = lambda;
    ia.o = interp3(fromA.o, toA.o, lambda);
    ib.o = interp3(tb.o, tb.o, lambda);
    pc.add(gjkPairDetector(sa, ia, sb, ib, hv, 3.4028235e38f, nullptr, false));
    if (pc.hasResult) {
      if (pc.distance < 0.0f) {
        fraction = lastLambda;
        n = pc.normal;
        normal = n;
        hitPoint = pc.point;
        return true;
      }
      c = pc.point;
      n = pc.normal;
      dist = pc.distance;
    } else {
      return false;
    }
  }
  // no time of impact for motion away from the contact normal
  if (dot(n, r) >= -allowedPenetration) return false;
  fraction = lambda;
  normal = n;
  hitPoint = c;
  return true;
}

namespace {
// the cast shape's AABB over the sweep: rotation-only transform of `from`, temporal AABB over the motion
void castAabb(const Shape& castShape, const Tr& from, const Tr& to, V3& mn, V3& mx) {
  V3 lv, av;
  calculateVelocity(from, to, 1.0f, lv, av);
  Tr R;
  R.b = m3FromQ(qFromMatrix(from.b));  // R.setRotation(from.getRotation())
  R.o = v3(0, 0, 0);
  temporalAabb(castShape, R, lv, av, 1.0f, mn, mx);
}

// one object of convexSweepTest: AABB expanded by the cast AABB, hit by the segment of the origins, then the
// cast; true with the cast's (normalised) result when objectQuerySingle would call addSingleResult
bool sweepObject(const World& w, const Shape& castShape, const Tr& from, const Tr& to, V3 cmn, V3 cmx, int i,
                 float closest, float& f, V3& nrm, V3& pt) {
  const Body& b = w.bodies[i];
  const Shape& sh = w.shapes[b.shape];
  V3 mn, mx;
  shapeAabb(sh, b.tr, mn, mx);
  mn = add(mn, cmn);  // AabbUtil2.aabbExpand
  mx = add(mx, cmx);
  float lambda = 1.0f;
  V3 hn;
  if (!rayAabb(from.o, to.o, mn, mx, lambda, hn)) return false;
  if (sh.type == JBGPU_SHAPE_PLANE) return false;  // no caster for StaticPlaneShape (the reference throws)
  const V3* hv = w.hullv.empty() ? nullptr : w.hullv.data();
  V3 n, p;
  float fr = 1.0f;
  if (!gjkConvexCast(castShape, from, to, sh, b.tr, hv, 0.04f, fr, n, p)) return false;  // allowedCcdPenetration
  if (!(len2(n) > 0.0001f)) return false;
  if (!(fr < closest)) return false;
  f = fr;
  nrm = nor(n);
  pt = p;
  return true;
}
}  // namespace

RayHit convexSweepClosest(const World& w, const Shape& castShape, const Tr& from, const Tr& to, int16_t group,
                          int16_t mask, int world) {
  RayHit out;
  out.object = -1;
  out.fraction = 1.0f;
  out.normal = v3(0, 0, 0);
  out.point = v3(0, 0, 0);
  V3 cmn, cmx;
  castAabb(castShape, from, to, cmn, cmx);
  for (size_t i = 0; i < w.bodies.size(); i++) {
    const Proxy& px = w.bp.proxies[i];
    if (px.dead || px.world != world) continue;
    bool collides = (((int)px.group & (int)mask) & 0xFFFF) != 0;  // ConvexResultCallback.needsCollision
    collides = collides && (((int)group & (int)px.mask) & 0xFFFF) != 0;
    if (!collides) continue;
    float f;
    V3 n, p;
    if (!sweepObject(w, castShape, from, to, cmn, cmx, (int)i, out.fraction, f, n, p)) continue;
    out.fraction = f;  // ClosestConvexResultCallback.addSingleResult
    out.object = (int)i;
    out.normal = n;
    out.point = p;
  }
  return out;
}

Tr ccdClamp(World& w, int me, const Tr& pred, float dt) {
  Body& b = w.bodies[me];
  const Shape cast = makeSphere(b.ccdSweptSphereRadius);  // new SphereShape(ccdSweptSphereRadius)
  const Tr from = b.tr;
  const Proxy& mp = w.bp.proxies[me];
  V3 cmn, cmx;
  castAabb(cast, from, pred, cmn, cmx);
  float closest = 1.0f;
  bool hit = false;
  const V3 relVel = sub(pred.o, from.o);  // ClosestNotMe: linVelA - linVelB with B at rest
  for (size_t i = 0; i < w.bodies.size(); i++) {
    if ((int)i == me) continue;  // needsCollision: not itself
    const Proxy& px = w.bp.proxies[i];
    if (px.dead || px.world != mp.world) continue;
    bool collides = (((int)px.group & (int)mp.mask) & 0xFFFF) != 0;
    collides = collides && (((int)mp.group & (int)px.mask) & 0xFFFF) != 0;
    if (!collides) continue;
    if (w.needsResponse(me, (int)i)) {  // no CCD against a body it already touches
      const uint64_t lo = (uint64_t)std::min(me, (int)i) + 1, hi = (uint64_t)std::max(me, (int)i) + 1;
      const uint64_t key = (lo << 32) | hi;
      size_t a = 0, e = w.pairs.size();
      while (a < e) {
        const size_t m = (a + e) / 2;
        if (w.pairs[m].key < key) a = m + 1; else e = m;
      }
      if (a < w.pairs.size() && w.pairs[a].key == key && w.pairs[a].m.n > 0) continue;
    }
    float f;
    V3 n, p;
    if (!sweepObject(w, cast, from, pred, cmn, cmx, (int)i, closest, f, n, p)) continue;
    // ClosestNotMeConvexResultCallback.addSingleResult: motion away from the normal is not a hit
    if (dot(n, relVel) >= -0.0f) continue;
    closest = f;
    hit = true;
  }
  if (hit && closest > 0.0001f) return integrateTransform(b.tr, b.lv, b.av, dt * closest);
  return pred;
}

}  // namespace jb
