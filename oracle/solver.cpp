// Oracle (test infrastructure only): SequentialImpulseConstraintSolver, the
// cache-friendly path (dynamics/constraintsolver/SequentialImpulseConstraintSolver.java:
// setup :418-786, iterations :788-890, writeback :892-950) with SolverBody
// (dynamics/constraintsolver/SolverBody.java) and TransformUtil.planeSpace1.
//
// Project additions, off unless requested: a residual r_k = ||l_k - l_{k-1}||_2 /
// max(1, ||l_k||_2) over all rows (double accumulation), a tolerance-driven
// iteration count, per-row warm-start overrides (learned warm start) and a
// per-row feature dump.
#include <cmath>
#include <cstdio>
#include <vector>

#include "oracle.h"

namespace jb {

static void planeSpace1(V3 n, V3& p, V3& q) {  // TransformUtil.java:45-61
  const float SIMDSQRT12 = 0.7071067811865475244008443621048490f;
  if (jabs(n.z) > SIMDSQRT12) {
    float a = n.y * n.y + n.z * n.z;
    float k = 1.0f / jsqrt(a);
    p = v3(0, -n.z * k, n.y * k);
    q = v3(a * k, -n.x * p.z, n.x * p.y);
  } else {
    float a = n.x * n.x + n.y * n.y;
    float k = 1.0f / jsqrt(a);
    p = v3(-n.y * k, n.x * k, 0);
    q = v3(-n.z * p.y, n.z * p.x, a * k);
  }
}

static inline void applyImpulse(SolverBody& b, V3 lc, V3 ac, float mag) {  // SolverBody.java:61-72
  if (b.invMass != 0.0f) {
    b.lv.x += mag * 1.0f * lc.x;
    b.lv.y += mag * 1.0f * lc.y;
    b.lv.z += mag * 1.0f * lc.z;
    b.av.x += mag * 1.0f * ac.x;
    b.av.y += mag * 1.0f * ac.y;
    b.av.z += mag * 1.0f * ac.z;
  }
}

static SolverRow makeFrictionRow(V3 axis, int ida, int idb, int fi, const MPoint& cp, V3 rp1, V3 rp2, const Body& b0,
                                 const Body& b1, float relaxation) {  // :349-416
  SolverRow r{};
  r.normal = axis;
  r.bodyA = ida;
  r.bodyB = idb;
  r.frictionIndex = fi;
  r.friction = cp.friction;
  r.applied = 0;
  r.pushApplied = 0;
  r.penetration = 0;
  V3 t = crs(rp1, r.normal);
  r.rel1xn = t;
  r.angA = mulMV(b0.invInertiaWorld, t);
  t = crs(rp2, r.normal);
  r.rel2xn = t;
  r.angB = mulMV(b1.invInertiaWorld, t);
  float d0 = b0.invMass + dot(axis, crs(r.angA, rp1));
  float d1 = b1.invMass + dot(axis, crs(r.angB, rp2));
  r.jacInv = relaxation / (d0 + d1);
  return r;
}

// ---- Point2PointConstraint ------------------------------------------------------------------------------
namespace {
Body g_fixed = [] {  // TypedConstraint.getFixed(): new RigidBody(0, null, null), setMassProps(0, (0, 0, 0))
  Body b{};
  b.tr.b = m3ident();
  b.tr.o = v3(0, 0, 0);
  b.invInertiaWorld = mscale(m3ident(), v3(0, 0, 0));
  b.flags = JBGPU_BODY_STATIC;
  return b;
}();
}  // namespace

static void hingeBuildJacobian(Joint& j, const std::vector<Body>& bodies);
static void hingeSolve(Joint& j, std::vector<Body>& bodies, float timeStep);

void jointBuildJacobian(Joint& j, const std::vector<Body>& bodies) {  // Point2PointConstraint.java:71-107
  if (j.type == 1) {
    hingeBuildJacobian(j, bodies);
    return;
  }
  j.applied = 0.0f;
  const Body& A = bodies[j.bodyA];
  const Body& B = j.bodyB >= 0 ? bodies[j.bodyB] : g_fixed;
  for (int i = 0; i < 3; i++) {
    V3 normal = v3(0, 0, 0);
    setcoord(normal, i, 1.0f);
    const M3 world2A = transpose(A.tr.b), world2B = transpose(B.tr.b);
    const V3 rel1 = sub(xform(A.tr, j.pivotA), A.tr.o);
    const V3 rel2 = sub(xform(B.tr, j.pivotB), B.tr.o);
    // JacobianEntry.init, two bodies (JacobianEntry.java:56-73)
    const V3 aJ = mulMV(world2A, crs(rel1, normal));
    const V3 bJ = mulMV(world2B, crs(rel2, scl(normal, -1.0f)));
    const V3 m0 = mulc(A.invInertiaLocal, aJ), m1 = mulc(B.invInertiaLocal, bJ);
    j.adiag[i] = A.invMass + dot(m0, aJ) + B.invMass + dot(m1, bJ);
  }
}

// RigidBody.applyImpulse (RigidBody.java:414-428) with linear and angular factors 1
static void applyImpulseRB(Body& b, V3 impulse, V3 relPos) {
  if (b.invMass == 0.0f) return;
  const V3 scaled = v3(impulse.x * 1.0f, impulse.y * 1.0f, impulse.z * 1.0f);  // applyCentralImpulse :395-402
  b.lv = add(b.lv, scl(scaled, b.invMass));
  V3 t = crs(relPos, v3(impulse.x * 1.0f, impulse.y * 1.0f, impulse.z * 1.0f));
  t = v3(t.x * 1.0f, t.y * 1.0f, t.z * 1.0f);
  b.av = add(b.av, mulMV(b.invInertiaWorld, t));  // applyTorqueImpulse :405-411
}

void jointSolve(Joint& j, std::vector<Body>& bodies, float timeStep) {  // Point2PointConstraint.java:109-177
  if (j.type == 1) {
    hingeSolve(j, bodies, timeStep);
    return;
  }
  Body fixed = g_fixed;
  Body& A = bodies[j.bodyA];
  Body& B = j.bodyB >= 0 ? bodies[j.bodyB] : fixed;
  const V3 pivotAInW = xform(A.tr, j.pivotA), pivotBInW = xform(B.tr, j.pivotB);
  for (int i = 0; i < 3; i++) {
    V3 normal = v3(0, 0, 0);
    setcoord(normal, i, 1.0f);
    const float jacDiagABInv = 1.0f / j.adiag[i];
    const V3 rel1 = sub(pivotAInW, A.tr.o), rel2 = sub(pivotBInW, B.tr.o);
    const V3 vel1 = add(crs(A.av, rel1), A.lv), vel2 = add(crs(B.av, rel2), B.lv);  // getVelocityInLocalPoint
    const V3 vel = sub(vel1, vel2);
    const float rel_vel = dot(normal, vel);
    const float depth = -dot(sub(pivotAInW, pivotBInW), normal);
    float impulse = depth * j.tau / timeStep * jacDiagABInv - j.damping * rel_vel * jacDiagABInv;
    if (j.impulseClamp > 0.0f) {
      if (impulse < -j.impulseClamp) impulse = -j.impulseClamp;
      if (impulse > j.impulseClamp) impulse = j.impulseClamp;
    }
    j.applied += impulse;
    const V3 iv = scl(normal, impulse);
    applyImpulseRB(A, iv, sub(pivotAInW, A.tr.o));
    applyImpulseRB(B, scl(iv, -1.0f), sub(pivotBInW, B.tr.o));
  }
}

// ---- HingeConstraint -----------------------------------------------------------------------------------
static V3 column(const M3& m, int c) { return v3(m.m[c], m.m[3 + c], m.m[6 + c]); }  // MatrixUtil.getColumn
static void setRow(M3& m, int r, float x, float y, float z) {                        // MatrixUtil.setRow
  m.m[3 * r] = x;
  m.m[3 * r + 1] = y;
  m.m[3 * r + 2] = z;
}
static Q shortestArcQuat(V3 v0, V3 v1) {  // QuaternionUtil.java:49-68
  const V3 c = crs(v0, v1);
  const float d = dot(v0, v1);
  if ((double)d < -1.0 + (double)1.19209290e-07f) return q4(0.0f, 1.0f, 0.0f, 0.0f);  // a double comparison
  const float s = (float)std::sqrt((double)((1.0f + d) * 2.0f));
  const float rs = 1.0f / s;
  return q4(c.x * rs, c.y * rs, c.z * rs, s * 0.5f);
}
static V3 quatRotate(Q r, V3 v) {  // QuaternionUtil.java:70-90: (r * v) * conj(r)
  Q q = q4(r.w * v.x + r.y * v.z - r.z * v.y, r.w * v.y + r.z * v.x - r.x * v.z, r.w * v.z + r.x * v.y - r.y * v.x,
           -r.x * v.x - r.y * v.y - r.z * v.z);
  q = qmul(q, q4(-r.x, -r.y, -r.z, r.w));
  return v3(q.x, q.y, q.z);
}

void hingeFramesFromAxes(const Tr& trA, V3 pivotA, V3 pivotB, V3 axisA, V3 axisB, bool fixedB, Tr& frameA,
                         Tr& frameB) {
  const float eps = 1.19209290e-07f;
  V3 a1 = column(trA.b, 0), a2;
  if (!fixedB) {  // HingeConstraint.java:84-130
    const float projection = dot(axisA, a1);
    if (projection >= 1.0f - eps) {
      a1 = scl(column(trA.b, 2), -1.0f);
      a2 = column(trA.b, 1);
    } else if (projection <= -1.0f + eps) {
      a1 = column(trA.b, 2);
      a2 = column(trA.b, 1);
    } else {
      a2 = crs(axisA, a1);
      a1 = crs(a2, axisA);
    }
  } else {  // :132-178, B is the fixed body: the axis in B is -axisA in world space
    const float projection = dot(a1, axisA);
    if (projection > eps) {
      a1 = scl(a1, projection);
      a1 = sub(a1, axisA);
    } else {
      a1 = column(trA.b, 1);
    }
    a2 = crs(axisA, a1);
    axisB = mulMV(trA.b, scl(axisA, -1.0f));  // Vector3.mul(Matrix3)
  }
  frameA.o = pivotA;
  setRow(frameA.b, 0, a1.x, a2.x, axisA.x);
  setRow(frameA.b, 1, a1.y, a2.y, axisA.y);
  setRow(frameA.b, 2, a1.z, a2.z, axisA.z);
  const Q arc = shortestArcQuat(axisA, axisB);
  const V3 b1 = quatRotate(arc, a1);
  const V3 b2 = crs(axisB, b1);
  if (!fixedB) {
    frameB.o = pivotB;
    setRow(frameB.b, 0, b1.x, b2.x, -axisB.x);
    setRow(frameB.b, 1, b1.y, b2.y, -axisB.y);
    setRow(frameB.b, 2, b1.z, b2.z, -axisB.z);
  } else {
    frameB.o = xform(trA, pivotA);
    setRow(frameB.b, 0, b1.x, b2.x, axisB.x);
    setRow(frameB.b, 1, b1.y, b2.y, axisB.y);
    setRow(frameB.b, 2, b1.z, b2.z, axisB.z);
  }
}

void hingeFramesFromFrames(const Tr& trA, const Tr& inA, const Tr& inB, bool fixedB, Tr& frameA, Tr& frameB) {
  frameA = inA;
  frameB = fixedB ? inA : inB;  // (rbA, rbAFrame): B's frame starts as A's (:212-238)
  frameB.b.m[2] *= -1.0f;       // flip axis: M02, M12, M22
  frameB.b.m[5] *= -1.0f;
  frameB.b.m[8] *= -1.0f;
  if (fixedB) frameB.o = xform(trA, inA.o);
}

float atan2Fast(float y, float x) {  // ScalarUtil.java:43-57
  const float simdPi = 6.283185307179586232f * 0.5f;
  const float coeff_1 = simdPi / 4.0f;
  const float coeff_2 = 3.0f * coeff_1;
  const float abs_y = jabs(y);
  float angle;
  if (x >= 0.0f) {
    const float r = (x - abs_y) / (x + abs_y);
    angle = coeff_1 - coeff_1 * r;
  } else {
    const float r = (x + abs_y) / (abs_y - x);
    angle = coeff_2 - coeff_1 * r;
  }
  return (y < 0.0f) ? -angle : angle;
}

float hingeAngle(const Joint& j, const std::vector<Body>& bodies) {  // HingeConstraint.java:530-549
  const Body& A = bodies[j.bodyA];
  const Body& B = j.bodyB >= 0 ? bodies[j.bodyB] : g_fixed;
  const V3 ref0 = mulMV(A.tr.b, column(j.frameA.b, 0));
  const V3 ref1 = mulMV(A.tr.b, column(j.frameA.b, 1));
  const V3 swing = mulMV(B.tr.b, column(j.frameB.b, 1));
  return atan2Fast(dot(swing, ref0), dot(swing, ref1));
}

// RigidBody.computeAngularImpulseDenominator (RigidBody.java:534-541): axis . (I_world^T axis)
static float angularDenominator(const Body& b, V3 axis) { return dot(axis, mulMtV(b.invInertiaWorld, axis)); }
// RigidBody.applyTorqueImpulse (RigidBody.java:405-411): no inverse-mass test
static void applyTorqueImpulse(Body& b, V3 torque) { b.av = add(b.av, mulMV(b.invInertiaWorld, torque)); }

static void hingeBuildJacobian(Joint& j, const std::vector<Body>& bodies) {  // HingeConstraint.java:240-333
  const Body& A = bodies[j.bodyA];
  const Body& B = j.bodyB >= 0 ? bodies[j.bodyB] : g_fixed;
  j.applied = 0.0f;
  if (!j.angularOnly) {
    const V3 pivotAInW = xform(A.tr, j.frameA.o), pivotBInW = xform(B.tr, j.frameB.o);
    const V3 relPos = sub(pivotBInW, pivotAInW);
    if (len2(relPos) > 1.19209290e-07f) j.linAxis[0] = nor(relPos);
    else j.linAxis[0] = v3(1.0f, 0.0f, 0.0f);
    planeSpace1(j.linAxis[0], j.linAxis[1], j.linAxis[2]);
    for (int i = 0; i < 3; i++) {  // JacobianEntry.init (JacobianEntry.java:56-73)
      const M3 world2A = transpose(A.tr.b), world2B = transpose(B.tr.b);
      const V3 rel1 = sub(pivotAInW, A.tr.o), rel2 = sub(pivotBInW, B.tr.o);
      const V3 n = j.linAxis[i];
      const V3 aJ = mulMV(world2A, crs(rel1, n));
      const V3 bJ = mulMV(world2B, crs(rel2, scl(n, -1.0f)));
      const V3 m0 = mulc(A.invInertiaLocal, aJ), m1 = mulc(B.invInertiaLocal, bJ);
      j.adiag[i] = A.invMass + dot(m0, aJ) + B.invMass + dot(m1, bJ);
    }
  }
  // the angular JacobianEntries (jacAng) are built by the reference but never read by solveConstraint
  const float angle = hingeAngle(j, bodies);
  j.correction = 0.0f;
  j.limitSign = 0.0f;
  j.solveLimit = false;
  j.accLimitImpulse = 0.0f;
  if (j.lowerLimit < j.upperLimit) {
    if (angle <= j.lowerLimit * j.limitSoftness) {
      j.correction = j.lowerLimit - angle;
      j.limitSign = 1.0f;
      j.solveLimit = true;
    } else if (angle >= j.upperLimit * j.limitSoftness) {
      j.correction = j.upperLimit - angle;
      j.limitSign = -1.0f;
      j.solveLimit = true;
    }
  }
  const V3 axisA = mulMV(A.tr.b, column(j.frameA.b, 2));
  j.kHinge = 1.0f / (angularDenominator(A, axisA) + angularDenominator(B, axisA));
}

static void hingeSolve(Joint& j, std::vector<Body>& bodies, float timeStep) {  // HingeConstraint.java:335-520
  Body fixed = g_fixed;
  Body& A = bodies[j.bodyA];
  Body& B = j.bodyB >= 0 ? bodies[j.bodyB] : fixed;
  const V3 pivotAInW = xform(A.tr, j.frameA.o), pivotBInW = xform(B.tr, j.frameB.o);
  const float tau = 0.3f;
  if (!j.angularOnly) {  // linear part: the relative velocity is taken once, before the three rows
    const V3 rel1 = sub(pivotAInW, A.tr.o), rel2 = sub(pivotBInW, B.tr.o);
    const V3 vel = sub(add(crs(A.av, rel1), A.lv), add(crs(B.av, rel2), B.lv));
    for (int i = 0; i < 3; i++) {
      const V3 n = j.linAxis[i];
      const float jacDiagABInv = 1.0f / j.adiag[i];
      const float rel_vel = dot(n, vel);
      const float depth = -dot(sub(pivotAInW, pivotBInW), n);
      const float impulse = depth * tau / timeStep * jacDiagABInv - rel_vel * jacDiagABInv;
      j.applied += impulse;
      const V3 iv = scl(n, impulse);
      applyImpulseRB(A, iv, sub(pivotAInW, A.tr.o));
      applyImpulseRB(B, scl(iv, -1.0f), sub(pivotBInW, B.tr.o));
    }
  }
  // angular part
  const V3 axisA = mulMV(A.tr.b, column(j.frameA.b, 2));
  const V3 axisB = mulMV(B.tr.b, column(j.frameB.b, 2));
  const V3 angVelA = A.av, angVelB = B.av;
  const V3 aroundA = scl(axisA, dot(axisA, angVelA));
  const V3 aroundB = scl(axisB, dot(axisB, angVelB));
  V3 velrelOrthog = sub(sub(angVelA, aroundA), sub(angVelB, aroundB));
  {
    const float relaxation = 1.0f;
    if (len(velrelOrthog) > 0.00001f) {
      const V3 n = nor(velrelOrthog);
      const float denom = angularDenominator(A, n) + angularDenominator(B, n);
      velrelOrthog = scl(velrelOrthog, (1.0f / denom) * j.relaxationFactor);
    }
    V3 angularError = crs(axisA, axisB);
    angularError = scl(angularError, -1.0f);
    angularError = scl(angularError, 1.0f / timeStep);
    if (len(angularError) > 0.00001f) {
      const V3 n2 = nor(angularError);
      const float denom2 = angularDenominator(A, n2) + angularDenominator(B, n2);
      angularError = scl(angularError, (1.0f / denom2) * relaxation);
    }
    applyTorqueImpulse(A, add(scl(velrelOrthog, -1.0f), angularError));
    applyTorqueImpulse(B, sub(velrelOrthog, angularError));
    if (j.solveLimit) {  // on the angular velocities read before this constraint's torques
      const float amplitude =
          (dot(sub(angVelB, angVelA), axisA) * j.relaxationFactor + j.correction * (1.0f / timeStep) * j.biasFactor) *
          j.limitSign;
      float impulseMag = amplitude * j.kHinge;
      const float temp = j.accLimitImpulse;
      j.accLimitImpulse = jmax(j.accLimitImpulse + impulseMag, 0.0f);
      impulseMag = j.accLimitImpulse - temp;
      const V3 impulse = scl(axisA, impulseMag * j.limitSign);
      applyTorqueImpulse(A, impulse);
      applyTorqueImpulse(B, scl(impulse, -1.0f));
    }
  }
  if (j.enableMotor) {
    const V3 angularLimit = v3(0.0f, 0.0f, 0.0f);
    const float projRelVel = dot(sub(aroundA, aroundB), axisA);
    const float motor_relvel = j.motorTarget - projRelVel;
    const float unclipped = j.kHinge * motor_relvel;
    float clipped = unclipped > j.maxMotorImpulse ? j.maxMotorImpulse : unclipped;
    clipped = clipped < -j.maxMotorImpulse ? -j.maxMotorImpulse : clipped;
    const V3 motorImp = scl(axisA, clipped);
    applyTorqueImpulse(A, add(motorImp, angularLimit));
    applyTorqueImpulse(B, sub(scl(motorImp, -1.0f), angularLimit));
  }
}

void shuffleOrder(std::vector<int>& order, JavaRandom& rng) {  // :805-821, one pool
  for (int j = 0; j < (int)order.size(); ++j) {
    int tmp = order[j];
    int swapi = rng.nextInt(j + 1);
    order[j] = order[swapi];
    order[swapi] = tmp;
  }
}

void solveGroup(std::vector<Body>& bodies, std::vector<Manifold*>& manifolds, const SolverInfo& info, bool shuffle,
                JavaRandom& rng, SolverStats* st, const float* learnedInit, std::vector<float>* features,
                std::vector<Joint*>* joints) {
  std::vector<SolverBody> sb;
  std::vector<SolverRow> rows, frows;
  std::vector<int> companion(bodies.size(), -1);
  const bool warm = (info.solverMode & 4) != 0;
  auto solverBodyFor = [&](int bi) -> int {
    const Body& b = bodies[bi];
    if (b.islandTag >= 0) {
      if (companion[bi] >= 0) return companion[bi];
      companion[bi] = (int)sb.size();
    }
    SolverBody s;  // initSolverBody (:151-177); static bodies get a fresh one per manifold
    s.av = b.av;
    s.lv = b.lv;
    s.invMass = b.invMass;
    s.body = bi;
    sb.push_back(s);
    return (int)sb.size() - 1;
  };
  // ---- setup
  for (Manifold* mp : manifolds) {
    Manifold& m = *mp;
    const Body& c0 = bodies[m.body0];
    const Body& c1 = bodies[m.body1];
    int idA = -1, idB = -1;
    if (m.n != 0) {
      idA = solverBodyFor(m.body0);
      idB = solverBodyFor(m.body1);
    }
    for (int j = 0; j < m.n; j++) {
      MPoint& cp = m.p[j];
      if (!(cp.distance <= 0.0f)) continue;
      V3 rp1 = sub(cp.worldA, c0.tr.o);
      V3 rp2 = sub(cp.worldB, c1.tr.o);
      const float relaxation = 1.0f;
      int frictionIndex = (int)rows.size();
      SolverRow r{};
      r.bodyA = idA;
      r.bodyB = idB;
      r.cp = &cp;
      V3 ta0 = crs(rp1, cp.normal);
      r.angA = mulMV(c0.invInertiaWorld, ta0);
      V3 ta1 = crs(rp2, cp.normal);
      r.angB = mulMV(c1.invInertiaWorld, ta1);
      float d0 = c0.invMass + dot(cp.normal, crs(r.angA, rp1));
      float d1 = c1.invMass + dot(cp.normal, crs(r.angB, rp2));
      r.jacInv = relaxation / (d0 + d1);
      r.normal = cp.normal;
      r.rel1xn = crs(rp1, cp.normal);
      r.rel2xn = crs(rp2, cp.normal);
      V3 vel1 = add(crs(c0.av, rp1), c0.lv);  // RigidBody.getVelocityInLocalPoint
      V3 vel2 = add(crs(c1.av, rp2), c1.lv);
      V3 vel = sub(vel1, vel2);
      float rel_vel = dot(cp.normal, vel);
      r.penetration = jmin(cp.distance + info.linearSlop, 0.0f);
      r.friction = cp.friction;
      r.restitution = cp.restitution * -rel_vel;  // restitutionCurve
      if (r.restitution <= 0.0f) r.restitution = 0.0f;
      float penVel = -r.penetration / info.timeStep;
      if (r.restitution > penVel) r.penetration = 0.0f;
      if (features) {  // learned warm start input (16 floats per normal row)
        const float f[16] = {cp.normal.x, cp.normal.y, cp.normal.z, vel.x, vel.y, vel.z,
                             c0.av.x - c1.av.x, c0.av.y - c1.av.y, c0.av.z - c1.av.z, cp.distance,
                             c0.invMass, c1.invMass, cp.applied, cp.lat1, cp.lat2, cp.friction};
        features->insert(features->end(), f, f + 16);
      }
      const int rowIdx = (int)rows.size();
      if (learnedInit) {
        r.applied = learnedInit[rowIdx * 3 + 0];
        applyImpulse(sb[r.bodyA], scl(r.normal, c0.invMass), r.angA, r.applied);
        applyImpulse(sb[r.bodyB], scl(r.normal, c1.invMass), r.angB, -r.applied);
      } else if (warm) {
        r.applied = cp.applied * info.warmstartingFactor;
        applyImpulse(sb[r.bodyA], scl(r.normal, c0.invMass), r.angA, r.applied);
        applyImpulse(sb[r.bodyB], scl(r.normal, c1.invMass), r.angB, -r.applied);
      } else {
        r.applied = 0.0f;
      }
      r.pushApplied = 0.0f;
      r.frictionIndex = (int)frows.size();
      if (!cp.latInit) {  // :678-708
        cp.latDir1 = sub(vel, scl(cp.normal, rel_vel));
        float lat_rel_vel = len2(cp.latDir1);
        if (lat_rel_vel > FLT_EPS) {
          cp.latDir1 = scl(cp.latDir1, 1.0f / jsqrt(lat_rel_vel));
          frows.push_back(makeFrictionRow(cp.latDir1, idA, idB, frictionIndex, cp, rp1, rp2, c0, c1, relaxation));
          cp.latDir2 = nor(crs(cp.latDir1, cp.normal));
          frows.push_back(makeFrictionRow(cp.latDir2, idA, idB, frictionIndex, cp, rp1, rp2, c0, c1, relaxation));
        } else {
          planeSpace1(cp.normal, cp.latDir1, cp.latDir2);
          frows.push_back(makeFrictionRow(cp.latDir1, idA, idB, frictionIndex, cp, rp1, rp2, c0, c1, relaxation));
          frows.push_back(makeFrictionRow(cp.latDir2, idA, idB, frictionIndex, cp, rp1, rp2, c0, c1, relaxation));
        }
        cp.latInit = true;
      } else {
        frows.push_back(makeFrictionRow(cp.latDir1, idA, idB, frictionIndex, cp, rp1, rp2, c0, c1, relaxation));
        frows.push_back(makeFrictionRow(cp.latDir2, idA, idB, frictionIndex, cp, rp1, rp2, c0, c1, relaxation));
      }
      for (int k = 0; k < 2; k++) {  // :710-747
        SolverRow& f = frows[r.frictionIndex + k];
        if (learnedInit) {
          f.applied = learnedInit[rowIdx * 3 + 1 + k];
        } else if (warm) {
          f.applied = (k == 0 ? cp.lat1 : cp.lat2) * info.warmstartingFactor;
        } else {
          f.applied = 0.0f;
          continue;
        }
        applyImpulse(sb[r.bodyA], scl(f.normal, c0.invMass), f.angA, f.applied);
        applyImpulse(sb[r.bodyB], scl(f.normal, c1.invMass), f.angB, -f.applied);
      }
      rows.push_back(r);
    }
  }
  if (joints)  // buildJacobian of every constraint after the contact rows (:757-763)
    for (Joint* jt : *joints) jointBuildJacobian(*jt, bodies);
  std::vector<int> order(rows.size()), forder(frows.size());
  for (size_t i = 0; i < rows.size(); i++) order[i] = (int)i;
  for (size_t i = 0; i < frows.size(); i++) forder[i] = (int)i;
  // ---- iterations (:788-890)
  int iters = info.numIterations;
  if (info.tolerance > 0.0f) iters = info.maxIterations;
  std::vector<float> prev(rows.size() + frows.size());
  int done = 0;
  float residual = 0.0f;
  double lastDD = 0.0, lastNN = 0.0;  // the last iteration's sums (whole-step residual over all islands)
  if (st) st->residualHistory.clear();
  for (int it = 0; it < iters; it++) {
    if (shuffle && (info.solverMode & 1) && ((it & 7) == 0)) {  // normal rows first, then friction rows
      shuffleOrder(order, rng);
      shuffleOrder(forder, rng);
    }
    for (size_t i = 0; i < rows.size(); i++) prev[i] = rows[i].applied;
    for (size_t i = 0; i < frows.size(); i++) prev[rows.size() + i] = frows[i].applied;
    if (joints) {  // constraints first, on the rigid bodies' velocities (:823-842)
      for (Joint* jt : *joints) {
        const int ids[2] = {jt->bodyA, jt->bodyB};
        for (int bi : ids)  // SolverBody.writebackVelocity of a body that has a solver body
          if (bi >= 0 && bodies[bi].islandTag >= 0 && companion[bi] >= 0 && sb[companion[bi]].invMass != 0.0f) {
            bodies[bi].lv = sb[companion[bi]].lv;
            bodies[bi].av = sb[companion[bi]].av;
          }
        jointSolve(*jt, bodies, info.timeStep);
        for (int bi : ids)  // SolverBody.readVelocity
          if (bi >= 0 && bodies[bi].islandTag >= 0 && companion[bi] >= 0 && sb[companion[bi]].invMass != 0.0f) {
            sb[companion[bi]].lv = bodies[bi].lv;
            sb[companion[bi]].av = bodies[bi].av;
          }
      }
    }
    for (int j = 0; j < (int)order.size(); j++) {  // resolveSingleCollisionCombinedCacheFriendly (:234-284)
      SolverRow& c = rows[order[j]];
      SolverBody& A = sb[c.bodyA];
      SolverBody& B = sb[c.bodyB];
      float v1 = dot(c.normal, A.lv) + dot(c.rel1xn, A.av);
      float v2 = dot(c.normal, B.lv) + dot(c.rel2xn, B.av);
      float rel_vel = v1 - v2;
      float positionalError = -c.penetration * info.erp / info.timeStep;
      float velocityError = c.restitution - rel_vel;
      float pi = positionalError * c.jacInv;
      float vi = velocityError * c.jacInv;
      float ni = pi + vi;
      float old = c.applied;
      float sum = old + ni;
      c.applied = 0.0f > sum ? 0.0f : sum;
      ni = c.applied - old;
      applyImpulse(A, scl(c.normal, A.invMass), c.angA, ni);
      applyImpulse(B, scl(c.normal, B.invMass), c.angB, -ni);
    }
    for (int j = 0; j < (int)forder.size(); j++) {  // resolveSingleFrictionCacheFriendly (:286-346)
      SolverRow& c = frows[forder[j]];
      float total = rows[c.frictionIndex].applied + rows[c.frictionIndex].pushApplied;
      float limit = total * c.friction;
      if (total > 0.0f) {
        SolverBody& A = sb[c.bodyA];
        SolverBody& B = sb[c.bodyB];
        float v1 = dot(c.normal, A.lv) + dot(c.rel1xn, A.av);
        float v2 = dot(c.normal, B.lv) + dot(c.rel2xn, B.av);
        float rel_vel = v1 - v2;
        float j1 = -rel_vel * c.jacInv;
        float old = c.applied;
        c.applied = old + j1;
        if (limit < c.applied) {
          c.applied = limit;
        } else if (c.applied < -limit) {
          c.applied = -limit;
        }
        j1 = c.applied - old;
        applyImpulse(A, scl(c.normal, A.invMass), c.angA, j1);
        applyImpulse(B, scl(c.normal, B.invMass), c.angB, -j1);
      }
    }
    done = it + 1;
    double dd = 0, nn = 0;
    for (size_t i = 0; i < rows.size(); i++) {
      double d = (double)rows[i].applied - (double)prev[i];
      dd += d * d;
      nn += (double)rows[i].applied * (double)rows[i].applied;
    }
    for (size_t i = 0; i < frows.size(); i++) {
      double d = (double)frows[i].applied - (double)prev[rows.size() + i];
      dd += d * d;
      nn += (double)frows[i].applied * (double)frows[i].applied;
    }
    double nrm = std::sqrt(nn);
    residual = (float)(std::sqrt(dd) / (nrm > 1.0 ? nrm : 1.0));
    lastDD = dd;
    lastNN = nn;
    if (st) st->residualHistory.push_back(residual);
    if (info.tolerance > 0.0f && residual <= info.tolerance) break;
  }
  // ---- writeback (:892-924): appliedImpulseLateral1 is written twice, the second
  // time with the second friction row; appliedImpulseLateral2 is never stored.
  for (auto& r : rows) {
    r.cp->applied = r.applied;
    r.cp->lat1 = frows[r.frictionIndex].applied;
    r.cp->lat1 = frows[r.frictionIndex + 1].applied;
  }
  for (auto& s : sb) {  // SolverBody.writebackVelocity (:87-93)
    if (s.invMass != 0.0f) {
      bodies[s.body].lv = s.lv;
      bodies[s.body].av = s.av;
    }
  }
  if (st) {
    st->iterations += done;
    st->residual = residual;
    st->finalDD += lastDD;
    st->finalNN += lastNN;
    st->rows += (int)rows.size();
    st->frictionRows += (int)frows.size();
  }
}

}  // namespace jb
