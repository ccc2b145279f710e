// Oracle (test infrastructure only — never linked into the product library).
// A single-threaded C++ restatement of the reference JBullet hot path:
// AABB update, DbvtBroadphase / AxisSweep3 / SimpleBroadphase pair sets,
// collision dispatch (sphere-sphere, convex-plane, GJK + GJK-EPA), the
// persistent manifold, the cache-friendly sequential-impulse solver, the
// integrator, activation and islands.  Every function cites the Java file:line
// it follows (paths relative to src/com/bulletphysics/).
//
// Parity pinning (what fixes each part, since the reference has no tests):
//   jrandom          KATs of java.util.Random, the shuffle permutation   tests/test_oracle_pins.py
//   jmath / libGDX   orthonormality, known rotations                     tests/test_oracle_pins.py
//   broadphase       O(N^2) brute force, closed-form quantised overlap, invariants
//   sphere/plane     closed forms
//   GJK / EPA        double-precision brute-force distance/depth on boxes & spheres
//   manifold         hand-computed add/replace/evict (sortCachedPoints) and refresh sequences
//   solver           resting-box impulse m*g*dt, head-on e=0, cone/complementarity, hand-computed
//                    warm start (0.85 factor, Lateral1 written twice, directions kept per point lifetime)
//   integrator       exact free fall, constant-omega rotation
// libGDX arithmetic (Quaternion.setFromMatrix / nor) is restated from libGDX's
// documented behaviour; its bit parity against a libGDX build is unpinned.
#pragma once
#include <cstdint>
#include <vector>

#include "jmath.h"
#include "jrandom.h"
#include "../include/jbgpu_format.h"

namespace jb {

// ---- shapes (collision/shapes/*.java) --------------------------------------
struct Shape {
  int type;       // JBGPU_SHAPE_*
  float margin;   // getMargin()
  V3 dims;        // implicitShapeDimensions
  int vfirst, vcount;  // hull points
  V3 lmin, lmax;  // hull local AABB (PolyhedralConvexShape.recalcLocalAabb)
  V3 pn;          // plane normal
  float pc;       // plane constant
};

V3 supportNoMargin(const Shape& s, const V3* hullv, V3 dir);   // localGetSupportingVertexWithoutMargin
V3 supportMargin(const Shape& s, const V3* hullv, V3 dir);     // localGetSupportingVertex
void shapeAabb(const Shape& s, const Tr& t, V3& mn, V3& mx);   // getAabb
V3 shapeLocalInertia(const Shape& s, const V3* hullv, float mass);  // calculateLocalInertia
Shape makeSphere(float r);
Shape makeBox(V3 halfExtents);
Shape makeCapsule(float r, float height);
Shape makeHull(const V3* pts, int first, int n);
Shape makePlane(V3 n, float c);

// ---- bodies (dynamics/RigidBody.java, collision/dispatch/CollisionObject.java)
struct Body {
  int shape;
  Tr tr;
  V3 lv, av;
  float invMass;
  V3 invInertiaLocal;
  M3 invInertiaWorld;
  V3 gravity;          // RigidBody.gravity (force, acceleration / invMass)
  float friction, restitution;
  int flags;           // JBGPU_BODY_STATIC
  int activation;      // CollisionObject.activationState1
  float deactivationTime;
  int islandTag, companionId;
  float ccdMotionThreshold = 0.0f, ccdSweptSphereRadius = 0.0f;  // CollisionObject.java:83-86 (0: no CCD)
  bool isStatic() const { return (flags & JBGPU_BODY_STATIC) != 0; }
  bool isActive() const { return activation != JBGPU_ISLAND_SLEEPING && activation != JBGPU_DISABLE_SIMULATION; }
};
void updateInertiaTensor(Body& b);  // RigidBody.java:449-459
Tr integrateTransform(const Tr& cur, V3 lv, V3 av, float dt);  // TransformUtil.integrateTransform (:64-103)

// ---- broadphase -----------------------------------------------------------
struct Aabb {
  V3 mn, mx;
};
inline bool intersect(const Aabb& a, const Aabb& b) {  // DbvtAabbMm.java:209-212
  return (a.mn.x <= b.mx.x) && (a.mx.x >= b.mn.x) && (a.mn.y <= b.mx.y) && (a.mx.y >= b.mn.y) &&
         (a.mn.z <= b.mx.z) && (a.mx.z >= b.mn.z);
}
struct Proxy {
  int stage;        // DbvtProxy.stage (0/1 dynamic set stages, 2 = fixed set)
  Aabb leaf;        // leaf.volume in the dbvt
  Aabb aabb;        // DbvtProxy.aabb (the stored, possibly fattened box)
  int16_t group, mask;
  int32_t q[6];     // AxisSweep3 quantised min xyz / max xyz
  int32_t world = 0;  // batched independent worlds: proxies of different worlds never pair
  bool dead = false;  // destroyProxy: out of the tree and the stage lists, never pairs again
};
struct Broadphase {
  int mode;  // JBGPU_BP_*
  int stageCurrent = 0;
  std::vector<Proxy> proxies;  // index = uid - 1
  // AxisSweep3
  V3 wmin, wmax, quant;
  uint32_t handleMask = 0, sentinel = 0, qmask = 0;
  void init(int mode, V3 wmin, V3 wmax);
  void createProxy(Aabb a, int16_t group, int16_t mask);
  void destroyProxy(int idx);  // DbvtBroadphase.java:184-194, AxisSweep3Internal.removeHandle, SimpleBroadphase
  void setAabb(int idx, Aabb a);
  void collide();  // stage rotation of DbvtBroadphase.collide
  // overlapping pair set, canonical (uid_lo, uid_hi), sorted ascending
  void pairs(std::vector<uint64_t>& out) const;
  void pairsBrute(std::vector<uint64_t>& out) const;
  bool overlap(int i, int j) const;
  void quantize(V3 p, int isMax, int32_t* out) const;
};
inline uint64_t pairKey(uint32_t uidLo, uint32_t uidHi) { return ((uint64_t)uidLo << 32) | uidHi; }

// ---- manifold (collision/narrowphase/ManifoldPoint.java, PersistentManifold.java)
struct MPoint {
  V3 localA, localB, worldA, worldB, normal;
  float distance, friction, restitution;
  float applied, lat1, lat2;
  bool latInit;
  V3 latDir1, latDir2;
  int lifetime;
};
struct Manifold {
  int body0, body1;  // body indices (manifold.getBody0/1)
  int n;
  MPoint p[4];
};
int manifoldCacheEntry(const Manifold& m, const MPoint& np);               // PersistentManifold.java:214-233
int manifoldAddPoint(Manifold& m, const MPoint& np);                       // :235-257
void manifoldReplacePoint(Manifold& m, const MPoint& np, int idx);         // :280-305
void manifoldRemovePoint(Manifold& m, int idx);                            // :259-278
void manifoldRefresh(Manifold& m, const Tr& trA, const Tr& trB);           // :312-372
int manifoldSortCachedPoints(const Manifold& m, const MPoint& pt);         // :83-156

// ---- narrowphase -----------------------------------------------------------
struct ContactCandidate {
  bool valid;
  V3 normalOnB, pointOnB;
  float depth;
};
struct GjkStats {
  int gjkIters = 0, epaCalls = 0, epaFails = 0;
};
ContactCandidate sphereSphere(const Shape& s0, const Tr& t0, const Shape& s1, const Tr& t1, bool& touching);
ContactCandidate convexPlane(const Shape& convex, const V3* hv, const Tr& tc, const Shape& plane, const Tr& tp);
// penetrationSolver false: GjkPairDetector without a penetration depth solver (the convex casts' use)
ContactCandidate gjkPairDetector(const Shape& sa, const Tr& ta, const Shape& sb, const Tr& tb, const V3* hv,
                                 float maxDist2, GjkStats* st, bool penetrationSolver = true);
bool gjkEpaPenDepth(const Shape& sa, const Tr& ta, const Shape& sb, const Tr& tb, const V3* hv, V3& wA, V3& wB,
                    int* epaIters);
// ManifoldResult.addContactPoint (ManifoldResult.java:92-157)
void addContactPoint(Manifold& m, const Body& r0, const Body& r1, int rbody0, const ContactCandidate& c);
// ManifoldResult.refreshContactPoints (:177-191)
void refreshContactPoints(Manifold& m, const Body* bodies);

// ---- solver (dynamics/constraintsolver/SequentialImpulseConstraintSolver.java)
struct SolverInfo {
  float timeStep = 1.0f / 60.0f;
  int numIterations = 10;
  float erp = 0.2f;
  float linearSlop = 0.0f;
  float warmstartingFactor = 0.85f;
  int solverMode = 1 | 8 | 4;  // RANDOMIZE | CACHE_FRIENDLY | WARMSTART (ContactSolverInfo.java:48)
  // project additions: run to a residual tolerance (0 = fixed iterations)
  float tolerance = 0.0f;
  int maxIterations = 100;
};
struct SolverStats {
  int iterations = 0;
  float residual = 0.0f;
  double finalDD = 0.0, finalNN = 0.0;  // last-iteration sums of (dl)^2 and l^2 over every island solved
  int rows = 0, frictionRows = 0;
  std::vector<float> residualHistory;
};
struct SolverRow {
  int bodyA, bodyB;  // solver body ids
  V3 normal, rel1xn, rel2xn, angA, angB;
  float jacInv, penetration, restitution, friction, applied, pushApplied;
  int frictionIndex;
  MPoint* cp;
};
struct SolverBody {
  V3 lv, av;
  float invMass;
  int body;  // original body index (-1 for none)
};
struct SolverState;  // opaque per-call temporaries
struct ImpulseInit { // learned warm start: override of the warm-start impulses per row
  const float* values = nullptr;  // [row][3] normal, lat1, lat2 (already scaled)
};

// Point2PointConstraint (dynamics/constraintsolver/Point2PointConstraint.java): a ball socket between pivotA in
// bodyA's frame and pivotB in bodyB's frame; bodyB == -1 is TypedConstraint's fixed body (static, identity
// transform, zero velocity), pivotB then a world point.  adiag / applied are the per-solve state of buildJacobian.
// HingeConstraint (dynamics/constraintsolver/HingeConstraint.java): frameA / frameB are rbAFrame / rbBFrame as
// the constructors leave them (B's third column already flipped); the linear part is three JacobianEntry rows
// along linAxis (buildJacobian), the angular part a torque on the orthogonal relative angular velocity and the
// axis misalignment, plus an optional limit and motor about the hinge axis.
struct Joint {
  int type = 0;  // 0 Point2PointConstraint, 1 HingeConstraint
  int bodyA, bodyB;
  V3 pivotA, pivotB;
  float tau = 0.3f, damping = 1.0f, impulseClamp = 0.0f;  // ConstraintSetting (:202-206)
  float adiag[3];
  float applied;
  Tr frameA, frameB;
  float lowerLimit = 1e30f, upperLimit = -1e30f, limitSoftness = 0.9f, biasFactor = 0.3f, relaxationFactor = 1.0f;
  bool angularOnly = false, enableMotor = false;
  float motorTarget = 0.0f, maxMotorImpulse = 0.0f;
  V3 linAxis[3];
  float kHinge = 0.0f, correction = 0.0f, limitSign = 0.0f, accLimitImpulse = 0.0f;
  bool solveLimit = false;
};
// The HingeConstraint constructors' frames from the body A transform at construction:
// (rbA, rbB, pivotInA, pivotInB, axisInA, axisInB) (HingeConstraint.java:84-130), and with fixedB
// (rbA, pivotInA, axisInA) (:132-178, pivotB / axisB unused).
void hingeFramesFromAxes(const Tr& trA, V3 pivotA, V3 pivotB, V3 axisA, V3 axisB, bool fixedB, Tr& frameA, Tr& frameB);
// (rbA, rbB, rbAFrame, rbBFrame) (:180-210) and, with fixedB, (rbA, rbAFrame) (:212-238).
void hingeFramesFromFrames(const Tr& trA, const Tr& inA, const Tr& inB, bool fixedB, Tr& frameA, Tr& frameB);
float hingeAngle(const Joint& j, const std::vector<Body>& bodies);  // getHingeAngle (:530-549)
float atan2Fast(float y, float x);                                   // ScalarUtil.atan2Fast (ScalarUtil.java:43-57)
// The row-order shuffle of one pool: j = 0..n-1 swaps order[j] with order[Random.nextInt(j + 1)]
// (SequentialImpulseConstraintSolver.java:805-821, randInt2 :138-149).
void shuffleOrder(std::vector<int>& order, JavaRandom& rng);
// One solveGroup over an ordered list of manifolds (SequentialImpulseConstraintSolver.java:954-969).
// shuffle: apply the Random shuffle at iterations with (iter & 7) == 0.
void solveGroup(std::vector<Body>& bodies, std::vector<Manifold*>& manifolds, const SolverInfo& info, bool shuffle,
                JavaRandom& rng, SolverStats* st, const float* learnedInit, std::vector<float>* features,
                std::vector<Joint*>* joints = nullptr);
// Point2PointConstraint.buildJacobian (:71-107) and solveConstraint (:109-177) on the bodies' own velocities
void jointBuildJacobian(Joint& j, const std::vector<Body>& bodies);
void jointSolve(Joint& j, std::vector<Body>& bodies, float timeStep);

// ---- islands (collision/dispatch/SimulationIslandManager.java, UnionFind.java)
struct UnionFind {
  std::vector<int> id, sz;
  void reset(int n);
  int find(int x);
  void unite(int p, int q);
  void sortIslands();  // id := find(i), sz := i, quicksort by id
};

// ---- world -----------------------------------------------------------------
struct WorldConfig {
  int bpMode = JBGPU_BP_DBVT;
  V3 worldMin, worldMax;  // AxisSweep3 bounds
  SolverInfo solver;
  bool deactivationDisabled = true;  // BulletGlobals.setDeactivationDisabled(true)
  V3 gravity;
  int solverOrder = JBGPU_ORDER_REFERENCE;
};
struct PairState {
  uint64_t key;
  Manifold m;
};
struct RayHit {  // ClosestRayResultCallback: collisionObject, closestHitFraction, hitNormalWorld, hitPointWorld
  int object;
  float fraction;
  V3 normal, point;
};
struct World {
  WorldConfig cfg;
  std::vector<Shape> shapes;
  std::vector<V3> hullv;
  std::vector<Body> bodies;
  Broadphase bp;
  std::vector<PairState> pairs;  // sorted by key
  JavaRandom rng{0};
  UnionFind uf;
  SolverStats lastSolve;
  GjkStats gjk;
  std::vector<int> explicitOrder;  // manifold order for JBGPU_ORDER_EXPLICIT (indices into pairs)
  std::vector<Joint> joints;       // DiscreteDynamicsWorld.addConstraint, in creation order
  std::vector<int> explicitJointOrder;  // joint order for JBGPU_ORDER_EXPLICIT
  const float* learnedInit = nullptr;
  std::vector<float>* features = nullptr;
  long stepCount = 0;

  void addBody(const Body& b);
  void step(float dt);
  // phases
  void predictUnconstraintMotion(float dt);
  void updateAabbs();
  void calculateOverlappingPairs();
  void dispatchAllCollisionPairs();
  void calculateSimulationIslands();
  void solveConstraints();
  void integrateTransforms(float dt);
  void updateActivationState(float dt);
  bool needsCollision(int a, int b) const;
  bool needsResponse(int a, int b) const;
};

// CollisionWorld.rayTest with a ClosestRayResultCallback over the bodies of `world`
RayHit rayTestClosest(const World& w, V3 from, V3 to, int16_t group, int16_t mask, int world);
// CollisionWorld.convexSweepTest with a ClosestConvexResultCallback over the bodies of `world` (RayHit fields:
// hitCollisionObject, closestHitFraction, hitNormalWorld, hitPointWorld).  Static planes are skipped: the
// reference's objectQuerySingle has no caster for them (it throws).
RayHit convexSweepClosest(const World& w, const Shape& castShape, const Tr& from, const Tr& to, int16_t group,
                          int16_t mask, int world);
// GjkConvexCast.calcTimeOfImpact (GjkConvexCast.java:67-193) for a static B (fromB == toB)
bool gjkConvexCast(const Shape& sa, const Tr& fromA, const Tr& toA, const Shape& sb, const Tr& tb, const V3* hv,
                   float allowedPenetration, float& fraction, V3& normal, V3& hitPoint);
// TransformUtil.calculateVelocity (TransformUtil.java:105-115, calculateDiffAxisAngle :117-151)
void calculateVelocity(const Tr& t0, const Tr& t1, float timeStep, V3& linVel, V3& angVel);
// CollisionShape.calculateTemporalAabb (CollisionShape.java:76-122) with getAngularMotionDisc (:44-72)
void temporalAabb(const Shape& s, const Tr& cur, V3 linvel, V3 angvel, float timeStep, V3& mn, V3& mx);
V3 interp3(V3 a, V3 b, float t);  // VectorUtil.setInterpolate3
// the CCD motion clamp of integrateTransforms (DiscreteDynamicsWorld.java:693-725) for body i with predicted
// transform `pred`: the clamped transform (pred itself when no sweep hit counts)
Tr ccdClamp(World& w, int i, const Tr& pred, float dt);
bool rayAabb(V3 from, V3 to, V3 mn, V3 mx, float& param, V3& normal);
bool subsimplexRayCast(const Shape& sb, const V3* hv, const Tr& fromA, const Tr& toA, const Tr& fromB, const Tr& toB,
                       float& fraction, V3& normal);

}  // namespace jb
