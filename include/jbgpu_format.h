/*
 * jbgpu_format.h — record layouts and enumerations shared by the jbgpu C ABI
 * (include/jbgpu.h) and every caller of it.  Plain C, no dependencies.
 *
 * All records are arrays of 32-bit words (float, or int32 stored in the same
 * word).  Matrices are row-major: basis[r*3+c] is element (r, c), which is
 * libGDX Matrix3.val[M_rc] in the reference (linearmath/Transform.java:43-49).
 */
#ifndef JBGPU_FORMAT_H
#define JBGPU_FORMAT_H

/* Shape types: the BroadphaseNativeType values the hot path dispatches on
 * (collision/broadphase/BroadphaseNativeType.java:34-112). */
#define JBGPU_SHAPE_SPHERE 0  /* SphereShape (collision/shapes/SphereShape.java) */
#define JBGPU_SHAPE_BOX 1     /* BoxShape (collision/shapes/BoxShape.java) */
#define JBGPU_SHAPE_CAPSULE 2 /* CapsuleShape, Y up (collision/shapes/CapsuleShape.java) */
#define JBGPU_SHAPE_HULL 3    /* ConvexHullShape (collision/shapes/ConvexHullShape.java) */
#define JBGPU_SHAPE_PLANE 4   /* StaticPlaneShape (collision/shapes/StaticPlaneShape.java) */

/* Shape record, JBGPU_SHAPE_WORDS words:
 *  [0] type (int)        [1] collision margin (getMargin())
 *  [2..4] implicitShapeDimensions (box: half extents minus margin; sphere: (r,0,0);
 *         capsule: (r, halfHeight, r))
 *  [5] first hull point (int)  [6] hull point count (int)
 *  [7..9] hull local AABB min  [10..12] hull local AABB max  (recalcLocalAabb)
 *  [13..15] plane normal       [16] plane constant */
#define JBGPU_SHAPE_WORDS 20

/* Body record, JBGPU_BODY_WORDS words:
 *  [0..8] basis (row-major)  [9..11] origin  [12..14] linear velocity
 *  [15..17] angular velocity [18] inverse mass  [19..21] inverse local inertia
 *  [22..24] gravity force (RigidBody.gravity)   [25] friction  [26] restitution
 *  [27] shape index (int)    [28] flags (int)   [29] activation state (int)
 *  [30] deactivation time */
#define JBGPU_BODY_WORDS 32
#define JBGPU_BODY_STATIC 1

/* CollisionObject activation states (collision/dispatch/CollisionObject.java:41-45) */
#define JBGPU_ACTIVE_TAG 1
#define JBGPU_ISLAND_SLEEPING 2
#define JBGPU_WANTS_DEACTIVATION 3
#define JBGPU_DISABLE_DEACTIVATION 4
#define JBGPU_DISABLE_SIMULATION 5

/* Broadphase semantics selected by a context (BroadphaseInterface implementations) */
#define JBGPU_BP_DBVT 0   /* DbvtBroadphase: stored-AABB fattening (bp/DbvtBroadphase.java:196-228) */
#define JBGPU_BP_SAP16 1  /* AxisSweep3, 16-bit quantisation (bp/AxisSweep3.java:51-55) */
#define JBGPU_BP_SAP32 2  /* AxisSweep3_32 (bp/AxisSweep3_32.java:48-52) */
#define JBGPU_BP_SIMPLE 3 /* SimpleBroadphase: raw AABBs (bp/SimpleBroadphase.java:81-110) */

/* Proxy record, JBGPU_PROXY_WORDS words:
 *  [0] stage (int; 0/1 dynamic-set stages, 2 fixed set)
 *  [1..3] leaf volume min  [4..6] leaf volume max  [7..9] stored AABB min
 *  [10..12] stored AABB max [13] filter group (int) [14] filter mask (int) */
#define JBGPU_PROXY_WORDS 16

/* Manifold point record, JBGPU_POINT_WORDS words (ManifoldPoint.java:36-62):
 *  [0..2] localPointA [3..5] localPointB [6..8] positionWorldOnA [9..11] positionWorldOnB
 *  [12..14] normalWorldOnB [15] distance [16] combined friction [17] combined restitution
 *  [18] appliedImpulse [19] appliedImpulseLateral1 [20] appliedImpulseLateral2
 *  [21..23] lateralFrictionDir1 [24..26] lateralFrictionDir2
 *  [27] lateralFrictionInitialized (int) [28] lifeTime (int)
 * The two friction directions are meaningful only while [27] is 1. */
#define JBGPU_POINT_WORDS 32
/* Manifold record: [0] point count [1] body0 index [2] body1 index [3] pair type,
 * then 4 point records. */
#define JBGPU_MANIFOLD_HDR 4
#define JBGPU_MANIFOLD_WORDS (JBGPU_MANIFOLD_HDR + 4 * JBGPU_POINT_WORDS)

/* Pair types produced by the dispatcher's double-dispatch table
 * (collision/dispatch/DefaultCollisionConfiguration.java:149-213) */
#define JBGPU_PAIR_SPHERE_SPHERE 0
#define JBGPU_PAIR_CONVEX_PLANE 1
#define JBGPU_PAIR_CONVEX_CONVEX 2
#define JBGPU_PAIR_NONE 3

/* Solver row order */
#define JBGPU_ORDER_REFERENCE 0 /* islands, Random(0) shuffle (SequentialImpulseConstraintSolver.java:805-821) */
#define JBGPU_ORDER_COLOR 1     /* graph-coloured batches */
#define JBGPU_ORDER_EXPLICIT 2  /* caller-supplied manifold order, no shuffle */

#endif /* JBGPU_FORMAT_H */
