/*
 * jbgpu.h — C ABI of libjbgpu, the MI355X (gfx950) rigid-body hot path.
 *
 * Plain C: opaque context handle, integer status codes (0 = OK, < 0 = error,
 * message via jbgpu_last_error), caller-owned host buffers borrowed only for
 * the duration of a call, records laid out as in jbgpu_format.h.  A context is
 * used by one host thread at a time and owns one HIP stream.
 *
 * Each group of entry points stands in for one plugin seam of the reference
 * (paths relative to src/com/bulletphysics/ in the reference):
 *   BroadphaseInterface      collision/broadphase/BroadphaseInterface.java:33-52
 *   Dispatcher (+ CollisionAlgorithm per pair)
 *                            collision/broadphase/Dispatcher.java:38-68,
 *                            collision/broadphase/CollisionAlgorithm.java:38-70
 *   ConstraintSolver         dynamics/constraintsolver/ConstraintSolver.java:37-55
 *   DynamicsWorld.stepSimulation (the fused device-resident step)
 *                            dynamics/DynamicsWorld.java:53-74,
 *                            dynamics/DiscreteDynamicsWorld.java:290-400
 */
#ifndef JBGPU_H
#define JBGPU_H

#include <stdint.h>

#include "jbgpu_format.h"

#ifdef __cplusplus
extern "C" {
#endif

#define JBGPU_OK 0
#define JBGPU_ERR_ARG (-1)
#define JBGPU_ERR_HIP (-2)
#define JBGPU_ERR_STATE (-3)
#define JBGPU_ERR_CAPACITY (-4)
#define JBGPU_ERR_NO_DEVICE (-5)

typedef struct jbgpu_ctx jbgpu_ctx;

typedef struct jbgpu_config {
  int32_t bp_mode;               /* JBGPU_BP_*: which BroadphaseInterface semantics */
  float world_min[3];            /* AxisSweep3 world bounds (AxisSweep3Internal.java:80-103) */
  float world_max[3];
  int32_t num_iterations;        /* ContactSolverInfo.numIterations (ContactSolverInfo.java:38) */
  int32_t solver_order;          /* JBGPU_ORDER_COLOR (the only order on the device; others: JBGPU_ERR_ARG) */
  int32_t deactivation_disabled; /* BulletGlobals.setDeactivationDisabled (BulletGlobals.java:114-116) */
  float tolerance;               /* > 0: iterate to residual <= tolerance (capped by max_iterations) */
  int32_t max_iterations;
  int32_t warmstart;             /* SolverMode.SOLVER_USE_WARMSTARTING */
  float erp;                     /* ContactSolverInfo.erp (0.2) */
  float warmstart_factor;        /* ContactSolverInfo.warmstartingFactor (0.85) */
} jbgpu_config;

typedef struct jbgpu_stats {
  int64_t pairs;        /* overlapping pairs after the last broadphase pass */
  int64_t contacts;     /* manifold points after the last narrowphase pass */
  int64_t manifolds;    /* manifolds that produced solver rows */
  int64_t colors;       /* colour batches of the last solve */
  int64_t epa_calls;    /* pairs that ran GJK-EPA */
  int64_t epa_fails;
  int64_t iterations;   /* solver iterations run in the last solve */
  float residual;       /* residual after the last iteration */
  int64_t huge_proxies; /* proxies handled outside the BVH (e.g. planes) */
  int64_t rows;         /* normal rows (contacts with distance <= 0) of the last solve */
  int64_t epa_retries;  /* EPA jobs of the last pass that outgrew the first-pass workspace and ran again */
  int64_t epa_huge;     /* ... of those, the ones that also outgrew the large LDS workspace (global-memory tier) */
  int64_t graph_redos;  /* graph-captured steps that outgrew their pair capacity and were finished on the ordinary
                           path (since the context was created) */
} jbgpu_stats;

/* ---- context ---------------------------------------------------------------- */
int jbgpu_version(void);
int jbgpu_device_count(int32_t* count);
void jbgpu_config_default(jbgpu_config* cfg);
/* New context on the current HIP device.  Replaces `new DbvtBroadphase()` / `new
 * AxisSweep3(_32)(min, max)` + `new CollisionDispatcher(cfg)` + `new
 * SequentialImpulseConstraintSolver()` (DiscreteDynamicsWorld.java:86-103). */
int jbgpu_ctx_create(const jbgpu_config* cfg, jbgpu_ctx** out);
int jbgpu_ctx_destroy(jbgpu_ctx* ctx);
/* Copies the last error message of the context (or of the last failed create when ctx is NULL). */
int jbgpu_last_error(const jbgpu_ctx* ctx, char* buf, int32_t len);

/* ---- shape helpers (host only) ------------------------------------------------ */
/* Shape record as built by the reference's constructors: SphereShape(r), BoxShape(he),
 * CapsuleShape(r, height), ConvexHullShape(points) (recalcLocalAabb), StaticPlaneShape(n, c).
 * params: sphere {r}; box {hx,hy,hz}; capsule {r,height}; hull {first,count} (points in verts);
 * plane {nx,ny,nz,c}.  (the shape constructors in collision/shapes/) */
int jbgpu_make_shape(int32_t type, const float* params, const float* verts, float* rec);
/* CollisionShape.calculateLocalInertia(mass, inertia) for a shape record. */
int jbgpu_shape_local_inertia(const float* rec, const float* verts, float mass, float* inertia3);

/* ---- BroadphaseInterface -------------------------------------------------------- */
/* createProxy for n AABBs (min xyz, max xyz per proxy), appended after the proxies created so far;
 * uid_out[i] = creation index + 1 in every mode, never reused (DbvtBroadphase.java:173-182 gid;
 * AxisSweep3 and SimpleBroadphase number their handles differently, the pair set is the same).
 * world ids may be NULL (all 0); pairs only form inside a world. */
int jbgpu_bp_create_proxies(jbgpu_ctx* ctx, int32_t n, const float* aabb6, const int16_t* group,
                            const int16_t* mask, const int32_t* world, int32_t* uid_out);
/* setAabb for every proxy (DbvtBroadphase.java:196-228 / AxisSweep3Internal.java:542-579).
 * aabb6 is a host pointer, or a device pointer when on_device != 0. */
int jbgpu_bp_set_aabbs(jbgpu_ctx* ctx, int32_t n, const float* aabb6, int32_t on_device);
/* calculateOverlappingPairs (DbvtBroadphase.java:230, collide :89-150). */
int jbgpu_bp_calculate_overlapping_pairs(jbgpu_ctx* ctx, int64_t* npairs);
/* getOverlappingPairCache().getOverlappingPairArray() as sorted keys (uid_lo << 32 | uid_hi). */
int jbgpu_bp_get_pairs(jbgpu_ctx* ctx, uint64_t* keys, int64_t cap, int64_t* n);
/* destroyProxy for n uids (DbvtBroadphase.java:184-194, AxisSweep3Internal.java:498-540,
 * SimpleBroadphase.destroyProxy): the proxies never pair again and, as
 * removeOverlappingPairsContainingProxy (HashedOverlappingPairCache.java), their pairs leave the current
 * pair set at once.  Later setAabb calls still pass one box per created proxy (ignored for destroyed
 * ones).  Broadphase-only contexts; a uid destroyed twice is an error. */
int jbgpu_bp_destroy_proxies(jbgpu_ctx* ctx, int32_t n, const int32_t* uids);
/* The pairs the last pair computation (calculate_overlapping_pairs or a world step) added and removed,
 * as sorted keys: what addOverlappingPair / removeOverlappingPair calls a pair cache would have seen
 * (HashedOverlappingPairCache.java:65-110), so a caller keeps its per-pair objects across steps.
 * Either array may be NULL (counts only); error if a count exceeds its capacity. */
int jbgpu_bp_get_pair_deltas(jbgpu_ctx* ctx, uint64_t* added, int64_t cap_added, int64_t* n_added,
                             uint64_t* removed, int64_t cap_removed, int64_t* n_removed);
int jbgpu_bp_get_proxies(jbgpu_ctx* ctx, float* rec, int32_t* stage_current);
int jbgpu_bp_set_proxies(jbgpu_ctx* ctx, const float* rec, int32_t stage_current);
int jbgpu_bp_get_quantized(jbgpu_ctx* ctx, int32_t* q6);

/* ---- DynamicsWorld ---------------------------------------------------------------- */
/* addRigidBody for every body record (DiscreteDynamicsWorld.java:426-440): creates the proxies
 * (static bodies: filter group 2, mask ~2) with the shapes' AABBs. */
int jbgpu_world_create(jbgpu_ctx* ctx, int32_t nshapes, const float* shapes, int32_t nverts, const float* verts,
                       int32_t nbodies, const float* bodies, const int32_t* world_ids);
/* CollisionWorld.rayTest with a ClosestRayResultCallback for each of n rays (CollisionWorld.java:552-591,
 * rayTestSingle :260-336, ClosestRayResultCallback :697-729) against the bodies' current transforms:
 * convex shapes by SubsimplexConvexCast, planes by their two triangles.  group/mask are the callback's
 * filter (reference defaults 1 and -1); world ids per ray may be NULL (world 0).  Per ray: obj = body
 * index of the closest hit or -1, fraction (1 without a hit), normal = hitNormalWorld, point =
 * hitPointWorld.  on_device != 0: every pointer is a device pointer (rays resident in HBM). */
int jbgpu_world_ray_test(jbgpu_ctx* ctx, int32_t n, const float* from3, const float* to3, int16_t group,
                         int16_t mask, const int32_t* world, int32_t* obj, float* fraction, float* normal3,
                         float* point3, int32_t on_device);
/* CollisionWorld.convexSweepTest with a ClosestConvexResultCallback for each of n queries of one convex shape
 * (CollisionWorld.java:596-646, objectQuerySingle's convex branch :392-424, ClosestConvexResultCallback
 * :766-810): cast_shape is a shape record (JBGPU_SHAPE_WORDS floats in host memory; sphere, box, capsule or a
 * hull over the world's point table) swept from from12 to to12 (12 floats per query: row-major basis, then
 * origin) against the bodies' current transforms by GjkConvexCast with allowedCcdPenetration 0.04.  Planes are
 * never hit (the reference has no caster for a StaticPlaneShape and throws, :455-485).  group/mask/world as for
 * jbgpu_world_ray_test.  Per query: obj = body index
 * of the closest hit or -1, fraction (1 without a hit), normal = hitNormalWorld (unit), point = hitPointWorld.
 * on_device != 0: every pointer but cast_shape is a device pointer. */
int jbgpu_world_convex_sweep_test(jbgpu_ctx* ctx, const float* cast_shape, int32_t n, const float* from12,
                                  const float* to12, int16_t group, int16_t mask, const int32_t* world, int32_t* obj,
                                  float* fraction, float* normal3, float* point3, int32_t on_device);
/* CollisionObject.setCcdMotionThreshold / setCcdSweptSphereRadius (CollisionObject.java:280-300) of n bodies
 * (all 0 when the world is created: no continuous collision detection).  In every step's integrateTransforms
 * (DiscreteDynamicsWorld.java:679-739) a convex, active, non-static body whose predicted step moves its origin
 * further than its threshold sweeps a sphere of its swept radius from its old to its predicted transform
 * (ClosestNotMeConvexResultCallback, :1129-1198: not itself, its own proxy filter, no body it already has contact
 * points with, no hit for motion along or away from the normal); a hit above 0.0001 of the step re-integrates
 * it over that fraction of the step.  Bodies are moved in index order, so a body's sweep sees the bodies before
 * it at their new transforms.  Thresholds and radii must be finite, radii not negative (JBGPU_ERR_ARG, nothing
 * changed). */
int jbgpu_world_set_ccd(jbgpu_ctx* ctx, int32_t n, const int32_t* bodies, const float* motion_threshold,
                        const float* swept_sphere_radius);
/* stepSimulation(dt, 0) nsteps times: exactly one substep per call.  JBGPU_ERR_CAPACITY with an "EPA workspace
 * overflow" message reports that a pair's GJK-EPA outgrew its workspace in the step just taken (that pair got no
 * contact point); the step has been applied and the world state has advanced, the following steps are not
 * affected. */
int jbgpu_world_step(jbgpu_ctx* ctx, float dt, int32_t nsteps);
/* enable != 0: jbgpu_world_step runs each step as a replayed hipGraph (no counterpart in the reference: a launch
 * form of the same stepSimulation).  Every launch of the captured step is shaped by capacities (bodies, a pair
 * capacity of 1.5x the last pair count + 4096) with the per-step counts kept on the device, so the steps of one
 * call run without host waits; results are bit-identical to the ordinary step.  A step that finds more pairs than
 * the capacity is finished on the ordinary path, the capacity raised and the step captured again.  Needs the
 * world-block solve (jbgpu_solver_set_world_blocks), deactivation disabled, no constraints, no tolerance mode,
 * no learned warm start and no phase timing (JBGPU_ERR_ARG naming what is missing). */
int jbgpu_world_set_step_graph(jbgpu_ctx* ctx, int32_t enable);
/* One step run up to and including phase p: 0 predictUnconstraintMotion, 1 AABB update +
 * calculateOverlappingPairs, 2 dispatchAllCollisionPairs, 3 solveConstraints, 4 full step. */
int jbgpu_world_step_until(jbgpu_ctx* ctx, float dt, int32_t phase);
int jbgpu_world_get_bodies(jbgpu_ctx* ctx, float* bodies);
int jbgpu_world_set_bodies(jbgpu_ctx* ctx, const float* bodies);
int jbgpu_world_num_pairs(jbgpu_ctx* ctx, int64_t* n);
/* pairs (sorted keys) with their persistent manifold records (JBGPU_MANIFOLD_WORDS each) */
int jbgpu_world_get_pairs(jbgpu_ctx* ctx, uint64_t* keys, float* manifolds);
/* Replaces the pair cache (the persistent manifolds of HashedOverlappingPairCache's pairs).  The records are
 * checked on the host first (jbgpu_check_pair_records); a rejected call changes nothing. */
int jbgpu_world_set_pairs(jbgpu_ctx* ctx, int64_t n, const uint64_t* keys, const float* manifolds);
/* Host-only check of n pair keys and manifold records for a world of nbodies bodies: keys strictly increasing
 * with 1 <= uid_lo < uid_hi <= nbodies (uid = body index + 1), point count in [0, 4], pair type in [0, 3],
 * manifold bodies equal to the key's two bodies (either order).  JBGPU_ERR_ARG with the reason in
 * jbgpu_last_error(NULL, ...) otherwise.  Needs no device. */
int jbgpu_check_pair_records(int32_t nbodies, int64_t n, const uint64_t* keys, const float* manifolds);
int jbgpu_world_stats(jbgpu_ctx* ctx, jbgpu_stats* out);
/* DiscreteDynamicsWorld.addConstraint (DiscreteDynamicsWorld.java:393-400) of n Point2PointConstraints
 * (dynamics/constraintsolver/Point2PointConstraint.java:35-69): a ball socket between pivot_a (3 floats each, in
 * body A's frame) and pivot_b (in body B's frame); body_b -1 is TypedConstraint's fixed body (static, identity
 * transform), pivot_b then a world point.  settings: per constraint (tau, damping, impulseClamp) of its
 * ConstraintSetting, or NULL for (0.3, 1, 0).  Constraint ids follow creation order from 0.  Every step solves
 * the constraints of the awake islands before the contact rows in each iteration (buildJacobian, solveConstraint;
 * SequentialImpulseConstraintSolver.java:757-763, 823-842), and they join islands
 * (DiscreteDynamicsWorld.calculateSimulationIslands :655-669).  Bodies out of range, or body A equal to body B,
 * give JBGPU_ERR_ARG and change nothing. */
int jbgpu_world_add_point2point(jbgpu_ctx* ctx, int32_t n, const int32_t* body_a, const int32_t* body_b,
                                const float* pivot_a, const float* pivot_b, const float* settings);
/* DiscreteDynamicsWorld.addConstraint of n HingeConstraints (dynamics/constraintsolver/HingeConstraint.java):
 * frame_a / frame_b are rbAFrame / rbBFrame as the constructors leave them (12 floats each: basis row-major, then
 * origin; B's third column already flipped; for the fixed body, body_b -1, B's frame is in world space), e.g. from
 * jbgpu_hinge_frames.  params per hinge: (lowerLimit, upperLimit, limitSoftness, biasFactor, relaxationFactor,
 * angularOnly, enableAngularMotor, motorTargetVelocity, maxMotorImpulse) as setLimit / setAngularOnly /
 * enableAngularMotor set them (:566-583), or NULL for a free hinge (1e30, -1e30, 0.9, 0.3, 1, 0, 0, 0, 0).
 * Solved with the other constraints (buildJacobian :240-333, solveConstraint :335-520); ids continue the
 * constraint numbering; errors as jbgpu_world_add_point2point. */
int jbgpu_world_add_hinge(jbgpu_ctx* ctx, int32_t n, const int32_t* body_a, const int32_t* body_b,
                          const float* frame_a, const float* frame_b, const float* params);
/* Host-only: the frames HingeConstraint's constructors build (HingeConstraint.java:84-238) from body A's current
 * transform tr_a (12 floats).  from_frames 0: the pivot/axis forms (rbA, rbB, pivotInA, pivotInB, axisInA,
 * axisInB), or with fixed_b (rbA, pivotInA, axisInA; pivot_b / axis_b unused); from_frames 1: the frame forms
 * (rbA, rbB, rbAFrame, rbBFrame), or with fixed_b (rbA, rbAFrame; frame_b_in unused).  Writes frame_a and
 * frame_b for jbgpu_world_add_hinge.  Needs no device. */
int jbgpu_hinge_frames(int32_t from_frames, int32_t fixed_b, const float* tr_a, const float* pivot_a,
                       const float* pivot_b, const float* axis_a, const float* axis_b, const float* frame_a_in,
                       const float* frame_b_in, float* frame_a, float* frame_b);
int jbgpu_world_num_joints(jbgpu_ctx* ctx, int32_t* n);
/* appliedImpulse of every constraint (TypedConstraint.getAppliedImpulse): the sum over the last solve that
 * processed it of its impulses on all three axes and iterations; 0 before its first solve. */
int jbgpu_world_joint_impulses(jbgpu_ctx* ctx, float* out);
/* The constraints the last step solved, in solve order (joint colour after joint colour, creation order inside
 * a colour): a SequentialImpulseConstraintSolver given this constraint order (and the manifold order of
 * jbgpu_solver_get_order) produces the same velocities and impulses.  *n = their count; at most cap written. */
int jbgpu_world_joint_order(jbgpu_ctx* ctx, int32_t* order, int32_t cap, int32_t* n);

/* ---- Dispatcher ----------------------------------------------------------------------- */
/* CollisionDispatcher.dispatchAllCollisionPairs(pairCache, dispatchInfo, dispatcher)
 * (collision/broadphase/Dispatcher.java:58, collision/dispatch/CollisionDispatcher.java:252-257) on the caller's
 * pair cache: n pairs (sorted keys uid_lo << 32 | uid_hi with uid = body index + 1, and their persistent
 * manifold records), the bodies as last set by jbgpu_world_set_bodies (or as the world's own steps left them).
 * Every pair is run through its collision algorithm (sphere-sphere, convex-plane, GJK/EPA; by shape types) and
 * its manifold refreshed and updated (ManifoldResult.addContactPoint / refreshContactPoints); the updated
 * records are written to manifolds_out and become the context's pair cache (jbgpu_world_get_pairs).  Records
 * are checked first (jbgpu_check_pair_records); a record with points must carry the dispatcher's header (body0
 * the lower uid, except a convex-plane pair whose lower uid is the plane, which lists the convex body first;
 * type as JBGPU_PAIR_*), a record without points gets that header written.  A rejected call changes
 * nothing.  manifolds_out may be manifolds_in. */
int jbgpu_np_dispatch(jbgpu_ctx* ctx, int64_t n, const uint64_t* keys, const float* manifolds_in,
                      float* manifolds_out);

/* EPA workspace tiers (no reference counterpart: the reference's GjkEpaSolver grows its lists): every EPA job
 * runs first with a small LDS workspace, a job that outgrows it reruns with a large LDS one, and one that
 * outgrows that too with a global-memory one (16,384 faces).  These set the face capacities the two LDS tiers
 * use (defaults and maxima 192 and 1024; at least 16): they move work between the tiers and never change a
 * result.  JBGPU_ERR_ARG outside the ranges. */
int jbgpu_np_set_epa_capacity(jbgpu_ctx* ctx, int32_t small_faces, int32_t large_faces);

/* The device radix sort the broadphase, the pair cache, the dispatcher's bins and the colour order use (no
 * reference counterpart: the reference sorts on the host with its own comparators), on caller host arrays, for
 * its tests: n keys of key_bytes (4 or 8) bytes, sorted in place by bits [bit_lo, bit_hi) as unsigned words,
 * stably (equal digits keep their order; bits outside the range are carried along), with the n int values
 * (NULL: none) moved alongside.  JBGPU_ERR_ARG for n outside [0, 2^28], a bad key size or bit range. */
int jbgpu_prim_sort(jbgpu_ctx* ctx, void* keys, int32_t key_bytes, int32_t* values, int64_t n, int32_t bit_lo,
                    int32_t bit_hi);

/* ---- ConstraintSolver -------------------------------------------------------------- */
/* The solver seam as DiscreteDynamicsWorld.solveConstraints drives it (dynamics/DiscreteDynamicsWorld.java:
 * 636-643): prepareSolve(numBodies, numManifolds) once (ConstraintSolver.java:41), solveGroup once per awake
 * island (ConstraintSolver.java:46), allSolved once (ConstraintSolver.java:49).  Here solveGroup only records the
 * island's manifolds (pair indices into the context's pair cache); allSolved colours and solves every recorded
 * manifold in one pass (islands share no dynamic body, so solving them together in colour order equals solving
 * each alone in that order) and leaves the velocities in the bodies (jbgpu_world_get_bodies) and the impulses in
 * the manifolds (jbgpu_world_get_pairs).  jbgpu_solver_get_order then gives the manifold order used. */
int jbgpu_solver_prepare(jbgpu_ctx* ctx, int32_t num_bodies, int64_t num_manifolds);
int jbgpu_solver_add_island(jbgpu_ctx* ctx, int64_t n, const int32_t* pair_index);
int jbgpu_solver_all_solved(jbgpu_ctx* ctx, float dt);
/* Manifold order of the last solve: pair indices, colour batch after colour batch; a
 * SequentialImpulseConstraintSolver fed this manifold order without shuffling produces
 * the same velocities and impulses.  offsets[k]..offsets[k+1] is colour k (may be NULL). */
int jbgpu_solver_get_order(jbgpu_ctx* ctx, int32_t* order, int64_t cap, int64_t* n, int64_t* offsets,
                           int32_t* ncolors);
/* Learned warm start: per solver contact slot (active manifold k, point j -> slot 4k+j) the
 * initial (normal, lateral1, lateral2) impulses, used instead of the 0.85 * cached impulses
 * in the next solve; NULL turns it off. */
int jbgpu_solver_set_warmstart(jbgpu_ctx* ctx, const float* init, int64_t nslots, int32_t on_device);
/* Per-contact 16-float feature vectors (normal, relative velocity, relative angular velocity,
 * depth, inverse masses, cached impulses, friction) of the last solve, by contact slot. */
int jbgpu_solver_enable_features(jbgpu_ctx* ctx, int32_t on);
int jbgpu_solver_get_features(jbgpu_ctx* ctx, float* out, int64_t nslots);
/* Final (normal, lateral1, lateral2) impulses of the last solve by contact slot, 3 floats per
 * slot, zeros for unused slots.  The manifold keeps only part of them: the reference stores the
 * second lateral impulse into appliedImpulseLateral1 and never sets appliedImpulseLateral2
 * (dynamics/constraintsolver/SequentialImpulseConstraintSolver.java:908-909). */
int jbgpu_solver_get_impulses(jbgpu_ctx* ctx, float* out, int64_t nslots);
/* tolerance > 0: iterate until the residual ||dlambda|| / max(1, ||lambda||) <= tolerance or
 * max_iterations; tolerance 0: exactly num_iterations (ContactSolverInfo.numIterations). */
int jbgpu_solver_set_params(jbgpu_ctx* ctx, float tolerance, int32_t max_iterations, int32_t num_iterations);
/* Launch shaping of the sweeps (no reference counterpart; results are identical for every value):
 * dynamic bodies are grouped into blocks of `block_size` consecutive Morton positions (at most 2048);
 * manifolds inside a block are swept by one workgroup per block with the block's velocities in LDS,
 * colour after colour, the others by one launch per colour.  0 (the default) disables blocks. */
int jbgpu_solver_set_blocks(jbgpu_ctx* ctx, int32_t block_size);
/* Batched worlds (no reference counterpart; results are identical for every value): with bodies_per_block > 0
 * every fixed-iteration solve runs one workgroup of `threads` threads (64, 128, 256 or 512) per block of whole
 * worlds (consecutive world ids, at most bodies_per_block dynamic bodies, in [16, 1024]; a larger world forms a
 * block of its own, up to 1024 dynamic bodies): the workgroup colours the block's manifolds, applies the warm
 * start and runs every iteration with the block's velocities in LDS, so a solve is one launch.  The manifold
 * order (jbgpu_solver_get_order) is then block after block, colour after colour inside a block, reported as one
 * batch.  0 (the default) solves with one launch per colour batch.  Tolerance-driven solves (tolerance > 0)
 * always use the per-colour launches.  pipelined != 0: while a colour is swept, the rows of the next colour are
 * already being loaded (more registers per thread, fewer waves per SIMD). */
int jbgpu_solver_set_world_blocks(jbgpu_ctx* ctx, int32_t bodies_per_block, int32_t threads, int32_t pipelined);
/* Colouring of the manifolds into batches (no reference counterpart): Jones-Plassmann rounds with hashed
 * priorities (the default) or priorities largest degree first (degree_first != 0; a manifold's degree is the
 * number of manifolds on its dynamic bodies), then recolour_passes iterated-greedy passes (classes from the last
 * colour to the first, each manifold moved to the smallest colour free at its bodies; 0 by default).  Fewer
 * colours would mean fewer sweep launches.  The order changes (jbgpu_solver_get_order reports it); every order is
 * a valid sequential order of the reference solver. */
int jbgpu_solver_set_colouring(jbgpu_ctx* ctx, int32_t degree_first, int32_t recolour_passes);
/* ContactSolverInfo.warmstartingFactor for the following solves (0.85 by default). */
int jbgpu_solver_set_warmstart_factor(jbgpu_ctx* ctx, float factor);
/* Launch shaping of the colour sweeps (no reference counterpart; results are identical for every
 * value): trailing colours of at most `threshold` manifolds are swept by one single-workgroup launch,
 * colour after colour; 0 gives every colour its own launch.  Default 1024. */
int jbgpu_solver_set_tail(jbgpu_ctx* ctx, int32_t threshold);
/* Residual after each iteration of the last solve (in fixed-iteration mode only the last one). */
int jbgpu_solver_residual_history(jbgpu_ctx* ctx, float* out, int32_t cap, int32_t* n);

/* ---- learned warm start ------------------------------------------------------------------ */
/* A 16-64-64-64-3 MLP (ReLU) maps the 16 contact features above to the initial (normal,
 * lateral1, lateral2) impulses of that contact; bf16 MFMA operands, fp32 accumulation.  No
 * counterpart exists in the reference: it replaces the constant warm-start factor applied in
 * SequentialImpulseConstraintSolver.solveGroupCacheFriendlySetup (dynamics/constraintsolver/
 * SequentialImpulseConstraintSolver.java:659-747).
 * params: JBGPU_MLP_PARAMS floats, W1[64][16] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
 * W4[3][64] b4[3]; norm: mean_in[16] inv_std_in[16] std_out[3] mean_out[3]. */
#define JBGPU_MLP_PARAMS 9603
#define JBGPU_MLP_NORM_WORDS 38
int jbgpu_mlp_set_params(jbgpu_ctx* ctx, const float* params, const float* norm);
int jbgpu_mlp_get_params(jbgpu_ctx* ctx, float* params, float* norm);
/* y[n][3] = model(x[n][16]) (host arrays). */
int jbgpu_mlp_infer(jbgpu_ctx* ctx, const float* x, int64_t n, float* y);
/* jbgpu_mlp_infer without any device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The features and predictions are staged through a fixed buffer the library keeps on every device
 * it runs on until the process ends (JBGPU_MLP_STATIC_MAX contacts, 4.98 MB); at most JBGPU_MLP_STATIC_MAX
 * contacts per call (JBGPU_ERR_CAPACITY beyond).  Calls that overlap on one GPU (several threads or contexts)
 * race on that buffer.  Same results as jbgpu_mlp_infer. */
#define JBGPU_MLP_STATIC_MAX 65536
int jbgpu_mlp_infer_static(jbgpu_ctx* ctx, const float* x, int64_t n, float* y);
/* Gradient (JBGPU_MLP_PARAMS floats) of mean((y' - t')^2) over the n x 3 normalised outputs. */
int jbgpu_mlp_grad(jbgpu_ctx* ctx, const float* x, const float* t, int64_t n, float* grad, float* loss);
/* Adam training from He-uniform weights (seeded), normalisation taken from the data;
 * loss_per_epoch may be NULL. */
int jbgpu_mlp_train(jbgpu_ctx* ctx, const float* x, const float* t, int64_t n, int32_t epochs, int32_t batch,
                    float lr, uint64_t seed, float* loss_per_epoch);
/* mode 1: every solve takes the warm-start impulses of every contact from the model; mode 2: only contacts
 * without cached impulses (all three zero, e.g. points new in this step) take the model's, the others keep the
 * warm-start factor times their cached impulses as SequentialImpulseConstraintSolver does; 0: off.  Needs a
 * model. */
int jbgpu_solver_use_mlp(jbgpu_ctx* ctx, int32_t on);

/* ---- measurement ---------------------------------------------------------------------- */
/* Accumulated time (ms, HIP events recorded on the context stream around the launches) per
 * region since enabling, JBGPU_TIMER_REGIONS entries: 0 predict, 1 setAabb, 2 BVH traversal,
 * 3 GJK, 4 manifold update, 5 solver setup, 6 solver iterations, 7 integrate, 8 whole step,
 * 9 EPA, 10 pair remap, 11 MLP inference, 12 BVH build (bounds, Morton sort, build, node boxes),
 * 13 pair-key sort, 14 narrowphase binning, 15 colouring (host waits included);
 * launches[r] counts the kernel launches inside region r. */
#define JBGPU_TIMER_REGIONS 16
int jbgpu_timing_enable(jbgpu_ctx* ctx, int32_t on);
int jbgpu_timing_get(jbgpu_ctx* ctx, double* ms, int64_t* launches, int64_t* steps);
/* Kernel-level timing of the solver sweeps (independent of jbgpu_timing_enable): on every `every`-th
 * solve after this call (0 = off) each sweep launch of every `stride`-th iteration (jbgpu_timing_kernels_stride,
 * default 5) carries its own start/stop events (hipExtLaunchKernel), so the figures are kernel execution time
 * without launch gaps: summed ms, timed launches and rows processed by them.  Every iteration sweeps the same
 * colours; the events cost about 10 us per launch, which a sampled solve pays. */
int jbgpu_timing_kernels(jbgpu_ctx* ctx, int32_t every);
/* The iteration stride of that timing (at least 1; 1 = every iteration). */
int jbgpu_timing_kernels_stride(jbgpu_ctx* ctx, int32_t stride);
/* four entries each: [0] normal rows, per-colour launches, [1] friction rows, per-colour launches,
 * [2] normal rows, block launches, [3] friction rows, block launches */
/* Device memory for callers that keep their own inputs resident in HBM (e.g. AABBs for
 * jbgpu_bp_set_aabbs with on_device = 1) without linking another GPU runtime; freed by the caller. */
int jbgpu_buffer_alloc(jbgpu_ctx* ctx, int64_t bytes, void** dev);
int jbgpu_buffer_upload(jbgpu_ctx* ctx, void* dev, const void* host, int64_t bytes);
int jbgpu_buffer_free(jbgpu_ctx* ctx, void* dev);
/* Device copy of `bytes` (read + write counted) repeated `reps` times on the context stream: the
 * measured HBM copy ceiling in GB/s (no reference counterpart; a measurement aid). */
int jbgpu_measure_copy(jbgpu_ctx* ctx, int64_t bytes, int32_t reps, double* gbs);
int jbgpu_timing_kernels_get(jbgpu_ctx* ctx, double* ms2, int64_t* launches2, int64_t* rows2);

#ifdef __cplusplus
}
#endif

#endif /* JBGPU_H */
