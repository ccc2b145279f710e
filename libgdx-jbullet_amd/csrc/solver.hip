// solver.hip — graph-coloured sequential-impulse contact solver for gfx950.
//
// The row arithmetic is SequentialImpulseConstraintSolver's cache-friendly path
// (dynamics/constraintsolver/SequentialImpulseConstraintSolver.java: setup
// :418-786 with addFrictionConstraint :349-416, normal rows
// resolveSingleCollisionCombinedCacheFriendly :234-284, friction rows
// resolveSingleFrictionCacheFriendly :286-346, writeback :892-924, SolverBody
// :61-93).  Instead of the reference's Random(0)-shuffled row order the
// manifolds that carry rows are coloured (Jones-Plassmann with hashed
// priorities, smallest free colour per round) so that manifolds of one colour
// share no dynamic body; one thread handles one manifold's points in order and
// a colour is one launch.  Processing colour after colour is exactly the
// sequential solver fed the manifold order "colour 0 manifolds, colour 1
// manifolds, ..." (any order inside a colour gives the same bits); the library
// reports that order so the CPU oracle can be fed the same order.
//
// Row storage is in colour-slot order (slot p = position in the colour order):
// a 16-byte header {point mask, body A, body B, pair index} and the rows of the
// slot's four contacts, so a sweep thread issues its header and row loads in one
// round trip and the two body velocities in a second.  Setup is split into a
// fully parallel row build (rows depend only on the pre-solve state) and a
// per-colour warm-start application (impulses are applied in manifold order).
#include <cmath>


#include "internal.h"
#include <algorithm>

#include "mlp.h"
#include "shapes.h"

namespace jbg {

namespace {

constexpr int RW = kRowWords;  // words per row
constexpr int NW = 4 * RW;     // normal rows of one slot (4 points), register staging and capacity
constexpr int FW = 8 * RW;     // friction rows of one slot (4 points), register staging and capacity

struct SolverParams {
  float dt, erp, warmFactor, linearSlop;
  int warm;
  int useLearned, dumpFeatures;
  float tol;
};

__device__ __forceinline__ bool isActiveState(int a) {
  return a != JBGPU_ISLAND_SLEEPING && a != JBGPU_DISABLE_SIMULATION;
}

// manifolds that produce rows: needsCollision (one body awake) and needsResponse,
// and at least one point with distance <= 0
__device__ __forceinline__ bool isStaticBody(const float* __restrict__ props, int i) {
  return (ldi(props + 8 * i + 7) & JBGPU_BODY_STATIC) != 0;
}

// ---- simulation islands (SimulationIslandManager.java:57-110,120-240): connected components of the
// non-static bodies over the broadphase pairs, as a parent forest with parent[i] <= i
__global__ void kIslandInit(int n, int* parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[i] = i;
}
__device__ __forceinline__ int islandRoot(const int* parent, int i) {
  int r = parent[i];
  while (r != parent[r]) r = parent[r];
  return r;
}
// findUnions (:57-70): every pair of island-merging (non-static) objects joins their islands; the larger
// root is hooked under the smaller, so parents only decrease and never form cycles
__global__ void kIslandHook(int64_t np, const uint64_t* __restrict__ keys, const float* __restrict__ props,
                            int* parent, int* changed) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int a = (int)(keys[k] >> 32) - 1, b = (int)(keys[k] & 0xffffffffu) - 1;
  if (isStaticBody(props, a) || isStaticBody(props, b)) return;
  int ra = islandRoot(parent, a), rb = islandRoot(parent, b);
  if (ra == rb) return;
  atomicMin(parent + max(ra, rb), min(ra, rb));
  *changed = 1;
}
__global__ void kIslandCompress(int n, int* parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[i] = islandRoot(parent, i);
}
// buildIslands (:120-240): an island whose bodies are all neither ACTIVE_TAG nor DISABLE_DEACTIVATION
// goes to sleep; otherwise its sleeping bodies want deactivation (i.e. wake up)
__global__ void kIslandAwake(int n, const int* __restrict__ parent, const float* __restrict__ props,
                             const int* __restrict__ act, int* awake) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || isStaticBody(props, i)) return;
  if (act[i] == JBGPU_ACTIVE_TAG || act[i] == JBGPU_DISABLE_DEACTIVATION) awake[parent[i]] = 1;
}
__global__ void kIslandApply(int n, const int* __restrict__ parent, const float* __restrict__ props,
                             const int* __restrict__ awake, int* act, int* inactive) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || isStaticBody(props, i)) return;
  int a = act[i];
  const bool settable = a != JBGPU_DISABLE_DEACTIVATION && a != JBGPU_DISABLE_SIMULATION;  // setActivationState
  if (!awake[parent[i]]) {
    if (settable) a = JBGPU_ISLAND_SLEEPING;
  } else if (a == JBGPU_ISLAND_SLEEPING) {
    a = JBGPU_WANTS_DEACTIVATION;
  }
  act[i] = a;
  if (!isActiveState(a)) inactive[parent[i]] = 1;  // buildAndProcessIslands skips such an island (:283-289)
}

// manifolds solved this step: response needed, not both bodies asleep, island without inactive bodies
// (island == nullptr: deactivation disabled, every island is awake)
__global__ void kMarkActive(int64_t npairs, const int* __restrict__ dn, const float* __restrict__ man, const int* __restrict__ ptype,
                            const int* __restrict__ act, const float* __restrict__ props,
                            const int* __restrict__ island, const int* __restrict__ inactive, int* flag) {
  if (dn) npairs = *dn;  // graph-captured steps: the count on the device
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * i;
  int n = ldi(m), b0 = ldi(m + 1), b1 = ldi(m + 2);
  int f = 0;
  if (ptype[i] != JBGPU_PAIR_NONE && n > 0) {
    bool s0 = isStaticBody(props, b0), s1 = isStaticBody(props, b1);
    bool awake = (act[b0] != JBGPU_ISLAND_SLEEPING) || (act[b1] != JBGPU_ISLAND_SLEEPING);
    bool solvable = isActiveState(act[b0]) || isActiveState(act[b1]);
    if (island && (!s0 || !s1)) solvable = !inactive[island[s0 ? b1 : b0]];
    if ((!s0 || !s1) && awake && solvable) {
      for (int k = 0; k < n; k++)
        if (m[JBGPU_MANIFOLD_HDR + JBGPU_POINT_WORDS * k + 15] <= 0.0f) f = 1;
    }
  }
  flag[i] = f;
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ bool dynamicBody(const float* props, int i) {
  return (ldi(props + 8 * i + 7) & JBGPU_BODY_STATIC) == 0;
}
// round (bits 52-63), then a 20-bit hash of the manifold's index, then the index (ties); colouring
// largest-degree-first, the manifold's degree (manifolds sharing one of its dynamic bodies) takes bits 40-51 and
// the hash keeps 8 bits (a hash with only 8 bits leaves long runs of index-ordered priorities, which colour worse)
__device__ __forceinline__ unsigned long long prio(int round, int k, const int* __restrict__ deg) {
  if (!deg)
    return ((unsigned long long)round << 52) | ((unsigned long long)(hash32(k) >> 12) << 32) | (unsigned long long)k;
  const unsigned long long d = (unsigned long long)min(deg[k], 4095);
  return ((unsigned long long)round << 52) | (d << 40) | ((unsigned long long)(hash32(k) >> 24) << 32) |
         (unsigned long long)k;
}

// largest-degree-first priorities: manifolds per dynamic body, then per manifold the sum over its bodies
__global__ void kBodyDegree(int n, const int2* __restrict__ ab, const float* __restrict__ props, int* bodyDeg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int2 b = ab[k];
  if (dynamicBody(props, b.x)) atomicAdd(bodyDeg + b.x, 1);
  if (dynamicBody(props, b.y)) atomicAdd(bodyDeg + b.y, 1);
}
__global__ void kManifoldDegree(int n, const int2* __restrict__ ab, const float* __restrict__ props,
                                const int* __restrict__ bodyDeg, int* deg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int2 b = ab[k];
  deg[k] = (dynamicBody(props, b.x) ? bodyDeg[b.x] : 0) + (dynamicBody(props, b.y) ? bodyDeg[b.y] : 0);
}

// body pair of active manifold k
__global__ void kActiveBodies(int n, const int* __restrict__ dn, const int* __restrict__ list, const float* __restrict__ man, int2* ab) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * list[k];
  ab[k] = make_int2(ldi(m + 1), ldi(m + 2));
}

// Jones-Plassmann rounds over the still uncoloured manifolds pend[0, m): a manifold wins a round when its
// priority is the largest among the uncoloured manifolds of both its dynamic bodies, and then takes the
// smallest colour free at both.  Priorities carry the round in their top bits, so a later round's proposal
// always replaces an earlier one under atomicMax and the body maxima need no reset between rounds.  The body
// maxima are double-buffered by round parity (bodyPrio[(round & 1) * nb + body]), so that one launch can settle
// round r (reading buffer r & 1, complete since the previous launch) and propose round r + 1 (into the other
// buffer) for the manifolds still uncoloured: one launch per round instead of two, the same colouring.
__device__ __forceinline__ void colourPropose(int k, int round, int2 b, const float* __restrict__ props, int nb,
                                              unsigned long long* bodyPrio, const int* __restrict__ deg) {
  unsigned long long* bp = bodyPrio + (size_t)(round & 1) * nb;
  const unsigned long long p = prio(round, k, deg);
  if (dynamicBody(props, b.x)) atomicMax(bp + b.x, p);
  if (dynamicBody(props, b.y)) atomicMax(bp + b.y, p);
}

__global__ void kColorPropose(int m, const int* __restrict__ pend, int round, const int2* __restrict__ ab,
                              const float* __restrict__ props, const int* __restrict__ color, int nb,
                              unsigned long long* bodyPrio, const int* __restrict__ deg) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int k = pend[t];
  if (color[k] >= 0) return;
  colourPropose(k, round, ab[k], props, nb, bodyPrio, deg);
}

// settle round `round`; with proposeNext the losers propose round + 1
__global__ void kColorRound(int m, const int* __restrict__ pend, int round, bool proposeNext,
                            const int2* __restrict__ ab, const float* __restrict__ props, int* color, int nb,
                            unsigned long long* bodyPrio, unsigned long long* bodyMask, const int* __restrict__ deg) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int k = pend[t];
  if (color[k] >= 0) return;
  const int2 b = ab[k];
  const unsigned long long p = prio(round, k, deg);
  const unsigned long long* bp = bodyPrio + (size_t)(round & 1) * nb;
  const bool da = dynamicBody(props, b.x), db = dynamicBody(props, b.y);
  const bool win = (!da || bp[b.x] == p) && (!db || bp[b.y] == p);
  if (win) {  // the round's winners share no dynamic body, so their mask updates do not race
    unsigned long long used = (da ? bodyMask[b.x] : 0ull) | (db ? bodyMask[b.y] : 0ull);
    int c;
    if (used == ~0ull) {
      c = 64 + round;
    } else {
      c = __ffsll((long long)~used) - 1;
      if (da) bodyMask[b.x] |= (1ull << c);
      if (db) bodyMask[b.y] |= (1ull << c);
    }
    color[k] = c;
  } else if (proposeNext) {
    colourPropose(k, round + 1, b, props, nb, bodyPrio, deg);
  }
}

// one colour class of the iterated-greedy pass: slots [off, off + cnt) of the colour-sorted order
__global__ void kRecolour(int off, int cnt, const int* __restrict__ order, const int2* __restrict__ ab,
                          const float* __restrict__ props, int* color, unsigned long long* bodyMask) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int k = order[off + t];
  const int c = color[k];
  if (c >= 64) return;  // colours past the masks stay where they are
  const int2 b = ab[k];
  const bool da = dynamicBody(props, b.x), db = dynamicBody(props, b.y);
  const unsigned long long bit = 1ull << c;
  const unsigned long long used = ((da ? bodyMask[b.x] : 0ull) | (db ? bodyMask[b.y] : 0ull)) & ~bit;
  const int nc = __ffsll((long long)~used) - 1;
  if (nc >= c) return;
  color[k] = nc;
  if (da) bodyMask[b.x] = (bodyMask[b.x] & ~bit) | (1ull << nc);
  if (db) bodyMask[b.y] = (bodyMask[b.y] & ~bit) | (1ull << nc);
}

__global__ void kUncoloured(int m, const int* __restrict__ pend, const int* __restrict__ color, int* flag) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) flag[t] = color[pend[t]] < 0;
}

__global__ void kIota(int n, int* out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = k;
}

// first slot of every colour in the colour-sorted key array
__global__ void kColorStarts(int n, const int* __restrict__ sorted, int* starts) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (k == 0 || sorted[k] != sorted[k - 1]) starts[sorted[k]] = k;
}

__device__ __forceinline__ M3 invInertiaWorld(const float* tr, const float* props, int i) {  // RigidBody.java:449-459
  M3 b;
  for (int k = 0; k < 9; k++) b.m[k] = tr[12 * i + k];
  V3 il = v3(props[8 * i + 1], props[8 * i + 2], props[8 * i + 3]);
  M3 m1;
  for (int r = 0; r < 3; r++) {
    m1.m[r * 3 + 0] = b.m[r * 3 + 0] * il.x;
    m1.m[r * 3 + 1] = b.m[r * 3 + 1] * il.y;
    m1.m[r * 3 + 2] = b.m[r * 3 + 2] * il.z;
  }
  return mulMM(m1, transpose(b));
}

// slot header: x = valid-point mask (bits 0-3) | first dense row << 4, y / z = bodies, w = pair index
__device__ __forceinline__ int slotMask(const int4& h) { return h.x & 15; }
__device__ __forceinline__ int slotRow(const int4& h) { return (int)((unsigned)h.x >> 4); }

struct SBody {
  V3 lv, av;
  float invMass;
};
__device__ __forceinline__ void applyImpulse(SBody& b, V3 lc, V3 ac, float mag) {  // SolverBody.java:61-72
  if (b.invMass != 0.0f) {
    b.lv.x += mag * 1.0f * lc.x;
    b.lv.y += mag * 1.0f * lc.y;
    b.lv.z += mag * 1.0f * lc.z;
    b.av.x += mag * 1.0f * ac.x;
    b.av.y += mag * 1.0f * ac.y;
    b.av.z += mag * 1.0f * ac.z;
  }
}
// solver body record svel[i][8]: linear velocity, inverse mass, angular velocity, 0 (one 32-byte read per body)
__device__ __forceinline__ SBody loadBody(const float* __restrict__ v, const float* __restrict__ props, int i) {
  (void)props;
  const float4* p = (const float4*)(v + 8 * i);
  float4 a = p[0], b = p[1];
  SBody s;
  s.lv = v3(a.x, a.y, a.z);
  s.av = v3(b.x, b.y, b.z);
  s.invMass = a.w;
  return s;
}
__device__ __forceinline__ void storeBody(float* v, int i, const SBody& s) {
  float4* p = (float4*)(v + 8 * i);
  p[0] = make_float4(s.lv.x, s.lv.y, s.lv.z, s.invMass);
  p[1] = make_float4(s.av.x, s.av.y, s.av.z, 0.0f);
}
// the solver body records of a solve: velocities of the step with each body's inverse mass beside them
__global__ void kSolverBodies(int n, const float* __restrict__ vel, const float* __restrict__ props, float* svel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4* s = (const float4*)(vel + 8 * i);
  float4 a = s[0], b = s[1];
  a.w = props[8 * i];
  ((float4*)(svel + 8 * i))[0] = a;
  ((float4*)(svel + 8 * i))[1] = b;
}

__device__ void frictionRow(float* R, V3 axis, V3 rp1, V3 rp2, const M3& I0, const M3& I1, float im0, float im1,
                            float friction) {  // addFrictionConstraint (:349-416)
  st3(R + 0, axis);
  V3 t = crs(rp1, axis);
  st3(R + 3, t);
  V3 angA = mulMV(I0, t);
  st3(R + 9, angA);
  t = crs(rp2, axis);
  st3(R + 6, t);
  V3 angB = mulMV(I1, t);
  st3(R + 12, angB);
  float d0 = im0 + dot(axis, crs(angA, rp1));
  float d1 = im1 + dot(axis, crs(angB, rp2));
  R[15] = 1.0f / (d0 + d1);
  R[16] = 0.0f;
  R[17] = 0.0f;
  R[18] = friction;
  R[19] = 0.0f;
}

__device__ __forceinline__ void planeSpace1(V3 n, V3& p, V3& q) {  // TransformUtil.java:45-61
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

// rows of every slot (solveGroupCacheFriendlySetup :494-747 without the impulse application);
// the warm-start impulse of each row is left in its applied word
// contact rows of each slot: points with distance <= 0 (the rows solveGroupCacheFriendlySetup creates)
__global__ void kSlotRows(int na, const int* __restrict__ dn, const int* __restrict__ corder, const int* __restrict__ list,
                          const float* __restrict__ man, int* count) {
  if (dn) na = *dn;  // graph-captured steps: the count on the device
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= na) return;
  const float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * list[corder[p]];
  const int n = ldi(m);
  int k = 0;
  for (int j = 0; j < n; j++)
    if (m[JBGPU_MANIFOLD_HDR + JBGPU_POINT_WORDS * j + 15] <= 0.0f) k++;
  count[p] = k;
}

// a 20-word row from registers to its 16-byte-aligned place (rows are 80 B)
__device__ __forceinline__ void storeRow(float* dst, const float* r) {
#pragma unroll
  for (int k = 0; k < RW / 4; k++)
    reinterpret_cast<float4*>(dst)[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
}

// rows of every slot, stored densely from the slot's first row (rowStart, exclusive prefix sum of the
// slot row counts): normal row r at nrow[r], its two friction rows at frow[2r], frow[2r+1]
// (solveGroupCacheFriendlySetup :494-747 without the impulse application); the warm-start impulse of each
// row is left in its applied word
__global__ void kBuildRows(int na, const int* __restrict__ dn, const int* __restrict__ corder, const int* __restrict__ list, SolverParams SP,
                           float* man, const float* __restrict__ tr, const float* __restrict__ vel,
                           const float* __restrict__ props, const int* __restrict__ rowStart, int4* hdr, float* nrow,
                           float* frow, const float* __restrict__ learned, float* features) {
  if (dn) na = *dn;  // graph-captured steps: the count on the device
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= na) return;
  int pi = list[corder[p]];
  float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * pi;
  int n = ldi(m), b0 = ldi(m + 1), b1 = ldi(m + 2);
  const int row0 = rowStart[p];
  int row = row0;
  M3 I0 = invInertiaWorld(tr, props, b0), I1 = invInertiaWorld(tr, props, b1);
  V3 o0 = ld3(tr + 12 * b0 + 9), o1 = ld3(tr + 12 * b1 + 9);
  V3 lv0 = ld3(vel + 8 * b0), av0 = ld3(vel + 8 * b0 + 4), lv1 = ld3(vel + 8 * b1), av1 = ld3(vel + 8 * b1 + 4);
  float im0 = props[8 * b0], im1 = props[8 * b1];
  int mask = 0;
  for (int j = 0; j < n; j++) {
    // the point in 16-byte words: [1] localB.yz, worldA.xy; [2] worldA.z, worldB; [3] normal, distance;
    // [4] friction, restitution, impulse, lateral impulse 1; [5] lateral impulse 2, lateral direction 1;
    // [6] lateral direction 2, the directions' flag
    float4* P = reinterpret_cast<float4*>(m + JBGPU_MANIFOLD_HDR + JBGPU_POINT_WORDS * j);
    const float4 q3 = P[3];
    if (!(q3.w <= 0.0f)) continue;
    const float4 q1 = P[1], q2 = P[2], q4 = P[4], q5 = P[5], q6 = P[6];
    mask |= 1 << j;
    float N[RW];  // the row is assembled in registers and stored as five 16-byte words (scalar stores of
                  // neighbouring slots' rows, 80 B apart, would touch a cache line per lane each)
    V3 normal = v3(q3.x, q3.y, q3.z);
    V3 rp1 = sub(v3(q1.z, q1.w, q2.x), o0);
    V3 rp2 = sub(v3(q2.y, q2.z, q2.w), o1);
    V3 angA = mulMV(I0, crs(rp1, normal));
    V3 angB = mulMV(I1, crs(rp2, normal));
    float d0 = im0 + dot(normal, crs(angA, rp1));
    float d1 = im1 + dot(normal, crs(angB, rp2));
    st3(N + 0, normal);
    st3(N + 3, crs(rp1, normal));
    st3(N + 6, crs(rp2, normal));
    st3(N + 9, angA);
    st3(N + 12, angB);
    N[15] = 1.0f / (d0 + d1);
    V3 vel1 = add(crs(av0, rp1), lv0);
    V3 vel2 = add(crs(av1, rp2), lv1);
    V3 vv = sub(vel1, vel2);
    float rel_vel = dot(normal, vv);
    float pen = jmin(q3.w + SP.linearSlop, 0.0f);
    float friction = q4.x;
    float rest = q4.y * -rel_vel;
    if (rest <= 0.0f) rest = 0.0f;
    float penVel = -pen / SP.dt;
    if (rest > penVel) pen = 0.0f;
    N[16] = pen;
    N[17] = rest;
    N[18] = friction;
    const int slot = p * 4 + j;
    if (SP.dumpFeatures) {
      float* F = features + (size_t)16 * slot;
      F[0] = normal.x; F[1] = normal.y; F[2] = normal.z;
      F[3] = vv.x; F[4] = vv.y; F[5] = vv.z;
      F[6] = av0.x - av1.x; F[7] = av0.y - av1.y; F[8] = av0.z - av1.z;
      F[9] = q3.w; F[10] = im0; F[11] = im1;
      F[12] = q4.z; F[13] = q4.w; F[14] = q5.x; F[15] = friction;
    }
    N[19] = SP.useLearned ? learned[3 * slot] : (SP.warm ? q4.z * SP.warmFactor : 0.0f);
    storeRow(nrow + (size_t)RW * row, N);
    // lateral friction directions (:678-708)
    V3 dir1 = v3(q5.y, q5.z, q5.w), dir2 = v3(q6.x, q6.y, q6.z);
    if (__builtin_bit_cast(int, q6.w) == 0) {
      V3 d1v = sub(vv, scl(normal, rel_vel));
      float lat = len2(d1v);
      V3 d2v;
      if (lat > kFltEps) {
        d1v = scl(d1v, 1.0f / jsqrt(lat));
        d2v = nor(crs(d1v, normal));
      } else {
        planeSpace1(normal, d1v, d2v);
      }
      P[5] = make_float4(q5.x, d1v.x, d1v.y, d1v.z);
      P[6] = make_float4(d2v.x, d2v.y, d2v.z, __builtin_bit_cast(float, 1));
      dir1 = d1v;
      dir2 = d2v;
    }
    float F1[RW], F2[RW];
    frictionRow(F1, dir1, rp1, rp2, I0, I1, im0, im1, friction);
    frictionRow(F2, dir2, rp1, rp2, I0, I1, im0, im1, friction);
    if (SP.useLearned) {
      F1[19] = learned[3 * slot + 1];
      F2[19] = learned[3 * slot + 2];
    } else if (SP.warm) {
      F1[19] = q4.w * SP.warmFactor;
      F2[19] = q5.x * SP.warmFactor;
    }
    storeRow(frow + (size_t)RW * (2 * row), F1);
    storeRow(frow + (size_t)RW * (2 * row + 1), F2);
    row++;
  }
  hdr[p] = make_int4(mask | (row0 << 4), b0, b1, pi);
}

// normal rows of the solve, read off the rows' prefix sum: out[0] all slots, out[1] slots below headEnd
__global__ void kRowTotals(int na, const int* __restrict__ dn, int headEnd, const int* __restrict__ count,
                           const int* __restrict__ start, unsigned long long* out) {
  if (dn) na = headEnd = *dn;  // graph-captured steps (world blocks: one batch)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long all = na > 0 ? (unsigned long long)(start[na - 1] + count[na - 1]) : 0ull;
  out[0] = all;
  out[1] = headEnd >= na ? all : (unsigned long long)start[max(headEnd, 0)];
}

// warm-start impulses of colour c in row order: per point normal, lateral 1, lateral 2 (:659-747)
// warm start of slot p: the applied impulses of its rows onto its two bodies (warm-start part of
// solveGroupCacheFriendlySetup, point after point, normal row then the two friction rows)
__device__ __forceinline__ void warmApplySlot(int p, const int4* __restrict__ hdr, const float* __restrict__ nrow,
                                              const float* __restrict__ frow, const float* __restrict__ props,
                                              float* svel) {
  int4 h = hdr[p];
  SBody A = loadBody(svel, props, h.y), B = loadBody(svel, props, h.z);
  const int rows = __popc(slotMask(h)), row0 = slotRow(h);
  for (int q = 0; q < rows; q++) {
    for (int r = 0; r < 3; r++) {
      const float* R = r == 0 ? nrow + (size_t)RW * (row0 + q) : frow + (size_t)RW * (2 * (row0 + q) + r - 1);
      V3 nrm = ld3(R);
      float ap = R[19];
      applyImpulse(A, scl(nrm, A.invMass), ld3(R + 9), ap);
      applyImpulse(B, scl(nrm, B.invMass), ld3(R + 12), -ap);
    }
  }
  if (A.invMass != 0.0f) storeBody(svel, h.y, A);
  if (B.invMass != 0.0f) storeBody(svel, h.z, B);
}

__global__ __launch_bounds__(64) void kWarmApply(int off, int cnt, const int4* __restrict__ hdr,
                                                 const float* __restrict__ nrow, const float* __restrict__ frow,
                                                 const float* __restrict__ props, float* svel) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < cnt) warmApplySlot(off + k, hdr, nrow, frow, props, svel);
}

// The warm start gathered per body in one launch.  A colour batch's slots share no dynamic body, so a body has at
// most one slot per batch: slotOf[k * nb + body] (k the batch, -1 for none) lists them, and each body applies its
// slots' rows batch after batch, point after point, normal row then the two friction rows: the same additions in
// the same order as the batch-by-batch launches give its velocity.
__global__ void kWarmSlotTable(int na, int nBatches, const int* __restrict__ colorOff, const int4* __restrict__ hdr,
                               const float* __restrict__ svel, int nb, int* slotOf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= na) return;
  int lo = 0, hi = nBatches;  // the batch of slot p: colorOff[lo] <= p < colorOff[lo + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (colorOff[mid] <= p) lo = mid; else hi = mid;
  }
  const int4 h = hdr[p];
  if (svel[8 * (size_t)h.y + 3] != 0.0f) slotOf[(size_t)lo * nb + h.y] = p;
  if (svel[8 * (size_t)h.z + 3] != 0.0f) slotOf[(size_t)lo * nb + h.z] = p;
}
__global__ void kWarmGather(int nb, int nBatches, const int* __restrict__ slotOf, const int4* __restrict__ hdr,
                            const float* __restrict__ nrow, const float* __restrict__ frow,
                            const float* __restrict__ props, float* svel) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nb) return;
  SBody X = loadBody(svel, props, x);
  if (X.invMass == 0.0f) return;
  bool any = false;
  // eight batches at a time: their table words, then the headers of the slots found, loaded together
  for (int k0 = 0; k0 < nBatches; k0 += 8) {
    int ps[8];
#pragma unroll
    for (int j = 0; j < 8; j++) ps[j] = k0 + j < nBatches ? slotOf[(size_t)(k0 + j) * nb + x] : -1;
    int4 hs[8];
#pragma unroll
    for (int j = 0; j < 8; j++) hs[j] = ps[j] >= 0 ? hdr[ps[j]] : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (ps[j] < 0) continue;
      any = true;
      const int4 h = hs[j];
      const bool isA = h.y == x;
      const int rows = __popc(slotMask(h)), row0 = slotRow(h);
      for (int q = 0; q < rows; q++) {
        // the point's three rows: direction, impulse and this body's angular component, loaded together
        const float* R0 = nrow + (size_t)RW * (row0 + q);
        const float* R1 = frow + (size_t)RW * (2 * (row0 + q));
        const float* R2 = R1 + RW;
        const int ao = isA ? 9 : 12;
        const V3 n0 = ld3(R0), n1 = ld3(R1), n2 = ld3(R2);
        const V3 c0 = ld3(R0 + ao), c1 = ld3(R1 + ao), c2 = ld3(R2 + ao);
        const float a0 = R0[19], a1 = R1[19], a2 = R2[19];
        applyImpulse(X, scl(n0, X.invMass), c0, isA ? a0 : -a0);
        applyImpulse(X, scl(n1, X.invMass), c1, isA ? a1 : -a1);
        applyImpulse(X, scl(n2, X.invMass), c2, isA ? a2 : -a2);
      }
    }
  }
  if (any) storeBody(svel, x, X);
}

// the small colours [c0, c1) in one workgroup, colour after colour (a barrier orders the workgroup's global
// accesses: its waves share one CU and its L1)
constexpr int kWarmTailThreads = 1024;
constexpr int kWarmGatherBatches = 256;  // the per-body warm start's table holds this many batches at most
__global__ __launch_bounds__(kWarmTailThreads) void kWarmApplyTail(int c0, int c1, const int* __restrict__ colorOff,
                                                                   const int4* __restrict__ hdr,
                                                                   const float* __restrict__ nrow,
                                                                   const float* __restrict__ frow,
                                                                   const float* __restrict__ props, float* svel) {
  for (int c = c0; c < c1; c++) {
    for (int p = colorOff[c] + (int)threadIdx.x; p < colorOff[c + 1]; p += kWarmTailThreads)
      warmApplySlot(p, hdr, nrow, frow, props, svel);
    __syncthreads();
  }
}

// residual terms of one wave into one of kResidStripes (dd, nn) accumulators (res == nullptr: not tracked)
__device__ __forceinline__ void accumulate(double dd, double nn, double* res) {
  if (!res) return;
  for (int off = 32; off > 0; off >>= 1) {
    dd += __shfl_xor(dd, off);
    nn += __shfl_xor(nn, off);
  }
  if ((threadIdx.x & 63) == 0 && (dd != 0.0 || nn != 0.0)) {
    double* r = res + 2 * (blockIdx.x & (kResidStripes - 1));
    atomicAdd(r, dd);
    atomicAdd(r + 1, nn);
  }
}

// where a sweep reads and writes the solver velocities of a slot's two bodies
struct GlobalBodies {  // the solver body array in global memory
  float* svel;
  const float* __restrict__ props;
  __device__ void load(int p, const int4& h, SBody& A, SBody& B) const {
    (void)p;
    A = loadBody(svel, props, h.y);
    B = loadBody(svel, props, h.z);
  }
  __device__ void store(int p, const int4& h, const SBody& A, const SBody& B) const {
    (void)p;
    if (A.invMass != 0.0f) storeBody(svel, h.y, A);
    if (B.invMass != 0.0f) storeBody(svel, h.z, B);
  }
};
struct LdsBodies {  // a block's dynamic bodies in LDS (8 floats each: lv, invMass, av, -), static ones in global
  float* lds;
  const int2* __restrict__ local;  // per slot: LDS positions of body A and B, -1 for a static body
  const float* __restrict__ svel;
  const float* __restrict__ props;
  __device__ static SBody get(const float* l) {
    SBody s;
    s.lv = v3(l[0], l[1], l[2]);
    s.invMass = l[3];
    s.av = v3(l[4], l[5], l[6]);
    return s;
  }
  __device__ static void put(float* l, const SBody& s) {
    l[0] = s.lv.x; l[1] = s.lv.y; l[2] = s.lv.z;
    l[4] = s.av.x; l[5] = s.av.y; l[6] = s.av.z;
  }
  __device__ void load(int p, const int4& h, SBody& A, SBody& B) const {
    const int2 q = local[p];
    A = q.x >= 0 ? get(lds + 8 * q.x) : loadBody(svel, props, h.y);
    B = q.y >= 0 ? get(lds + 8 * q.y) : loadBody(svel, props, h.z);
  }
  __device__ void store(int p, const int4& h, const SBody& A, const SBody& B) const {
    (void)h;
    const int2 q = local[p];
    if (q.x >= 0) put(lds + 8 * q.x, A);
    if (q.y >= 0) put(lds + 8 * q.y, B);
  }
};

// normal rows of one manifold slot p: resolveSingleCollisionCombinedCacheFriendly (:234-284)
template <class Bodies>
__device__ __forceinline__ void normalSlot(int p, const SolverParams& SP, const int4* __restrict__ hdr, float* nrow,
                                           const Bodies& io, double& dd, double& nn) {
  const int4 h = hdr[p];
  const int mask = slotMask(h), row0 = slotRow(h);
  float r[NW];  // the slot's rows in point order; only the valid points' rows are loaded
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (!((mask >> j) & 1)) continue;
    const float4* src = (const float4*)(nrow + (size_t)RW * (row0 + __popc(mask & ((1 << j) - 1))));
#pragma unroll
    for (int q = 0; q < RW / 4; q++) {
      float4 v = src[q];
      r[RW * j + 4 * q] = v.x; r[RW * j + 4 * q + 1] = v.y; r[RW * j + 4 * q + 2] = v.z; r[RW * j + 4 * q + 3] = v.w;
    }
  }
  SBody A, B;
  io.load(p, h, A, B);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (!((mask >> j) & 1)) continue;
    const float* N = r + RW * j;
    V3 nrm = v3(N[0], N[1], N[2]);
    float v1 = dot(nrm, A.lv) + dot(v3(N[3], N[4], N[5]), A.av);
    float v2 = dot(nrm, B.lv) + dot(v3(N[6], N[7], N[8]), B.av);
    float rel_vel = v1 - v2;
    float positionalError = -N[16] * SP.erp / SP.dt;
    float velocityError = N[17] - rel_vel;
    float pimp = positionalError * N[15];
    float vimp = velocityError * N[15];
    float ni = pimp + vimp;
    float old = N[19];
    float sum = old + ni;
    float nw = 0.0f > sum ? 0.0f : sum;
    ni = nw - old;
    applyImpulse(A, scl(nrm, A.invMass), v3(N[9], N[10], N[11]), ni);
    applyImpulse(B, scl(nrm, B.invMass), v3(N[12], N[13], N[14]), -ni);
    nrow[(size_t)RW * (row0 + __popc(mask & ((1 << j) - 1))) + 19] = nw;
    double d = (double)nw - (double)old;
    dd += d * d;
    nn += (double)nw * (double)nw;
  }
  io.store(p, h, A, B);
}

// friction rows of one manifold slot p: resolveSingleFrictionCacheFriendly (:286-346); the limit uses the
// normal impulse of the row's own contact from this iteration
template <class Bodies>
__device__ __forceinline__ void frictionSlot(int p, const SolverParams& SP, const int4* __restrict__ hdr,
                                             const float* __restrict__ nrow, float* frow, const Bodies& io,
                                             double& dd, double& nn) {
  const int4 h = hdr[p];
  const int mask = slotMask(h), row0 = slotRow(h);
  float r[FW];  // the slot's friction rows in point order; only the valid points' rows are loaded
  float napp[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (!((mask >> j) & 1)) continue;
    const int row = row0 + __popc(mask & ((1 << j) - 1));
    const float4* src = (const float4*)(frow + (size_t)RW * (2 * row));
#pragma unroll
    for (int q = 0; q < 2 * RW / 4; q++) {
      float4 v = src[q];
      r[2 * RW * j + 4 * q] = v.x; r[2 * RW * j + 4 * q + 1] = v.y; r[2 * RW * j + 4 * q + 2] = v.z;
      r[2 * RW * j + 4 * q + 3] = v.w;
    }
    napp[j] = nrow[(size_t)RW * row + 19];
  }
  SBody A, B;
  io.load(p, h, A, B);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (!((mask >> j) & 1)) continue;
    float total = napp[j] + 0.0f;  // appliedImpulse + appliedPushImpulse (0)
#pragma unroll
    for (int f = 0; f < 2; f++) {
      const float* N = r + RW * (2 * j + f);
      float old = N[19];
      float ap = old;
      float limit = total * N[18];
      if (total > 0.0f) {
        V3 nrm = v3(N[0], N[1], N[2]);
        float v1 = dot(nrm, A.lv) + dot(v3(N[3], N[4], N[5]), A.av);
        float v2 = dot(nrm, B.lv) + dot(v3(N[6], N[7], N[8]), B.av);
        float rel_vel = v1 - v2;
        float j1 = -rel_vel * N[15];
        ap = old + j1;
        if (limit < ap) {
          ap = limit;
        } else if (ap < -limit) {
          ap = -limit;
        }
        j1 = ap - old;
        applyImpulse(A, scl(nrm, A.invMass), v3(N[9], N[10], N[11]), j1);
        applyImpulse(B, scl(nrm, B.invMass), v3(N[12], N[13], N[14]), -j1);
        frow[(size_t)RW * (2 * (row0 + __popc(mask & ((1 << j) - 1))) + f) + 19] = ap;
      }
      double d = (double)ap - (double)old;
      dd += d * d;
      nn += (double)ap * (double)ap;
    }
  }
  io.store(p, h, A, B);
}

// ---- sweeps with four lanes per manifold slot --------------------------------------------------------
// Lane q of a slot's quad holds the rows of the slot's q-th valid point (rows are dense, so they are rows
// row0 + q): a quad's loads are contiguous and a lane stages one point's rows, not four.  The points are
// applied in order: in round t lane t updates the quad's copy of the two bodies, then every lane takes
// lane t's copy (a DPP quad broadcast), so each body sees exactly the sequence of a one-thread sweep.
template <int L>
__device__ __forceinline__ float quadBcast(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), L * 0x55, 0xf, 0xf, false));
}
template <int L>
__device__ __forceinline__ void quadBcast(SBody& s) {
  s.lv = v3(quadBcast<L>(s.lv.x), quadBcast<L>(s.lv.y), quadBcast<L>(s.lv.z));
  s.av = v3(quadBcast<L>(s.av.x), quadBcast<L>(s.av.y), quadBcast<L>(s.av.z));
}

// one normal row: resolveSingleCollisionCombinedCacheFriendly (:234-284)
__device__ __forceinline__ void normalRow(const float* N, const SolverParams& SP, SBody& A, SBody& B, float* out,
                                          double& dd, double& nn) {
  V3 nrm = v3(N[0], N[1], N[2]);
  float v1 = dot(nrm, A.lv) + dot(v3(N[3], N[4], N[5]), A.av);
  float v2 = dot(nrm, B.lv) + dot(v3(N[6], N[7], N[8]), B.av);
  float rel_vel = v1 - v2;
  float positionalError = -N[16] * SP.erp / SP.dt;
  float velocityError = N[17] - rel_vel;
  float pimp = positionalError * N[15];
  float vimp = velocityError * N[15];
  float ni = pimp + vimp;
  float old = N[19];
  float sum = old + ni;
  float nw = 0.0f > sum ? 0.0f : sum;
  ni = nw - old;
  applyImpulse(A, scl(nrm, A.invMass), v3(N[9], N[10], N[11]), ni);
  applyImpulse(B, scl(nrm, B.invMass), v3(N[12], N[13], N[14]), -ni);
  *out = nw;
  double d = (double)nw - (double)old;
  dd += d * d;
  nn += (double)nw * (double)nw;
}

// one friction row: resolveSingleFrictionCacheFriendly (:286-346), limit from the contact's normal impulse
__device__ __forceinline__ void frictionRow(const float* N, float total, SBody& A, SBody& B, float* out, double& dd,
                                            double& nn) {
  float old = N[19];
  float ap = old;
  float limit = total * N[18];
  if (total > 0.0f) {
    V3 nrm = v3(N[0], N[1], N[2]);
    float v1 = dot(nrm, A.lv) + dot(v3(N[3], N[4], N[5]), A.av);
    float v2 = dot(nrm, B.lv) + dot(v3(N[6], N[7], N[8]), B.av);
    float rel_vel = v1 - v2;
    float j1 = -rel_vel * N[15];
    ap = old + j1;
    if (limit < ap) {
      ap = limit;
    } else if (ap < -limit) {
      ap = -limit;
    }
    j1 = ap - old;
    applyImpulse(A, scl(nrm, A.invMass), v3(N[9], N[10], N[11]), j1);
    applyImpulse(B, scl(nrm, B.invMass), v3(N[12], N[13], N[14]), -j1);
    *out = ap;
  }
  double d = (double)ap - (double)old;
  dd += d * d;
  nn += (double)ap * (double)ap;
}

__device__ __forceinline__ void loadRow(float* r, const float* src) {
#pragma unroll
  for (int k = 0; k < RW / 4; k++) {
    const float4 v = ((const float4*)src)[k];
    r[4 * k] = v.x; r[4 * k + 1] = v.y; r[4 * k + 2] = v.z; r[4 * k + 3] = v.w;
  }
}

template <int L>
__device__ __forceinline__ void normalRound(int q, bool mine, const float* N, const SolverParams& SP, SBody& A,
                                            SBody& B, float* out, double& dd, double& nn) {
  if (q == L && mine) normalRow(N, SP, A, B, out, dd, nn);
  quadBcast<L>(A);
  quadBcast<L>(B);
}
// normal rows of slot p, lane q of its quad
template <class Bodies>
__device__ __forceinline__ void normalQuad(int p, int q, const SolverParams& SP, const int4* __restrict__ hdr,
                                           float* nrow, const Bodies& io, double& dd, double& nn) {
  const int4 h = hdr[p];
  const int nr = __popc(slotMask(h));
  const bool mine = q < nr;
  float* row = nrow + (size_t)RW * (slotRow(h) + q);
  float N[RW];
  if (mine) loadRow(N, row);
  SBody A, B;
  io.load(p, h, A, B);
  normalRound<0>(q, mine, N, SP, A, B, row + 19, dd, nn);
  normalRound<1>(q, mine, N, SP, A, B, row + 19, dd, nn);
  normalRound<2>(q, mine, N, SP, A, B, row + 19, dd, nn);
  normalRound<3>(q, mine, N, SP, A, B, row + 19, dd, nn);
  if (q == 0) io.store(p, h, A, B);
}

template <int L>
__device__ __forceinline__ void frictionRound(int q, bool mine, const float* F, float total, SBody& A, SBody& B,
                                              float* out, double& dd, double& nn) {
  if (q == L && mine) {
    frictionRow(F, total, A, B, out + 19, dd, nn);
    frictionRow(F + RW, total, A, B, out + RW + 19, dd, nn);
  }
  quadBcast<L>(A);
  quadBcast<L>(B);
}
// friction rows of slot p, lane q of its quad (the point's two lateral rows in order)
template <class Bodies>
__device__ __forceinline__ void frictionQuad(int p, int q, const SolverParams& SP, const int4* __restrict__ hdr,
                                             const float* __restrict__ nrow, float* frow, const Bodies& io,
                                             double& dd, double& nn) {
  (void)SP;
  const int4 h = hdr[p];
  const int nr = __popc(slotMask(h));
  const bool mine = q < nr;
  const int r = slotRow(h) + q;
  float* rows = frow + (size_t)RW * (2 * r);
  float F[2 * RW];
  float total = 0.0f;
  if (mine) {
    loadRow(F, rows);
    loadRow(F + RW, rows + RW);
    total = nrow[(size_t)RW * r + 19] + 0.0f;  // appliedImpulse + appliedPushImpulse (0)
  }
  SBody A, B;
  io.load(p, h, A, B);
  frictionRound<0>(q, mine, F, total, A, B, rows, dd, nn);
  frictionRound<1>(q, mine, F, total, A, B, rows, dd, nn);
  frictionRound<2>(q, mine, F, total, A, B, rows, dd, nn);
  frictionRound<3>(q, mine, F, total, A, B, rows, dd, nn);
  if (q == 0) io.store(p, h, A, B);
}

// one colour: slots [off, off + cnt), four lanes each
constexpr int kSweepThreads = 256;
__global__ __launch_bounds__(kSweepThreads) void kNormal(int off, int cnt, SolverParams SP,
                                                         const int4* __restrict__ hdr, float* nrow,
                                                         const float* __restrict__ props, float* svel, double* res,
                                                         const int* __restrict__ done) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 2;
  double dd = 0.0, nn = 0.0;
  const int stop = done ? *done : 0;  // tolerance mode only
  if (k < cnt && !stop) normalQuad(off + k, t & 3, SP, hdr, nrow, GlobalBodies{svel, props}, dd, nn);
  accumulate(dd, nn, res);
}
__global__ __launch_bounds__(kSweepThreads) void kFriction(int off, int cnt, SolverParams SP,
                                                           const int4* __restrict__ hdr,
                                                           const float* __restrict__ nrow, float* frow,
                                                           const float* __restrict__ props, float* svel, double* res,
                                                           const int* __restrict__ done) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 2;
  double dd = 0.0, nn = 0.0;
  const int stop = done ? *done : 0;
  if (k < cnt && !stop) frictionQuad(off + k, t & 3, SP, hdr, nrow, frow, GlobalBodies{svel, props}, dd, nn);
  accumulate(dd, nn, res);
}

// the small colours [c0, c1) of a sweep in one workgroup: colour after colour, separated by workgroup
// barriers (all waves of a workgroup share one CU and its L1, so a barrier orders their global accesses)
constexpr int kTailThreads = 1024;
__global__ __launch_bounds__(kTailThreads) void kNormalTail(int c0, int c1, const int* __restrict__ colorOff,
                                                            SolverParams SP, const int4* __restrict__ hdr,
                                                            float* nrow, const float* __restrict__ props,
                                                            float* svel, double* res, const int* __restrict__ done) {
  double dd = 0.0, nn = 0.0;
  const int stop = done ? *done : 0;
  if (!stop)
    for (int c = c0; c < c1; c++) {
      const int a = colorOff[c], b = colorOff[c + 1];
      for (int p = a + (int)(threadIdx.x >> 2); p < b; p += kTailThreads / 4)
        normalQuad(p, threadIdx.x & 3, SP, hdr, nrow, GlobalBodies{svel, props}, dd, nn);
      __syncthreads();
    }
  accumulate(dd, nn, res);
}
__global__ __launch_bounds__(kTailThreads) void kFrictionTail(int c0, int c1, const int* __restrict__ colorOff,
                                                              SolverParams SP, const int4* __restrict__ hdr,
                                                              const float* __restrict__ nrow, float* frow,
                                                              const float* __restrict__ props, float* svel,
                                                              double* res, const int* __restrict__ done) {
  double dd = 0.0, nn = 0.0;
  const int stop = done ? *done : 0;
  if (!stop)
    for (int c = c0; c < c1; c++) {
      const int a = colorOff[c], b = colorOff[c + 1];
      for (int p = a + (int)(threadIdx.x >> 2); p < b; p += kTailThreads / 4)
        frictionQuad(p, threadIdx.x & 3, SP, hdr, nrow, frow, GlobalBodies{svel, props}, dd, nn);
      __syncthreads();
    }
  accumulate(dd, nn, res);
}

// ---- block-partitioned sweeps ---------------------------------------------------------------------
// position of every body in this step's Morton order (its block is position / B), -1 for static bodies
__global__ void kBlockPos(int n, const int* __restrict__ sorted, const float* __restrict__ props, int* pos) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int body = sorted[q];
  pos[body] = isStaticBody(props, body) ? -1 : q;
}
// per active manifold: its block when its dynamic bodies share one (interior), else -1 (boundary)
__global__ void kClassify(int na, int B, const int2* __restrict__ ab, const int* __restrict__ pos, int* phase) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= na) return;
  const int2 b = ab[k];
  const int pa = pos[b.x], pb = pos[b.y];
  int blk;
  if (pa < 0) blk = pb / B;
  else if (pb < 0) blk = pa / B;
  else blk = (pa / B == pb / B) ? pa / B : -1;
  phase[k] = blk;
}
__global__ void kPassFlag(int na, const int* __restrict__ phase, int interior, int* flag) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < na) flag[k] = interior ? (phase[k] >= 0) : (phase[k] < 0);
}
// slot order key: boundary bit, colour, block (interior slots: colour-major, then block)
__global__ void kSlotKeys(int na, const int* __restrict__ phase, const int* __restrict__ color,
                          unsigned long long* key, int* idx) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= na) return;
  const int ph = phase ? phase[k] : 0;
  key[k] = (ph < 0 ? (1ull << 52) : 0ull) | ((unsigned long long)color[k] << 20) | (unsigned long long)(ph < 0 ? 0 : ph);
  idx[k] = k;
}
// first slot of every (boundary bit, colour) batch
__global__ void kBatchFlags(int na, const unsigned long long* __restrict__ key, int* flag) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < na) flag[p] = (p == 0) || ((key[p] >> 20) != (key[p - 1] >> 20));
}
// 1 for the batches made of boundary slots
// (boundary bit, colour) batches of a solve: colour indices stay below 1024 (checked on the host)
constexpr int kMaxBatches = 2048;
__global__ void kBatchBoundary(const int* __restrict__ nbatch, const int* __restrict__ batchStart,
                               const unsigned long long* __restrict__ key, int* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *nbatch) out[i] = (int)((key[batchStart[i]] >> 52) & 1ull);
}
// slot range of every (interior batch, block) run; batchStart holds the first slot of each batch
__global__ void kRunBounds(int interiorEnd, int nBatchInterior, const int* __restrict__ batchStart, int nBlocks,
                           const unsigned long long* __restrict__ key, int* runStart, int* runEnd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= interiorEnd) return;
  int lo = 0, hi = nBatchInterior;  // batch = last start <= p
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (batchStart[mid] <= p) lo = mid; else hi = mid;
  }
  const int blk = (int)(key[p] & 0xfffffull);
  const int r = lo * nBlocks + blk;
  if (p == 0 || key[p] != key[p - 1]) runStart[r] = p;
  if (p == interiorEnd - 1 || key[p] != key[p + 1]) runEnd[r] = p + 1;
}
// LDS positions of an interior slot's bodies (-1: static)
__global__ void kSlotLocal(int n, int B, const int* __restrict__ corder, const int2* __restrict__ ab,
                           const int* __restrict__ pos, int2* local) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int2 b = ab[corder[p]];
  const int pa = pos[b.x], pb = pos[b.y];
  local[p] = make_int2(pa < 0 ? -1 : pa % B, pb < 0 ? -1 : pb % B);
}

// One workgroup per block: the block's dynamic bodies are held in LDS while the workgroup sweeps the
// block's interior slots colour after colour (workgroup barriers between colours); blocks share no
// dynamic body, so the result equals sweeping the interior slots in slot order.
constexpr int kBlockThreads = 256;
template <bool Friction>
__global__ __launch_bounds__(kBlockThreads) void kSweepBlocks(int nBatches, int nBlocks, int B, int nSorted,
                                                              const int* __restrict__ sorted,
                                                              const int* __restrict__ runStart,
                                                              const int* __restrict__ runEnd, SolverParams SP,
                                                              const int4* __restrict__ hdr,
                                                              const int2* __restrict__ local, float* nrow, float* frow,
                                                              const float* __restrict__ props, float* svel,
                                                              double* res, const int* __restrict__ done) {
  extern __shared__ float lds[];
  const int b = blockIdx.x, t = threadIdx.x;
  const int q0 = b * B, q1 = min(q0 + B, nSorted);
  const int stop = done ? *done : 0;
  for (int q = q0 + t; q < q1; q += kBlockThreads) {
    const int body = sorted[q];
    const float4* v = (const float4*)(svel + 8 * body);
    float4 lv = v[0], av = v[1];  // lv.w: inverse mass
    ((float4*)lds)[2 * (q - q0)] = lv;
    ((float4*)lds)[2 * (q - q0) + 1] = av;
  }
  __syncthreads();
  double dd = 0.0, nn = 0.0;
  const LdsBodies io{lds, local, svel, props};
  if (!stop)
    for (int ci = 0; ci < nBatches; ci++) {
      const int s0 = runStart[ci * nBlocks + b];
      if (s0 >= 0) {
        const int s1 = runEnd[ci * nBlocks + b];
        for (int p = s0 + t; p < s1; p += kBlockThreads) {
          if (Friction)
            frictionSlot(p, SP, hdr, nrow, frow, io, dd, nn);
          else
            normalSlot(p, SP, hdr, nrow, io, dd, nn);
        }
      }
      __syncthreads();
    }
  for (int q = q0 + t; q < q1; q += kBlockThreads) {
    const int body = sorted[q];
    const float4 lv = ((const float4*)lds)[2 * (q - q0)];
    if (lv.w == 0.0f) continue;  // static bodies keep their (unchanged) velocities
    const float4 av = ((const float4*)lds)[2 * (q - q0) + 1];
    float4* v = (float4*)(svel + 8 * body);
    v[0] = lv;
    v[1] = make_float4(av.x, av.y, av.z, 0.0f);
  }
  accumulate(dd, nn, res);
}

// ---- batched worlds: one workgroup per block of whole worlds ----------------------------------------------
// Worlds never share a body, so a block made of whole worlds is solved on its own: the workgroup keeps the
// block's dynamic bodies in LDS, colours the block's manifolds itself (first free colour, slot order), then
// applies the warm start and runs every iteration colour after colour with workgroup barriers in between, and
// finally writes the velocities back.  The whole solve is one launch.  The order used (block after block,
// colour after colour inside a block, slot order inside a colour) is written to colorOrder; a sequential solver
// fed that manifold order gives the same bits (blocks are independent, a colour's manifolds share no dynamic
// body).
constexpr int kWbMaxColours = 128;
struct WorldLdsBodies {  // a block's dynamic bodies in LDS (lv, invMass, av, -), static bodies from svel
  float* lds;
  const int* __restrict__ local;  // per body: position in its block, -1 for static bodies
  const float* __restrict__ svel;
  const float* __restrict__ props;
  __device__ void load(int p, const int4& h, SBody& A, SBody& B) const {
    (void)p;
    const int la = local[h.y], lb = local[h.z];
    A = la >= 0 ? LdsBodies::get(lds + 8 * la) : loadBody(svel, props, h.y);
    B = lb >= 0 ? LdsBodies::get(lds + 8 * lb) : loadBody(svel, props, h.z);
  }
  __device__ void store(int p, const int4& h, const SBody& A, const SBody& B) const {
    (void)p;
    const int la = local[h.y], lb = local[h.z];
    if (la >= 0) LdsBodies::put(lds + 8 * la, A);
    if (lb >= 0) LdsBodies::put(lds + 8 * lb, B);
  }
};

// warm start of slot p onto bodies held by io (the warm-start part of solveGroupCacheFriendlySetup, :659-747)
template <class Bodies>
__device__ __forceinline__ void warmApplySlotIo(int p, const int4* __restrict__ hdr, const float* __restrict__ nrow,
                                                const float* __restrict__ frow, const Bodies& io) {
  const int4 h = hdr[p];
  SBody A, B;
  io.load(p, h, A, B);
  const int rows = __popc(slotMask(h)), row0 = slotRow(h);
  for (int q = 0; q < rows; q++) {
    for (int r = 0; r < 3; r++) {
      const float* R = r == 0 ? nrow + (size_t)RW * (row0 + q) : frow + (size_t)RW * (2 * (row0 + q) + r - 1);
      V3 nrm = ld3(R);
      float ap = R[19];
      applyImpulse(A, scl(nrm, A.invMass), ld3(R + 9), ap);
      applyImpulse(B, scl(nrm, B.invMass), ld3(R + 12), -ap);
    }
  }
  io.store(p, h, A, B);
}

// blockOf[k] for active manifold k: the block of its dynamic body (both bodies of a manifold with two dynamic
// bodies are in one world, hence in one block)
__global__ void kManifoldBlocks(int na, const int* __restrict__ dn, const int2* __restrict__ ab, const int* __restrict__ bodyBlock, int* key,
                                int* idx, int* info) {
  if (dn) na = *dn;  // graph-captured steps: the count on the device
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= na) return;
  const int2 b = ab[k];
  const int ba = bodyBlock[b.x], bb = bodyBlock[b.y];
  if (ba >= 0 && bb >= 0 && ba != bb) atomicOr(info, 4);  // a manifold across two blocks: not worlds
  key[k] = ba >= 0 ? ba : bb;
  idx[k] = k;
}
// slotStart[b] = first slot of block b in the block-sorted slot order (slotStart[nBlocks] = na)
__global__ void kBlockSlotStarts(int na, const int* __restrict__ dn, int nBlocks, const int* __restrict__ key, int* slotStart) {
  if (dn) na = *dn;  // graph-captured steps: the count on the device
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > na) return;
  const int lo = p == 0 ? -1 : key[p - 1];
  const int hi = p == na ? nBlocks - 1 : key[p];
  for (int b = lo + 1; b <= hi; b++) slotStart[b] = p;
  if (p == na && lo < nBlocks) slotStart[nBlocks] = na;
}

// Per solve position of a block: (valid-point mask | first row << 4, body A, body B, LDS positions of A and B
// packed as 16-bit fields, 0xffff for a static body).  One 16-byte load names everything a sweep step needs.
__device__ __forceinline__ int infoLocal(const int4& i, int which) {
  const int v = which ? ((unsigned)i.w >> 16) : (i.w & 0xffff);
  return v == 0xffff ? -1 : v;
}
struct WorldInfoBodies {  // bodies of a solve position: LDS positions from the packed info, static ones from svel
  float* lds;
  const float* __restrict__ svel;
  const float* __restrict__ props;
  __device__ void load(const int4& i, SBody& A, SBody& B) const {
    const int la = infoLocal(i, 0), lb = infoLocal(i, 1);
    A = la >= 0 ? LdsBodies::get(lds + 8 * la) : loadBody(svel, props, i.y);
    B = lb >= 0 ? LdsBodies::get(lds + 8 * lb) : loadBody(svel, props, i.z);
  }
  __device__ void store(const int4& i, const SBody& A, const SBody& B) const {
    const int la = infoLocal(i, 0), lb = infoLocal(i, 1);
    if (la >= 0) LdsBodies::put(lds + 8 * la, A);
    if (lb >= 0) LdsBodies::put(lds + 8 * lb, B);
  }
};
// the rows a quad lane holds for one solve position (loaded ahead of their colour's step)
struct NormalStage {
  float N[RW];
};
struct FrictionStage {
  float F[2 * RW];
  float total;
};
__device__ __forceinline__ void stageNormal(const int4& i, int q, const float* __restrict__ nrow, NormalStage& st) {
  if (q < __popc(i.x & 15)) loadRow(st.N, nrow + (size_t)RW * (((unsigned)i.x >> 4) + q));
}
__device__ __forceinline__ void stageFriction(const int4& i, int q, const float* __restrict__ nrow,
                                              const float* __restrict__ frow, FrictionStage& st) {
  if (q < __popc(i.x & 15)) {
    const int r = (int)((unsigned)i.x >> 4) + q;
    loadRow(st.F, frow + (size_t)RW * (2 * r));
    loadRow(st.F + RW, frow + (size_t)RW * (2 * r + 1));
    st.total = nrow[(size_t)RW * r + 19] + 0.0f;  // appliedImpulse + appliedPushImpulse (0)
  }
}
// normalQuad / frictionQuad on staged rows
__device__ __forceinline__ void normalQuadStaged(const int4& i, int q, const SolverParams& SP, const NormalStage& st,
                                                 float* nrow, const WorldInfoBodies& io, double& dd, double& nn) {
  const bool mine = q < __popc(i.x & 15);
  float* out = nrow + (size_t)RW * (((unsigned)i.x >> 4) + q) + 19;
  SBody A, B;
  io.load(i, A, B);
  normalRound<0>(q, mine, st.N, SP, A, B, out, dd, nn);
  normalRound<1>(q, mine, st.N, SP, A, B, out, dd, nn);
  normalRound<2>(q, mine, st.N, SP, A, B, out, dd, nn);
  normalRound<3>(q, mine, st.N, SP, A, B, out, dd, nn);
  if (q == 0) io.store(i, A, B);
}
__device__ __forceinline__ void frictionQuadStaged(const int4& i, int q, const FrictionStage& st, float* frow,
                                                   const WorldInfoBodies& io, double& dd, double& nn) {
  const bool mine = q < __popc(i.x & 15);
  float* rows = frow + (size_t)RW * (2 * (((unsigned)i.x >> 4) + q));
  SBody A, B;
  io.load(i, A, B);
  frictionRound<0>(q, mine, st.F, st.total, A, B, rows, dd, nn);
  frictionRound<1>(q, mine, st.F, st.total, A, B, rows, dd, nn);
  frictionRound<2>(q, mine, st.F, st.total, A, B, rows, dd, nn);
  frictionRound<3>(q, mine, st.F, st.total, A, B, rows, dd, nn);
  if (q == 0) io.store(i, A, B);
}

template <int T, bool Pipelined>
__global__ __launch_bounds__(T) void kSolveWorldBlocks(int iters, SolverParams SP, int warm,
                                                       const int* __restrict__ bodyList,
                                                       const int* __restrict__ bodyStart,
                                                       const int* __restrict__ bodyLocal,
                                                       const int* __restrict__ slotStart,
                                                       const int* __restrict__ slotManifold, int* slotColour,
                                                       int* solveSlot, int* colorOrder, const int4* __restrict__ hdr,
                                                       int4* sinfo, float* nrow, float* frow,
                                                       const float* __restrict__ props, const float* __restrict__ svel,
                                                       float* vel, double* res, int* info,
                                                       const int* __restrict__ abortFlag) {
  extern __shared__ float lds[];  // [nbod][8] bodies, then [nbod][2] colour masks (64-bit words)
  if (abortFlag && *abortFlag) return;  // an aborted graph-captured step changes nothing
  __shared__ int colourStart[kWbMaxColours + 1];
  __shared__ int count[kWbMaxColours];
  __shared__ int nColours;
  const int b = blockIdx.x, t = threadIdx.x;
  const int q0 = bodyStart[b], nbod = bodyStart[b + 1] - q0;
  const int s0 = slotStart[b], ns = slotStart[b + 1] - s0;
  if (ns == 0) return;  // nothing touches this block's bodies: their velocities stay as they are
  unsigned long long* masks = reinterpret_cast<unsigned long long*>(lds + 8 * nbod);
  for (int i = t; i < nbod; i += T) {
    const float4* v = reinterpret_cast<const float4*>(svel + 8 * bodyList[q0 + i]);
    reinterpret_cast<float4*>(lds)[2 * i] = v[0];  // .w: inverse mass
    reinterpret_cast<float4*>(lds)[2 * i + 1] = v[1];
    masks[2 * i] = 0ull;
    masks[2 * i + 1] = 0ull;
  }
  __syncthreads();
  // colouring: first free colour of the slot's dynamic bodies, slot after slot (one thread; the slots' body
  // positions are loaded eight at a time ahead of the dependent mask updates)
  if (t == 0) {
    for (int c = 0; c < kWbMaxColours; c++) count[c] = 0;
    int ncol = 0, overflow = 0;
    for (int i0 = 0; i0 < ns; i0 += 8) {
      int la[8], lb[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int4 h = i0 + k < ns ? hdr[s0 + i0 + k] : make_int4(0, 0, 0, 0);
        la[k] = i0 + k < ns ? bodyLocal[h.y] : -1;
        lb[k] = i0 + k < ns ? bodyLocal[h.z] : -1;
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (i0 + k >= ns) break;
        unsigned long long u0 = 0ull, u1 = 0ull;
        if (la[k] >= 0) { u0 |= masks[2 * la[k]]; u1 |= masks[2 * la[k] + 1]; }
        if (lb[k] >= 0) { u0 |= masks[2 * lb[k]]; u1 |= masks[2 * lb[k] + 1]; }
        int c;
        if (~u0) c = __ffsll((long long)~u0) - 1;
        else if (~u1) c = 64 + __ffsll((long long)~u1) - 1;
        else { c = kWbMaxColours - 1; overflow = 1; }
        const unsigned long long bit = 1ull << (c & 63);
        if (la[k] >= 0) masks[2 * la[k] + (c >> 6)] |= bit;
        if (lb[k] >= 0) masks[2 * lb[k] + (c >> 6)] |= bit;
        slotColour[s0 + i0 + k] = c;
        count[c]++;
        ncol = max(ncol, c + 1);
      }
    }
    int acc = 0;
    for (int c = 0; c < ncol; c++) {
      colourStart[c] = acc;
      acc += count[c];
    }
    colourStart[ncol] = acc;
    // stable placement: position of every slot in its colour, in slot order
    for (int c = 0; c < ncol; c++) count[c] = colourStart[c];
    for (int i = 0; i < ns; i++) {
      const int pos = count[slotColour[s0 + i]]++;
      solveSlot[s0 + pos] = s0 + i;
    }
    nColours = ncol;
    if (overflow) atomicOr(info, 2);
    atomicMax(info + 1, ncol);
  }
  __syncthreads();
  const int ncol = nColours;
  // solve positions: manifold order for the caller, packed slot info for the sweeps
  for (int j = t; j < ns; j += T) {
    const int p = solveSlot[s0 + j];
    colorOrder[s0 + j] = slotManifold[p];
    const int4 h = hdr[p];
    const int la = bodyLocal[h.y], lb = bodyLocal[h.z];
    sinfo[s0 + j] = make_int4(h.x, h.y, h.z, (la & 0xffff) | ((lb & 0xffff) << 16));
  }
  __syncthreads();
  const WorldLdsBodies io{lds, bodyLocal, svel, props};
  const WorldInfoBodies iio{lds, svel, props};
  if (warm) {
    for (int c = 0; c < ncol; c++) {
      for (int j = colourStart[c] + t; j < colourStart[c + 1]; j += T) warmApplySlotIo(solveSlot[s0 + j], hdr, nrow, frow, io);
      __syncthreads();
    }
  }
  // Sweeps, software-pipelined over the colours: while colour c is swept, the quad's slot info of colour c + 2
  // and its rows of colour c + 1 are already on their way (rows do not depend on velocities; a colour's rows are
  // written only by its own step).  Quads past the first round of a large colour load as they go.
  const int q = t & 3, g = t >> 2;
  auto infoAt = [&](int c) -> int4 {
    if (c >= ncol) return make_int4(0, 0, 0, -1);
    const int j = colourStart[c] + g;
    return j < colourStart[c + 1] ? sinfo[s0 + j] : make_int4(0, 0, 0, -1);
  };
  double dd = 0.0, nn = 0.0;
  for (int it = 0; it < iters && !Pipelined; it++) {  // one step per colour, each loading its own rows
    double dd1 = 0.0, nn1 = 0.0;
    for (int c = 0; c < ncol; c++) {
      for (int j = colourStart[c] + g; j < colourStart[c + 1]; j += T / 4) {
        const int4 ij = sinfo[s0 + j];
        NormalStage sj;
        stageNormal(ij, q, nrow, sj);
        normalQuadStaged(ij, q, SP, sj, nrow, iio, dd1, nn1);
      }
      __syncthreads();
    }
    for (int c = 0; c < ncol; c++) {
      for (int j = colourStart[c] + g; j < colourStart[c + 1]; j += T / 4) {
        const int4 ij = sinfo[s0 + j];
        FrictionStage sj;
        stageFriction(ij, q, nrow, frow, sj);
        frictionQuadStaged(ij, q, sj, frow, iio, dd1, nn1);
      }
      __syncthreads();
    }
    if (it == iters - 1) {
      dd = dd1;
      nn = nn1;
    }
  }
  for (int it = 0; it < iters && Pipelined; it++) {
    double dd1 = 0.0, nn1 = 0.0;
    {
      int4 i0 = infoAt(0), i1 = infoAt(1);
      NormalStage cur;
      stageNormal(i0, q, nrow, cur);
      for (int c = 0; c < ncol; c++) {
        NormalStage nxt;
        stageNormal(i1, q, nrow, nxt);
        const int4 i2 = infoAt(c + 2);
        if ((i0.x & 15) != 0) normalQuadStaged(i0, q, SP, cur, nrow, iio, dd1, nn1);
        for (int j = colourStart[c] + g + T / 4; j < colourStart[c + 1]; j += T / 4) {
          const int4 ij = sinfo[s0 + j];
          NormalStage sj;
          stageNormal(ij, q, nrow, sj);
          normalQuadStaged(ij, q, SP, sj, nrow, iio, dd1, nn1);
        }
        __syncthreads();
        cur = nxt;
        i0 = i1;
        i1 = i2;
      }
    }
    {
      int4 i0 = infoAt(0), i1 = infoAt(1);
      FrictionStage cur;
      stageFriction(i0, q, nrow, frow, cur);
      for (int c = 0; c < ncol; c++) {
        FrictionStage nxt;
        stageFriction(i1, q, nrow, frow, nxt);
        const int4 i2 = infoAt(c + 2);
        if ((i0.x & 15) != 0) frictionQuadStaged(i0, q, cur, frow, iio, dd1, nn1);
        for (int j = colourStart[c] + g + T / 4; j < colourStart[c + 1]; j += T / 4) {
          const int4 ij = sinfo[s0 + j];
          FrictionStage sj;
          stageFriction(ij, q, nrow, frow, sj);
          frictionQuadStaged(ij, q, sj, frow, iio, dd1, nn1);
        }
        __syncthreads();
        cur = nxt;
        i0 = i1;
        i1 = i2;
      }
    }
    if (it == iters - 1) {
      dd = dd1;
      nn = nn1;
    }
  }
  for (int i = t; i < nbod; i += T) {
    const float4 lv = reinterpret_cast<const float4*>(lds)[2 * i];
    const float4 av = reinterpret_cast<const float4*>(lds)[2 * i + 1];
    float4* d = reinterpret_cast<float4*>(vel + 8 * bodyList[q0 + i]);
    d[0] = make_float4(lv.x, lv.y, lv.z, 0.0f);
    d[1] = make_float4(av.x, av.y, av.z, 0.0f);
  }
  accumulate(dd, nn, res);
}

// one wave: sum the stripes of iteration `it` into hist[2*it..], then r = ||dl|| / max(1, ||l||)
__global__ __launch_bounds__(64) void kResidual(int it, float tol, const double* __restrict__ stripes, double* hist,
                                                int* done, float* out) {
  const int t = threadIdx.x;
  double dd = stripes[2 * t], nn = stripes[2 * t + 1];
  for (int off = 32; off > 0; off >>= 1) {
    dd += __shfl_xor(dd, off);
    nn += __shfl_xor(nn, off);
  }
  if (t != 0 || *done) return;
  hist[2 * it] = dd;
  hist[2 * it + 1] = nn;
  double nrm = sqrt(nn);
  float r = (float)(sqrt(dd) / (nrm > 1.0 ? nrm : 1.0));
  out[0] = r;
  out[1] = (float)(it + 1);
  if (tol > 0.0f && r <= tol) *done = 1;
}

__global__ void kWriteback(int n, const int* __restrict__ dn, const int4* __restrict__ hdr, const float* __restrict__ nrow,
                           const float* __restrict__ frow, float* man) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int4 h = hdr[p];
  float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * h.w;
  const int mask = slotMask(h);
  int row = slotRow(h);
  for (int j = 0; j < 4; j++) {
    if (!((mask >> j) & 1)) continue;
    float* P = m + JBGPU_MANIFOLD_HDR + JBGPU_POINT_WORDS * j;
    P[18] = nrow[(size_t)RW * row + 19];
    P[19] = frow[(size_t)RW * (2 * row) + 19];
    P[19] = frow[(size_t)RW * (2 * row + 1) + 19];  // Lateral1 written twice; Lateral2 never stored (:908-909)
    row++;
  }
}

// final (normal, lateral 1, lateral 2) impulses of every colour slot; zeros for unused slots
__global__ void kGatherImpulses(int64_t nslots, const int4* __restrict__ hdr, const float* __restrict__ nrow,
                                const float* __restrict__ frow, float* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nslots) return;
  const int64_t p = q >> 2;
  const int j = (int)(q & 3);
  const int4 h = hdr[p];
  const int mask = slotMask(h);
  const bool ok = (mask >> j) & 1;
  const size_t row = (size_t)slotRow(h) + __popc(mask & ((1 << j) - 1));
  out[3 * q] = ok ? nrow[RW * row + 19] : 0.0f;
  out[3 * q + 1] = ok ? frow[RW * (2 * row) + 19] : 0.0f;
  out[3 * q + 2] = ok ? frow[RW * (2 * row + 1) + 19] : 0.0f;
}

__global__ void kCopyVel(int n, const float* __restrict__ src, float* dst, const float* __restrict__ props) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (props[8 * i] != 0.0f) {
    const float4* s = (const float4*)(src + 8 * i);
    float4* d = (float4*)(dst + 8 * i);
    float4 a = s[0];
    a.w = 0.0f;  // the solver record's inverse mass stays out of the velocity array
    d[0] = a;
    d[1] = s[1];
  }
}


// ---- Point2PointConstraint (dynamics/constraintsolver/Point2PointConstraint.java) ---------------------------
// One thread per constraint.  The constraints are solved on the solver body records: the reference copies a
// body's solver velocities into the rigid body before each constraint and back after it (SolverBody.
// writebackVelocity / readVelocity, SequentialImpulseConstraintSolver.java:823-842), and a body without solver
// rows keeps its rigid-body velocities, which its solver record here equals.  The fixed body (bodyB == -1,
// TypedConstraint.getFixed) has the identity transform, zero velocity, zero inverse mass and inertia.
__device__ __forceinline__ int jointType(const float* R) { return ldi(R); }
__device__ __forceinline__ int jointBodyA(const float* R) { return ldi(R + 1); }
__device__ __forceinline__ int jointBodyB(const float* R) { return ldi(R + 2); }
__device__ __forceinline__ Tr bodyTr(const float* __restrict__ tr, int i) {
  Tr t;
  if (i < 0) {
    t.b = mident();
    t.o = v3(0.0f, 0.0f, 0.0f);
    return t;
  }
  for (int k = 0; k < 9; k++) t.b.m[k] = tr[12 * i + k];
  t.o = v3(tr[12 * i + 9], tr[12 * i + 10], tr[12 * i + 11]);
  return t;
}
__device__ __forceinline__ V3 unitAxis(int i) {  // new Vector3f(0, 0, 0) with coordinate i set to 1
  V3 n = v3(0.0f, 0.0f, 0.0f);
  if (i == 0) n.x = 1.0f;
  else if (i == 1) n.y = 1.0f;
  else n.z = 1.0f;
  return n;
}
// the island merge of a constraint (DiscreteDynamicsWorld.calculateSimulationIslands :655-669): two
// non-static bodies, one of them active
__global__ void kIslandHookJoints(int nj, const float* __restrict__ rec, const float* __restrict__ props,
                                  const int* __restrict__ act, int* parent, int* changed) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  const float* R = rec + kJointWords * j;
  const int a = jointBodyA(R), b = jointBodyB(R);
  if (b < 0 || isStaticBody(props, a) || isStaticBody(props, b)) return;
  if (!isActiveState(act[a]) && !isActiveState(act[b])) return;
  int ra = islandRoot(parent, a), rb = islandRoot(parent, b);
  if (ra == rb) return;
  atomicMin(parent + max(ra, rb), min(ra, rb));
  *changed = 1;
}
// constraints solved this step: the island of body A (of body B when A is static; getConstraintIslandId
// :550-556) is processed, i.e. it has no inactive body (buildAndProcessIslands :283-289); without islands
// (deactivation disabled) its non-static bodies are active
__global__ void kJointMark(int nj, const float* __restrict__ rec, const float* __restrict__ props,
                           const int* __restrict__ act, const int* __restrict__ island,
                           const int* __restrict__ inactive, int* solve) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  const float* R = rec + kJointWords * j;
  const int a = jointBodyA(R), b = jointBodyB(R);
  const bool sa = isStaticBody(props, a), sb = b < 0 || isStaticBody(props, b);
  int f = 0;
  if (!(sa && sb)) {
    if (island) f = !inactive[island[sa ? b : a]];
    else f = (sa || isActiveState(act[a])) && (sb || isActiveState(act[b]));
  }
  solve[j] = f;
}
__device__ __forceinline__ V3 column(const M3& m, int c) { return v3(m.m[c], m.m[3 + c], m.m[6 + c]); }
__device__ __forceinline__ float atan2Fast(float y, float x) {  // ScalarUtil.java:43-57
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
__device__ __forceinline__ M3 loadM3(const float* p) {
  M3 m;
  for (int k = 0; k < 9; k++) m.m[k] = p[k];
  return m;
}
// RigidBody.computeAngularImpulseDenominator (RigidBody.java:534-541): axis . (I_world^T axis)
__device__ __forceinline__ float angularDenominator(const M3& I, V3 axis) { return dot(axis, mulMtV(I, axis)); }
// buildJacobian with appliedImpulse = 0: Point2PointConstraint.java:71-107 (linear rows along the world axes),
// HingeConstraint.java:240-333 (linear rows along the pivot separation and its plane space, the limit state from
// getHingeAngle :530-549, kHinge); JacobianEntry.java:56-73.  The state keeps what solveConstraint recomputes from
// the unchanged transforms: 1 / diagonal per row, the pivots' arms, pivotAInW - pivotBInW, both bodies' world
// inverse inertia and, for a hinge, the world hinge axes
__global__ void kJointBuild(int nj, const float* __restrict__ rec, const int* __restrict__ solve,
                            const float* __restrict__ tr, const float* __restrict__ props, float* st) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj || !solve[j]) return;
  const float* R = rec + kJointWords * j;
  float* S = st + kJointStateWords * j;
  const int type = jointType(R), a = jointBodyA(R), b = jointBodyB(R);
  const Tr ta = bodyTr(tr, a), tb = bodyTr(tr, b);
  const float imA = props[8 * a], imB = b >= 0 ? props[8 * b] : 0.0f;
  const V3 ilA = v3(props[8 * a + 1], props[8 * a + 2], props[8 * a + 3]);
  const V3 ilB = b >= 0 ? v3(props[8 * b + 1], props[8 * b + 2], props[8 * b + 3]) : v3(0.0f, 0.0f, 0.0f);
  const M3 IA = invInertiaWorld(tr, props, a);
  M3 IB;
  if (b >= 0) IB = invInertiaWorld(tr, props, b);
  else
    for (int k = 0; k < 9; k++) IB.m[k] = 0.0f;
  for (int k = 0; k < 9; k++) {
    S[14 + k] = IA.m[k];
    S[23 + k] = IB.m[k];
  }
  S[3] = 0.0f;
  S[13] = 0.0f;
  const bool hinge = type == kJointHinge;
  const bool linear = !hinge || R[33] == 0.0f;  // a hinge with angularOnly has no linear rows
  if (linear) {
    // the pivots: a hinge's frame origins, a point-to-point's pivots
    const V3 pivotAInW = xform(ta, ld3(R + (hinge ? 13 : 4))), pivotBInW = xform(tb, ld3(R + (hinge ? 25 : 7)));
    const V3 rel1 = sub(pivotAInW, ta.o), rel2 = sub(pivotBInW, tb.o);
    V3 n[3];
    if (hinge) {
      const V3 relPos = sub(pivotBInW, pivotAInW);
      n[0] = len2(relPos) > 1.19209290e-07f ? nor(relPos) : v3(1.0f, 0.0f, 0.0f);
      planeSpace1(n[0], n[1], n[2]);
    } else {
      for (int i = 0; i < 3; i++) n[i] = unitAxis(i);
    }
    const M3 world2A = transpose(ta.b), world2B = transpose(tb.b);
    for (int i = 0; i < 3; i++) {
      const V3 aJ = mulMV(world2A, crs(rel1, n[i]));
      const V3 bJ = mulMV(world2B, crs(rel2, scl(n[i], -1.0f)));
      const V3 m0 = mulc(ilA, aJ), m1 = mulc(ilB, bJ);
      const float diag = imA + dot(m0, aJ) + imB + dot(m1, bJ);
      S[i] = 1.0f / diag;
      if (hinge) st3(S + 32 + 3 * i, n[i]);
    }
    st3(S + 4, rel1);
    st3(S + 7, rel2);
    st3(S + 10, sub(pivotAInW, pivotBInW));
  }
  if (!hinge) return;
  const M3 fa = loadM3(R + 4), fb = loadM3(R + 16);
  // getHingeAngle: B's second frame axis against A's first two, in world space
  const V3 ref0 = mulMV(ta.b, column(fa, 0)), ref1 = mulMV(ta.b, column(fa, 1));
  const V3 swing = mulMV(tb.b, column(fb, 1));
  const float angle = atan2Fast(dot(swing, ref0), dot(swing, ref1));
  const float lower = R[28], upper = R[29], softness = R[30];
  float correction = 0.0f, limitSign = 0.0f, solveLimit = 0.0f;
  if (lower < upper) {
    if (angle <= lower * softness) {
      correction = lower - angle;
      limitSign = 1.0f;
      solveLimit = 1.0f;
    } else if (angle >= upper * softness) {
      correction = upper - angle;
      limitSign = -1.0f;
      solveLimit = 1.0f;
    }
  }
  const V3 axisA = mulMV(ta.b, column(fa, 2)), axisB = mulMV(tb.b, column(fb, 2));
  st3(S + 41, axisA);
  st3(S + 44, axisB);
  S[47] = 1.0f / (angularDenominator(IA, axisA) + angularDenominator(IB, axisA));
  S[48] = correction;
  S[49] = limitSign;
  S[50] = solveLimit;
  S[51] = 0.0f;
}
// RigidBody.applyImpulse (RigidBody.java:414-428, linear and angular factors 1): applyCentralImpulse, then
// applyTorqueImpulse(relPos x impulse) through the world inverse inertia
__device__ __forceinline__ void applyImpulseRB(SBody& b, V3 impulse, V3 relPos, const float* I) {
  if (b.invMass == 0.0f) return;
  b.lv = add(b.lv, scl(impulse, b.invMass));
  M3 m;
  for (int k = 0; k < 9; k++) m.m[k] = I[k];
  b.av = add(b.av, mulMV(m, crs(relPos, impulse)));
}
// RigidBody.applyTorqueImpulse (RigidBody.java:405-411): no inverse-mass test
__device__ __forceinline__ void applyTorqueImpulse(SBody& b, V3 torque, const float* I) {
  M3 m;
  for (int k = 0; k < 9; k++) m.m[k] = I[k];
  b.av = add(b.av, mulMV(m, torque));
}
// Point2PointConstraint.solveConstraint (Point2PointConstraint.java:109-177): per world axis the positional error
// (tau / dt) and the damped relative velocity along it, clamped when impulseClamp > 0
__device__ __forceinline__ void p2pSolve(const float* R, float* S, float dt, SBody& A, SBody& B) {
  const float tau = R[10], damping = R[11], clamp = R[12];
  const V3 rel1 = ld3(S + 4), rel2 = ld3(S + 7), dpiv = ld3(S + 10);
  float applied = S[3];
  for (int i = 0; i < 3; i++) {
    const V3 normal = unitAxis(i);
    const float jacDiagABInv = S[i];
    const V3 vel1 = add(crs(A.av, rel1), A.lv), vel2 = add(crs(B.av, rel2), B.lv);  // getVelocityInLocalPoint
    const V3 vel = sub(vel1, vel2);
    const float rel_vel = dot(normal, vel);
    const float depth = -dot(dpiv, normal);
    float impulse = depth * tau / dt * jacDiagABInv - damping * rel_vel * jacDiagABInv;
    if (clamp > 0.0f) {
      if (impulse < -clamp) impulse = -clamp;
      if (impulse > clamp) impulse = clamp;
    }
    applied += impulse;
    const V3 iv = scl(normal, impulse);
    applyImpulseRB(A, iv, rel1, S + 14);
    applyImpulseRB(B, scl(iv, -1.0f), rel2, S + 23);
  }
  S[3] = applied;
}
// HingeConstraint.solveConstraint (HingeConstraint.java:335-520): the linear rows on the relative pivot velocity
// taken once before them; then, on the angular velocities read after them, the torque that removes the relative
// angular velocity orthogonal to the hinge axis and the one that realigns the axes, the limit (accumulated impulse
// kept >= 0) and the motor (clipped to its maximum impulse, not accumulated)
__device__ __forceinline__ void hingeSolve(const float* R, float* S, float dt, SBody& A, SBody& B) {
  const float tau = 0.3f;
  const float* IA = S + 14;
  const float* IB = S + 23;
  if (R[33] == 0.0f) {
    const V3 rel1 = ld3(S + 4), rel2 = ld3(S + 7), dpiv = ld3(S + 10);
    const V3 vel = sub(add(crs(A.av, rel1), A.lv), add(crs(B.av, rel2), B.lv));
    float applied = S[3];
    for (int i = 0; i < 3; i++) {
      const V3 n = ld3(S + 32 + 3 * i);
      const float jacDiagABInv = S[i];
      const float rel_vel = dot(n, vel);
      const float depth = -dot(dpiv, n);
      const float impulse = depth * tau / dt * jacDiagABInv - rel_vel * jacDiagABInv;
      applied += impulse;
      const V3 iv = scl(n, impulse);
      applyImpulseRB(A, iv, rel1, IA);
      applyImpulseRB(B, scl(iv, -1.0f), rel2, IB);
    }
    S[3] = applied;
  }
  const V3 axisA = ld3(S + 41), axisB = ld3(S + 44);
  const float relaxationFactor = R[32];
  const M3 mA = loadM3(IA), mB = loadM3(IB);
  const V3 angVelA = A.av, angVelB = B.av;
  const V3 aroundA = scl(axisA, dot(axisA, angVelA));
  const V3 aroundB = scl(axisB, dot(axisB, angVelB));
  V3 velrelOrthog = sub(sub(angVelA, aroundA), sub(angVelB, aroundB));
  const float relaxation = 1.0f;
  if (len(velrelOrthog) > 0.00001f) {
    const V3 n = nor(velrelOrthog);
    const float denom = angularDenominator(mA, n) + angularDenominator(mB, n);
    velrelOrthog = scl(velrelOrthog, (1.0f / denom) * relaxationFactor);
  }
  V3 angularError = crs(axisA, axisB);
  angularError = scl(angularError, -1.0f);
  angularError = scl(angularError, 1.0f / dt);
  if (len(angularError) > 0.00001f) {
    const V3 n2 = nor(angularError);
    const float denom2 = angularDenominator(mA, n2) + angularDenominator(mB, n2);
    angularError = scl(angularError, (1.0f / denom2) * relaxation);
  }
  applyTorqueImpulse(A, add(scl(velrelOrthog, -1.0f), angularError), IA);
  applyTorqueImpulse(B, sub(velrelOrthog, angularError), IB);
  if (S[50] != 0.0f) {
    const float amplitude =
        (dot(sub(angVelB, angVelA), axisA) * relaxationFactor + S[48] * (1.0f / dt) * R[31]) * S[49];
    float impulseMag = amplitude * S[47];
    const float temp = S[51];
    S[51] = jmax(S[51] + impulseMag, 0.0f);
    impulseMag = S[51] - temp;
    const V3 impulse = scl(axisA, impulseMag * S[49]);
    applyTorqueImpulse(A, impulse, IA);
    applyTorqueImpulse(B, scl(impulse, -1.0f), IB);
  }
  if (R[34] != 0.0f) {
    const V3 angularLimit = v3(0.0f, 0.0f, 0.0f);
    const float projRelVel = dot(sub(aroundA, aroundB), axisA);
    const float motor_relvel = R[35] - projRelVel;
    const float unclipped = S[47] * motor_relvel;
    const float maxImp = R[36];
    float clipped = unclipped > maxImp ? maxImp : unclipped;
    clipped = clipped < -maxImp ? -maxImp : clipped;
    const V3 motorImp = scl(axisA, clipped);
    applyTorqueImpulse(A, add(motorImp, angularLimit), IA);
    applyTorqueImpulse(B, sub(scl(motorImp, -1.0f), angularLimit), IB);
  }
}
__device__ void jointSolve(int j, float dt, const float* __restrict__ rec, float* st, float* svel) {
  const float* R = rec + kJointWords * j;
  float* S = st + kJointStateWords * j;
  const int a = jointBodyA(R), b = jointBodyB(R);
  SBody A = loadBody(svel, nullptr, a), B;
  if (b >= 0) B = loadBody(svel, nullptr, b);
  else {
    B.lv = v3(0.0f, 0.0f, 0.0f);
    B.av = v3(0.0f, 0.0f, 0.0f);
    B.invMass = 0.0f;
  }
  if (jointType(R) == kJointHinge) hingeSolve(R, S, dt, A, B);
  else p2pSolve(R, S, dt, A, B);
  // a body without inverse mass keeps its record (the reference's read-back skips it as well)
  if (A.invMass != 0.0f) storeBody(svel, a, A);
  if (b >= 0 && B.invMass != 0.0f) storeBody(svel, b, B);
}
// one joint colour per launch: joints [off, off + cnt) of the order
__global__ void kJointSweep(int off, int cnt, float dt, const int* __restrict__ order, const int* __restrict__ solve,
                            const float* __restrict__ rec, float* st, float* svel, const int* __restrict__ done) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt || (done && *done)) return;
  const int j = order[off + k];
  if (solve[j]) jointSolve(j, dt, rec, st, svel);
}
// every joint colour in one workgroup, colour after colour, separated by workgroup barriers
constexpr int kJointTailThreads = 1024;
__global__ __launch_bounds__(kJointTailThreads) void kJointSweepAll(int ncol, const int* __restrict__ colourOff,
                                                                    float dt, const int* __restrict__ order,
                                                                    const int* __restrict__ solve,
                                                                    const float* __restrict__ rec, float* st,
                                                                    float* svel, const int* __restrict__ done) {
  if (done && *done) return;
  for (int c = 0; c < ncol; c++) {
    const int a = colourOff[c], e = colourOff[c + 1];
    for (int k = a + (int)threadIdx.x; k < e; k += kJointTailThreads) {
      const int j = order[k];
      if (solve[j]) jointSolve(j, dt, rec, st, svel);
    }
    __syncthreads();
  }
}

}  // namespace

void solverGetImpulses(Ctx& c, float* out, int64_t nslots) {
  if (nslots > 4 * c.nActive) throw std::length_error("more slots than the last solve had");
  if (nslots <= 0) return;
  hipStream_t s = c.stream;
  c.mlpY.ensure((size_t)3 * nslots);
  hipLaunchKernelGGL(kGatherImpulses, dim3(gridFor(nslots)), dim3(256), 0, s, nslots, c.hdr.p, c.nrow.p, c.frow.p,
                     c.mlpY.p);
  JB_CHECK(hipGetLastError());
  JB_CHECK(hipMemcpyAsync(out, c.mlpY.p, 3 * (size_t)nslots * sizeof(float), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
}

void jointsChanged(Ctx& c) { c.jointOrderReady = false; }

// The joints' solve order: greedy colours in creation order (a joint takes the smallest colour that no earlier
// joint on one of its moving bodies has; a body moves when it is not static or has an inverse mass), then colour
// after colour.  Solving the colours one after the other is the sequential solver fed this joint order.
static void jointsMakeOrder(Ctx& c) {
  if (c.jointOrderReady) return;
  const int nj = c.nj;
  std::vector<int> colour(nj);
  std::vector<std::vector<int>> used(c.nb);
  auto moves = [&](int b) { return b >= 0 && (!c.bodyStaticHost[b] || c.bodyInvMassHost[b] != 0.0f); };
  auto taken = [&](int b, int k) {
    if (!moves(b)) return false;
    for (int u : used[b])
      if (u == k) return true;
    return false;
  };
  int ncol = 0;
  for (int j = 0; j < nj; j++) {
    const int a = ldi(&c.jointHost[(size_t)kJointWords * j + 1]), b = ldi(&c.jointHost[(size_t)kJointWords * j + 2]);
    int k = 0;
    while (taken(a, k) || taken(b, k)) k++;
    colour[j] = k;
    if (moves(a)) used[a].push_back(k);
    if (moves(b)) used[b].push_back(k);
    ncol = std::max(ncol, k + 1);
  }
  c.jointColourOff.assign(ncol + 1, 0);
  for (int j = 0; j < nj; j++) c.jointColourOff[colour[j] + 1]++;
  for (int k = 0; k < ncol; k++) c.jointColourOff[k + 1] += c.jointColourOff[k];
  std::vector<int> order(std::max(nj, 1)), fill(c.jointColourOff.begin(), c.jointColourOff.end() - 1);
  for (int j = 0; j < nj; j++) order[fill[colour[j]]++] = j;
  c.jointOrder.ensure(std::max(nj, 1));
  c.jointColourOffDev.ensure(ncol + 1);
  c.jointSolve.ensure(std::max(nj, 1));
  JB_CHECK(hipMemcpy(c.jointOrder.p, order.data(), (size_t)nj * sizeof(int), hipMemcpyHostToDevice));
  JB_CHECK(hipMemcpy(c.jointColourOffDev.p, c.jointColourOff.data(), (ncol + 1) * sizeof(int),
                     hipMemcpyHostToDevice));
  JB_CHECK(hipMemset(c.jointSolve.p, 0, (size_t)std::max(nj, 1) * sizeof(int)));  // none solved since the change
  c.jointOrderReady = true;
}

void jointsImpulses(Ctx& c, float* out) {
  if (c.nj == 0) return;
  jointsMakeOrder(c);
  std::vector<float> st((size_t)kJointStateWords * c.nj);
  JB_CHECK(hipMemcpy(st.data(), c.jointState.p, st.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int j = 0; j < c.nj; j++) out[j] = st[(size_t)kJointStateWords * j + 3];
}

int jointsOrderSolved(Ctx& c, int32_t* out, int cap) {
  if (c.nj == 0) return 0;
  jointsMakeOrder(c);
  std::vector<int> order(c.nj), solve(c.nj);
  JB_CHECK(hipMemcpy(order.data(), c.jointOrder.p, (size_t)c.nj * sizeof(int), hipMemcpyDeviceToHost));
  JB_CHECK(hipMemcpy(solve.data(), c.jointSolve.p, (size_t)c.nj * sizeof(int), hipMemcpyDeviceToHost));
  int n = 0;
  for (int k = 0; k < c.nj; k++)
    if (solve[order[k]]) {
      if (out && n < cap) out[n] = order[k];
      n++;
    }
  return n;
}

namespace {
// the step's joint work, in the order the solve needs it
struct JointPass {
  Ctx& c;
  bool on;
  void hook(int* parent, int* changed) {  // island merges, inside the island labelling loop
    if (on)
      hipLaunchKernelGGL(kIslandHookJoints, dim3(gridFor(c.nj)), dim3(256), 0, c.stream, c.nj, (const float*)c.joints.p,
                         (const float*)c.props.p, (const int*)c.act.p, parent, changed);
  }
  void mark(const int* island, const int* inactive) {  // which joints this solve processes
    if (on)
      hipLaunchKernelGGL(kJointMark, dim3(gridFor(c.nj)), dim3(256), 0, c.stream, c.nj, (const float*)c.joints.p,
                         (const float*)c.props.p, (const int*)c.act.p, island, inactive, c.jointSolve.p);
  }
  void build() {
    if (on)
      hipLaunchKernelGGL(kJointBuild, dim3(gridFor(c.nj)), dim3(256), 0, c.stream, c.nj, (const float*)c.joints.p,
                         (const int*)c.jointSolve.p, (const float*)c.tr.p, (const float*)c.props.p, c.jointState.p);
  }
  int sweep(float dt, const int* done) {  // one iteration's constraints; returns the launches
    if (!on) return 0;
    const int ncol = (int)c.jointColourOff.size() - 1;
    if (c.nj <= 4 * kJointTailThreads) {
      hipLaunchKernelGGL(kJointSweepAll, dim3(1), dim3(kJointTailThreads), 0, c.stream, ncol,
                         (const int*)c.jointColourOffDev.p, dt, (const int*)c.jointOrder.p,
                         (const int*)c.jointSolve.p, (const float*)c.joints.p, c.jointState.p, c.svel.p, done);
      return 1;
    }
    for (int k = 0; k < ncol; k++) {
      const int cnt = c.jointColourOff[k + 1] - c.jointColourOff[k];
      hipLaunchKernelGGL(kJointSweep, dim3(gridFor(cnt)), dim3(256), 0, c.stream, c.jointColourOff[k], cnt, dt,
                         (const int*)c.jointOrder.p, (const int*)c.jointSolve.p, (const float*)c.joints.p,
                         c.jointState.p, c.svel.p, done);
    }
    return ncol;
  }
};
}  // namespace

namespace {
// Block tables of the world-block solve: dynamic bodies grouped by world (world id, then index), whole worlds
// packed into blocks of at most wbBodies dynamic bodies in world order (a larger world forms a block of its
// own, up to the LDS limit of 1024 bodies).
void worldBlocksPrepare(Ctx& c) {
  const int nb = c.nb, B = c.wbBodies;
  std::vector<int> ids;
  for (int i = 0; i < nb; i++)
    if (!c.bodyStaticHost[i]) ids.push_back(i);
  std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return c.worldHost[a] < c.worldHost[b]; });
  std::vector<int> start(1, 0), block(nb, -1), local(nb, -1);
  int cur = 0, maxBodies = 0;
  for (size_t i = 0; i < ids.size();) {
    size_t j = i;
    while (j < ids.size() && c.worldHost[ids[j]] == c.worldHost[ids[i]]) j++;
    const int w = (int)(j - i);
    if (w > 1024) throw std::length_error("a world has more than 1024 dynamic bodies: too many for one block");
    if (cur > 0 && cur + w > B) {
      start.push_back((int)i);
      cur = 0;
    }
    cur += w;
    i = j;
  }
  start.push_back((int)ids.size());
  const int nBlocks = (int)start.size() - 1;
  for (int b = 0; b < nBlocks; b++) {
    maxBodies = std::max(maxBodies, start[b + 1] - start[b]);
    for (int q = start[b]; q < start[b + 1]; q++) {
      block[ids[q]] = b;
      local[ids[q]] = q - start[b];
    }
  }
  hipStream_t s = c.stream;
  c.wbBodyList.ensure(std::max<size_t>(ids.size(), 1));
  c.wbBodyStart.ensure(start.size());
  c.wbBodyBlock.ensure(nb);
  c.wbBodyLocal.ensure(nb);
  if (!ids.empty()) JB_CHECK(hipMemcpyAsync(c.wbBodyList.p, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.wbBodyStart.p, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.wbBodyBlock.p, block.data(), (size_t)nb * sizeof(int), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.wbBodyLocal.p, local.data(), (size_t)nb * sizeof(int), hipMemcpyHostToDevice, s));
  JB_CHECK(hipStreamSynchronize(s));
  c.wbBlocks = nBlocks;
  c.wbMaxBodies = maxBodies;
  c.wbReady = true;
}

int bitsForBlocks(int n) {
  int b = 1;
  while ((1 << b) <= n) b++;
  return b;
}

// the whole solve of batched worlds: manifolds grouped by block, rows built in that order, one launch per solve.
// dna: the graph-captured form (na is then the capacity, *dna the active manifolds; no host-side statistics)
void solveWorldBlocks(Ctx& c, float dt, int na, const int* dna = nullptr) {
  hipStream_t s = c.stream;
  float* man = c.man[c.cur].p;
  const bool graph = dna != nullptr;
  const int* abortFlag = graph ? c.gc.p + kGcAbort : nullptr;
  auto clear = [&](void* p, size_t bytes) {  // a captured step clears these with its first fill
    if (!graph) JB_CHECK(hipMemsetAsync(p, 0, bytes, s));
  };
  if (!c.wbReady) worldBlocksPrepare(c);
  const int nBlocks = c.wbBlocks;
  tBegin(c, T_COLOR);
  c.activeBodies.ensure(na);
  c.wbKey.ensure(na);
  c.wbKeyAlt.ensure(na);
  c.wbIdx.ensure(na);
  c.colorOrderAlt.ensure(na);
  c.wbSlotStart.ensure((size_t)nBlocks + 1);
  c.wbInfo.ensure(4);
  clear(c.wbInfo.p, 4 * sizeof(int));
  hipLaunchKernelGGL(kActiveBodies, dim3(gridFor(na)), dim3(256), 0, s, na, dna, c.mActive.p, man, c.activeBodies.p);
  hipLaunchKernelGGL(kManifoldBlocks, dim3(gridFor(na)), dim3(256), 0, s, na, dna, (const int2*)c.activeBodies.p,
                     (const int*)c.wbBodyBlock.p, c.wbKey.p, c.wbIdx.p, c.wbInfo.p);
  // slot order: active manifolds by block (stable: pair-key order inside a block)
  Db<int> dk(c.wbKey.p, c.wbKeyAlt.p), dv(c.wbIdx.p, c.colorOrderAlt.p);
  const int keyBits = bitsForBlocks(nBlocks);
  if (graph)
    primSortPairsDev(c, dk, dv, dna, na, 0, keyBits);
  else
    primSortPairs(c, dk, dv, na, 0, keyBits);
  const int* slotManifold = dv.Current();  // slot p -> active manifold
  hipLaunchKernelGGL(kBlockSlotStarts, dim3(gridFor((int64_t)na + 1)), dim3(256), 0, s, na, dna, nBlocks,
                     (const int*)dk.Current(), c.wbSlotStart.p);
  tEnd(c, T_COLOR);
  // setup: rows of every slot (the warm start is applied inside the block solve)
  tBegin(c, T_SOLVE_SETUP);
  c.hdr.ensure(na);
  c.nrow.ensure((size_t)NW * na);
  c.frow.ensure((size_t)FW * na);
  c.svel.ensure((size_t)8 * c.nb);
  hipLaunchKernelGGL(kSolverBodies, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, (const float*)c.vel.p,
                     (const float*)c.props.p, c.svel.p);
  const bool learnedHere = c.useMlp;
  SolverParams SP{dt, c.erp, c.warmFactor, c.linearSlop, c.warmstart, (c.useLearned && !c.useMlp) ? 1 : 0,
                  (c.dumpFeatures || learnedHere) ? 1 : 0, 0.0f};
  if (SP.dumpFeatures) c.features.ensure((size_t)16 * 4 * na);
  if ((int64_t)na * 4 >= (1ll << 27)) throw std::length_error("too many active manifolds for the slot header");
  c.slotCount.ensure(na);
  c.slotStart.ensure(na);
  hipLaunchKernelGGL(kSlotRows, dim3(gridFor(na)), dim3(256), 0, s, na, dna, slotManifold, c.mActive.p, man,
                     c.slotCount.p);
  if (graph)
    primScanDev(c, c.slotCount.p, c.slotStart.p, dna, na);
  else
    primScan(c, c.slotCount.p, c.slotStart.p, na);
  hipLaunchKernelGGL(kBuildRows, dim3(gridFor(na, 128)), dim3(128), 0, s, na, dna, slotManifold, c.mActive.p, SP, man,
                     c.tr.p, c.vel.p, c.props.p, (const int*)c.slotStart.p, c.hdr.p, c.nrow.p, c.frow.p,
                     SP.useLearned ? c.learned.p : nullptr, SP.dumpFeatures ? c.features.p : nullptr);
  c.rowCount.ensure(2);
  hipLaunchKernelGGL(kRowTotals, dim3(1), dim3(64), 0, s, na, dna, na, (const int*)c.slotCount.p, (const int*)c.slotStart.p,
                     c.rowCount.p);
  if (learnedHere) mlpWarmStart(c, na);
  tEnd(c, T_SOLVE_SETUP);
  // the block solve: colouring, warm start, every iteration, velocities
  const int iters = c.numIterations;
  const size_t histWords = 2 * (size_t)iters + 4, stripeWords = 2 * (size_t)kResidStripes * iters;
  c.resid.ensure(histWords + stripeWords);
  c.done.ensure(4);
  clear(c.resid.p, (histWords + stripeWords) * sizeof(double));
  clear(c.done.p, 4 * sizeof(int));
  float* resOut = (float*)(c.counter.p + 6);
  clear(resOut, 2 * sizeof(float));
  double* res = iters > 0 ? c.resid.p + histWords + 2 * (size_t)kResidStripes * (iters - 1) : nullptr;
  c.wbColour.ensure(na);
  c.wbSolveSlot.ensure(na);
  c.wbSlotInfo.ensure(na);
  c.colorOrder.ensure(na);
  const bool apply = c.warmstart || c.useLearned || c.useMlp;
  const size_t ldsBytes = (size_t)std::max(c.wbMaxBodies, 1) * (8 * sizeof(float) + 2 * sizeof(unsigned long long));
  tBegin(c, T_SOLVE_ITER);
#define JB_WB_LAUNCH(T, P)                                                                                        \
  hipLaunchKernelGGL((kSolveWorldBlocks<T, P>), dim3(nBlocks), dim3(T), ldsBytes, s, iters, SP, apply ? 1 : 0,    \
                     (const int*)c.wbBodyList.p, (const int*)c.wbBodyStart.p, (const int*)c.wbBodyLocal.p,        \
                     (const int*)c.wbSlotStart.p, slotManifold, c.wbColour.p, c.wbSolveSlot.p, c.colorOrder.p,     \
                     (const int4*)c.hdr.p, c.wbSlotInfo.p, c.nrow.p, c.frow.p, (const float*)c.props.p,            \
                     (const float*)c.svel.p, c.vel.p, res, c.wbInfo.p, abortFlag)
  if (c.wbPipelined) {
    switch (c.wbThreads) {
      case 64: JB_WB_LAUNCH(64, true); break;
      case 128: JB_WB_LAUNCH(128, true); break;
      case 512: JB_WB_LAUNCH(512, true); break;
      default: JB_WB_LAUNCH(256, true); break;
    }
  } else {
    switch (c.wbThreads) {
      case 64: JB_WB_LAUNCH(64, false); break;
      case 128: JB_WB_LAUNCH(128, false); break;
      case 512: JB_WB_LAUNCH(512, false); break;
      default: JB_WB_LAUNCH(256, false); break;
    }
  }
#undef JB_WB_LAUNCH
  JB_CHECK(hipGetLastError());
  if (iters > 0)
    hipLaunchKernelGGL(kResidual, dim3(1), dim3(64), 0, s, iters - 1, 0.0f, res, c.resid.p, c.done.p, resOut);
  tEnd(c, T_SOLVE_ITER, iters > 0 ? 2 : 1);
  hipLaunchKernelGGL(kWriteback, dim3(gridFor(na)), dim3(256), 0, s, na, dna, c.hdr.p, c.nrow.p, c.frow.p, man);
  // the order is reported as one batch (block after block, colour after colour inside each block)
  c.colorOff.assign({0, (int64_t)na});
  c.ncolors = 1;
  c.headColors = 1;
  c.wbLastSolve = true;
  if (graph) return;  // statistics: solverStatsCopy after the graph's launch
  // statistics: read after the step's final wait
  c.solveRo.ensure(2);
  c.solveRh.ensure(2 * (size_t)std::max(iters, 1));
  c.solveRows.ensure(2);
  c.wbInfoHost.ensure(4);
  JB_CHECK(hipMemcpyAsync(c.solveRo.p, resOut, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
  if (iters > 0)
    JB_CHECK(hipMemcpyAsync(c.solveRh.p, c.resid.p, 2 * (size_t)iters * sizeof(double), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(c.solveRows.p, c.rowCount.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(c.wbInfoHost.p, c.wbInfo.p, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
  c.solveStat.pending = true;
  c.solveStat.tolMode = false;
  c.solveStat.blockSweeps = false;
  c.solveStat.sampleK = false;
  c.solveStat.iters = iters;
  c.solveStat.kUsed = 0;
  c.wbLastSolve = true;
}
}  // namespace

// A solve with constraints and no contact rows (solveGroup with numManifolds == 0): the constraints' iterations
// on the solver records, then the velocities back into the bodies.  In tolerance mode the residual over the
// (absent) contact rows is 0, so one iteration runs.
static void solveJointsOnly(Ctx& c, float dt, JointPass& jp) {
  hipStream_t s = c.stream;
  c.svel.ensure((size_t)8 * c.nb);
  tBegin(c, T_SOLVE_SETUP);
  hipLaunchKernelGGL(kSolverBodies, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, (const float*)c.vel.p,
                     (const float*)c.props.p, c.svel.p);
  jp.build();
  tEnd(c, T_SOLVE_SETUP, 2);
  const int iters = c.tolerance > 0.0f ? 1 : c.numIterations;
  tBegin(c, T_SOLVE_ITER);
  int launches = 0;
  for (int itn = 0; itn < iters; itn++) launches += jp.sweep(dt, nullptr);
  tEnd(c, T_SOLVE_ITER, launches);
  hipLaunchKernelGGL(kCopyVel, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, c.svel.p, c.vel.p, c.props.p);
  JB_CHECK(hipGetLastError());
  c.st.iterations = iters;
}

void solverSolve(Ctx& c, float dt, const std::vector<int>* explicitList) {
  hipStream_t s = c.stream;
  const int64_t np = c.npairs;
  c.nActive = 0;
  c.ncolors = 0;
  c.colorOff.assign(1, 0);
  c.st.rows = 0;
  c.st.colors = 0;
  c.st.iterations = 0;
  c.st.residual = 0.0f;
  c.residHist.clear();
  // the world's constraints are solved with its steps (a caller's island list carries manifolds only)
  JointPass jp{c, c.nj > 0 && !explicitList};
  if (jp.on) jointsMakeOrder(c);
  if (np == 0 && !jp.on) return;
  float* man = c.man[c.cur].p;
  // active manifolds (stable compaction keeps pair-key order)
  c.flags.ensure(np);
  c.mActive.ensure(np);
  c.counter.ensure(8);
  const int* island = nullptr;
  int na = 0;
  if (explicitList) {  // the caller's islands: exactly these manifolds, in the caller's order
    na = (int)explicitList->size();
    c.mActive.ensure(std::max(na, 1));
    if (na) JB_CHECK(hipMemcpyAsync(c.mActive.p, explicitList->data(), (size_t)na * sizeof(int), hipMemcpyHostToDevice, s));
  } else {
  if (!c.deactivationDisabled) {  // islands decide sleeping; with deactivation disabled every body stays awake
    const int nb = c.nb;
    c.island.ensure(nb);
    c.islandFlag.ensure(2 * (size_t)nb + 1);
    int* awake = c.islandFlag.p;
    int* inactive = awake + nb;
    int* changed = inactive + nb;
    hipLaunchKernelGGL(kIslandInit, dim3(gridFor(nb)), dim3(256), 0, s, nb, c.island.p);
    for (int round = 0;; round++) {
      JB_CHECK(hipMemsetAsync(changed, 0, sizeof(int), s));
      if (np > 0)
        hipLaunchKernelGGL(kIslandHook, dim3(gridFor(np)), dim3(256), 0, s, np, (const uint64_t*)c.keys[c.cur].p,
                           (const float*)c.props.p, c.island.p, changed);
      jp.hook(c.island.p, changed);
      hipLaunchKernelGGL(kIslandCompress, dim3(gridFor(nb)), dim3(256), 0, s, nb, c.island.p);
      int h = 0;
      JB_CHECK(hipMemcpyAsync(&h, changed, sizeof(int), hipMemcpyDeviceToHost, s));
      JB_CHECK(hipStreamSynchronize(s));
      if (!h) break;
      if (round > 64) throw std::runtime_error("island labelling did not converge");
    }
    JB_CHECK(hipMemsetAsync(awake, 0, 2 * (size_t)nb * sizeof(int), s));
    hipLaunchKernelGGL(kIslandAwake, dim3(gridFor(nb)), dim3(256), 0, s, nb, (const int*)c.island.p,
                       (const float*)c.props.p, (const int*)c.act.p, awake);
    hipLaunchKernelGGL(kIslandApply, dim3(gridFor(nb)), dim3(256), 0, s, nb, (const int*)c.island.p,
                       (const float*)c.props.p, (const int*)awake, c.act.p, inactive);
    island = c.island.p;
  }
  jp.mark(island, island ? (const int*)(c.islandFlag.p + c.nb) : nullptr);
  if (np > 0) {
    hipLaunchKernelGGL(kMarkActive, dim3(gridFor(np)), dim3(256), 0, s, np, (const int*)nullptr, man, c.ptype.p, c.act.p, c.props.p,
                       island, island ? (const int*)(c.islandFlag.p + c.nb) : nullptr, c.flags.p);
    JB_CHECK(hipMemsetAsync(c.counter.p, 0, 8 * sizeof(unsigned long long), s));
    primSelectIndex(c, c.flags.p, (int)np, c.mActive.p, (int*)c.counter.p);
    JB_CHECK(hipMemcpyAsync(&na, c.counter.p, sizeof(int), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipStreamSynchronize(s));
  }
  }
  c.nActive = na;
  c.wbLastSolve = false;
  if (na == 0) {
    if (jp.on) solveJointsOnly(c, dt, jp);
    return;
  }
  // batched worlds: one launch per solve (scenes without constraints)
  if (!explicitList && c.wbBodies > 0 && !(c.tolerance > 0.0f) && !jp.on) {
    if (c.useMlp && !c.mlpReady) throw std::logic_error("learned warm start enabled without a model");
    if (!c.useMlp && c.useLearned && c.learnedSlots < 4 * (int64_t)na)
      throw std::length_error("learned warm start has fewer slots than 4 x active manifolds");
    solveWorldBlocks(c, dt, na);
    return;
  }
  if (c.useMlp && !c.mlpReady) throw std::logic_error("learned warm start enabled without a model");
  if (!c.useMlp && c.useLearned && c.learnedSlots < 4 * (int64_t)na)
    throw std::length_error("learned warm start has fewer slots than 4 x active manifolds");
  // colouring
  c.color.ensure(na);
  c.colorAlt.ensure(na);
  c.colorOrder.ensure(na);
  c.colorOrderAlt.ensure(na);
  c.activeBodies.ensure(na);
  c.bodyPrio.ensure(2 * (size_t)c.nb);
  c.bodyMask.ensure(c.nb);
  tBegin(c, T_COLOR);
  hipLaunchKernelGGL(kActiveBodies, dim3(gridFor(na)), dim3(256), 0, s, na, (const int*)nullptr, c.mActive.p, man, c.activeBodies.p);
  const int* deg = nullptr;  // largest-degree-first priorities (null: hashed priorities only)
  if (c.colourByDegree) {
    c.bodyDeg.ensure(c.nb);
    c.mDeg.ensure(na);
    JB_CHECK(hipMemsetAsync(c.bodyDeg.p, 0, (size_t)c.nb * sizeof(int), s));
    hipLaunchKernelGGL(kBodyDegree, dim3(gridFor(na)), dim3(256), 0, s, na, (const int2*)c.activeBodies.p,
                       (const float*)c.props.p, c.bodyDeg.p);
    hipLaunchKernelGGL(kManifoldDegree, dim3(gridFor(na)), dim3(256), 0, s, na, (const int2*)c.activeBodies.p,
                       (const float*)c.props.p, (const int*)c.bodyDeg.p, c.mDeg.p);
    deg = c.mDeg.p;
  }
  JB_CHECK(hipMemsetAsync(c.color.p, 0xff, (size_t)na * sizeof(int), s));
  JB_CHECK(hipMemsetAsync(c.bodyPrio.p, 0, 2 * (size_t)c.nb * sizeof(unsigned long long), s));
  JB_CHECK(hipMemsetAsync(c.bodyMask.p, 0, (size_t)c.nb * sizeof(unsigned long long), s));
  // block partition of the dynamic bodies along this step's Morton order: interior / boundary manifolds
  const bool blocked = !explicitList && c.blockSize > 0 && c.sortedOrder != nullptr && c.np == c.nb;
  const int B = c.blockSize;
  const int* phase = nullptr;
  c.nBlocks = 0;
  if (blocked) {
    c.blockLocal.ensure(c.nb);
    c.mPhase.ensure(na);
    hipLaunchKernelGGL(kBlockPos, dim3(gridFor(c.np)), dim3(256), 0, s, c.np, c.sortedOrder, (const float*)c.props.p,
                       c.blockLocal.p);
    hipLaunchKernelGGL(kClassify, dim3(gridFor(na)), dim3(256), 0, s, na, B, (const int2*)c.activeBodies.p,
                       (const int*)c.blockLocal.p, c.mPhase.p);
    phase = c.mPhase.p;
    c.nBlocks = (c.np + B - 1) / B;
    if (c.nBlocks >= (1 << 20)) throw std::length_error("too many sweep blocks for the slot key");
  }
  // colouring: batches of rounds over the pending list, which is compacted between batches (the host waits
  // only there); with blocks, interior and boundary manifolds are coloured separately
  c.pend.ensure(na);
  c.pendAlt.ensure(na);
  c.pendFlag.ensure(na);
  int round = 0;
  // the first batch's length adapts to the scene (one round shorter after a step whose first batch
  // coloured everything, the rounds then needed after one that did not)
  const int firstRounds = std::max(12, c.colourRounds);
  int batches = 0;
  int* remDev = (int*)(c.counter.p + 3);  // manifolds a batch left uncoloured
  // one batch of rounds over pend[0, m): the manifolds still uncoloured are compacted into pendAlt and counted
  // into *remDev, and the lists swap
  auto colourBatch = [&](int m, int rounds) {
    hipLaunchKernelGGL(kColorPropose, dim3(gridFor(m)), dim3(256), 0, s, m, (const int*)c.pend.p, round,
                       (const int2*)c.activeBodies.p, (const float*)c.props.p, (const int*)c.color.p, c.nb,
                       c.bodyPrio.p, deg);
    for (int r = 0; r < rounds; r++, round++)
      hipLaunchKernelGGL(kColorRound, dim3(gridFor(m)), dim3(256), 0, s, m, (const int*)c.pend.p, round,
                         r + 1 < rounds, (const int2*)c.activeBodies.p, (const float*)c.props.p, c.color.p, c.nb,
                         c.bodyPrio.p, c.bodyMask.p, deg);
    hipLaunchKernelGGL(kUncoloured, dim3(gridFor(m)), dim3(256), 0, s, m, (const int*)c.pend.p, (const int*)c.color.p,
                       c.pendFlag.p);
    primSelectValues<int>(c, c.pend.p, c.pendFlag.p, m, c.pendAlt.p, remDev);
    std::swap(c.pend.p, c.pendAlt.p);
    std::swap(c.pend.n, c.pendAlt.n);
    batches++;
    if (round > 4000) throw std::runtime_error("colouring did not converge");
  };
  // batches until every manifold of pend[0, m) has a colour, with a host wait after each
  auto colourPass = [&](int m) {
    for (int batch = 0; m > 0; batch++) {
      colourBatch(m, batch == 0 ? firstRounds : 6);
      int rem = 0;
      JB_CHECK(hipMemcpyAsync(&rem, remDev, sizeof(int), hipMemcpyDeviceToHost, s));
      JB_CHECK(hipStreamSynchronize(s));
      m = rem;
    }
  };
  if (blocked) {
    for (int interior = 1; interior >= 0; interior--) {
      hipLaunchKernelGGL(kPassFlag, dim3(gridFor(na)), dim3(256), 0, s, na, phase, interior, c.pendFlag.p);
      primSelectIndex(c, c.pendFlag.p, na, c.pend.p, (int*)c.counter.p);
      int m = 0;
      JB_CHECK(hipMemcpyAsync(&m, c.counter.p, sizeof(int), hipMemcpyDeviceToHost, s));
      JB_CHECK(hipMemsetAsync(c.bodyPrio.p, 0, 2 * (size_t)c.nb * sizeof(unsigned long long), s));
      JB_CHECK(hipMemsetAsync(c.bodyMask.p, 0, (size_t)c.nb * sizeof(unsigned long long), s));
      JB_CHECK(hipStreamSynchronize(s));
      colourPass(m);
    }
  }
  // slot order and colour batches from the colours: interior slots by (colour, block), then boundary slots by
  // colour; count, first slots and boundary flags come back in one readback, with the first colouring
  // batch's uncoloured count (returned)
  c.slotKey.ensure(na);
  c.slotKeyAlt.ensure(na);
  c.permAlt.ensure(na);
  c.batchFlag.ensure(na + (size_t)kMaxBatches);
  c.colorStart.ensure(std::max<size_t>(na + 1, kMaxBatches));
  c.batchHost.ensure(2 + 2 * (size_t)kMaxBatches);
  int* hb = c.batchHost.p;
  const unsigned long long* skey = nullptr;  // the sorted slot keys
  auto order = [&]() {
    if (round >= (1 << 10) - 64) throw std::runtime_error("colour index exceeds the slot key");
    hipLaunchKernelGGL(kSlotKeys, dim3(gridFor(na)), dim3(256), 0, s, na, phase, (const int*)c.color.p, c.slotKey.p,
                       c.colorOrder.p);
    Db<unsigned long long> dk(c.slotKey.p, c.slotKeyAlt.p);
    Db<int> dv(c.colorOrder.p, c.permAlt.p);
    // key bits in use: without blocks only the colour (bits 20-29; stable sort, so equal keys keep slot order)
    const int keyLo = blocked ? 0 : 20, keyHi = blocked ? 53 : 30;
    primSortPairs(c, dk, dv, na, keyLo, keyHi);
    if (dv.Current() != c.colorOrder.p)
      JB_CHECK(hipMemcpyAsync(c.colorOrder.p, dv.Current(), (size_t)na * sizeof(int), hipMemcpyDeviceToDevice, s));
    skey = dk.Current();
    hipLaunchKernelGGL(kBatchFlags, dim3(gridFor(na)), dim3(256), 0, s, na, skey, c.batchFlag.p);
    primSelectIndex(c, c.batchFlag.p, na, c.colorStart.p, (int*)c.counter.p);
    int* bnd = c.batchFlag.p + na;
    hipLaunchKernelGGL(kBatchBoundary, dim3(kMaxBatches / 256), dim3(256), 0, s, (const int*)c.counter.p,
                       (const int*)c.colorStart.p, skey, bnd);
    JB_CHECK(hipMemcpyAsync(hb, c.counter.p, sizeof(int), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(hb + 1, remDev, sizeof(int), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(hb + 2, c.colorStart.p, kMaxBatches * sizeof(int), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipMemcpyAsync(hb + 2 + kMaxBatches, bnd, kMaxBatches * sizeof(int), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipStreamSynchronize(s));
    return hb[1];
  };
  if (!blocked) {
    // no host wait after the first batch: its uncoloured count comes back with the batches; in the rare case
    // that it is not zero, the remaining batches run (waiting after each) and the order is made again
    hipLaunchKernelGGL(kIota, dim3(gridFor(na)), dim3(256), 0, s, na, c.pend.p);
    colourBatch(na, firstRounds);
    const int rem = order();
    if (rem > 0) {
      colourPass(rem);
      order();
    }
    c.colourRounds = batches <= 1 ? firstRounds - 1 : round;
    for (int pass = 0; pass < c.recolourPasses; pass++) {
      // iterated greedy: the colour classes from the last to the first, every manifold of a class moved to the
      // smallest colour free at its bodies (a class shares no dynamic body, so its moves do not race); the
      // colour count never grows and the empty classes disappear when the order is made again
      const int nb0 = hb[0];
      std::vector<int> st(hb + 2, hb + 2 + nb0);
      st.push_back(na);
      for (int k = nb0 - 1; k >= 0; k--)
        hipLaunchKernelGGL(kRecolour, dim3(gridFor(st[k + 1] - st[k])), dim3(256), 0, s, st[k], st[k + 1] - st[k],
                           (const int*)c.colorOrder.p, (const int2*)c.activeBodies.p, (const float*)c.props.p,
                           c.color.p, c.bodyMask.p);
      order();
    }
  } else {
    order();
  }
  const int nbatch = hb[0];
  if (nbatch < 1 || nbatch > kMaxBatches) throw std::runtime_error("colour batch count out of range");
  const int* starts = hb + 2;
  const int* boundary = hb + 2 + kMaxBatches;
  int nInterior = 0;
  for (int k = 0; k < nbatch; k++) {
    if (k > 0) c.colorOff.push_back(starts[k]);
    if (blocked && !boundary[k]) nInterior++;
  }
  c.colorOff.push_back(na);
  c.ncolors = (int)c.colorOff.size() - 1;
  c.st.colors = c.ncolors;
  c.interiorColors = nInterior;
  c.interiorEnd = c.colorOff[nInterior];
  if (nInterior > 0) {  // per (interior batch, block): slot range; per interior slot: LDS body positions
    const size_t nr = (size_t)nInterior * c.nBlocks;
    c.runStart.ensure(nr);
    c.runEnd.ensure(nr);
    JB_CHECK(hipMemsetAsync(c.runStart.p, 0xff, nr * sizeof(int), s));
    hipLaunchKernelGGL(kRunBounds, dim3(gridFor(c.interiorEnd)), dim3(256), 0, s, (int)c.interiorEnd, nInterior,
                       (const int*)c.colorStart.p, c.nBlocks, skey, c.runStart.p, c.runEnd.p);
    c.slotLocal.ensure(c.interiorEnd);
    hipLaunchKernelGGL(kSlotLocal, dim3(gridFor(c.interiorEnd)), dim3(256), 0, s, (int)c.interiorEnd, B,
                       (const int*)c.colorOrder.p, (const int2*)c.activeBodies.p, (const int*)c.blockLocal.p,
                       c.slotLocal.p);
  }
  tEnd(c, T_COLOR);
  // setup: rows for every slot, then the warm-start impulses colour by colour
  tBegin(c, T_SOLVE_SETUP);
  c.hdr.ensure(na);
  c.nrow.ensure((size_t)NW * na);
  c.frow.ensure((size_t)FW * na);
  c.svel.ensure((size_t)8 * c.nb);
  hipLaunchKernelGGL(kSolverBodies, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, (const float*)c.vel.p,
                     (const float*)c.props.p, c.svel.p);
  jp.build();  // buildJacobian of the constraints after the contact rows (:757-763)
  const bool learnedHere = c.useMlp;
  SolverParams SP{dt, c.erp, c.warmFactor, c.linearSlop, c.warmstart, (c.useLearned && !c.useMlp) ? 1 : 0,
                  (c.dumpFeatures || learnedHere) ? 1 : 0, c.tolerance};
  if (SP.dumpFeatures) c.features.ensure((size_t)16 * 4 * na);
  if ((int64_t)na * 4 >= (1ll << 27)) throw std::length_error("too many active manifolds for the slot header");
  c.slotCount.ensure(na);
  c.slotStart.ensure(na);
  hipLaunchKernelGGL(kSlotRows, dim3(gridFor(na)), dim3(256), 0, s, na, (const int*)nullptr, c.colorOrder.p, c.mActive.p, man,
                     c.slotCount.p);
  primScan(c, c.slotCount.p, c.slotStart.p, na);
  hipLaunchKernelGGL(kBuildRows, dim3(gridFor(na, 128)), dim3(128), 0, s, na, (const int*)nullptr, c.colorOrder.p, c.mActive.p, SP, man,
                     c.tr.p, c.vel.p, c.props.p, (const int*)c.slotStart.p, c.hdr.p, c.nrow.p, c.frow.p,
                     SP.useLearned ? c.learned.p : nullptr, SP.dumpFeatures ? c.features.p : nullptr);
  // per-colour launches: with blocks, the boundary colours [firstColor, ncolors); without, colours
  // [0, headColors) get a launch each and the trailing small ones share one single-workgroup launch
  const int firstColor = c.interiorColors;
  const bool blockSweeps = firstColor > 0;
  int headColors = c.ncolors;
  if (!blocked && c.tailThreshold > 0) {
    while (headColors > 0 && c.colorOff[headColors] - c.colorOff[headColors - 1] <= c.tailThreshold) headColors--;
    if (c.ncolors - headColors < 2) headColors = c.ncolors;
  }
  c.headColors = headColors;
  if (headColors < c.ncolors) {
    std::vector<int> offs(c.colorOff.begin(), c.colorOff.end());
    c.colorOffDev.ensure(offs.size());
    JB_CHECK(hipMemcpyAsync(c.colorOffDev.p, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice, s));
  }
  c.rowCount.ensure(2);
  // rows: all slots, and the slots swept by the timed per-colour launches' counterpart (block sweeps: the
  // interior slots; otherwise the head colours)
  hipLaunchKernelGGL(kRowTotals, dim3(1), dim3(64), 0, s, na, (const int*)nullptr,
                     blockSweeps ? (int)c.interiorEnd : (int)c.colorOff[headColors], (const int*)c.slotCount.p,
                     (const int*)c.slotStart.p, c.rowCount.p);
  if (learnedHere) mlpWarmStart(c, na);  // overwrites the applied words of the valid rows
  const bool apply = c.warmstart || c.useLearned || c.useMlp;
  int setupLaunches = 1;
  if (apply && !blockSweeps && c.ncolors <= kWarmGatherBatches) {  // warm start gathered per body
    c.warmSlotOf.ensure((size_t)c.ncolors * c.nb);
    if (c.headColors >= c.ncolors) {  // the batch offsets on the device (already there when a tail is swept)
      std::vector<int> offs(c.colorOff.begin(), c.colorOff.end());
      c.colorOffDev.ensure(offs.size());
      JB_CHECK(hipMemcpyAsync(c.colorOffDev.p, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice, s));
    }
    JB_CHECK(hipMemsetAsync(c.warmSlotOf.p, 0xff, (size_t)c.ncolors * c.nb * sizeof(int), s));
    hipLaunchKernelGGL(kWarmSlotTable, dim3(gridFor(na)), dim3(256), 0, s, na, c.ncolors, (const int*)c.colorOffDev.p,
                       (const int4*)c.hdr.p, (const float*)c.svel.p, c.nb, c.warmSlotOf.p);
    hipLaunchKernelGGL(kWarmGather, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, c.ncolors,
                       (const int*)c.warmSlotOf.p, (const int4*)c.hdr.p, (const float*)c.nrow.p,
                       (const float*)c.frow.p, (const float*)c.props.p, c.svel.p);
    setupLaunches += 2;
  } else if (apply) {  // warm start colour by colour: the head colours one launch each, the tail in one workgroup
    for (int k = 0; k < headColors; k++) {
      int cnt = (int)(c.colorOff[k + 1] - c.colorOff[k]);
      hipLaunchKernelGGL(kWarmApply, dim3(gridFor(cnt, 64)), dim3(64), 0, s, (int)c.colorOff[k], cnt, c.hdr.p,
                         c.nrow.p, c.frow.p, c.props.p, c.svel.p);
    }
    if (headColors < c.ncolors)
      hipLaunchKernelGGL(kWarmApplyTail, dim3(1), dim3(kWarmTailThreads), 0, s, headColors, c.ncolors,
                         (const int*)c.colorOffDev.p, c.hdr.p, c.nrow.p, c.frow.p, c.props.p, c.svel.p);
    setupLaunches += headColors + (headColors < c.ncolors ? 1 : 0);
  }
  tEnd(c, T_SOLVE_SETUP, setupLaunches);
  // iterations
  // The residual is accumulated only where it is used: every iteration in tolerance mode, else the last one.
  const bool tolMode = c.tolerance > 0.0f;
  int iters = tolMode ? c.maxIterations : c.numIterations;
  const size_t histWords = 2 * (size_t)iters + 4, stripeWords = 2 * (size_t)kResidStripes * iters;
  c.resid.ensure(histWords + stripeWords);
  c.done.ensure(4);
  JB_CHECK(hipMemsetAsync(c.resid.p, 0, (histWords + stripeWords) * sizeof(double), s));
  JB_CHECK(hipMemsetAsync(c.done.p, 0, 4 * sizeof(int), s));
  float* resOut = (float*)(c.counter.p + 6);
  JB_CHECK(hipMemsetAsync(resOut, 0, 2 * sizeof(float), s));
  int* done = tolMode ? c.done.p : nullptr;
  // kernel-level timing of this step's sweeps (sampled): on a sampled step, the launches of every
  // kernelStride-th iteration get their own events (every iteration sweeps the same colours and rows; the events
  // cost the step about 10 us per launch, so timing all of them would stretch the sampled step)
  const bool sampleK = c.kernelEvery > 0 && c.kernelSteps++ % c.kernelEvery == 0;
  bool sampleIt = false;
  int sampledIters = 0;
  size_t kUsed = 0;
  auto kPair = [&](int kind) -> hipEvent_t* {
    if (kUsed + 2 > c.kEv.size()) {
      const size_t old = c.kEv.size();
      c.kEv.resize(old + 256);
      for (size_t e = old; e < c.kEv.size(); e++) JB_CHECK(hipEventCreate(&c.kEv[e]));
    }
    c.kEvKind.resize(c.kEv.size() / 2);
    c.kEvKind[kUsed / 2] = kind;
    kUsed += 2;
    return &c.kEv[kUsed - 2];
  };
  const size_t blockLds = (size_t)B * 8 * sizeof(float);
  auto sweepBlocks = [&](bool friction, double* res) {
    if (!blockSweeps) return;
    const int nb = c.nBlocks, nbt = c.interiorColors;
    const dim3 grid(nb), block(kBlockThreads);
    const int4* hd = c.hdr.p;
    const int2* lo = c.slotLocal.p;
    const float* pr = c.props.p;
    const int* rs = c.runStart.p;
    const int* re = c.runEnd.p;
    const int* dn = done;
    if (sampleIt) {
      hipEvent_t* ev = kPair(friction ? 3 : 2);
      if (friction)
        hipExtLaunchKernelGGL(kSweepBlocks<true>, grid, block, blockLds, s, ev[0], ev[1], 0, nbt, nb, B, c.np,
                              c.sortedOrder, rs, re, SP, hd, lo, c.nrow.p, c.frow.p, pr, c.svel.p, res, dn);
      else
        hipExtLaunchKernelGGL(kSweepBlocks<false>, grid, block, blockLds, s, ev[0], ev[1], 0, nbt, nb, B, c.np,
                              c.sortedOrder, rs, re, SP, hd, lo, c.nrow.p, c.frow.p, pr, c.svel.p, res, dn);
    } else if (friction) {
      hipLaunchKernelGGL(kSweepBlocks<true>, grid, block, blockLds, s, nbt, nb, B, c.np, c.sortedOrder, rs, re, SP, hd,
                         lo, c.nrow.p, c.frow.p, pr, c.svel.p, res, dn);
    } else {
      hipLaunchKernelGGL(kSweepBlocks<false>, grid, block, blockLds, s, nbt, nb, B, c.np, c.sortedOrder, rs, re, SP,
                         hd, lo, c.nrow.p, c.frow.p, pr, c.svel.p, res, dn);
    }
  };
  tBegin(c, T_SOLVE_ITER);
  int jointLaunches = 0;
  int launchedIters = iters;
  if (tolMode) c.doneHost.ensure(1);
  for (int itn = 0; itn < iters; itn++) {
    const bool track = tolMode || itn == iters - 1;
    sampleIt = sampleK && itn % c.kernelStride == 0;
    sampledIters += sampleIt ? 1 : 0;
    double* res = track ? c.resid.p + histWords + 2 * (size_t)kResidStripes * itn : nullptr;
    jointLaunches += jp.sweep(dt, done);  // the constraints first in every iteration (:823-842)
    sweepBlocks(false, res);
    for (int k = firstColor; k < headColors; k++) {
      int cnt = (int)(c.colorOff[k + 1] - c.colorOff[k]);
      if (sampleIt) {
        hipEvent_t* ev = kPair(0);
        hipExtLaunchKernelGGL(kNormal, dim3(gridFor(4 * (int64_t)cnt, kSweepThreads)), dim3(kSweepThreads), 0, s, ev[0], ev[1], 0, (int)c.colorOff[k],
                              cnt, SP, (const int4*)c.hdr.p, c.nrow.p, (const float*)c.props.p, c.svel.p, res,
                              (const int*)done);
      } else {
        hipLaunchKernelGGL(kNormal, dim3(gridFor(4 * (int64_t)cnt, kSweepThreads)), dim3(kSweepThreads), 0, s, (int)c.colorOff[k], cnt, SP, c.hdr.p,
                           c.nrow.p, c.props.p, c.svel.p, res, done);
      }
    }
    if (headColors < c.ncolors)
      hipLaunchKernelGGL(kNormalTail, dim3(1), dim3(kTailThreads), 0, s, headColors, c.ncolors,
                         (const int*)c.colorOffDev.p, SP, (const int4*)c.hdr.p, c.nrow.p, (const float*)c.props.p,
                         c.svel.p, res, (const int*)done);
    sweepBlocks(true, res);
    for (int k = firstColor; k < headColors; k++) {
      int cnt = (int)(c.colorOff[k + 1] - c.colorOff[k]);
      if (sampleIt) {
        hipEvent_t* ev = kPair(1);
        hipExtLaunchKernelGGL(kFriction, dim3(gridFor(4 * (int64_t)cnt, kSweepThreads)), dim3(kSweepThreads), 0, s, ev[0], ev[1], 0,
                              (int)c.colorOff[k], cnt, SP, (const int4*)c.hdr.p, (const float*)c.nrow.p, c.frow.p,
                              (const float*)c.props.p, c.svel.p, res, (const int*)done);
      } else {
        hipLaunchKernelGGL(kFriction, dim3(gridFor(4 * (int64_t)cnt, kSweepThreads)), dim3(kSweepThreads), 0, s, (int)c.colorOff[k], cnt, SP,
                           c.hdr.p, c.nrow.p, c.frow.p, c.props.p, c.svel.p, res, done);
      }
    }
    if (headColors < c.ncolors)
      hipLaunchKernelGGL(kFrictionTail, dim3(1), dim3(kTailThreads), 0, s, headColors, c.ncolors,
                         (const int*)c.colorOffDev.p, SP, (const int4*)c.hdr.p, (const float*)c.nrow.p, c.frow.p,
                         (const float*)c.props.p, c.svel.p, res, (const int*)done);
    if (track)
      hipLaunchKernelGGL(kResidual, dim3(1), dim3(64), 0, s, itn, c.tolerance, res, c.resid.p, c.done.p, resOut);
    // tolerance mode: every few iterations the host reads the done flag and stops launching once the residual
    // is reached (the kernels of later iterations would only read the flag and exit)
    if (tolMode && c.tolCheckEvery > 0 && (itn + 1) % c.tolCheckEvery == 0 && itn + 1 < iters) {
      JB_CHECK(hipMemcpyAsync(c.doneHost.p, c.done.p, sizeof(int), hipMemcpyDeviceToHost, s));
      JB_CHECK(hipStreamSynchronize(s));
      if (*c.doneHost.p) {
        launchedIters = itn + 1;
        break;
      }
    }
  }
  JB_CHECK(hipGetLastError());
  tEnd(c, T_SOLVE_ITER,
       (int64_t)launchedIters * 2 *
               (headColors - firstColor + (headColors < c.ncolors ? 1 : 0) + (blockSweeps ? 1 : 0)) +
           jointLaunches);
  hipLaunchKernelGGL(kWriteback, dim3(gridFor(na)), dim3(256), 0, s, na, (const int*)nullptr, c.hdr.p, c.nrow.p, c.frow.p, man);
  hipLaunchKernelGGL(kCopyVel, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, c.svel.p, c.vel.p, c.props.p);
  // statistics: copied into pinned memory now, read by solverCollectStats after the step's final wait
  c.solveRo.ensure(2);
  c.solveRh.ensure(2 * (size_t)std::max(iters, 1));
  c.solveRows.ensure(2);
  JB_CHECK(hipMemcpyAsync(c.solveRo.p, resOut, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
  if (iters > 0)
    JB_CHECK(hipMemcpyAsync(c.solveRh.p, c.resid.p, 2 * (size_t)iters * sizeof(double), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(c.solveRows.p, c.rowCount.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  c.solveStat.pending = true;
  c.solveStat.tolMode = tolMode;
  c.solveStat.blockSweeps = blockSweeps;
  c.solveStat.sampleK = sampleK;
  c.solveStat.sampledIters = sampledIters;
  c.solveStat.iters = iters;
  c.solveStat.kUsed = kUsed;
}

void solverCollectStats(Ctx& c) {
  if (!c.solveStat.pending) return;
  c.solveStat.pending = false;
  const bool tolMode = c.solveStat.tolMode, blockSweeps = c.solveStat.blockSweeps;
  const int iters = c.solveStat.iters;
  const float* ro = c.solveRo.p;
  const double* rh = c.solveRh.p;
  const unsigned long long* rowsHT = c.solveRows.p;
  c.st.rows = (int64_t)rowsHT[0];
  // rows swept per sweep by the per-colour launches (kinds 0, 1) and by the block launches (kinds 2, 3)
  const int64_t colourRows = blockSweeps ? (int64_t)(rowsHT[0] - rowsHT[1]) : (int64_t)rowsHT[1];
  const int64_t blockRows = blockSweeps ? (int64_t)rowsHT[1] : 0;
  if (c.solveStat.sampleK) {
    for (size_t e = 0; e < c.solveStat.kUsed; e += 2) {
      float ms = 0.0f;
      JB_CHECK(hipEventElapsedTime(&ms, c.kEv[e], c.kEv[e + 1]));
      const int kind = c.kEvKind[e / 2];
      c.kMs[kind] += ms;
      c.kLaunches[kind] += 1;
    }
    const int si = c.solveStat.sampledIters;
    c.kRows[0] += colourRows * si;
    c.kRows[1] += 2 * colourRows * si;
    c.kRows[2] += blockRows * si;
    c.kRows[3] += 2 * blockRows * si;
  }
  c.st.residual = ro[0];
  c.st.iterations = (int64_t)ro[1];
  if (c.wbLastSolve) {
    const int* wi = c.wbInfoHost.p;
    c.st.colors = wi[1];
    if (wi[0] & 2) throw std::length_error("a block needed more than 128 colours");
    if (wi[0] & 4) throw std::logic_error("a manifold joins two world blocks (bodies of different worlds)");
  }
  for (int k = tolMode ? 0 : std::max(iters - 1, 0); k < (int)c.st.iterations; k++) {
    double nrm = std::sqrt(rh[2 * k + 1]);
    c.residHist.push_back((float)(std::sqrt(rh[2 * k]) / (nrm > 1.0 ? nrm : 1.0)));
  }
}

// ---- the solve of a graph-captured step: batched worlds, one block solve, counts on the device ------------
void solverGraphReserve(Ctx& c) {
  if (!c.wbReady) worldBlocksPrepare(c);
  const size_t cap = (size_t)std::max<int64_t>(c.pairCap, 1);
  const int iters = c.numIterations;
  c.flags.ensure(cap);
  c.mActive.ensure(cap);
  c.counter.ensure(8);
  c.activeBodies.ensure(cap);
  c.wbKey.ensure(cap);
  c.wbKeyAlt.ensure(cap);
  c.wbIdx.ensure(cap);
  c.colorOrderAlt.ensure(cap);
  c.wbSlotStart.ensure((size_t)c.wbBlocks + 1);
  c.wbInfo.ensure(4);
  c.hdr.ensure(cap);
  c.nrow.ensure((size_t)NW * cap);
  c.frow.ensure((size_t)FW * cap);
  c.svel.ensure((size_t)8 * c.nb);
  c.slotCount.ensure(cap);
  c.slotStart.ensure(cap);
  c.rowCount.ensure(2);
  c.resid.ensure(2 * (size_t)iters + 4 + 2 * (size_t)kResidStripes * iters);
  c.done.ensure(4);
  c.wbColour.ensure(cap);
  c.wbSolveSlot.ensure(cap);
  c.wbSlotInfo.ensure(cap);
  c.colorOrder.ensure(cap);
  c.solveRo.ensure(2);
  c.solveRh.ensure(2 * (size_t)std::max(iters, 1));
  c.solveRows.ensure(2);
  c.wbInfoHost.ensure(4);
  primReserve(c, (int)cap);
  if ((int64_t)cap * 4 >= (1ll << 27)) throw std::length_error("pair capacity too large for the slot header");
}

void solverSolveGraph(Ctx& c, float dt) {
  hipStream_t s = c.stream;
  const int cap = (int)c.pairCap;
  int* gc = c.gc.p;
  float* man = c.man[c.cur].p;
  c.nActive = 0;
  hipLaunchKernelGGL(kMarkActive, dim3(gridFor(cap)), dim3(256), 0, s, (int64_t)cap, (const int*)(gc + kGcPairs), man,
                     c.ptype.p, c.act.p, c.props.p, (const int*)nullptr, (const int*)nullptr, c.flags.p);
  primSelectIndexDev(c, c.flags.p, gc + kGcPairs, cap, c.mActive.p, gc + kGcActive);
  solveWorldBlocks(c, dt, cap, gc + kGcActive);
}

void solverStatsCopy(Ctx& c) {
  hipStream_t s = c.stream;
  const int iters = c.numIterations;
  JB_CHECK(hipMemcpyAsync(c.solveRo.p, (float*)(c.counter.p + 6), 2 * sizeof(float), hipMemcpyDeviceToHost, s));
  if (iters > 0)
    JB_CHECK(hipMemcpyAsync(c.solveRh.p, c.resid.p, 2 * (size_t)iters * sizeof(double), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(c.solveRows.p, c.rowCount.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(c.wbInfoHost.p, c.wbInfo.p, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
  c.solveStat.pending = true;
  c.solveStat.tolMode = false;
  c.solveStat.blockSweeps = false;
  c.solveStat.sampleK = false;
  c.solveStat.iters = iters;
  c.solveStat.kUsed = 0;
  c.wbLastSolve = true;
}

}  // namespace jbg
