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
// share no dynamic body; one thread handles one manifold's points in order,
// and a colour is one launch.  Processing colour after colour is then exactly
// the sequential solver fed the manifold order "colour 0 manifolds, colour 1
// manifolds, ..." (any order inside a colour gives the same bits), which the
// library reports so the CPU oracle can be fed the same order.
#include <cmath>

#include <hipcub/hipcub.hpp>

#include "internal.h"
#include "shapes.h"

namespace jbg {

namespace {

constexpr int RW = kRowWords;          // words per row
constexpr int CW = kContactRowWords;   // words per contact (normal + 2 friction rows)
constexpr int MW = kManifoldRowWords;  // words per manifold

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
__global__ void kMarkActive(int64_t npairs, const float* __restrict__ man, const int* __restrict__ ptype,
                            const int* __restrict__ act, const float* __restrict__ props, int* flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * i;
  int n = ldi(m), b0 = ldi(m + 1), b1 = ldi(m + 2);
  int f = 0;
  if (ptype[i] != JBGPU_PAIR_NONE && n > 0) {
    bool s0 = (ldi(props + 8 * b0 + 7) & JBGPU_BODY_STATIC) != 0, s1 = (ldi(props + 8 * b1 + 7) & JBGPU_BODY_STATIC) != 0;
    bool awake = (act[b0] != JBGPU_ISLAND_SLEEPING) || (act[b1] != JBGPU_ISLAND_SLEEPING);
    bool solvable = isActiveState(act[b0]) || isActiveState(act[b1]);
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

__device__ __forceinline__ void bodiesOf(const float* man, const int* list, int k, int& a, int& b) {
  const float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * list[k];
  a = ldi(m + 1);
  b = ldi(m + 2);
}
__device__ __forceinline__ bool dynamicBody(const float* props, int i) {
  return (ldi(props + 8 * i + 7) & JBGPU_BODY_STATIC) == 0;
}

__global__ void kColorPropose(int n, int round, const int* __restrict__ list, const float* __restrict__ man,
                              const float* __restrict__ props, const int* __restrict__ color,
                              unsigned long long* bodyPrio) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || color[k] >= 0) return;
  int a, b;
  bodiesOf(man, list, k, a, b);
  unsigned long long p = ((unsigned long long)round << 52) | ((unsigned long long)(hash32(k) >> 12) << 32) |
                         (unsigned long long)k;
  if (dynamicBody(props, a)) atomicMax(bodyPrio + a, p);
  if (dynamicBody(props, b)) atomicMax(bodyPrio + b, p);
}

__global__ void kColorResolve(int n, int round, const int* __restrict__ list, const float* __restrict__ man,
                              const float* __restrict__ props, int* color, const unsigned long long* bodyPrio,
                              unsigned long long* bodyMask, unsigned long long* remaining) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || color[k] >= 0) return;
  int a, b;
  bodiesOf(man, list, k, a, b);
  unsigned long long p = ((unsigned long long)round << 52) | ((unsigned long long)(hash32(k) >> 12) << 32) |
                         (unsigned long long)k;
  bool da = dynamicBody(props, a), db = dynamicBody(props, b);
  bool win = (!da || bodyPrio[a] == p) && (!db || bodyPrio[b] == p);
  if (!win) {
    atomicAdd(remaining, 1ull);
    return;
  }
  unsigned long long used = (da ? bodyMask[a] : 0ull) | (db ? bodyMask[b] : 0ull);
  int c;
  if (used == ~0ull) {
    c = 64 + round;
  } else {
    c = __ffsll((long long)~used) - 1;
    if (da) bodyMask[a] |= (1ull << c);
    if (db) bodyMask[b] |= (1ull << c);
  }
  color[k] = c;
}

__global__ void kIota(int n, int* out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = k;
}

__global__ void kColorCount(int n, const int* __restrict__ color, unsigned long long* counts) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  atomicAdd(counts + color[k], 1ull);
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

struct SBody {
  V3 lv, av;
  float invMass;
};
__device__ __forceinline__ SBody loadS(const float* v, const float* props, int i) {
  SBody s;
  s.lv = ld3(v + 8 * i);
  s.av = ld3(v + 8 * i + 4);
  s.invMass = props[8 * i];
  return s;
}
__device__ __forceinline__ void storeS(float* v, int i, const SBody& s) {
  st3(v + 8 * i, s.lv);
  st3(v + 8 * i + 4, s.av);
}
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

// one thread per manifold of colour c (solveGroupCacheFriendlySetup :494-747)
__global__ void kSetup(int cnt, const int* __restrict__ corder, const int* __restrict__ list, SolverParams SP,
                       float* man, const float* __restrict__ tr, const float* __restrict__ vel,
                       const float* __restrict__ props, float* svel, float* rows, const float* __restrict__ learned,
                       float* features) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  int ai = corder[k];
  int pi = list[ai];
  float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * pi;
  int n = ldi(m), b0 = ldi(m + 1), b1 = ldi(m + 2);
  float* R = rows + (size_t)MW * ai;
  M3 I0 = invInertiaWorld(tr, props, b0), I1 = invInertiaWorld(tr, props, b1);
  V3 o0 = ld3(tr + 12 * b0 + 9), o1 = ld3(tr + 12 * b1 + 9);
  V3 lv0 = ld3(vel + 8 * b0), av0 = ld3(vel + 8 * b0 + 4), lv1 = ld3(vel + 8 * b1), av1 = ld3(vel + 8 * b1 + 4);
  float im0 = props[8 * b0], im1 = props[8 * b1];
  SBody A = loadS(svel, props, b0), B = loadS(svel, props, b1);
  int mask = 0;
  for (int j = 0; j < n; j++) {
    float* P = m + JBGPU_MANIFOLD_HDR + JBGPU_POINT_WORDS * j;
    if (!(P[15] <= 0.0f)) continue;
    mask |= 1 << j;
    float* N = R + CW * j;
    V3 normal = ld3(P + 12);
    V3 rp1 = sub(ld3(P + 6), o0);
    V3 rp2 = sub(ld3(P + 9), o1);
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
    float pen = jmin(P[15] + SP.linearSlop, 0.0f);
    float friction = P[16];
    float rest = P[17] * -rel_vel;
    if (rest <= 0.0f) rest = 0.0f;
    float penVel = -pen / SP.dt;
    if (rest > penVel) pen = 0.0f;
    N[16] = pen;
    N[17] = rest;
    N[18] = friction;
    const int slot = ai * 4 + j;
    if (SP.dumpFeatures) {
      float* F = features + (size_t)16 * slot;
      F[0] = normal.x; F[1] = normal.y; F[2] = normal.z;
      F[3] = vv.x; F[4] = vv.y; F[5] = vv.z;
      F[6] = av0.x - av1.x; F[7] = av0.y - av1.y; F[8] = av0.z - av1.z;
      F[9] = P[15]; F[10] = im0; F[11] = im1;
      F[12] = P[18]; F[13] = P[19]; F[14] = P[20]; F[15] = friction;
    }
    float applied;
    if (SP.useLearned) applied = learned[3 * slot];
    else if (SP.warm) applied = P[18] * SP.warmFactor;
    else applied = 0.0f;
    N[19] = applied;
    if (SP.useLearned || SP.warm) {
      applyImpulse(A, scl(normal, im0), angA, applied);
      applyImpulse(B, scl(normal, im1), angB, -applied);
    }
    // lateral friction directions (:678-708)
    float* F1 = N + RW;
    float* F2 = N + 2 * RW;
    if (ldi(P + 27) == 0) {
      V3 d1v = sub(vv, scl(normal, rel_vel));
      float lat = len2(d1v);
      V3 d2v;
      if (lat > kFltEps) {
        d1v = scl(d1v, 1.0f / jsqrt(lat));
        d2v = nor(crs(d1v, normal));
      } else {
        planeSpace1(normal, d1v, d2v);
      }
      st3(P + 21, d1v);
      st3(P + 24, d2v);
      sti(P + 27, 1);
    }
    frictionRow(F1, ld3(P + 21), rp1, rp2, I0, I1, im0, im1, friction);
    frictionRow(F2, ld3(P + 24), rp1, rp2, I0, I1, im0, im1, friction);
    for (int f = 0; f < 2; f++) {  // friction warm start (:710-747)
      float* FR = f == 0 ? F1 : F2;
      float fa;
      if (SP.useLearned) fa = learned[3 * slot + 1 + f];
      else if (SP.warm) fa = P[19 + f] * SP.warmFactor;
      else continue;
      FR[19] = fa;
      applyImpulse(A, scl(ld3(FR), im0), ld3(FR + 9), fa);
      applyImpulse(B, scl(ld3(FR), im1), ld3(FR + 12), -fa);
    }
  }
  R[4 * CW + 0] = __builtin_bit_cast(float, mask);
  R[4 * CW + 1] = __builtin_bit_cast(float, b0);
  R[4 * CW + 2] = __builtin_bit_cast(float, b1);
  if (A.invMass != 0.0f) storeS(svel, b0, A);
  if (B.invMass != 0.0f) storeS(svel, b1, B);
}

__device__ __forceinline__ void accumulate(double dd, double nn, double* res) {
  for (int off = 32; off > 0; off >>= 1) {
    dd += __shfl_xor(dd, off);
    nn += __shfl_xor(nn, off);
  }
  if ((threadIdx.x & 63) == 0 && (dd != 0.0 || nn != 0.0)) {
    atomicAdd(res, dd);
    atomicAdd(res + 1, nn);
  }
}

// normal rows of colour c: resolveSingleCollisionCombinedCacheFriendly (:234-284)
__global__ void kNormal(int cnt, const int* __restrict__ corder, SolverParams SP, float* rows,
                        const float* __restrict__ props, float* svel, double* res, const int* done) {
  if (*done) return;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  double dd = 0.0, nn = 0.0;
  if (k < cnt) {
    float* R = rows + (size_t)MW * corder[k];
    int mask = __builtin_bit_cast(int, R[4 * CW]);
    int b0 = __builtin_bit_cast(int, R[4 * CW + 1]), b1 = __builtin_bit_cast(int, R[4 * CW + 2]);
    SBody A = loadS(svel, props, b0), B = loadS(svel, props, b1);
    for (int j = 0; j < 4; j++) {
      if (!(mask & (1 << j))) continue;
      float* N = R + CW * j;
      V3 nrm = ld3(N);
      float v1 = dot(nrm, A.lv) + dot(ld3(N + 3), A.av);
      float v2 = dot(nrm, B.lv) + dot(ld3(N + 6), B.av);
      float rel_vel = v1 - v2;
      float positionalError = -N[16] * SP.erp / SP.dt;
      float velocityError = N[17] - rel_vel;
      float pi = positionalError * N[15];
      float vi = velocityError * N[15];
      float ni = pi + vi;
      float old = N[19];
      float sum = old + ni;
      float nw = 0.0f > sum ? 0.0f : sum;
      N[19] = nw;
      ni = nw - old;
      applyImpulse(A, scl(nrm, A.invMass), ld3(N + 9), ni);
      applyImpulse(B, scl(nrm, B.invMass), ld3(N + 12), -ni);
      double d = (double)nw - (double)old;
      dd += d * d;
      nn += (double)nw * (double)nw;
    }
    if (A.invMass != 0.0f) storeS(svel, b0, A);
    if (B.invMass != 0.0f) storeS(svel, b1, B);
  }
  accumulate(dd, nn, res);
}

// friction rows of colour c: resolveSingleFrictionCacheFriendly (:286-346)
__global__ void kFriction(int cnt, const int* __restrict__ corder, SolverParams SP, float* rows,
                          const float* __restrict__ props, float* svel, double* res, const int* done) {
  if (*done) return;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  double dd = 0.0, nn = 0.0;
  if (k < cnt) {
    float* R = rows + (size_t)MW * corder[k];
    int mask = __builtin_bit_cast(int, R[4 * CW]);
    int b0 = __builtin_bit_cast(int, R[4 * CW + 1]), b1 = __builtin_bit_cast(int, R[4 * CW + 2]);
    SBody A = loadS(svel, props, b0), B = loadS(svel, props, b1);
    for (int j = 0; j < 4; j++) {
      if (!(mask & (1 << j))) continue;
      float total = R[CW * j + 19];  // appliedImpulse of this contact's normal row (+ push impulse 0)
      total = total + 0.0f;
      for (int f = 1; f <= 2; f++) {
        float* N = R + CW * j + RW * f;
        float old = N[19];
        float limit = total * N[18];
        if (total > 0.0f) {
          V3 nrm = ld3(N);
          float v1 = dot(nrm, A.lv) + dot(ld3(N + 3), A.av);
          float v2 = dot(nrm, B.lv) + dot(ld3(N + 6), B.av);
          float rel_vel = v1 - v2;
          float j1 = -rel_vel * N[15];
          float ap = old + j1;
          if (limit < ap) {
            ap = limit;
          } else if (ap < -limit) {
            ap = -limit;
          }
          N[19] = ap;
          j1 = ap - old;
          applyImpulse(A, scl(nrm, A.invMass), ld3(N + 9), j1);
          applyImpulse(B, scl(nrm, B.invMass), ld3(N + 12), -j1);
        }
        double d = (double)N[19] - (double)old;
        dd += d * d;
        nn += (double)N[19] * (double)N[19];
      }
    }
    if (A.invMass != 0.0f) storeS(svel, b0, A);
    if (B.invMass != 0.0f) storeS(svel, b1, B);
  }
  accumulate(dd, nn, res);
}

// residual after iteration `it`: r = ||dl|| / max(1, ||l||); stop when r <= tol
__global__ void kResidual(int it, float tol, double* res, int* done, float* out) {
  if (*done) return;
  double dd = res[2 * it], nn = res[2 * it + 1];
  double nrm = sqrt(nn);
  float r = (float)(sqrt(dd) / (nrm > 1.0 ? nrm : 1.0));
  out[0] = r;
  out[1] = (float)(it + 1);
  if (tol > 0.0f && r <= tol) *done = 1;
}

__global__ void kWriteback(int n, const int* __restrict__ list, const float* __restrict__ rows, float* man) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float* R = rows + (size_t)MW * k;
  int mask = __builtin_bit_cast(int, R[4 * CW]);
  float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * list[k];
  for (int j = 0; j < 4; j++) {
    if (!(mask & (1 << j))) continue;
    float* P = m + JBGPU_MANIFOLD_HDR + JBGPU_POINT_WORDS * j;
    P[18] = R[CW * j + 19];
    P[19] = R[CW * j + RW + 19];
    P[19] = R[CW * j + 2 * RW + 19];  // appliedImpulseLateral1 written twice; Lateral2 never stored
  }
}

__global__ void kCopyVel(int n, const float* __restrict__ src, float* dst, const float* __restrict__ props) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (props[8 * i] != 0.0f) {
    const float4* s = (const float4*)(src + 8 * i);
    float4* d = (float4*)(dst + 8 * i);
    d[0] = s[0];
    d[1] = s[1];
  }
}

}  // namespace

void solverSolve(Ctx& c, float dt) {
  hipStream_t s = c.stream;
  const int64_t np = c.npairs;
  c.nActive = 0;
  c.ncolors = 0;
  c.st.rows = 0;
  c.st.colors = 0;
  c.st.iterations = 0;
  c.st.residual = 0.0f;
  if (np == 0) return;
  const float* man = c.man[c.cur].p;
  // active manifolds (stable compaction keeps pair-key order)
  c.flags.ensure(np);
  c.mActive.ensure(np);
  c.counter.ensure(8);
  hipLaunchKernelGGL(kMarkActive, dim3(gridFor(np)), dim3(256), 0, s, np, man, c.ptype.p, c.act.p, c.props.p,
                     c.flags.p);
  size_t tb = 0;
  hipcub::CountingInputIterator<int> it(0);
  JB_CHECK(hipcub::DeviceSelect::Flagged(nullptr, tb, it, c.flags.p, c.mActive.p, (int*)c.counter.p, (int)np, s));
  c.cubTmp.ensure(tb);
  JB_CHECK(hipMemsetAsync(c.counter.p, 0, 8 * sizeof(unsigned long long), s));
  JB_CHECK(hipcub::DeviceSelect::Flagged(c.cubTmp.p, tb, it, c.flags.p, c.mActive.p, (int*)c.counter.p, (int)np, s));
  int na = 0;
  JB_CHECK(hipMemcpyAsync(&na, c.counter.p, sizeof(int), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  c.nActive = na;
  if (na == 0) return;
  if (c.useLearned && c.learnedSlots < 4 * (int64_t)na)
    throw std::length_error("learned warm start has fewer slots than 4 x active manifolds");
  // colouring
  c.color.ensure(na);
  c.colorAlt.ensure(na);
  c.colorOrder.ensure(na);
  c.colorOrderAlt.ensure(na);
  c.bodyPrio.ensure(c.nb);
  c.bodyMask.ensure(c.nb);
  JB_CHECK(hipMemsetAsync(c.color.p, 0xff, (size_t)na * sizeof(int), s));
  JB_CHECK(hipMemsetAsync(c.bodyPrio.p, 0, (size_t)c.nb * sizeof(unsigned long long), s));
  JB_CHECK(hipMemsetAsync(c.bodyMask.p, 0, (size_t)c.nb * sizeof(unsigned long long), s));
  int round = 0;
  for (;;) {
    JB_CHECK(hipMemsetAsync(c.counter.p + 1, 0, sizeof(unsigned long long), s));
    for (int r = 0; r < 4; r++, round++) {
      hipLaunchKernelGGL(kColorPropose, dim3(gridFor(na)), dim3(256), 0, s, na, round, c.mActive.p, man, c.props.p,
                         c.color.p, c.bodyPrio.p);
      if (r == 3) JB_CHECK(hipMemsetAsync(c.counter.p + 1, 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(kColorResolve, dim3(gridFor(na)), dim3(256), 0, s, na, round, c.mActive.p, man, c.props.p,
                         c.color.p, c.bodyPrio.p, c.bodyMask.p, c.counter.p + 1);
    }
    unsigned long long rem = 0;
    JB_CHECK(hipMemcpyAsync(&rem, c.counter.p + 1, sizeof(rem), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipStreamSynchronize(s));
    if (rem == 0) break;
    if (round > 4000) throw std::runtime_error("colouring did not converge");
  }
  // order by colour, stable in active index
  int maxColor = 64 + round + 1;
  std::vector<unsigned long long> counts(maxColor, 0);
  DBuf<unsigned long long>& cc = c.bodyPrio;  // reuse as colour histogram
  cc.ensure(std::max<int64_t>(c.nb, maxColor));
  JB_CHECK(hipMemsetAsync(cc.p, 0, maxColor * sizeof(unsigned long long), s));
  hipLaunchKernelGGL(kColorCount, dim3(gridFor(na)), dim3(256), 0, s, na, c.color.p, cc.p);
  {
    hipLaunchKernelGGL(kIota, dim3(gridFor(na)), dim3(256), 0, s, na, c.colorOrder.p);
    JB_CHECK(hipMemcpyAsync(c.colorAlt.p, c.color.p, (size_t)na * sizeof(int), hipMemcpyDeviceToDevice, s));
    int bitsC = 1;
    while ((1 << bitsC) <= maxColor) bitsC++;
    c.typeKeyAlt.ensure(na);
    c.permAlt.ensure(na);
    hipcub::DoubleBuffer<int> dk(c.colorAlt.p, c.typeKeyAlt.p);
    hipcub::DoubleBuffer<int> dv(c.colorOrder.p, c.permAlt.p);
    size_t tb2 = 0;
    JB_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, dk, dv, na, 0, bitsC, s));
    c.cubTmp.ensure(tb2);
    JB_CHECK(hipcub::DeviceRadixSort::SortPairs(c.cubTmp.p, tb2, dk, dv, na, 0, bitsC, s));
    if (dv.Current() != c.colorOrder.p)
      JB_CHECK(hipMemcpyAsync(c.colorOrder.p, dv.Current(), (size_t)na * sizeof(int), hipMemcpyDeviceToDevice, s));
  }
  JB_CHECK(hipMemcpyAsync(counts.data(), cc.p, maxColor * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  c.colorOff.assign(1, 0);
  for (int k = 0; k < maxColor; k++)
    if (counts[k]) c.colorOff.push_back(c.colorOff.back() + (int64_t)counts[k]);
  c.ncolors = (int)c.colorOff.size() - 1;
  c.st.colors = c.ncolors;
  // setup
  tBegin(c, T_SOLVE_SETUP);
  c.rows.ensure((size_t)MW * na);
  c.svel.ensure((size_t)8 * c.nb);
  JB_CHECK(hipMemcpyAsync(c.svel.p, c.vel.p, (size_t)8 * c.nb * sizeof(float), hipMemcpyDeviceToDevice, s));
  SolverParams SP{dt, c.erp, c.warmFactor, c.linearSlop, c.warmstart, c.useLearned ? 1 : 0, c.dumpFeatures ? 1 : 0,
                  c.tolerance};
  if (c.dumpFeatures) c.features.ensure((size_t)16 * 4 * na);
  for (int k = 0; k < c.ncolors; k++) {
    int cnt = (int)(c.colorOff[k + 1] - c.colorOff[k]);
    hipLaunchKernelGGL(kSetup, dim3(gridFor(cnt, 128)), dim3(128), 0, s, cnt, c.colorOrder.p + c.colorOff[k],
                       c.mActive.p, SP, c.man[c.cur].p, c.tr.p, c.vel.p, c.props.p, c.svel.p, c.rows.p,
                       c.useLearned ? c.learned.p : nullptr, c.dumpFeatures ? c.features.p : nullptr);
  }
  tEnd(c, T_SOLVE_SETUP, c.ncolors);
  // iterations
  int iters = c.tolerance > 0.0f ? c.maxIterations : c.numIterations;
  c.resid.ensure(2 * (size_t)iters + 4);
  c.done.ensure(4);
  JB_CHECK(hipMemsetAsync(c.resid.p, 0, (2 * (size_t)iters + 4) * sizeof(double), s));
  JB_CHECK(hipMemsetAsync(c.done.p, 0, 4 * sizeof(int), s));
  float* resOut = (float*)(c.counter.p + 6);
  tBegin(c, T_SOLVE_ITER);
  for (int itn = 0; itn < iters; itn++) {
    double* res = c.resid.p + 2 * itn;
    for (int k = 0; k < c.ncolors; k++) {
      int cnt = (int)(c.colorOff[k + 1] - c.colorOff[k]);
      hipLaunchKernelGGL(kNormal, dim3(gridFor(cnt, 128)), dim3(128), 0, s, cnt, c.colorOrder.p + c.colorOff[k], SP,
                         c.rows.p, c.props.p, c.svel.p, res, c.done.p);
    }
    for (int k = 0; k < c.ncolors; k++) {
      int cnt = (int)(c.colorOff[k + 1] - c.colorOff[k]);
      hipLaunchKernelGGL(kFriction, dim3(gridFor(cnt, 128)), dim3(128), 0, s, cnt, c.colorOrder.p + c.colorOff[k],
                         SP, c.rows.p, c.props.p, c.svel.p, res, c.done.p);
    }
    hipLaunchKernelGGL(kResidual, dim3(1), dim3(1), 0, s, itn, c.tolerance, c.resid.p, c.done.p, resOut);
  }
  tEnd(c, T_SOLVE_ITER, (int64_t)iters * 2 * c.ncolors);
  hipLaunchKernelGGL(kWriteback, dim3(gridFor(na)), dim3(256), 0, s, na, c.mActive.p, c.rows.p, c.man[c.cur].p);
  hipLaunchKernelGGL(kCopyVel, dim3(gridFor(c.nb)), dim3(256), 0, s, c.nb, c.svel.p, c.vel.p, c.props.p);
  float ro[2] = {0, 0};
  std::vector<double> rh(2 * (size_t)iters);
  JB_CHECK(hipMemcpyAsync(ro, resOut, sizeof(ro), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(rh.data(), c.resid.p, rh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  c.st.residual = ro[0];
  c.st.iterations = (int64_t)ro[1];
  c.residHist.clear();
  for (int k = 0; k < (int)c.st.iterations; k++) {
    double nrm = std::sqrt(rh[2 * k + 1]);
    c.residHist.push_back((float)(std::sqrt(rh[2 * k]) / (nrm > 1.0 ? nrm : 1.0)));
  }
  c.st.rows = 0;
}

}  // namespace jbg
