// np.hip — narrowphase kernel set for gfx950: the collision dispatcher's work
// over the overlapping-pair list, binned by shape-pair type.
//
//   sphere-sphere  SphereSphereCollisionAlgorithm.processCollision (collision/dispatch/
//                  SphereSphereCollisionAlgorithm.java:73-134)
//   convex-plane   ConvexPlaneCollisionAlgorithm.processCollision (:75-136)
//   convex-convex  GjkPairDetector.getClosestPoints (collision/narrowphase/
//                  GjkPairDetector.java:73-316) with the Voronoi simplex solver
//                  (VoronoiSimplexSolver.java), falling back to GjkEpaSolver
//                  (GjkEpaSolver.java:347-911) for penetrations
//   manifold       ManifoldResult.addContactPoint / refreshContactPoints
//                  (ManifoldResult.java:92-191) on PersistentManifold
//                  (PersistentManifold.java:83-372)
// One thread per pair; the expanding-polytope stage runs as a second kernel over
// a queue of the pairs that need it, each with a face/vertex pool in a global
// workspace slot.  Candidates (valid, normal on B, point on B, depth) are
// written per pair and folded into the manifold by a final per-pair kernel.
// World queries live here too: rayTest, convexSweepTest (GjkConvexCast over the
// query tree's candidates) and the continuous-collision motion clamp of
// integrateTransforms.

#include <cstring>

#include "internal.h"
#include "shapes.h"

namespace jbg {

namespace {

struct Cand {  // 8 floats per pair
  float valid, nx, ny, nz, px, py, pz, depth;
};

__device__ __forceinline__ Tr loadTr(const float* tr, int i) {
  Tr t;
  const float* p = tr + 12 * i;
  for (int k = 0; k < 9; k++) t.b.m[k] = p[k];
  t.o = v3(p[9], p[10], p[11]);
  return t;
}
__device__ __forceinline__ Shape bodyShape(const float* shapes, const float* props, int i) {
  return loadShape(shapes + JBGPU_SHAPE_WORDS * ldi(props + 8 * i + 6));
}
__device__ __forceinline__ bool isActiveState(int a) {
  return a != JBGPU_ISLAND_SLEEPING && a != JBGPU_DISABLE_SIMULATION;
}
__device__ __forceinline__ void putCand(float* out, int i, bool valid, V3 n, V3 p, float d) {
  float* c = out + 8 * i;
  c[0] = valid ? 1.0f : 0.0f;
  c[1] = n.x; c[2] = n.y; c[3] = n.z;
  c[4] = p.x; c[5] = p.y; c[6] = p.z;
  c[7] = d;
}

// ------------------------------------------------------------------ binning
// sort key of a pair: its type (which kernel handles it) above the class of its two shape types, so that inside a
// bin the lanes of a wave mostly run the same support functions.  Class: the other shape's type for a pair with
// a plane, else lower convex type * 4 + higher (convex types are 0-3): 4 bits, one radix pass with 64 buckets
constexpr int kBinClassBits = 4;
__global__ void kBinKeys(int64_t n, const int* __restrict__ dn, const int* __restrict__ ptype, const uint64_t* __restrict__ keys,
                         const int* __restrict__ act, const float* __restrict__ shapes,
                         const float* __restrict__ props, int* tkey, int* idx) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int a = (int)(keys[i] >> 32) - 1, b = (int)(keys[i] & 0xffffffffu) - 1;
  int t = ptype[i];
  if (!(isActiveState(act[a]) || isActiveState(act[b]))) t = JBGPU_PAIR_NONE;  // needsCollision
  const int sa = ldi(shapes + JBGPU_SHAPE_WORDS * ldi(props + 8 * a + 6));
  const int sb = ldi(shapes + JBGPU_SHAPE_WORDS * ldi(props + 8 * b + 6));
  const int cls = sa == JBGPU_SHAPE_PLANE ? sb : sb == JBGPU_SHAPE_PLANE ? sa : min(sa, sb) * 4 + max(sa, sb);
  tkey[i] = (t << kBinClassBits) | (cls & ((1 << kBinClassBits) - 1));
  idx[i] = (int)i;
}

// first position of each pair type in the sorted keys (types absent from the step keep -1)
__global__ void kBinStarts(int64_t n, const int* __restrict__ dn, const int* __restrict__ sorted, int* starts) {
  if (dn) n = *dn;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = sorted[i] >> kBinClassBits;
  if (i == 0 || (sorted[i - 1] >> kBinClassBits) != t) starts[t] = (int)i;
}

// ------------------------------------------------------------------ sphere-sphere
// per pair type t: range[2t] first sorted position, range[2t + 1] count (types absent from the sort: 0, 0)
__global__ void kBinRanges(int64_t n, const int* __restrict__ dn, const int* __restrict__ starts, int* range) {
  if (dn) n = *dn;
  const int t = threadIdx.x;
  if (t >= 4) return;
  int end = (int)n;
  for (int u = t + 1; u < 4; u++)
    if (starts[u] >= 0) {
      end = starts[u];
      break;
    }
  range[2 * t] = starts[t] >= 0 ? starts[t] : 0;
  range[2 * t + 1] = starts[t] >= 0 ? end - starts[t] : 0;
}

__global__ void kSphereSphere(const int* __restrict__ range, const int* __restrict__ list, const uint64_t* __restrict__ keys,
                              const float* __restrict__ shapes, const float* __restrict__ props,
                              const float* __restrict__ tr, float* cand) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= range[1]) return;
  int i = list[range[0] + k];
  int a = (int)(keys[i] >> 32) - 1, b = (int)(keys[i] & 0xffffffffu) - 1;
  Shape s0 = bodyShape(shapes, props, a), s1 = bodyShape(shapes, props, b);
  V3 o0 = ld3(tr + 12 * a + 9), o1 = ld3(tr + 12 * b + 9);
  V3 diff = sub(o0, o1);
  float l = len(diff);
  float r0 = s0.dims.x, r1 = s1.dims.x;
  if (l > (r0 + r1)) {
    putCand(cand, i, false, v3(0, 0, 0), v3(0, 0, 0), 0.0f);
    return;
  }
  float dist = l - (r0 + r1);
  V3 nrm = v3(1, 0, 0);
  if (l > kFltEps) nrm = scl(diff, 1.0f / l);
  V3 pos1 = add(o1, scl(nrm, r1));
  putCand(cand, i, true, nrm, pos1, dist);
}

// ------------------------------------------------------------------ convex-plane
__global__ void kConvexPlane(const int* __restrict__ range, const int* __restrict__ list, const uint64_t* __restrict__ keys,
                             const float* __restrict__ shapes, const float* __restrict__ verts,
                             const float* __restrict__ props, const float* __restrict__ tr, float* cand) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= range[1]) return;
  int i = list[range[0] + k];
  int a = (int)(keys[i] >> 32) - 1, b = (int)(keys[i] & 0xffffffffu) - 1;
  Shape sa = bodyShape(shapes, props, a), sb = bodyShape(shapes, props, b);
  int cv = (sb.type == JBGPU_SHAPE_PLANE) ? a : b;
  int pl = (cv == a) ? b : a;
  Shape convex = (cv == a) ? sa : sb;
  Shape plane = (cv == a) ? sb : sa;
  Tr tc = loadTr(tr, cv), tp = loadTr(tr, pl);
  V3 pn = plane.pn;
  Tr planeInConvex = trMul(trInverse(tc), tp);
  Tr convexInPlane = trMul(trInverse(tp), tc);
  V3 dir = mulMV(planeInConvex.b, scl(pn, -1.0f));
  V3 vtx = supportMargin(convex, verts, dir);
  V3 vip = xform(convexInPlane, vtx);
  float distance = (dot(pn, vip) - plane.pc);
  V3 proj = sub(vip, scl(pn, distance));
  V3 world = xform(tp, proj);
  bool has = distance < kContactThreshold;
  putCand(cand, i, has, mulMV(tp.b, pn), world, distance);
}

// ------------------------------------------------------------------ Voronoi simplex solver
struct Sub {
  V3 closest;
  uint32_t used;  // bit0 A .. bit3 D
  float bc[4];
  bool degenerate;
  __device__ void reset() {
    degenerate = false;
    bc[0] = bc[1] = bc[2] = bc[3] = 0.0f;
    used = 0;
  }
  __device__ bool isValid() const { return (bc[0] >= 0.0f) && (bc[1] >= 0.0f) && (bc[2] >= 0.0f) && (bc[3] >= 0.0f); }
  __device__ void setBC(float a, float b, float c, float d) { bc[0] = a; bc[1] = b; bc[2] = c; bc[3] = d; }
};

__device__ bool triClosest(V3 p, V3 a, V3 b, V3 c, Sub& r) {
  r.used = 0;
  V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) {
    r.closest = a; r.used = 1; r.setBC(1, 0, 0, 0);
    return true;
  }
  V3 bp = sub(p, b);
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) {
    r.closest = b; r.used = 2; r.setBC(0, 1, 0, 0);
    return true;
  }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float v = d1 / (d1 - d3);
    r.closest = v3(v * ab.x + a.x, v * ab.y + a.y, v * ab.z + a.z);
    r.used = 3;
    r.setBC(1.0f - v, v, 0, 0);
    return true;
  }
  V3 cp = sub(p, c);
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) {
    r.closest = c; r.used = 4; r.setBC(0, 0, 1, 0);
    return true;
  }
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    r.closest = v3(w * ac.x + a.x, w * ac.y + a.y, w * ac.z + a.z);
    r.used = 5;
    r.setBC(1.0f - w, 0, w, 0);
    return true;
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    V3 t = sub(c, b);
    r.closest = v3(w * t.x + b.x, w * t.y + b.y, w * t.z + b.z);
    r.used = 6;
    r.setBC(0, 1.0f - w, w, 0);
    return true;
  }
  float denom = 1.0f / (va + vb + vc);
  float v = vb * denom, w = vc * denom;
  V3 t1 = scl(ab, v), t2 = scl(ac, w);
  r.closest = v3(a.x + t1.x + t2.x, a.y + t1.y + t2.y, a.z + t1.z + t2.z);
  r.used = 7;
  r.setBC(1.0f - v - w, v, w, 0);
  return true;
}

__device__ __forceinline__ int outsidePlane(V3 p, V3 a, V3 b, V3 c, V3 d) {
  V3 normal = crs(sub(b, a), sub(c, a));
  float signp = dot(sub(p, a), normal);
  float signd = dot(sub(d, a), normal);
  if (signd * signd < ((1e-4f) * (1e-4f))) return -1;
  return (signp * signd < 0.0f) ? 1 : 0;
}

__device__ __forceinline__ uint32_t bit(uint32_t m, int k) { return (m >> k) & 1u; }

__device__ bool tetClosest(V3 p, V3 a, V3 b, V3 c, V3 d, Sub& fr) {
  Sub tr;
  tr.reset();
  fr.closest = p;
  fr.used = 15;
  int oABC = outsidePlane(p, a, b, c, d);
  int oACD = outsidePlane(p, a, c, d, b);
  int oADB = outsidePlane(p, a, d, b, c);
  int oBDC = outsidePlane(p, b, d, c, a);
  if (oABC < 0 || oACD < 0 || oADB < 0 || oBDC < 0) {
    fr.degenerate = true;
    return false;
  }
  if (oABC == 0 && oACD == 0 && oADB == 0 && oBDC == 0) return false;
  float best = 3.40282347e+38f;
  if (oABC != 0) {
    triClosest(p, a, b, c, tr);
    V3 t = sub(tr.closest, p);
    float sq = dot(t, t);
    if (sq < best) {
      best = sq;
      fr.closest = tr.closest;
      fr.used = bit(tr.used, 0) | (bit(tr.used, 1) << 1) | (bit(tr.used, 2) << 2);
      fr.setBC(tr.bc[0], tr.bc[1], tr.bc[2], 0);
    }
  }
  if (oACD != 0) {
    triClosest(p, a, c, d, tr);
    V3 t = sub(tr.closest, p);
    float sq = dot(t, t);
    if (sq < best) {
      best = sq;
      fr.closest = tr.closest;
      fr.used = bit(tr.used, 0) | (bit(tr.used, 1) << 2) | (bit(tr.used, 2) << 3);
      fr.setBC(tr.bc[0], 0, tr.bc[1], tr.bc[2]);
    }
  }
  if (oADB != 0) {
    triClosest(p, a, d, b, tr);
    V3 t = sub(tr.closest, p);
    float sq = dot(t, t);
    if (sq < best) {
      best = sq;
      fr.closest = tr.closest;
      fr.used = bit(tr.used, 0) | (bit(tr.used, 2) << 1) | (bit(tr.used, 1) << 3);
      fr.setBC(tr.bc[0], tr.bc[2], 0, tr.bc[1]);
    }
  }
  if (oBDC != 0) {
    triClosest(p, b, d, c, tr);
    V3 t = sub(tr.closest, p);
    float sq = dot(t, t);
    if (sq < best) {
      best = sq;
      fr.closest = tr.closest;
      fr.used = (bit(tr.used, 0) << 1) | (bit(tr.used, 2) << 2) | (bit(tr.used, 1) << 3);
      fr.setBC(0, tr.bc[0], tr.bc[2], tr.bc[1]);
    }
  }
  return true;
}

// Vertex storage of the simplex (w, p, q of up to 5 vertices) in LDS, [row][thread] with row = 15 * array
// + 3 * vertex + component and a row stride of S (the block size): indexed by the vertex count, it would
// otherwise sit in scratch memory
constexpr int kGjkBlock = 128;
constexpr int kSimplexRows = 45;

template <int S>
struct SimplexT {
  int nv;
  float* lw;  // this thread's column of the LDS block
  V3 cP1, cP2, cV, lastW;
  bool validClosest, needsUpdate;
  Sub bc;

  __device__ V3 get(int arr, int i) const {
    const float* p = lw + (15 * arr + 3 * i) * S;
    return v3(p[0], p[S], p[2 * S]);
  }
  __device__ void set(int arr, int i, V3 v) {
    float* p = lw + (15 * arr + 3 * i) * S;
    p[0] = v.x;
    p[S] = v.y;
    p[2 * S] = v.z;
  }
  __device__ V3 W(int i) const { return get(0, i); }
  __device__ V3 P(int i) const { return get(1, i); }
  __device__ V3 Q(int i) const { return get(2, i); }

  __device__ void reset() {
    validClosest = false;
    nv = 0;
    needsUpdate = true;
    lastW = v3(1e30f, 1e30f, 1e30f);
    bc.reset();
  }
  __device__ void removeVertex(int idx) {
    nv--;
    set(0, idx, W(nv));
    set(1, idx, P(nv));
    set(2, idx, Q(nv));
  }
  __device__ void reduce(uint32_t used) {
    if ((nv >= 4) && !bit(used, 3)) removeVertex(3);
    if ((nv >= 3) && !bit(used, 2)) removeVertex(2);
    if ((nv >= 2) && !bit(used, 1)) removeVertex(1);
    if ((nv >= 1) && !bit(used, 0)) removeVertex(0);
  }
  __device__ void addVertex(V3 w, V3 p, V3 q) {
    lastW = w;
    needsUpdate = true;
    set(0, nv, w);
    set(1, nv, p);
    set(2, nv, q);
    nv++;
  }
  __device__ bool inSimplex(V3 w) const {
    bool found = false;
    for (int i = 0; i < nv; i++)
      if (veq(W(i), w)) found = true;
    if (veq(w, lastW)) return true;
    return found;
  }
  __device__ bool update() {
    if (needsUpdate) {
      bc.reset();
      needsUpdate = false;
      switch (nv) {
        case 0:
          validClosest = false;
          break;
        case 1:
          cP1 = P(0);
          cP2 = Q(0);
          cV = sub(cP1, cP2);
          bc.reset();
          bc.setBC(1, 0, 0, 0);
          validClosest = bc.isValid();
          break;
        case 2: {
          V3 from = W(0), to = W(1);
          V3 diff = sub(v3(0, 0, 0), from);
          V3 v = sub(to, from);
          float t = dot(v, diff);
          if (t > 0) {
            float dotVV = dot(v, v);
            if (t < dotVV) {
              t /= dotVV;
              diff = sub(diff, scl(v, t));
              bc.used |= 3;
            } else {
              t = 1;
              diff = sub(diff, v);
              bc.used |= 2;
            }
          } else {
            t = 0;
            bc.used |= 1;
          }
          bc.setBC(1.0f - t, t, 0, 0);
          cP1 = add(P(0), scl(sub(P(1), P(0)), t));
          cP2 = add(Q(0), scl(sub(Q(1), Q(0)), t));
          cV = sub(cP1, cP2);
          reduce(bc.used);
          validClosest = bc.isValid();
          break;
        }
        case 3: {
          triClosest(v3(0, 0, 0), W(0), W(1), W(2), bc);
          V3 t1 = scl(P(0), bc.bc[0]), t2 = scl(P(1), bc.bc[1]), t3 = scl(P(2), bc.bc[2]);
          cP1 = v3(t1.x + t2.x + t3.x, t1.y + t2.y + t3.y, t1.z + t2.z + t3.z);
          t1 = scl(Q(0), bc.bc[0]); t2 = scl(Q(1), bc.bc[1]); t3 = scl(Q(2), bc.bc[2]);
          cP2 = v3(t1.x + t2.x + t3.x, t1.y + t2.y + t3.y, t1.z + t2.z + t3.z);
          cV = sub(cP1, cP2);
          reduce(bc.used);
          validClosest = bc.isValid();
          break;
        }
        case 4: {
          bool sep = tetClosest(v3(0, 0, 0), W(0), W(1), W(2), W(3), bc);
          if (sep) {
            V3 t1 = scl(P(0), bc.bc[0]), t2 = scl(P(1), bc.bc[1]), t3 = scl(P(2), bc.bc[2]), t4 = scl(P(3), bc.bc[3]);
            cP1 = v3(t1.x + t2.x + t3.x + t4.x, t1.y + t2.y + t3.y + t4.y, t1.z + t2.z + t3.z + t4.z);
            t1 = scl(Q(0), bc.bc[0]); t2 = scl(Q(1), bc.bc[1]); t3 = scl(Q(2), bc.bc[2]); t4 = scl(Q(3), bc.bc[3]);
            cP2 = v3(t1.x + t2.x + t3.x + t4.x, t1.y + t2.y + t3.y + t4.y, t1.z + t2.z + t3.z + t4.z);
            cV = sub(cP1, cP2);
            reduce(bc.used);
          } else {
            if (bc.degenerate) {
              validClosest = false;
            } else {
              validClosest = true;
              cV = v3(0, 0, 0);
            }
            break;
          }
          validClosest = bc.isValid();
          break;
        }
        default:
          validClosest = false;
      }
    }
    return validClosest;
  }
};
using Simplex = SimplexT<kGjkBlock>;

// ------------------------------------------------------------------ GJK pair detector
// GjkPairDetector.getClosestPoints up to the penetration test (GjkPairDetector.java:76-260): the closest points
// of a and b (cores plus margins) in the frame shifted by off = (ta.o + tb.o) / 2, md the maximum distance
// squared; checkPenetration and the degenerate-simplex code decide whether a penetration solver would run
struct GjkOut {
  bool isValid, checkPenetration;
  int degen;
  float distance, margin;
  V3 normalInB, pointOnA, pointOnB, off;
};
template <int S>
__device__ __forceinline__ void gjkClosest(const Shape& sa, const Tr& ta, const Shape& sb, const Tr& tb,
                                           const float* __restrict__ verts, float md, SimplexT<S>& vs, GjkOut& out) {
  const float REL_ERROR2 = 1.0e-6f;
  float distance = 0.0f;
  V3 normalInB = v3(0, 0, 0), pointOnA = v3(0, 0, 0), pointOnB = v3(0, 0, 0);
  Tr la = ta, lb = tb;
  V3 off = scl(add(la.o, lb.o), 0.5f);
  la.o = sub(la.o, off);
  lb.o = sub(lb.o, off);
  float marginA = sa.margin, marginB = sb.margin;
  int curIter = 0;
  V3 sep = v3(0, 1, 0);
  bool isValid = false, checkSimplex = false, checkPenetration = true;
  int degen = 0;
  float sqd = kInf;
  float margin = marginA + marginB;
  vs.reset();
  for (;;) {
    V3 sepA = mulMtV(ta.b, scl(sep, -1.0f));
    V3 sepB = mulMtV(tb.b, sep);
    V3 pInA = supportNoMargin(sa, verts, sepA);
    V3 qInB = supportNoMargin(sb, verts, sepB);
    V3 pW = xform(la, pInA);
    V3 qW = xform(lb, qInB);
    V3 w = sub(pW, qW);
    float delta = dot(sep, w);
    if ((delta > 0.0f) && (delta * delta > sqd * md)) {
      checkPenetration = false;
      break;
    }
    if (vs.inSimplex(w)) {
      degen = 1;
      checkSimplex = true;
      break;
    }
    float f0 = sqd - delta;
    float f1 = sqd * REL_ERROR2;
    if (f0 <= f1) {
      if (f0 <= 0.0f) degen = 2;
      checkSimplex = true;
      break;
    }
    vs.addVertex(w, pW, qW);
    bool ok = vs.update();
    sep = vs.cV;
    if (!ok) {
      degen = 3;
      checkSimplex = true;
      break;
    }
    if (len2(sep) < REL_ERROR2) {
      degen = 6;
      checkSimplex = true;
      break;
    }
    float prev = sqd;
    sqd = len2(sep);
    if (prev - sqd <= kFltEps * prev) {
      sep = vs.cV;
      checkSimplex = true;
      break;
    }
    if (curIter++ > 1000) break;
    if (vs.nv == 4) {
      sep = vs.cV;
      break;
    }
  }
  if (checkSimplex) {
    vs.update();
    pointOnA = vs.cP1;
    pointOnB = vs.cP2;
    normalInB = sub(pointOnA, pointOnB);
    float lenSqr = len2(sep);
    if (lenSqr < 0.0001f) degen = 5;
    if (lenSqr > kFltEps * kFltEps) {
      float rlen = 1.0f / jsqrt(lenSqr);
      normalInB = scl(normalInB, rlen);
      float s = jsqrt(sqd);
      pointOnA = sub(pointOnA, scl(sep, marginA / s));
      pointOnB = add(pointOnB, scl(sep, marginB / s));
      distance = ((1.0f / rlen) - margin);
      isValid = true;
    }
  }
  out.isValid = isValid;
  out.checkPenetration = checkPenetration;
  out.degen = degen;
  out.distance = distance;
  out.margin = margin;
  out.normalInB = normalInB;
  out.pointOnA = pointOnA;
  out.pointOnB = pointOnB;
  out.off = off;
}

// epaState per pair: [needEpa, isValid, distance, pA xyz, pB xyz, nB xyz]
// epaState per pair: [needEpa, isValid, distance, pA xyz, pB xyz, nB xyz]
__global__ __launch_bounds__(kGjkBlock) void kGjk(const int* __restrict__ range, const int* __restrict__ list, const uint64_t* __restrict__ keys,
                                            const float* __restrict__ shapes, const float* __restrict__ verts,
                                            const float* __restrict__ props, const float* __restrict__ tr,
                                            float* cand, float* epaState, int* epaQueue,
                                            unsigned long long* epaCount) {
  __shared__ float simplexLds[kSimplexRows * kGjkBlock];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= range[1]) return;
  int i = list[range[0] + k];
  int a = (int)(keys[i] >> 32) - 1, b = (int)(keys[i] & 0xffffffffu) - 1;
  Shape sa = bodyShape(shapes, props, a), sb = bodyShape(shapes, props, b);
  Tr ta = loadTr(tr, a), tb = loadTr(tr, b);
  float md = sa.margin + sb.margin + kContactThreshold;  // ConvexConvexAlgorithm.java:122-123
  md *= md;
  Simplex vs;
  vs.lw = simplexLds + threadIdx.x;
  GjkOut g;
  gjkClosest(sa, ta, sb, tb, verts, md, vs, g);
  bool catchDegen = (g.degen != 0 && ((g.distance + g.margin) < 0.01f));
  float* es = epaState + 12 * i;
  if (g.checkPenetration && (!g.isValid || catchDegen)) {
    es[0] = 1.0f;
    es[1] = g.isValid ? 1.0f : 0.0f;
    es[2] = g.distance;
    st3(es + 3, g.pointOnA);
    st3(es + 6, g.pointOnB);
    st3(es + 9, g.normalInB);
    unsigned long long q = atomicAdd(epaCount, 1ull);
    epaQueue[q] = i;
    putCand(cand, i, false, v3(0, 0, 0), v3(0, 0, 0), 0.0f);
    return;
  }
  es[0] = 0.0f;
  putCand(cand, i, g.isValid, g.normalInB, add(g.pointOnB, g.off), g.distance);
}

// ------------------------------------------------------------------ GJK-EPA (GjkEpaSolver)
constexpr int GJK_maxiterations = 128;
constexpr int GJK_hashmask = 63;
constexpr float GJK_insimplex_eps = 0.0001f;
constexpr float GJK_sqinsimplex_eps = GJK_insimplex_eps * GJK_insimplex_eps;
constexpr int EPA_maxiterations = 256;
constexpr float EPA_inface_eps = 0.01f;
constexpr float EPA_accuracy = 0.001f;

// hull support map spread over a wave's lanes (all lanes call it with the same arguments): per lane the
// first maximum of its stride, then the largest dot product with the lowest index among equals, i.e. the
// vertex hullSupport's strict '>' scan keeps
__device__ V3 hullSupportWave(const float* verts, int first, int n, V3 vec) {
  float l2 = len2(vec);
  if (l2 < 0.0001f)
    vec = v3(1, 0, 0);
  else
    vec = scl(vec, 1.0f / jsqrt(l2));
  const float* p = verts + 3 * first;
  float best = -1e30f;
  int bi = -1;
  for (int i = threadIdx.x & 63; i < n; i += 64) {
    V3 vtx = v3(p[3 * i] * 1.0f, p[3 * i + 1] * 1.0f, p[3 * i + 2] * 1.0f);
    float d = dot(vec, vtx);
    if (d > best) {
      best = d;
      bi = i;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float od = __shfl_xor(best, o);
    const int oi = __shfl_xor(bi, o);
    if (oi >= 0 && (bi < 0 || od > best || (od == best && oi < bi))) {
      best = od;
      bi = oi;
    }
  }
  return bi >= 0 ? v3(p[3 * bi] * 1.0f, p[3 * bi + 1] * 1.0f, p[3 * bi + 2] * 1.0f) : v3(0, 0, 0);
}
__device__ V3 supportMarginWave(const Shape& s, const float* verts, V3 vec) {
  if (s.type != JBGPU_SHAPE_HULL) return supportMargin(s, verts, vec);
  V3 sup = hullSupportWave(verts, s.vfirst, s.vcount, vec);
  if (s.margin != 0.0f) {
    V3 vn = vec;
    if (len2(vn) < (kFltEps * kFltEps)) vn = v3(-1, -1, -1);
    vn = nor(vn);
    vn = scl(vn, s.margin);
    sup = add(sup, vn);
  }
  return sup;
}

// initial polytopes of EvaluatePD: faces and links of the tetrahedron and of the hexahedron
// (GjkEpaSolver.java EPA.EvaluatePD tables)
__constant__ int kTetF[4][3] = {{2, 1, 0}, {3, 0, 1}, {3, 1, 2}, {3, 2, 0}};
__constant__ int kTetE[6][4] = {{0, 0, 2, 1}, {0, 1, 1, 1}, {0, 2, 3, 1}, {1, 0, 3, 2}, {2, 0, 1, 2}, {3, 0, 2, 2}};
__constant__ int kHexF[6][3] = {{2, 0, 4}, {4, 1, 2}, {1, 4, 0}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
__constant__ int kHexE[9][4] = {{0, 0, 4, 0}, {0, 1, 2, 1}, {0, 2, 1, 2}, {1, 1, 5, 2}, {1, 0, 2, 0},
                                {2, 2, 3, 2}, {3, 1, 5, 0}, {3, 0, 4, 2}, {5, 1, 4, 1}};

struct Mkv {
  V3 w, r;
};

// workspace slot layout (floats): faces [capFace*16] | vertices [capVtx*6] | he [kEpaHe*4]
struct EpaWork {
  float* face;
  float* vtx;
  float* he;
  int nface, nvtx, nhe;
  int capFace, capVtx;  // capacities of this workspace
  bool overflow;
  // face fields: 0-2 v, 3-5 f, 6-8 e, 9-11 n, 12 d, 13 mark, 14 prev, 15 next
  __device__ int& fi(int f, int k) { return *reinterpret_cast<int*>(face + kEpaFaceWords * f + k); }
  __device__ float& ff(int f, int k) { return face[kEpaFaceWords * f + k]; }
  __device__ V3 fn(int f) { return ld3(face + kEpaFaceWords * f + 9); }
  __device__ V3 vw(int v) { return ld3(vtx + 6 * v); }
  __device__ V3 vr(int v) { return ld3(vtx + 6 * v + 3); }
};

struct EpaSolver {
  Shape sh[2];
  M3 wrot[2];
  V3 pos[2];
  const float* verts;
  float margin;
  // shared by the wave's lanes (all of them run the job in lockstep on identical values): LDS
  int* table;     // 64 hash heads
  Mkv* simplex;   // 5 vertices
  int* horizon;   // BuildHorizon stack: capStack faces, then capStack edges
  int capStack;
  V3 ray;
  int order, iterations;
  bool failed;
  EpaWork W;
  int root, nfaces, epaIter;
  V3 nearest[2];

  __device__ __forceinline__ static int Hash(V3 v) {
    int h = jint(v.x * 15461) ^ jint(v.y * 83003) ^ jint(v.z * 15473);
    return (int)((uint32_t)h * 169639u) & GJK_hashmask;
  }
  __device__ __forceinline__ V3 LocalSupport(V3 d, int i) {
    V3 t = mulMtV(wrot[i], d);
    V3 o = supportMarginWave(sh[i], verts, t);
    o = mulMV(wrot[i], o);
    return add(o, pos[i]);
  }
  __device__ __forceinline__ void Support(V3 d, Mkv& v) {
    v.r = d;
    V3 t1 = LocalSupport(d, 0);
    V3 t2 = LocalSupport(scl(d, -1.0f), 1);
    v.w = sub(t1, t2);
    v.w.x += margin * d.x;
    v.w.y += margin * d.y;
    v.w.z += margin * d.z;
  }
  __device__ __forceinline__ bool FetchSupport() {
    int h = Hash(ray);
    int e = table[h];
    int guardHops = 0;
    while (e >= 0) {
      if (veq(ld3(W.he + 4 * e), ray)) {
        --order;
        return false;
      }
      e = __builtin_bit_cast(int, W.he[4 * e + 3]);
      if (++guardHops > kEpaHe) {  // chains are acyclic; bound the walk regardless
        W.overflow = true;
        return false;
      }
    }
    if (W.nhe >= kEpaHe) {
      W.overflow = true;
      return false;
    }
    int ne = W.nhe++;
    st3(W.he + 4 * ne, ray);
    W.he[4 * ne + 3] = __builtin_bit_cast(float, table[h]);
    table[h] = ne;
    Support(ray, simplex[++order]);
    return (dot(ray, simplex[order].w) > 0);
  }
  __device__ __forceinline__ bool SolveSimplex2(V3 ao, V3 ab) {
    if (dot(ab, ao) >= 0) {
      V3 cabo = crs(ab, ao);
      if (len2(cabo) > GJK_sqinsimplex_eps)
        ray = crs(cabo, ab);
      else
        return true;
    } else {
      order = 0;
      simplex[0] = simplex[1];
      ray = ao;
    }
    return false;
  }
  __device__ __forceinline__ bool SolveSimplex3a(V3 ao, V3 ab, V3 ac, V3 cabc) {
    V3 t = crs(cabc, ab);
    V3 t2 = crs(cabc, ac);
    if (dot(t, ao) < -GJK_insimplex_eps) {
      order = 1;
      simplex[0] = simplex[1];
      simplex[1] = simplex[2];
      return SolveSimplex2(ao, ab);
    } else if (dot(t2, ao) > +GJK_insimplex_eps) {
      order = 1;
      simplex[1] = simplex[2];
      return SolveSimplex2(ao, ac);
    } else {
      float d = dot(cabc, ao);
      if (jabs(d) > GJK_insimplex_eps) {
        if (d > 0) {
          ray = cabc;
        } else {
          ray = scl(cabc, -1.0f);
          Mkv sw = simplex[0];
          simplex[0] = simplex[1];
          simplex[1] = sw;
        }
        return false;
      }
      return true;
    }
  }
  __device__ __forceinline__ bool SolveSimplex4(V3 ao, V3 ab, V3 ac, V3 ad) {
    V3 t = crs(ab, ac), t2 = crs(ac, ad), t3 = crs(ad, ab);
    if (dot(t, ao) > GJK_insimplex_eps) {
      order = 2;
      simplex[0] = simplex[1];
      simplex[1] = simplex[2];
      simplex[2] = simplex[3];
      return SolveSimplex3a(ao, ab, ac, t);
    } else if (dot(t2, ao) > GJK_insimplex_eps) {
      order = 2;
      simplex[2] = simplex[3];
      return SolveSimplex3a(ao, ac, ad, t2);
    } else if (dot(t3, ao) > GJK_insimplex_eps) {
      order = 2;
      simplex[1] = simplex[0];
      simplex[0] = simplex[2];
      simplex[2] = simplex[3];
      return SolveSimplex3a(ao, ad, ab, t3);
    }
    return true;
  }
  __device__ __forceinline__ bool SearchOrigin() {
    iterations = 0;
    order = -1;
    failed = false;
    ray = nor(v3(1, 0, 0));
    table[threadIdx.x & 63] = -1;  // one entry per lane
    W.nhe = 0;
    FetchSupport();
    ray = scl(simplex[0].w, -1.0f);
    for (; iterations < GJK_maxiterations; ++iterations) {
      if (W.overflow) return false;
      float rl = len(ray);
      ray = scl(ray, 1.0f / (rl > 0.0f ? rl : 1.0f));
      if (FetchSupport()) {
        bool found = false;
        switch (order) {
          case 1:
            found = SolveSimplex2(scl(simplex[1].w, -1.0f), sub(simplex[0].w, simplex[1].w));
            break;
          case 2: {
            V3 ab = sub(simplex[1].w, simplex[2].w), ac = sub(simplex[0].w, simplex[2].w);
            found = SolveSimplex3a(scl(simplex[2].w, -1.0f), ab, ac, crs(ab, ac));
            break;
          }
          case 3:
            found = SolveSimplex4(scl(simplex[3].w, -1.0f), sub(simplex[2].w, simplex[3].w),
                                  sub(simplex[1].w, simplex[3].w), sub(simplex[0].w, simplex[3].w));
            break;
        }
        if (found) return true;
      } else {
        return false;
      }
    }
    failed = true;
    return false;
  }
  __device__ __forceinline__ bool EncloseOrigin() {
    switch (order) {
      case 0:
        break;
      case 1: {
        V3 ab = sub(simplex[1].w, simplex[0].w);
        V3 b0 = crs(ab, v3(1, 0, 0)), b1 = crs(ab, v3(0, 1, 0)), b2 = crs(ab, v3(0, 0, 1));
        float m0 = len2(b0), m1 = len2(b1), m2 = len2(b2);
        V3 ax = nor(ab);
        float ang = k2Pi / 3.0f;
        float dd = len(ax);
        float s = (float)sin((double)(ang * 0.5f)) / dd;
        Q4 q = Q4{ax.x * s, ax.y * s, ax.z * s, (float)cos((double)(ang * 0.5f))};
        M3 r = mFromQ(q);
        int sel = m0 > m1 ? (m0 > m2 ? 0 : 2) : (m1 > m2 ? 1 : 2);
        V3 w = sel == 0 ? b0 : (sel == 1 ? b1 : b2);
        Support(nor(w), simplex[4]);
        w = mulMV(r, w);
        Support(nor(w), simplex[2]);
        w = mulMV(r, w);
        Support(nor(w), simplex[3]);
        order = 4;
        return true;
      }
      case 2: {
        V3 n = nor(crs(sub(simplex[1].w, simplex[0].w), sub(simplex[2].w, simplex[0].w)));
        Support(n, simplex[3]);
        Support(scl(n, -1.0f), simplex[4]);
        order = 4;
        return true;
      }
      case 3:
      case 4:
        return true;
    }
    return false;
  }
  __device__ __forceinline__ int addVertex(const Mkv& m) {
    if (W.nvtx >= W.capVtx) {
      W.overflow = true;
      return -1;
    }
    int v = W.nvtx++;
    st3(W.vtx + 6 * v, m.w);
    st3(W.vtx + 6 * v + 3, m.r);
    return v;
  }
  __device__ __forceinline__ bool Set(int f, int a, int b, int c) {
    V3 aw = W.vw(a), bw = W.vw(b), cw = W.vw(c);
    V3 nrm = crs(sub(bw, aw), sub(cw, aw));
    float l = len(nrm);
    V3 t1 = crs(aw, bw), t2 = crs(bw, cw), t3 = crs(cw, aw);
    bool valid = (dot(t1, nrm) >= -EPA_inface_eps) && (dot(t2, nrm) >= -EPA_inface_eps) &&
                 (dot(t3, nrm) >= -EPA_inface_eps);
    W.fi(f, 0) = a;
    W.fi(f, 1) = b;
    W.fi(f, 2) = c;
    W.fi(f, 13) = 0;
    V3 n = scl(nrm, 1.0f / (l > 0.0f ? l : kInf));
    st3(W.face + kEpaFaceWords * f + 9, n);
    W.ff(f, 12) = jmax(0.0f, -dot(n, aw));
    return valid;
  }
  __device__ __forceinline__ int NewFace(int a, int b, int c) {
    if (W.nface >= W.capFace) {
      W.overflow = true;
      return -1;
    }
    int pf = W.nface++;
    W.fi(pf, 3) = W.fi(pf, 4) = W.fi(pf, 5) = -1;
    if (Set(pf, a, b, c)) {
      if (root >= 0) W.fi(root, 14) = pf;
      W.fi(pf, 14) = -1;
      W.fi(pf, 15) = root;
      root = pf;
      ++nfaces;
    } else {
      W.fi(pf, 14) = W.fi(pf, 15) = -1;
    }
    return pf;
  }
  __device__ __forceinline__ void Detach(int f) {
    int pv = W.fi(f, 14), nx = W.fi(f, 15);
    if (pv >= 0 || nx >= 0) {
      --nfaces;
      if (f == root) {
        root = nx;
        if (root >= 0) W.fi(root, 14) = -1;
      } else {
        if (nx < 0) {
          W.fi(pv, 15) = -1;
        } else {
          W.fi(pv, 15) = nx;
          W.fi(nx, 14) = pv;
        }
      }
      W.fi(f, 14) = W.fi(f, 15) = -1;
    }
  }
  __device__ __forceinline__ void Link(int f0, int e0, int f1, int e1) {
    W.fi(f0, 3 + e0) = f1;
    W.fi(f1, 6 + e1) = e0;
    W.fi(f1, 3 + e1) = f0;
    W.fi(f0, 6 + e0) = e1;
  }
  // BuildHorizon (GjkEpaSolver.java:686-706) as an explicit-stack traversal in the same visiting order
  __device__ __forceinline__ int BuildHorizon(int markid, int w, int f0, int e0, int* cf, int* ff) {
    const int mod3[5] = {0, 1, 2, 0, 1};
    int* stkF = horizon;
    int* stkE = horizon + capStack;
    int sp = 0;
    int ne = 0;
    stkF[sp] = f0;
    stkE[sp] = e0;
    sp++;
    while (sp > 0) {
      if (W.overflow) return ne;
      sp--;
      int f = stkF[sp], e = stkE[sp];
      if (f < 0 || f >= W.nface || e < 0 || e > 2) {  // unlinked edge: treat as a pool failure
        W.overflow = true;
        return ne;
      }
      if (W.fi(f, 13) == markid) continue;
      int e1 = mod3[e + 1];
      if ((dot(W.fn(f), W.vw(w)) + W.ff(f, 12)) > 0) {
        int nf = NewFace(W.fi(f, e1), W.fi(f, e), w);
        if (nf < 0) return ne;
        Link(nf, 0, f, e);
        if (*cf >= 0)
          Link(*cf, 1, nf, 2);
        else
          *ff = nf;
        *cf = nf;
        ne++;
      } else {
        int e2 = mod3[e + 2];
        Detach(f);
        W.fi(f, 13) = markid;
        if (sp + 2 > capStack) {
          W.overflow = true;
          return ne;
        }
        // recursion visits f.f[e1] first, then f.f[e2]: push in reverse
        stkF[sp] = W.fi(f, 3 + e2);
        stkE[sp] = W.fi(f, 6 + e2);
        sp++;
        stkF[sp] = W.fi(f, 3 + e1);
        stkE[sp] = W.fi(f, 6 + e1);
        sp++;
      }
    }
    return ne;
  }
  // the list holds the live faces newest first (NewFace inserts at the root), so the serial walk's first
  // smallest distance is the newest face among the smallest: every lane scans a stride of the pool, then a
  // wave reduction (smaller distance, equal distance: higher index) picks the same face
  __device__ __forceinline__ int FindBest() {
    if (root < 0) return -1;
    const int lane = threadIdx.x & 63;
    float bd = kInf;
    int bf = -1;
    for (int f = lane; f < W.nface; f += 64) {
      if (f != root && W.fi(f, 14) < 0) continue;  // not in the list: only the root has no predecessor
      const float d = W.ff(f, 12);
      if (d < bd || (d == bd && bf >= 0)) {
        bd = d;
        bf = f;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const float od = __shfl_xor(bd, o);
      const int of = __shfl_xor(bf, o);
      if (of >= 0 && (bf < 0 || od < bd || (od == bd && of > bf))) {
        bd = od;
        bf = of;
      }
    }
    return bf;
  }
  __device__ __forceinline__ float EvaluatePD() {
    const int(*tet_f)[3] = kTetF;
    const int(*tet_e)[4] = kTetE;
    const int(*hex_f)[3] = kHexF;
    const int(*hex_e)[4] = kHexE;
    int bestface = -1;
    int markid = 1;
    float depth = -kInf;
    root = -1;
    nfaces = 0;
    epaIter = 0;
    W.nface = 0;
    W.nvtx = 0;
    if (EncloseOrigin()) {
      int nf = 0, ne = 0;
      if (order == 3) { nf = 4; ne = 6; }
      else if (order == 4) { nf = 6; ne = 9; }
      for (int i = 0; i <= order; ++i) addVertex(simplex[i]);
      int base[6];
      for (int i = 0; i < nf; ++i) {
        const int* fx = (order == 3) ? tet_f[i] : hex_f[i];
        base[i] = NewFace(fx[0], fx[1], fx[2]);
      }
      for (int i = 0; i < ne; ++i) {
        const int* ex = (order == 3) ? tet_e[i] : hex_e[i];
        Link(base[ex[0]], ex[1], base[ex[2]], ex[3]);
      }
    }
    if (0 == nfaces) return depth;
    for (; epaIter < EPA_maxiterations; ++epaIter) {
      if (W.overflow) return depth;
      int bf = FindBest();
      if (bf >= 0) {
        Mkv wv;
        Support(scl(W.fn(bf), -1.0f), wv);
        int w = addVertex(wv);
        if (w < 0) return depth;
        float d = dot(W.fn(bf), W.vw(w)) + W.ff(bf, 12);
        bestface = bf;
        if (d < -EPA_accuracy) {
          int cf = -1, ff = -1;
          int nf = 0;
          Detach(bf);
          W.fi(bf, 13) = ++markid;
          for (int i = 0; i < 3; ++i) nf += BuildHorizon(markid, w, W.fi(bf, 3 + i), W.fi(bf, 6 + i), &cf, &ff);
          if (W.overflow) return depth;
          if (nf <= 2) break;
          Link(cf, 1, ff, 2);
        } else {
          break;
        }
      } else {
        break;
      }
    }
    if (bestface >= 0) {
      V3 n = W.fn(bestface);
      float fd = W.ff(bestface, 12);
      V3 o = scl(n, -fd);
      int v0 = W.fi(bestface, 0), v1 = W.fi(bestface, 1), v2 = W.fi(bestface, 2);
      float a0 = len(crs(sub(W.vw(v0), o), sub(W.vw(v1), o)));
      float a1 = len(crs(sub(W.vw(v1), o), sub(W.vw(v2), o)));
      float a2 = len(crs(sub(W.vw(v2), o), sub(W.vw(v0), o)));
      float sm = a0 + a1 + a2;
      V3 bcw = scl(v3(a1, a2, a0), 1.0f / (sm > 0.0f ? sm : 1.0f));
      depth = jmax(0.0f, fd);
      int vv[3] = {v0, v1, v2};
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        float s = i != 0 ? -1.0f : 1.0f;
        V3 f0 = LocalSupport(scl(W.vr(vv[0]), s), i);
        V3 f1 = LocalSupport(scl(W.vr(vv[1]), s), i);
        V3 f2 = LocalSupport(scl(W.vr(vv[2]), s), i);
        V3 t1 = scl(f0, bcw.x), t2 = scl(f1, bcw.y), t3 = scl(f2, bcw.z);
        nearest[i] = v3(t1.x + t2.x + t3.x, t1.y + t2.y + t3.y, t1.z + t2.z + t3.z);
      }
    }
    return depth;
  }
};

// one EPA job (pair i), run by all lanes of a wave on identical values: the polytope workspace (faces,
// vertices, hash entries) and the solver's arrays (hash heads, horizon stack of capStack entries, simplex: aux) are
// in LDS (in global memory for the last tier), the face search and hull support maps are spread over the lanes;
// the GJK partial result is merged as in GjkPairDetector.java:263-303
// retry != nullptr: a job that overflows this workspace writes nothing and is queued for the next larger one
// instead (the algorithm does not depend on the capacities until it overflows, so the rerun gives the same result)
// faceLimit (<= capFace) lowers the face capacity this tier uses; a job that needs more moves to the next tier,
// so the limits shape the work between tiers and never the results
template <int capFace, int capVtx, int capStack = 64>
__device__ void epaJob(int i, const uint64_t* __restrict__ keys, const float* __restrict__ shapes,
                       const float* __restrict__ verts, const float* __restrict__ props, const float* __restrict__ tr,
                       const float* __restrict__ epaState, float* cand, unsigned long long* fails, float* slot,
                       int* aux, int* retry, unsigned long long* retryCount, int faceLimit = capFace) {
  int a = (int)(keys[i] >> 32) - 1, b = (int)(keys[i] & 0xffffffffu) - 1;
  EpaSolver S;
  S.table = aux;
  S.horizon = aux + 64;
  S.capStack = capStack;
  S.simplex = reinterpret_cast<Mkv*>(aux + 64 + 2 * capStack);
  S.sh[0] = bodyShape(shapes, props, a);
  S.sh[1] = bodyShape(shapes, props, b);
  Tr ta = loadTr(tr, a), tb = loadTr(tr, b);
  V3 off = scl(add(ta.o, tb.o), 0.5f);
  S.wrot[0] = ta.b;
  S.pos[0] = sub(ta.o, off);
  S.wrot[1] = tb.b;
  S.pos[1] = sub(tb.o, off);
  S.verts = verts;
  S.margin = 0.0f + EPA_accuracy;
  S.W.face = slot;
  S.W.vtx = slot + capFace * kEpaFaceWords;
  S.W.he = S.W.vtx + capVtx * 6;
  S.W.nface = S.W.nvtx = S.W.nhe = 0;
  S.W.capFace = faceLimit < capFace ? faceLimit : capFace;
  S.W.capVtx = capVtx;
  S.W.overflow = false;
  const float* es = epaState + 12 * i;
  bool isValid = es[1] != 0.0f;
  float distance = es[2];
  V3 pointOnA = ld3(es + 3), pointOnB = ld3(es + 6), normalInB = ld3(es + 9);
  bool ok = false;
  V3 wA, wB;
  if (S.SearchOrigin()) {
    float pd = S.EvaluatePD();
    if (pd > 0 && !S.W.overflow) {
      ok = true;
      wA = S.nearest[0];
      wB = S.nearest[1];
    }
  }
  const bool lane0 = (threadIdx.x & 63) == 0;
  if (S.W.overflow && retry) {
    if (lane0) retry[atomicAdd(retryCount, 1ull)] = i;
    return;
  }
  if (S.W.overflow && lane0) atomicAdd(fails + 1, 1ull);
  if (ok) {
    V3 tn = sub(wB, wA);
    float lenSqr = len2(tn);
    if (lenSqr > (kFltEps * kFltEps)) {
      tn = scl(tn, 1.0f / jsqrt(lenSqr));
      float distance2 = -len(sub(wA, wB));
      if (!isValid || (distance2 < distance)) {
        distance = distance2;
        pointOnA = wA;
        pointOnB = wB;
        normalInB = tn;
        isValid = true;
      }
    }
  } else if (lane0) {
    atomicAdd(fails, 1ull);
  }
  if (lane0) putCand(cand, i, isValid, normalInB, add(pointOnB, off), distance);
}

// EPA jobs in a fixed grid: every block (one wave) steps through the queue (its length read on the device)
// until it is exhausted; all lanes run each job together, lane 0 writes its result.  Every job first runs with
// a small workspace (many jobs resident per CU); the few that outgrow it are queued and rerun by kEpa with the
// large one.
__global__ __launch_bounds__(64) void kEpaSmall(const unsigned long long* __restrict__ count,
                                                const int* __restrict__ queue, const uint64_t* __restrict__ keys,
                                                const float* __restrict__ shapes, const float* __restrict__ verts,
                                                const float* __restrict__ props, const float* __restrict__ tr,
                                                const float* __restrict__ epaState, float* cand,
                                                unsigned long long* fails, int* retry,
                                                unsigned long long* retryCount, int faceLimit) {
  __shared__ float slot[kEpaSmallSlotWords];
  __shared__ int aux[192 + 5 * 6];  // hash heads, horizon stack, simplex
  const int jobs = (int)*count;
  for (int k = blockIdx.x; k < jobs; k += gridDim.x)
    epaJob<kEpaSmallFaces, kEpaSmallVerts>(queue[k], keys, shapes, verts, props, tr, epaState, cand, fails, slot, aux,
                                           retry, retryCount, faceLimit);
}
__global__ __launch_bounds__(64) void kEpa(const unsigned long long* __restrict__ count, const int* __restrict__ queue,
                                           const uint64_t* __restrict__ keys, const float* __restrict__ shapes,
                                           const float* __restrict__ verts, const float* __restrict__ props,
                                           const float* __restrict__ tr, const float* __restrict__ epaState,
                                           float* cand, unsigned long long* fails, int* retry,
                                           unsigned long long* retryCount, int faceLimit) {
  __shared__ float slot[kEpaSlotWords];
  __shared__ int aux[192 + 5 * 6];  // hash heads, horizon stack, simplex
  const int jobs = (int)*count;
  for (int k = blockIdx.x; k < jobs; k += gridDim.x)
    epaJob<kEpaFaces, kEpaVerts>(queue[k], keys, shapes, verts, props, tr, epaState, cand, fails, slot, aux, retry,
                                 retryCount, faceLimit);
}
// the last tier: the rare jobs that outgrow the large LDS workspace (long EPA runs on detailed hulls: up to 256
// iterations, each adding a horizon's worth of faces; or a horizon deeper than 64), rerun with a workspace in
// global memory, one per block: kEpaHugeFaces faces and a kEpaHugeStack-entry horizon stack
__global__ __launch_bounds__(64) void kEpaHuge(const unsigned long long* __restrict__ count,
                                               const int* __restrict__ queue, const uint64_t* __restrict__ keys,
                                               const float* __restrict__ shapes, const float* __restrict__ verts,
                                               const float* __restrict__ props, const float* __restrict__ tr,
                                               const float* __restrict__ epaState, float* cand,
                                               unsigned long long* fails, float* work, int* auxWork) {
  float* slot = work + (size_t)kEpaHugeSlotWords * blockIdx.x;
  int* aux = auxWork + (size_t)kEpaHugeAuxInts * blockIdx.x;
  const int jobs = (int)*count;
  for (int k = blockIdx.x; k < jobs; k += gridDim.x)
    epaJob<kEpaHugeFaces, kEpaVerts, kEpaHugeStack>(queue[k], keys, shapes, verts, props, tr, epaState, cand, fails,
                                                    slot, aux, nullptr, nullptr);
}

// ------------------------------------------------------------------ manifold update
struct Pt {
  float* p;
  __device__ V3 g3(int o) const { return ld3(p + o); }
  __device__ void s3(int o, V3 v) { st3(p + o, v); }
};

__device__ __forceinline__ void copyPoint(float* dst, const float* src) {
  for (int k = 0; k < JBGPU_POINT_WORDS; k++) dst[k] = src[k];
}

__device__ void removePoint(float* m, int idx) {  // PersistentManifold.java:259-278
  int n = ldi(m);
  int last = n - 1;
  float* P = m + JBGPU_MANIFOLD_HDR;
  if (idx != last) {
    copyPoint(P + JBGPU_POINT_WORDS * idx, P + JBGPU_POINT_WORDS * last);
    float* L = P + JBGPU_POINT_WORDS * last;
    L[18] = 0.0f;
    sti(L + 27, 0);
    L[19] = 0.0f;
    L[20] = 0.0f;
    sti(L + 28, 0);
  }
  sti(m, n - 1);
}

__global__ void kManifold(int64_t npairs, const int* __restrict__ dn, const uint64_t* __restrict__ keys, const int* __restrict__ ptype,
                          const int* __restrict__ act, const float* __restrict__ props, const float* __restrict__ tr,
                          const float* __restrict__ cand, float* man) {
  if (dn) npairs = *dn;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  int a = (int)(keys[i] >> 32) - 1, b = (int)(keys[i] & 0xffffffffu) - 1;
  if (ptype[i] == JBGPU_PAIR_NONE) return;
  if (!(isActiveState(act[a]) || isActiveState(act[b]))) return;  // needsCollision
  float* m = man + (size_t)JBGPU_MANIFOLD_WORDS * i;
  float* P = m + JBGPU_MANIFOLD_HDR;
  int body0 = ldi(m + 1), body1 = ldi(m + 2);
  const float* c = cand + 8 * i;
  Tr t0 = loadTr(tr, body0), t1 = loadTr(tr, body1);
  if (c[0] != 0.0f && !(c[7] > kContactThreshold)) {
    // ManifoldResult.addContactPoint (ManifoldResult.java:92-157); result body0 = a
    V3 nB = v3(c[1], c[2], c[3]), pB = v3(c[4], c[5], c[6]);
    float depth = c[7];
    V3 pointA = add(scl(nB, depth), pB);
    // localA on the manifold's body0, localB on body1 (swapped when body0 != a)
    V3 localA = invXform(t0, pointA);
    V3 localB = invXform(t1, pB);
    float np_[JBGPU_POINT_WORDS];
    for (int k = 0; k < JBGPU_POINT_WORDS; k++) np_[k] = 0.0f;
    st3(np_ + 0, localA);
    st3(np_ + 3, localB);
    st3(np_ + 6, pointA);
    st3(np_ + 9, pB);
    st3(np_ + 12, nB);
    np_[15] = depth;
    // getCacheEntry (PersistentManifold.java:214-233)
    int n = ldi(m);
    float shortest = kContactThreshold * kContactThreshold;
    int idx = -1;
    for (int k = 0; k < n; k++) {
      V3 d = sub(ld3(P + JBGPU_POINT_WORDS * k), localA);
      float dd = dot(d, d);
      if (dd < shortest) {
        shortest = dd;
        idx = k;
      }
    }
    float fr = props[8 * a + 4] * props[8 * b + 4];
    if (fr < -10.0f) fr = -10.0f;
    if (fr > 10.0f) fr = 10.0f;
    np_[16] = fr;
    np_[17] = props[8 * a + 5] * props[8 * b + 5];
    if (idx >= 0) {  // replaceContactPoint (:280-305)
      float* O = P + JBGPU_POINT_WORDS * idx;
      float ai = O[18], l1 = O[19], l2 = O[20];
      int life = ldi(O + 28);
      copyPoint(O, np_);
      O[18] = ai;
      O[19] = l1;
      O[20] = l2;
      sti(O + 28, life);
    } else {  // addManifoldPoint (:235-257)
      int ins = n;
      if (ins == 4) {  // sortCachedPoints (:83-156)
        int maxPen = -1;
        float maxPenV = depth;
        for (int k = 0; k < 4; k++) {
          float dk = P[JBGPU_POINT_WORDS * k + 15];
          if (dk < maxPenV) {
            maxPen = k;
            maxPenV = dk;
          }
        }
        V3 l0 = ld3(P), l1v = ld3(P + JBGPU_POINT_WORDS), l2v = ld3(P + 2 * JBGPU_POINT_WORDS),
           l3v = ld3(P + 3 * JBGPU_POINT_WORDS);
        float r0 = 0, r1 = 0, r2 = 0, r3 = 0;
        if (maxPen != 0) r0 = len2(crs(sub(localA, l1v), sub(l3v, l2v)));
        if (maxPen != 1) r1 = len2(crs(sub(localA, l0), sub(l3v, l2v)));
        if (maxPen != 2) r2 = len2(crs(sub(localA, l0), sub(l3v, l1v)));
        if (maxPen != 3) r3 = len2(crs(sub(localA, l0), sub(l2v, l1v)));
        float v[4] = {jabs(r0), jabs(r1), jabs(r2), jabs(r3)};
        int mi = -1;
        float mv = -1e30f;
        for (int k = 0; k < 4; k++)
          if (v[k] > mv) {
            mi = k;
            mv = v[k];
          }
        ins = mi;
      } else {
        sti(m, n + 1);
      }
      copyPoint(P + JBGPU_POINT_WORDS * ins, np_);
    }
  }
  // refreshContactPoints (PersistentManifold.java:312-372) if the manifold has points
  int n = ldi(m);
  if (n == 0) return;
  // a point in 16-byte words (records and points are 16-byte aligned): [0] localA, localB.x; [1] localB.yz,
  // worldA.xy; [2] worldA.z, worldB; [3] normal, distance; [7] lifetime and three spare words. The keep/drop
  // test of point k reads only its own refreshed values, and removing a point k moves a point of a higher index
  // (already tested) into k, so the tests are taken in the refresh pass and the removals follow, from the top
  unsigned int drop = 0;
  for (int k = n - 1; k >= 0; k--) {
    float4* Q = reinterpret_cast<float4*>(P + JBGPU_POINT_WORDS * k);
    const float4 q0 = Q[0], q1 = Q[1], q3 = Q[3], q7 = Q[7];
    const V3 wa = xform(t0, v3(q0.x, q0.y, q0.z));
    const V3 wb = xform(t1, v3(q0.w, q1.x, q1.y));
    const V3 nrm = v3(q3.x, q3.y, q3.z);
    const float d = dot(sub(wa, wb), nrm);
    Q[1] = make_float4(q1.x, q1.y, wa.x, wa.y);
    Q[2] = make_float4(wa.z, wb.x, wb.y, wb.z);
    Q[3] = make_float4(q3.x, q3.y, q3.z, d);
    Q[7] = make_float4(__builtin_bit_cast(float, __builtin_bit_cast(int, q7.x) + 1), q7.y, q7.z, q7.w);
    bool remove;
    if (!(d <= kContactThreshold)) {
      remove = true;
    } else {
      const V3 proj = sub(wa, scl(nrm, d));
      const V3 pd = sub(wb, proj);
      remove = dot(pd, pd) > kContactThreshold * kContactThreshold;
    }
    if (remove) drop |= 1u << k;
  }
  for (int k = n - 1; k >= 0; k--)
    if ((drop >> k) & 1u) removePoint(m, k);
}

// contact points over all manifolds (per-block sums, one atomic per block); run only when statistics are read
__global__ __launch_bounds__(256) void kCountPoints(int64_t n, const float* __restrict__ man,
                                                    unsigned long long* out) {
  __shared__ unsigned long long part[4];
  unsigned long long k = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    k += (unsigned long long)ldi(man + (size_t)JBGPU_MANIFOLD_WORDS * i);
  for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(out, t);
  }
}


// ------------------------------------------------------------------ rayTest (CollisionWorld.java:552-591)
__device__ __forceinline__ int rayOutcode(V3 p, V3 h) {  // AabbUtil2.java:40-43
  return (p.x < -h.x ? 0x01 : 0x0) | (p.x > h.x ? 0x08 : 0x0) | (p.y < -h.y ? 0x02 : 0x0) | (p.y > h.y ? 0x10 : 0x0) |
         (p.z < -h.z ? 0x4 : 0x0) | (p.z > h.z ? 0x20 : 0x0);
}
// AabbUtil2.rayAabb (AabbUtil2.java:45-104): param in, entry fraction out
__device__ bool rayAabb(V3 from, V3 to, V3 mn, V3 mx, float& param) {
  V3 he = scl(sub(mx, mn), 0.5f);
  V3 c = scl(add(mx, mn), 0.5f);
  V3 source = sub(from, c), target = sub(to, c);
  const int so = rayOutcode(source, he), tc = rayOutcode(target, he);
  if ((so & tc) != 0x0) return false;
  float enter = 0.0f, exit = param;
  V3 r = sub(target, source);
  float normSign = 1.0f;
  int bit = 1;
  for (int j = 0; j < 2; j++) {
    for (int i = 0; i != 3; ++i) {
      if ((so & bit) != 0) {
        float l = (-coord(source, i) - coord(he, i) * normSign) / coord(r, i);
        if (enter <= l) enter = l;
      } else if ((tc & bit) != 0) {
        float l = (-coord(source, i) - coord(he, i) * normSign) / coord(r, i);
        exit = jmin(exit, l);
      }
      bit <<= 1;
    }
    normSign = -1.0f;
  }
  if (enter <= exit) {
    param = enter;
    return true;
  }
  return false;
}
// conservative segment-box test for tree nodes (double precision, box widened): never rejects a box that the
// exact float test above accepts
__device__ __forceinline__ bool segmentHitsBox(V3 F, V3 T, float4 mn, float4 mx) {
  double t0 = 0.0, t1 = 1.0;
  const float fo[3] = {F.x, F.y, F.z}, ft[3] = {T.x, T.y, T.z};
  const float lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
  for (int a = 0; a < 3; a++) {
    const double l = lo[a], h = hi[a];
    const double eps = 1e-5 * (fabs(l) + fabs(h) + fabs((double)fo[a]) + fabs((double)ft[a])) + 1e-6;
    const double o = fo[a], d = (double)ft[a] - (double)fo[a];
    if (d == 0.0) {
      if (o < l - eps || o > h + eps) return false;
    } else {
      double ta = (l - eps - o) / d, tb = (h + eps - o) / d;
      if (ta > tb) {
        const double x = ta;
        ta = tb;
        tb = x;
      }
      t0 = ta > t0 ? ta : t0;
      t1 = tb < t1 ? tb : t1;
      if (t0 > t1) return false;
    }
  }
  return true;
}
__device__ __forceinline__ bool rayNeeds(int filt, int group, int mask) {  // RayResultCallback.needsCollision
  const int pg = (int)(short)(filt & 0xffff), pm = (int)(short)(filt >> 16);
  return ((pg & mask) & 0xFFFF) != 0 && ((group & pm) & 0xFFFF) != 0;
}

// candidate objects of every ray: filter and rayAabb with the callback's initial fraction (1); the count pass
// stores the number per ray, the emit pass writes (ray << 32 | object) at the ray's offset.  With `expand`
// (min xyz, max xyz: a convex sweep's cast-shape AABB) every object box is widened by it first, as
// AabbUtil2.aabbExpand does (node boxes alike: float addition keeps them around their leaves).  A query's
// from, to and expand values start at stride * query.
constexpr int kRayStack = 64;
template <bool Emit>
__global__ void kRayCandidates(int nr, const float* __restrict__ from, const float* __restrict__ to,
                               const float* __restrict__ expand, int stride,
                               const int* __restrict__ rworld, int group, int mask, int nn,
                               const float4* __restrict__ rec, const int4* __restrict__ leafExact,
                               const int* __restrict__ hugeRange, const int* __restrict__ hugeList,
                               const float* __restrict__ aabb, const int* __restrict__ filt,
                               const int* __restrict__ wid, int maxWorld, int* count,
                               const int64_t* __restrict__ offset, uint64_t* keys, int* err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const float* fp = from + (size_t)stride * r;
  const float* tp = to + (size_t)stride * r;
  const V3 F = v3(fp[0], fp[1], fp[2]), T = v3(tp[0], tp[1], tp[2]);
  V3 emn = v3(0, 0, 0), emx = v3(0, 0, 0);
  if (expand) {
    const float* ep = expand + (size_t)stride * r;
    emn = v3(ep[0], ep[1], ep[2]);
    emx = v3(ep[3], ep[4], ep[5]);
  }
  const int w = rworld ? rworld[r] : 0;
  if (w < 0 || w > maxWorld) {  // a world id the proxies never had: no candidates, and the call reports it
    if (!Emit) {
      count[r] = 0;
      atomicOr(err, 2);
    }
    return;
  }
  int cnt = 0;
  const int64_t base = Emit ? offset[r] : 0;
  const int room = Emit ? count[r] : 0;  // the emit pass writes no more than the count pass found
  auto consider = [&](int o, V3 mn, V3 mx, int f) {
    if (wid[o] != w || !rayNeeds(f, group, mask)) return;  // other worlds are other collision-object lists
    if (expand) {
      mn = add(mn, emn);
      mx = add(mx, emx);
    }
    float param = 1.0f;
    if (!rayAabb(F, T, mn, mx, param)) return;
    if (Emit && cnt < room) keys[base + cnt] = ((uint64_t)(uint32_t)r << 32) | (uint32_t)o;
    cnt++;
  };
  for (int k = hugeRange[2 * w]; k < hugeRange[2 * w + 1]; k++) {
    const int o = hugeList[k];
    const float* A = aabb + 6 * o;
    consider(o, v3(A[0], A[1], A[2]), v3(A[3], A[4], A[5]), filt[o]);
  }
  auto leaf = [&](int t) {
    const int4 e0 = leafExact[2 * t], e1 = leafExact[2 * t + 1];
    consider(e1.w, v3(__int_as_float(e0.x), __int_as_float(e0.y), __int_as_float(e0.z)),
             v3(__int_as_float(e0.w), __int_as_float(e1.x), __int_as_float(e1.y)), e1.z);
  };
  if (nn == 1) {
    leaf(0);
  } else if (nn >= 2) {
    int stack[kRayStack];
    int sp = 0, node = 0;
    while (true) {
      const float4* R = rec + kRecF4 * node;
      const float4 m = R[4];
      int next = -1;
      for (int c = 0; c < 2; c++) {
        const int ch = __float_as_int(c == 0 ? m.x : m.y);
        float4 bmn = R[2 * c], bmx = R[2 * c + 1];
        if (w < __float_as_int(bmn.w) || w > __float_as_int(bmx.w)) continue;
        if (expand) {
          bmn.x += emn.x; bmn.y += emn.y; bmn.z += emn.z;
          bmx.x += emx.x; bmx.y += emx.y; bmx.z += emx.z;
        }
        if (!segmentHitsBox(F, T, bmn, bmx)) continue;
        if (ch < 0) {
          leaf(-ch - 1);
        } else if (next < 0) {
          next = ch;
        } else if (sp < kRayStack) {
          stack[sp++] = ch;
        } else {
          atomicOr(err, 1);
        }
      }
      if (next >= 0) node = next;
      else if (sp > 0) node = stack[--sp];
      else break;
    }
  }
  if (!Emit) count[r] = cnt;
}

// planeSpace1 (TransformUtil.java:45-61)
__device__ __forceinline__ void planeSpace1(V3 n, V3& p, V3& q) {
  if (jabs(n.z) > 0.7071067811865475244008443621048490f) {
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
__device__ __forceinline__ V3 interp3(V3 a, V3 b, float t) {  // VectorUtil.setInterpolate3
  const float s = 1.0f - t;
  return v3(s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z);
}
// TriangleRaycastCallback.processTriangle (TriangleRaycastCallback.java:51-124) without its `distance <
// hitFraction` test (that one is made where the fraction is known; nothing else depends on it)
__device__ bool rayTriangleGeom(V3 a, V3 b, V3 cc, V3 from, V3 to, float& distance, V3& normal) {
  V3 tn = crs(sub(b, a), sub(cc, a));
  float dist = dot(a, tn);
  float da = dot(tn, from);
  da -= dist;
  float db = dot(tn, to);
  db -= dist;
  if (da * db >= 0.0f) return false;
  float proj = da - db;
  distance = da / proj;
  float tol = len2(tn);
  tol *= -0.0001f;
  V3 point = interp3(from, to, distance);
  V3 v0p = sub(a, point), v1p = sub(b, point);
  if (!(dot(crs(v0p, v1p), tn) >= tol)) return false;
  V3 v2p = sub(cc, point);
  if (!(dot(crs(v1p, v2p), tn) >= tol)) return false;
  if (!(dot(crs(v2p, v0p), tn) >= tol)) return false;
  normal = da > 0.0f ? tn : scl(tn, -1.0f);
  return true;
}

// SubsimplexConvexCast.calcTimeOfImpact (SubsimplexConvexCast.java:61-196) for the ray's point (SphereShape(0),
// margin 0: support at its transform's origin) against shape sb at tb
__device__ bool subsimplexRayCast(Simplex& ss, const Shape& sb, const float* verts, V3 F, V3 T, const Tr& tb,
                                  float& fraction, V3& normal) {
  ss.reset();
  Tr fromA;
  fromA.b = mident();
  fromA.o = F;
  V3 linVelA = sub(T, F), linVelB = sub(tb.o, tb.o);
  float lambda = 0.0f;
  Tr iA = fromA, iB = tb;
  V3 r = sub(linVelA, linVelB);
  V3 supA = xform(fromA, v3(0, 0, 0));
  V3 supB = xform(tb, supportMargin(sb, verts, mulMtV(tb.b, r)));
  V3 v = sub(supA, supB);
  int maxIter = 32;
  V3 n = v3(0, 0, 0);
  float dist2 = len2(v);
  const float epsilon = 0.0001f;
  while ((dist2 > epsilon) && (maxIter--) != 0) {
    supA = xform(iA, v3(0, 0, 0));
    supB = xform(iB, supportMargin(sb, verts, mulMtV(iB.b, v)));
    V3 w = sub(supA, supB);
    float VdotW = dot(v, w);
    if (lambda > 1.0f) return false;
    if (VdotW > 0.0f) {
      float VdotR = dot(v, r);
      if (VdotR >= -(kFltEps * kFltEps)) return false;
      lambda = lambda - VdotW / VdotR;
      const float s = 1.0f - lambda;
      iA.o = v3(s * F.x + lambda * T.x, s * F.y + lambda * T.y, s * F.z + lambda * T.z);
      iB.o = v3(s * tb.o.x + lambda * tb.o.x, s * tb.o.y + lambda * tb.o.y, s * tb.o.z + lambda * tb.o.z);
      w = sub(supA, supB);
      n = v;
    }
    ss.addVertex(w, supA, supB);
    const bool ok = ss.update();
    v = ss.cV;
    dist2 = ok ? len2(v) : 0.0f;
  }
  fraction = lambda;
  normal = len2(n) >= kFltEps * kFltEps ? nor(n) : v3(0, 0, 0);
  if (dot(normal, r) >= -0.0f) return false;  // CastResult.allowedPenetration = 0
  return true;
}

// per candidate (sorted by ray, then object): the object's hit, independent of the ray's running fraction.
// record: [flags (bit0 hit / first triangle, bit1 second triangle, bit2 plane), f0, n0 xyz, f1, n1 xyz, -]
constexpr int kRayRec = 10;
__global__ __launch_bounds__(kGjkBlock) void kRayCast(int64_t n, const uint64_t* __restrict__ keys,
                                                      const float* __restrict__ from, const float* __restrict__ to,
                                                      const float* __restrict__ shapes,
                                                      const float* __restrict__ props, const float* __restrict__ tr,
                                                      const float* __restrict__ verts, float* out) {
  __shared__ float simplexLds[kSimplexRows * kGjkBlock];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int r = (int)(k >> 32), o = (int)(k & 0xffffffffu);
  const V3 F = v3(from[3 * r], from[3 * r + 1], from[3 * r + 2]), T = v3(to[3 * r], to[3 * r + 1], to[3 * r + 2]);
  const Shape sb = loadShape(shapes + JBGPU_SHAPE_WORDS * ldi(props + 8 * o + 6));
  const Tr tb = loadTr(tr, o);
  float* R = out + kRayRec * i;
  int flags = 0;
  float f0 = 0.0f, f1 = 0.0f;
  V3 n0 = v3(0, 0, 0), n1 = v3(0, 0, 0);
  if (sb.type != JBGPU_SHAPE_PLANE) {
    Simplex ss;
    ss.lw = simplexLds + threadIdx.x;
    V3 n;
    if (subsimplexRayCast(ss, sb, verts, F, T, tb, f0, n) && len2(n) > 0.0001f) {
      flags = 1;
      n0 = nor(mulMV(mident(), n));  // castResult.normal.mul(rayFromTrans.basis), then nor()
    }
  } else {  // concave branch: processAllTriangles in the plane's frame (StaticPlaneShape.java:60-121)
    flags = 4;
    const Tr w2o = trInverse(tb);
    const V3 fl = xform(w2o, F), tl = xform(w2o, T);
    const V3 rmin = v3(jmin(fl.x, tl.x), jmin(fl.y, tl.y), jmin(fl.z, tl.z));
    const V3 rmax = v3(jmax(fl.x, tl.x), jmax(fl.y, tl.y), jmax(fl.z, tl.z));
    const float radius = len(scl(rmin, 0.5f));  // halfExtents set from aabbMax, then from aabbMin (kept)
    const V3 center = scl(add(rmax, rmin), 0.5f);
    V3 t0, t1;
    planeSpace1(sb.pn, t0, t1);
    const V3 pc = sub(center, scl(sb.pn, dot(sb.pn, center) - sb.pc));
    const V3 a = scl(t0, radius), b = scl(t1, radius);
    const V3 ppp = v3(pc.x + a.x + b.x, pc.y + a.y + b.y, pc.z + a.z + b.z);
    V3 nl;
    if (rayTriangleGeom(ppp, add(pc, sub(a, b)), sub(pc, sub(a, b)), fl, tl, f0, nl)) {
      flags |= 1;
      n0 = mulMV(tb.b, nl);  // addSingleResult with the normal in the object's frame
    }
    if (rayTriangleGeom(sub(pc, sub(a, b)), sub(pc, add(a, b)), ppp, fl, tl, f1, nl)) {
      flags |= 2;
      n1 = mulMV(tb.b, nl);
    }
  }
  R[0] = __int_as_float(flags);
  R[1] = f0;
  R[2] = n0.x; R[3] = n0.y; R[4] = n0.z;
  R[5] = f1;
  R[6] = n1.x; R[7] = n1.y; R[8] = n1.z;
  R[9] = 0.0f;
}

// ClosestRayResultCallback over a ray's candidates in world order: rayAabb with the running fraction, then
// the object's hit if its fraction is below it (CollisionWorld.java:566-589, :282-298, :710-727)
__global__ void kRayFold(int nr, const int* __restrict__ count, const int64_t* __restrict__ offset,
                         const uint64_t* __restrict__ keys, const float* __restrict__ recs,
                         const float* __restrict__ from, const float* __restrict__ to,
                         const float* __restrict__ aabb, int* obj, float* frac, float* normal, float* point) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const V3 F = v3(from[3 * r], from[3 * r + 1], from[3 * r + 2]), T = v3(to[3 * r], to[3 * r + 1], to[3 * r + 2]);
  float best = 1.0f;
  int hit = -1;
  V3 hn = v3(0, 0, 0), hp = v3(0, 0, 0);
  const int64_t b = offset[r], e = b + count[r];
  for (int64_t k = b; k < e; k++) {
    if (best == 0.0f) break;
    const int o = (int)(keys[k] & 0xffffffffu);
    const float* A = aabb + 6 * o;
    float param = best;
    if (!rayAabb(F, T, v3(A[0], A[1], A[2]), v3(A[3], A[4], A[5]), param)) continue;
    const float* R = recs + kRayRec * k;
    const int flags = __float_as_int(R[0]);
    if ((flags & 1) && R[1] < best) {
      best = R[1];
      hit = o;
      hn = v3(R[2], R[3], R[4]);
      hp = interp3(F, T, best);
    }
    if ((flags & 2) && R[5] < best) {  // the plane's second triangle, against the updated fraction
      best = R[5];
      hit = o;
      hn = v3(R[6], R[7], R[8]);
      hp = interp3(F, T, best);
    }
  }
  obj[r] = hit;
  frac[r] = best;
  normal[3 * r] = hn.x; normal[3 * r + 1] = hn.y; normal[3 * r + 2] = hn.z;
  point[3 * r] = hp.x; point[3 * r + 1] = hp.y; point[3 * r + 2] = hp.z;
}

// ------------------------------------------------------------------ convex sweeps (CollisionWorld.java:596-646)
// MatrixUtil.invert (MatrixUtil.java:366-394): cofactors over the determinant
__device__ __forceinline__ float cofac(const M3& m, int r1, int c1, int r2, int c2) {
  return m.m[3 * r1 + c1] * m.m[3 * r2 + c2] - m.m[3 * r1 + c2] * m.m[3 * r2 + c1];
}
__device__ M3 invertM(const M3& a) {
  const float cx = cofac(a, 1, 1, 2, 2), cy = cofac(a, 1, 2, 2, 0), cz = cofac(a, 1, 0, 2, 1);
  const float det = a.m[0] * cx + a.m[1] * cy + a.m[2] * cz;
  const float s = 1.0f / det;
  M3 r;
  r.m[0] = cx * s;
  r.m[1] = cofac(a, 0, 2, 2, 1) * s;
  r.m[2] = cofac(a, 0, 1, 1, 2) * s;
  r.m[3] = cy * s;
  r.m[4] = cofac(a, 0, 0, 2, 2) * s;
  r.m[5] = cofac(a, 0, 2, 1, 0) * s;
  r.m[6] = cz * s;
  r.m[7] = cofac(a, 0, 1, 2, 0) * s;
  r.m[8] = cofac(a, 0, 0, 1, 1) * s;
  return r;
}
// TransformUtil.calculateVelocity (TransformUtil.java:105-140): origin difference over the step; the rotation
// basis1 * basis0^-1 as a normalised quaternion gives angle 2 acos(w) about its vector part (x when that vanishes)
__device__ void calcVelocity(const Tr& t0, const Tr& t1, float timeStep, V3& lin, V3& ang) {
  lin = scl(sub(t1.o, t0.o), 1.0f / timeStep);
  const Q4 d = qnor(qFromMatrix(mulMM(t1.b, invertM(t0.b))));
  const float angle = 2.0f * (float)acos((double)d.w);
  V3 axis = v3(d.x, d.y, d.z);
  const float l = len2(axis);
  if (l < kFltEps * kFltEps)
    axis = v3(1.0f, 0.0f, 0.0f);
  else
    axis = scl(axis, 1.0f / (float)sqrt((double)l));
  ang = scl(axis, angle / timeStep);
}
// CollisionShape.calculateTemporalAabb and getAngularMotionDisc (CollisionShape.java:65-100): the AABB grown on
// the side of the linear motion, then on all sides by |angvel| dt times the half diagonal of the identity AABB
// plus the distance of its centre
__device__ void temporalAabb(const Shape& sh, const Tr& cur, V3 lv, V3 av, float timeStep, V3& mn, V3& mx) {
  shapeAabb(sh, cur, mn, mx);
  const V3 lm = scl(lv, timeStep);
  if (lm.x > 0.0f) mx.x += lm.x; else mn.x += lm.x;
  if (lm.y > 0.0f) mx.y += lm.y; else mn.y += lm.y;
  if (lm.z > 0.0f) mx.z += lm.z; else mn.z += lm.z;
  Tr id;
  id.b = mident();
  id.o = v3(0, 0, 0);
  V3 a0, a1;
  shapeAabb(sh, id, a0, a1);
  const float disc = len(sub(a1, a0)) * 0.5f + len(scl(add(a0, a1), 0.5f));
  const float am = len(av) * disc * timeStep;
  mn = sub(mn, v3(am, am, am));
  mx = add(mx, v3(am, am, am));
}
// the cast shape's AABB over a sweep (CollisionWorld.java:608-617): velocities over a unit step, the shape at
// from's rotation (through a quaternion) and the origin
__device__ void castAabb(const Shape& cs, const Tr& from, const Tr& to, V3& mn, V3& mx) {
  V3 lv, av;
  calcVelocity(from, to, 1.0f, lv, av);
  Tr R;
  R.b = mFromQ(qFromMatrix(from.b));
  R.o = v3(0, 0, 0);
  temporalAabb(cs, R, lv, av, 1.0f, mn, mx);
}
__device__ __forceinline__ Shape sphereShape(float r) {  // SphereShape(r): margin = radius
  Shape s;
  s.type = JBGPU_SHAPE_SPHERE;
  s.margin = r * 1.0f;
  s.dims = v3(r, 0, 0);
  s.vfirst = s.vcount = 0;
  s.lmin = s.lmax = s.pn = v3(0, 0, 0);
  s.pc = 0.0f;
  return s;
}

// GjkConvexCast.calcTimeOfImpact (GjkConvexCast.java:67-175) for a moving against b at rest: conservative
// advancement along the GJK normal (no penetration solver, unlimited distance) while the distance exceeds 0.001,
// at most 32 steps; the point collector keeps the smallest distance over all its calls (PointCollector.java)
template <int S>
__device__ bool gjkConvexCast(const Shape& sa, const Tr& fromA, const Tr& toA, const Shape& sb, const Tr& tb,
                              const float* __restrict__ verts, float allowedPenetration, SimplexT<S>& vs,
                              float& fraction, V3& normal, V3& hitPoint) {
  const V3 r = sub(sub(toA.o, fromA.o), sub(tb.o, tb.o));
  const float radius = 0.001f;
  float lambda = 0.0f, lastLambda = 0.0f;
  int numIter = 0;
  bool pcHas = false;
  float pcDist = 1e30f;
  V3 pcN = v3(0, 0, 0), pcP = v3(0, 0, 0);
  auto collect = [&](const GjkOut& g) {
    if (g.isValid && g.distance < pcDist) {
      pcHas = true;
      pcN = g.normalInB;
      pcP = add(g.pointOnB, g.off);
      pcDist = g.distance;
    }
  };
  Tr ia = fromA, ib = tb;
  GjkOut g;
  gjkClosest(sa, ia, sb, ib, verts, kInf, vs, g);
  collect(g);
  if (!pcHas) return false;
  V3 c = pcP, n = pcN;
  float dist = pcDist;
  while (dist > radius) {
    numIter++;
    if (numIter > 32) return false;
    const float dLambda = dist / dot(r, n);
    lambda = lambda - dLambda;
    if (lambda > 1.0f) return false;
    if (lambda < 0.0f) return false;
    if (lambda <= lastLambda) return false;
    lastLambda = lambda;
    ia.o = interp3(fromA.o, toA.o, lambda);
    ib.o = interp3(tb.o, tb.o, lambda);
    gjkClosest(sa, ia, sb, ib, verts, kInf, vs, g);
    collect(g);
    if (!pcHas) return false;
    if (pcDist < 0.0f) {
      fraction = lastLambda;
      normal = pcN;
      hitPoint = pcP;
      return true;
    }
    c = pcP;
    n = pcN;
    dist = pcDist;
  }
  if (dot(n, r) >= -allowedPenetration) return false;  // no time of impact for motion away from the normal
  fraction = lambda;
  normal = n;
  hitPoint = c;
  return true;
}

// per query: the origins of from and to, and the cast shape's AABB over the sweep
__global__ void kSweepPrep(int nq, const float* __restrict__ castRec, const float* __restrict__ from,
                           const float* __restrict__ to, float* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const Shape cs = loadShape(castRec);
  const Tr f = loadTr(from, q), t = loadTr(to, q);
  V3 mn, mx;
  castAabb(cs, f, t, mn, mx);
  float* o = out + 12 * (size_t)q;
  st3(o, f.o);
  st3(o + 3, t.o);
  st3(o + 6, mn);
  st3(o + 9, mx);
}

// per candidate (sorted by query, then object): objectQuerySingle's convex branch (CollisionWorld.java:392-424)
// up to the fraction test: the cast with allowedCcdPenetration 0.04 and a normal longer than 0.01, normalised.
// Planes are not convex and have no caster there: never a hit.  record: [hit, fraction, normal xyz, point xyz]
constexpr int kSweepRec = 8;
__global__ __launch_bounds__(kGjkBlock) void kSweepCast(int64_t n, const uint64_t* __restrict__ keys,
                                                        const float* __restrict__ castRec,
                                                        const float* __restrict__ from, const float* __restrict__ to,
                                                        const float* __restrict__ shapes,
                                                        const float* __restrict__ props, const float* __restrict__ tr,
                                                        const float* __restrict__ verts, float* out) {
  __shared__ float simplexLds[kSimplexRows * kGjkBlock];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int q = (int)(k >> 32), o = (int)(k & 0xffffffffu);
  const Shape cs = loadShape(castRec), sb = bodyShape(shapes, props, o);
  float* R = out + kSweepRec * i;
  float fr = 1.0f;
  V3 nrm = v3(0, 0, 0), pt = v3(0, 0, 0);
  bool hit = false;
  if (sb.type != JBGPU_SHAPE_PLANE) {
    Simplex vs;
    vs.lw = simplexLds + threadIdx.x;
    if (gjkConvexCast(cs, loadTr(from, q), loadTr(to, q), sb, loadTr(tr, o), verts, 0.04f, vs, fr, nrm, pt) &&
        len2(nrm) > 0.0001f) {
      hit = true;
      nrm = nor(nrm);
    }
  }
  R[0] = hit ? 1.0f : 0.0f;
  R[1] = fr;
  st3(R + 2, nrm);
  st3(R + 5, pt);
}

// ClosestConvexResultCallback over a query's candidates in world order: a hit below the running fraction
// replaces it (CollisionWorld.java:766-810)
__global__ void kSweepFold(int nq, const int* __restrict__ count, const int64_t* __restrict__ offset,
                           const uint64_t* __restrict__ keys, const float* __restrict__ recs, int* obj, float* frac,
                           float* normal, float* point) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  float best = 1.0f;
  int hit = -1;
  V3 hn = v3(0, 0, 0), hp = v3(0, 0, 0);
  const int64_t b = offset[q], e = b + count[q];
  for (int64_t k = b; k < e; k++) {
    const float* R = recs + kSweepRec * k;
    if (R[0] != 0.0f && R[1] < best) {
      best = R[1];
      hit = (int)(keys[k] & 0xffffffffu);
      hn = v3(R[2], R[3], R[4]);
      hp = v3(R[5], R[6], R[7]);
    }
  }
  obj[q] = hit;
  frac[q] = best;
  st3(normal + 3 * (size_t)q, hn);
  st3(point + 3 * (size_t)q, hp);
}

// ------------------------------------------------------------------ CCD motion clamping
// DiscreteDynamicsWorld.integrateTransforms (DiscreteDynamicsWorld.java:679-739) for the bodies with a motion
// threshold, in body order, after the integrate kernel moved every body: a body whose predicted step exceeds its
// threshold sweeps a sphere of its swept radius from its old to its predicted origin through the world as it is
// at that point of the loop (earlier bodies at their new transforms, later ones at their old), with
// ClosestNotMeConvexResultCallback (:1129-1198): not itself, its own filter, no body it already has contact
// points with, no hit for motion along or away from the normal.  A hit above 0.0001 of the step re-integrates
// the body over that fraction of the step.  One workgroup walks the list; its threads split the objects.
constexpr int kCcdBlock = 512;
__device__ __forceinline__ bool pairHasPoints(int a, int b, int64_t np, const uint64_t* __restrict__ keys,
                                              const float* __restrict__ man) {
  const uint64_t lo = (uint64_t)min(a, b) + 1, hi = (uint64_t)max(a, b) + 1;
  const uint64_t key = (lo << 32) | hi;
  int64_t l = 0, e = np;
  while (l < e) {
    const int64_t m = (l + e) / 2;
    if (keys[m] < key) l = m + 1; else e = m;
  }
  return l < np && keys[l] == key && ldi(man + (size_t)JBGPU_MANIFOLD_WORDS * l) > 0;
}
__global__ __launch_bounds__(kCcdBlock) void kCcd(int nlist, const int* __restrict__ list,
                                                  const float* __restrict__ ccd, int nb, float dt,
                                                  const float* __restrict__ trOld, float* tr,
                                                  const float* __restrict__ vel, const float* __restrict__ props,
                                                  const int* __restrict__ act, const float* __restrict__ shapes,
                                                  const float* __restrict__ verts, const int* __restrict__ filt,
                                                  const int* __restrict__ wid, int64_t np, const int* __restrict__ dnp,
                                                  const uint64_t* __restrict__ pkeys, const float* __restrict__ man,
                                                  const int* __restrict__ abortFlag) {
  __shared__ float simplexLds[kSimplexRows * kCcdBlock];
  if (abortFlag && *abortFlag) return;  // an aborted graph-captured step changes nothing
  if (dnp) np = *dnp;
  __shared__ float red[kCcdBlock / 64];
  SimplexT<kCcdBlock> vs;
  vs.lw = simplexLds + threadIdx.x;
  for (int k = 0; k < nlist; k++) {
    const int i = list[k];
    if ((ldi(props + 8 * i + 7) & JBGPU_BODY_STATIC) != 0 || !isActiveState(act[i])) continue;  // not integrated
    const Shape si = bodyShape(shapes, props, i);
    if (si.type == JBGPU_SHAPE_PLANE) continue;  // only convex shapes are clamped
    const Tr from = loadTr(trOld, i), pred = loadTr(tr, i);
    const float thr = ccd[2 * i];
    const float sq = thr * thr;
    const float squareMotion = len2(sub(pred.o, from.o));
    if (!(sq != 0.0f && sq < squareMotion)) continue;
    const Shape cast = sphereShape(ccd[2 * i + 1]);
    V3 cmn, cmx;
    castAabb(cast, from, pred, cmn, cmx);
    const V3 relVel = sub(sub(pred.o, from.o), v3(0, 0, 0));
    const int fi = filt[i];
    const int group = (int)(short)(fi & 0xffff), mask = (int)(short)(fi >> 16), w = wid[i];
    float best = 1.0f;
    for (int j = threadIdx.x; j < nb; j += kCcdBlock) {
      if (j == i || wid[j] != w || !rayNeeds(filt[j], group, mask)) continue;
      const Tr tj = loadTr(j < i ? (const float*)tr : trOld, j);
      const Shape sj = bodyShape(shapes, props, j);
      V3 mn, mx;
      shapeAabb(sj, tj, mn, mx);
      mn = add(mn, cmn);
      mx = add(mx, cmx);
      float lambda = 1.0f;
      if (!rayAabb(from.o, pred.o, mn, mx, lambda)) continue;
      if (sj.type == JBGPU_SHAPE_PLANE) continue;
      if (pairHasPoints(i, j, np, pkeys, man)) continue;  // needsResponse holds: i is not static
      float fr;
      V3 nrm, pt;
      if (!gjkConvexCast(cast, from, pred, sj, tj, verts, 0.04f, vs, fr, nrm, pt)) continue;
      if (!(len2(nrm) > 0.0001f)) continue;
      if (dot(nor(nrm), relVel) >= -0.0f) continue;
      if (fr < best) best = fr;
    }
    for (int o = 32; o > 0; o >>= 1) best = fminf(best, __shfl_xor(best, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = red[0];
      for (int q = 1; q < kCcdBlock / 64; q++) m = fminf(m, red[q]);
      if (m < 1.0f && m > 0.0001f) {  // hasHit and the closestHitFraction test
        const Tr r = integrateTr(from, ld3(vel + 8 * i), ld3(vel + 8 * i + 4), dt * m);
        float* T = tr + 12 * i;
        for (int q = 0; q < 9; q++) T[q] = r.b.m[q];
        T[9] = r.o.x;
        T[10] = r.o.y;
        T[11] = r.o.z;
      }
    }
    __syncthreads();
  }
}

}  // namespace

int64_t npCountContacts(Ctx& c) {
  if (c.npairs == 0 || !c.man[c.cur].p) return 0;
  c.counter.ensure(8);
  unsigned long long* out = c.counter.p + 7;  // words 0-6 hold the step's own counters
  JB_CHECK(hipMemsetAsync(out, 0, sizeof(*out), c.stream));
  hipLaunchKernelGGL(kCountPoints, dim3(std::min(gridFor(c.npairs), 1024)), dim3(256), 0, c.stream, c.npairs,
                     (const float*)c.man[c.cur].p, out);
  unsigned long long k = 0;
  JB_CHECK(hipMemcpyAsync(&k, out, sizeof(k), hipMemcpyDeviceToHost, c.stream));
  JB_CHECK(hipStreamSynchronize(c.stream));
  return (int64_t)k;
}

void npDispatchAll(Ctx& c) {
  hipStream_t s = c.stream;
  const int64_t n = c.npairs;
  c.st.epaCalls = c.st.epaFails = c.st.epaRetries = 0;
  if (n == 0) return;
  c.typeKey.ensure(n);
  c.typeKeyAlt.ensure(n);
  c.perm.ensure(n);
  c.permAlt.ensure(n);
  c.candC.ensure(8 * n);
  c.epaState.ensure(12 * n);
  c.epaQueue.ensure(n);
  c.counter.ensure(8);
  const uint64_t* keys = c.keys[c.cur].p;
  float* man = c.man[c.cur].p;
  JB_CHECK(hipMemsetAsync(c.counter.p, 0, 8 * sizeof(unsigned long long), s));
  tBegin(c, T_BIN);
  hipLaunchKernelGGL(kBinKeys, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)nullptr, c.ptype.p, keys, c.act.p,
                     (const float*)c.shapes.p, (const float*)c.props.p, c.typeKey.p, c.perm.p);
  Db<int> kk(c.typeKey.p, c.typeKeyAlt.p), vv(c.perm.p, c.permAlt.p);
  const int keyBits = 2 + kBinClassBits;
  primSortPairs(c, kk, vv, (int)n, 0, keyBits);
  int* starts = (int*)(c.counter.p + 4);  // 4 ints
  JB_CHECK(hipMemsetAsync(starts, 0xff, 4 * sizeof(int), s));
  hipLaunchKernelGGL(kBinStarts, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)nullptr, (const int*)kk.Current(), starts);
  // per-type ranges stay on the device: each type's kernel covers every pair position and reads its range
  c.binRange.ensure(8);
  int* range = c.binRange.p;
  hipLaunchKernelGGL(kBinRanges, dim3(1), dim3(64), 0, s, n, (const int*)nullptr, (const int*)starts, range);
  tEnd(c, T_BIN);
  const int* list = vv.Current();
  // EPA queue length, failures, workspace overflows, large-tier and last-tier queue lengths: read back with the
  // step's final synchronisation
  c.npStat.ensure(5);
  JB_CHECK(hipMemsetAsync(c.npStat.p, 0, 5 * sizeof(unsigned long long), s));
  hipLaunchKernelGGL(kSphereSphere, dim3(gridFor(n)), dim3(256), 0, s, (const int*)range + 2 * JBGPU_PAIR_SPHERE_SPHERE,
                     list, keys, c.shapes.p, c.props.p, c.tr.p, c.candC.p);
  hipLaunchKernelGGL(kConvexPlane, dim3(gridFor(n)), dim3(256), 0, s, (const int*)range + 2 * JBGPU_PAIR_CONVEX_PLANE,
                     list, keys, c.shapes.p, c.verts.p, c.props.p, c.tr.p, c.candC.p);
  tBegin(c, T_GJK);
  hipLaunchKernelGGL(kGjk, dim3(gridFor(n, kGjkBlock)), dim3(kGjkBlock), 0, s,
                     (const int*)range + 2 * JBGPU_PAIR_CONVEX_CONVEX, list, keys, c.shapes.p, c.verts.p, c.props.p,
                     c.tr.p, c.candC.p, c.epaState.p, c.epaQueue.p, c.npStat.p);
  tEnd(c, T_GJK);
  tBegin(c, T_EPA);
  c.epaRetry.ensure(n);
  hipLaunchKernelGGL(kEpaSmall, dim3(kEpaSmallGrid), dim3(64), 0, s, (const unsigned long long*)c.npStat.p,
                     (const int*)c.epaQueue.p, keys, c.shapes.p, c.verts.p, c.props.p, c.tr.p, c.epaState.p, c.candC.p,
                     c.npStat.p + 1, c.epaRetry.p, c.npStat.p + 3, c.epaSmallFaces);
  c.epaRetryHuge.ensure(n);
  c.epaHugeWork.ensure((size_t)kEpaHugeGrid * kEpaHugeSlotWords);
  c.epaHugeAux.ensure((size_t)kEpaHugeGrid * kEpaHugeAuxInts);
  hipLaunchKernelGGL(kEpa, dim3(kEpaGrid), dim3(64), 0, s, (const unsigned long long*)(c.npStat.p + 3),
                     (const int*)c.epaRetry.p, keys, c.shapes.p, c.verts.p, c.props.p, c.tr.p, c.epaState.p,
                     c.candC.p, c.npStat.p + 1, c.epaRetryHuge.p, c.npStat.p + 4, c.epaLargeFaces);
  hipLaunchKernelGGL(kEpaHuge, dim3(kEpaHugeGrid), dim3(64), 0, s, (const unsigned long long*)(c.npStat.p + 4),
                     (const int*)c.epaRetryHuge.p, keys, c.shapes.p, c.verts.p, c.props.p, c.tr.p, c.epaState.p,
                     c.candC.p, c.npStat.p + 1, c.epaHugeWork.p, c.epaHugeAux.p);
  tEnd(c, T_EPA);
  tBegin(c, T_MANIFOLD);
  hipLaunchKernelGGL(kManifold, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)nullptr, keys, c.ptype.p, c.act.p, c.props.p, c.tr.p,
                     c.candC.p, man);
  tEnd(c, T_MANIFOLD);
  JB_CHECK(hipMemcpyAsync(c.npStatHost, c.npStat.p, sizeof(c.npStatHost), hipMemcpyDeviceToHost, s));
  c.npStatPending = true;
}

// the narrowphase of a graph-captured step: launches shaped by the pair capacity, the pair count on the device
void npGraphReserve(Ctx& c) {
  const size_t n = (size_t)std::max<int64_t>(c.pairCap, 1);
  c.typeKey.ensure(n);
  c.typeKeyAlt.ensure(n);
  c.perm.ensure(n);
  c.permAlt.ensure(n);
  c.candC.ensure(8 * n);
  c.epaState.ensure(12 * n);
  c.epaQueue.ensure(n);
  c.epaRetry.ensure(n);
  c.epaRetryHuge.ensure(n);
  c.epaHugeWork.ensure((size_t)kEpaHugeGrid * kEpaHugeSlotWords);
  c.epaHugeAux.ensure((size_t)kEpaHugeGrid * kEpaHugeAuxInts);
  c.counter.ensure(8);
  c.binRange.ensure(8);
  c.npStat.ensure(5);
}

void npDispatchGraph(Ctx& c) {
  hipStream_t s = c.stream;
  const int64_t n = c.pairCap;
  const int* dn = c.gc.p + kGcPairs;
  const uint64_t* keys = c.keys[c.cur].p;
  float* man = c.man[c.cur].p;
  hipLaunchKernelGGL(kBinKeys, dim3(gridFor(n)), dim3(256), 0, s, n, dn, c.ptype.p, keys, c.act.p,
                     (const float*)c.shapes.p, (const float*)c.props.p, c.typeKey.p, c.perm.p);
  Db<int> kk(c.typeKey.p, c.typeKeyAlt.p), vv(c.perm.p, c.permAlt.p);
  primSortPairsDev(c, kk, vv, dn, (int)n, 0, 2 + kBinClassBits);
  int* starts = (int*)(c.counter.p + 4);  // -1 from the step's first fill
  hipLaunchKernelGGL(kBinStarts, dim3(gridFor(n)), dim3(256), 0, s, n, dn, (const int*)kk.Current(), starts);
  int* range = c.binRange.p;
  hipLaunchKernelGGL(kBinRanges, dim3(1), dim3(64), 0, s, n, dn, (const int*)starts, range);
  const int* list = vv.Current();
  hipLaunchKernelGGL(kSphereSphere, dim3(gridFor(n)), dim3(256), 0, s, (const int*)range + 2 * JBGPU_PAIR_SPHERE_SPHERE,
                     list, keys, c.shapes.p, c.props.p, c.tr.p, c.candC.p);
  hipLaunchKernelGGL(kConvexPlane, dim3(gridFor(n)), dim3(256), 0, s, (const int*)range + 2 * JBGPU_PAIR_CONVEX_PLANE,
                     list, keys, c.shapes.p, c.verts.p, c.props.p, c.tr.p, c.candC.p);
  hipLaunchKernelGGL(kGjk, dim3(gridFor(n, kGjkBlock)), dim3(kGjkBlock), 0, s,
                     (const int*)range + 2 * JBGPU_PAIR_CONVEX_CONVEX, list, keys, c.shapes.p, c.verts.p, c.props.p,
                     c.tr.p, c.candC.p, c.epaState.p, c.epaQueue.p, c.npStat.p);
  hipLaunchKernelGGL(kEpaSmall, dim3(kEpaSmallGrid), dim3(64), 0, s, (const unsigned long long*)c.npStat.p,
                     (const int*)c.epaQueue.p, keys, c.shapes.p, c.verts.p, c.props.p, c.tr.p, c.epaState.p, c.candC.p,
                     c.npStat.p + 1, c.epaRetry.p, c.npStat.p + 3, c.epaSmallFaces);
  hipLaunchKernelGGL(kEpa, dim3(kEpaGrid), dim3(64), 0, s, (const unsigned long long*)(c.npStat.p + 3),
                     (const int*)c.epaRetry.p, keys, c.shapes.p, c.verts.p, c.props.p, c.tr.p, c.epaState.p,
                     c.candC.p, c.npStat.p + 1, c.epaRetryHuge.p, c.npStat.p + 4, c.epaLargeFaces);
  hipLaunchKernelGGL(kEpaHuge, dim3(kEpaHugeGrid), dim3(64), 0, s, (const unsigned long long*)(c.npStat.p + 4),
                     (const int*)c.epaRetryHuge.p, keys, c.shapes.p, c.verts.p, c.props.p, c.tr.p, c.epaState.p,
                     c.candC.p, c.npStat.p + 1, c.epaHugeWork.p, c.epaHugeAux.p);
  hipLaunchKernelGGL(kManifold, dim3(gridFor(n)), dim3(256), 0, s, n, dn, keys, c.ptype.p, c.act.p, c.props.p, c.tr.p,
                     c.candC.p, man);
}

void npStatsCopy(Ctx& c) {
  JB_CHECK(hipMemcpyAsync(c.npStatHost, c.npStat.p, sizeof(c.npStatHost), hipMemcpyDeviceToHost, c.stream));
  c.npStatPending = true;
}

void npCollectStats(Ctx& c) {  // after the step's final synchronisation
  if (!c.npStatPending) return;
  c.npStatPending = false;
  c.st.epaCalls = (int64_t)c.npStatHost[0];
  c.st.epaFails = (int64_t)c.npStatHost[1];
  c.st.epaRetries = (int64_t)c.npStatHost[3];
  c.st.epaHuge = (int64_t)c.npStatHost[4];
  c.epaOverflow = c.npStatHost[2] != 0;  // this pass only: the next step starts clean
}


void rayTest(Ctx& c, int n, const float* dFrom, const float* dTo, int16_t group, int16_t mask, const int* dWorld,
             int* dObj, float* dFrac, float* dNormal, float* dPoint) {
  if (n <= 0) return;
  rayBuildTree(c);
  RayTree& t = c.rt;
  hipStream_t s = c.stream;
  t.count.ensure((size_t)n + 1);  // count[n] = 0 closes the scan: offset[n] is the total
  t.offset.ensure((size_t)n + 1);
  int* err = (int*)(c.counter.p + 6);
  JB_CHECK(hipMemsetAsync(err, 0, sizeof(int), s));
  hipLaunchKernelGGL(kRayCandidates<false>, dim3(gridFor(n)), dim3(256), 0, s, n, dFrom, dTo, (const float*)nullptr, 3,
                     dWorld, (int)group,
                     (int)mask, t.nn, (const float4*)t.nodeRec.p, (const int4*)t.leafExact.p, (const int*)t.hugeRange.p,
                     (const int*)t.hugeList.p, (const float*)t.aabb.p, (const int*)c.filt.p,
                     (const int*)c.world.p, c.maxWorld, t.count.p, (const int64_t*)nullptr, (uint64_t*)nullptr, err);
  JB_CHECK(hipMemsetAsync(t.count.p + n, 0, sizeof(int), s));
  primScanWide(c, t.count.p, t.offset.p, n + 1);
  int64_t total = 0;
  int errh = 0;
  JB_CHECK(hipMemcpyAsync(&total, t.offset.p + n, sizeof(total), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(&errh, err, sizeof(int), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  if (errh & 2) throw std::invalid_argument("ray world id out of range (no such world among the proxies)");
  if (errh & 1) throw std::runtime_error("ray traversal stack overflow");
  if (total > 0) {
    t.keys.ensure(total);
    t.keysAlt.ensure(total);
    t.rec.ensure((size_t)kRayRec * total);
    hipLaunchKernelGGL(kRayCandidates<true>, dim3(gridFor(n)), dim3(256), 0, s, n, dFrom, dTo, (const float*)nullptr, 3,
                       dWorld, (int)group,
                       (int)mask, t.nn, (const float4*)t.nodeRec.p, (const int4*)t.leafExact.p,
                       (const int*)t.hugeRange.p, (const int*)t.hugeList.p, (const float*)t.aabb.p,
                       (const int*)c.filt.p, (const int*)c.world.p, c.maxWorld, t.count.p, (const int64_t*)t.offset.p,
                       t.keys.p, err);
    // world order inside every ray's segment: sort by (ray, object)
    int bits = 32;
    while (bits < 64 && (1ll << (bits - 32)) < (int64_t)n) bits++;
    if (total >= (int64_t)1 << 31) throw std::length_error("too many ray candidates in one call");
    Db<uint64_t> kk(t.keys.p, t.keysAlt.p);
    primSortKeys(c, kk, (int)total, 0, bits);
    hipLaunchKernelGGL(kRayCast, dim3(gridFor(total, kGjkBlock)), dim3(kGjkBlock), 0, s, total,
                       (const uint64_t*)kk.Current(), dFrom, dTo, (const float*)c.shapes.p, (const float*)c.props.p,
                       (const float*)c.tr.p, (const float*)c.verts.p, t.rec.p);
    hipLaunchKernelGGL(kRayFold, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)t.count.p,
                       (const int64_t*)t.offset.p, (const uint64_t*)kk.Current(), (const float*)t.rec.p, dFrom, dTo,
                       (const float*)t.aabb.p, dObj, dFrac, dNormal, dPoint);
  } else {
    hipLaunchKernelGGL(kRayFold, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)t.count.p,
                       (const int64_t*)t.offset.p, (const uint64_t*)nullptr, (const float*)nullptr, dFrom, dTo,
                       (const float*)t.aabb.p, dObj, dFrac, dNormal, dPoint);
  }
}

void sweepTest(Ctx& c, int n, const float* castShape, const float* dFrom, const float* dTo, int16_t group,
               int16_t mask, const int* dWorld, int* dObj, float* dFrac, float* dNormal, float* dPoint) {
  if (n <= 0) return;
  rayBuildTree(c);
  RayTree& t = c.rt;
  hipStream_t s = c.stream;
  std::memcpy(t.sweepShapeHost, castShape, sizeof(t.sweepShapeHost));
  t.sweepShape.ensure(JBGPU_SHAPE_WORDS);
  JB_CHECK(hipMemcpyAsync(t.sweepShape.p, t.sweepShapeHost, sizeof(t.sweepShapeHost), hipMemcpyHostToDevice, s));
  t.sweep.ensure(12 * (size_t)n);
  hipLaunchKernelGGL(kSweepPrep, dim3(gridFor(n)), dim3(256), 0, s, n, (const float*)t.sweepShape.p, dFrom, dTo,
                     t.sweep.p);
  const float* F = t.sweep.p;
  t.count.ensure((size_t)n + 1);
  t.offset.ensure((size_t)n + 1);
  int* err = (int*)(c.counter.p + 6);
  JB_CHECK(hipMemsetAsync(err, 0, sizeof(int), s));
  hipLaunchKernelGGL(kRayCandidates<false>, dim3(gridFor(n)), dim3(256), 0, s, n, F, F + 3, F + 6, 12, dWorld,
                     (int)group, (int)mask, t.nn, (const float4*)t.nodeRec.p, (const int4*)t.leafExact.p,
                     (const int*)t.hugeRange.p, (const int*)t.hugeList.p, (const float*)t.aabb.p,
                     (const int*)c.filt.p, (const int*)c.world.p, c.maxWorld, t.count.p, (const int64_t*)nullptr,
                     (uint64_t*)nullptr, err);
  JB_CHECK(hipMemsetAsync(t.count.p + n, 0, sizeof(int), s));
  primScanWide(c, t.count.p, t.offset.p, n + 1);
  int64_t total = 0;
  int errh = 0;
  JB_CHECK(hipMemcpyAsync(&total, t.offset.p + n, sizeof(total), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(&errh, err, sizeof(int), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  if (errh & 2) throw std::invalid_argument("sweep world id out of range (no such world among the proxies)");
  if (errh & 1) throw std::runtime_error("sweep traversal stack overflow");
  if (total >= (int64_t)1 << 31) throw std::length_error("too many sweep candidates in one call");
  const uint64_t* keys = nullptr;
  if (total > 0) {
    t.keys.ensure(total);
    t.keysAlt.ensure(total);
    t.rec.ensure((size_t)kSweepRec * total);
    hipLaunchKernelGGL(kRayCandidates<true>, dim3(gridFor(n)), dim3(256), 0, s, n, F, F + 3, F + 6, 12, dWorld,
                       (int)group, (int)mask, t.nn, (const float4*)t.nodeRec.p, (const int4*)t.leafExact.p,
                       (const int*)t.hugeRange.p, (const int*)t.hugeList.p, (const float*)t.aabb.p,
                       (const int*)c.filt.p, (const int*)c.world.p, c.maxWorld, t.count.p,
                       (const int64_t*)t.offset.p, t.keys.p, err);
    int bits = 32;
    while (bits < 64 && (1ll << (bits - 32)) < (int64_t)n) bits++;
    Db<uint64_t> kk(t.keys.p, t.keysAlt.p);
    primSortKeys(c, kk, (int)total, 0, bits);
    keys = kk.Current();
    hipLaunchKernelGGL(kSweepCast, dim3(gridFor(total, kGjkBlock)), dim3(kGjkBlock), 0, s, total, keys,
                       (const float*)t.sweepShape.p, dFrom, dTo, (const float*)c.shapes.p, (const float*)c.props.p,
                       (const float*)c.tr.p, (const float*)c.verts.p, t.rec.p);
  }
  hipLaunchKernelGGL(kSweepFold, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)t.count.p,
                     (const int64_t*)t.offset.p, keys, (const float*)t.rec.p, dObj, dFrac, dNormal, dPoint);
}

void ccdClamp(Ctx& c, float dt, const int* abortFlag, const int* dnp) {
  if (c.nccd == 0) return;
  hipLaunchKernelGGL(kCcd, dim3(1), dim3(kCcdBlock), 0, c.stream, c.nccd, (const int*)c.ccdList.p,
                     (const float*)c.ccd.p, c.nb, dt, (const float*)c.trOld.p, c.tr.p, (const float*)c.vel.p,
                     (const float*)c.props.p, (const int*)c.act.p, (const float*)c.shapes.p, (const float*)c.verts.p,
                     (const int*)c.filt.p, (const int*)c.world.p, c.npairs, dnp, (const uint64_t*)c.keys[c.cur].p,
                     (const float*)c.man[c.cur].p, abortFlag);
}

}  // namespace jbg
