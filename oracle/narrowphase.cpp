// Oracle (test infrastructure only): collision algorithms and the persistent
// manifold of the hot path, restated from
//   collision/dispatch/SphereSphereCollisionAlgorithm.java
//   collision/dispatch/ConvexPlaneCollisionAlgorithm.java
//   collision/dispatch/ConvexConvexAlgorithm.java, ManifoldResult.java
//   collision/narrowphase/GjkPairDetector.java, VoronoiSimplexSolver.java,
//   collision/narrowphase/GjkEpaSolver.java, GjkEpaPenetrationDepthSolver.java,
//   collision/narrowphase/PersistentManifold.java, ManifoldPoint.java
#include <cstdio>
#include <vector>

#include "oracle.h"

namespace jb {

// ============================ VoronoiSimplexSolver =========================
struct SubResult {  // VoronoiSimplexSolver.SubSimplexClosestResult (:664-696)
  V3 closest;
  bool uA, uB, uC, uD;
  float bc[4];
  bool degenerate;
  void resetUsed() { uA = uB = uC = uD = false; }
  void reset() {
    degenerate = false;
    setBC(0, 0, 0, 0);
    resetUsed();
  }
  bool isValid() const { return (bc[0] >= 0.0f) && (bc[1] >= 0.0f) && (bc[2] >= 0.0f) && (bc[3] >= 0.0f); }
  void setBC(float a, float b, float c, float d) { bc[0] = a; bc[1] = b; bc[2] = c; bc[3] = d; }
};

struct Voronoi {
  int numVertices;
  V3 W[5], P[5], Q[5];
  V3 cachedP1, cachedP2, cachedV, lastW;
  bool cachedValidClosest;
  SubResult cachedBC;
  bool needsUpdate;

  void reset() {  // :574-580
    cachedValidClosest = false;
    numVertices = 0;
    needsUpdate = true;
    lastW = v3(1e30f, 1e30f, 1e30f);
    cachedBC.reset();
  }
  void removeVertex(int index) {  // :82-88
    numVertices--;
    W[index] = W[numVertices];
    P[index] = P[numVertices];
    Q[index] = Q[numVertices];
  }
  void reduceVertices(const SubResult& u) {  // :90-98
    if ((numVertices >= 4) && (!u.uD)) removeVertex(3);
    if ((numVertices >= 3) && (!u.uC)) removeVertex(2);
    if ((numVertices >= 2) && (!u.uB)) removeVertex(1);
    if ((numVertices >= 1) && (!u.uA)) removeVertex(0);
  }
  void addVertex(V3 w, V3 p, V3 q) {  // :582-591
    lastW = w;
    needsUpdate = true;
    W[numVertices] = w;
    P[numVertices] = p;
    Q[numVertices] = q;
    numVertices++;
  }
  bool inSimplex(V3 w) const {  // :611-630
    bool found = false;
    for (int i = 0; i < numVertices; i++)
      if (veq(W[i], w)) found = true;
    if (veq(w, lastW)) return true;
    return found;
  }

  // closestPtPointTriangle (:267-389) with p = origin passed explicitly
  static bool closestPtPointTriangle(V3 p, V3 a, V3 b, V3 c, SubResult& r) {
    r.resetUsed();
    V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
      r.closest = a; r.uA = true; r.setBC(1, 0, 0, 0);
      return true;
    }
    V3 bp = sub(p, b);
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
      r.closest = b; r.uB = true; r.setBC(0, 1, 0, 0);
      return true;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
      float v = d1 / (d1 - d3);
      r.closest = v3(v * ab.x + a.x, v * ab.y + a.y, v * ab.z + a.z);
      r.uA = r.uB = true;
      r.setBC(1.0f - v, v, 0, 0);
      return true;
    }
    V3 cp = sub(p, c);
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
      r.closest = c; r.uC = true; r.setBC(0, 0, 1, 0);
      return true;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
      float w = d2 / (d2 - d6);
      r.closest = v3(w * ac.x + a.x, w * ac.y + a.y, w * ac.z + a.z);
      r.uA = r.uC = true;
      r.setBC(1.0f - w, 0, w, 0);
      return true;
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
      float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      V3 t = sub(c, b);
      r.closest = v3(w * t.x + b.x, w * t.y + b.y, w * t.z + b.z);
      r.uB = r.uC = true;
      r.setBC(0, 1.0f - w, w, 0);
      return true;
    }
    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom;
    float w = vc * denom;
    V3 t1 = scl(ab, v), t2 = scl(ac, w);
    r.closest = v3(a.x + t1.x + t2.x, a.y + t1.y + t2.y, a.z + t1.z + t2.z);
    r.uA = r.uB = r.uC = true;
    r.setBC(1.0f - v - w, v, w, 0);
    return true;
  }

  static int pointOutsideOfPlane(V3 p, V3 a, V3 b, V3 c, V3 d) {  // :393-425
    V3 normal = crs(sub(b, a), sub(c, a));
    float signp = dot(sub(p, a), normal);
    float signd = dot(sub(d, a), normal);
    if (signd * signd < ((1e-4f) * (1e-4f))) return -1;
    return (signp * signd < 0.0f) ? 1 : 0;
  }

  bool closestPtPointTetrahedron(V3 p, V3 a, V3 b, V3 c, V3 d, SubResult& fr) {  // :428-561
    SubResult tr;
    tr.reset();
    fr.closest = p;
    fr.resetUsed();
    fr.uA = fr.uB = fr.uC = fr.uD = true;
    int oABC = pointOutsideOfPlane(p, a, b, c, d);
    int oACD = pointOutsideOfPlane(p, a, c, d, b);
    int oADB = pointOutsideOfPlane(p, a, d, b, c);
    int oBDC = pointOutsideOfPlane(p, b, d, c, a);
    if (oABC < 0 || oACD < 0 || oADB < 0 || oBDC < 0) {
      fr.degenerate = true;
      return false;
    }
    if (oABC == 0 && oACD == 0 && oADB == 0 && oBDC == 0) return false;
    float best = 3.40282347e+38f;  // Float.MAX_VALUE
    if (oABC != 0) {
      closestPtPointTriangle(p, a, b, c, tr);
      V3 q = tr.closest;
      V3 t = sub(q, p);
      float sq = dot(t, t);
      if (sq < best) {
        best = sq;
        fr.closest = q;
        fr.resetUsed();
        fr.uA = tr.uA; fr.uB = tr.uB; fr.uC = tr.uC;
        fr.setBC(tr.bc[0], tr.bc[1], tr.bc[2], 0);
      }
    }
    if (oACD != 0) {
      closestPtPointTriangle(p, a, c, d, tr);
      V3 q = tr.closest;
      V3 t = sub(q, p);
      float sq = dot(t, t);
      if (sq < best) {
        best = sq;
        fr.closest = q;
        fr.resetUsed();
        fr.uA = tr.uA; fr.uC = tr.uB; fr.uD = tr.uC;
        fr.setBC(tr.bc[0], 0, tr.bc[1], tr.bc[2]);
      }
    }
    if (oADB != 0) {
      closestPtPointTriangle(p, a, d, b, tr);
      V3 q = tr.closest;
      V3 t = sub(q, p);
      float sq = dot(t, t);
      if (sq < best) {
        best = sq;
        fr.closest = q;
        fr.resetUsed();
        fr.uA = tr.uA; fr.uB = tr.uC; fr.uD = tr.uB;
        fr.setBC(tr.bc[0], tr.bc[2], 0, tr.bc[1]);
      }
    }
    if (oBDC != 0) {
      closestPtPointTriangle(p, b, d, c, tr);
      V3 q = tr.closest;
      V3 t = sub(q, p);
      float sq = dot(t, t);
      if (sq < best) {
        best = sq;
        fr.closest = q;
        fr.resetUsed();
        fr.uB = tr.uA; fr.uC = tr.uC; fr.uD = tr.uB;
        fr.setBC(0, tr.bc[0], tr.bc[2], tr.bc[1]);
      }
    }
    return true;
  }

  bool updateClosestVectorAndPoints() {  // :100-264
    if (needsUpdate) {
      cachedBC.reset();
      needsUpdate = false;
      switch (numVertices) {
        case 0:
          cachedValidClosest = false;
          break;
        case 1:
          cachedP1 = P[0];
          cachedP2 = Q[0];
          cachedV = sub(cachedP1, cachedP2);
          cachedBC.reset();
          cachedBC.setBC(1, 0, 0, 0);
          cachedValidClosest = cachedBC.isValid();
          break;
        case 2: {
          V3 from = W[0], to = W[1];
          V3 diff = sub(v3(0, 0, 0), from);
          V3 v = sub(to, from);
          float t = dot(v, diff);
          if (t > 0) {
            float dotVV = dot(v, v);
            if (t < dotVV) {
              t /= dotVV;
              diff = sub(diff, scl(v, t));
              cachedBC.uA = true;
              cachedBC.uB = true;
            } else {
              t = 1;
              diff = sub(diff, v);
              cachedBC.uB = true;
            }
          } else {
            t = 0;
            cachedBC.uA = true;
          }
          cachedBC.setBC(1.0f - t, t, 0, 0);
          cachedP1 = add(P[0], scl(sub(P[1], P[0]), t));
          cachedP2 = add(Q[0], scl(sub(Q[1], Q[0]), t));
          cachedV = sub(cachedP1, cachedP2);
          reduceVertices(cachedBC);
          cachedValidClosest = cachedBC.isValid();
          break;
        }
        case 3: {
          closestPtPointTriangle(v3(0, 0, 0), W[0], W[1], W[2], cachedBC);
          V3 t1 = scl(P[0], cachedBC.bc[0]), t2 = scl(P[1], cachedBC.bc[1]), t3 = scl(P[2], cachedBC.bc[2]);
          cachedP1 = v3(t1.x + t2.x + t3.x, t1.y + t2.y + t3.y, t1.z + t2.z + t3.z);
          t1 = scl(Q[0], cachedBC.bc[0]); t2 = scl(Q[1], cachedBC.bc[1]); t3 = scl(Q[2], cachedBC.bc[2]);
          cachedP2 = v3(t1.x + t2.x + t3.x, t1.y + t2.y + t3.y, t1.z + t2.z + t3.z);
          cachedV = sub(cachedP1, cachedP2);
          reduceVertices(cachedBC);
          cachedValidClosest = cachedBC.isValid();
          break;
        }
        case 4: {
          bool sep = closestPtPointTetrahedron(v3(0, 0, 0), W[0], W[1], W[2], W[3], cachedBC);
          if (sep) {
            V3 t1 = scl(P[0], cachedBC.bc[0]), t2 = scl(P[1], cachedBC.bc[1]), t3 = scl(P[2], cachedBC.bc[2]),
               t4 = scl(P[3], cachedBC.bc[3]);
            cachedP1 = v3(t1.x + t2.x + t3.x + t4.x, t1.y + t2.y + t3.y + t4.y, t1.z + t2.z + t3.z + t4.z);
            t1 = scl(Q[0], cachedBC.bc[0]); t2 = scl(Q[1], cachedBC.bc[1]); t3 = scl(Q[2], cachedBC.bc[2]);
            t4 = scl(Q[3], cachedBC.bc[3]);
            cachedP2 = v3(t1.x + t2.x + t3.x + t4.x, t1.y + t2.y + t3.y + t4.y, t1.z + t2.z + t3.z + t4.z);
            cachedV = sub(cachedP1, cachedP2);
            reduceVertices(cachedBC);
          } else {
            if (cachedBC.degenerate) {
              cachedValidClosest = false;
            } else {
              cachedValidClosest = true;
              cachedV = v3(0, 0, 0);
            }
            break;
          }
          cachedValidClosest = cachedBC.isValid();
          break;
        }
        default:
          cachedValidClosest = false;
      }
    }
    return cachedValidClosest;
  }
  bool closest(V3& v) {  // :594-598
    bool s = updateClosestVectorAndPoints();
    v = cachedV;
    return s;
  }
  void computePoints(V3& p1, V3& p2) {  // :639-643
    updateClosestVectorAndPoints();
    p1 = cachedP1;
    p2 = cachedP2;
  }
};

// ============================ GjkPairDetector ==============================
ContactCandidate gjkPairDetector(const Shape& sa, const Tr& ta, const Shape& sb, const Tr& tb, const V3* hv,
                                 float maxDist2, GjkStats* st, bool penetrationSolver) {
  ContactCandidate out{};
  const float REL_ERROR2 = 1.0e-6f;  // GjkPairDetector.java:44
  float distance = 0.0f;
  V3 normalInB = v3(0, 0, 0);
  V3 pointOnA = v3(0, 0, 0), pointOnB = v3(0, 0, 0);
  Tr la = ta, lb = tb;
  V3 positionOffset = scl(add(la.o, lb.o), 0.5f);  // :84-87
  la.o = sub(la.o, positionOffset);
  lb.o = sub(lb.o, positionOffset);
  float marginA = sa.margin, marginB = sb.margin;
  int curIter = 0;
  const int gGjkMaxIter = 1000;
  V3 cachedSep = v3(0, 1, 0);  // :102
  bool isValid = false, checkSimplex = false, checkPenetration = true;
  int degenerateSimplex = 0;
  float squaredDistance = SIMD_INF;
  float delta = 0.0f;
  float margin = marginA + marginB;
  Voronoi vs;
  vs.reset();
  for (;;) {  // :132-228
    V3 sepA = mulMtV(ta.b, scl(cachedSep, -1.0f));
    V3 sepB = mulMtV(tb.b, cachedSep);
    V3 pInA = supportNoMargin(sa, hv, sepA);
    V3 qInB = supportNoMargin(sb, hv, sepB);
    V3 pWorld = xform(la, pInA);
    V3 qWorld = xform(lb, qInB);
    V3 w = sub(pWorld, qWorld);
    delta = dot(cachedSep, w);
    if ((delta > 0.0f) && (delta * delta > squaredDistance * maxDist2)) {
      checkPenetration = false;
      break;
    }
    if (vs.inSimplex(w)) {
      degenerateSimplex = 1;
      checkSimplex = true;
      break;
    }
    float f0 = squaredDistance - delta;
    float f1 = squaredDistance * REL_ERROR2;
    if (f0 <= f1) {
      if (f0 <= 0.0f) degenerateSimplex = 2;
      checkSimplex = true;
      break;
    }
    vs.addVertex(w, pWorld, qWorld);
    if (!vs.closest(cachedSep)) {
      degenerateSimplex = 3;
      checkSimplex = true;
      break;
    }
    if (len2(cachedSep) < REL_ERROR2) {
      degenerateSimplex = 6;
      checkSimplex = true;
      break;
    }
    float previousSquaredDistance = squaredDistance;
    squaredDistance = len2(cachedSep);
    if (previousSquaredDistance - squaredDistance <= FLT_EPS * previousSquaredDistance) {
      cachedSep = vs.cachedV;  // backup_closest
      checkSimplex = true;
      break;
    }
    if (curIter++ > gGjkMaxIter) break;
    bool check = (vs.numVertices != 4);
    if (!check) {
      cachedSep = vs.cachedV;
      break;
    }
  }
  if (st) st->gjkIters += curIter;
  if (checkSimplex) {  // :230-257
    vs.computePoints(pointOnA, pointOnB);
    normalInB = sub(pointOnA, pointOnB);
    float lenSqr = len2(cachedSep);
    if (lenSqr < 0.0001f) degenerateSimplex = 5;
    if (lenSqr > FLT_EPS * FLT_EPS) {
      float rlen = 1.0f / jsqrt(lenSqr);
      normalInB = scl(normalInB, rlen);
      float s = jsqrt(squaredDistance);
      pointOnA = sub(pointOnA, scl(cachedSep, marginA / s));
      pointOnB = add(pointOnB, scl(cachedSep, marginB / s));
      distance = ((1.0f / rlen) - margin);
      isValid = true;
    }
  }
  bool catchDegeneratePenetrationCase = (degenerateSimplex != 0 && ((distance + margin) < 0.01f));  // :259-260
  if (penetrationSolver && checkPenetration && (!isValid || catchDegeneratePenetrationCase)) {  // :263-303
    V3 tA, tB;
    int it = 0;
    if (st) st->epaCalls++;
    bool ok = gjkEpaPenDepth(sa, la, sb, lb, hv, tA, tB, &it);
    if (ok) {
      V3 tn = sub(tB, tA);
      float lenSqr = len2(tn);
      if (lenSqr > (FLT_EPS * FLT_EPS)) {
        tn = scl(tn, 1.0f / jsqrt(lenSqr));
        float distance2 = -len(sub(tA, tB));
        if (!isValid || (distance2 < distance)) {
          distance = distance2;
          pointOnA = tA;
          pointOnB = tB;
          normalInB = tn;
          isValid = true;
        }
      }
    } else if (st) {
      st->epaFails++;
    }
  }
  if (isValid) {  // :306-314
    out.valid = true;
    out.normalOnB = normalInB;
    out.pointOnB = add(pointOnB, positionOffset);
    out.depth = distance;
  }
  return out;
}

// ============================ GjkEpaSolver =================================
namespace epa {
static const int GJK_maxiterations = 128;
static const int GJK_hashsize = 1 << 6;
static const int GJK_hashmask = GJK_hashsize - 1;
static const float GJK_insimplex_eps = 0.0001f;
static const float GJK_sqinsimplex_eps = GJK_insimplex_eps * GJK_insimplex_eps;
static const int EPA_maxiterations = 256;
static const float EPA_inface_eps = 0.01f;
static const float EPA_accuracy = 0.001f;
static const float cstInf = SIMD_INF;
static const float cst2Pi = SIMD_2_PI;

struct Mkv {
  V3 w, r;
};
struct He {
  V3 v;
  int n;  // next index, -1 = null
};
struct Face {
  int v[3];  // Mkv indices into the vertex pool
  int f[3];  // adjacent faces
  int e[3];
  V3 n;
  float d;
  int mark;
  int prev, next;
};

struct Solver {
  const Shape* shapes[2];
  M3 wrot[2];
  V3 pos[2];
  const V3* hv;
  float margin;
  // GJK
  int table[GJK_hashsize];
  std::vector<He> he;
  Mkv simplex[5];
  V3 ray;
  int order, iterations;
  bool failed;
  // EPA
  std::vector<Mkv> mkv;  // EPA vertex pool (basemkv + supports)
  std::vector<Face> faces;
  int root, nfaces, epaIterations;
  V3 nearest[2];
  V3 normal;
  float depth;
  bool epaFailed;

  static int Hash(V3 v) {  // GjkEpaSolver.java:188-191
    int h = jint(v.x * 15461) ^ jint(v.y * 83003) ^ jint(v.z * 15473);
    return (int)((uint32_t)h * 169639u) & GJK_hashmask;
  }
  V3 LocalSupport(V3 d, int i) {  // :193-203
    V3 t = mulMtV(wrot[i], d);
    V3 o = supportMargin(*shapes[i], hv, t);
    o = mulMV(wrot[i], o);
    return add(o, pos[i]);
  }
  void Support(V3 d, Mkv& v) {  // :205-218
    v.r = d;
    V3 t1 = LocalSupport(d, 0);
    V3 t2 = LocalSupport(scl(d, -1.0f), 1);
    v.w = sub(t1, t2);
    v.w.x += margin * d.x;
    v.w.y += margin * d.y;
    v.w.z += margin * d.z;
  }
  bool FetchSupport() {  // :220-238
    int h = Hash(ray);
    int e = table[h];
    while (e >= 0) {
      if (veq(he[e].v, ray)) {
        --order;
        return false;
      }
      e = he[e].n;
    }
    He ne;
    ne.v = ray;
    ne.n = table[h];
    he.push_back(ne);
    table[h] = (int)he.size() - 1;
    Support(ray, simplex[++order]);
    return (dot(ray, simplex[order].w) > 0);
  }
  bool SolveSimplex2(V3 ao, V3 ab) {  // :240-256
    if (dot(ab, ao) >= 0) {
      V3 cabo = crs(ab, ao);
      if (len2(cabo) > GJK_sqinsimplex_eps) {
        ray = crs(cabo, ab);
      } else {
        return true;
      }
    } else {
      order = 0;
      simplex[0] = simplex[1];
      ray = ao;
    }
    return false;
  }
  bool SolveSimplex3(V3 ao, V3 ab, V3 ac) { return SolveSimplex3a(ao, ab, ac, crs(ab, ac)); }
  bool SolveSimplex3a(V3 ao, V3 ab, V3 ac, V3 cabc) {  // :267-306
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
  bool SolveSimplex4(V3 ao, V3 ab, V3 ac, V3 ad) {  // :308-345
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
  bool SearchOrigin() {  // :347-420, initial ray (1,0,0)
    iterations = 0;
    order = -1;
    failed = false;
    ray = nor(v3(1, 0, 0));
    for (int i = 0; i < GJK_hashsize; i++) table[i] = -1;
    he.clear();
    FetchSupport();
    ray = scl(simplex[0].w, -1.0f);
    for (; iterations < GJK_maxiterations; ++iterations) {
      float rl = len(ray);
      ray = scl(ray, 1.0f / (rl > 0.0f ? rl : 1.0f));
      if (FetchSupport()) {
        bool found = false;
        switch (order) {
          case 1:
            found = SolveSimplex2(scl(simplex[1].w, -1.0f), sub(simplex[0].w, simplex[1].w));
            break;
          case 2:
            found = SolveSimplex3(scl(simplex[2].w, -1.0f), sub(simplex[1].w, simplex[2].w),
                                  sub(simplex[0].w, simplex[2].w));
            break;
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
  bool EncloseOrigin() {  // :422-496
    switch (order) {
      case 0:
        break;
      case 1: {
        V3 ab = sub(simplex[1].w, simplex[0].w);
        V3 b[3] = {crs(ab, v3(1, 0, 0)), crs(ab, v3(0, 1, 0)), crs(ab, v3(0, 0, 1))};
        float m[3] = {len2(b[0]), len2(b[1]), len2(b[2])};
        Q q = qFromAxisAngle(nor(ab), cst2Pi / 3.0f);
        M3 r = m3FromQ(q);
        V3 w = b[m[0] > m[1] ? (m[0] > m[2] ? 0 : 2) : (m[1] > m[2] ? 1 : 2)];
        Support(nor(w), simplex[4]);
        w = mulMV(r, w);
        Support(nor(w), simplex[2]);
        w = mulMV(r, w);
        Support(nor(w), simplex[3]);
        w = mulMV(r, w);
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

  // ---- EPA (:533-856); faces/vertices in pools, indices instead of references
  V3 GetCoordinates(const Face& f) {  // :562-593
    V3 o = scl(f.n, -f.d);
    float a[3];
    a[0] = len(crs(sub(mkv[f.v[0]].w, o), sub(mkv[f.v[1]].w, o)));
    a[1] = len(crs(sub(mkv[f.v[1]].w, o), sub(mkv[f.v[2]].w, o)));
    a[2] = len(crs(sub(mkv[f.v[2]].w, o), sub(mkv[f.v[0]].w, o)));
    float sm = a[0] + a[1] + a[2];
    return scl(v3(a[1], a[2], a[0]), 1.0f / (sm > 0.0f ? sm : 1.0f));
  }
  int FindBest() {  // :595-608
    int bf = -1;
    if (root >= 0) {
      int cf = root;
      float bd = cstInf;
      do {
        if (faces[cf].d < bd) {
          bd = faces[cf].d;
          bf = cf;
        }
      } while (-1 != (cf = faces[cf].next));
    }
    return bf;
  }
  bool Set(Face& f, int a, int b, int c) {  // :610-634
    V3 nrm = crs(sub(mkv[b].w, mkv[a].w), sub(mkv[c].w, mkv[a].w));
    float l = len(nrm);
    V3 t1 = crs(mkv[a].w, mkv[b].w), t2 = crs(mkv[b].w, mkv[c].w), t3 = crs(mkv[c].w, mkv[a].w);
    bool valid = (dot(t1, nrm) >= -EPA_inface_eps) && (dot(t2, nrm) >= -EPA_inface_eps) &&
                 (dot(t3, nrm) >= -EPA_inface_eps);
    f.v[0] = a; f.v[1] = b; f.v[2] = c;
    f.mark = 0;
    f.n = scl(nrm, 1.0f / (l > 0.0f ? l : cstInf));
    f.d = jmax(0.0f, -dot(f.n, mkv[a].w));
    return valid;
  }
  int NewFace(int a, int b, int c) {  // :636-651
    Face nf{};
    nf.f[0] = nf.f[1] = nf.f[2] = -1;
    faces.push_back(nf);
    int pf = (int)faces.size() - 1;
    if (Set(faces[pf], a, b, c)) {
      if (root >= 0) faces[root].prev = pf;
      faces[pf].prev = -1;
      faces[pf].next = root;
      root = pf;
      ++nfaces;
    } else {
      faces[pf].prev = faces[pf].next = -1;
    }
    return pf;
  }
  void Detach(int fi) {  // :653-670
    Face& face = faces[fi];
    if (face.prev >= 0 || face.next >= 0) {
      --nfaces;
      if (fi == root) {
        root = face.next;
        faces[root].prev = -1;
      } else {
        if (face.next < 0) {
          faces[face.prev].next = -1;
        } else {
          faces[face.prev].next = face.next;
          faces[face.next].prev = face.prev;
        }
      }
      face.prev = face.next = -1;
    }
  }
  void Link(int f0, int e0, int f1, int e1) {  // :672-677
    faces[f0].f[e0] = f1;
    faces[f1].e[e1] = e0;
    faces[f1].f[e1] = f0;
    faces[f0].e[e0] = e1;
  }
  int BuildHorizon(int markid, int w, int f, int e, int* cf, int* ff) {  // :686-706
    static const int mod3[5] = {0, 1, 2, 0, 1};
    int ne = 0;
    if (faces[f].mark != markid) {
      int e1 = mod3[e + 1];
      if ((dot(faces[f].n, mkv[w].w) + faces[f].d) > 0) {
        int nf = NewFace(faces[f].v[e1], faces[f].v[e], w);
        Link(nf, 0, f, e);
        if (*cf >= 0)
          Link(*cf, 1, nf, 2);
        else
          *ff = nf;
        *cf = nf;
        ne = 1;
      } else {
        int e2 = mod3[e + 2];
        Detach(f);
        faces[f].mark = markid;
        int fa = faces[f].f[e1], ea = faces[f].e[e1];
        ne += BuildHorizon(markid, w, fa, ea, cf, ff);
        int fb = faces[f].f[e2], eb = faces[f].e[e2];
        ne += BuildHorizon(markid, w, fb, eb, cf, ff);
      }
    }
    return ne;
  }
  float EvaluatePD(float accuracy) {  // :712-856
    static const int tet_f[4][3] = {{2, 1, 0}, {3, 0, 1}, {3, 1, 2}, {3, 2, 0}};
    static const int tet_e[6][4] = {{0, 0, 2, 1}, {0, 1, 1, 1}, {0, 2, 3, 1}, {1, 0, 3, 2}, {2, 0, 1, 2}, {3, 0, 2, 2}};
    static const int hex_f[6][3] = {{2, 0, 4}, {4, 1, 2}, {1, 4, 0}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
    static const int hex_e[9][4] = {{0, 0, 4, 0}, {0, 1, 2, 1}, {0, 2, 1, 2}, {1, 1, 5, 2}, {1, 0, 2, 0},
                                    {2, 2, 3, 2}, {3, 1, 5, 0}, {3, 0, 4, 2}, {5, 1, 4, 1}};
    int bestface = -1;
    int markid = 1;
    depth = -cstInf;
    normal = v3(0, 0, 0);
    root = -1;
    nfaces = 0;
    epaIterations = 0;
    epaFailed = false;
    mkv.clear();
    faces.clear();
    if (EncloseOrigin()) {
      const int(*pf)[3] = nullptr;
      const int(*pe)[4] = nullptr;
      int nf = 0, ne = 0;
      if (order == 3) { pf = tet_f; nf = 4; pe = tet_e; ne = 6; }
      else if (order == 4) { pf = hex_f; nf = 6; pe = hex_e; ne = 9; }
      for (int i = 0; i <= order; ++i) mkv.push_back(simplex[i]);
      int basefaces[6];
      for (int i = 0; i < nf; ++i) basefaces[i] = NewFace(pf[i][0], pf[i][1], pf[i][2]);
      for (int i = 0; i < ne; ++i) Link(basefaces[pe[i][0]], pe[i][1], basefaces[pe[i][2]], pe[i][3]);
    }
    if (0 == nfaces) return depth;
    for (; epaIterations < EPA_maxiterations; ++epaIterations) {
      int bf = FindBest();
      if (bf >= 0) {
        Mkv wv;
        Support(scl(faces[bf].n, -1.0f), wv);
        mkv.push_back(wv);
        int w = (int)mkv.size() - 1;
        float d = dot(faces[bf].n, mkv[w].w) + faces[bf].d;
        bestface = bf;
        if (d < -accuracy) {
          int cf = -1, ff = -1;
          int nf = 0;
          Detach(bf);
          faces[bf].mark = ++markid;
          for (int i = 0; i < 3; ++i) nf += BuildHorizon(markid, w, faces[bf].f[i], faces[bf].e[i], &cf, &ff);
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
      const Face& bfc = faces[bestface];
      V3 b = GetCoordinates(bfc);
      normal = bfc.n;
      depth = jmax(0.0f, bfc.d);
      V3 feat[2][3];
      for (int i = 0; i < 2; ++i) {
        float s = i != 0 ? -1.0f : 1.0f;
        for (int j = 0; j < 3; ++j) feat[i][j] = LocalSupport(scl(mkv[bfc.v[j]].r, s), i);
      }
      for (int i = 0; i < 2; i++) {
        V3 t1 = scl(feat[i][0], b.x), t2 = scl(feat[i][1], b.y), t3 = scl(feat[i][2], b.z);
        nearest[i] = v3(t1.x + t2.x + t3.x, t1.y + t2.y + t3.y, t1.z + t2.z + t3.z);
      }
    } else {
      epaFailed = true;
    }
    return depth;
  }
};
}  // namespace epa

// GjkEpaPenetrationDepthSolver.calcPenDepth -> GjkEpaSolver.collide (radialmargin 0)
bool gjkEpaPenDepth(const Shape& sa, const Tr& ta, const Shape& sb, const Tr& tb, const V3* hv, V3& wA, V3& wB,
                    int* epaIters) {
  epa::Solver s;
  s.shapes[0] = &sa;
  s.shapes[1] = &sb;
  s.wrot[0] = ta.b;
  s.pos[0] = ta.o;
  s.wrot[1] = tb.b;
  s.pos[1] = tb.o;
  s.hv = hv;
  s.margin = 0.0f + epa::EPA_accuracy;  // GjkEpaSolver.java:880
  bool collide = s.SearchOrigin();
  if (collide) {
    float pd = s.EvaluatePD(epa::EPA_accuracy);
    if (epaIters) *epaIters = s.epaIterations;
    if (pd > 0) {
      wA = s.nearest[0];
      wB = s.nearest[1];
      return true;
    }
  }
  return false;
}

// ============================ simple algorithms ============================
// SphereSphereCollisionAlgorithm.processCollision (:73-134). touching=false means
// "no new point, refresh only".
ContactCandidate sphereSphere(const Shape& s0, const Tr& t0, const Shape& s1, const Tr& t1, bool& touching) {
  ContactCandidate c{};
  V3 diff = sub(t0.o, t1.o);
  float l = len(diff);
  float r0 = s0.dims.x, r1 = s1.dims.x;
  touching = false;
  if (l > (r0 + r1)) return c;
  touching = true;
  float dist = l - (r0 + r1);
  V3 n = v3(1, 0, 0);
  if (l > FLT_EPS) n = scl(diff, 1.0f / l);
  V3 pos1 = add(t1.o, scl(n, r1));
  c.valid = true;
  c.normalOnB = n;
  c.pointOnB = pos1;
  c.depth = dist;
  return c;
}

// ConvexPlaneCollisionAlgorithm.processCollision (:75-136), convex = manifold body0
ContactCandidate convexPlane(const Shape& convex, const V3* hv, const Tr& tc, const Shape& plane, const Tr& tp) {
  ContactCandidate c{};
  V3 pn = plane.pn;
  float pc = plane.pc;
  Tr planeInConvex = trmul(inverse(tc), tp);
  Tr convexInPlane = trmul(inverse(tp), tc);
  V3 dir = mulMV(planeInConvex.b, scl(pn, -1.0f));
  V3 vtx = supportMargin(convex, hv, dir);
  V3 vtxInPlane = xform(convexInPlane, vtx);
  float distance = (dot(pn, vtxInPlane) - pc);
  V3 vproj = sub(vtxInPlane, scl(pn, distance));
  V3 vworld = xform(tp, vproj);
  if (distance < CONTACT_BREAKING_THRESHOLD) {
    c.valid = true;
    c.normalOnB = mulMV(tp.b, pn);
    c.pointOnB = vworld;
    c.depth = distance;
  }
  return c;
}

// ============================ manifold =====================================
int manifoldCacheEntry(const Manifold& m, const MPoint& np) {  // PersistentManifold.java:214-233
  float shortest = CONTACT_BREAKING_THRESHOLD * CONTACT_BREAKING_THRESHOLD;
  int nearest = -1;
  for (int i = 0; i < m.n; i++) {
    V3 d = sub(m.p[i].localA, np.localA);
    float dd = dot(d, d);
    if (dd < shortest) {
      shortest = dd;
      nearest = i;
    }
  }
  return nearest;
}

int manifoldSortCachedPoints(const Manifold& m, const MPoint& pt) {  // :83-156
  int maxPenetrationIndex = -1;
  float maxPenetration = pt.distance;
  for (int i = 0; i < 4; i++) {
    if (m.p[i].distance < maxPenetration) {
      maxPenetrationIndex = i;
      maxPenetration = m.p[i].distance;
    }
  }
  float res0 = 0, res1 = 0, res2 = 0, res3 = 0;
  if (maxPenetrationIndex != 0) res0 = len2(crs(sub(pt.localA, m.p[1].localA), sub(m.p[3].localA, m.p[2].localA)));
  if (maxPenetrationIndex != 1) res1 = len2(crs(sub(pt.localA, m.p[0].localA), sub(m.p[3].localA, m.p[2].localA)));
  if (maxPenetrationIndex != 2) res2 = len2(crs(sub(pt.localA, m.p[0].localA), sub(m.p[3].localA, m.p[1].localA)));
  if (maxPenetrationIndex != 3) res3 = len2(crs(sub(pt.localA, m.p[0].localA), sub(m.p[2].localA, m.p[1].localA)));
  // VectorUtil.closestAxis4 -> maxAxis4 of absolute values (VectorUtil.java:60-90)
  float v[4] = {jabs(res0), jabs(res1), jabs(res2), jabs(res3)};
  int maxIndex = -1;
  float maxVal = -1e30f;
  for (int i = 0; i < 4; i++)
    if (v[i] > maxVal) {
      maxIndex = i;
      maxVal = v[i];
    }
  return maxIndex;
}

int manifoldAddPoint(Manifold& m, const MPoint& np) {  // :235-257
  int idx = m.n;
  if (idx == 4)
    idx = manifoldSortCachedPoints(m, np);
  else
    m.n++;
  m.p[idx] = np;
  return idx;
}

void manifoldReplacePoint(Manifold& m, const MPoint& np, int idx) {  // :280-305
  int life = m.p[idx].lifetime;
  float ai = m.p[idx].applied, l1 = m.p[idx].lat1, l2 = m.p[idx].lat2;
  m.p[idx] = np;
  m.p[idx].applied = ai;
  m.p[idx].lat1 = l1;
  m.p[idx].lat2 = l2;
  m.p[idx].lifetime = life;
}

void manifoldRemovePoint(Manifold& m, int idx) {  // :259-278
  int last = m.n - 1;
  if (idx != last) {
    m.p[idx] = m.p[last];
    m.p[last].applied = 0;
    m.p[last].latInit = false;
    m.p[last].lat1 = 0;
    m.p[last].lat2 = 0;
    m.p[last].lifetime = 0;
  }
  m.n--;
}

void manifoldRefresh(Manifold& m, const Tr& trA, const Tr& trB) {  // :312-372
  for (int i = m.n - 1; i >= 0; i--) {
    MPoint& p = m.p[i];
    p.worldA = xform(trA, p.localA);
    p.worldB = xform(trB, p.localB);
    p.distance = dot(sub(p.worldA, p.worldB), p.normal);
    p.lifetime++;
  }
  const float thr = CONTACT_BREAKING_THRESHOLD;
  for (int i = m.n - 1; i >= 0; i--) {
    MPoint& p = m.p[i];
    if (!(p.distance <= thr)) {
      manifoldRemovePoint(m, i);
    } else {
      V3 proj = sub(p.worldA, scl(p.normal, p.distance));
      V3 pd = sub(p.worldB, proj);
      float d2 = dot(pd, pd);
      if (d2 > thr * thr) manifoldRemovePoint(m, i);
    }
  }
}

// ManifoldResult.addContactPoint (ManifoldResult.java:92-157).  rbody0 is the
// ManifoldResult's body0 (the pair's proxy0); r0/r1 its two bodies in that order.
void addContactPoint(Manifold& m, const Body& r0, const Body& r1, int rbody0, const ContactCandidate& c) {
  if (c.depth > CONTACT_BREAKING_THRESHOLD) return;
  bool isSwapped = m.body0 != rbody0;
  V3 pointA = add(scl(c.normalOnB, c.depth), c.pointOnB);
  V3 localA, localB;
  if (isSwapped) {
    localA = invXform(r1.tr, pointA);
    localB = invXform(r0.tr, c.pointOnB);
  } else {
    localA = invXform(r0.tr, pointA);
    localB = invXform(r1.tr, c.pointOnB);
  }
  MPoint np{};
  np.localA = localA;
  np.localB = localB;
  np.normal = c.normalOnB;
  np.distance = c.depth;
  np.worldA = pointA;
  np.worldB = c.pointOnB;
  int idx = manifoldCacheEntry(m, np);
  // calculateCombinedFriction / Restitution (:160-175)
  float fr = r0.friction * r1.friction;
  if (fr < -10.0f) fr = -10.0f;
  if (fr > 10.0f) fr = 10.0f;
  np.friction = fr;
  np.restitution = r0.restitution * r1.restitution;
  if (idx >= 0)
    manifoldReplacePoint(m, np, idx);
  else
    manifoldAddPoint(m, np);
}

void refreshContactPoints(Manifold& m, const Body* bodies) {  // ManifoldResult.java:177-191
  if (m.n == 0) return;
  manifoldRefresh(m, bodies[m.body0].tr, bodies[m.body1].tr);
}


// ============================ rayTest ======================================
// CollisionWorld.rayTest (CollisionWorld.java:552-591) with rayTestSingle (:260-336) for convex shapes
// (SubsimplexConvexCast) and StaticPlaneShape (concave: processAllTriangles + TriangleRaycastCallback), reported
// to a ClosestRayResultCallback (:697-729).
static int outcode(V3 p, V3 h) {  // AabbUtil2.java:40-43
  return (p.x < -h.x ? 0x01 : 0x0) | (p.x > h.x ? 0x08 : 0x0) | (p.y < -h.y ? 0x02 : 0x0) | (p.y > h.y ? 0x10 : 0x0) |
         (p.z < -h.z ? 0x4 : 0x0) | (p.z > h.z ? 0x20 : 0x0);
}

bool rayAabb(V3 from, V3 to, V3 mn, V3 mx, float& param, V3& normal) {  // AabbUtil2.java:45-104
  V3 he = scl(sub(mx, mn), 0.5f);
  V3 c = scl(add(mx, mn), 0.5f);
  V3 source = sub(from, c), target = sub(to, c);
  int so = outcode(source, he), tc = outcode(target, he);
  if ((so & tc) == 0x0) {
    float enter = 0.0f, exit = param;
    V3 r = sub(target, source);
    float normSign = 1.0f;
    V3 hn = v3(0, 0, 0);
    int bit = 1;
    for (int j = 0; j < 2; j++) {
      for (int i = 0; i != 3; ++i) {
        if ((so & bit) != 0) {
          float l = (-coord(source, i) - coord(he, i) * normSign) / coord(r, i);
          if (enter <= l) {
            enter = l;
            hn = v3(0, 0, 0);
            setcoord(hn, i, normSign);
          }
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
      normal = hn;
      return true;
    }
  }
  return false;
}

// SubsimplexConvexCast.calcTimeOfImpact (SubsimplexConvexCast.java:61-196), shape A the ray's point
// (SphereShape(0), margin 0: its support is the origin of its transform)
bool subsimplexRayCast(const Shape& sb, const V3* hv, const Tr& fromA, const Tr& toA, const Tr& fromB, const Tr& toB,
                       float& fraction, V3& normal) {
  Voronoi ss;
  ss.reset();
  V3 linVelA = sub(toA.o, fromA.o), linVelB = sub(toB.o, fromB.o);
  float lambda = 0.0f;
  Tr iA = fromA, iB = fromB;
  V3 r = sub(linVelA, linVelB);
  V3 supA = xform(fromA, v3(0, 0, 0));
  V3 supB = xform(fromB, supportMargin(sb, hv, mulMtV(fromB.b, r)));
  V3 v = sub(supA, supB);
  int maxIter = 32;
  V3 n = v3(0, 0, 0);
  float dist2 = len2(v);
  const float epsilon = 0.0001f;
  while ((dist2 > epsilon) && (maxIter--) != 0) {
    supA = xform(iA, v3(0, 0, 0));
    supB = xform(iB, supportMargin(sb, hv, mulMtV(iB.b, v)));
    V3 w = sub(supA, supB);
    float VdotW = dot(v, w);
    if (lambda > 1.0f) return false;
    if (VdotW > 0.0f) {
      float VdotR = dot(v, r);
      if (VdotR >= -(FLT_EPS * FLT_EPS)) return false;
      lambda = lambda - VdotW / VdotR;
      const float s = 1.0f - lambda;  // VectorUtil.setInterpolate3
      iA.o = v3(s * fromA.o.x + lambda * toA.o.x, s * fromA.o.y + lambda * toA.o.y, s * fromA.o.z + lambda * toA.o.z);
      iB.o = v3(s * fromB.o.x + lambda * toB.o.x, s * fromB.o.y + lambda * toB.o.y, s * fromB.o.z + lambda * toB.o.z);
      w = sub(supA, supB);
      n = v;
    }
    ss.addVertex(w, supA, supB);
    if (ss.closest(v))
      dist2 = len2(v);
    else
      dist2 = 0.0f;
  }
  fraction = lambda;
  if (len2(n) >= FLT_EPS * FLT_EPS)
    normal = nor(n);
  else
    normal = v3(0, 0, 0);
  if (dot(normal, r) >= -0.0f) return false;  // CastResult.allowedPenetration = 0
  return true;
}

static void planeSpace1(V3 n, V3& p, V3& q) {  // TransformUtil.java:45-61
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

V3 interp3(V3 a, V3 b, float t) {  // VectorUtil.setInterpolate3
  float s = 1.0f - t;
  return v3(s * a.x + t * b.x, s * a.y + t * b.y, s * a.z + t * b.z);
}

// TriangleRaycastCallback.processTriangle (TriangleRaycastCallback.java:51-124); returns true and the hit
// (distance, unnormalised local normal) when it would call reportHit
bool rayTriangle(const V3 tri[3], V3 from, V3 to, float hitFraction, float& distance, V3& normal) {
  V3 v10 = sub(tri[1], tri[0]), v20 = sub(tri[2], tri[0]);
  V3 tn = crs(v10, v20);
  float dist = dot(tri[0], tn);
  float da = dot(tn, from);
  da -= dist;
  float db = dot(tn, to);
  db -= dist;
  if (da * db >= 0.0f) return false;
  float proj = da - db;
  distance = da / proj;
  if (!(distance < hitFraction)) return false;
  float tol = len2(tn);
  tol *= -0.0001f;
  V3 point = interp3(from, to, distance);
  V3 v0p = sub(tri[0], point), v1p = sub(tri[1], point);
  if (!(dot(crs(v0p, v1p), tn) >= tol)) return false;
  V3 v2p = sub(tri[2], point);
  if (!(dot(crs(v1p, v2p), tn) >= tol)) return false;
  if (!(dot(crs(v2p, v0p), tn) >= tol)) return false;
  normal = da > 0.0f ? tn : scl(tn, -1.0f);
  return true;
}

// StaticPlaneShape.processAllTriangles (StaticPlaneShape.java:60-121) for the ray's local box; the two
// triangles in order.  The reference sets halfExtents from aabbMax and then from aabbMin, so the radius is
// half the length of aabbMin (kept).
void planeTriangles(const Shape& s, V3 aabbMin, V3 aabbMax, V3 tri[2][3]) {
  V3 he = scl(aabbMin, 0.5f);
  float radius = len(he);
  V3 center = scl(add(aabbMax, aabbMin), 0.5f);
  V3 t0, t1;
  planeSpace1(s.pn, t0, t1);
  V3 pc = sub(center, scl(s.pn, dot(s.pn, center) - s.pc));
  V3 a = scl(t0, radius), b = scl(t1, radius);
  tri[0][0] = v3(pc.x + a.x + b.x, pc.y + a.y + b.y, pc.z + a.z + b.z);
  tri[0][1] = add(pc, sub(a, b));
  tri[0][2] = sub(pc, sub(a, b));
  tri[1][0] = sub(pc, sub(a, b));
  tri[1][1] = sub(pc, add(a, b));
  tri[1][2] = v3(pc.x + a.x + b.x, pc.y + a.y + b.y, pc.z + a.z + b.z);
}

RayHit rayTestClosest(const World& w, V3 from, V3 to, int16_t group, int16_t mask, int world) {
  RayHit out;
  out.object = -1;
  out.fraction = 1.0f;  // RayResultCallback.closestHitFraction
  out.normal = v3(0, 0, 0);
  out.point = v3(0, 0, 0);
  const V3* hv = w.hullv.empty() ? nullptr : w.hullv.data();
  Tr rayFrom, rayTo;
  rayFrom.b = m3ident();
  rayFrom.o = from;
  rayTo.b = m3ident();
  rayTo.o = to;
  for (size_t i = 0; i < w.bodies.size(); i++) {
    if (out.fraction == 0.0f) break;
    const Proxy& px = w.bp.proxies[i];
    if (px.dead || px.world != world) continue;  // separate worlds: separate collision object lists
    bool collides = (((int)px.group & (int)mask) & 0xFFFF) != 0;  // RayResultCallback.needsCollision
    collides = collides && (((int)group & (int)px.mask) & 0xFFFF) != 0;
    if (!collides) continue;
    const Body& b = w.bodies[i];
    const Shape& sh = w.shapes[b.shape];
    V3 mn, mx;
    shapeAabb(sh, b.tr, mn, mx);
    float lambda = out.fraction;
    V3 hn;
    if (!rayAabb(from, to, mn, mx, lambda, hn)) continue;
    if (sh.type != JBGPU_SHAPE_PLANE) {
      float f;
      V3 n;
      if (subsimplexRayCast(sh, hv, rayFrom, rayTo, b.tr, b.tr, f, n)) {
        if (len2(n) > 0.0001f && f < out.fraction) {
          n = nor(mulMV(rayFrom.b, n));
          out.fraction = f;
          out.object = (int)i;
          out.normal = n;
          out.point = interp3(from, to, f);
        }
      }
    } else {
      Tr w2o = inverse(b.tr);
      V3 fl = xform(w2o, from), tl = xform(w2o, to);
      V3 rmin = v3(jmin(fl.x, tl.x), jmin(fl.y, tl.y), jmin(fl.z, tl.z));
      V3 rmax = v3(jmax(fl.x, tl.x), jmax(fl.y, tl.y), jmax(fl.z, tl.z));
      V3 tri[2][3];
      planeTriangles(sh, rmin, rmax, tri);
      float hitFraction = out.fraction;
      for (int t = 0; t < 2; t++) {
        float d;
        V3 nl;
        if (rayTriangle(tri[t], fl, tl, hitFraction, d, nl)) {
          out.fraction = d;  // BridgeTriangleRaycastCallback.reportHit -> addSingleResult(normal in local space)
          out.object = (int)i;
          out.normal = mulMV(b.tr.b, nl);
          out.point = interp3(from, to, d);
          hitFraction = d;
        }
      }
    }
  }
  return out;
}

}  // namespace jb
