// Oracle (test infrastructure only): broadphase pair-set semantics.
//
// DbvtBroadphase keeps, per proxy, a stored AABB that setAabb fattens in place
// (bp/DbvtBroadphase.java:196-228, bp/Dbvt.java:157-169, bp/DbvtAabbMm.java:
// 117-145).  After collide() the overlapping-pair cache holds exactly
// {a<b : Intersect(aabb_a, aabb_b) && filter} (tree collide adds, cleanup at
// :127-148 removes; leaf volumes always contain the stored AABB).
// AxisSweep3 keeps pairs whose quantised intervals overlap on all axes
// (bp/AxisSweep3Internal.java:174-216); min edges are even, max edges odd.
// SimpleBroadphase tests the raw AABBs (bp/SimpleBroadphase.java:81-110).
// Pair ids are canonical (uid_lo < uid_hi), HashedOverlappingPairCache.java:292-296,
// filtered by needsBroadphaseCollision (:179-188).
//
// The set is computed here with a sort-based uniform grid (exact: every
// candidate is re-tested with the reference predicate) so that 1M proxies
// finish in seconds; pairsBrute() is the O(N^2) SimpleBroadphase loop that
// pins it.
#include <algorithm>
#include <cstdio>

#include "oracle.h"

namespace jb {

static const int STAGECOUNT = 2;               // DbvtBroadphase.java:39
static const float DBVT_BP_MARGIN = 0.05f;     // DbvtBroadphase.java:35
static const float PREDICTED_FRAMES = 2.0f;    // DbvtBroadphase.java:73

void Broadphase::init(int m, V3 mn, V3 mx) {
  mode = m;
  stageCurrent = 0;
  proxies.clear();
  wmin = mn;
  wmax = mx;
  if (mode == JBGPU_BP_SAP16) {  // AxisSweep3.java:51-55, getMask :67-69
    handleMask = 0xfffe;
    sentinel = 0xffff;
    qmask = 0xffff;
  } else {  // AxisSweep3_32.java:48-52, getMask :64-66
    handleMask = 0xfffffffe;
    sentinel = 0x7fffffff;
    qmask = 0xffffffff;
  }
  // AxisSweep3Internal.java:98-103: quantize = maxInt / aabbSize (int -> float division)
  V3 size = sub(wmax, wmin);
  float maxInt = (float)(int32_t)sentinel;
  quant = v3(maxInt / size.x, maxInt / size.y, maxInt / size.z);
}

// AxisSweep3Internal.quantize (AxisSweep3Internal.java:201-216)
void Broadphase::quantize(V3 p, int isMax, int32_t* out) const {
  V3 c = p;
  c.x = jmax(c.x, wmin.x); c.y = jmax(c.y, wmin.y); c.z = jmax(c.z, wmin.z);
  c.x = jmin(c.x, wmax.x); c.y = jmin(c.y, wmax.y); c.z = jmin(c.z, wmax.z);
  V3 v = mulc(sub(c, wmin), quant);
  out[0] = (int32_t)((((uint32_t)jint(v.x) & handleMask) | (uint32_t)isMax) & qmask);
  out[1] = (int32_t)((((uint32_t)jint(v.y) & handleMask) | (uint32_t)isMax) & qmask);
  out[2] = (int32_t)((((uint32_t)jint(v.z) & handleMask) | (uint32_t)isMax) & qmask);
}

void Broadphase::createProxy(Aabb a, int16_t group, int16_t mask) {
  Proxy p{};
  p.group = group;
  p.mask = mask;
  p.aabb = a;   // DbvtAabbMm.FromMM (DbvtBroadphase.java:175)
  p.leaf = a;   // sets[0].insert(proxy.aabb) :176
  p.stage = stageCurrent;  // :177
  if (mode == JBGPU_BP_SAP16 || mode == JBGPU_BP_SAP32) {  // addHandle :449-496
    quantize(a.mn, 0, p.q);
    quantize(a.mx, 1, p.q + 3);
  }
  proxies.push_back(p);
}

// destroyProxy (DbvtBroadphase.java:184-194; AxisSweep3Internal.removeHandle :498-540;
// SimpleBroadphase.destroyProxy): the proxy leaves every structure and
// removeOverlappingPairsContainingProxy drops its pairs; ids of the remaining proxies are unchanged
void Broadphase::destroyProxy(int idx) { proxies[idx].dead = true; }

void Broadphase::setAabb(int idx, Aabb A) {
  Proxy& p = proxies[idx];
  if (mode == JBGPU_BP_SAP16 || mode == JBGPU_BP_SAP32) {  // updateHandle :542-579
    quantize(A.mn, 0, p.q);
    quantize(A.mx, 1, p.q + 3);
    p.aabb = A;
    return;
  }
  if (mode == JBGPU_BP_SIMPLE) {  // SimpleBroadphase.setAabb
    p.aabb = A;
    return;
  }
  // DbvtBroadphase.setAabb (DbvtBroadphase.java:196-228)
  if (p.stage == STAGECOUNT) {
    p.leaf = A;  // fixed -> dynamic set: sets[0].insert(aabb)
  } else {
    if (intersect(p.leaf, A)) {
      // delta = ((min+max)*0.5 - Center(proxy.aabb)) * predictedframes
      V3 delta = scl(add(A.mn, A.mx), 0.5f);
      V3 c = scl(add(p.aabb.mn, p.aabb.mx), 0.5f);
      delta = sub(delta, c);
      delta = scl(delta, PREDICTED_FRAMES);
      // Dbvt.update(leaf, volume, velocity, margin) (Dbvt.java:157-169)
      bool contained = (p.leaf.mn.x <= A.mn.x) && (p.leaf.mn.y <= A.mn.y) && (p.leaf.mn.z <= A.mn.z) &&
                       (p.leaf.mx.x >= A.mx.x) && (p.leaf.mx.y >= A.mx.y) && (p.leaf.mx.z >= A.mx.z);
      if (!contained) {
        V3 m = v3(DBVT_BP_MARGIN, DBVT_BP_MARGIN, DBVT_BP_MARGIN);
        A.mn = sub(A.mn, m);  // Expand (DbvtAabbMm.java:117-120)
        A.mx = add(A.mx, m);
        // SignedExpand (DbvtAabbMm.java:122-138)
        if (delta.x > 0) A.mx.x += delta.x; else A.mn.x += delta.x;
        if (delta.y > 0) A.mx.y += delta.y; else A.mn.y += delta.y;
        if (delta.z > 0) A.mx.z += delta.z; else A.mn.z += delta.z;
        p.leaf = A;
      }
    } else {
      p.leaf = A;  // teleporting: sets[0].update(leaf, aabb)
    }
  }
  p.aabb = A;
  p.stage = stageCurrent;
}

// DbvtBroadphase.collide stage rotation (DbvtBroadphase.java:96-113): proxies not
// updated for STAGECOUNT collides move to the fixed set, leaf := stored aabb.
void Broadphase::collide() {
  if (mode != JBGPU_BP_DBVT) return;
  stageCurrent = (stageCurrent + 1) % STAGECOUNT;
  for (auto& p : proxies) {
    if (p.stage == stageCurrent) {
      p.leaf = p.aabb;
      p.stage = STAGECOUNT;
    }
  }
}

static inline bool filterPass(const Proxy& a, const Proxy& b) {  // HashedOverlappingPairCache.java:184-185
  // separate DynamicsWorld instances (one broadphase each) never see each other's proxies;
  // destroyed proxies are in no structure any more
  return !a.dead && !b.dead && a.world == b.world && ((a.group & b.mask) != 0) && ((b.group & a.mask) != 0);
}

bool Broadphase::overlap(int i, int j) const {
  const Proxy& a = proxies[i];
  const Proxy& b = proxies[j];
  if (mode == JBGPU_BP_SAP16 || mode == JBGPU_BP_SAP32) {
    // testOverlap on edge order == value order (min even, max odd, no ties)
    for (int ax = 0; ax < 3; ax++) {
      uint32_t amin = (uint32_t)a.q[ax], amax = (uint32_t)a.q[3 + ax];
      uint32_t bmin = (uint32_t)b.q[ax], bmax = (uint32_t)b.q[3 + ax];
      if ((int32_t)amax < (int32_t)bmin || (int32_t)bmax < (int32_t)amin) return false;
    }
    return true;
  }
  return intersect(a.aabb, b.aabb);
}

void Broadphase::pairsBrute(std::vector<uint64_t>& out) const {
  out.clear();
  int n = (int)proxies.size();
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (overlap(i, j) && filterPass(proxies[i], proxies[j])) out.push_back(pairKey(i + 1, j + 1));
}

struct CellEntry {  // proxies of different worlds never pair, so the grid is per world
  int world;
  uint64_t cell;
  int idx;
};

void Broadphase::pairs(std::vector<uint64_t>& out) const {
  out.clear();
  int n = (int)proxies.size();
  if (n < 2) return;
  bool sap = (mode == JBGPU_BP_SAP16 || mode == JBGPU_BP_SAP32);
  // candidate boxes (SAP: widened by two quanta so that every quantised overlap is a candidate)
  std::vector<Aabb> box(n);
  double pad[3] = {0, 0, 0};
  if (sap) {
    V3 size = sub(wmax, wmin);
    pad[0] = 2.5 * size.x / (double)sentinel;
    pad[1] = 2.5 * size.y / (double)sentinel;
    pad[2] = 2.5 * size.z / (double)sentinel;
  }
  double ext = 0;
  int normal = 0;
  std::vector<char> huge(n, 0);
  for (int i = 0; i < n; i++) {
    Aabb a = proxies[i].aabb;
    if (sap) {
      // quantised coordinates back to world space, widened (clamped boxes included)
      for (int ax = 0; ax < 3; ax++) {
        double q0 = (double)(uint32_t)proxies[i].q[ax], q1 = (double)(uint32_t)proxies[i].q[3 + ax];
        double s = (double)coord(quant, ax), w0 = (double)coord(wmin, ax);
        setcoord(a.mn, ax, (float)(w0 + q0 / s - pad[ax]));
        setcoord(a.mx, ax, (float)(w0 + (q1 + 1.0) / s + pad[ax]));
      }
    }
    box[i] = a;
    double e = std::max({(double)a.mx.x - a.mn.x, (double)a.mx.y - a.mn.y, (double)a.mx.z - a.mn.z});
    if (!(e < 1e6)) {
      huge[i] = 1;
    } else {
      ext += e;
      normal++;
    }
  }
  double cell = normal ? 2.0 * ext / normal : 1.0;
  if (!(cell > 0)) cell = 1.0;
  for (int i = 0; i < n; i++) {  // boxes covering very many cells are treated like huge ones
    if (huge[i]) continue;
    double cx = ((double)box[i].mx.x - box[i].mn.x) / cell + 1, cy = ((double)box[i].mx.y - box[i].mn.y) / cell + 1,
           cz = ((double)box[i].mx.z - box[i].mn.z) / cell + 1;
    if (cx * cy * cz > 4096.0) huge[i] = 1;
  }
  // huge proxies (e.g. StaticPlaneShape's +-1e30 box) are tested against every proxy of their world
  std::vector<int> byWorld(n);
  for (int i = 0; i < n; i++) byWorld[i] = i;
  std::stable_sort(byWorld.begin(), byWorld.end(),
                   [&](int a, int b) { return proxies[a].world < proxies[b].world; });
  std::vector<int> worldStart(n), worldEnd(n);  // per proxy: its world's range in byWorld
  for (int s0 = 0, e0 = 0; s0 < n; s0 = e0) {
    for (e0 = s0 + 1; e0 < n && proxies[byWorld[e0]].world == proxies[byWorld[s0]].world; e0++) {
    }
    for (int k = s0; k < e0; k++) {
      worldStart[byWorld[k]] = s0;
      worldEnd[byWorld[k]] = e0;
    }
  }
  for (int i = 0; i < n; i++) {
    if (!huge[i]) continue;
    for (int k = worldStart[i]; k < worldEnd[i]; k++) {
      const int j = byWorld[k];
      if (j == i || (huge[j] && j < i)) continue;  // huge-huge pairs once
      int a = std::min(i, j), b = std::max(i, j);
      if (overlap(a, b) && filterPass(proxies[a], proxies[b])) out.push_back(pairKey(a + 1, b + 1));
    }
  }
  // grid: every proxy is listed in each cell it covers; a pair is reported only in
  // the cell holding max(min_a, min_b), which both boxes cover when they overlap
  auto cidx = [&](float f) -> int64_t { return (int64_t)std::floor((double)f / cell) + (1 << 20); };
  std::vector<CellEntry> ent;
  ent.reserve((size_t)n * 4);
  for (int i = 0; i < n; i++) {
    if (huge[i]) continue;
    int64_t x0 = cidx(box[i].mn.x), x1 = cidx(box[i].mx.x);
    int64_t y0 = cidx(box[i].mn.y), y1 = cidx(box[i].mx.y);
    int64_t z0 = cidx(box[i].mn.z), z1 = cidx(box[i].mx.z);
    for (int64_t x = x0; x <= x1; x++)
      for (int64_t y = y0; y <= y1; y++)
        for (int64_t z = z0; z <= z1; z++)
          ent.push_back({proxies[i].world, ((uint64_t)x << 42) | ((uint64_t)y << 21) | (uint64_t)z, i});
  }
  std::sort(ent.begin(), ent.end(), [](const CellEntry& a, const CellEntry& b) {
    if (a.world != b.world) return a.world < b.world;
    return a.cell < b.cell || (a.cell == b.cell && a.idx < b.idx);
  });
  size_t s = 0;
  while (s < ent.size()) {
    size_t e = s + 1;
    while (e < ent.size() && ent[e].cell == ent[s].cell && ent[e].world == ent[s].world) e++;
    for (size_t u = s; u < e; u++) {
      for (size_t v = u + 1; v < e; v++) {
        int a = ent[u].idx, b = ent[v].idx;  // a < b by sort order
        const Aabb& A = box[a];
        const Aabb& B = box[b];
        if (!(A.mn.x <= B.mx.x && B.mn.x <= A.mx.x && A.mn.y <= B.mx.y && B.mn.y <= A.mx.y && A.mn.z <= B.mx.z &&
              B.mn.z <= A.mx.z))
          continue;
        uint64_t own = ((uint64_t)cidx(std::max(A.mn.x, B.mn.x)) << 42) | ((uint64_t)cidx(std::max(A.mn.y, B.mn.y)) << 21) |
                       (uint64_t)cidx(std::max(A.mn.z, B.mn.z));
        if (own != ent[s].cell) continue;
        if (overlap(a, b) && filterPass(proxies[a], proxies[b])) out.push_back(pairKey(a + 1, b + 1));
      }
    }
    s = e;
  }
  std::sort(out.begin(), out.end());
}

}  // namespace jb
