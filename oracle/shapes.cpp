// Oracle (test infrastructure only): convex shapes of the hot path.
// Support mappings, world AABBs and local inertia, restated from
// collision/shapes/{BoxShape,SphereShape,CapsuleShape,ConvexHullShape,
// ConvexInternalShape,PolyhedralConvexShape,StaticPlaneShape}.java and
// linearmath/AabbUtil2.java.
#include "oracle.h"

namespace jb {

Shape makeSphere(float r) {  // SphereShape.java:38-41 (margin = radius)
  Shape s{};
  s.type = JBGPU_SHAPE_SPHERE;
  s.dims = v3(r, 0, 0);
  s.margin = r * 1.0f;  // getMargin() = getRadius() = implicitShapeDimensions.x * localScaling.x
  return s;
}

Shape makeBox(V3 he) {  // BoxShape.java:46-50
  Shape s{};
  s.type = JBGPU_SHAPE_BOX;
  s.margin = CONVEX_DISTANCE_MARGIN;
  V3 m = v3(s.margin, s.margin, s.margin);
  s.dims = sub(mulc(he, v3(1, 1, 1)), m);
  return s;
}

Shape makeCapsule(float r, float height) {  // CapsuleShape.java:54-57, Y up
  Shape s{};
  s.type = JBGPU_SHAPE_CAPSULE;
  s.margin = CONVEX_DISTANCE_MARGIN;
  s.dims = v3(r, 0.5f * height, r);
  return s;
}

Shape makePlane(V3 n, float c) {  // StaticPlaneShape.java:44-47
  Shape s{};
  s.type = JBGPU_SHAPE_PLANE;
  s.margin = CONVEX_DISTANCE_MARGIN;
  s.pn = nor(n);
  s.pc = c;
  return s;
}

// ConvexHullShape.localGetSupportingVertexWithoutMargin (ConvexHullShape.java:75-102)
static V3 hullSupport(const V3* pts, int first, int n, V3 vec) {
  V3 sup = v3(0, 0, 0);
  float maxDot = -1e30f;
  float l2 = len2(vec);
  if (l2 < 0.0001f)
    vec = v3(1, 0, 0);
  else
    vec = scl(vec, 1.0f / jsqrt(l2));
  for (int i = 0; i < n; i++) {
    V3 vtx = mulc(pts[first + i], v3(1, 1, 1));
    float d = dot(vec, vtx);
    if (d > maxDot) {
      maxDot = d;
      sup = vtx;
    }
  }
  return sup;
}

Shape makeHull(const V3* pts, int first, int n) {  // ConvexHullShape.java:45-53 + recalcLocalAabb
  Shape s{};
  s.type = JBGPU_SHAPE_HULL;
  s.margin = CONVEX_DISTANCE_MARGIN;
  s.vfirst = first;
  s.vcount = n;
  // PolyhedralConvexShape.recalcLocalAabb (PolyhedralConvexShape.java:177-201), batched support over
  // the six unit axes with the strict '>' first-max rule (ConvexHullShape.java:105-131)
  static const V3 dirs[6] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  V3 sup[6];
  float w[6];
  for (int j = 0; j < 6; j++) { w[j] = -1e30f; sup[j] = v3(0, 0, 0); }
  for (int i = 0; i < n; i++) {
    V3 vtx = mulc(pts[first + i], v3(1, 1, 1));
    for (int j = 0; j < 6; j++) {
      float d = dot(dirs[j], vtx);
      if (d > w[j]) { sup[j] = vtx; w[j] = d; }
    }
  }
  for (int i = 0; i < 3; i++) {
    setcoord(s.lmax, i, coord(sup[i], i) + s.margin);
    setcoord(s.lmin, i, coord(sup[i + 3], i) - s.margin);
  }
  return s;
}

V3 supportNoMargin(const Shape& s, const V3* hullv, V3 vec) {
  switch (s.type) {
    case JBGPU_SHAPE_BOX:  // BoxShape.java:88-97
      return v3(fsel(vec.x, s.dims.x, -s.dims.x), fsel(vec.y, s.dims.y, -s.dims.y), fsel(vec.z, s.dims.z, -s.dims.z));
    case JBGPU_SHAPE_SPHERE:  // SphereShape.java:44-47
      return v3(0, 0, 0);
    case JBGPU_SHAPE_CAPSULE: {  // CapsuleShape.java:60-111
      V3 sup = v3(0, 0, 0);
      float maxDot = -1e30f;
      float l2 = len2(vec);
      if (l2 < 0.0001f)
        vec = v3(1, 0, 0);
      else
        vec = scl(vec, 1.0f / jsqrt(l2));
      float radius = s.dims.x;
      float hh = s.dims.y;
      for (int k = 0; k < 2; k++) {
        V3 pos = v3(0, k == 0 ? hh : -hh, 0);
        V3 t1 = scl(mulc(vec, v3(1, 1, 1)), radius);
        V3 t2 = scl(vec, s.margin);
        V3 vtx = sub(add(pos, t1), t2);
        float d = dot(vec, vtx);
        if (d > maxDot) {
          maxDot = d;
          sup = vtx;
        }
      }
      return sup;
    }
    case JBGPU_SHAPE_HULL:
      return hullSupport(hullv, s.vfirst, s.vcount, vec);
  }
  return v3(0, 0, 0);
}

V3 supportMargin(const Shape& s, const V3* hullv, V3 vec) {
  if (s.type == JBGPU_SHAPE_BOX) {  // BoxShape.java:73-85
    V3 he = s.dims;
    float m = s.margin;
    he.x += m; he.y += m; he.z += m;
    return v3(fsel(vec.x, he.x, -he.x), fsel(vec.y, he.y, -he.y), fsel(vec.z, he.z, -he.z));
  }
  // ConvexInternalShape.java:85-100 (ConvexHullShape.java:133-149 is identical)
  V3 sup = supportNoMargin(s, hullv, vec);
  if (s.margin != 0.0f) {
    V3 vn = vec;
    if (len2(vn) < (FLT_EPS * FLT_EPS)) vn = v3(-1, -1, -1);
    vn = nor(vn);
    vn = scl(vn, s.margin);
    sup = add(sup, vn);
  }
  return sup;
}

// AabbUtil2.transformAabb(halfExtents, margin, t) (AabbUtil2.java:133-163)
static void transformAabbHE(V3 he, float margin, const Tr& t, V3& mn, V3& mx) {
  V3 hem;
  hem.x = he.x + margin;
  hem.y = he.y + margin;
  hem.z = he.z + margin;
  M3 ab = mabs(t.b);
  V3 c = t.o;
  V3 e = v3(dot(row(ab, 0), hem), dot(row(ab, 1), hem), dot(row(ab, 2), hem));
  mn = sub(c, e);
  mx = add(c, e);
}

void shapeAabb(const Shape& s, const Tr& t, V3& mn, V3& mx) {
  switch (s.type) {
    case JBGPU_SHAPE_BOX:  // BoxShape.java:147-151
      transformAabbHE(s.dims, s.margin, t, mn, mx);
      return;
    case JBGPU_SHAPE_SPHERE: {  // SphereShape.java:57-65
      V3 e = v3(s.margin, s.margin, s.margin);
      mn = sub(t.o, e);
      mx = add(t.o, e);
      return;
    }
    case JBGPU_SHAPE_CAPSULE: {  // CapsuleShape.java:161-190
      float r = s.dims.x;
      V3 he = v3(r, r, r);
      he.y = r + s.dims.y;
      he.x += s.margin; he.y += s.margin; he.z += s.margin;
      M3 ab = mabs(t.b);
      V3 e = v3(dot(row(ab, 0), he), dot(row(ab, 1), he), dot(row(ab, 2), he));
      mn = sub(t.o, e);
      mx = add(t.o, e);
      return;
    }
    case JBGPU_SHAPE_HULL: {  // PolyhedralConvexShape.java:168-171 -> AabbUtil2.java:165-208
      V3 lhe = scl(sub(s.lmax, s.lmin), 0.5f);
      lhe.x += s.margin; lhe.y += s.margin; lhe.z += s.margin;
      V3 lc = scl(add(s.lmax, s.lmin), 0.5f);
      M3 ab = mabs(t.b);
      V3 c = xform(t, lc);
      V3 e = v3(dot(row(ab, 0), lhe), dot(row(ab, 1), lhe), dot(row(ab, 2), lhe));
      mn = sub(c, e);
      mx = add(c, e);
      return;
    }
    case JBGPU_SHAPE_PLANE:  // StaticPlaneShape.java:125-128
      mn = v3(-1e30f, -1e30f, -1e30f);
      mx = v3(1e30f, 1e30f, 1e30f);
      return;
  }
}

V3 shapeLocalInertia(const Shape& s, const V3* hullv, float mass) {
  (void)hullv;
  switch (s.type) {
    case JBGPU_SHAPE_BOX: {  // BoxShape.java:154-167
      V3 he = v3(s.dims.x + s.margin, s.dims.y + s.margin, s.dims.z + s.margin);
      float lx = 2.0f * he.x, ly = 2.0f * he.y, lz = 2.0f * he.z;
      return v3(mass / 12.0f * (ly * ly + lz * lz), mass / 12.0f * (lx * lx + lz * lz), mass / 12.0f * (lx * lx + ly * ly));
    }
    case JBGPU_SHAPE_SPHERE: {  // SphereShape.java:73-76
      float e = 0.4f * mass * s.margin * s.margin;
      return v3(e, e, e);
    }
    case JBGPU_SHAPE_CAPSULE: {  // CapsuleShape.java:126-153
      float r = s.dims.x;
      V3 he = v3(r, r, r);
      he.y = r + s.dims.y;
      float m = CONVEX_DISTANCE_MARGIN;
      float lx = 2.0f * (he.x + m), ly = 2.0f * (he.y + m), lz = 2.0f * (he.z + m);
      float x2 = lx * lx, y2 = ly * ly, z2 = lz * lz;
      float sm = mass * 0.08333333f;
      return v3(sm * (y2 + z2), sm * (x2 + z2), sm * (x2 + y2));
    }
    case JBGPU_SHAPE_HULL: {  // PolyhedralConvexShape.java:134-159
      Tr id;
      id.b = m3ident();
      id.o = v3(0, 0, 0);
      V3 mn, mx;
      shapeAabb(s, id, mn, mx);
      V3 he = scl(sub(mx, mn), 0.5f);
      float m = s.margin;
      float lx = 2.0f * (he.x + m), ly = 2.0f * (he.y + m), lz = 2.0f * (he.z + m);
      float x2 = lx * lx, y2 = ly * ly, z2 = lz * lz;
      float sm = mass * 0.08333333f;
      return v3((y2 + z2) * sm, (x2 + z2) * sm, (x2 + y2) * sm);
    }
  }
  return v3(0, 0, 0);
}

}  // namespace jb
