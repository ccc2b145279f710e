// shapes.h — convex support mappings, world AABBs and local inertia of the
// shapes on the hot path, for device kernels and the library's host helpers.
// Semantics: collision/shapes/BoxShape.java:73-167, SphereShape.java:38-76,
// CapsuleShape.java:54-190, ConvexHullShape.java:45-149,
// ConvexInternalShape.java:85-100, PolyhedralConvexShape.java:134-201,
// StaticPlaneShape.java:44-128, linearmath/AabbUtil2.java:133-208.
#pragma once
#include "../../include/jbgpu_format.h"
#include "jbmath.h"

namespace jbg {

struct Shape {
  int type;
  float margin;
  V3 dims;
  int vfirst, vcount;
  V3 lmin, lmax;
  V3 pn;
  float pc;
};

JBD V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
JBD void st3(float* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
JBD int32_t ldi(const float* p) { return __builtin_bit_cast(int32_t, *p); }
JBD void sti(float* p, int32_t i) { *p = __builtin_bit_cast(float, i); }

JBD Shape loadShape(const float* r) {
  Shape s;
  s.type = ldi(r + 0);
  s.margin = r[1];
  s.dims = ld3(r + 2);
  s.vfirst = ldi(r + 5);
  s.vcount = ldi(r + 6);
  s.lmin = ld3(r + 7);
  s.lmax = ld3(r + 10);
  s.pn = ld3(r + 13);
  s.pc = r[16];
  return s;
}
JBD void storeShape(const Shape& s, float* r) {
  for (int i = 0; i < JBGPU_SHAPE_WORDS; i++) r[i] = 0.0f;
  sti(r + 0, s.type);
  r[1] = s.margin;
  st3(r + 2, s.dims);
  sti(r + 5, s.vfirst);
  sti(r + 6, s.vcount);
  st3(r + 7, s.lmin);
  st3(r + 10, s.lmax);
  st3(r + 13, s.pn);
  r[16] = s.pc;
}

// hull support: normalised direction, strict '>' keeps the first maximum
JBD V3 hullSupport(const float* verts, int first, int n, V3 vec) {
  V3 sup = v3(0, 0, 0);
  float maxDot = -1e30f;
  float l2 = len2(vec);
  if (l2 < 0.0001f)
    vec = v3(1, 0, 0);
  else
    vec = scl(vec, 1.0f / jsqrt(l2));
  const float* p = verts + 3 * first;
  for (int i = 0; i < n; i++) {
    V3 vtx = v3(p[3 * i] * 1.0f, p[3 * i + 1] * 1.0f, p[3 * i + 2] * 1.0f);
    float d = dot(vec, vtx);
    if (d > maxDot) {
      maxDot = d;
      sup = vtx;
    }
  }
  return sup;
}

JBD V3 supportNoMargin(const Shape& s, const float* verts, V3 vec) {
  switch (s.type) {
    case JBGPU_SHAPE_BOX:
      return v3(fsel(vec.x, s.dims.x, -s.dims.x), fsel(vec.y, s.dims.y, -s.dims.y), fsel(vec.z, s.dims.z, -s.dims.z));
    case JBGPU_SHAPE_CAPSULE: {
      V3 sup = v3(0, 0, 0);
      float maxDot = -1e30f;
      float l2 = len2(vec);
      if (l2 < 0.0001f)
        vec = v3(1, 0, 0);
      else
        vec = scl(vec, 1.0f / jsqrt(l2));
      const float radius = s.dims.x, hh = s.dims.y;
      for (int k = 0; k < 2; k++) {
        V3 pos = v3(0, k == 0 ? hh : -hh, 0);
        V3 t1 = scl(v3(vec.x * 1.0f, vec.y * 1.0f, vec.z * 1.0f), radius);
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
      return hullSupport(verts, s.vfirst, s.vcount, vec);
    default:  // sphere: a point
      return v3(0, 0, 0);
  }
}

JBD V3 supportMargin(const Shape& s, const float* verts, V3 vec) {
  if (s.type == JBGPU_SHAPE_BOX) {
    V3 he = s.dims;
    he.x += s.margin;
    he.y += s.margin;
    he.z += s.margin;
    return v3(fsel(vec.x, he.x, -he.x), fsel(vec.y, he.y, -he.y), fsel(vec.z, he.z, -he.z));
  }
  V3 sup = supportNoMargin(s, verts, vec);
  if (s.margin != 0.0f) {
    V3 vn = vec;
    if (len2(vn) < (kFltEps * kFltEps)) vn = v3(-1, -1, -1);
    vn = nor(vn);
    vn = scl(vn, s.margin);
    sup = add(sup, vn);
  }
  return sup;
}

JBD void shapeAabb(const Shape& s, const Tr& t, V3& mn, V3& mx) {
  switch (s.type) {
    case JBGPU_SHAPE_BOX: {
      V3 hem = v3(s.dims.x + s.margin, s.dims.y + s.margin, s.dims.z + s.margin);
      V3 e = v3(dot(absRow(t.b, 0), hem), dot(absRow(t.b, 1), hem), dot(absRow(t.b, 2), hem));
      mn = sub(t.o, e);
      mx = add(t.o, e);
      return;
    }
    case JBGPU_SHAPE_SPHERE: {
      V3 e = v3(s.margin, s.margin, s.margin);
      mn = sub(t.o, e);
      mx = add(t.o, e);
      return;
    }
    case JBGPU_SHAPE_CAPSULE: {
      float r = s.dims.x;
      V3 he = v3(r, r + s.dims.y, r);
      he.x += s.margin;
      he.y += s.margin;
      he.z += s.margin;
      V3 e = v3(dot(absRow(t.b, 0), he), dot(absRow(t.b, 1), he), dot(absRow(t.b, 2), he));
      mn = sub(t.o, e);
      mx = add(t.o, e);
      return;
    }
    case JBGPU_SHAPE_HULL: {
      V3 lhe = scl(sub(s.lmax, s.lmin), 0.5f);
      lhe.x += s.margin;
      lhe.y += s.margin;
      lhe.z += s.margin;
      V3 lc = scl(add(s.lmax, s.lmin), 0.5f);
      V3 c = xform(t, lc);
      V3 e = v3(dot(absRow(t.b, 0), lhe), dot(absRow(t.b, 1), lhe), dot(absRow(t.b, 2), lhe));
      mn = sub(c, e);
      mx = add(c, e);
      return;
    }
    default:  // plane
      mn = v3(-1e30f, -1e30f, -1e30f);
      mx = v3(1e30f, 1e30f, 1e30f);
      return;
  }
}

// TransformUtil.integrateTransform (linearmath/TransformUtil.java:64-103): the origin advanced by lv dt, the
// orientation by the exponential map of av dt (angle limited to a quarter turn per step), renormalised
JBD Tr integrateTr(const Tr& t, V3 lv, V3 av, float dt) {
  const float ANGULAR_MOTION_THRESHOLD = 0.5f * kHalfPi;
  Tr r;
  r.o = v3(t.o.x + dt * lv.x, t.o.y + dt * lv.y, t.o.z + dt * lv.z);
  float fAngle = len(av);
  if (fAngle * dt > ANGULAR_MOTION_THRESHOLD) fAngle = ANGULAR_MOTION_THRESHOLD / dt;
  V3 axis;
  if (fAngle < 0.001f)
    axis = scl(av, 0.5f * dt - (dt * dt * dt) * (0.020833333333f) * fAngle * fAngle);
  else
    axis = scl(av, (float)sin((double)(0.5f * fAngle * dt)) / fAngle);
  Q4 dorn = Q4{axis.x, axis.y, axis.z, (float)cos((double)(fAngle * dt * 0.5f))};
  Q4 orn0 = qFromMatrix(t.b);
  Q4 pq = qnor(qmul(dorn, orn0));
  r.b = mFromQ(pq);
  return r;
}

}  // namespace jbg
