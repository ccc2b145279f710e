// jbmath.h — float arithmetic of the JBullet hot path for gfx950 (and the host
// helpers of libjbgpu).  Every operation is one IEEE binary32 op in Java's
// left-to-right order; the library is built with -ffp-contract=off,
// -fhip-fp32-correctly-rounded-divide-sqrt and without denormal flushing, so
// results are bit-identical to Java float code with the same op order.
//
// Vector/matrix/quaternion conventions follow libGDX as used by the reference
// (linearmath/Transform.java, MatrixUtil.java, TransformUtil.java):
//   Vector3.mul(Matrix3) = M*v, Matrix3.mul(m) = this*m,
//   Vector3.nor() returns early when len2 is 0 or 1,
//   Quaternion.nor() skips when |len2 - 1| <= 1e-6.
// Matrices are row-major m[r*3+c].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define JBD __host__ __device__ __forceinline__

namespace jbg {

constexpr float kFltEps = 1.19209290e-07f;         // BulletGlobals.FLT_EPSILON
constexpr float kInf = 3.40282347e+38f;            // Float.MAX_VALUE
constexpr float k2Pi = 6.283185307179586232f;      // BulletGlobals.SIMD_2_PI
constexpr float kHalfPi = 6.283185307179586232f * 0.25f;
constexpr float kContactThreshold = 0.02f;         // BulletGlobals.contactBreakingThreshold

struct V3 {
  float x, y, z;
};
struct M3 {
  float m[9];
};
struct Q4 {
  float x, y, z, w;
};
struct Tr {
  M3 b;
  V3 o;
};

JBD uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
JBD bool feqBits(float a, float b) {  // Float.floatToIntBits equality
  uint32_t ua = (a != a) ? 0x7fc00000u : fbits(a);
  uint32_t ub = (b != b) ? 0x7fc00000u : fbits(b);
  return ua == ub;
}
JBD int32_t jint(float f) {  // Java (int) cast
  if (f != f) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (int32_t)0x80000000;
  return (int32_t)f;
}
JBD float jmax(float a, float b) {
  if (a != a) return a;
  if (a == 0.0f && b == 0.0f && fbits(a) == 0x80000000u) return b;
  return (a >= b) ? a : b;
}
JBD float jmin(float a, float b) {
  if (a != a) return a;
  if (a == 0.0f && b == 0.0f && fbits(b) == 0x80000000u) return b;
  return (a <= b) ? a : b;
}
JBD float jabs(float a) { return __builtin_bit_cast(float, fbits(a) & 0x7fffffffu); }
JBD float jsqrt(float a) { return __builtin_sqrtf(a); }
JBD float fsel(float a, float b, float c) { return a >= 0 ? b : c; }

JBD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
JBD V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
JBD V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
JBD V3 scl(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
JBD V3 mulc(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
JBD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
JBD V3 crs(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
JBD float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
JBD float len(V3 a) { return jsqrt(len2(a)); }
JBD V3 nor(V3 a) {
  float l2 = len2(a);
  if (l2 == 0.0f || l2 == 1.0f) return a;
  return scl(a, 1.0f / jsqrt(l2));
}
JBD bool veq(V3 a, V3 b) { return feqBits(a.x, b.x) && feqBits(a.y, b.y) && feqBits(a.z, b.z); }
JBD float coord(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

JBD M3 mident() {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = (i % 4 == 0) ? 1.0f : 0.0f;
  return r;
}
JBD V3 mulMV(const M3& a, V3 v) {  // M*v (Vector3.mul(Matrix3))
  return v3(v.x * a.m[0] + v.y * a.m[1] + v.z * a.m[2], v.x * a.m[3] + v.y * a.m[4] + v.z * a.m[5],
            v.x * a.m[6] + v.y * a.m[7] + v.z * a.m[8]);
}
JBD V3 mulMtV(const M3& a, V3 v) {  // M^T*v (MatrixUtil.transposeTransform)
  return v3(a.m[0] * v.x + a.m[3] * v.y + a.m[6] * v.z, a.m[1] * v.x + a.m[4] * v.y + a.m[7] * v.z,
            a.m[2] * v.x + a.m[5] * v.y + a.m[8] * v.z);
}
JBD M3 mulMM(const M3& a, const M3& b) {
  M3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r.m[i * 3 + j] = a.m[i * 3 + 0] * b.m[j] + a.m[i * 3 + 1] * b.m[3 + j] + a.m[i * 3 + 2] * b.m[6 + j];
  return r;
}
JBD M3 transpose(const M3& a) {
  M3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r.m[i * 3 + j] = a.m[j * 3 + i];
  return r;
}
JBD V3 row(const M3& a, int r) { return v3(a.m[r * 3], a.m[r * 3 + 1], a.m[r * 3 + 2]); }
JBD V3 absRow(const M3& a, int r) { return v3(jabs(a.m[r * 3]), jabs(a.m[r * 3 + 1]), jabs(a.m[r * 3 + 2])); }

JBD Q4 qmul(Q4 a, Q4 o) {
  return Q4{a.w * o.x + a.x * o.w + a.y * o.z - a.z * o.y, a.w * o.y + a.y * o.w + a.z * o.x - a.x * o.z,
            a.w * o.z + a.z * o.w + a.x * o.y - a.y * o.x, a.w * o.w - a.x * o.x - a.y * o.y - a.z * o.z};
}
JBD Q4 qnor(Q4 q) {
  float l = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  if (l != 0.0f && !(jabs(l - 1.0f) <= 0.000001f)) {
    l = jsqrt(l);
    q.w /= l;
    q.x /= l;
    q.y /= l;
    q.z /= l;
  }
  return q;
}
// Quaternion.setFromMatrix(true, M): axes normalised, trace / largest-diagonal branches
JBD Q4 qFromMatrix(const M3& a) {
  float xx = a.m[0], xy = a.m[1], xz = a.m[2], yx = a.m[3], yy = a.m[4], yz = a.m[5], zx = a.m[6], zy = a.m[7],
        zz = a.m[8];
  const float lx = 1.0f / jsqrt(xx * xx + xy * xy + xz * xz);
  const float ly = 1.0f / jsqrt(yx * yx + yy * yy + yz * yz);
  const float lz = 1.0f / jsqrt(zx * zx + zy * zy + zz * zz);
  xx *= lx; xy *= lx; xz *= lx;
  yx *= ly; yy *= ly; yz *= ly;
  zx *= lz; zy *= lz; zz *= lz;
  const float t = xx + yy + zz;
  Q4 q;
  if (t >= 0) {
    float s = jsqrt(t + 1);
    q.w = 0.5f * s;
    s = 0.5f / s;
    q.x = (zy - yz) * s;
    q.y = (xz - zx) * s;
    q.z = (yx - xy) * s;
  } else if ((xx > yy) && (xx > zz)) {
    float s = (float)__builtin_sqrt(1.0 + (double)xx - (double)yy - (double)zz);
    q.x = s * 0.5f;
    s = 0.5f / s;
    q.y = (yx + xy) * s;
    q.z = (xz + zx) * s;
    q.w = (zy - yz) * s;
  } else if (yy > zz) {
    float s = (float)__builtin_sqrt(1.0 + (double)yy - (double)xx - (double)zz);
    q.y = s * 0.5f;
    s = 0.5f / s;
    q.x = (yx + xy) * s;
    q.z = (zy + yz) * s;
    q.w = (xz - zx) * s;
  } else {
    float s = (float)__builtin_sqrt(1.0 + (double)zz - (double)xx - (double)yy);
    q.z = s * 0.5f;
    s = 0.5f / s;
    q.x = (xz + zx) * s;
    q.y = (zy + yz) * s;
    q.w = (yx - xy) * s;
  }
  return q;
}
JBD M3 mFromQ(Q4 q) {  // MatrixUtil.setRotation
  float d = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  float s = 2.0f / d;
  float xs = q.x * s, ys = q.y * s, zs = q.z * s;
  float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
  float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
  float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
  M3 r;
  r.m[0] = 1.0f - (yy + zz); r.m[1] = xy - wz; r.m[2] = xz + wy;
  r.m[3] = xy + wz; r.m[4] = 1.0f - (xx + zz); r.m[5] = yz - wx;
  r.m[6] = xz - wy; r.m[7] = yz + wx; r.m[8] = 1.0f - (xx + yy);
  return r;
}

JBD V3 xform(const Tr& t, V3 v) { return add(mulMV(t.b, v), t.o); }
JBD V3 invXform(const Tr& t, V3 v) { return mulMV(transpose(t.b), sub(v, t.o)); }
JBD Tr trInverse(const Tr& t) {
  Tr r;
  r.b = transpose(t.b);
  r.o = mulMV(r.b, scl(t.o, -1.0f));
  return r;
}
JBD Tr trMul(const Tr& a, const Tr& b) {
  Tr r;
  r.o = xform(a, b.o);
  r.b = mulMM(a.b, b.b);
  return r;
}

}  // namespace jbg
