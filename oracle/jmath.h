// Oracle (test infrastructure only): Java float semantics and the libGDX math
// operations that the JBullet hot path calls, restated in C++ for the CPU oracle.
//
// Compile with -ffp-contract=off -fno-fast-math: every float op below is one
// IEEE-754 binary32 operation, evaluated in Java's left-to-right order.
//
// libGDX (com.badlogic.gdx.math) is not vendored in the reference and its
// version is unpinned; the semantics restated here are libGDX's documented
// behaviour (Vector3.mul(Matrix3) = M*v, Matrix3.mul = this*m, Vector3.nor early
// return on len2 in {0,1}, Quaternion.nor tolerance 1e-6, setFromAxes trace /
// largest-diagonal branches).  No reference test pins this boundary: parity of
// those functions against libGDX is unpinned; they are pinned here against
// closed forms (orthonormality, rotation of known axes) in tests/.
//
// Matrix storage: m[r*3+c] is element (row r, column c), i.e. libGDX's
// Matrix3.val[M_rc].
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace jb {

static const float FLT_EPS = 1.19209290e-07f;          // BulletGlobals.java:40
static const float SIMD_INF = 3.40282347e+38f;         // Float.MAX_VALUE, BulletGlobals.java:46
static const float SIMD_2_PI = 6.283185307179586232f;   // BulletGlobals.java:43
static const float SIMD_PI = SIMD_2_PI * 0.5f;
static const float SIMD_HALF_PI = SIMD_2_PI * 0.25f;
static const float CONVEX_DISTANCE_MARGIN = 0.04f;     // BulletGlobals.java:39
static const float CONTACT_BREAKING_THRESHOLD = 0.02f; // BulletGlobals.java:63

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// Float.floatToIntBits equality (canonical NaN, -0 != +0): libGDX Vector3.equals.
inline bool feq_bits(float a, float b) {
  uint32_t ua = (a != a) ? 0x7fc00000u : fbits(a);
  uint32_t ub = (b != b) ? 0x7fc00000u : fbits(b);
  return ua == ub;
}

// Java (int) cast of a float: truncation, NaN -> 0, saturation at int range.
inline int32_t jint(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (int32_t)0x80000000;
  return (int32_t)f;
}

// Math.max / Math.min (float) with Java's NaN and signed-zero rules.
inline float jmax(float a, float b) {
  if (a != a) return a;
  if (a == 0.0f && b == 0.0f && (fbits(a) == 0x80000000u)) return b;
  return (a >= b) ? a : b;
}
inline float jmin(float a, float b) {
  if (a != a) return a;
  if (a == 0.0f && b == 0.0f && (fbits(b) == 0x80000000u)) return b;
  return (a <= b) ? a : b;
}
inline float jabs(float a) { return std::fabs(a); }
inline float jsqrt(float a) { return std::sqrt(a); }  // (float)Math.sqrt(float) == sqrtf
// ScalarUtil.fsel (ScalarUtil.java:35-37)
inline float fsel(float a, float b, float c) { return a >= 0 ? b : c; }

struct V3 {
  float x, y, z;
};
inline V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 scl(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 neg(V3 a) { return scl(a, -1.0f); }  // Vector3.scl(-1): multiply, not negate
inline V3 mulc(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Vector3.crs: (y*v.z - z*v.y, z*v.x - x*v.z, x*v.y - y*v.x)
inline V3 crs(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float len(V3 a) { return jsqrt(len2(a)); }
// Vector3.nor: early return when len2 is exactly 0 or 1, else scale by 1/sqrt(len2)
inline V3 nor(V3 a) {
  float l2 = len2(a);
  if (l2 == 0.0f || l2 == 1.0f) return a;
  return scl(a, 1.0f / jsqrt(l2));
}
inline bool veq(V3 a, V3 b) { return feq_bits(a.x, b.x) && feq_bits(a.y, b.y) && feq_bits(a.z, b.z); }
inline float coord(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline void setcoord(V3& a, int i, float f) { if (i == 0) a.x = f; else if (i == 1) a.y = f; else a.z = f; }

struct M3 {
  float m[9];  // row-major: m[r*3+c]
};
inline float at(const M3& a, int r, int c) { return a.m[r * 3 + c]; }
inline M3 m3ident() { M3 r; for (int i = 0; i < 9; i++) r.m[i] = (i % 4 == 0) ? 1.0f : 0.0f; return r; }
// Vector3.mul(Matrix3): M*v, row dot v evaluated x,y,z
inline V3 mulMV(const M3& a, V3 v) {
  return v3(v.x * a.m[0] + v.y * a.m[1] + v.z * a.m[2],
            v.x * a.m[3] + v.y * a.m[4] + v.z * a.m[5],
            v.x * a.m[6] + v.y * a.m[7] + v.z * a.m[8]);
}
// MatrixUtil.transposeTransform (MatrixUtil.java:292-314): M^T * v, mat element first
inline V3 mulMtV(const M3& a, V3 v) {
  return v3(a.m[0] * v.x + a.m[3] * v.y + a.m[6] * v.z,
            a.m[1] * v.x + a.m[4] * v.y + a.m[7] * v.z,
            a.m[2] * v.x + a.m[5] * v.y + a.m[8] * v.z);
}
// Matrix3.mul(m): this*m
inline M3 mulMM(const M3& a, const M3& b) {
  M3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r.m[i * 3 + j] = a.m[i * 3 + 0] * b.m[0 * 3 + j] + a.m[i * 3 + 1] * b.m[1 * 3 + j] + a.m[i * 3 + 2] * b.m[2 * 3 + j];
  return r;
}
inline M3 transpose(const M3& a) {
  M3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r.m[i * 3 + j] = a.m[j * 3 + i];
  return r;
}
inline M3 mabs(const M3& a) { M3 r; for (int i = 0; i < 9; i++) r.m[i] = jabs(a.m[i]); return r; }
// MatrixUtil.scale (MatrixUtil.java:227-237): column j scaled by s[j]
inline M3 mscale(const M3& a, V3 s) {
  M3 r;
  for (int i = 0; i < 3; i++) {
    r.m[i * 3 + 0] = a.m[i * 3 + 0] * s.x;
    r.m[i * 3 + 1] = a.m[i * 3 + 1] * s.y;
    r.m[i * 3 + 2] = a.m[i * 3 + 2] * s.z;
  }
  return r;
}
inline V3 row(const M3& a, int r) { return v3(a.m[r * 3], a.m[r * 3 + 1], a.m[r * 3 + 2]); }

struct Q {
  float x, y, z, w;
};
inline Q q4(float x, float y, float z, float w) { Q r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
// Quaternion.mul(other): this (x) other
inline Q qmul(Q a, Q o) {
  return q4(a.w * o.x + a.x * o.w + a.y * o.z - a.z * o.y,
            a.w * o.y + a.y * o.w + a.z * o.x - a.x * o.z,
            a.w * o.z + a.z * o.w + a.x * o.y - a.y * o.x,
            a.w * o.w - a.x * o.x - a.y * o.y - a.z * o.z);
}
// Quaternion.nor: skip when len2 == 0 or |len2 - 1| <= 1e-6 (MathUtils.isEqual)
inline Q qnor(Q q) {
  float l = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  if (l != 0.0f && !(jabs(l - 1.0f) <= 0.000001f)) {
    l = jsqrt(l);
    q.w /= l; q.x /= l; q.y /= l; q.z /= l;
  }
  return q;
}
// Quaternion.setFromMatrix(true, M) -> setFromAxes(true, M00,M01,M02, M10,M11,M12, M20,M21,M22)
inline Q qFromMatrix(const M3& a) {
  float xx = a.m[0], xy = a.m[1], xz = a.m[2];
  float yx = a.m[3], yy = a.m[4], yz = a.m[5];
  float zx = a.m[6], zy = a.m[7], zz = a.m[8];
  const float lx = 1.0f / jsqrt(xx * xx + xy * xy + xz * xz);
  const float ly = 1.0f / jsqrt(yx * yx + yy * yy + yz * yz);
  const float lz = 1.0f / jsqrt(zx * zx + zy * zy + zz * zz);
  xx *= lx; xy *= lx; xz *= lx;
  yx *= ly; yy *= ly; yz *= ly;
  zx *= lz; zy *= lz; zz *= lz;
  const float t = xx + yy + zz;
  Q q;
  if (t >= 0) {
    float s = jsqrt(t + 1);
    q.w = 0.5f * s;
    s = 0.5f / s;
    q.x = (zy - yz) * s;
    q.y = (xz - zx) * s;
    q.z = (yx - xy) * s;
  } else if ((xx > yy) && (xx > zz)) {
    float s = (float)std::sqrt(1.0 + (double)xx - (double)yy - (double)zz);
    q.x = s * 0.5f;
    s = 0.5f / s;
    q.y = (yx + xy) * s;
    q.z = (xz + zx) * s;
    q.w = (zy - yz) * s;
  } else if (yy > zz) {
    float s = (float)std::sqrt(1.0 + (double)yy - (double)xx - (double)zz);
    q.y = s * 0.5f;
    s = 0.5f / s;
    q.x = (yx + xy) * s;
    q.z = (zy + yz) * s;
    q.w = (xz - zx) * s;
  } else {
    float s = (float)std::sqrt(1.0 + (double)zz - (double)xx - (double)yy);
    q.z = s * 0.5f;
    s = 0.5f / s;
    q.x = (xz + zx) * s;
    q.y = (zy + yz) * s;
    q.w = (yx - xy) * s;
  }
  return q;
}
// MatrixUtil.setRotation (MatrixUtil.java:316-335)
inline M3 m3FromQ(Q q) {
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
// QuaternionUtil.setRotation (QuaternionUtil.java:41-46): Math.sin/cos in double
inline Q qFromAxisAngle(V3 axis, float angle) {
  float d = len(axis);
  float s = (float)std::sin((double)(angle * 0.5f)) / d;
  return q4(axis.x * s, axis.y * s, axis.z * s, (float)std::cos((double)(angle * 0.5f)));
}

struct Tr {
  M3 b;
  V3 o;
};
// Transform.transform(v): basis*v + origin
inline V3 xform(const Tr& t, V3 v) { return add(mulMV(t.b, v), t.o); }
// Transform.invXform (Transform.java:133-140): basis^T * (v - origin), via an explicit transpose
inline V3 invXform(const Tr& t, V3 v) { return mulMV(transpose(t.b), sub(v, t.o)); }
// Transform.inverse(): basis^T, origin = basis^T * (-origin)
inline Tr inverse(const Tr& t) {
  Tr r;
  r.b = transpose(t.b);
  r.o = mulMV(r.b, scl(t.o, -1.0f));
  return r;
}
// Transform.mul(tr1, tr2)
inline Tr trmul(const Tr& a, const Tr& b) {
  Tr r;
  V3 v = xform(a, b.o);
  r.b = mulMM(a.b, b.b);
  r.o = v;
  return r;
}

}  // namespace jb
