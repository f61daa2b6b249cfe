// d2o_math.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Small fixed-size linear algebra for the CPU oracle of Delta2's
// contact-detection path.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load the oracle; the product never does.
//
// The reference uses Eigen 3.4.  This file restates the Eigen semantics the
// path depends on, with one fixed evaluation order per operation:
//   * dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z
//   * normalized(v) returns v unchanged when |v|^2 == 0 (Eigen behaviour)
//   * M*v and M*M sum their four products left to right
//   * the 4x4 inverse is the cofactor (Laplace expansion) inverse
//   * quaternion product is the Hamilton product, toRotationMatrix is the
//     standard unit-quaternion formula (common/basic_utils.h:79-86 uses it)
// Eigen's own operation order is not pinned by anything available here; the
// product's GPU code uses these same orders so the two agree bit for bit.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace d2o {

template <class R> struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  template <class S> V3<S> cast() const { return V3<S>(S(x), S(y), S(z)); }
};

template <class R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(const V3<R>& a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, const V3<R>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(const V3<R>& a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R norm(const V3<R>& a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalized(const V3<R>& a) {
  R z = dot(a, a);
  if (z > R(0)) return a / std::sqrt(z);
  return a;
}
// Eigen's operator!= on vectors: true when any component differs (NaN differs).
template <class R> inline bool neq(const V3<R>& a, const V3<R>& b) { return !(a.x == b.x && a.y == b.y && a.z == b.z); }
template <class R> inline bool eq(const V3<R>& a, const V3<R>& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// common/basic_utils.h:10-19
template <class R> inline R clamp01(R t) {
  if (t < 0) return R(0);
  if (t > 1) return R(1);
  return t;
}
// common/basic_utils.h:22-33
template <class R> inline V3<R> lerp(const V3<R>& A, const V3<R>& B, R t) { return A * (R(1) - t) + B * t; }
template <class R> inline R lerp(R A, R B, R t) { return A * (R(1) - t) + B * t; }

// Row-major 4x4.
template <class R> struct M4 {
  R m[16];
  R& operator()(int r, int c) { return m[r * 4 + c]; }
  R operator()(int r, int c) const { return m[r * 4 + c]; }
  static M4 identity() {
    M4 a;
    for (int i = 0; i < 16; ++i) a.m[i] = R(0);
    a.m[0] = a.m[5] = a.m[10] = a.m[15] = R(1);
    return a;
  }
  template <class S> M4<S> cast() const {
    M4<S> o;
    for (int i = 0; i < 16; ++i) o.m[i] = S(m[i]);
    return o;
  }
};

template <class R> inline M4<R> mul(const M4<R>& a, const M4<R>& b) {
  M4<R> c;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      c(i, j) = ((a(i, 0) * b(0, j) + a(i, 1) * b(1, j)) + a(i, 2) * b(2, j)) + a(i, 3) * b(3, j);
  return c;
}

// M * (x, y, z, w)
template <class R> inline void mul4(const M4<R>& a, R x, R y, R z, R w, R out[4]) {
  for (int i = 0; i < 4; ++i) out[i] = ((a(i, 0) * x + a(i, 1) * y) + a(i, 2) * z) + a(i, 3) * w;
}

// common/basic_utils.h:37-42: homogeneous transform with divide by w.
template <class R> inline V3<R> transform(const V3<R>& p, const M4<R>& M) {
  R t[4];
  mul4(M, p.x, p.y, p.z, R(1), t);
  return {t[0] / t[3], t[1] / t[3], t[2] / t[3]};
}

// (1-t)*A + t*B elementwise (continuous_comparison.h:58,86).
template <class R> inline M4<R> lerpM(const M4<R>& A, const M4<R>& B, R t) {
  M4<R> o;
  R s = R(1) - t;
  for (int i = 0; i < 16; ++i) o.m[i] = s * A.m[i] + t * B.m[i];
  return o;
}

// General 4x4 inverse by cofactors (Laplace expansion over 2x2 minors of
// rows 0-1 and rows 2-3).  Used where the reference calls Matrix4::inverse()
// (continuous_comparison.h:43,45).
template <class R> inline M4<R> inverse(const M4<R>& A) {
  const R* a = A.m;
  R s0 = a[0] * a[5] - a[4] * a[1];
  R s1 = a[0] * a[6] - a[4] * a[2];
  R s2 = a[0] * a[7] - a[4] * a[3];
  R s3 = a[1] * a[6] - a[5] * a[2];
  R s4 = a[1] * a[7] - a[5] * a[3];
  R s5 = a[2] * a[7] - a[6] * a[3];
  R c5 = a[10] * a[15] - a[14] * a[11];
  R c4 = a[9] * a[15] - a[13] * a[11];
  R c3 = a[9] * a[14] - a[13] * a[10];
  R c2 = a[8] * a[15] - a[12] * a[11];
  R c1 = a[8] * a[14] - a[12] * a[10];
  R c0 = a[8] * a[13] - a[12] * a[9];
  R det = ((((s0 * c5 - s1 * c4) + s2 * c3) + s3 * c2) - s4 * c1) + s5 * c0;
  R id = R(1) / det;
  M4<R> b;
  b.m[0] = ((a[5] * c5 - a[6] * c4) + a[7] * c3) * id;
  b.m[1] = ((-a[1] * c5 + a[2] * c4) - a[3] * c3) * id;
  b.m[2] = ((a[13] * s5 - a[14] * s4) + a[15] * s3) * id;
  b.m[3] = ((-a[9] * s5 + a[10] * s4) - a[11] * s3) * id;
  b.m[4] = ((-a[4] * c5 + a[6] * c2) - a[7] * c1) * id;
  b.m[5] = ((a[0] * c5 - a[2] * c2) + a[3] * c1) * id;
  b.m[6] = ((-a[12] * s5 + a[14] * s2) - a[15] * s1) * id;
  b.m[7] = ((a[8] * s5 - a[10] * s2) + a[11] * s1) * id;
  b.m[8] = ((a[4] * c4 - a[5] * c2) + a[7] * c0) * id;
  b.m[9] = ((-a[0] * c4 + a[1] * c2) - a[3] * c0) * id;
  b.m[10] = ((a[12] * s4 - a[13] * s2) + a[15] * s0) * id;
  b.m[11] = ((-a[8] * s4 + a[9] * s2) - a[11] * s0) * id;
  b.m[12] = ((-a[4] * c3 + a[5] * c1) - a[6] * c0) * id;
  b.m[13] = ((a[0] * c3 - a[1] * c1) + a[2] * c0) * id;
  b.m[14] = ((-a[12] * s3 + a[13] * s1) - a[14] * s0) * id;
  b.m[15] = ((a[8] * s3 - a[9] * s1) + a[10] * s0) * id;
  return b;
}

template <class R> struct M3 {
  R m[9];
  R& operator()(int r, int c) { return m[r * 3 + c]; }
  R operator()(int r, int c) const { return m[r * 3 + c]; }
};
template <class R> inline M3<R> mul(const M3<R>& a, const M3<R>& b) {
  M3<R> c;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c(i, j) = (a(i, 0) * b(0, j) + a(i, 1) * b(1, j)) + a(i, 2) * b(2, j);
  return c;
}
template <class R> inline M3<R> transpose(const M3<R>& a) {
  M3<R> t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t(i, j) = a(j, i);
  return t;
}
template <class R> inline V3<R> mul(const M3<R>& a, const V3<R>& v) {
  return {(a(0, 0) * v.x + a(0, 1) * v.y) + a(0, 2) * v.z, (a(1, 0) * v.x + a(1, 1) * v.y) + a(1, 2) * v.z,
          (a(2, 0) * v.x + a(2, 1) * v.y) + a(2, 2) * v.z};
}

template <class R> struct Quat {
  R w, x, y, z;
  static Quat identity() { return {R(1), R(0), R(0), R(0)}; }
};
// Hamilton product.
template <class R> inline Quat<R> qmul(const Quat<R>& a, const Quat<R>& b) {
  return {((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z, ((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y,
          ((a.w * b.y + a.y * b.w) + a.z * b.x) - a.x * b.z, ((a.w * b.z + a.z * b.w) + a.x * b.y) - a.y * b.x};
}
template <class R> inline M3<R> toRotation(const Quat<R>& q) {
  R tx = R(2) * q.x, ty = R(2) * q.y, tz = R(2) * q.z;
  R twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
  R txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
  R tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
  M3<R> r;
  r(0, 0) = R(1) - (tyy + tzz);
  r(0, 1) = txy - twz;
  r(0, 2) = txz + twy;
  r(1, 0) = txy + twz;
  r(1, 1) = R(1) - (txx + tzz);
  r(1, 2) = tyz - twx;
  r(2, 0) = txz - twy;
  r(2, 1) = tyz + twx;
  r(2, 2) = R(1) - (txx + tyy);
  return r;
}
// common/basic_utils.h:79-86
template <class R> inline M4<R> transformationMatrix(const Quat<R>& q, const V3<R>& t) {
  M3<R> r = toRotation(q);
  M4<R> o = M4<R>::identity();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o(i, j) = r(i, j);
  o(0, 3) = t.x;
  o(1, 3) = t.y;
  o(2, 3) = t.z;
  return o;
}
// common/basic_utils.h:70-77: AngleAxis(x,X) * AngleAxis(y,Y) * AngleAxis(z,Z)
template <class R> inline Quat<R> angleAxis(R ang, int axis) {
  R h = ang / R(2);
  Quat<R> q{std::cos(h), R(0), R(0), R(0)};
  R s = std::sin(h);
  if (axis == 0) q.x = s;
  if (axis == 1) q.y = s;
  if (axis == 2) q.z = s;
  return q;
}
template <class R> inline Quat<R> eulerToQuat(const V3<R>& e) {
  return qmul(qmul(angleAxis(e.x, 0), angleAxis(e.y, 1)), angleAxis(e.z, 2));
}
template <class R> inline M4<R> transformationMatrixEuler(const V3<R>& e, const V3<R>& t) {
  return transformationMatrix(eulerToQuat(e), t);
}

// common/basic_utils.cpp:36-44: quaternion exponential of a 3-vector.
inline Quat<double> qexp(const V3<double>& d) {
  if (norm(d) == 0.0) return Quat<double>::identity();
  double n = norm(d);
  V3<double> v = (d / n) * std::sin(n);
  return {std::cos(n), v.x, v.y, v.z};
}

// PCG32, O'Neill's published minimal generator (reference behaviour:
// common/pcg_basic.cpp:42-109, seeded in common/cli_options.h:46).
struct Pcg32 {
  uint64_t state = 0, inc = 0;
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
  }
  void seed(uint64_t initstate, uint64_t initseq) {
    state = 0U;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  uint32_t bounded(uint32_t bound) {
    uint32_t threshold = (uint32_t)(-bound) % bound;
    for (;;) {
      uint32_t r = next();
      if (r >= threshold) return r % bound;
    }
  }
};

}  // namespace d2o
