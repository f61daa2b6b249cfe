// d2_geom.hpp -- device geometry of the contact-detection path (gfx950).
//
// Straight-line, branch-light forms of Delta2's primitives.  Arithmetic is
// written in a fixed evaluation order (sums left to right, no contraction:
// the library is built with -ffp-contract=off) so that the FP64 kernels return
// the same bits as a sequential CPU evaluation of the reference algorithm.
// Reference citations are relative to the reference's src/.
//
// The includer defines D2_HD: the kernels use __host__ __device__
// __forceinline__; the CPU cross-check build (tests/cpu_geom_check.cpp) uses
// plain inline.
#pragma once
#include <cmath>
#include <cstdint>

#ifndef D2_HD
#error "define D2_HD before including d2_geom.hpp"
#endif

namespace d2 {

template <class R> struct v3 {
  R x, y, z;
};
template <class R> D2_HD v3<R> mk(R x, R y, R z) { return v3<R>{x, y, z}; }
template <class R> D2_HD v3<R> add(v3<R> a, v3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> D2_HD v3<R> sub(v3<R> a, v3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> D2_HD v3<R> scl(v3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> D2_HD v3<R> sclL(R s, v3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> D2_HD v3<R> dvs(v3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> D2_HD R dot(v3<R> a, v3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> D2_HD v3<R> cross(v3<R> a, v3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> D2_HD R vlen(v3<R> a) { return sqrt(dot(a, a)); }
// Eigen normalized(): a zero vector is returned unchanged.
template <class R> D2_HD v3<R> unit(v3<R> a) {
  R z = dot(a, a);
  return z > R(0) ? dvs(a, (R)sqrt(z)) : a;
}
// vlen(d) < search without the square root in all but borderline cases: the
// squared comparison decides whenever |d|^2 is more than 1e-12 (relative)
// away from search^2, far beyond the rounding of either side.
D2_HD bool near_mid(const v3<double>& d, double search) {
  const double s2 = dot(d, d), q = search * search;
  if (q > 1e-280 && q < 1e280) {
    if (s2 < q * (1.0 - 1e-12)) return true;
    if (s2 > q * (1.0 + 1e-12)) return false;
  }
  return sqrt(s2) < search;
}
template <class R> D2_HD bool same(v3<R> a, v3<R> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
template <class R> D2_HD R clamp01(R t) { return t < R(0) ? R(0) : (t > R(1) ? R(1) : t); }
template <class R> D2_HD v3<R> lerp(v3<R> a, v3<R> b, R t) { return add(scl(a, R(1) - t), scl(b, t)); }
template <class R> D2_HD R lerp(R a, R b, R t) { return a * (R(1) - t) + b * t; }

// Row-major 4x4 helpers over plain arrays (m[r*4+c]).
template <class R> D2_HD void mul4v(const R* m, R x, R y, R z, R w, R* o) {
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = ((m[4 * i] * x + m[4 * i + 1] * y) + m[4 * i + 2] * z) + m[4 * i + 3] * w;
}
// basic_utils.h:37-42: homogeneous transform, divide by w.  A rigid pose
// gives w = 1 exactly, and x / 1 is x itself (IEEE division is correctly
// rounded), so the division is skipped there: the same bits, three FP64
// divisions fewer per point.
template <class R> D2_HD v3<R> hdiv(const R* h) {
  if (h[3] == R(1)) return {h[0], h[1], h[2]};
  return {h[0] / h[3], h[1] / h[3], h[2] / h[3]};
}
template <class R> D2_HD v3<R> xform(const R* m, v3<R> p) {
  R h[4];
  mul4v(m, p.x, p.y, p.z, R(1), h);
  return hdiv(h);
}
template <class R> D2_HD void mul44(const R* a, const R* b, R* c) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      c[4 * i + j] = ((a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j]) + a[4 * i + 2] * b[8 + j]) + a[4 * i + 3] * b[12 + j];
}
// Cofactor inverse over the 2x2 minors of rows (0,1) and (2,3).
template <class R> D2_HD void inv44(const R* a, R* b) {
  R s0 = a[0] * a[5] - a[4] * a[1], s1 = a[0] * a[6] - a[4] * a[2], s2 = a[0] * a[7] - a[4] * a[3];
  R s3 = a[1] * a[6] - a[5] * a[2], s4 = a[1] * a[7] - a[5] * a[3], s5 = a[2] * a[7] - a[6] * a[3];
  R c5 = a[10] * a[15] - a[14] * a[11], c4 = a[9] * a[15] - a[13] * a[11], c3 = a[9] * a[14] - a[13] * a[10];
  R c2 = a[8] * a[15] - a[12] * a[11], c1 = a[8] * a[14] - a[12] * a[10], c0 = a[8] * a[13] - a[12] * a[9];
  R det = ((((s0 * c5 - s1 * c4) + s2 * c3) + s3 * c2) - s4 * c1) + s5 * c0;
  R id = R(1) / det;
  b[0] = ((a[5] * c5 - a[6] * c4) + a[7] * c3) * id;
  b[1] = ((-a[1] * c5 + a[2] * c4) - a[3] * c3) * id;
  b[2] = ((a[13] * s5 - a[14] * s4) + a[15] * s3) * id;
  b[3] = ((-a[9] * s5 + a[10] * s4) - a[11] * s3) * id;
  b[4] = ((-a[4] * c5 + a[6] * c2) - a[7] * c1) * id;
  b[5] = ((a[0] * c5 - a[2] * c2) + a[3] * c1) * id;
  b[6] = ((-a[12] * s5 + a[14] * s2) - a[15] * s1) * id;
  b[7] = ((a[8] * s5 - a[10] * s2) + a[11] * s1) * id;
  b[8] = ((a[4] * c4 - a[5] * c2) + a[7] * c0) * id;
  b[9] = ((-a[0] * c4 + a[1] * c2) - a[3] * c0) * id;
  b[10] = ((a[12] * s4 - a[13] * s2) + a[15] * s0) * id;
  b[11] = ((-a[8] * s4 + a[9] * s2) - a[11] * s0) * id;
  b[12] = ((-a[4] * c3 + a[5] * c1) - a[6] * c0) * id;
  b[13] = ((a[0] * c3 - a[1] * c1) + a[2] * c0) * id;
  b[14] = ((-a[12] * s3 + a[13] * s1) - a[14] * s0) * id;
  b[15] = ((a[8] * s3 - a[9] * s1) + a[10] * s0) * id;
}

// A leaf triangle's reach beyond the search distance: its bucket eps plus the
// largest distance a float-barycentric "inside" test can accept a point
// outside it, 64 eps_f kappa L (kappa = d00 d11 / (d00 d11 - d01^2), L the
// longest edge; 4x the 16 eps_f kappa L rounding bound).
D2_HD double leaf_slack(const double* a, const double* b, const double* c, double eps) {
  double e0[3], e1[3], e2[3];
  for (int d = 0; d < 3; ++d) e0[d] = b[d] - a[d], e1[d] = c[d] - a[d], e2[d] = c[d] - b[d];
  const double d00 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
  const double d11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
  const double d01 = e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2];
  const double d22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
  const double den = d00 * d11 - d01 * d01;
  const double kappa = den > 0.0 ? d00 * d11 / den : 1e30;
  const double L2 = d00 > d11 ? (d00 > d22 ? d00 : d22) : (d11 > d22 ? d11 : d22);
  return eps + 64.0 * 5.9604645e-08 * kappa * sqrt(L2);
}
// Whether no leaf triangle under node a (pose Ta) comes within reach of one
// under node b (pose Tb), so that the node pair's subtree holds no contact:
// its discrete leaf test accepts a triangle pair when the computed distance
// is below search + eps_i + eps_j (comparison_dist.h:233), and the computed
// distance undercuts the true one by at most the float-barycentric
// acceptance bound, both folded into each node's slack E.  The lower bound on
// the true distance is the separating-axis gap of the two oriented leaf
// boxes over the 15 axes of two OBBs (3 + 3 face normals, 9 edge
// directions, each gap divided by the axis length), in FP64 with a 1e-9
// relative and absolute margin.  Pruning such a node pair leaves every
// contact of the traversal unchanged: the reference reaches no contact
// below it.
D2_HD bool nodes_apart(const double* Ta, const double* Tb, const float* ba, const float* bb,
                         double search) {
  const double thr = ((search + (double)ba[6]) + (double)bb[6]) * (1.0 + 1e-9) + 1e-9;
  double ca[3], ea[3], cb[3], eb[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ca[k] = 0.5 * ((double)ba[k] + (double)ba[3 + k]), ea[k] = 0.5 * ((double)ba[3 + k] - (double)ba[k]);
    cb[k] = 0.5 * ((double)bb[k] + (double)bb[3 + k]), eb[k] = 0.5 * ((double)bb[3 + k] - (double)bb[k]);
  }
  // centres in the world, then R = Ra^T Rb and t = Ra^T (Cb - Ca)
  double d[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double wa = ((Ta[4 * r] * ca[0] + Ta[4 * r + 1] * ca[1]) + Ta[4 * r + 2] * ca[2]) + Ta[4 * r + 3];
    const double wb = ((Tb[4 * r] * cb[0] + Tb[4 * r + 1] * cb[1]) + Tb[4 * r + 2] * cb[2]) + Tb[4 * r + 3];
    d[r] = wb - wa;
  }
  double R[3][3], A[3][3], t[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    t[i] = (Ta[i] * d[0] + Ta[4 + i] * d[1]) + Ta[8 + i] * d[2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R[i][j] = (Ta[i] * Tb[j] + Ta[4 + i] * Tb[4 + j]) + Ta[8 + i] * Tb[8 + j];
      A[i][j] = fabs(R[i][j]);
    }
  }
  double gap = -1.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)  // a's face normals
    gap = fmax(gap, fabs(t[i]) - (ea[i] + ((eb[0] * A[i][0] + eb[1] * A[i][1]) + eb[2] * A[i][2])));
#pragma unroll
  for (int j = 0; j < 3; ++j) {  // b's face normals
    const double tj = (t[0] * R[0][j] + t[1] * R[1][j]) + t[2] * R[2][j];
    gap = fmax(gap, fabs(tj) - (((ea[0] * A[0][j] + ea[1] * A[1][j]) + ea[2] * A[2][j]) + eb[j]));
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {  // a_i x b_j, length sqrt(1 - R_ij^2)
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      const double len2 = 1.0 - R[i][j] * R[i][j];
      if (!(len2 > 1e-12)) continue;  // (nearly) parallel edges: the face axes cover them
      const double ra = ea[i1] * A[i2][j] + ea[i2] * A[i1][j];
      const double rb = eb[j1] * A[i][j2] + eb[j2] * A[i][j1];
      const double s = fabs(t[i2] * R[i1][j] - t[i1] * R[i2][j]) - (ra + rb);
      gap = fmax(gap, s / sqrt(len2));
    }
  }
  return gap > thr;
}
// nodes_apart over the six face axes only (each box's three face normals):
// fewer live registers, for the FP32 form's mask kernel, which tests every
// node pair on arrival.  The face axes find nearly all of the separations
// the edge axes add to the spheres' (on a 4096-body funnel pile 98% for
// funnel pairs, 92% for body pairs).  b's offset along its own face normal
// j (Rb's column j) is taken as that column times the centre difference.
D2_HD bool boxes_apart_faces(const double* Ta, const double* Tb, const float* ba, const float* bb,
                             double search) {
  const double thr = ((search + (double)ba[6]) + (double)bb[6]) * (1.0 + 1e-9) + 1e-9;
  double d[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double wa = ((Ta[4 * r] * (0.5 * ((double)ba[0] + (double)ba[3])) +
                        Ta[4 * r + 1] * (0.5 * ((double)ba[1] + (double)ba[4]))) +
                       Ta[4 * r + 2] * (0.5 * ((double)ba[2] + (double)ba[5]))) +
                      Ta[4 * r + 3];
    const double wb = ((Tb[4 * r] * (0.5 * ((double)bb[0] + (double)bb[3])) +
                        Tb[4 * r + 1] * (0.5 * ((double)bb[1] + (double)bb[4]))) +
                       Tb[4 * r + 2] * (0.5 * ((double)bb[2] + (double)bb[5]))) +
                      Tb[4 * r + 3];
    d[r] = wb - wa;
  }
  double A[3][3], ta[3], tb[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ta[i] = (Ta[i] * d[0] + Ta[4 + i] * d[1]) + Ta[8 + i] * d[2];
    tb[i] = (Tb[i] * d[0] + Tb[4 + i] * d[1]) + Tb[8 + i] * d[2];
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = fabs((Ta[i] * Tb[j] + Ta[4 + i] * Tb[4 + j]) + Ta[8 + i] * Tb[8 + j]);
  }
  double gap = -1.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double eai = 0.5 * ((double)ba[3 + i] - (double)ba[i]), ebi = 0.5 * ((double)bb[3 + i] - (double)bb[i]);
    double rb = 0.0, ra = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      rb += 0.5 * ((double)bb[3 + j] - (double)bb[j]) * A[i][j];
      ra += 0.5 * ((double)ba[3 + j] - (double)ba[j]) * A[j][i];
    }
    gap = fmax(gap, fmax(fabs(ta[i]) - (eai + rb), fabs(tb[i]) - (ra + ebi)));
  }
  return gap > thr;
}
// The same question answered with the nodes' bounding spheres (centre
// b[8..10], radius b[7] around the same leaf vertices): |Ca - Cb| - ra - rb
// against the same threshold, in FP64.  Cheaper than the boxes (a few dozen
// operations and registers); either test proving a pair apart is enough.
D2_HD bool spheres_apart(const double* Ta, const double* Tb, const float* ba, const float* bb, double search) {
  const double thr = ((search + (double)ba[6]) + (double)bb[6]) * (1.0 + 1e-9) + 1e-9;
  double s2 = 0.0;
  for (int r = 0; r < 3; ++r) {
    const double wa = ((Ta[4 * r] * (double)ba[8] + Ta[4 * r + 1] * (double)ba[9]) + Ta[4 * r + 2] * (double)ba[10]) + Ta[4 * r + 3];
    const double wb = ((Tb[4 * r] * (double)bb[8] + Tb[4 * r + 1] * (double)bb[9]) + Tb[4 * r + 2] * (double)bb[10]) + Tb[4 * r + 3];
    s2 += (wb - wa) * (wb - wa);
  }
  return sqrt(s2) - ((double)ba[7] + (double)bb[7]) > thr;
}
// spheres_apart in FP32 for the FP32 form's mask kernel (it runs at five
// wavefronts per SIMD; the FP64 poses would spill there).  Each world
// coordinate w = R c + t is off by at most 8 u (|R c|_1 + |t|) after the
// rounding of the pose, the product and the sums (u = 2^-24), so the centre
// distance by at most 16 u m + 5 u |d| with m the sum of |w| over both
// centres; the test demands 32 u (m + |d|) beyond the FP64 threshold.
D2_HD bool spheres_apart_f32(const double* Ta, const double* Tb, const float* ba, const float* bb, double search) {
  const double thr = ((search + (double)ba[6]) + (double)bb[6]) * (1.0 + 1e-9) + 1e-9;
  float s2 = 0.0f, mag = 0.0f;
  for (int r = 0; r < 3; ++r) {
    const float wa = (((float)Ta[4 * r] * ba[8] + (float)Ta[4 * r + 1] * ba[9]) + (float)Ta[4 * r + 2] * ba[10]) +
                     (float)Ta[4 * r + 3];
    const float wb = (((float)Tb[4 * r] * bb[8] + (float)Tb[4 * r + 1] * bb[9]) + (float)Tb[4 * r + 2] * bb[10]) +
                     (float)Tb[4 * r + 3];
    s2 += (wb - wa) * (wb - wa);
    mag += fabsf(wa) + fabsf(wb) + fabsf((float)Ta[4 * r + 3]) + fabsf((float)Tb[4 * r + 3]);
  }
  const double dist = (double)sqrtf(s2);
  return dist - ((double)ba[7] + (double)bb[7]) > thr + 32.0 * 5.9604645e-08 * ((double)mag + dist);
}
// ---- rigid-body state (model/state.cpp) -----------------------------------
struct quat {
  double w, x, y, z;
};
D2_HD quat qmul(quat a, quat b) {
  return {((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z, ((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y,
          ((a.w * b.y + a.y * b.w) + a.z * b.x) - a.x * b.z, ((a.w * b.z + a.z * b.w) + a.x * b.y) - a.y * b.x};
}
D2_HD void qrot(quat q, double* r) {
  double tx = 2.0 * q.x, ty = 2.0 * q.y, tz = 2.0 * q.z;
  double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
  double txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
  double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
  r[0] = 1.0 - (tyy + tzz);
  r[1] = txy - twz;
  r[2] = txz + twy;
  r[3] = txy + twz;
  r[4] = 1.0 - (txx + tzz);
  r[5] = tyz - twx;
  r[6] = txz - twy;
  r[7] = tyz + twx;
  r[8] = 1.0 - (txx + tyy);
}
D2_HD void mul33(const double* a, const double* b, double* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c[3 * i + j] = (a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j]) + a[3 * i + 2] * b[6 + j];
}
D2_HD v3<double> mul33v(const double* a, v3<double> v) {
  return {(a[0] * v.x + a[1] * v.y) + a[2] * v.z, (a[3] * v.x + a[4] * v.y) + a[5] * v.z,
          (a[6] * v.x + a[7] * v.y) + a[8] * v.z};
}
// World inverse inertia R * I^-1 * R^T  (state.cpp:95-97, particle.cpp:73-75).
D2_HD void world_inv_inertia(quat q, const double* ib, double* out) {
  double R[9], Rt[9], t[9];
  qrot(q, R);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Rt[3 * i + j] = R[3 * j + i];
  mul33(R, ib, t);
  mul33(t, Rt, out);
}
// Pose matrix (basic_utils.h:79-86).
D2_HD void pose(quat q, v3<double> t, double* m) {
  double r[9];
  qrot(q, r);
  m[0] = r[0], m[1] = r[1], m[2] = r[2], m[3] = t.x;
  m[4] = r[3], m[5] = r[4], m[6] = r[5], m[7] = t.y;
  m[8] = r[6], m[9] = r[7], m[10] = r[8], m[11] = t.z;
  m[12] = 0.0, m[13] = 0.0, m[14] = 0.0, m[15] = 1.0;
}
// State::extrapolate (state.cpp:90-102): translation t*v, rotation times
// exp(0.5*omega*t) on the right; basic_utils.cpp:36-44 for exp.
struct body_state {
  quat q;
  v3<double> t, v, L;
  double time;
};
D2_HD body_state extrapolate(const body_state& s, double dt, const double* ib) {
  body_state r;
  r.time = s.time + dt;
  r.v = s.v;
  r.t = add(s.t, sclL(dt, s.v));
  double Iw[9];
  world_inv_inertia(s.q, ib, Iw);
  v3<double> om = mul33v(Iw, s.L);
  v3<double> d = scl(sclL(0.5, om), dt);
  quat e{1.0, 0.0, 0.0, 0.0};
  if (!(vlen(d) == 0.0)) {
    double n = vlen(d);
    v3<double> u = scl(dvs(d, n), (double)sin(n));
    e = quat{(double)cos(n), u.x, u.y, u.z};
  }
  r.q = qmul(s.q, e);
  r.L = s.L;
  return r;
}
D2_HD body_state load_state(const double* r) {
  body_state s;
  s.q = quat{r[0], r[1], r[2], r[3]};
  s.t = v3<double>{r[4], r[5], r[6]};
  s.v = v3<double>{r[7], r[8], r[9]};
  s.L = v3<double>{r[10], r[11], r[12]};
  s.time = r[13];
  return s;
}
D2_HD void store_state(const body_state& s, double* r) {
  r[0] = s.q.w, r[1] = s.q.x, r[2] = s.q.y, r[3] = s.q.z;
  r[4] = s.t.x, r[5] = s.t.y, r[6] = s.t.z;
  r[7] = s.v.x, r[8] = s.v.y, r[9] = s.v.z;
  r[10] = s.L.x, r[11] = s.L.y, r[12] = s.L.z;
  r[13] = s.time;
}

// ---- triangles --------------------------------------------------------------
template <class R> struct tri {
  v3<R> a, b, c;
};
template <class R> D2_HD v3<R> corner(const tri<R>& t, int k) { return k == 0 ? t.a : (k == 1 ? t.b : t.c); }
template <class R> D2_HD tri<R> xform(const R* m, const tri<R>& t) { return {xform(m, t.a), xform(m, t.b), xform(m, t.c)}; }
template <class R> D2_HD v3<R> normal(const tri<R>& t) { return unit(cross(sub(t.b, t.a), sub(t.c, t.a))); }

// triangle.h:90-103 -- dot products in R, the solve in float.  The caller
// receives u, v widened back to R (the reference adds them in R).
template <class R> D2_HD void bary(const tri<R>& T, v3<R> P, R& u, R& v) {
  v3<R> v0 = sub(T.b, T.a), v1 = sub(T.c, T.a), v2 = sub(P, T.a);
  float d00 = (float)dot(v0, v0), d01 = (float)dot(v0, v1), d11 = (float)dot(v1, v1);
  float d20 = (float)dot(v2, v0), d21 = (float)dot(v2, v1);
  float den = d00 * d11 - d01 * d01;
  float fv = (d11 * d20 - d01 * d21) / den;
  float fw = (d00 * d21 - d01 * d20) / den;
  float fu = 1.0f - fv - fw;
  u = (R)fu;
  v = (R)fv;
}
template <class R> D2_HD bool inside(R u, R v) { return R(0) < u && u < R(1) && R(0) < v && v < R(1) && u + v < R(1); }

// triangle.h:133-141
template <class R> D2_HD v3<R> project_plane(const tri<R>& T, v3<R> p, bool& in) {
  v3<R> n = normal(T);
  v3<R> on = sub(p, sclL(dot(sub(p, T.a), n), n));
  R u, v;
  bary(T, on, u, v);
  in = inside(u, v);
  return on;
}
// triangle.h:105-131.  Returns t; sets tinf (|p2| <= epsilon) and hit.
template <class R> D2_HD R segment_hit(const tri<R>& T, v3<R> s, v3<R> e, bool& tinf, bool& hit) {
  v3<R> n = normal(T);
  v3<R> D = sub(e, s);
  R p1 = dot(sub(T.a, s), n);
  R p2 = dot(D, n);
  tinf = false;
  hit = false;
  const R eps = sizeof(R) == 8 ? (R)2.220446049250313080847e-16 : (R)1.1920928955078125e-07f;
  if (fabs(p2) <= eps) {
    tinf = true;
    return R(0);
  }
  R t = p1 / p2;
  if (R(0) < t && t < R(1)) {
    R u, v;
    bary(T, add(s, scl(D, t)), u, v);
    hit = inside(u, v);
  }
  return t;
}

// edge.h:61-65
template <class R> D2_HD v3<R> onto_segment(v3<R> A, v3<R> B, v3<R> p) {
  v3<R> ba = sub(B, A);
  R t = dot(sub(p, A), ba) / dot(ba, ba);
  return lerp(A, B, clamp01(t));
}
// edge.h:40-59: closest points of segments [A,B] (P side) and [C,D] (Q side)
// via projection into the plane orthogonal to CD; exact-equality parallel test.
template <class R> D2_HD R seg_closest(v3<R> A, v3<R> B, v3<R> C, v3<R> D, v3<R>& P, v3<R>& Q) {
  v3<R> dc = sub(D, C);
  R m = dot(dc, dc);
  v3<R> ipA = sub(A, scl(dc, dot(sub(A, C), dc) / m));
  v3<R> ipB = sub(B, scl(dc, dot(sub(B, C), dc) / m));
  v3<R> ipBA = sub(ipB, ipA);
  R t = dot(sub(C, ipA), ipBA) / dot(ipBA, ipBA);
  t = same(ipA, ipB) ? R(0) : t;
  v3<R> onAB = lerp(A, B, clamp01(t));
  Q = onto_segment(C, D, onAB);
  P = onto_segment(A, B, Q);
  return vlen(sub(Q, P));
}
template <class R> D2_HD R seg_param(v3<R> A, v3<R> B, v3<R> p) {
  v3<R> ab = sub(B, A);
  return dot(sub(p, A), ab) / dot(ab, ab);
}

// comparison_dist.h:12-189: closest points of two world-space triangles.
template <class R> D2_HD R tri_tri_dist(const tri<R>& a, const tri<R>& b, v3<R>& P, v3<R>& Q) {
  R best = R(1) / R(0);
  bool in;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    v3<R> q = corner(b, k);
    v3<R> p = project_plane(a, q, in);
    if (in) {
      R d = vlen(sub(p, q));
      if (d < best) P = p, Q = q, best = d;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    v3<R> p = corner(a, k);
    v3<R> q = project_plane(b, p, in);
    if (in) {
      R d = vlen(sub(p, q));
      if (d < best) P = p, Q = q, best = d;
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      v3<R> p, q;
      R d = seg_closest(corner(a, i), corner(a, (i + 1) % 3), corner(b, j), corner(b, (j + 1) % 3), p, q);
      if (d < best) P = p, Q = q, best = d;
    }
  bool tinf, hit;
  for (int j = 0; j < 3; ++j) {
    v3<R> s = corner(b, j), e = corner(b, (j + 1) % 3);
    R t = segment_hit(a, s, e, tinf, hit);
    if (hit) P = Q = lerp(s, e, t);
  }
  for (int i = 0; i < 3; ++i) {
    v3<R> s = corner(a, i), e = corner(a, (i + 1) % 3);
    R t = segment_hit(b, s, e, tinf, hit);
    if (hit) P = Q = lerp(s, e, t);
  }
  return vlen(sub(P, Q));
}

// ---- CCD (continuous_comparison.h) -------------------------------------------
struct ccd_out {
  double d, toc;
  v3<double> P, Q;
};

// continuous_comparison.h:41-98.  m0 = inverse(tri_T_start) * pt_T_start and
// m1 = inverse(tri_T_end) * pt_T_end are precomputed per body pair.
template <class R>
D2_HD R point_tri(v3<R> pt, const R* m0, const R* m1, const tri<R>& T, const R* Ts, const R* Te, v3<R>& P, v3<R>& Q,
                  R& toc) {
  R h[4];
  mul4v(m0, pt.x, pt.y, pt.z, R(1), h);
  const v3<R> ps = hdiv(h);
  mul4v(m1, pt.x, pt.y, pt.z, R(1), h);
  const v3<R> pe = hdiv(h);
  toc = R(1);
  bool tinf, hit;
  R t = segment_hit(T, ps, pe, tinf, hit);
  R M[16];
  if (R(0) < t && t < R(1) && hit) {
    R s = R(1) - t;
#pragma unroll
    for (int i = 0; i < 16; ++i) M[i] = s * Ts[i] + t * Te[i];
    P = Q = xform(M, add(scl(ps, R(1) - t), scl(pe, t)));
    toc = t;
    return R(0);
  }
  if (tinf) {
    bool above;
    v3<R> cl = lerp(ps, pe, R(0));
    v3<R> on = project_plane(T, cl, above);
    if (above) {
      P = xform(Ts, cl);
      Q = xform(Ts, on);
      toc = R(0);
      return vlen(sub(P, Q));
    }
    cl = lerp(ps, pe, R(1));
    on = project_plane(T, cl, above);
    if (above) {
      P = xform(Te, cl);
      Q = xform(Te, on);
      toc = R(1);
      return vlen(sub(P, Q));
    }
    return R(1) / R(0);
  }
  t = clamp01(t);
  R s = R(1) - t;
#pragma unroll
  for (int i = 0; i < 16; ++i) M[i] = s * Ts[i] + t * Te[i];
  v3<R> cl = lerp(ps, pe, t);
  bool above;
  v3<R> on = project_plane(T, cl, above);
  if (above) {
    P = xform(M, cl);
    Q = xform(M, on);
    toc = t;
    return vlen(sub(P, Q));
  }
  return R(1) / R(0);
}

// continuous_comparison.h:565-673: coplanarity bisection, then 10 Newton steps
// on |D(t)| where D is the reference's generated expression (its a_end.B +
// a_start.B*(1-t) term included), written as f' = D.dD/|D|,
// f'' = dD.dD/|D| - (D.dD)^2/|D|^3.
template <class R>
D2_HD R edge_edge(v3<R> asA, v3<R> asB, v3<R> aeA, v3<R> aeB, v3<R> bsA, v3<R> bsB, v3<R> beA, v3<R> beB, v3<R>& P,
                  v3<R>& Q, R& toc) {
  R best = sizeof(R) == 8 ? (R)1.7976931348623157e308 : (R)3.40282346638528859812e+38f;
  R dot_s = dot(sub(bsB, asA), cross(sub(asB, asA), sub(bsA, asA)));
  R dot_e = dot(sub(beB, aeA), cross(sub(aeB, aeA), sub(beA, aeA)));
  if (dot_s * dot_e < R(0)) {
    R lo = R(0), hi = R(1);
    for (int it = 0; it < 20; ++it) {
      R mid = (lo + hi) / R(2);
      v3<R> a0 = lerp(asA, aeA, mid), a1 = lerp(asB, aeB, mid);
      v3<R> b0 = lerp(bsA, beA, mid), b1 = lerp(bsB, beB, mid);
      R dm = dot(sub(b1, a0), cross(sub(a1, a0), sub(b0, a0)));
      if ((double)fabs(dm) < 1e-8) {
        lo = hi = mid;
        break;
      }
      if (dot_s * dm < R(0)) hi = mid;
      else lo = mid;
    }
    R mid = (lo + hi) / R(2);
    best = seg_closest(lerp(asA, aeA, mid), lerp(asB, aeB, mid), lerp(bsA, beA, mid), lerp(bsB, beB, mid), P, Q);
    toc = mid;
  }
  if ((double)best > 1e-8) {
    R t = R(0);
    for (int it = 0; it < 10; ++it) {
      v3<R> a0 = lerp(asA, aeA, t), a1 = lerp(asB, aeB, t);
      v3<R> b0 = lerp(bsA, beA, t), b1 = lerp(bsB, beB, t);
      v3<R> p, q;
      R d = seg_closest(a0, a1, b0, b1, p, q);
      if (d < best) best = d, P = p, Q = q, toc = t;
      R pa = seg_param(a0, a1, p);
      R pb = seg_param(b0, b1, q);
      R u = R(1) - t, ca = R(1) - pa, cb = R(1) - pb;
      R D[3], dD[3];
      const R sA[3] = {asA.x, asA.y, asA.z}, sB[3] = {asB.x, asB.y, asB.z};
      const R eA[3] = {aeA.x, aeA.y, aeA.z}, eB[3] = {aeB.x, aeB.y, aeB.z};
      const R tA[3] = {bsA.x, bsA.y, bsA.z}, tB[3] = {bsB.x, bsB.y, bsB.z};
      const R fA[3] = {beA.x, beA.y, beA.z}, fB[3] = {beB.x, beB.y, beB.z};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        D[k] = ((pa * (eB[k] + sB[k] * u) - pb * (fB[k] + tB[k] * u)) + ca * (eA[k] * t + sA[k] * u)) -
               cb * (fA[k] * t + tA[k] * u);
        dD[k] = ((-sB[k] * pa + tB[k] * pb) + ca * (eA[k] - sA[k])) + (fA[k] - tA[k]) * (pb - R(1));
      }
      R S = (D[0] * D[0] + D[1] * D[1]) + D[2] * D[2];
      R N = (D[0] * dD[0] + D[1] * dD[1]) + D[2] * dD[2];
      R Mm = (dD[0] * dD[0] + dD[1] * dD[1]) + dD[2] * dD[2];
      R rS = sqrt(S);
      R f1 = N / rS;
      R f2 = Mm / rS - (N * N) / (S * rS);
      t = clamp01(t - f1 / f2);
    }
  }
  return best;
}

}  // namespace d2
