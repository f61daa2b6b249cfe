// d2o_geom.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Geometric primitives, the discrete triangle-triangle distance and the
// linear-interpolation CCD kernels of Delta2, restated without Eigen.  Every
// function cites the reference lines it follows.  Quirks of the reference are
// kept on purpose (float barycentrics, exact-equality parallel test, the
// a_end.B + a_start.B*(1-t) term of the Newton step, P/Q orders).
#pragma once
#include <tuple>

#include "d2o_math.hpp"

namespace d2o {

// common/edge.h:8-75
template <class R> struct Edge {
  V3<R> A, B;
  Edge() {}
  Edge(const V3<R>& a, const V3<R>& b) : A(a), B(b) {}
  V3<R> lerpAt(R t) const { return d2o::lerp(A, B, t); }  // edge.h:67-69
  // edge.h:61-65
  V3<R> constrainToSegment(const V3<R>& p) const {
    V3<R> ba = B - A;
    R t = dot(p - A, ba) / dot(ba, ba);
    return lerpAt(clamp01(t));
  }
  // edge.h:71-74
  R project(const V3<R>& v) const {
    V3<R> ab = B - A;
    return dot(v - A, ab) / dot(ab, ab);
  }
  // edge.h:40-59: closest points by projecting this segment into the plane
  // orthogonal to E; the parallel test is exact vector equality (edge.h:49).
  R closest(const Edge& E, V3<R>& P_out, V3<R>& Q_out) const {
    V3<R> segDC = E.B - E.A;
    R lineDirSqrMag = dot(segDC, segDC);
    V3<R> segAC = A - E.A;
    V3<R> segBC = B - E.A;
    V3<R> inPlaneA = A - ((dot(segAC, segDC) / lineDirSqrMag) * segDC);
    V3<R> inPlaneB = B - ((dot(segBC, segDC) / lineDirSqrMag) * segDC);
    V3<R> inPlaneBA = inPlaneB - inPlaneA;
    R t = dot(E.A - inPlaneA, inPlaneBA) / dot(inPlaneBA, inPlaneBA);
    t = neq(inPlaneA, inPlaneB) ? t : R(0);
    V3<R> segABtoLineCD = lerpAt(clamp01(t));
    V3<R> segCDtoSegAB = E.constrainToSegment(segABtoLineCD);
    V3<R> segABtoSegCD = constrainToSegment(segCDtoSegAB);
    Q_out = segCDtoSegAB;
    P_out = segABtoSegCD;
    return norm(Q_out - P_out);
  }
  R dist(const Edge& E) const {
    V3<R> p, q;
    return closest(E, p, q);
  }
};

// common/bbox.h:8-46
template <class R> inline bool overlappingRange(R a_lower, R a_upper, R b_lower, R b_upper) {
  if (b_lower < a_lower) return b_upper > a_lower;
  if (b_lower < a_upper) return true;
  return false;
}
template <class R> struct BBox {
  V3<R> lo, hi;
  bool overlap(const BBox& o) const {
    bool x = overlappingRange(lo.x, hi.x, o.lo.x, o.hi.x);
    bool y = overlappingRange(lo.y, hi.y, o.lo.y, o.hi.y);
    bool z = overlappingRange(lo.z, hi.z, o.lo.z, o.hi.z);
    return x && y && z;
  }
  BBox expand(const BBox& o) const {
    return {{std::min(lo.x, o.lo.x), std::min(lo.y, o.lo.y), std::min(lo.z, o.lo.z)},
            {std::max(hi.x, o.hi.x), std::max(hi.y, o.hi.y), std::max(hi.z, o.hi.z)}};
  }
  BBox expand(R e) const { return {{lo.x - e, lo.y - e, lo.z - e}, {hi.x + e, hi.y + e, hi.z + e}}; }
};

// common/triangle.h:12-502
template <class R> struct Tri {
  V3<R> A, B, C;
  Tri() {}
  Tri(const V3<R>& a, const V3<R>& b, const V3<R>& c) : A(a), B(b), C(c) {}
  const V3<R>& operator[](int i) const { return i == 0 ? A : (i == 1 ? B : C); }
  R area() const { return R(0.5) * norm(cross(A - C, B - C)); }  // triangle.h:21-23
  V3<R> normal() const { return normalized(cross(B - A, C - A)); }  // triangle.h:43-45
  Edge<R> AB() const { return Edge<R>(A, B); }
  Edge<R> BC() const { return Edge<R>(B, C); }
  Edge<R> CA() const { return Edge<R>(C, A); }
  // triangle.h:78-88
  BBox<R> bbox() const {
    return {{std::min(std::min(A.x, B.x), C.x), std::min(std::min(A.y, B.y), C.y), std::min(std::min(A.z, B.z), C.z)},
            {std::max(std::max(A.x, B.x), C.x), std::max(std::max(A.y, B.y), C.y), std::max(std::max(A.z, B.z), C.z)}};
  }
  // triangle.h:90-103: dot products in R, everything after them in float.
  void calcBarycentric(const V3<R>& P, R& u_out, R& v_out, R& w_out) const {
    V3<R> v0 = B - A, v1 = C - A, v2 = P - A;
    float d00 = (float)dot(v0, v0);
    float d01 = (float)dot(v0, v1);
    float d11 = (float)dot(v1, v1);
    float d20 = (float)dot(v2, v0);
    float d21 = (float)dot(v2, v1);
    float denom = d00 * d11 - d01 * d01;
    float v = (d11 * d20 - d01 * d21) / denom;
    float w = (d00 * d21 - d01 * d20) / denom;
    float u = 1.0f - v - w;
    u_out = R(u);
    v_out = R(v);
    w_out = R(w);
  }
  // triangle.h:105-131
  V3<R> intersectSegment(const Edge<R>& E, R& t, bool& t_inf, bool& isIntersecting) const {
    const V3<R>& start = E.A;
    const V3<R>& end = E.B;
    V3<R> AP = A - start;
    V3<R> n = normal();
    V3<R> D = end - start;
    R p1 = dot(AP, n);
    R p2 = dot(D, n);
    t_inf = false;
    if (std::fabs(p2) <= std::numeric_limits<R>::epsilon()) {
      t = R(0);
      t_inf = true;
      isIntersecting = false;
      return V3<R>(0, 0, 0);
    }
    t = p1 / p2;
    if (R(0) < t && t < R(1)) {
      R u, v, w;
      calcBarycentric(start + D * t, u, v, w);
      isIntersecting = R(0) < u && u < R(1) && R(0) < v && v < R(1) && u + v < R(1);
      return start * (R(1) - t) + end * t;
    }
    isIntersecting = false;
    return V3<R>(0, 0, 0);
  }
  // triangle.h:133-141
  V3<R> projectToPlane(const V3<R>& pt, bool& isInTriangle) const {
    V3<R> AP = pt - A;
    V3<R> n = normal();
    V3<R> onPlane = pt - dot(AP, n) * n;
    R u, v, w;
    calcBarycentric(onPlane, u, v, w);
    isInTriangle = R(0) < u && u < R(1) && R(0) < v && v < R(1) && u + v < R(1);
    return onPlane;
  }
  // triangle.h:280-471: closest point on a triangle (Eberly's region method).
  R distToPoint(const V3<R>& X) const {
    V3<R> AX = X - A, AB_ = B - A, AC_ = C - A;
    R a00 = dot(AB_, AB_), a01 = dot(AB_, AC_), a11 = dot(AC_, AC_);
    R b0 = -dot(AX, AB_), b1 = -dot(AX, AC_);
    R det = a00 * a11 - a01 * a01;
    R t0 = a01 * b1 - a11 * b0;
    R t1 = a01 * b0 - a00 * b1;
    if (t0 + t1 <= det) {
      if (t0 < R(0)) {
        if (t1 < R(0)) {  // region 4
          if (b0 < R(0)) {
            t1 = R(0);
            t0 = (-b0 >= a00) ? R(1) : -b0 / a00;
          } else {
            t0 = R(0);
            if (b1 >= R(0)) t1 = R(0);
            else if (-b1 >= a11) t1 = R(1);
            else t1 = -b1 / a11;
          }
        } else {  // region 3
          t0 = R(0);
          if (b1 >= R(0)) t1 = R(0);
          else if (-b1 >= a11) t1 = R(1);
          else t1 = -b1 / a11;
        }
      } else if (t1 < R(0)) {  // region 5
        t1 = R(0);
        if (b0 >= R(0)) t0 = R(0);
        else if (-b0 >= a00) t0 = R(1);
        else t0 = -b0 / a00;
      } else {  // region 0
        R invDet = R(1) / det;
        t0 *= invDet;
        t1 *= invDet;
      }
    } else {
      R tmp0, tmp1, numer, denom;
      if (t0 < R(0)) {  // region 2
        tmp0 = a01 + b0;
        tmp1 = a11 + b1;
        if (tmp1 > tmp0) {
          numer = tmp1 - tmp0;
          denom = a00 - R(2) * a01 + a11;
          if (numer >= denom) {
            t0 = R(1);
            t1 = R(0);
          } else {
            t0 = numer / denom;
            t1 = R(1) - t0;
          }
        } else {
          t0 = R(0);
          if (tmp1 <= R(0)) t1 = R(1);
          else if (b1 >= R(0)) t1 = R(0);
          else t1 = -b1 / a11;
        }
      } else if (t1 < R(0)) {  // region 6
        tmp0 = a01 + b1;
        tmp1 = a00 + b0;
        if (tmp1 > tmp0) {
          numer = tmp1 - tmp0;
          denom = a00 - R(2) * a01 + a11;
          if (numer >= denom) {
            t1 = R(1);
            t0 = R(0);
          } else {
            t1 = numer / denom;
            t0 = R(1) - t1;
          }
        } else {
          t1 = R(0);
          if (tmp1 <= R(0)) t0 = R(1);
          else if (b0 >= R(0)) t0 = R(0);
          else t0 = -b0 / a00;
        }
      } else {  // region 1
        numer = a11 + b1 - a01 - b0;
        if (numer <= R(0)) {
          t0 = R(0);
          t1 = R(1);
        } else {
          denom = a00 - R(2) * a01 + a11;
          if (numer >= denom) {
            t0 = R(1);
            t1 = R(0);
          } else {
            t0 = numer / denom;
            t1 = R(1) - t0;
          }
        }
      }
    }
    V3<R> diff = X - ((A + t0 * AB_) + t1 * AC_);
    return norm(diff);
  }
  // triangle.h:472-487
  Tri transformed(const M4<R>& T) const { return Tri(transform(A, T), transform(B, T), transform(C, T)); }
};

// ---------------------------------------------------------------------------
// Discrete triangle-triangle closest points: comparison_dist.h:12-189.
// Order: vertex->face (b onto a, then a onto b), 9 edge-edge, then 6
// segment-triangle intersections where the last intersecting one wins.
template <class R> inline R comparisonDist(const Tri<R>& a, const Tri<R>& b, V3<R>& P, V3<R>& Q) {
  V3<R> Pc, Qc, Pt, Qt;
  R best = std::numeric_limits<R>::infinity();
  bool in;
  const V3<R>* bv[3] = {&b.A, &b.B, &b.C};
  for (int k = 0; k < 3; ++k) {
    Pt = a.projectToPlane(*bv[k], in);
    if (in) {
      R d = norm(Pt - *bv[k]);
      if (d < best) {
        Pc = Pt;
        Qc = *bv[k];
        best = d;
      }
    }
  }
  const V3<R>* av[3] = {&a.A, &a.B, &a.C};
  for (int k = 0; k < 3; ++k) {
    Qt = b.projectToPlane(*av[k], in);
    if (in) {
      R d = norm(*av[k] - Qt);
      if (d < best) {
        Pc = *av[k];
        Qc = Qt;
        best = d;
      }
    }
  }
  Edge<R> ae[3] = {a.AB(), a.BC(), a.CA()};
  Edge<R> be[3] = {b.AB(), b.BC(), b.CA()};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R d = ae[i].closest(be[j], Pt, Qt);
      if (d < best) {
        Pc = Pt;
        Qc = Qt;
        best = d;
      }
    }
  R t;
  bool tinf, isect;
  for (int j = 0; j < 3; ++j) {
    a.intersectSegment(be[j], t, tinf, isect);
    if (isect) {
      Pc = be[j].lerpAt(t);
      Qc = be[j].lerpAt(t);
      best = R(0);
    }
  }
  for (int i = 0; i < 3; ++i) {
    b.intersectSegment(ae[i], t, tinf, isect);
    if (isect) {
      Pc = ae[i].lerpAt(t);
      Qc = ae[i].lerpAt(t);
      best = R(0);
    }
  }
  (void)best;
  P = Pc;
  Q = Qc;
  return norm(Pc - Qc);
}

// continuous_comparison.h:12-36
template <class R>
inline R sphereSphereFirstIntersect(const V3<R>& a_pos, const V3<R>& a_vel, R a_rad, const V3<R>& b_pos, const V3<R>& b_vel,
                                    R b_rad) {
  V3<R> ab = a_pos - b_pos;
  V3<R> rel = a_vel - b_vel;
  R r = a_rad + b_rad;
  R c = dot(ab, ab) - r * r;
  if (c < R(0)) return R(0);
  R a = dot(rel, rel);
  R b = dot(rel, ab);
  if (b >= R(0)) return R(-1);
  R d = b * b - a * c;
  if (d < R(0)) return R(-1);
  return (-b - std::sqrt(d)) / a;
}

// continuous_comparison.h:41-98.  The point is taken into the triangle's
// local frame at start and end (inverse(tri_T) * pt_T * pt), the local
// segment is intersected with the triangle, outputs go back to world through
// the elementwise blend of the two triangle matrices.
template <class R>
inline R pointTriCCDLinear(const V3<R>& pt, const M4<R>& pt_T_start, const M4<R>& pt_T_end, const Tri<R>& tri,
                           const M4<R>& tri_T_start, const M4<R>& tri_T_end, V3<R>& P_out, V3<R>& Q_out, R& TOC_out) {
  M4<R> m0 = mul(inverse(tri_T_start), pt_T_start);
  M4<R> m1 = mul(inverse(tri_T_end), pt_T_end);
  R h[4];
  mul4(m0, pt.x, pt.y, pt.z, R(1), h);
  V3<R> ps(h[0] / h[3], h[1] / h[3], h[2] / h[3]);
  mul4(m1, pt.x, pt.y, pt.z, R(1), h);
  V3<R> pe(h[0] / h[3], h[1] / h[3], h[2] / h[3]);
  Edge<R> e(ps, pe);
  TOC_out = R(1);
  bool isect, tinf;
  R t;
  V3<R> hit = tri.intersectSegment(e, t, tinf, isect);
  if (R(0) < t && t < R(1) && isect) {
    M4<R> tT = lerpM(tri_T_start, tri_T_end, t);
    P_out = transform(hit, tT);
    Q_out = transform(hit, tT);
    TOC_out = t;
    return R(0);
  } else if (tinf) {
    t = R(0);
    V3<R> cl = e.lerpAt(t);
    bool above;
    V3<R> on = tri.projectToPlane(cl, above);
    if (above) {
      P_out = transform(cl, tri_T_start);
      Q_out = transform(on, tri_T_start);
      TOC_out = t;
      return norm(P_out - Q_out);
    }
    t = R(1);
    cl = e.lerpAt(t);
    on = tri.projectToPlane(cl, above);
    if (above) {
      P_out = transform(cl, tri_T_end);
      Q_out = transform(on, tri_T_end);
      TOC_out = t;
      return norm(P_out - Q_out);
    }
  } else {
    t = clamp01(t);
    M4<R> tT = lerpM(tri_T_start, tri_T_end, t);
    V3<R> cl = e.lerpAt(t);
    bool above;
    V3<R> on = tri.projectToPlane(cl, above);
    if (above) {
      P_out = transform(cl, tT);
      Q_out = transform(on, tT);
      TOC_out = t;
      return norm(P_out - Q_out);
    }
  }
  return std::numeric_limits<R>::infinity();
}

// continuous_comparison.h:565-673.  Sign change of the coplanarity triple
// product -> 20-step bisection; then (if still apart) 10 Newton steps on the
// distance between the lerped edges' closest points.  The Newton derivative is
// the reference's generated expression in factored form: per axis
//   D  = pa*(aeB + asB*(1-t)) - pb*(beB + bsB*(1-t))
//        + (1-pa)*(aeA*t + asA*(1-t)) - (1-pb)*(beA*t + bsA*(1-t))
//   dD = -asB*pa + bsB*pb + (1-pa)*(aeA-asA) + (beA-bsA)*(pb-1)
// (the aeB term without a factor t is the reference's, continuous_comparison.h:630)
//   f' = (D.dD)/|D|,  f'' = (dD.dD)/|D| - (D.dD)^2/|D|^3.
template <class R>
inline R edgeEdgeCCD(const Edge<R>& as, const Edge<R>& ae, const Edge<R>& bs, const Edge<R>& be, V3<R>& P_out,
                     V3<R>& Q_out, R& toc_out) {
  R min_dist = std::numeric_limits<R>::max();
  V3<R> n_start = cross(as.B - as.A, bs.A - as.A);
  R dot_start = dot(bs.B - as.A, n_start);
  V3<R> n_end = cross(ae.B - ae.A, be.A - ae.A);
  R dot_end = dot(be.B - ae.A, n_end);
  if (dot_start * dot_end < R(0)) {
    R lo = R(0), hi = R(1);
    for (int it = 0; it < 20; ++it) {
      R mid = (lo + hi) / R(2);
      Edge<R> a_t(lerp(as.A, ae.A, mid), lerp(as.B, ae.B, mid));
      Edge<R> b_t(lerp(bs.A, be.A, mid), lerp(bs.B, be.B, mid));
      V3<R> n_mid = cross(a_t.B - a_t.A, b_t.A - a_t.A);
      R dot_mid = dot(b_t.B - a_t.A, n_mid);
      if ((double)std::fabs(dot_mid) < 1e-8) {
        lo = mid;
        hi = mid;
        break;
      } else if (dot_start * dot_mid < R(0)) {
        hi = mid;
      } else {
        lo = mid;
      }
    }
    R mid = (lo + hi) / R(2);
    Edge<R> a_t(lerp(as.A, ae.A, mid), lerp(as.B, ae.B, mid));
    Edge<R> b_t(lerp(bs.A, be.A, mid), lerp(bs.B, be.B, mid));
    min_dist = a_t.closest(b_t, P_out, Q_out);
    toc_out = mid;
  }
  if ((double)min_dist > 1e-8) {
    R t = R(0);
    for (int it = 0; it < 10; ++it) {
      Edge<R> a_t(lerp(as.A, ae.A, t), lerp(as.B, ae.B, t));
      Edge<R> b_t(lerp(bs.A, be.A, t), lerp(bs.B, be.B, t));
      V3<R> P, Q;
      R d = a_t.closest(b_t, P, Q);
      if (d < min_dist) {
        min_dist = d;
        P_out = P;
        Q_out = Q;
        toc_out = t;
      }
      R pa = a_t.project(P);
      R pb = b_t.project(Q);
      R u = R(1) - t;
      R ca = R(1) - pa, cb = R(1) - pb;
      R D[3], dD[3];
      const R asA[3] = {as.A.x, as.A.y, as.A.z}, asB[3] = {as.B.x, as.B.y, as.B.z};
      const R aeA[3] = {ae.A.x, ae.A.y, ae.A.z}, aeB[3] = {ae.B.x, ae.B.y, ae.B.z};
      const R bsA[3] = {bs.A.x, bs.A.y, bs.A.z}, bsB[3] = {bs.B.x, bs.B.y, bs.B.z};
      const R beA[3] = {be.A.x, be.A.y, be.A.z}, beB[3] = {be.B.x, be.B.y, be.B.z};
      for (int k = 0; k < 3; ++k) {
        D[k] = ((pa * (aeB[k] + asB[k] * u) - pb * (beB[k] + bsB[k] * u)) + ca * (aeA[k] * t + asA[k] * u)) -
               cb * (beA[k] * t + bsA[k] * u);
        dD[k] = ((-asB[k] * pa + bsB[k] * pb) + ca * (aeA[k] - asA[k])) + (beA[k] - bsA[k]) * (pb - R(1));
      }
      R S = (D[0] * D[0] + D[1] * D[1]) + D[2] * D[2];
      R N = (D[0] * dD[0] + D[1] * dD[1]) + D[2] * dD[2];
      R M = (dD[0] * dD[0] + dD[1] * dD[1]) + dD[2] * dD[2];
      R rS = std::sqrt(S);
      R f1 = N / rS;
      R f2 = M / rS - (N * N) / (S * rS);
      t = t - f1 / f2;
      t = clamp01(t);
    }
  }
  return min_dist;
}

// continuous_comparison.h:680-779.  Selection: strictly smaller distance, or
// equal distance with an earlier toc.  P/Q of the b->a point test are not
// swapped (continuous_comparison.h:721).
template <class R>
inline R triTriCCDLinear(const Tri<R>& a, const M4<R>& aTs, const M4<R>& aTe, const Tri<R>& b, const M4<R>& bTs,
                         const M4<R>& bTe, V3<R>& P_out, V3<R>& Q_out, R& TOC_out) {
  R best = std::numeric_limits<R>::infinity();
  TOC_out = R(1);
  Tri<R> a0 = a.transformed(aTs), a1 = a.transformed(aTe), b0 = b.transformed(bTs), b1 = b.transformed(bTe);
  for (int i = 0; i < 3; ++i) {
    V3<R> P, Q;
    R tc;
    R d = pointTriCCDLinear(a[i], aTs, aTe, b, bTs, bTe, P, Q, tc);
    if (d < best || (d == best && tc < TOC_out)) {
      P_out = P;
      Q_out = Q;
      best = d;
      TOC_out = tc;
    }
    d = pointTriCCDLinear(b[i], bTs, bTe, a, aTs, aTe, P, Q, tc);
    if (d < best || (d == best && tc < TOC_out)) {
      P_out = P;
      Q_out = Q;
      best = d;
      TOC_out = tc;
    }
    int i1 = (i + 1) % 3;
    Edge<R> aes(a0[i], a0[i1]), aee(a1[i], a1[i1]);
    for (int j = 0; j < 3; ++j) {
      int j1 = (j + 1) % 3;
      Edge<R> bes(b0[j], b0[j1]), bee(b1[j], b1[j1]);
      R t;
      d = edgeEdgeCCD(aes, aee, bes, bee, P, Q, t);
      if (d < best || (d == best && t < TOC_out)) {
        best = d;
        P_out = P;
        Q_out = Q;
        TOC_out = t;
      }
    }
  }
  return best;
}

}  // namespace d2o
