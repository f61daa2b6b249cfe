// d2_cull.hpp -- FP32 culls of the CCD sub-tests (gfx950 kernels and the CPU
// soundness check in tests/cpu_geom_check.cpp).
//
// A cull decides only whether an FP64 sub-test runs: a culled sub-test
// provably reports no distance below the search distance, so it cannot be the
// winner of a hit, and results are identical with the culls on or off.
//   point-triangle (pointTriCCDLinear, continuous_comparison.h:41-98): the
//     point moves on the segment p0 -> p1 in the triangle's frame and the
//     reported distance is 0 (the segment crosses the triangle), the plane
//     distance at t = 0 or 1 (projection inside), or infinite.  Culled when
//     p0 and p1 lie on one side of the plane farther than the search distance
//     plus margin (no crossing; both plane distances large), or both lie
//     outside one edge's half-plane by more than the margin plus the
//     barycentric acceptance bound (no projection or crossing inside).
//   edge-edge (edgeEdgeCCD, :565-673): the reported distance is |P - Q| for
//     points of the two lerped segments at one t, so it is at least the
//     separation of the two segments along an axis that separates them on the
//     same side at t = 0 and at t = 1 (point offsets are linear in t).  Axes:
//     the midpoint offsets and the edge-direction cross products at both ends.
// margin = 1e-6 m + 64 eps_f |coords| (FP32 transforms and dot products),
// where |coords| bounds every magnitude the transforms involve: the results,
// the body coordinates transformed and the matrices' translations (a frame
// change between two far-apart body origins cancels large terms, and its
// rounding scales with them, not with the small result).
//
// The includer defines D2_HD (see d2_geom.hpp, included first).
#pragma once

namespace d2 {

// FP32 culls only decide which FP64 tests run, so their arithmetic may fuse:
// an FMA rounds once where a multiply-add rounds twice, so the error bounds
// the margins cover (gamma_n |a||b| for an n-term dot product) still hold.
D2_HD float f3dot(const float* a, const float* b) {
  return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}
// Poses as float 3x4 matrices, translations taken relative to an origin
// (b's start position) so that world coordinates stay small.
struct M34f {
  float m[12];
};
D2_HD M34f to_m34f(const double* m, const double* org) {
  M34f o;
  for (int r = 0; r < 3; ++r) {
    o.m[4 * r] = (float)m[4 * r], o.m[4 * r + 1] = (float)m[4 * r + 1], o.m[4 * r + 2] = (float)m[4 * r + 2];
    o.m[4 * r + 3] = (float)(m[4 * r + 3] - (org ? org[4 * r + 3] : 0.0));
  }
  return o;
}
D2_HD void xf_f(const M34f& M, const float* p, float* o) {  // p: body coordinates rounded to float
  const float x = p[0], y = p[1], z = p[2];
  o[0] = (M.m[0] * x + M.m[1] * y) + (M.m[2] * z + M.m[3]);
  o[1] = (M.m[4] * x + M.m[5] * y) + (M.m[6] * z + M.m[7]);
  o[2] = (M.m[8] * x + M.m[9] * y) + (M.m[10] * z + M.m[11]);
}
D2_HD void xf(const M34f& M, const v3<double>& p, float* o) {
  const float x = (float)p.x, y = (float)p.y, z = (float)p.z;
  o[0] = (M.m[0] * x + M.m[1] * y) + (M.m[2] * z + M.m[3]);
  o[1] = (M.m[4] * x + M.m[5] * y) + (M.m[6] * z + M.m[7]);
  o[2] = (M.m[8] * x + M.m[9] * y) + (M.m[10] * z + M.m[11]);
}
D2_HD float amax3(const float* v) { return fmaxf(fabsf(v[0]), fmaxf(fabsf(v[1]), fabsf(v[2]))); }
// The largest translation entry of a pose, and of a body-coordinate point:
// magnitudes the FP32 transform's rounding scales with.
D2_HD float m34_mag(const M34f& M) { return fmaxf(fabsf(M.m[3]), fmaxf(fabsf(M.m[7]), fabsf(M.m[11]))); }
D2_HD float v3_mag(const v3<double>& p) { return fmaxf(fabsf((float)p.x), fmaxf(fabsf((float)p.y), fabsf((float)p.z))); }
// Separation of segment Q above segment P along u at one instant (negative
// when none), scaled by 1/|u|.
D2_HD float seg_sep(const float* u, float nu, const float* p0, const float* p1, const float* q0, const float* q1) {
  const float a0 = f3dot(u, p0), a1 = f3dot(u, p1), b0 = f3dot(u, q0), b1 = f3dot(u, q1);
  return (fminf(b0, b1) - fmaxf(a0, a1)) / nu;
}
// extra: the magnitudes of the transforms' inputs (body coordinates,
// translations) that made these points.
D2_HD bool cull_edge_edge(const float* as0, const float* as1, const float* ae0, const float* ae1, const float* bs0,
                          const float* bs1, const float* be0, const float* be1, float thr, float extra) {
  const float EPSF = 5.9604645e-08f;
  const float mag = fmaxf(fmaxf(fmaxf(fmaxf(amax3(as0), amax3(as1)), fmaxf(amax3(ae0), amax3(ae1))),
                                fmaxf(fmaxf(amax3(bs0), amax3(bs1)), fmaxf(amax3(be0), amax3(be1)))),
                          extra);
  const float lim = thr + 1e-6f + 64.f * EPSF * mag;
  float ax[4][3];
  for (int r = 0; r < 3; ++r) {
    ax[0][r] = 0.5f * ((bs0[r] + bs1[r]) - (as0[r] + as1[r]));
    ax[1][r] = 0.5f * ((be0[r] + be1[r]) - (ae0[r] + ae1[r]));
  }
  const float da[3] = {as1[0] - as0[0], as1[1] - as0[1], as1[2] - as0[2]};
  const float db[3] = {bs1[0] - bs0[0], bs1[1] - bs0[1], bs1[2] - bs0[2]};
  const float ea[3] = {ae1[0] - ae0[0], ae1[1] - ae0[1], ae1[2] - ae0[2]};
  const float eb[3] = {be1[0] - be0[0], be1[1] - be0[1], be1[2] - be0[2]};
  ax[2][0] = da[1] * db[2] - da[2] * db[1], ax[2][1] = da[2] * db[0] - da[0] * db[2], ax[2][2] = da[0] * db[1] - da[1] * db[0];
  ax[3][0] = ea[1] * eb[2] - ea[2] * eb[1], ax[3][1] = ea[2] * eb[0] - ea[0] * eb[2], ax[3][2] = ea[0] * eb[1] - ea[1] * eb[0];
  for (int k = 0; k < 4; ++k) {
    const float nu = sqrtf(f3dot(ax[k], ax[k]));
    if (!(nu > 1e-20f)) continue;
    const float up = fminf(seg_sep(ax[k], nu, as0, as1, bs0, bs1), seg_sep(ax[k], nu, ae0, ae1, be0, be1));
    const float dn = fminf(seg_sep(ax[k], nu, bs0, bs1, as0, as1), seg_sep(ax[k], nu, be0, be1, ae0, ae1));
    if (fmaxf(up, dn) > lim) return true;
  }
  return false;
}

// All nine edge-edge sub-tests of a triangle pair at once: the vertices of
// triangle a (a0 at t = 0, a1 at t = 1, world coordinates as in
// cull_edge_edge) and of b.  A point of any edge of a at time t is a convex
// combination of a's lerped vertices, so along a fixed axis u its projection
// lies within the lerp of a's extremes at t = 0 and t = 1, and likewise for b;
// an axis that separates the two vertex sets on the same side at both ends
// by more than cull_edge_edge's limit therefore culls every edge pair.  Axes:
// the centroid offset over the step and both triangles' normals at both ends.
D2_HD bool cull_tri_edges(const float (*a0)[3], const float (*a1)[3], const float (*b0)[3], const float (*b1)[3],
                          float thr, float extra) {
  const float EPSF = 5.9604645e-08f;
  float mag = extra;
  for (int k = 0; k < 3; ++k)
    mag = fmaxf(mag, fmaxf(fmaxf(amax3(a0[k]), amax3(a1[k])), fmaxf(amax3(b0[k]), amax3(b1[k]))));
  const float lim = thr + 1e-6f + 64.f * EPSF * mag;
  float ax[5][3];
  for (int r = 0; r < 3; ++r)
    ax[0][r] = ((b0[0][r] + b0[1][r]) + (b0[2][r] + b1[0][r]) + (b1[1][r] + b1[2][r])) -
               ((a0[0][r] + a0[1][r]) + (a0[2][r] + a1[0][r]) + (a1[1][r] + a1[2][r]));
  const float (*T[4])[3] = {a0, a1, b0, b1};
  for (int q = 0; q < 4; ++q) {
    const float (*p)[3] = T[q];
    const float e0[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
    const float e1[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
    ax[1 + q][0] = e0[1] * e1[2] - e0[2] * e1[1], ax[1 + q][1] = e0[2] * e1[0] - e0[0] * e1[2];
    ax[1 + q][2] = e0[0] * e1[1] - e0[1] * e1[0];
  }
  for (int k = 0; k < 5; ++k) {
    const float nu = sqrtf(f3dot(ax[k], ax[k]));
    if (!(nu > 1e-20f)) continue;
    float amin0 = 3.4e38f, amax0 = -3.4e38f, amin1 = 3.4e38f, amax1 = -3.4e38f;
    float bmin0 = 3.4e38f, bmax0 = -3.4e38f, bmin1 = 3.4e38f, bmax1 = -3.4e38f;
    for (int v = 0; v < 3; ++v) {
      float x = f3dot(ax[k], a0[v]);
      amin0 = fminf(amin0, x), amax0 = fmaxf(amax0, x);
      x = f3dot(ax[k], a1[v]);
      amin1 = fminf(amin1, x), amax1 = fmaxf(amax1, x);
      x = f3dot(ax[k], b0[v]);
      bmin0 = fminf(bmin0, x), bmax0 = fmaxf(bmax0, x);
      x = f3dot(ax[k], b1[v]);
      bmin1 = fminf(bmin1, x), bmax1 = fmaxf(bmax1, x);
    }
    const float up = fminf(bmin0 - amax0, bmin1 - amax1), dn = fminf(amin0 - bmax0, amin1 - bmax1);
    if (fmaxf(up, dn) / nu > lim) return true;
  }
  return false;
}

// A triangle's plane and outward edge normals in its body frame, computed
// in FP64 and rounded once to float: unit normal n (the reference's normal()
// bit for bit before the rounding), offset d = n.A, in-plane edge normals
// m[q] pointing away from the opposite corner with offsets o[q], and eo =
// 64 eps_f kappa L (kappa = d00 d11 / (d00 d11 - d01^2), L the longest edge),
// the acceptance bound of the reference's float-barycentric inside test
// (huge when the system is ill-conditioned, denominator <= 1e-4 d00 d11).  A
// degenerate triangle gets n = 0 and o = +huge: its sub-tests are never
// culled.
struct SoupTri {
  float n[3], d, m[9], o[3], eo;
};
D2_HD SoupTri soup_tri(const tri<double>& T) {
  SoupTri s;
  for (int r = 0; r < 3; ++r) s.n[r] = 0.f, s.m[r] = s.m[3 + r] = s.m[6 + r] = 0.f;
  s.d = 0.f;
  s.o[0] = s.o[1] = s.o[2] = 3.4e38f;  // never outside
  s.eo = 3.4e38f;
  const v3<double> e0 = sub(T.b, T.a), e1 = sub(T.c, T.a), e2 = sub(T.c, T.b);
  const v3<double> cr = cross(e0, e1);
  const double cl = vlen(cr);
  const double d00 = dot(e0, e0), d11 = dot(e1, e1), d01 = dot(e0, e1), d22 = dot(e2, e2);
  const double den = d00 * d11 - d01 * d01;
  if (!(cl > 0.0) || !(den > 0.0)) return s;  // degenerate: n = 0 keeps every sub-test
  const double L = sqrt(d00 > d11 ? (d00 > d22 ? d00 : d22) : (d11 > d22 ? d11 : d22));
  // the reference's barycentrics are float (triangle.h:92-100): an
  // ill-conditioned system gets no lateral cull, the others 64 eps_f kappa L
  if (den > 1e-4 * d00 * d11) s.eo = (float)(64.0 * 5.9604644775390625e-08 * (d00 * d11 / den) * L * (1.0 + 1e-6));
  const v3<double> n = normal(T);  // the reference's normal() bit for bit, rounded once to float below
  s.n[0] = (float)n.x, s.n[1] = (float)n.y, s.n[2] = (float)n.z;
  s.d = (float)dot(n, T.a);
  for (int q = 0; q < 3; ++q) {
    const v3<double> x = corner(T, q), y = corner(T, (q + 1) % 3), z = corner(T, (q + 2) % 3);
    v3<double> mm = cross(sub(y, x), n);
    if (dot(mm, sub(z, x)) > 0.0) mm = v3<double>{-mm.x, -mm.y, -mm.z};
    const double ml = vlen(mm);
    if (!(ml > 0.0)) continue;
    mm = dvs(mm, ml);
    s.m[3 * q] = (float)mm.x, s.m[3 * q + 1] = (float)mm.y, s.m[3 * q + 2] = (float)mm.z;
    s.o[q] = (float)dot(mm, x);
  }
  return s;
}
D2_HD bool soup_point_keep(const float* p0, const float* p1, const SoupTri& t, float thr, float margin) {
  if (t.n[0] == 0.f && t.n[1] == 0.f && t.n[2] == 0.f) return true;
  const float h0 = f3dot(t.n, p0) - t.d, h1 = f3dot(t.n, p1) - t.d;
  const float lim = thr + margin;
  if ((h0 > lim && h1 > lim) || (h0 < -lim && h1 < -lim)) return false;
  const float lat = margin + t.eo;
  for (int q = 0; q < 3; ++q)
    if (f3dot(t.m + 3 * q, p0) - t.o[q] > lat && f3dot(t.m + 3 * q, p1) - t.o[q] > lat) return false;
  return true;
}

// The surviving sub-tests of one soup triangle pair (ta of body a, tb of b,
// body coordinates) as a 15-bit mask in triTriCCDLinear's slot order
// (continuous_comparison.h:680-779): slot 5k is corner k of ta against tb,
// 5k+1 corner k of tb against ta, 5k+2+l edge k of ta against edge l of tb.
// Poses: Ta0/Ta1, Tb0/Tb1 (start / end), Mab = inverse(Tb) Ta and Mba =
// inverse(Ta) Tb at both ends (the frame changes of the point sub-tests).
// soup_cull_mask composes the three parts below; the kernel computes the
// same parts, with the per-edge culls packed over a wavefront.

// The corners of a triangle rounded to float, corner by corner.
D2_HD void corners_f(const tri<double>& T, float* c9) {
  const v3<double> v[3] = {T.a, T.b, T.c};
  for (int k = 0; k < 3; ++k) c9[3 * k] = (float)v[k].x, c9[3 * k + 1] = (float)v[k].y, c9[3 * k + 2] = (float)v[k].z;
}
// The triangles' corners (ca, cb: rounded to float, as corners_f gives them)
// in world coordinates relative to b's start position at both ends (w[0..2]
// a at t = 0, w[3..5] a at t = 1, w[6..8] b at t = 0, w[9..11] b at t = 1);
// returns the magnitude of the transforms' inputs (the edge culls' extra).
D2_HD float soup_world(const double* Ta0, const double* Ta1, const double* Tb0, const double* Tb1, const float* ca,
                       const float* cb, float (*w)[3]) {
  const M34f A0 = to_m34f(Ta0, Tb0), A1 = to_m34f(Ta1, Tb0);
  const M34f B0 = to_m34f(Tb0, Tb0), B1 = to_m34f(Tb1, Tb0);
  float extra = fmaxf(fmaxf(m34_mag(A0), m34_mag(A1)), m34_mag(B1));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    xf_f(A0, ca + 3 * k, w[k]), xf_f(A1, ca + 3 * k, w[3 + k]);
    xf_f(B0, cb + 3 * k, w[6 + k]), xf_f(B1, cb + 3 * k, w[9 + k]);
    extra = fmaxf(extra, fmaxf(amax3(ca + 3 * k), amax3(cb + 3 * k)));
  }
  return extra;
}
// The point-triangle slots: a's corners in b's frame against tb (its planes
// sb, soup_tri), b's corners in a's frame against ta (sa).
D2_HD unsigned soup_point_mask(const double* Mab0, const double* Mab1, const double* Mba0, const double* Mba1,
                               const float* ca, const float* cb, const SoupTri& sa, const SoupTri& sb, float thr) {
  unsigned mask = 0u;
  float pa0[3][3], pa1[3][3], pb0[3][3], pb1[3][3];
  const M34f F0 = to_m34f(Mab0, nullptr), F1 = to_m34f(Mab1, nullptr);
  const M34f G0 = to_m34f(Mba0, nullptr), G1 = to_m34f(Mba1, nullptr);
  float mag = fmaxf(fmaxf(m34_mag(F0), m34_mag(F1)), fmaxf(m34_mag(G0), m34_mag(G1)));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    xf_f(F0, ca + 3 * k, pa0[k]), xf_f(F1, ca + 3 * k, pa1[k]);
    xf_f(G0, cb + 3 * k, pb0[k]), xf_f(G1, cb + 3 * k, pb1[k]);
    mag = fmaxf(mag, fmaxf(fmaxf(amax3(pa0[k]), amax3(pa1[k])), fmaxf(amax3(pb0[k]), amax3(pb1[k]))));
    mag = fmaxf(mag, fmaxf(amax3(ca + 3 * k), amax3(cb + 3 * k)));
  }
  const float margin = 1e-6f + 64.f * 5.9604645e-08f * mag;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (soup_point_keep(pa0[k], pa1[k], sb, thr, margin)) mask |= 1u << (5 * k);
    if (soup_point_keep(pb0[k], pb1[k], sa, thr, margin)) mask |= 1u << (5 * k + 1);
  }
  return mask;
}
// Edge pair (k, l)'s slot survives cull_edge_edge (world corners w and their
// inputs' magnitude extra as soup_world gives them).
D2_HD bool soup_edge_keep(const float (*w)[3], int k, int l, float thr, float extra) {
  const int k1 = k == 2 ? 0 : k + 1, l1 = l == 2 ? 0 : l + 1;
  return !cull_edge_edge(w[k], w[k1], w[3 + k], w[3 + k1], w[6 + l], w[6 + l1], w[9 + l], w[9 + l1], thr, extra);
}
D2_HD unsigned soup_cull_mask(const double* Ta0, const double* Ta1, const double* Tb0, const double* Tb1,
                              const double* Mab0, const double* Mab1, const double* Mba0, const double* Mba1,
                              const tri<double>& ta, const tri<double>& tb, float thr) {
  float ca[9], cb[9];
  corners_f(ta, ca), corners_f(tb, cb);
  unsigned mask = soup_point_mask(Mab0, Mab1, Mba0, Mba1, ca, cb, soup_tri(ta), soup_tri(tb), thr);
  float w[12][3];
  const float extra = soup_world(Ta0, Ta1, Tb0, Tb1, ca, cb, w);
  if (cull_tri_edges(w, w + 3, w + 6, w + 9, thr, extra)) return mask;  // all nine edge pairs culled at once
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l)
      if (soup_edge_keep(w, k, l, thr, extra)) mask |= 1u << (5 * k + 2 + l);
  return mask;
}

// ---------------------------------------------- discrete triangle-pair cull
// Classifies one triangle pair of the discrete phase in FP32:
//   0 = certain miss, 1 = needs the FP64 test, 2 = certain hit.
// Lower bound: the reference's distance is the length of a pair of points on
// the two triangles, except that a point projected onto a plane is accepted
// as "inside" by float barycentrics (triangle.h:92-100), which can accept a
// point up to eps_out ~ 16*eps_f*kappa*L outside (kappa = d00*d11/denom of the
// barycentric system, L the longest edge).  So dist_ref >= true distance -
// eps_out, and the true distance is >= the projection gap of the two vertex
// sets on any axis (here both normals and the centroid direction).
// Upper bound: every edge-edge candidate (comparison_dist.h:80-140) returns
// |P - Q| with Q the closest point of one segment to a point X of the other
// and P the closest point to Q, so it is <= max over the four endpoint pairs;
// dist_ref <= the smallest such bound over the 9 edge pairs.
// margin = 64*eps_f*kappa*L (4x the barycentric bound) + 16*eps_f*|coords|
// (FP32 rounding of the relative coordinates and projections) + 1e-6 m.
// Ill-conditioned barycentric systems (denominator <= 1e-4 d00 d11) get no
// cull.
//
// The per-triangle terms are computed once (cull_prep): vertices in the
// frame of the pair (b's position as origin), normal, kappa, longest squared
// edge, own-normal extent, centroid, largest coordinate (of the results and
// of the transform's inputs), and whether the barycentric system is well
// conditioned.
struct alignas(16) CullTri {
  float c3[3], kap, L2, crd, ok;
  float v[9], n[3], nn, pmin, pmax;
  float pad[4];  // 112 bytes: 28 dwords, so 8 consecutive records start in 8 distinct 4-bank groups of LDS
  double eps;    // the triangle's eps from its bucket (read with the triangle)
};
// From the triangle's corners rounded to float (body coordinates, corner
// by corner: a.x a.y a.z b.x ...), as the float copy of the buckets holds them.
D2_HD void cull_prep_f(const float* rot, const float* tr, const float* c9, CullTri& o) {
  // the magnitudes the transform's rounding scales with: inputs and results
  float crd = fmaxf(fabsf(tr[0]), fmaxf(fabsf(tr[1]), fabsf(tr[2])));
  for (int k = 0; k < 9; ++k) crd = fmaxf(crd, fabsf(c9[k]));
  for (int v = 0; v < 3; ++v) {
    const float x = c9[3 * v], y = c9[3 * v + 1], z = c9[3 * v + 2];
    for (int r = 0; r < 3; ++r) {
      o.v[3 * v + r] = fmaf(rot[3 * r + 2], z, fmaf(rot[3 * r + 1], y, fmaf(rot[3 * r], x, tr[r])));
      crd = fmaxf(crd, fabsf(o.v[3 * v + r]));
    }
  }
  float e0[3], e1[3], e2[3];
  for (int r = 0; r < 3; ++r)
    e0[r] = o.v[3 + r] - o.v[r], e1[r] = o.v[6 + r] - o.v[r], e2[r] = o.v[6 + r] - o.v[3 + r];
  o.n[0] = e0[1] * e1[2] - e0[2] * e1[1], o.n[1] = e0[2] * e1[0] - e0[0] * e1[2], o.n[2] = e0[0] * e1[1] - e0[1] * e1[0];
  const float d00 = f3dot(e0, e0), d11 = f3dot(e1, e1), d01 = f3dot(e0, e1);
  const float den = d00 * d11 - d01 * d01;
  o.ok = (den > 1e-4f * d00 * d11) ? 1.f : 0.f;
  o.kap = o.ok != 0.f ? d00 * d11 / den : 0.f;
  o.L2 = fmaxf(fmaxf(d00, d11), f3dot(e2, e2));
  o.nn = f3dot(o.n, o.n);
  const float p0 = f3dot(o.v, o.n), p1 = f3dot(o.v + 3, o.n), p2 = f3dot(o.v + 6, o.n);
  o.pmin = fminf(p0, fminf(p1, p2)), o.pmax = fmaxf(p0, fmaxf(p1, p2));
  for (int r = 0; r < 3; ++r) o.c3[r] = ((o.v[r] + o.v[3 + r]) + o.v[6 + r]) * (1.f / 3.f);
  o.crd = crd;
}
D2_HD void cull_prep(const float* rot, const float* tr, const tri<double>& t, CullTri& o) {
  const float c9[9] = {(float)t.a.x, (float)t.a.y, (float)t.a.z, (float)t.b.x, (float)t.b.y,
                       (float)t.b.z, (float)t.c.x, (float)t.c.y, (float)t.c.z};
  cull_prep_f(rot, tr, c9, o);
}
D2_HD float cull_margin(const CullTri& A, const CullTri& B, float thr) {
  const float EPSF = 5.9604645e-08f;
  const float kappa = fmaxf(A.kap, B.kap);
  const float L = sqrtf(fmaxf(A.L2, B.L2));
  return 1e-6f + 64.f * EPSF * kappa * L + 16.f * EPSF * fmaxf(A.crd, B.crd) + EPSF * thr;
}
// 0 certain miss, 1 needs the FP64 test, 2 certain hit.  The lower bound
// compares g^2 > T^2 |n|^2 per axis instead of g / |n| > T, with T = (thr +
// margin)(1 + 2^-20) covering the few roundings of the squared form.  The upper bound
// (certain hit) is only computed when want_hit: a certain hit only matters
// for node pairs with an inner node.
D2_HD int cull_pair(const CullTri& A, const CullTri& B, float thr, bool want_hit = true) {
  if (A.ok == 0.f || B.ok == 0.f) return 1;
  const float margin = cull_margin(A, B, thr);
  const float T = (thr + margin) * (1.f + 9.5367432e-07f);
  const float T2 = T * T;
  bool sep = false;
  {  // axis nA
    const float b0 = f3dot(B.v, A.n), b1 = f3dot(B.v + 3, A.n), b2 = f3dot(B.v + 6, A.n);
    const float g = fmaxf(fminf(b0, fminf(b1, b2)) - A.pmax, A.pmin - fmaxf(b0, fmaxf(b1, b2)));
    sep = sep || (A.nn > 0.f && g > 0.f && g * g > T2 * A.nn);
  }
  {  // axis nB
    const float a0 = f3dot(A.v, B.n), a1 = f3dot(A.v + 3, B.n), a2 = f3dot(A.v + 6, B.n);
    const float g = fmaxf(B.pmin - fmaxf(a0, fmaxf(a1, a2)), fminf(a0, fminf(a1, a2)) - B.pmax);
    sep = sep || (B.nn > 0.f && g > 0.f && g * g > T2 * B.nn);
  }
  {  // centroid direction
    const float cd[3] = {B.c3[0] - A.c3[0], B.c3[1] - A.c3[1], B.c3[2] - A.c3[2]};
    const float nn = f3dot(cd, cd);
    const float a0 = f3dot(A.v, cd), a1 = f3dot(A.v + 3, cd), a2 = f3dot(A.v + 6, cd);
    const float b0 = f3dot(B.v, cd), b1 = f3dot(B.v + 3, cd), b2 = f3dot(B.v + 6, cd);
    const float g = fmaxf(fminf(b0, fminf(b1, b2)) - fmaxf(a0, fmaxf(a1, a2)),
                          fminf(a0, fminf(a1, a2)) - fmaxf(b0, fmaxf(b1, b2)));
    sep = sep || (nn > 0.f && g > 0.f && g * g > T2 * nn);
  }
  if (sep) return 0;
  if (!want_hit) return 1;
  float ub2 = 3.4e38f;
  float dd[3][3];
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) {
      const float t[3] = {A.v[3 * x] - B.v[3 * y], A.v[3 * x + 1] - B.v[3 * y + 1], A.v[3 * x + 2] - B.v[3 * y + 2]};
      dd[x][y] = f3dot(t, t);
    }
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) {
      const int x1 = x == 2 ? 0 : x + 1, y1 = y == 2 ? 0 : y + 1;
      ub2 = fminf(ub2, fmaxf(fmaxf(dd[x][y], dd[x][y1]), fmaxf(dd[x1][y], dd[x1][y1])));
    }
  const float lim = thr - margin;
  if (lim > 0.f && ub2 < lim * lim) return 2;
  return 1;
}

}  // namespace d2
