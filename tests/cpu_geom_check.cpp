// CPU build of the device geometry (delta2_amd/csrc/d2_geom.hpp is
// __host__ __device__) checked bit-for-bit against the oracle on random
// inputs.  Built and run by tests/test_geom_cpu.py; prints one line per
// primitive: "<name> <cases> <mismatches>".
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#define D2_HD inline
#include <algorithm>
using std::sqrt;
using std::fabs;
using std::sin;
using std::cos;
#include "../delta2_amd/csrc/d2_geom.hpp"
#include "../delta2_amd/csrc/d2_cull.hpp"
#include "../oracle/d2o_tree.hpp"

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

static bool same_bits(double a, double b) { return std::memcmp(&a, &b, 8) == 0 || (a != a && b != b); }
static bool same_v(d2::v3<double> a, d2o::V3<double> b) { return same_bits(a.x, b.x) && same_bits(a.y, b.y) && same_bits(a.z, b.z); }

static void rand_pose(double* m, double tscale, double rscale) {
  d2o::V3<double> e(U(-rscale, rscale), U(-rscale, rscale), U(-rscale, rscale)), t(U(-tscale, tscale), U(-tscale, tscale), U(-tscale, tscale));
  d2o::M4<double> M = d2o::transformationMatrixEuler(e, t);
  std::memcpy(m, M.m, sizeof(M.m));
}
static d2o::Tri<double> rand_tri(double s, double off) {
  return d2o::Tri<double>(d2o::V3<double>(U(-s, s) + off, U(-s, s), U(-s, s)), d2o::V3<double>(U(-s, s) + off, U(-s, s), U(-s, s)),
                          d2o::V3<double>(U(-s, s) + off, U(-s, s), U(-s, s)));
}
static d2::tri<double> cvt(const d2o::Tri<double>& t) {
  return {{t.A.x, t.A.y, t.A.z}, {t.B.x, t.B.y, t.B.z}, {t.C.x, t.C.y, t.C.z}};
}
static d2o::M4<double> m4(const double* m) {
  d2o::M4<double> M;
  std::memcpy(M.m, m, sizeof(M.m));
  return M;
}

int main() {
  const int N = 20000;
  // tri-tri closest points
  int bad = 0;
  for (int k = 0; k < N; ++k) {
    d2o::Tri<double> a = rand_tri(1.0, 0.0), b = rand_tri(1.0, U(-1.5, 1.5));
    d2o::V3<double> P, Q;
    double d0 = d2o::comparisonDist(a, b, P, Q);
    d2::v3<double> p, q;
    double d1 = d2::tri_tri_dist(cvt(a), cvt(b), p, q);
    if (!same_bits(d0, d1) || !same_v(p, P) || !same_v(q, Q)) ++bad;
  }
  printf("tri_tri_dist %d %d\n", N, bad);
  // point-triangle CCD with precomputed frame changes
  bad = 0;
  for (int k = 0; k < N; ++k) {
    double Ta0[16], Ta1[16], Tb0[16], Tb1[16], inv[16], M0[16], M1[16];
    rand_pose(Ta0, 1.0, 3.0);
    rand_pose(Ta1, 1.0, 3.0);
    rand_pose(Tb0, 1.0, 3.0);
    rand_pose(Tb1, 1.0, 3.0);
    d2::inv44(Tb0, inv);
    d2::mul44(inv, Ta0, M0);
    d2::inv44(Tb1, inv);
    d2::mul44(inv, Ta1, M1);
    d2o::Tri<double> tb = rand_tri(1.0, 0.0);
    d2o::V3<double> pt(U(-1, 1), U(-1, 1), U(-1, 1));
    d2o::V3<double> P, Q;
    double toc0;
    double d0 = d2o::pointTriCCDLinear(pt, m4(Ta0), m4(Ta1), tb, m4(Tb0), m4(Tb1), P, Q, toc0);
    d2::v3<double> p{0, 0, 0}, q{0, 0, 0};
    double toc1;
    double d1 = d2::point_tri(d2::v3<double>{pt.x, pt.y, pt.z}, M0, M1, cvt(tb), Tb0, Tb1, p, q, toc1);
    bool ok = same_bits(d0, d1) && same_bits(toc0, toc1);
    if (ok && d0 != std::numeric_limits<double>::infinity()) ok = same_v(p, P) && same_v(q, Q);
    if (!ok) ++bad;
  }
  printf("point_tri %d %d\n", N, bad);
  // edge-edge CCD
  bad = 0;
  int bisect = 0;
  for (int k = 0; k < N; ++k) {
    d2o::V3<double> v[8];
    for (int i = 0; i < 8; ++i) v[i] = d2o::V3<double>(U(-1, 1), U(-1, 1), U(-1, 1));
    d2o::Edge<double> as(v[0], v[1]), ae(v[2], v[3]), bs(v[4], v[5]), be(v[6], v[7]);
    d2o::V3<double> P, Q;
    double toc0 = -1;
    double d0 = d2o::edgeEdgeCCD(as, ae, bs, be, P, Q, toc0);
    auto c = [](const d2o::V3<double>& x) { return d2::v3<double>{x.x, x.y, x.z}; };
    d2::v3<double> p, q;
    double toc1 = -1;
    double d1 = d2::edge_edge(c(v[0]), c(v[1]), c(v[2]), c(v[3]), c(v[4]), c(v[5]), c(v[6]), c(v[7]), p, q, toc1);
    if (!same_bits(d0, d1) || !same_bits(toc0, toc1) || !same_v(p, P) || !same_v(q, Q)) ++bad;
    d2o::V3<double> ns = d2o::cross(as.B - as.A, bs.A - as.A), ne = d2o::cross(ae.B - ae.A, be.A - ae.A);
    if (d2o::dot(bs.B - as.A, ns) * d2o::dot(be.B - ae.A, ne) < 0) ++bisect;
  }
  printf("edge_edge %d %d (bisection cases %d)\n", N, bad, bisect);
  // state extrapolation (host libm on both sides here)
  bad = 0;
  for (int k = 0; k < N; ++k) {
    double r[14];
    double q[4] = {U(-1, 1), U(-1, 1), U(-1, 1), U(-1, 1)};
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i) r[i] = q[i] / n;
    for (int i = 4; i < 13; ++i) r[i] = U(-3, 3);
    r[13] = 0;
    double I[9] = {U(0.5, 2), 0, 0, 0, U(0.5, 2), 0, 0, 0, U(0.5, 2)};
    I[1] = I[3] = U(-0.1, 0.1);
    double dt = U(0, 0.1);
    d2::body_state f = d2::extrapolate(d2::load_state(r), dt, I);
    d2o::State s;
    s.rot = {r[0], r[1], r[2], r[3]};
    s.trans = d2o::V3<double>(r[4], r[5], r[6]);
    s.vel = d2o::V3<double>(r[7], r[8], r[9]);
    s.L = d2o::V3<double>(r[10], r[11], r[12]);
    d2o::M3<double> Im;
    std::memcpy(Im.m, I, sizeof(I));
    d2o::State g = s.extrapolate(dt, Im);
    double pm[16];
    d2::pose(f.q, f.t, pm);
    d2o::M4<double> gm = g.transformation();
    bool ok = true;
    for (int i = 0; i < 16; ++i) ok = ok && same_bits(pm[i], gm.m[i]);
    if (!ok) ++bad;
  }
  printf("extrapolate %d %d\n", N, bad);

  // the filter's distance test without the square root equals vlen(d) < search,
  // also when search sits within a few ulps of |d|
  bad = 0;
  int cases = 0;
  for (int k = 0; k < N; ++k) {
    d2::v3<double> d{U(-1, 1), U(-1, 1), U(-1, 1)};
    const double sc = std::pow(10.0, U(-6, 3));
    d = d2::scl(d, sc);
    const double len = d2::vlen(d);
    for (int j = -6; j <= 6; ++j) {
      const double search = j == 6 ? U(0, 2) * len : std::nextafter(len, j < 0 ? 0.0 : 1e300) * (1.0 + j * 1e-16);
      ++cases;
      if (d2::near_mid(d, search) != (d2::vlen(d) < search)) ++bad;
    }
  }
  printf("near_mid %d %d\n", cases, bad);

  // nodes_apart, boxes_apart_faces, spheres_apart and spheres_apart_f32
  // never prune a node pair
  // that holds a leaf contact: two random
  // clusters of 1-8 triangles (some thin) with random leaf eps, their float
  // boxes and slack as d2_trees_upload computes them, poses placing them
  // around the contact distance; a prune while some triangle pair's oracle
  // distance is below search + eps_i + eps_j is a violation.
  {
    int cases = 0, violations = 0, pruned = 0, pruned_by[4] = {0, 0, 0, 0};
    const double search = (double)(float)0.02;
    for (int k = 0; k < 40000; ++k) {
      const double sz = std::pow(10.0, U(-2.0, 0.5));
      std::vector<d2o::Tri<double>> tri[2];
      std::vector<double> eps[2];
      float box[2][12];
      for (int s = 0; s < 2; ++s) {
        const int n = 1 + (int)(rng() % 8);
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, E = 0.0;
        for (int i = 0; i < n; ++i) {
          d2o::V3<double> a(U(-sz, sz), U(-sz, sz), U(-sz, sz)), b(U(-sz, sz), U(-sz, sz), U(-sz, sz));
          d2o::V3<double> c = (rng() % 4 == 0) ? a + (b - a) * U(0.2, 0.8) + d2o::V3<double>(U(-1e-4, 1e-4), U(-1e-4, 1e-4), U(-1e-4, 1e-4)) * sz
                                               : d2o::V3<double>(U(-sz, sz), U(-sz, sz), U(-sz, sz));
          tri[s].push_back(d2o::Tri<double>(a, b, c));
          const double e = (rng() % 2) ? 0.0 : U(0.0, 0.01);
          eps[s].push_back(e);
          const double va[3] = {a.x, a.y, a.z}, vb[3] = {b.x, b.y, b.z}, vc[3] = {c.x, c.y, c.z};
          E = std::max(E, d2::leaf_slack(va, vb, vc, e));
          for (const double* v : {va, vb, vc})
            for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], v[d]), hi[d] = std::max(hi[d], v[d]);
        }
        for (int d = 0; d < 3; ++d) {
          float fl = (float)lo[d], fh = (float)hi[d];
          if ((double)fl > lo[d]) fl = std::nextafter(fl, -INFINITY);
          if ((double)fh < hi[d]) fh = std::nextafter(fh, INFINITY);
          box[s][d] = fl, box[s][3 + d] = fh;
        }
        float fe = (float)E;
        if ((double)fe < E) fe = std::nextafter(fe, INFINITY);
        box[s][6] = fe;
        // the sphere as d2_trees_upload forms it: the vertices' centroid
        // (rounded to float) and the largest distance from it, rounded up
        double cen[3] = {0.0, 0.0, 0.0};
        for (const auto& t : tri[s])
          for (const auto* v : {&t.A, &t.B, &t.C}) cen[0] += v->x, cen[1] += v->y, cen[2] += v->z;
        float fc[3];
        for (int d = 0; d < 3; ++d) fc[d] = (float)(cen[d] / (3.0 * tri[s].size()));
        double rad = 0.0;
        for (const auto& t : tri[s])
          for (const auto* v : {&t.A, &t.B, &t.C}) {
            const double x = v->x - (double)fc[0], y = v->y - (double)fc[1], z = v->z - (double)fc[2];
            rad = std::max(rad, std::sqrt(x * x + y * y + z * z));
          }
        float fr = (float)(rad * (1.0 + 1e-12));
        if ((double)fr < rad) fr = std::nextafter(fr, INFINITY);
        box[s][7] = fr, box[s][8] = fc[0], box[s][9] = fc[1], box[s][10] = fc[2], box[s][11] = 0.0f;
      }
      double Ta[16], Tb[16];
      rand_pose(Ta, 5.0, 3.2);
      // b near a: mostly just beyond the two boxes' bounding radii (gaps from
      // -2 to +6 cm, around the contact distance), else anywhere within a few
      // cluster sizes
      double ra = 0.0, rb = 0.0;
      for (int d = 0; d < 3; ++d) {
        const double ha = std::max(std::fabs(box[0][d]), std::fabs(box[0][3 + d]));
        const double hb = std::max(std::fabs(box[1][d]), std::fabs(box[1][3 + d]));
        ra += ha * ha, rb += hb * hb;
      }
      d2o::V3<double> dir(U(-1, 1), U(-1, 1), U(-1, 1));
      dir = dir / d2o::norm(dir);
      const double dist = (rng() % 4) ? std::sqrt(ra) + std::sqrt(rb) + U(-0.02, 0.06) : U(0.0, 3.0) * sz;
      d2o::M4<double> Mb = d2o::transformationMatrixEuler(d2o::V3<double>(U(-3.2, 3.2), U(-3.2, 3.2), U(-3.2, 3.2)),
                                                          d2o::V3<double>(Ta[3] + dir.x * dist, Ta[7] + dir.y * dist,
                                                                          Ta[11] + dir.z * dist));
      std::memcpy(Tb, Mb.m, sizeof(Tb));
      ++cases;
      const bool apart[4] = {d2::nodes_apart(Ta, Tb, box[0], box[1], search),
                             d2::spheres_apart(Ta, Tb, box[0], box[1], search),
                             d2::spheres_apart_f32(Ta, Tb, box[0], box[1], search),
                             d2::boxes_apart_faces(Ta, Tb, box[0], box[1], search)};
      if (!apart[0] && !apart[1] && !apart[2] && !apart[3]) continue;
      bool contact = false;
      for (size_t i = 0; i < tri[0].size() && !contact; ++i)
        for (size_t j = 0; j < tri[1].size() && !contact; ++j) {
          d2o::V3<double> P, Q;
          const double d = d2o::comparisonDist(tri[0][i].transformed(m4(Ta)), tri[1][j].transformed(m4(Tb)), P, Q);
          contact = d < (search + eps[0][i]) + eps[1][j];
        }
      for (int v = 0; v < 4; ++v) pruned_by[v] += apart[v], violations += apart[v] && contact;
      pruned += 1;
    }
    printf("nodes_apart %d %d (pruned %d: boxes %d, spheres %d, spheres f32 %d, face axes %d)\n", cases, violations,
           pruned, pruned_by[0], pruned_by[1], pruned_by[2], pruned_by[3]);
  }
  // FP32 culls of the CCD soup sub-tests (soup_cull_mask): a culled sub-test
  // never reports a distance below the search distance.  Triangle pairs of
  // two nearby bodies (world offsets up to 40 m), small motions over the
  // step, search distances 2-50 mm.
  {
    long long subtests = 0, culled = 0, hits = 0, viol = 0, tri_cut = 0, tri_viol = 0;
    for (int k = 0; k < N; ++k) {
      double Ta0[16], Ta1[16], Tb0[16], Tb1[16], rel[16], da[16], db[16], inv[16];
      double Mab0[16], Mab1[16], Mba0[16], Mba1[16];
      rand_pose(Ta0, (k % 4 == 0) ? 40.0 : 2.0, 3.2);
      rand_pose(rel, (k % 2) ? 0.6 : 0.15, 3.2);  // half of the pairs nearly touching
      // a third of the pairs with one triangle 10-30 m from its body's origin
      // (a funnel wall) and the other body's origin next to it: the frame
      // changes then cancel large terms
      const int far = k % 3;  // 1: a's triangle far out, 2: b's
      const double off = far ? U(10.0, 30.0) : 0.0;
      if (far == 1) {
        rel[3] = off + U(-0.3, 0.3), rel[7] = U(-0.3, 0.3), rel[11] = U(-0.3, 0.3);
      } else if (far == 2) {
        for (int r = 0; r < 3; ++r) rel[4 * r + 3] = -rel[4 * r] * off + U(-0.3, 0.3);
      }
      d2::mul44(Ta0, rel, Tb0);
      rand_pose(da, 0.15, 0.2);
      rand_pose(db, 0.15, 0.2);
      d2::mul44(Ta0, da, Ta1);
      d2::mul44(Tb0, db, Tb1);
      d2::inv44(Tb0, inv), d2::mul44(inv, Ta0, Mab0);
      d2::inv44(Tb1, inv), d2::mul44(inv, Ta1, Mab1);
      d2::inv44(Ta0, inv), d2::mul44(inv, Tb0, Mba0);
      d2::inv44(Ta1, inv), d2::mul44(inv, Tb1, Mba1);
      const double sa = U(0.05, 0.5), sb = U(0.05, 0.5);
      d2o::Tri<double> oa = rand_tri(sa, U(-0.3, 0.3) + (far == 1 ? off : 0.0));
      d2o::Tri<double> ob = rand_tri(sb, U(-0.3, 0.3) + (far == 2 ? off : 0.0));
      if (k % 8 == 0) ob.C = ob.A + (ob.B - ob.A) * U(0.0, 1.0);  // degenerate (collinear) triangles too
      if (k % 8 == 4) {  // and thin ones (kappa up to ~1e4), whose float barycentrics stray the most
        const d2o::V3<double> e = ob.B - ob.A;
        const d2o::V3<double> q = d2o::cross(e, d2o::V3<double>(U(-1, 1), U(-1, 1), U(-1, 1)));
        ob.C = ob.A + e * U(0.1, 0.9) + q * (std::pow(10.0, U(-2.5, -1.0)) / std::max(1e-30, d2o::norm(q)) * d2o::norm(e));
      }
      const d2::tri<double> ta = cvt(oa), tb = cvt(ob);
      const float thr = (float)U(0.002, 0.05);
      const double thr_d = (double)thr;
      const unsigned mask = d2::soup_cull_mask(Ta0, Ta1, Tb0, Tb1, Mab0, Mab1, Mba0, Mba1, ta, tb, thr);
      // the triangle-level edge cull (cull_tri_edges) on the same float world points
      float w[12][3], ca9[9], cb9[9];
      d2::corners_f(ta, ca9), d2::corners_f(tb, cb9);
      const float extra = d2::soup_world(Ta0, Ta1, Tb0, Tb1, ca9, cb9, w);
      const bool tcut = d2::cull_tri_edges(w, w + 3, w + 6, w + 9, thr, extra);
      tri_cut += tcut;
      for (int s = 0; s < 15; ++s) {
        const int kk = s / 5, r = s % 5;
        d2::v3<double> P, Q;
        double toc = 1.0, d;
        if (r == 0) {
          d = d2::point_tri(d2::corner(ta, kk), Mab0, Mab1, tb, Tb0, Tb1, P, Q, toc);
        } else if (r == 1) {
          d = d2::point_tri(d2::corner(tb, kk), Mba0, Mba1, ta, Ta0, Ta1, P, Q, toc);
        } else {
          const int l = r - 2;
          const d2::v3<double> a0 = d2::corner(ta, kk), a1 = d2::corner(ta, (kk + 1) % 3);
          const d2::v3<double> b0 = d2::corner(tb, l), b1 = d2::corner(tb, (l + 1) % 3);
          d = d2::edge_edge(d2::xform(Ta0, a0), d2::xform(Ta0, a1), d2::xform(Ta1, a0), d2::xform(Ta1, a1),
                            d2::xform(Tb0, b0), d2::xform(Tb0, b1), d2::xform(Tb1, b0), d2::xform(Tb1, b1), P, Q, toc);
        }
        const bool hit = d < thr_d, keep = (mask >> s) & 1u;
        ++subtests, hits += hit, culled += !keep, viol += hit && !keep;
        tri_viol += r >= 2 && tcut && hit;
      }
    }
    printf("soup_culls %lld %lld (culled %lld, hits %lld)\n", subtests, viol, culled, hits);
    printf("tri_edges %d %lld (culled %lld)\n", N, tri_viol, tri_cut);
  }
  // FP32 classification of discrete triangle pairs (cull_prep, cull_pair): a
  // certain miss never hits in FP64, a certain hit
  // always does.  Pairs placed at the hit threshold within a relative 1e-5
  // (moved along their closest-point direction), skinny triangles, world
  // offsets up to 40 m.
  {
    long long cases = 0, miss = 0, two = 0, hits = 0, viol = 0;
    for (int k = 0; k < 2 * N; ++k) {
      double Ta[16], Tb[16];
      rand_pose(Ta, (k % 4 == 0) ? 40.0 : 2.0, 3.2);
      double rel[16];
      rand_pose(rel, (k % 3 == 0) ? 0.01 : 0.5, 3.2);
      const double off = (k % 5 == 1) ? U(10.0, 30.0) : 0.0;  // a's triangle far from a's origin, b's origin near it
      if (off > 0.0) rel[3] += off;
      d2::mul44(Ta, rel, Tb);
      const double sz = (k % 3 == 0) ? U(0.002, 0.01) : U(0.02, 0.5);  // small ones give certain hits
      d2o::Tri<double> oa = rand_tri(sz, off), ob = rand_tri(sz, 0.0);
      if (k % 7 == 0) oa.C = oa.A + (oa.B - oa.A) * U(0.0, 1.0) + d2o::V3<double>(U(-1, 1), U(-1, 1), U(-1, 1)) * 1e-3;
      const double search = U(0.001, 0.03), ea = U(0.0, 0.01), eb = U(0.0, 0.01);
      const double thr_d = (search + ea) + eb;
      d2::tri<double> ta = cvt(oa), tb = cvt(ob);
      d2::v3<double> P, Q;
      double d = d2::tri_tri_dist(d2::xform(Ta, ta), d2::xform(Tb, tb), P, Q);
      if (k % 2 == 0 && d > 0.0) {  // move b so that the distance sits at the threshold
        const double target = thr_d * (1.0 + U(-1e-5, 1e-5));
        const d2::v3<double> dir = d2::dvs(d2::sub(Q, P), d);
        const double mv = target - d;
        for (int r = 0; r < 3; ++r) Tb[4 * r + 3] += mv * (r == 0 ? dir.x : r == 1 ? dir.y : dir.z);
        d = d2::tri_tri_dist(d2::xform(Ta, ta), d2::xform(Tb, tb), P, Q);
      }
      const bool hit = d < thr_d;
      d2::CullTri A, B;
      float rot[9], tr[3];
      for (int r = 0; r < 3; ++r) {
        for (int q = 0; q < 3; ++q) rot[3 * r + q] = (float)Ta[4 * r + q];
        tr[r] = (float)(Ta[4 * r + 3] - Tb[4 * r + 3]);
      }
      d2::cull_prep(rot, tr, ta, A);
      for (int r = 0; r < 3; ++r) {
        for (int q = 0; q < 3; ++q) rot[3 * r + q] = (float)Tb[4 * r + q];
        tr[r] = 0.f;
      }
      d2::cull_prep(rot, tr, tb, B);
      const float thr = (float)thr_d;
      const int cls = d2::cull_pair(A, B, thr, true);
      ++cases, hits += hit, miss += cls == 0, two += cls == 2;
      viol += (cls == 0 && hit) || (cls == 2 && !hit);
    }
    printf("disc_culls %lld %lld (miss %lld, hit %lld, fp64 hits %lld)\n", cases, viol, miss, two, hits);
  }
  // The point-triangle cull's lateral bound at the edges of thin triangles:
  // a point just outside an edge, near the plane, standing still in the
  // triangle's frame; the reference's float barycentrics may accept it as
  // inside (then it reports the plane distance), so a cull there must hold
  // the float acceptance bound.
  {
    long long cases = 0, culled = 0, hits = 0, viol = 0;
    double I4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    for (int k = 0; k < N; ++k) {
      d2o::Tri<double> t = rand_tri(U(0.05, 1.0), U(-2.0, 2.0));
      if (k % 2) {  // thin: kappa up to ~1e6
        const d2o::V3<double> e = t.B - t.A;
        const d2o::V3<double> q = d2o::cross(e, d2o::V3<double>(U(-1, 1), U(-1, 1), U(-1, 1)));
        t.C = t.A + e * U(0.1, 0.9) + q * (std::pow(10.0, U(-3.0, -1.0)) / std::max(1e-30, d2o::norm(q)) * d2o::norm(e));
      }
      const d2o::V3<double> n = t.normal();
      const int q = (int)(rng() % 3);
      const d2o::V3<double> x = t[q], y = t[(q + 1) % 3], z = t[(q + 2) % 3];
      d2o::V3<double> out = d2o::cross(y - x, n);
      if (d2o::dot(out, z - x) > 0) out = out * -1.0;
      out = out / d2o::norm(out);
      const double L = std::max(d2o::norm(y - x), std::max(d2o::norm(z - x), d2o::norm(z - y)));
      const float thr = (float)U(0.002, 0.05);
      const d2o::V3<double> P = x + (y - x) * U(0.05, 0.95) + out * (L * std::pow(10.0, U(-7.0, -2.0))) + n * (U(-0.5, 0.5) * thr);
      const d2::tri<double> T = cvt(t);
      d2::v3<double> PP, QQ;
      double toc = 1.0;
      const double d = d2::point_tri(d2::v3<double>{P.x, P.y, P.z}, I4, I4, T, I4, I4, PP, QQ, toc);
      const float p0[3] = {(float)P.x, (float)P.y, (float)P.z};
      float mag = d2::amax3(p0);
      for (int v = 0; v < 3; ++v) mag = std::max(mag, d2::v3_mag(d2::corner(T, v)));
      const float margin = 1e-6f + 64.f * 5.9604645e-08f * mag;
      const bool keep = d2::soup_point_keep(p0, p0, d2::soup_tri(T), thr, margin);
      const bool hit = d < (double)thr;
      ++cases, culled += !keep, hits += hit, viol += hit && !keep;
    }
    printf("edge_points %lld %lld (culled %lld, hits %lld)\n", cases, viol, culled, hits);
  }
  return 0;
}
