// d2_tree.cpp -- host construction of Delta2 surrogate trees in the flat
// node/bucket format of include/d2hip.h.
//
// The tree is the reference's 8-ary k-means tree (model/surrogate_tree.cpp:39-112,
// model/construct_surrogate_tree.h:8-185, model/fit_surrogate_triangle.h:246-298),
// including its PCG32 consumption order, so a mesh built here gives the same
// nodes and buckets as the reference would.  Two things differ in *how* it
// is computed, not in what: the per-iteration point-to-surrogate distances are
// evaluated once per distinct vertex (the reference evaluates each vertex once
// per incident triangle; the values are identical), and the surrogate fit
// skips the eps it computes but never uses inside divideAndFit.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <stdexcept>
#include <vector>

#include "../../include/d2hip.h"

namespace {

struct P3 {
  double x, y, z;
};
inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 mulsv(double s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(P3 a, P3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct Tri3 {
  P3 a, b, c;
};

inline double area(const Tri3& t) {
  P3 n = cross(t.a - t.c, t.b - t.c);
  return 0.5 * std::sqrt(dot(n, n));
}

// Closest-point distance from X to triangle t (region classification of
// common/triangle.h:280-471).
double tri_point_dist(const Tri3& t, P3 X) {
  P3 AX = X - t.a, AB = t.b - t.a, AC = t.c - t.a;
  double a00 = dot(AB, AB), a01 = dot(AB, AC), a11 = dot(AC, AC);
  double b0 = -dot(AX, AB), b1 = -dot(AX, AC);
  double det = a00 * a11 - a01 * a01;
  double s = a01 * b1 - a11 * b0;
  double u = a01 * b0 - a00 * b1;
  if (s + u <= det) {
    if (s < 0.0) {
      if (u < 0.0) {
        if (b0 < 0.0) {
          u = 0.0;
          s = (-b0 >= a00) ? 1.0 : -b0 / a00;
        } else {
          s = 0.0;
          u = (b1 >= 0.0) ? 0.0 : ((-b1 >= a11) ? 1.0 : -b1 / a11);
        }
      } else {
        s = 0.0;
        u = (b1 >= 0.0) ? 0.0 : ((-b1 >= a11) ? 1.0 : -b1 / a11);
      }
    } else if (u < 0.0) {
      u = 0.0;
      s = (b0 >= 0.0) ? 0.0 : ((-b0 >= a00) ? 1.0 : -b0 / a00);
    } else {
      double inv = 1.0 / det;
      s *= inv;
      u *= inv;
    }
  } else if (s < 0.0) {
    double t0 = a01 + b0, t1 = a11 + b1;
    if (t1 > t0) {
      double num = t1 - t0, den = a00 - 2.0 * a01 + a11;
      if (num >= den) s = 1.0, u = 0.0;
      else s = num / den, u = 1.0 - s;
    } else {
      s = 0.0;
      u = (t1 <= 0.0) ? 1.0 : ((b1 >= 0.0) ? 0.0 : -b1 / a11);
    }
  } else if (u < 0.0) {
    double t0 = a01 + b1, t1 = a00 + b0;
    if (t1 > t0) {
      double num = t1 - t0, den = a00 - 2.0 * a01 + a11;
      if (num >= den) u = 1.0, s = 0.0;
      else u = num / den, s = 1.0 - u;
    } else {
      u = 0.0;
      s = (t1 <= 0.0) ? 1.0 : ((b0 >= 0.0) ? 0.0 : -b0 / a00);
    }
  } else {
    double num = a11 + b1 - a01 - b0;
    if (num <= 0.0) {
      s = 0.0, u = 1.0;
    } else {
      double den = a00 - 2.0 * a01 + a11;
      if (num >= den) s = 1.0, u = 0.0;
      else s = num / den, u = 1.0 - s;
    }
  }
  P3 d = X - ((t.a + mulsv(s, AB)) + mulsv(u, AC));
  return std::sqrt(dot(d, d));
}

struct Pcg {
  uint64_t state, inc;
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
  }
  uint32_t bounded(uint32_t bound) {
    uint32_t threshold = (uint32_t)(-bound) % bound;
    for (;;) {
      uint32_t r = next();
      if (r >= threshold) return r % bound;
    }
  }
};

struct Builder {
  const std::vector<P3>& V;
  const std::vector<int>& F;  // 3 per face
  Pcg& rng;
  std::vector<int> mark;  // scratch: vertex -> slot in the current cache
  Builder(const std::vector<P3>& v, const std::vector<int>& f, Pcg& r) : V(v), F(f), rng(r), mark(v.size(), -1) {}

  Tri3 face(int t) const { return {V[F[3 * t]], V[F[3 * t + 1]], V[F[3 * t + 2]]}; }

  // Random-restart fit of a group's surrogate (fit_surrogate_triangle.h:246-298).
  // `best` starts as the group's current surrogate: the reference starts from
  // an uninitialised triangle, whose value is undefined when every trial is
  // degenerate (possible for two triangles sharing an edge).
  Tri3 fit(const std::vector<int>& members, const Tri3& current) {
    int n = 3 * (int)members.size();
    auto pt = [&](int k) { return V[F[3 * members[k / 3] + k % 3]]; };
    Tri3 best = current;
    if (n > 3) {
      double best_area = -1;
      int attempts = (int)(std::log2(n) * 10);
      for (int i = 0; i < attempts; ++i) {
        int a = (int)rng.bounded((uint32_t)n), b = (int)rng.bounded((uint32_t)n), c = (int)rng.bounded((uint32_t)n);
        Tri3 tr{pt(a), pt(b), pt(c)};
        double ar = area(tr);
        if (ar > best_area && ar > 0.0) best = tr, best_area = ar;
      }
    } else {
      best = {pt(0), pt(1), pt(2)};
    }
    if (!(area(best) > 0)) throw std::runtime_error("No valid surrogate triangle was found");
    return best;
  }

  // k-means of construct_surrogate_tree.h:8-168.
  void divide(const std::vector<int>& ids, int branching, bool force, std::vector<Tri3>& sur,
              std::vector<std::vector<int>>& grp) {
    if (branching > (int)ids.size()) throw std::invalid_argument("There must be more triangles than branches");
    sur.assign(branching, Tri3{});
    grp.assign(branching, {});
    for (int g = 0; g < branching; ++g) sur[g] = face(g);  // seeds: first faces of the mesh
    // distinct vertices of this id set
    std::vector<int> verts;
    for (int t : ids)
      for (int k = 0; k < 3; ++k) {
        int v = F[3 * t + k];
        if (mark[v] < 0) mark[v] = (int)verts.size(), verts.push_back(v);
      }
    std::vector<int> slot(3 * ids.size());
    for (size_t i = 0; i < ids.size(); ++i)
      for (int k = 0; k < 3; ++k) slot[3 * i + k] = mark[F[3 * ids[i] + k]];
    for (int v : verts) mark[v] = -1;
    std::vector<double> dist(verts.size() * branching);
    for (int iter = 0; iter < 20; ++iter) {
      for (int g = 0; g < branching; ++g) grp[g].clear();
      for (size_t v = 0; v < verts.size(); ++v)
        for (int g = 0; g < branching; ++g) dist[v * branching + g] = tri_point_dist(sur[g], V[verts[v]]);
      for (size_t i = 0; i < ids.size(); ++i) {
        double best = std::numeric_limits<double>::infinity();
        int bg = 0;
        for (int g = 0; g < branching; ++g) {
          double da = dist[slot[3 * i] * branching + g], db = dist[slot[3 * i + 1] * branching + g],
                 dc = dist[slot[3 * i + 2] * branching + g];
          double m = da < db ? db : da;
          m = m < dc ? dc : m;
          if (m < best) bg = g, best = m;
        }
        grp[bg].push_back(ids[i]);
      }
      for (int g = 0; g < branching; ++g)
        if (!grp[g].empty()) sur[g] = fit(grp[g], sur[g]);
      if (!force) continue;
      bool improving = true, met = false;
      while (improving && !met) {
        int largest = 0;
        size_t lsize = 0;
        std::vector<int> zeros;
        for (int g = 0; g < branching; ++g) {
          if (grp[g].size() > lsize) lsize = grp[g].size(), largest = g;
          if (grp[g].empty()) zeros.push_back(g);
        }
        if (zeros.empty()) {
          met = true;
          break;
        }
        std::vector<Tri3> sub_s;
        std::vector<std::vector<int>> sub_g;
        int k = (int)std::min(grp[largest].size(), zeros.size() + 1);
        std::vector<int> members = grp[largest];
        divide(members, k, false, sub_s, sub_g);
        int nonempty = 0;
        for (int s = 0; s < (int)sub_s.size(); ++s) {
          int gi = s == 0 ? largest : zeros[s - 1];
          if (!sub_g[s].empty()) ++nonempty, grp[gi] = sub_g[s], sur[gi] = sub_s[s];
        }
        improving = nonempty > 1;
      }
      if (!met) {
        std::vector<int> zeros;
        do {
          int largest = 0;
          size_t lsize = 0;
          zeros.clear();
          for (int g = 0; g < branching; ++g) {
            if (grp[g].size() > lsize) lsize = grp[g].size(), largest = g;
            if (grp[g].empty()) zeros.push_back(g);
          }
          for (size_t z = 0; z < std::min(zeros.size(), grp[largest].size() - 1); ++z) {
            int t = grp[largest].back();
            grp[zeros[z]].push_back(t);
            grp[largest].pop_back();
            sur[zeros[z]] = face(t);
          }
          sur[largest] = fit(grp[largest], sur[largest]);
        } while (!zeros.empty());
      }
    }
  }
};

}  // namespace

extern "C" int d2_pcg_seed(uint64_t initstate, uint64_t initseq, uint64_t* rng_out) {
  if (!rng_out) return D2_ERR_ARG;
  Pcg r{0u, (initseq << 1u) | 1u};
  r.next();
  r.state += initstate;
  r.next();
  rng_out[0] = r.state;
  rng_out[1] = r.inc;
  return D2_OK;
}

extern "C" int d2_tree_build(const double* Vp, int nv, const int32_t* Fp, int nf, uint64_t* rng, int max_nodes,
                             int32_t* nodes_out, double* buckets_out, int* n_nodes_out) {
  if (!Vp || !Fp || !rng || nv <= 0 || nf <= 0 || !n_nodes_out) return D2_ERR_ARG;
  for (int i = 0; i < 3 * nf; ++i)
    if (Fp[i] < 0 || Fp[i] >= nv) return D2_ERR_ARG;
  std::vector<P3> V(nv);
  for (int i = 0; i < nv; ++i) V[i] = {Vp[3 * i], Vp[3 * i + 1], Vp[3 * i + 2]};
  std::vector<int> F(Fp, Fp + 3 * nf);
  Pcg r{rng[0], rng[1]};
  Builder B(V, F, r);
  struct Item {
    std::vector<int> ids;
    int depth, parent;
  };
  struct NodeOut {
    int32_t rec[8];
    double bucket[88];
  };
  std::vector<NodeOut> out;
  std::deque<Item> q;
  std::vector<int> all(nf);
  for (int i = 0; i < nf; ++i) all[i] = i;
  q.push_back({all, 0, -1});
  try {
    while (!q.empty()) {
      Item it = std::move(q.front());
      q.pop_front();
      NodeOut n;
      std::memset(&n, 0, sizeof(n));
      int id = (int)out.size();
      int ntris;
      std::vector<Tri3> tris;
      std::vector<double> eps, inner;
      if (it.ids.size() < 8) {
        ntris = (int)it.ids.size();
        for (int t : it.ids) tris.push_back(B.face(t));
        eps.assign(ntris, 0.0);
        inner.assign(ntris, 0.0);
        n.rec[2] = 0;
        n.rec[6] = 0;
      } else {
        std::vector<std::vector<int>> grp;
        B.divide(it.ids, 8, true, tris, grp);
        ntris = (int)tris.size();
        for (int k = 0; k < ntris; ++k) {  // calcEpss, construct_surrogate_tree.h:172-185
          double in = std::numeric_limits<double>::infinity(), ex = 0.0;
          for (int t : grp[k]) {
            Tri3 f = B.face(t);
            double a = tri_point_dist(tris[k], f.a), b = tri_point_dist(tris[k], f.b), c = tri_point_dist(tris[k], f.c);
            double mn = std::min(a, std::min(b, c)), mx = std::max(a, std::max(b, c));
            in = std::min(in, mn);
            ex = std::max(ex, mx);
          }
          inner.push_back(in);
          eps.push_back(ex);
        }
        n.rec[2] = 1;
        n.rec[6] = ntris;
        for (int k = 0; k < ntris; ++k) q.push_back({grp[k], it.depth + 1, id});
      }
      n.rec[0] = id;
      n.rec[1] = id;
      n.rec[3] = it.depth;
      n.rec[4] = it.parent;
      n.rec[5] = -1;
      n.rec[7] = ntris;
      for (int k = 0; k < ntris; ++k) {
        const P3* vs[3] = {&tris[k].a, &tris[k].b, &tris[k].c};
        for (int v = 0; v < 3; ++v) {
          n.bucket[(3 * v + 0) * 8 + k] = vs[v]->x;
          n.bucket[(3 * v + 1) * 8 + k] = vs[v]->y;
          n.bucket[(3 * v + 2) * 8 + k] = vs[v]->z;
        }
        n.bucket[72 + k] = eps[k];
        n.bucket[80 + k] = inner[k];
      }
      out.push_back(n);
      if (it.parent != -1 && out[it.parent].rec[5] == -1) out[it.parent].rec[5] = id;
    }
  } catch (const std::invalid_argument&) {
    return D2_ERR_ARG;
  } catch (const std::runtime_error&) {
    return D2_ERR_RUNTIME;
  }
  *n_nodes_out = (int)out.size();
  // Too small an output leaves the generator untouched so the call can be
  // repeated; 2*n_faces records always suffice.
  if ((int)out.size() > max_nodes || !nodes_out || !buckets_out) return D2_ERR_CAPACITY;
  rng[0] = r.state;
  rng[1] = r.inc;
  for (size_t i = 0; i < out.size(); ++i) {
    std::memcpy(nodes_out + 8 * i, out[i].rec, sizeof(out[i].rec));
    std::memcpy(buckets_out + 88 * i, out[i].bucket, sizeof(out[i].bucket));
  }
  return D2_OK;
}
