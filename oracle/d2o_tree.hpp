// d2o_tree.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Surrogate-tree construction, rigid-body states, the two tree traversals
// (discrete and continuous), the contact filter and the TOI -> step rule of
// Delta2, restated without Eigen.  Cites the reference lines it follows.
#pragma once
#include <algorithm>
#include <deque>
#include <stdexcept>
#include <vector>

#include "d2o_geom.hpp"

namespace d2o {

using V = V3<double>;
using T = Tri<double>;
using M = M4<double>;

// model/surrogate_components.h:11-20
struct Node {
  uint32_t node_id = 0, bucket_id = 0;
  bool is_inner = false;
  uint32_t depth = 0;
  int32_t parent_id = -1, child_id_first = -1;
  uint32_t num_children = 0;
};
// model/surrogate_components.h:66-88 (BucketSoup)
struct Bucket {
  T tris[8];
  double eps[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double inner_eps[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int n = 0;
};
struct Tree {
  std::vector<Node> nodes;
  std::vector<Bucket> buckets;
};

// model/fit_surrogate_triangle.h:246-298
inline void selectSurrogateTriangle(const std::vector<V>& pts, Pcg32& rng, T& result_out, double& eps_out) {
  double best_seen = -1;
  int num_pts = (int)pts.size();
  if (num_pts > 3) {
    int attempts = (int)(std::log2(num_pts) * 10);
    for (int i = 0; i < attempts; ++i) {
      int Ai = (int)rng.bounded((uint32_t)num_pts);
      int Bi = (int)rng.bounded((uint32_t)num_pts);
      int Ci = (int)rng.bounded((uint32_t)num_pts);
      T trial(pts[Ai], pts[Bi], pts[Ci]);
      if (trial.area() > best_seen && trial.area() > 0.0) {
        result_out = trial;
        best_seen = trial.area();
      }
    }
  } else {
    result_out = T(pts[0], pts[1], pts[2]);
  }
  if (!(result_out.area() > 0)) throw std::runtime_error("No valid surrogate triangle was found");
  eps_out = 0;
  for (const V& p : pts) eps_out = std::max(eps_out, result_out.distToPoint(p));
}

// model/construct_surrogate_tree.h:8-168 (k-means over triangles; seeds are
// the first `branching` triangles of the whole mesh, construct_surrogate_tree.h:23-25)
inline void divideAndFit(const std::vector<T>& tris, const std::vector<int>& tri_ids, int branching, bool force, Pcg32& rng,
                         std::vector<T>& tris_out, std::vector<std::vector<int>>& ids_out) {
  if (branching > (int)tri_ids.size()) throw std::invalid_argument("There must be more triangles than branches");
  tris_out.assign(branching, T());
  ids_out.clear();
  ids_out.resize(branching);
  for (int i = 0; i < branching; ++i) tris_out[i] = tris[i];
  auto fit_group = [&](int g) {
    std::vector<V> pts;
    for (int id : ids_out[g]) {
      pts.push_back(tris[id].A);
      pts.push_back(tris[id].B);
      pts.push_back(tris[id].C);
    }
    // The reference passes an uninitialised triangle here
    // (construct_surrogate_tree.h:64-67); when no random trial has positive
    // area its value is undefined.  Both this oracle and the library keep the
    // group's previous surrogate in that case.
    T sel = tris_out[g];
    double e;
    selectSurrogateTriangle(pts, rng, sel, e);
    tris_out[g] = sel;
  };
  for (int iteration = 0; iteration < 20; ++iteration) {
    for (int g = 0; g < branching; ++g) ids_out[g].clear();
    for (int t : tri_ids) {
      double best = std::numeric_limits<double>::infinity();
      int bg = 0;
      for (int g = 0; g < branching; ++g) {
        const T& gt = tris_out[g];
        double da = gt.distToPoint(tris[t].A), db = gt.distToPoint(tris[t].B), dc = gt.distToPoint(tris[t].C);
        double mx = std::max(std::max(da, db), dc);
        if (mx < best) {
          bg = g;
          best = mx;
        }
      }
      ids_out[bg].push_back(t);
    }
    for (int g = 0; g < branching; ++g)
      if (!ids_out[g].empty()) fit_group(g);
    if (!force) continue;
    bool improving = true, met = false;
    while (improving && !met) {
      int largest = 0;
      size_t lsize = 0;
      std::vector<int> zeros;
      for (int g = 0; g < branching; ++g) {
        if (ids_out[g].size() > lsize) {
          lsize = ids_out[g].size();
          largest = g;
        }
        if (ids_out[g].empty()) zeros.push_back(g);
      }
      if (!zeros.empty()) {
        std::vector<T> sub;
        std::vector<std::vector<int>> subids;
        int num = (int)std::min(ids_out[largest].size(), zeros.size() + 1);
        std::vector<int> members = ids_out[largest];
        divideAndFit(tris, members, num, false, rng, sub, subids);
        int nonempty = 0;
        for (int s = 0; s < (int)sub.size(); ++s) {
          int gi = (s == 0) ? largest : zeros[s - 1];
          if (!subids[s].empty()) {
            nonempty++;
            ids_out[gi] = subids[s];
            tris_out[gi] = sub[s];
          }
        }
        improving = nonempty > 1;
      } else {
        met = true;
      }
    }
    if (!met) {
      std::vector<int> zeros;
      do {
        int largest = 0;
        size_t lsize = 0;
        zeros.clear();
        for (int g = 0; g < branching; ++g) {
          if (ids_out[g].size() > lsize) {
            lsize = ids_out[g].size();
            largest = g;
          }
          if (ids_out[g].empty()) zeros.push_back(g);
        }
        for (size_t zi = 0; zi < std::min(zeros.size(), ids_out[largest].size() - 1); ++zi) {
          int z = zeros[zi];
          int t = ids_out[largest].back();
          ids_out[z].push_back(t);
          ids_out[largest].pop_back();
          tris_out[z] = tris[t];
        }
        fit_group(largest);
      } while (!zeros.empty());
    }
  }
}

// model/construct_surrogate_tree.h:172-185
inline void calcEpss(const T& s, const std::vector<T>& tris, const std::vector<int>& ids, double& inner, double& eps) {
  inner = std::numeric_limits<double>::infinity();
  eps = 0.0;
  for (int id : ids) {
    double a = s.distToPoint(tris[id].A), b = s.distToPoint(tris[id].B), c = s.distToPoint(tris[id].C);
    inner = std::min(inner, std::min(a, std::min(b, c)));
    eps = std::max(eps, std::max(a, std::max(b, c)));
  }
}

// model/surrogate_tree.cpp:39-112 (BFS build; node i owns bucket i)
inline Tree buildSurrogateTree(const std::vector<V>& Vtx, const std::vector<int>& F, Pcg32& rng) {
  Tree tr;
  std::vector<T> tris;
  for (size_t f = 0; f + 2 < F.size(); f += 3) tris.emplace_back(Vtx[F[f]], Vtx[F[f + 1]], Vtx[F[f + 2]]);
  std::vector<int> all(tris.size());
  for (size_t i = 0; i < tris.size(); ++i) all[i] = (int)i;
  struct Item {
    std::vector<int> ids;
    int depth, parent;
  };
  std::deque<Item> q;
  q.push_back({all, 0, -1});
  while (!q.empty()) {
    Item it = q.front();
    q.pop_front();
    Node n;
    Bucket b;
    if (it.ids.size() < 8) {
      b.n = (int)it.ids.size();
      for (int k = 0; k < b.n; ++k) b.tris[k] = tris[it.ids[k]];
      n.is_inner = false;
      n.num_children = 0;
    } else {
      std::vector<T> st;
      std::vector<std::vector<int>> gids;
      divideAndFit(tris, it.ids, 8, true, rng, st, gids);
      b.n = (int)st.size();
      for (int k = 0; k < b.n; ++k) {
        b.tris[k] = st[k];
        calcEpss(st[k], tris, gids[k], b.inner_eps[k], b.eps[k]);
      }
      n.is_inner = true;
      n.num_children = (uint32_t)st.size();
      for (size_t k = 0; k < st.size(); ++k) q.push_back({gids[k], it.depth + 1, (int)tr.nodes.size()});
    }
    n.depth = (uint32_t)it.depth;
    n.parent_id = it.parent;
    n.child_id_first = -1;
    n.node_id = (uint32_t)tr.nodes.size();
    n.bucket_id = (uint32_t)tr.buckets.size();
    tr.buckets.push_back(b);
    tr.nodes.push_back(n);
    if (it.parent != -1 && tr.nodes[it.parent].child_id_first == -1) tr.nodes[it.parent].child_id_first = (int32_t)n.node_id;
  }
  return tr;
}

// ---------------------------------------------------------------------------
// model/state.cpp
struct State {
  double time = 0;
  Quat<double> rot = Quat<double>::identity();
  V L, trans, vel;
  M transformation() const { return transformationMatrix(rot, trans); }  // state.cpp:86-88
  // state.cpp:90-102 (rotation delta multiplied on the right)
  State extrapolate(double t, const M3<double>& inv_inertia) const {
    State r;
    r.time = time + t;
    r.vel = vel;
    r.trans = trans + t * vel;
    M3<double> R = toRotation(rot);
    M3<double> Iinv = mul(mul(R, inv_inertia), transpose(R));
    V omega = mul(Iinv, L);
    Quat<double> d = qexp((0.5 * omega) * t);
    r.rot = qmul(rot, d);
    r.L = L;
    return r;
  }
};

// model/particle.h:9-64 (only what the path reads)
struct Body {
  const Tree* tree = nullptr;
  State current, future;
  M3<double> inv_inertia;  // body frame, already divided by density
  double geo_eps = 0.01;
  bool is_static = false;
  double last_time_step_size = 1.0;
  int id = 0;
  void projectFutureState(double t) { future = current.extrapolate(t, inv_inertia); }  // particle.cpp:64-66
  // particle.cpp:72-78
  V futurePointVelocity(const V& pt) const {
    M3<double> R = toRotation(future.rot);
    M3<double> Iinv = mul(mul(R, inv_inertia), transpose(R));
    V omega = mul(Iinv, future.L);
    return future.vel + cross(omega, pt - future.trans);
  }
};

// Contact records (collision_detection/contact.h:9-33, 96-108) plus the index
// tuple that identifies each record for set comparisons:
//   kind 0: tri pair (ia, ib) of the node pair (discrete, or CCD soup)
//   kind 1: CCD leaf pair, vertex ia of a against triangle ib of b
//   kind 2: CCD leaf pair, vertex ib of b against triangle ia of a
//   kind 3: CCD leaf pair, edge ia of a against edge ib of b
struct Contact {
  V A, B;
  double eps_a = 0, eps_b = 0, eps_inner_a = 0, eps_inner_b = 0, toc = 0;
  uint32_t node_a = 0, node_b = 0;
  int kind = 0, ia = 0, ib = 0;
  int level = 0;
};

// comparison_dist.h:205-240 for one node pair; appends hits in (a_i, b_i) order.
inline void bucketCompare(const Body& a, const Body& b, const Node& na, const Node& nb, std::vector<Contact>& hits) {
  float search_dist = (float)(a.geo_eps + b.geo_eps);
  const Bucket& ba = a.tree->buckets[na.bucket_id];
  const Bucket& bb = b.tree->buckets[nb.bucket_id];
  M Ta = a.future.transformation(), Tb = b.future.transformation();
  for (int i = 0; i < ba.n; ++i) {
    T A = ba.tris[i].transformed(Ta);
    for (int j = 0; j < bb.n; ++j) {
      T B = bb.tris[j].transformed(Tb);
      V P, Q;
      double d = comparisonDist(A, B, P, Q);
      if (d < ((double)search_dist + ba.eps[i]) + bb.eps[j]) {
        Contact c;
        c.A = P;
        c.B = Q;
        c.eps_a = a.geo_eps + ba.eps[i];
        c.eps_b = b.geo_eps + bb.eps[j];
        c.eps_inner_a = ba.inner_eps[i];
        c.eps_inner_b = bb.inner_eps[j];
        c.node_a = na.node_id;
        c.node_b = nb.node_id;
        c.kind = 0;
        c.ia = i;
        c.ib = j;
        hits.push_back(c);
      }
    }
  }
}

struct NodePair {
  const Node* a;
  const Node* b;
};

// full_tree_comparison.h:13-67
inline std::vector<Contact> compareTreesFull(const Body& a, const Body& b) {
  std::vector<Contact> result;
  std::vector<NodePair> dc{{&a.tree->nodes[0], &b.tree->nodes[0]}}, next;
  int level = 0;
  while (!dc.empty()) {
    for (const NodePair& np : dc) {
      std::vector<Contact> hits;
      bucketCompare(a, b, *np.a, *np.b, hits);
      if (hits.empty()) continue;
      bool adv_a = false, adv_b = false;
      if (!np.a->is_inner && !np.b->is_inner) {
        for (Contact& c : hits) {
          c.level = level;
          result.push_back(c);
        }
      } else if (!np.a->is_inner) {
        adv_b = true;
      } else if (!np.b->is_inner) {
        adv_a = true;
      } else if (np.a->depth < np.b->depth) {
        adv_a = true;
      } else {
        adv_b = true;
      }
      if (adv_a)
        for (uint32_t c = 0; c < np.a->num_children; ++c) next.push_back({&a.tree->nodes[np.a->child_id_first + c], np.b});
      else if (adv_b)
        for (uint32_t c = 0; c < np.b->num_children; ++c) next.push_back({np.a, &b.tree->nodes[np.b->child_id_first + c]});
    }
    dc.swap(next);
    next.clear();
    ++level;
  }
  return result;
}

// contact.h:35-43
inline bool compareContacts(const Contact& x, const Contact& y) {
  double xe = x.eps_a + x.eps_b, ye = y.eps_a + y.eps_b;
  if (xe == ye) return norm(x.A - x.B) < norm(y.A - y.B);
  return xe < ye;
}
// contact.h:44-69 (std::sort is libstdc++'s here, as in the reference build)
inline std::vector<Contact> filterContacts(std::vector<Contact>& unfiltered, double margin) {
  std::vector<Contact> kept;
  std::sort(unfiltered.begin(), unfiltered.end(), compareContacts);
  for (const Contact& ct : unfiltered) {
    bool fresh = true;
    for (const Contact& h : kept) {
      V m1 = (ct.A + ct.B) / 2.0;
      V m2 = (h.A + h.B) / 2.0;
      double d = norm(m1 - m2);
      double search = (ct.eps_a + ct.eps_b) + margin;
      if (d < search) {
        fresh = false;
        break;
      }
    }
    if (fresh) kept.push_back(ct);
  }
  return kept;
}

// contact.h:112-124 (sort on eps_a + eps_b only, then the same greedy pass)
inline std::vector<Contact> filterContinuousContacts(std::vector<Contact> unfiltered, double margin) {
  std::vector<Contact> kept;
  std::sort(unfiltered.begin(), unfiltered.end(),
            [](const Contact& x, const Contact& y) { return (x.eps_a + x.eps_b) < (y.eps_a + y.eps_b); });
  for (const Contact& ct : unfiltered) {
    bool fresh = true;
    for (const Contact& h : kept) {
      double d = norm((ct.A + ct.B) / 2.0 - (h.A + h.B) / 2.0);
      if (d < (ct.eps_a + ct.eps_b) + margin) {
        fresh = false;
        break;
      }
    }
    if (fresh) kept.push_back(ct);
  }
  return kept;
}

// continuous_comparison.h:783-828 for one soup node pair.
inline void bucketCompareContinuousSoup(const Body& a, const Body& b, const Node& na, const Node& nb,
                                        std::vector<Contact>& hits) {
  float search_dist = (float)(a.geo_eps + b.geo_eps);
  M aTs = a.current.transformation(), aTe = a.future.transformation();
  M bTs = b.current.transformation(), bTe = b.future.transformation();
  const Bucket& ba = a.tree->buckets[na.bucket_id];
  const Bucket& bb = b.tree->buckets[nb.bucket_id];
  for (int i = 0; i < ba.n; ++i) {
    const T& at = ba.tris[i];
    BBox<double> abox = at.transformed(aTs).bbox().expand(at.transformed(aTe).bbox()).expand(a.geo_eps);
    for (int j = 0; j < bb.n; ++j) {
      const T& bt = bb.tris[j];
      BBox<double> bbox_ = bt.transformed(bTs).bbox().expand(bt.transformed(bTe).bbox()).expand(b.geo_eps);
      if (!abox.overlap(bbox_)) continue;
      V P, Q;
      double toc;
      double d = triTriCCDLinear(at, aTs, aTe, bt, bTs, bTe, P, Q, toc);
      if (d < (double)search_dist) {
        Contact c;
        c.A = P;
        c.B = Q;
        c.eps_a = a.geo_eps;
        c.eps_b = b.geo_eps;
        c.toc = toc;
        c.node_a = na.node_id;
        c.node_b = nb.node_id;
        c.kind = 0;
        c.ia = i;
        c.ib = j;
        hits.push_back(c);
      }
    }
  }
}

// continuous_comparison.h:830-910 for one leaf-leaf node pair.
inline void bucketCompareContinuousConnected(const Body& a, const Body& b, const Node& na, const Node& nb,
                                             std::vector<Contact>& hits) {
  float search_dist = (float)(a.geo_eps + b.geo_eps);
  const Bucket& ba = a.tree->buckets[na.bucket_id];
  const Bucket& bb = b.tree->buckets[nb.bucket_id];
  M aTs = a.current.transformation(), aTe = a.future.transformation();
  M bTs = b.current.transformation(), bTe = b.future.transformation();
  auto push = [&](const V& A, const V& B, double toc, int kind, int ia, int ib) {
    Contact c;
    c.A = A;
    c.B = B;
    c.eps_a = a.geo_eps;
    c.eps_b = b.geo_eps;
    c.toc = toc;
    c.node_a = na.node_id;
    c.node_b = nb.node_id;
    c.kind = kind;
    c.ia = ia;
    c.ib = ib;
    hits.push_back(c);
  };
  for (int v = 0; v < 3 * ba.n; ++v) {
    const V& p = ba.tris[v / 3][v % 3];
    for (int j = 0; j < bb.n; ++j) {
      V P, Q;
      double toc;
      double d = pointTriCCDLinear(p, aTs, aTe, bb.tris[j], bTs, bTe, P, Q, toc);
      if (d < (double)search_dist) push(P, Q, toc, 1, v, j);
    }
  }
  for (int v = 0; v < 3 * bb.n; ++v) {
    const V& p = bb.tris[v / 3][v % 3];
    for (int i = 0; i < ba.n; ++i) {
      V P, Q;
      double toc;
      double d = pointTriCCDLinear(p, bTs, bTe, ba.tris[i], aTs, aTe, P, Q, toc);
      if (d < (double)search_dist) push(Q, P, toc, 2, i, v);
    }
  }
  for (int ea = 0; ea < 3 * ba.n; ++ea) {
    const T& ta = ba.tris[ea / 3];
    int k = ea % 3;
    V eA = ta[k], eB = ta[(k + 1) % 3];
    Edge<double> as(transform(eA, aTs), transform(eB, aTs)), ae(transform(eA, aTe), transform(eB, aTe));
    for (int eb = 0; eb < 3 * bb.n; ++eb) {
      const T& tb = bb.tris[eb / 3];
      int l = eb % 3;
      V fA = tb[l], fB = tb[(l + 1) % 3];
      Edge<double> bs(transform(fA, bTs), transform(fB, bTs)), be(transform(fA, bTe), transform(fB, bTe));
      V P, Q;
      double toc;
      double d = edgeEdgeCCD(as, ae, bs, be, P, Q, toc);
      if (d < (double)search_dist) push(Q, P, toc, 3, ea, eb);
    }
  }
}

// full_tree_comparison.h:69-219.  Overwrites both bodies' future state.
inline std::vector<Contact> compareTreesFullContinuous(Body& a, Body& b, double max_time, double& min_time_out) {
  std::vector<Contact> result;
  min_time_out = max_time;
  a.projectFutureState(max_time);
  b.projectFutureState(max_time);
  std::vector<NodePair> soup{{&a.tree->nodes[0], &b.tree->nodes[0]}}, conn, soup_next, conn_next;
  int level = 0;
  auto expand = [&](const NodePair& np) {
    bool adv_a = false, adv_b = false;
    if (!np.a->is_inner) adv_b = true;
    else if (!np.b->is_inner) adv_a = true;
    else if (np.a->depth < np.b->depth) adv_a = true;
    else adv_b = true;
    auto route = [&](const Node* x, const Node* y) {
      if (x->is_inner || y->is_inner) soup_next.push_back({x, y});
      else conn_next.push_back({x, y});
    };
    if (adv_a)
      for (uint32_t c = 0; c < np.a->num_children; ++c) route(&a.tree->nodes[np.a->child_id_first + c], np.b);
    else if (adv_b)
      for (uint32_t c = 0; c < np.b->num_children; ++c) route(np.a, &b.tree->nodes[np.b->child_id_first + c]);
  };
  while (!soup.empty() || !conn.empty()) {
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<NodePair>& list = pass == 0 ? soup : conn;
      std::vector<Contact> hits;
      std::vector<size_t> counts;
      for (const NodePair& np : list) {
        if (pass == 0) bucketCompareContinuousSoup(a, b, *np.a, *np.b, hits);
        else bucketCompareContinuousConnected(a, b, *np.a, *np.b, hits);
        counts.push_back(hits.size());
      }
      size_t last = 0;
      for (size_t i = 0; i < list.size(); ++i) {
        const NodePair& np = list[i];
        if (counts[i] > last) {
          if (!np.a->is_inner && !np.b->is_inner) {
            for (size_t h = last; h < counts[i]; ++h) {
              Contact c = hits[h];
              c.level = 2 * level + pass;
              result.push_back(c);
            }
            for (const Contact& c : hits) min_time_out = std::min(min_time_out, c.toc * max_time);
          } else {
            expand(np);
          }
        }
        last = counts[i];
      }
    }
    soup.swap(soup_next);
    soup_next.clear();
    conn.swap(conn_next);
    conn_next.clear();
    ++level;
  }
  return result;
}

// separate_clusters.cpp:496-564: the step scale one pair's CCD contacts ask for.
inline double pairStepScale(const std::vector<Contact>& Ccs, const Body& a, const Body& b, double max_time) {
  double min_scaling = 1.0;
  for (const Contact& c : Ccs) {
    V n = c.A - c.B;
    double interaction = c.eps_a + c.eps_b;
    V rel = a.futurePointVelocity(c.A) - b.futurePointVelocity(c.B);
    if (dot(normalized(n), normalized(rel)) <= 0.0 || norm(n) < 1e-4) {
      if (norm(n) < 1e-4) {
        if (c.toc == 0.0) continue;
        double start = std::max(0.0, c.toc - (norm(rel) * max_time) / interaction);
        double s = lerp(start, c.toc, 0.75);
        min_scaling = std::min(min_scaling, s);
      } else if (c.toc < 1.0) {
        double proj_vel = -(dot(normalized(n), rel));
        double depth = interaction - norm(n);
        double proj_dist = (proj_vel * c.toc) * max_time;
        if (proj_dist > 1e-6) {
          if (depth < proj_dist) {
            double s = lerp(c.toc * (1.0 - depth / proj_dist), c.toc, 0.5);
            min_scaling = std::min(min_scaling, s);
          } else {
            double fp = c.toc + ((1.0 - c.toc) * interaction) / (proj_vel * max_time);
            double s = lerp(c.toc, std::min(fp, 1.0), 0.8);
            min_scaling = std::min(min_scaling, s);
          }
        }
      }
    }
  }
  return min_scaling;
}

}  // namespace d2o
