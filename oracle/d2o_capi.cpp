// d2o_capi.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Flat C entry points over the oracle restatement, for the pytest suite, the
// smoke check and the CPU-baseline leg of bench.py.  Nothing in the product
// links or loads this library.
//
// Flat formats are the ones the product's C ABI uses (include/d2hip.h):
//   body record   40 doubles: current state (q w,x,y,z | t xyz | v xyz | L xyz | time)
//                 future state (same 14), body-frame inverse inertia (3x3 row-major),
//                 geo_eps, is_static, last_time_step_size
//   node record    8 int32 : node_id, bucket_id, is_inner, depth, parent_id,
//                            child_id_first, num_children, num_tris
//   bucket record 88 double: Ax[8] Ay[8] Az[8] Bx[8] By[8] Bz[8] Cx[8] Cy[8] Cz[8]
//                            eps[8] inner_eps[8]
//   contact ints   8 int32 : pair, node_a, node_b, kind, ia, ib, level, 0
//   contact reals 11 double: A xyz, B xyz, eps_a, eps_b, eps_inner_a, eps_inner_b, toc
#include <chrono>
#include <cstring>
#include <memory>
#include <thread>
#include <atomic>

#include "d2o_broad.hpp"
#include "d2o_solver.hpp"
#include "d2o_tree.hpp"

using namespace d2o;

namespace {

template <class R> Tri<R> triFrom(const double* p) {
  return Tri<R>(V3<R>(R(p[0]), R(p[1]), R(p[2])), V3<R>(R(p[3]), R(p[4]), R(p[5])), V3<R>(R(p[6]), R(p[7]), R(p[8])));
}
template <class R> M4<R> matFrom(const double* p) {
  M4<R> m;
  for (int i = 0; i < 16; ++i) m.m[i] = R(p[i]);
  return m;
}
template <class R> V3<R> vFrom(const double* p) { return V3<R>(R(p[0]), R(p[1]), R(p[2])); }
template <class R> void vTo(const V3<R>& v, double* o) {
  o[0] = (double)v.x;
  o[1] = (double)v.y;
  o[2] = (double)v.z;
}

State stateFrom(const double* r) {
  State s;
  s.rot = {r[0], r[1], r[2], r[3]};
  s.trans = V(r[4], r[5], r[6]);
  s.vel = V(r[7], r[8], r[9]);
  s.L = V(r[10], r[11], r[12]);
  s.time = r[13];
  return s;
}
void stateTo(const State& s, double* r) {
  r[0] = s.rot.w;
  r[1] = s.rot.x;
  r[2] = s.rot.y;
  r[3] = s.rot.z;
  vTo(s.trans, r + 4);
  vTo(s.vel, r + 7);
  vTo(s.L, r + 10);
  r[13] = s.time;
}

struct Scene {
  std::vector<std::unique_ptr<Tree>> trees;
  std::vector<Body> bodies;
  std::vector<int32_t> out_i;
  std::vector<double> out_r;
  std::vector<int32_t> out_off;
  std::vector<double> out_pair;  // per pair: step scale (CCD) / min_time_out
  std::vector<SolverBody> props;  // per body: mass, friction coefficient (the solver)
  WarmStore store;                // the solver's warm start, kept across solves
  std::vector<SolverOut> sout;    // per hit of the last solve
};

void exportContacts(Scene& sc, int pair, const std::vector<Contact>& cs) {
  for (const Contact& c : cs) {
    int32_t iv[8] = {pair, (int32_t)c.node_a, (int32_t)c.node_b, c.kind, c.ia, c.ib, c.level, 0};
    sc.out_i.insert(sc.out_i.end(), iv, iv + 8);
    double rv[11] = {c.A.x, c.A.y, c.A.z, c.B.x, c.B.y, c.B.z, c.eps_a, c.eps_b, c.eps_inner_a, c.eps_inner_b, c.toc};
    sc.out_r.insert(sc.out_r.end(), rv, rv + 11);
  }
}

}  // namespace


// ----------------------------------------------------------- primitives ----
// Every primitive comes in an f64 and an f32 flavour (the reference tests run
// both template instantiations).  Inputs are passed as doubles and rounded to
// the working type on entry.
template <class R> static void transformationMatrixEuler_impl(const double* e, const double* t, double* out) {
  M4<R> m = transformationMatrixEuler(vFrom<R>(e), vFrom<R>(t));
  for (int i = 0; i < 16; ++i) out[i] = (double)m.m[i];
}
extern "C" void d2o_transformation_matrix_euler_f64(const double* e, const double* t, double* out) {
  transformationMatrixEuler_impl<double>(e, t, out);
}
extern "C" void d2o_transformation_matrix_euler_f32(const double* e, const double* t, double* out) {
  transformationMatrixEuler_impl<float>(e, t, out);
}

template <class R> static void transformPoint_impl(const double* p, const double* m, double* out) {
  vTo(transform(vFrom<R>(p), matFrom<R>(m)), out);
}
extern "C" void d2o_transform_point_f64(const double* p, const double* m, double* out) { transformPoint_impl<double>(p, m, out); }
extern "C" void d2o_transform_point_f32(const double* p, const double* m, double* out) { transformPoint_impl<float>(p, m, out); }

template <class R> static void comparisonDist_impl(const double* a, const double* b, double* out) {
  V3<R> P, Q;
  R d = comparisonDist(triFrom<R>(a), triFrom<R>(b), P, Q);
  out[0] = (double)d;
  vTo(P, out + 1);
  vTo(Q, out + 4);
}
extern "C" void d2o_comparison_dist_f64(const double* a, const double* b, double* out) { comparisonDist_impl<double>(a, b, out); }
extern "C" void d2o_comparison_dist_f32(const double* a, const double* b, double* out) { comparisonDist_impl<float>(a, b, out); }

template <class R> static void edgeClosest_impl(const double* e, const double* f, double* out) {
  Edge<R> a(vFrom<R>(e), vFrom<R>(e + 3)), b(vFrom<R>(f), vFrom<R>(f + 3));
  V3<R> P, Q;
  out[0] = (double)a.closest(b, P, Q);
  vTo(P, out + 1);
  vTo(Q, out + 4);
}
extern "C" void d2o_edge_closest_f64(const double* e, const double* f, double* out) { edgeClosest_impl<double>(e, f, out); }
extern "C" void d2o_edge_closest_f32(const double* e, const double* f, double* out) { edgeClosest_impl<float>(e, f, out); }

template <class R> static void barycentric_impl(const double* tri, const double* p, double* out) {
  R u, v, w;
  triFrom<R>(tri).calcBarycentric(vFrom<R>(p), u, v, w);
  out[0] = (double)u;
  out[1] = (double)v;
  out[2] = (double)w;
}
extern "C" void d2o_barycentric_f64(const double* tri, const double* p, double* out) { barycentric_impl<double>(tri, p, out); }
extern "C" void d2o_barycentric_f32(const double* tri, const double* p, double* out) { barycentric_impl<float>(tri, p, out); }

// out: t, t_inf, isIntersecting, point xyz
template <class R> static void intersectSegment_impl(const double* tri, const double* e, double* out) {
  R t;
  bool tinf, is;
  V3<R> p = triFrom<R>(tri).intersectSegment(Edge<R>(vFrom<R>(e), vFrom<R>(e + 3)), t, tinf, is);
  out[0] = (double)t;
  out[1] = tinf;
  out[2] = is;
  vTo(p, out + 3);
}
extern "C" void d2o_intersect_segment_f64(const double* tri, const double* e, double* out) { intersectSegment_impl<double>(tri, e, out); }
extern "C" void d2o_intersect_segment_f32(const double* tri, const double* e, double* out) { intersectSegment_impl<float>(tri, e, out); }

// out: inTriangle, point xyz
template <class R> static void projectToPlane_impl(const double* tri, const double* p, double* out) {
  bool in;
  V3<R> q = triFrom<R>(tri).projectToPlane(vFrom<R>(p), in);
  out[0] = in;
  vTo(q, out + 1);
}
extern "C" void d2o_project_to_plane_f64(const double* tri, const double* p, double* out) { projectToPlane_impl<double>(tri, p, out); }
extern "C" void d2o_project_to_plane_f32(const double* tri, const double* p, double* out) { projectToPlane_impl<float>(tri, p, out); }

extern "C" double d2o_dist_to_point_f64(const double* tri, const double* p) { return triFrom<double>(tri).distToPoint(vFrom<double>(p)); }
extern "C" double d2o_dist_to_point_f32(const double* tri, const double* p) { return triFrom<float>(tri).distToPoint(vFrom<float>(p)); }

extern "C" double d2o_sphere_sphere_f64(const double* ap, const double* av, double ar, const double* bp, const double* bv, double br) {
  return sphereSphereFirstIntersect(vFrom<double>(ap), vFrom<double>(av), ar, vFrom<double>(bp), vFrom<double>(bv), br);
}
extern "C" double d2o_sphere_sphere_f32(const double* ap, const double* av, double ar, const double* bp, const double* bv, double br) {
  return sphereSphereFirstIntersect(vFrom<float>(ap), vFrom<float>(av), (float)ar, vFrom<float>(bp), vFrom<float>(bv), (float)br);
}

// out: dist, P xyz, Q xyz, toc
template <class R>
static void pointTri_impl(const double* pt, const double* pTs, const double* pTe, const double* tri, const double* tTs,
                          const double* tTe, double* out) {
  V3<R> P, Q;
  R toc;
  R d = pointTriCCDLinear(vFrom<R>(pt), matFrom<R>(pTs), matFrom<R>(pTe), triFrom<R>(tri), matFrom<R>(tTs), matFrom<R>(tTe), P,
                          Q, toc);
  out[0] = (double)d;
  vTo(P, out + 1);
  vTo(Q, out + 4);
  out[7] = (double)toc;
}
extern "C" void d2o_point_tri_ccd_f64(const double* pt, const double* a, const double* b, const double* tri, const double* c,
                           const double* d, double* out) {
  pointTri_impl<double>(pt, a, b, tri, c, d, out);
}
extern "C" void d2o_point_tri_ccd_f32(const double* pt, const double* a, const double* b, const double* tri, const double* c,
                           const double* d, double* out) {
  pointTri_impl<float>(pt, a, b, tri, c, d, out);
}

template <class R>
static void edgeEdge_impl(const double* as, const double* ae, const double* bs, const double* be, double* out) {
  V3<R> P, Q;
  R toc = R(-1);
  R d = edgeEdgeCCD(Edge<R>(vFrom<R>(as), vFrom<R>(as + 3)), Edge<R>(vFrom<R>(ae), vFrom<R>(ae + 3)),
                    Edge<R>(vFrom<R>(bs), vFrom<R>(bs + 3)), Edge<R>(vFrom<R>(be), vFrom<R>(be + 3)), P, Q, toc);
  out[0] = (double)d;
  vTo(P, out + 1);
  vTo(Q, out + 4);
  out[7] = (double)toc;
}
extern "C" void d2o_edge_edge_ccd_f64(const double* as, const double* ae, const double* bs, const double* be, double* out) {
  edgeEdge_impl<double>(as, ae, bs, be, out);
}
extern "C" void d2o_edge_edge_ccd_f32(const double* as, const double* ae, const double* bs, const double* be, double* out) {
  edgeEdge_impl<float>(as, ae, bs, be, out);
}

template <class R>
static void triTri_impl(const double* a, const double* aTs, const double* aTe, const double* b, const double* bTs,
                        const double* bTe, double* out) {
  V3<R> P, Q;
  R toc;
  R d = triTriCCDLinear(triFrom<R>(a), matFrom<R>(aTs), matFrom<R>(aTe), triFrom<R>(b), matFrom<R>(bTs), matFrom<R>(bTe), P, Q,
                        toc);
  out[0] = (double)d;
  vTo(P, out + 1);
  vTo(Q, out + 4);
  out[7] = (double)toc;
}
extern "C" void d2o_tri_tri_ccd_f64(const double* a, const double* b, const double* c, const double* d, const double* e,
                         const double* f, double* out) {
  triTri_impl<double>(a, b, c, d, e, f, out);
}
extern "C" void d2o_tri_tri_ccd_f32(const double* a, const double* b, const double* c, const double* d, const double* e,
                         const double* f, double* out) {
  triTri_impl<float>(a, b, c, d, e, f, out);
}

// State helpers (model/state.cpp).
extern "C" void d2o_extrapolate(const double* state14, const double* inv_inertia9, double t, double* out14) {
  M3<double> I;
  for (int i = 0; i < 9; ++i) I.m[i] = inv_inertia9[i];
  stateTo(stateFrom(state14).extrapolate(t, I), out14);
}
extern "C" void d2o_state_transformation(const double* state14, double* out16) {
  M m = stateFrom(state14).transformation();
  for (int i = 0; i < 16; ++i) out16[i] = m.m[i];
}
// Particle::futurePointVelocity (model/particle.cpp:72-78) of a body record
// (its future state and body-frame inverse inertia) at world point pt.
extern "C" void d2o_future_point_velocity(const double* body40, const double* pt, double* out3) {
  Body b;
  b.future = stateFrom(body40 + 14);
  for (int k = 0; k < 9; ++k) b.inv_inertia.m[k] = body40[28 + k];
  vTo(b.futurePointVelocity(vFrom<double>(pt)), out3);
}

// PCG32 (common/pcg_basic.cpp) for seeding checks: writes n bounded draws.
extern "C" void d2o_pcg_draws(uint64_t s0, uint64_t s1, uint32_t bound, int n, uint32_t* out) {
  Pcg32 r;
  r.seed(s0, s1);
  for (int i = 0; i < n; ++i) out[i] = bound ? r.bounded(bound) : r.next();
}

// --------------------------------------------------------------- trees ----
// Builds one surrogate tree from V (nv x 3) and F (nf x 3) with the PCG state
// rng[2] = {state, inc}, which is advanced in place (the reference shares one
// generator across all meshes it builds).
extern "C" void* d2o_tree_build(const double* Vp, int nv, const int32_t* Fp, int nf, uint64_t* rng) {
  std::vector<V> Vt(nv);
  for (int i = 0; i < nv; ++i) Vt[i] = V(Vp[3 * i], Vp[3 * i + 1], Vp[3 * i + 2]);
  std::vector<int> F(Fp, Fp + 3 * nf);
  Pcg32 r;
  r.state = rng[0];
  r.inc = rng[1];
  Tree* t = new Tree(buildSurrogateTree(Vt, F, r));
  rng[0] = r.state;
  rng[1] = r.inc;
  return t;
}
extern "C" void d2o_pcg_seed(uint64_t s0, uint64_t s1, uint64_t* rng) {
  Pcg32 r;
  r.seed(s0, s1);
  rng[0] = r.state;
  rng[1] = r.inc;
}
extern "C" int d2o_tree_num_nodes(void* h) { return (int)static_cast<Tree*>(h)->nodes.size(); }
extern "C" void d2o_tree_export(void* h, int32_t* nodes, double* buckets) {
  Tree* t = static_cast<Tree*>(h);
  for (size_t i = 0; i < t->nodes.size(); ++i) {
    const Node& n = t->nodes[i];
    int32_t* o = nodes + 8 * i;
    o[0] = (int32_t)n.node_id;
    o[1] = (int32_t)n.bucket_id;
    o[2] = n.is_inner;
    o[3] = (int32_t)n.depth;
    o[4] = n.parent_id;
    o[5] = n.child_id_first;
    o[6] = (int32_t)n.num_children;
    o[7] = t->buckets[n.bucket_id].n;
    double* b = buckets + 88 * i;
    std::memset(b, 0, 88 * sizeof(double));
    const Bucket& bk = t->buckets[n.bucket_id];
    for (int k = 0; k < bk.n; ++k) {
      const T& tr = bk.tris[k];
      const V* vs[3] = {&tr.A, &tr.B, &tr.C};
      for (int v = 0; v < 3; ++v) {
        b[(3 * v + 0) * 8 + k] = vs[v]->x;
        b[(3 * v + 1) * 8 + k] = vs[v]->y;
        b[(3 * v + 2) * 8 + k] = vs[v]->z;
      }
      b[72 + k] = bk.eps[k];
      b[80 + k] = bk.inner_eps[k];
    }
  }
}
extern "C" void* d2o_tree_import(const int32_t* nodes, const double* buckets, int n) {
  Tree* t = new Tree();
  t->nodes.resize(n);
  t->buckets.resize(n);
  for (int i = 0; i < n; ++i) {
    const int32_t* o = nodes + 8 * i;
    Node& nd = t->nodes[i];
    nd.node_id = (uint32_t)o[0];
    nd.bucket_id = (uint32_t)o[1];
    nd.is_inner = o[2] != 0;
    nd.depth = (uint32_t)o[3];
    nd.parent_id = o[4];
    nd.child_id_first = o[5];
    nd.num_children = (uint32_t)o[6];
    Bucket& bk = t->buckets[nd.bucket_id];
    bk.n = o[7];
    const double* b = buckets + 88 * i;
    for (int k = 0; k < bk.n; ++k) {
      bk.tris[k] = T(V(b[0 * 8 + k], b[1 * 8 + k], b[2 * 8 + k]), V(b[3 * 8 + k], b[4 * 8 + k], b[5 * 8 + k]),
                     V(b[6 * 8 + k], b[7 * 8 + k], b[8 * 8 + k]));
      bk.eps[k] = b[72 + k];
      bk.inner_eps[k] = b[80 + k];
    }
  }
  return t;
}
extern "C" void d2o_tree_free(void* h) { delete static_cast<Tree*>(h); }

// --------------------------------------------------------------- scenes ---
// A scene owns imported trees and bodies.  tree_of_body maps each body to one
// of the trees (trees may be shared).
extern "C" void* d2o_scene_create(int n_trees, void** trees, int n_bodies, const int32_t* tree_of_body, const double* body_records) {
  Scene* sc = new Scene();
  for (int i = 0; i < n_trees; ++i) sc->trees.emplace_back(static_cast<Tree*>(trees[i]));
  sc->bodies.resize(n_bodies);
  for (int i = 0; i < n_bodies; ++i) {
    const double* r = body_records + 40 * i;
    Body& b = sc->bodies[i];
    b.tree = sc->trees[tree_of_body[i]].get();
    b.current = stateFrom(r);
    b.future = stateFrom(r + 14);
    for (int k = 0; k < 9; ++k) b.inv_inertia.m[k] = r[28 + k];
    b.geo_eps = r[37];
    b.is_static = r[38] != 0.0;
    b.last_time_step_size = r[39];
    b.id = i;
  }
  return sc;
}
extern "C" void d2o_scene_free(void* h) { delete static_cast<Scene*>(h); }
extern "C" void d2o_scene_get_bodies(void* h, double* body_records) {
  Scene* sc = static_cast<Scene*>(h);
  for (size_t i = 0; i < sc->bodies.size(); ++i) {
    double* r = body_records + 40 * i;
    stateTo(sc->bodies[i].current, r);
    stateTo(sc->bodies[i].future, r + 14);
  }
}

// Replaces every body's states, last step and flags from records (trees and
// the solver's warm store stay).
extern "C" void d2o_scene_set_bodies(void* h, const double* body_records) {
  Scene* sc = static_cast<Scene*>(h);
  for (size_t i = 0; i < sc->bodies.size(); ++i) {
    const double* r = body_records + 40 * i;
    Body& b = sc->bodies[i];
    b.current = stateFrom(r);
    b.future = stateFrom(r + 14);
    for (int k = 0; k < 9; ++k) b.inv_inertia.m[k] = r[28 + k];
    b.geo_eps = r[37];
    b.is_static = r[38] != 0.0;
    b.last_time_step_size = r[39];
  }
}

// Discrete narrow phase over a pair list (PDE_explicit.cpp:77-95): for each
// pair compareTreesFull, then filterContacts(.., 0.0) when filter != 0.
// Results are kept in the scene; see d2o_scene_result_*.  Returns the number
// of contacts.
extern "C" int d2o_narrow(void* h, int n_pairs, const int32_t* pairs, int filter) {
  Scene* sc = static_cast<Scene*>(h);
  sc->out_i.clear();
  sc->out_r.clear();
  sc->out_off.assign(1, 0);
  sc->out_pair.assign(n_pairs, 0.0);
  for (int p = 0; p < n_pairs; ++p) {
    Body& a = sc->bodies[pairs[2 * p]];
    Body& b = sc->bodies[pairs[2 * p + 1]];
    std::vector<Contact> cs = compareTreesFull(a, b);
    if (filter) cs = filterContacts(cs, 0.0);
    exportContacts(*sc, p, cs);
    sc->out_off.push_back((int32_t)(sc->out_i.size() / 8));
  }
  return (int)(sc->out_i.size() / 8);
}

// CCD over a pair list (separate_clusters.cpp:479-566 without the lock):
// compareTreesFullContinuous with max_time[p], then the step rule.  Per pair
// the result keeps [scale, min_time_out].  Future states of the bodies are
// left as the reference leaves them (projected by the last pair that touched
// them).
extern "C" int d2o_ccd(void* h, int n_pairs, const int32_t* pairs, const double* max_time) {
  Scene* sc = static_cast<Scene*>(h);
  sc->out_i.clear();
  sc->out_r.clear();
  sc->out_off.assign(1, 0);
  sc->out_pair.assign(2 * (size_t)n_pairs, 0.0);
  for (int p = 0; p < n_pairs; ++p) {
    Body& a = sc->bodies[pairs[2 * p]];
    Body& b = sc->bodies[pairs[2 * p + 1]];
    double mt;
    std::vector<Contact> cs = compareTreesFullContinuous(a, b, max_time[p], mt);
    sc->out_pair[2 * p] = cs.empty() ? 1.0 : pairStepScale(cs, a, b, max_time[p]);
    sc->out_pair[2 * p + 1] = mt;
    exportContacts(*sc, p, cs);
    sc->out_off.push_back((int32_t)(sc->out_i.size() / 8));
  }
  return (int)(sc->out_i.size() / 8);
}

// filterContacts(unfiltered, margin) (contact.h:44-69) over n contacts given
// as reals (11 per contact): sorted_out = the inputs' order after the
// reference's in-place std::sort, kept_out = the kept inputs in that order.
// Returns the number kept.
extern "C" int d2o_filter_order(const double* reals, int n, double margin, int32_t* sorted_out, int32_t* kept_out) {
  std::vector<Contact> cs(n);
  for (int i = 0; i < n; ++i) {
    const double* r = reals + 11 * i;
    cs[i].A = V(r[0], r[1], r[2]);
    cs[i].B = V(r[3], r[4], r[5]);
    cs[i].eps_a = r[6], cs[i].eps_b = r[7], cs[i].eps_inner_a = r[8], cs[i].eps_inner_b = r[9], cs[i].toc = r[10];
    cs[i].ia = i;  // the input position travels with the contact
  }
  std::vector<Contact> kept = filterContacts(cs, margin);  // sorts cs in place
  for (int i = 0; i < n; ++i) sorted_out[i] = cs[i].ia;
  for (size_t k = 0; k < kept.size(); ++k) kept_out[k] = kept[k].ia;
  return (int)kept.size();
}

// Every leaf of body a against every leaf of body b, without the traversal:
// mode 0 runs the discrete bucket comparison (comparison_dist.h:205-240) at the
// future states; mode 1 projects both future states at max_time as
// compareTreesFullContinuous does (full_tree_comparison.h:78-79) and runs the
// connected CCD bucket comparison (continuous_comparison.h:830-910).  Leaf
// pairs in (leaf of a, leaf of b) node order.  The result replaces the
// scene's last result (one pair); returns the number of hits.
extern "C" int d2o_exhaustive_leaves(void* h, int ia, int ib, int mode, double max_time) {
  Scene* sc = static_cast<Scene*>(h);
  Body& a = sc->bodies[ia];
  Body& b = sc->bodies[ib];
  if (mode == 1) {
    a.projectFutureState(max_time);
    b.projectFutureState(max_time);
  }
  std::vector<Contact> hits;
  for (const Node& na : a.tree->nodes) {
    if (na.is_inner) continue;
    for (const Node& nb : b.tree->nodes) {
      if (nb.is_inner) continue;
      if (mode == 0) bucketCompare(a, b, na, nb, hits);
      else bucketCompareContinuousConnected(a, b, na, nb, hits);
    }
  }
  sc->out_i.clear();
  sc->out_r.clear();
  sc->out_off.assign(1, 0);
  sc->out_pair.assign(2, 0.0);
  exportContacts(*sc, 0, hits);
  sc->out_off.push_back((int32_t)(sc->out_i.size() / 8));
  return (int)(sc->out_i.size() / 8);
}

extern "C" void d2o_scene_result(void* h, int32_t* offsets, int32_t* ints, double* reals, double* per_pair) {
  Scene* sc = static_cast<Scene*>(h);
  if (offsets) std::memcpy(offsets, sc->out_off.data(), sc->out_off.size() * sizeof(int32_t));
  if (ints) std::memcpy(ints, sc->out_i.data(), sc->out_i.size() * sizeof(int32_t));
  if (reals) std::memcpy(reals, sc->out_r.data(), sc->out_r.size() * sizeof(double));
  if (per_pair) std::memcpy(per_pair, sc->out_pair.data(), sc->out_pair.size() * sizeof(double));
}

// fineCollisionClustersWithTimeStepSelection + TimeStepSelectionDynamicContinuous
// (separate_clusters.cpp:462-599, time_step_selection_dynamic_continuous.cpp:17-21)
// for one cluster given as a list of pairs over scene bodies.  dt is the
// float Options::time_step_size.  Returns the clamped cluster step.
extern "C" double d2o_cluster_step(void* h, int n_pairs, const int32_t* pairs, float dt) {
  Scene* sc = static_cast<Scene*>(h);
  std::vector<int> local;  // cluster.particles order: first appearance
  auto lid = [&](int g) {
    for (size_t i = 0; i < local.size(); ++i)
      if (local[i] == g) return (int)i;
    local.push_back(g);
    return (int)local.size() - 1;
  };
  for (int p = 0; p < n_pairs; ++p) {
    lid(pairs[2 * p]);
    lid(pairs[2 * p + 1]);
  }
  std::vector<double> step(local.size());
  for (size_t i = 0; i < local.size(); ++i) step[i] = sc->bodies[local[i]].last_time_step_size;
  double cluster_step = -1.0;
  for (int p = 0; p < n_pairs; ++p) {
    int ia = lid(pairs[2 * p]), ib = lid(pairs[2 * p + 1]);
    Body& a = sc->bodies[local[ia]];
    Body& b = sc->bodies[local[ib]];
    double max_time = std::min(a.last_time_step_size, b.last_time_step_size);
    double mt;
    std::vector<Contact> cs = compareTreesFullContinuous(a, b, max_time, mt);
    if (!cs.empty()) {
      double scale = pairStepScale(cs, a, b, max_time);
      double use = std::min(max_time * scale, std::min(step[ia], step[ib]));
      if (!a.is_static) step[ia] = use;
      if (!b.is_static) step[ib] = use;
    }
  }
  for (int p = 0; p < n_pairs; ++p) {
    int ia = lid(pairs[2 * p]), ib = lid(pairs[2 * p + 1]);
    bool sa = sc->bodies[local[ia]].is_static, sb = sc->bodies[local[ib]].is_static;
    int id;
    if (!sa && !sb) id = ia;
    else if (sa && sb) continue;
    else if (sa) id = ib;
    else id = ia;
    if (cluster_step < 0.0) cluster_step = step[ia];
    else cluster_step = std::min(cluster_step, step[id]);
  }
  if (cluster_step < 0.0) cluster_step = local.empty() ? (double)dt : sc->bodies[local[0]].last_time_step_size;
  return std::max(1e-3, std::min(cluster_step, (double)dt));
}

// Sequential impulses (sequential_impulses.cpp:104-638) for one cluster.
// mass/friction per scene body must be set first (d2o_solver_props).  The
// cluster's hits are gathered as PDEExplicit::step gathers them
// (PDE_explicit.cpp:77-95): per pair in order, compareTreesFull then
// filterContacts(.., 0.0), at the bodies' current future states.  Returns 1
// (solved: future states replaced, warm store updated) or 0 (the reference
// returns false), and the hit count in *n_hits; d2o_solve_results reads the
// per-hit values (11 doubles: normal, impulse, impulse offset, friction
// impulse, tangential velocity) in gathered order.
extern "C" void d2o_solver_props(void* h, const double* mass, const double* friction) {
  Scene* sc = static_cast<Scene*>(h);
  sc->props.resize(sc->bodies.size());
  for (size_t i = 0; i < sc->bodies.size(); ++i) sc->props[i].mass = mass[i], sc->props[i].friction = friction[i];
}
extern "C" int d2o_solve_cluster(void* h, int n_bodies, const int32_t* bodies, int n_pairs, const int32_t* pairs,
                                 double t, const double* g, int* n_hits) {
  Scene* sc = static_cast<Scene*>(h);
  std::vector<SolverHit> hits;
  for (int p = 0; p < n_pairs; ++p) {
    Body& a = sc->bodies[pairs[2 * p]];
    Body& b = sc->bodies[pairs[2 * p + 1]];
    std::vector<Contact> cs = compareTreesFull(a, b);
    cs = filterContacts(cs, 0.0);
    for (const Contact& c : cs) {
      SolverHit sh;
      sh.ct = c, sh.a = pairs[2 * p], sh.b = pairs[2 * p + 1], sh.row = (int)hits.size();
      hits.push_back(sh);
    }
  }
  std::vector<int> cluster(bodies, bodies + n_bodies);
  sc->sout.assign(hits.size(), SolverOut());
  if (n_hits) *n_hits = (int)hits.size();
  return solveCluster(sc->bodies, sc->props, cluster, hits, t, V(g[0], g[1], g[2]), sc->store, sc->sout) ? 1 : 0;
}
extern "C" void d2o_solve_results(void* h, double* out) {
  Scene* sc = static_cast<Scene*>(h);
  for (size_t i = 0; i < sc->sout.size(); ++i) {
    const SolverOut& o = sc->sout[i];
    double v[11] = {o.normal.x, o.normal.y, o.normal.z, o.impulse, o.impulse_offset, o.friction.x, o.friction.y,
                    o.friction.z, o.v_tangent.x, o.v_tangent.y, o.v_tangent.z};
    std::memcpy(out + 11 * i, v, sizeof(v));
  }
}
// The warm store's value for the body pair (a, b); returns 0 when absent.
extern "C" int d2o_warm_get(void* h, int a, int b, double* value) {
  Scene* sc = static_cast<Scene*>(h);
  auto it = sc->store.find(std::make_pair(std::min(a, b), std::max(a, b)));
  if (it == sc->store.end()) return 0;
  *value = it->second;
  return 1;
}
// PDEExplicit::step's tail for the dynamic bodies (PDE_explicit.cpp:117-131):
// current <- future, future <- current.extrapolate(step); step per body.
extern "C" void d2o_commit(void* h, const double* step) {
  Scene* sc = static_cast<Scene*>(h);
  for (size_t i = 0; i < sc->bodies.size(); ++i) {
    Body& b = sc->bodies[i];
    if (b.is_static) continue;
    b.current = b.future;
    b.last_time_step_size = step[i];
    b.projectFutureState(step[i]);
  }
}

// Timed CPU baseline: runs CCD + step rule, then discrete + filter, for every
// pair, `repeats` times on `threads` threads (OpenMP-free: std::thread over a
// shared pair counter).  Returns the median wall time in seconds of one pass.
// max_time[p] is the CCD interval; the discrete phase uses the bodies' stored
// future states.
extern "C" double d2o_time_pass(void* h, int n_pairs, const int32_t* pairs, const double* max_time, int threads, int repeats,
                     double* contact_count) {
  Scene* sc = static_cast<Scene*>(h);
  std::vector<double> times;
  double total_contacts = 0;
  for (int r = 0; r < repeats; ++r) {
    std::atomic<int> next(0);
    std::atomic<long> ncont(0);
    auto worker = [&]() {
      long local_c = 0;
      for (;;) {
        int p = next.fetch_add(1);
        if (p >= n_pairs) break;
        // Private copies: the CCD overwrites future states, so parallel calls
        // sharing a body need their own Body objects.
        Body a = sc->bodies[pairs[2 * p]];
        Body b = sc->bodies[pairs[2 * p + 1]];
        double mt;
        std::vector<Contact> cc = compareTreesFullContinuous(a, b, max_time[p], mt);
        if (!cc.empty()) (void)pairStepScale(cc, a, b, max_time[p]);
        std::vector<Contact> cs = compareTreesFull(sc->bodies[pairs[2 * p]], sc->bodies[pairs[2 * p + 1]]);
        cs = filterContacts(cs, 0.0);
        local_c += (long)cc.size() + (long)cs.size();
      }
      ncont += local_c;
    };
    auto t0 = std::chrono::steady_clock::now();
    if (threads <= 1) {
      worker();
    } else {
      std::vector<std::thread> ts;
      for (int t = 0; t < threads; ++t) ts.emplace_back(worker);
      for (auto& t : ts) t.join();
    }
    auto t1 = std::chrono::steady_clock::now();
    times.push_back(std::chrono::duration<double>(t1 - t0).count());
    total_contacts = (double)ncont.load();
  }
  std::sort(times.begin(), times.end());
  if (contact_count) *contact_count = total_contacts;
  return times[times.size() / 2];
}


// Applies filterContacts (mode 0) or filterContinuousContacts (mode 1) to the
// scene's last result, in place.  Returns the new number of contacts.
extern "C" int d2o_scene_filter_result(void* h, int mode) {
  Scene* sc = static_cast<Scene*>(h);
  size_t n = sc->out_i.size() / 8;
  std::vector<Contact> cs(n);
  for (size_t i = 0; i < n; ++i) {
    const int32_t* iv = &sc->out_i[8 * i];
    const double* r = &sc->out_r[11 * i];
    Contact& c = cs[i];
    c.A = V(r[0], r[1], r[2]);
    c.B = V(r[3], r[4], r[5]);
    c.eps_a = r[6];
    c.eps_b = r[7];
    c.eps_inner_a = r[8];
    c.eps_inner_b = r[9];
    c.toc = r[10];
    c.node_a = (uint32_t)iv[1];
    c.node_b = (uint32_t)iv[2];
    c.kind = iv[3];
    c.ia = iv[4];
    c.ib = iv[5];
    c.level = iv[6];
  }
  int pair = n ? sc->out_i[0] : 0;
  cs = mode == 0 ? filterContacts(cs, 0.0) : filterContinuousContacts(cs, 0.0);
  sc->out_i.clear();
  sc->out_r.clear();
  exportContacts(*sc, pair, cs);
  sc->out_off.assign({0, (int32_t)cs.size()});
  return (int)cs.size();
}

// The TOI -> step rule (separate_clusters.cpp:496-564) on explicit contacts
// (reals layout of the contact records) for bodies a and b (body records;
// their future states are used for point velocities).
extern "C" double d2o_step_scale(const double* reals, int n, const double* body_a, const double* body_b, double max_time) {
  Body a, b;
  a.future = stateFrom(body_a + 14);
  b.future = stateFrom(body_b + 14);
  for (int k = 0; k < 9; ++k) {
    a.inv_inertia.m[k] = body_a[28 + k];
    b.inv_inertia.m[k] = body_b[28 + k];
  }
  std::vector<Contact> cs(n);
  for (int i = 0; i < n; ++i) {
    const double* r = reals + 11 * i;
    cs[i].A = V(r[0], r[1], r[2]);
    cs[i].B = V(r[3], r[4], r[5]);
    cs[i].eps_a = r[6];
    cs[i].eps_b = r[7];
    cs[i].toc = r[10];
  }
  return pairStepScale(cs, a, b, max_time);
}

// ---------------------------------------------------- broad phase, clusters --
// Body inputs of d2o_broad_phase / d2o_clusters: n body records (40 doubles),
// last states (14 doubles each, state layout; NULL: last = current), per body
// the mesh's float AABB (lo xyz, hi xyz) and bounding radius, and a group per
// body (NULL: one group).
namespace {
std::vector<BpBody> bpBodies(int n, const double* records, const double* last14, const float* aabb6,
                             const double* radius, const int32_t* groups) {
  std::vector<BpBody> bs(n);
  for (int i = 0; i < n; ++i) {
    const double* r = records + 40 * (size_t)i;
    BpBody& b = bs[i];
    b.current = stateFrom(r);
    b.future = stateFrom(r + 14);
    b.last = last14 ? stateFrom(last14 + 14 * (size_t)i) : b.current;
    b.geo_eps = r[37];
    b.is_static = r[38] != 0.0;
    b.last_time_step_size = r[39];
    if (aabb6)
      for (int k = 0; k < 3; ++k) b.mesh.aabb.lo[k] = aabb6[6 * i + k], b.mesh.aabb.hi[k] = aabb6[6 * i + 3 + k];
    b.mesh.radius = radius ? radius[i] : 0.0;
    b.group = groups ? groups[i] : 0;
  }
  return bs;
}
}  // namespace

// MeshData::initBoudingData (model/mesh.cpp:44-51) of one vertex list.
extern "C" void d2o_mesh_bounds(const double* Vp, int nv, float* aabb6, double* radius) {
  std::vector<V> vs(nv);
  for (int i = 0; i < nv; ++i) vs[i] = V(Vp[3 * i], Vp[3 * i + 1], Vp[3 * i + 2]);
  MeshBounds m = meshBounds(vs);
  for (int k = 0; k < 3; ++k) aabb6[k] = m.aabb.lo[k], aabb6[3 + k] = m.aabb.hi[k];
  *radius = m.radius;
}

// broadPhaseEmbree: mode 0 = (particles, t), mode 1 = (particles).  Writes at
// most max_pairs pairs (a < b, sorted) and returns their total count, or -1
// when State::interpolate throws.
extern "C" int d2o_broad_phase(int n, const double* records, const double* last14, const float* aabb6,
                               const double* radius, const int32_t* groups, int mode, double t, int max_pairs,
                               int32_t* pairs_out) {
  std::vector<BpBody> bs = bpBodies(n, records, last14, aabb6, radius, groups);
  std::vector<std::pair<int, int>> ps;
  try {
    ps = broadPhase(bs, mode, t);
  } catch (const std::runtime_error&) {
    return -1;
  }
  for (int i = 0; i < (int)ps.size() && i < max_pairs; ++i) pairs_out[2 * i] = ps[i].first, pairs_out[2 * i + 1] = ps[i].second;
  return (int)ps.size();
}

// separateCollisionClusters over a pair list.  Outputs (arrays sized n_bodies
// or n_pairs): cluster of each body (-1 static) and pair (-1 static-static),
// per cluster its first particle, smallest current time and smallest last
// step; the current states after the roll-back (14 doubles per body).
// Returns the cluster count, or -1 when State::interpolate throws.
extern "C" int d2o_clusters(int n, const double* records, const double* last14, int n_pairs, const int32_t* pairs,
                            int32_t* of_body, int32_t* of_pair, int32_t* first, double* min_time, double* step,
                            double* current14_out) {
  std::vector<BpBody> bs = bpBodies(n, records, last14, nullptr, nullptr, nullptr);
  std::vector<std::pair<int, int>> ps(n_pairs);
  for (int p = 0; p < n_pairs; ++p) ps[p] = {pairs[2 * p], pairs[2 * p + 1]};
  Clusters C;
  try {
    C = separateClusters(bs, ps);
  } catch (const std::runtime_error&) {
    return -1;
  }
  for (int i = 0; i < n; ++i) of_body[i] = C.of_body[i], stateTo(bs[i].current, current14_out + 14 * (size_t)i);
  for (int p = 0; p < n_pairs; ++p) of_pair[p] = C.of_pair[p];
  for (int c = 0; c < C.n; ++c) first[c] = C.first[c], min_time[c] = C.min_current_time[c], step[c] = C.step_size[c];
  return C.n;
}

// State::interpolate (model/state.cpp:55-70) of one state; returns 0, or -1
// when it throws.
extern "C" int d2o_interpolate(const double* cur14, const double* last14, double time, double* out14) {
  try {
    stateTo(interpolate(stateFrom(cur14), stateFrom(last14), time), out14);
  } catch (const std::runtime_error&) {
    return -1;
  }
  return 0;
}
