// d2_kernels.hip -- gfx950 kernels of the contact-detection path and the C ABI
// that drives them (include/d2hip.h).
//
// Work is processed breadth-first over all body pairs at once, one tree level
// per step, so every launch sees a flat, load-balanced list regardless of how
// uneven the pairs are:
//
//   CCD level (compareTreesFullContinuous, full_tree_comparison.h:69-219)
//     soup_cull    one wavefront per soup node pair, lane = triangle pair
//                  (8x8): swept-AABB test, ballot/popcount compaction of the
//                  overlapping triangle pairs into an item list
//     soup_point   one thread per (item, point-triangle sub-test)  6 per item
//     soup_edge    one thread per (item, edge-edge sub-test)       9 per item
//     soup_reduce  one thread per item: the reference's ordered selection of
//                  the 15 sub-results, hit test, leaf-pair contact emission
//     expand       one thread per node pair: children of hit node pairs
//     leaf_enum    one wavefront per leaf-leaf node pair: enumerates its
//                  vertex-triangle and edge-edge sub-tests
//     leaf_point / leaf_edge  one thread per sub-test, hit -> contact
//   Discrete level (compareTreesFull, full_tree_comparison.h:13-67)
//     tri_enum / tri_dist / expand
//   Post-processing: per-pair ordering of contacts by their breadth-first key,
//   the contact filter (contact.h:44-69, libstdc++ std::sort permutation) and
//   the TOI -> step rule (separate_clusters.cpp:496-564).
//
// Contacts are emitted with a 64-bit key (level, list, child-index path,
// sub-test) whose order is the reference's emission order, so allocation can
// use atomics and the order is restored per pair afterwards.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/d2hip.h"
#define D2_HD __host__ __device__ __forceinline__
#include "d2_geom.hpp"

namespace d2 {

// ------------------------------------------------------------------ records
struct NodeRec {
  int32_t id, bucket, inner, depth, parent, child_first, nchild, ntris;
};
struct NP {  // node pair work entry
  uint32_t pair, na, nb, pad;
  unsigned long long key;  // base-8 path of child indices from the root pair
};
struct Item {
  uint32_t np, sub;
};
struct ContactRec {
  unsigned long long key;
  int32_t pair, na, nb, kind, ia, ib, level, pad;
  double r[11];
};
static_assert(sizeof(ContactRec) == 128, "contact record size");

// Pair record (doubles): poses at start/end, the point->triangle frame
// changes, future-state kinematics for the step rule.
enum : int {
  PR_TA0 = 0, PR_TA1 = 16, PR_TB0 = 32, PR_TB1 = 48,
  PR_MAB0 = 64, PR_MAB1 = 80, PR_MBA0 = 96, PR_MBA1 = 112,
  PR_VA = 128, PR_WA = 131, PR_XA = 134, PR_VB = 137, PR_WB = 140, PR_XB = 143,
  PR_GEOA = 146, PR_GEOB = 147, PR_MAXT = 148, PR_SEARCH = 149,
  PR_STRIDE = 152
};
constexpr int BUCKET = 88;
constexpr int RES_SLOTS = 15;  // sub-results per soup triangle pair
enum Counter : int {
  C_SOUP = 0, C_CONN = 1, C_SOUP_NEXT = 2, C_CONN_NEXT = 3, C_ITEMS = 4, C_PT = 5, C_EE = 6,
  C_CONTACTS = 7, C_OVERFLOW = 8, C_LVL_ITEMS = 9, C_COUNT = 16
};
// Key layout: level(5) | list(1) | path(48) | sub(10)
__device__ __forceinline__ unsigned long long make_key(int level, int list, unsigned long long path, int sub) {
  return ((unsigned long long)level << 59) | ((unsigned long long)list << 58) | (path << 10) | (unsigned long long)sub;
}

// std::min / std::max semantics (first argument on ties and NaN).
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ unsigned long long lanes_below() { return (1ull << lane_id()) - 1ull; }

// One atomic per wavefront; returns this lane's slot (or -1 when !pred).
__device__ __forceinline__ long long wave_alloc(bool pred, unsigned long long* counter) {
  unsigned long long m = __ballot(pred);
  long long base = 0;
  int leader = m ? __ffsll((long long)m) - 1 : 0;
  if (m && lane_id() == leader) base = (long long)atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  return pred ? base + __popcll(m & lanes_below()) : -1;
}

__device__ __forceinline__ void atomic_min_pos(unsigned long long* addr, double v) {
  // Non-negative doubles order like their bit patterns.
  atomicMin(addr, (unsigned long long)__double_as_longlong(v));
}

template <class R> __device__ __forceinline__ tri<R> load_tri(const double* bucket, int k) {
  return tri<R>{{(R)bucket[0 * 8 + k], (R)bucket[1 * 8 + k], (R)bucket[2 * 8 + k]},
                {(R)bucket[3 * 8 + k], (R)bucket[4 * 8 + k], (R)bucket[5 * 8 + k]},
                {(R)bucket[6 * 8 + k], (R)bucket[7 * 8 + k], (R)bucket[8 * 8 + k]}};
}

struct Dev {
  // trees
  const NodeRec* nodes;
  const double* buckets;
  // pairs
  const int32_t* pair_info;  // a, b, node offset a, node offset b
  const double* pair_rec;
  // work lists
  NP* soup;
  NP* conn;
  NP* soup_next;
  NP* conn_next;
  Item* items;
  Item* pt_items;
  Item* ee_items;
  double* res;  // [items][15][8]
  int* np_hit;
  ContactRec* contacts;
  unsigned long long* cnt;
  unsigned long long* pass_min;  // per pair, min toc of this pass (bits)
  int* pass_flag;                // per pair, a leaf pair had hits this pass
  long long cap_np, cap_items, cap_contacts;
  int level;
  int discrete;
};

// ----------------------------------------------------------------- set-up
// Per pair: poses, frame changes, kinematics (CCD: future = extrapolate(max_time)).
__global__ void k_setup(int n_pairs, const int32_t* pairs, const double* max_time, const double* bodies,
                        const int32_t* body_tree, const int32_t* tree_off, int32_t* pair_info, double* pair_rec,
                        int ccd) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    int a = pairs[2 * p], b = pairs[2 * p + 1];
    const double* ra = bodies + 40 * (size_t)a;
    const double* rb = bodies + 40 * (size_t)b;
    body_state ca = load_state(ra), cb = load_state(rb);
    body_state fa, fb;
    double mt = ccd ? max_time[p] : 0.0;
    if (ccd) {
      fa = extrapolate(ca, mt, ra + 28);
      fb = extrapolate(cb, mt, rb + 28);
    } else {
      fa = load_state(ra + 14);
      fb = load_state(rb + 14);
    }
    double* pr = pair_rec + (size_t)PR_STRIDE * p;
    pose(ca.q, ca.t, pr + PR_TA0);
    pose(fa.q, fa.t, pr + PR_TA1);
    pose(cb.q, cb.t, pr + PR_TB0);
    pose(fb.q, fb.t, pr + PR_TB1);
    if (ccd) {
      double inv[16];
      inv44(pr + PR_TB0, inv);
      mul44(inv, pr + PR_TA0, pr + PR_MAB0);
      inv44(pr + PR_TB1, inv);
      mul44(inv, pr + PR_TA1, pr + PR_MAB1);
      inv44(pr + PR_TA0, inv);
      mul44(inv, pr + PR_TB0, pr + PR_MBA0);
      inv44(pr + PR_TA1, inv);
      mul44(inv, pr + PR_TB1, pr + PR_MBA1);
      double Iw[9];
      world_inv_inertia(fa.q, ra + 28, Iw);
      v3<double> wa = mul33v(Iw, fa.L);
      world_inv_inertia(fb.q, rb + 28, Iw);
      v3<double> wb = mul33v(Iw, fb.L);
      pr[PR_VA] = fa.v.x, pr[PR_VA + 1] = fa.v.y, pr[PR_VA + 2] = fa.v.z;
      pr[PR_WA] = wa.x, pr[PR_WA + 1] = wa.y, pr[PR_WA + 2] = wa.z;
      pr[PR_XA] = fa.t.x, pr[PR_XA + 1] = fa.t.y, pr[PR_XA + 2] = fa.t.z;
      pr[PR_VB] = fb.v.x, pr[PR_VB + 1] = fb.v.y, pr[PR_VB + 2] = fb.v.z;
      pr[PR_WB] = wb.x, pr[PR_WB + 1] = wb.y, pr[PR_WB + 2] = wb.z;
      pr[PR_XB] = fb.t.x, pr[PR_XB + 1] = fb.t.y, pr[PR_XB + 2] = fb.t.z;
    }
    pr[PR_GEOA] = ra[37];
    pr[PR_GEOB] = rb[37];
    pr[PR_MAXT] = mt;
    pr[PR_SEARCH] = (double)(float)(ra[37] + rb[37]);  // a float in the reference
    pair_info[4 * p + 0] = a;
    pair_info[4 * p + 1] = b;
    pair_info[4 * p + 2] = tree_off[body_tree[a]];
    pair_info[4 * p + 3] = tree_off[body_tree[b]];
  }
}

__global__ void k_roots(int n_pairs, NP* soup, unsigned long long* cnt, unsigned long long* pass_min, int* pass_flag,
                        double* min_time, const double* pair_rec, double* scale) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    soup[p] = NP{(uint32_t)p, 0u, 0u, 0u, 0ull};
    pass_min[p] = ~0ull;
    pass_flag[p] = 0;
    min_time[p] = pair_rec[(size_t)PR_STRIDE * p + PR_MAXT];
    scale[p] = 1.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int i = 0; i < C_COUNT; ++i) cnt[i] = 0;
    cnt[C_SOUP] = (unsigned long long)n_pairs;
  }
}

// Per pass: fold "min toc of every hit of this pass" into min_time when a
// leaf-leaf node pair of the pair had hits (full_tree_comparison.h:106-108,167-169).
__global__ void k_fold_min(int n_pairs, unsigned long long* pass_min, int* pass_flag, double* min_time,
                           const double* pair_rec) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    if (pass_flag[p] && pass_min[p] != ~0ull) {
      double t = __longlong_as_double((long long)pass_min[p]) * pair_rec[(size_t)PR_STRIDE * p + PR_MAXT];
      if (t < min_time[p]) min_time[p] = t;
    }
    pass_min[p] = ~0ull;
    pass_flag[p] = 0;
  }
}

// --------------------------------------------------------- CCD soup level
__global__ void __launch_bounds__(256) k_soup_cull(Dev D) {
  const long long n = (long long)D.cnt[C_SOUP];
  const int lane = lane_id();
  const long long wstride = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < n; w += wstride) {
    NP np = D.soup[w];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    int i = lane >> 3, j = lane & 7;
    bool ok = i < na.ntris && j < nb.ntris;
    bool hit = false;
    if (ok) {
      const double* ba = D.buckets + (size_t)BUCKET * (pi[2] + na.bucket);
      const double* bb = D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket);
      tri<double> ta = load_tri<double>(ba, i), tb = load_tri<double>(bb, j);
      // bbox(start) expanded by bbox(end), then by geo_eps (continuous_comparison.h:804-812)
      double lo[3], hi[3], lo2[3], hi2[3];
      auto box = [](const tri<double>& t, const double* m, double* l, double* h) {
        tri<double> w = xform(m, t);
        l[0] = smin(smin(w.a.x, w.b.x), w.c.x), h[0] = smax(smax(w.a.x, w.b.x), w.c.x);
        l[1] = smin(smin(w.a.y, w.b.y), w.c.y), h[1] = smax(smax(w.a.y, w.b.y), w.c.y);
        l[2] = smin(smin(w.a.z, w.b.z), w.c.z), h[2] = smax(smax(w.a.z, w.b.z), w.c.z);
      };
      double ga = pr[PR_GEOA], gb = pr[PR_GEOB];
      box(ta, pr + PR_TA0, lo, hi);
      box(ta, pr + PR_TA1, lo2, hi2);
      double al[3], ah[3], bl[3], bh[3];
      for (int k = 0; k < 3; ++k) al[k] = smin(lo[k], lo2[k]) - ga, ah[k] = smax(hi[k], hi2[k]) + ga;
      box(tb, pr + PR_TB0, lo, hi);
      box(tb, pr + PR_TB1, lo2, hi2);
      for (int k = 0; k < 3; ++k) bl[k] = smin(lo[k], lo2[k]) - gb, bh[k] = smax(hi[k], hi2[k]) + gb;
      hit = true;
      for (int k = 0; k < 3; ++k) {  // bbox.h:8-23, a is "this"
        bool o = (bl[k] < al[k]) ? (bh[k] > al[k]) : (bl[k] < ah[k]);
        hit = hit && o;
      }
    }
    long long slot = wave_alloc(hit, &D.cnt[C_ITEMS]);
    if (hit) {
      if (slot < D.cap_items) D.items[slot] = Item{(uint32_t)w, (uint32_t)lane};
      else D.cnt[C_OVERFLOW] = 1;
    }
    if (lane == 0) D.np_hit[w] = 0;
  }
}

__device__ __forceinline__ void store_res(double* r, double d, v3<double> P, v3<double> Q, double toc) {
  r[0] = d, r[1] = toc, r[2] = P.x, r[3] = P.y, r[4] = P.z, r[5] = Q.x, r[6] = Q.y, r[7] = Q.z;
}

// Point-triangle sub-tests of soup items; slot 5k+dir (k: vertex index,
// dir 0: vertex of a vs triangle of b, dir 1: vertex of b vs triangle of a).
__global__ void __launch_bounds__(256) k_soup_point(Dev D, long long first, long long count) {
  const long long total = count * 6;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    long long it = first + t / 6;
    int s = (int)(t % 6), k = s >> 1, dir = s & 1;
    Item item = D.items[it];
    NP np = D.soup[item.np];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    int i = item.sub >> 3, j = item.sub & 7;
    tri<double> ta = load_tri<double>(D.buckets + (size_t)BUCKET * (pi[2] + na.bucket), i);
    tri<double> tb = load_tri<double>(D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket), j);
    v3<double> P, Q;
    double toc = 1.0, d;
    if (dir == 0) d = point_tri(corner(ta, k), pr + PR_MAB0, pr + PR_MAB1, tb, pr + PR_TB0, pr + PR_TB1, P, Q, toc);
    else d = point_tri(corner(tb, k), pr + PR_MBA0, pr + PR_MBA1, ta, pr + PR_TA0, pr + PR_TA1, P, Q, toc);
    store_res(D.res + ((size_t)(it - first) * RES_SLOTS + 5 * k + dir) * 8, d, P, Q, toc);
  }
}

// Edge-edge sub-tests of soup items; slot 5k+2+l (edge k of a, edge l of b).
__global__ void __launch_bounds__(256) k_soup_edge(Dev D, long long first, long long count) {
  const long long total = count * 9;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    long long it = first + t / 9;
    int s = (int)(t % 9), k = s / 3, l = s % 3;
    Item item = D.items[it];
    NP np = D.soup[item.np];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    int i = item.sub >> 3, j = item.sub & 7;
    tri<double> ta = load_tri<double>(D.buckets + (size_t)BUCKET * (pi[2] + na.bucket), i);
    tri<double> tb = load_tri<double>(D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket), j);
    v3<double> a0 = corner(ta, k), a1 = corner(ta, (k + 1) % 3);
    v3<double> b0 = corner(tb, l), b1 = corner(tb, (l + 1) % 3);
    v3<double> P, Q;
    double toc = 1.0;
    double d = edge_edge(xform(pr + PR_TA0, a0), xform(pr + PR_TA0, a1), xform(pr + PR_TA1, a0), xform(pr + PR_TA1, a1),
                         xform(pr + PR_TB0, b0), xform(pr + PR_TB0, b1), xform(pr + PR_TB1, b0), xform(pr + PR_TB1, b1),
                         P, Q, toc);
    store_res(D.res + ((size_t)(it - first) * RES_SLOTS + 5 * k + 2 + l) * 8, d, P, Q, toc);
  }
}

// triTriCCDLinear's selection over its 15 sub-results in the reference order
// (continuous_comparison.h:680-779), hit test (continuous_comparison.h:817),
// contact emission for leaf-leaf node pairs.
__global__ void __launch_bounds__(256) k_soup_reduce(Dev D, long long first, long long count) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long long)gridDim.x * blockDim.x) {
    long long it = first + t;
    Item item = D.items[it];
    const double* r = D.res + (size_t)t * RES_SLOTS * 8;
    double best = 1.0 / 0.0, toc = 1.0;
    int win = -1;
    for (int s = 0; s < RES_SLOTS; ++s) {
      double d = r[8 * s], tc = r[8 * s + 1];
      if (d < best || (d == best && tc < toc)) best = d, toc = tc, win = s;
    }
    NP np = D.soup[item.np];
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    bool hit = best < pr[PR_SEARCH];
    if (!hit) continue;
    atomicOr(&D.np_hit[item.np], 1);
    atomic_min_pos(&D.pass_min[np.pair], toc);
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    if (na.inner || nb.inner) continue;
    D.pass_flag[np.pair] = 1;
    unsigned long long slot = atomicAdd(&D.cnt[C_CONTACTS], 1ull);
    if ((long long)slot >= D.cap_contacts) {
      D.cnt[C_OVERFLOW] = 1;
      continue;
    }
    ContactRec c;
    c.key = make_key(2 * D.level, 0, np.key, (int)item.sub);
    c.pair = (int32_t)np.pair, c.na = (int32_t)np.na, c.nb = (int32_t)np.nb, c.kind = 0;
    c.ia = (int32_t)(item.sub >> 3), c.ib = (int32_t)(item.sub & 7), c.level = 2 * D.level, c.pad = 0;
    const double* w = r + 8 * win;
    c.r[0] = w[2], c.r[1] = w[3], c.r[2] = w[4], c.r[3] = w[5], c.r[4] = w[6], c.r[5] = w[7];
    c.r[6] = pr[PR_GEOA], c.r[7] = pr[PR_GEOB], c.r[8] = 0.0, c.r[9] = 0.0, c.r[10] = toc;
    D.contacts[slot] = c;
  }
}

// ---------------------------------------------------- CCD leaf-leaf level
// Sub-test numbering of a leaf pair with n (a) and m (b) triangles, in the
// reference's order (continuous_comparison.h:855-907):
//   [0, 3nm)        vertex v = s/m of a   vs triangle s%m of b
//   [3nm, 6nm)      vertex v = s'/n of b  vs triangle s'%n of a
//   edge list       edge s/(3m) of a      vs edge s%(3m) of b   (offset 6nm)
__global__ void __launch_bounds__(256) k_leaf_enum(Dev D) {
  const long long n = (long long)D.cnt[C_CONN];
  const int lane = lane_id();
  const long long wstride = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < n; w += wstride) {
    NP np = D.conn[w];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    int na = D.nodes[pi[2] + np.na].ntris, nb = D.nodes[pi[3] + np.nb].ntris;
    int npt = 6 * na * nb, nee = 9 * na * nb;
    long long bpt = 0, bee = 0;
    if (lane == 0) {
      bpt = (long long)atomicAdd(&D.cnt[C_PT], (unsigned long long)npt);
      bee = (long long)atomicAdd(&D.cnt[C_EE], (unsigned long long)nee);
    }
    bpt = __shfl(bpt, 0);
    bee = __shfl(bee, 0);
    if (bpt + npt > D.cap_items || bee + nee > D.cap_items) {
      if (lane == 0) D.cnt[C_OVERFLOW] = 1;
      continue;
    }
    for (int s = lane; s < npt; s += 64) D.pt_items[bpt + s] = Item{(uint32_t)w, (uint32_t)s};
    for (int s = lane; s < nee; s += 64) D.ee_items[bee + s] = Item{(uint32_t)w, (uint32_t)s};
    if (lane == 0) D.np_hit[w] = 0;
  }
}

__device__ __forceinline__ void emit_leaf(Dev& D, const NP& np, const double* pr, int kind, int ia, int ib, int sub,
                                          v3<double> A, v3<double> B, double toc) {
  unsigned long long slot = atomicAdd(&D.cnt[C_CONTACTS], 1ull);
  if ((long long)slot >= D.cap_contacts) {
    D.cnt[C_OVERFLOW] = 1;
    return;
  }
  ContactRec c;
  c.key = make_key(2 * D.level + 1, 0, np.key, sub);
  c.pair = (int32_t)np.pair, c.na = (int32_t)np.na, c.nb = (int32_t)np.nb, c.kind = kind;
  c.ia = ia, c.ib = ib, c.level = 2 * D.level + 1, c.pad = 0;
  c.r[0] = A.x, c.r[1] = A.y, c.r[2] = A.z, c.r[3] = B.x, c.r[4] = B.y, c.r[5] = B.z;
  c.r[6] = pr[PR_GEOA], c.r[7] = pr[PR_GEOB], c.r[8] = 0.0, c.r[9] = 0.0, c.r[10] = toc;
  D.contacts[slot] = c;
}

__global__ void __launch_bounds__(256) k_leaf_point(Dev D, long long first, long long count) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long long)gridDim.x * blockDim.x) {
    Item item = D.pt_items[first + t];
    NP np = D.conn[item.np];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    const double* ba = D.buckets + (size_t)BUCKET * (pi[2] + na.bucket);
    const double* bb = D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket);
    int s = (int)item.sub, n = na.ntris, m = nb.ntris;
    v3<double> P, Q;
    double toc = 1.0, d;
    int kind, ia, ib;
    if (s < 3 * n * m) {
      int v = s / m, j = s % m;
      tri<double> ta = load_tri<double>(ba, v / 3), tb = load_tri<double>(bb, j);
      d = point_tri(corner(ta, v % 3), pr + PR_MAB0, pr + PR_MAB1, tb, pr + PR_TB0, pr + PR_TB1, P, Q, toc);
      kind = 1, ia = v, ib = j;
    } else {
      int s2 = s - 3 * n * m, v = s2 / n, i = s2 % n;
      tri<double> tb = load_tri<double>(bb, v / 3), ta = load_tri<double>(ba, i);
      d = point_tri(corner(tb, v % 3), pr + PR_MBA0, pr + PR_MBA1, ta, pr + PR_TA0, pr + PR_TA1, P, Q, toc);
      kind = 2, ia = i, ib = v;
    }
    if (d < pr[PR_SEARCH]) {
      D.np_hit[item.np] = 1;
      D.pass_flag[np.pair] = 1;
      atomic_min_pos(&D.pass_min[np.pair], toc);
      if (kind == 1) emit_leaf(D, np, pr, kind, ia, ib, s, P, Q, toc);
      else emit_leaf(D, np, pr, kind, ia, ib, s, Q, P, toc);  // continuous_comparison.h:876
    }
  }
}

__global__ void __launch_bounds__(256) k_leaf_edge(Dev D, long long first, long long count) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long long)gridDim.x * blockDim.x) {
    Item item = D.ee_items[first + t];
    NP np = D.conn[item.np];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    int s = (int)item.sub, n = na.ntris, m = nb.ntris;
    int ea = s / (3 * m), eb = s % (3 * m);
    tri<double> ta = load_tri<double>(D.buckets + (size_t)BUCKET * (pi[2] + na.bucket), ea / 3);
    tri<double> tb = load_tri<double>(D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket), eb / 3);
    v3<double> a0 = corner(ta, ea % 3), a1 = corner(ta, (ea % 3 + 1) % 3);
    v3<double> b0 = corner(tb, eb % 3), b1 = corner(tb, (eb % 3 + 1) % 3);
    v3<double> P, Q;
    double toc = 1.0;
    double d = edge_edge(xform(pr + PR_TA0, a0), xform(pr + PR_TA0, a1), xform(pr + PR_TA1, a0), xform(pr + PR_TA1, a1),
                         xform(pr + PR_TB0, b0), xform(pr + PR_TB0, b1), xform(pr + PR_TB1, b0), xform(pr + PR_TB1, b1),
                         P, Q, toc);
    if (d < pr[PR_SEARCH]) {
      D.np_hit[item.np] = 1;
      D.pass_flag[np.pair] = 1;
      atomic_min_pos(&D.pass_min[np.pair], toc);
      emit_leaf(D, np, pr, 3, ea, eb, 6 * n * m + s, Q, P, toc);  // continuous_comparison.h:903
    }
  }
}

// ------------------------------------------- ordered enumeration/expansion
// ------------------------------------------------ FP32 conservative cull
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
__device__ __forceinline__ float f3dot(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
__device__ int cull_class(const double* pa, const double* pb, const tri<double>& la, const tri<double>& lt, double thr_d) {
  const float EPSF = 5.9604645e-08f;
  float ra[9], rb[9], ta[3];
  for (int r = 0; r < 3; ++r) {
    for (int k = 0; k < 3; ++k) ra[3 * r + k] = (float)pa[4 * r + k], rb[3 * r + k] = (float)pb[4 * r + k];
    ta[r] = (float)(pa[4 * r + 3] - pb[4 * r + 3]);
  }
  float A[3][3], B[3][3];
  const v3<double> av[3] = {la.a, la.b, la.c}, bv[3] = {lt.a, lt.b, lt.c};
  float coord = 0.f;
  for (int v = 0; v < 3; ++v) {
    float x = (float)av[v].x, y = (float)av[v].y, z = (float)av[v].z;
    float u = (float)bv[v].x, w = (float)bv[v].y, q = (float)bv[v].z;
    for (int r = 0; r < 3; ++r) {
      A[v][r] = ((ra[3 * r] * x + ra[3 * r + 1] * y) + ra[3 * r + 2] * z) + ta[r];
      B[v][r] = (rb[3 * r] * u + rb[3 * r + 1] * w) + rb[3 * r + 2] * q;
      coord = fmaxf(coord, fmaxf(fabsf(A[v][r]), fabsf(B[v][r])));
    }
  }
  float ea0[3], ea1[3], eb0[3], eb1[3], nA[3], nB[3], cd[3];
  for (int r = 0; r < 3; ++r) {
    ea0[r] = A[1][r] - A[0][r], ea1[r] = A[2][r] - A[0][r];
    eb0[r] = B[1][r] - B[0][r], eb1[r] = B[2][r] - B[0][r];
    cd[r] = (B[0][r] + B[1][r] + B[2][r]) - (A[0][r] + A[1][r] + A[2][r]);
  }
  auto crs = [](const float* a, const float* b, float* o) {
    o[0] = a[1] * b[2] - a[2] * b[1], o[1] = a[2] * b[0] - a[0] * b[2], o[2] = a[0] * b[1] - a[1] * b[0];
  };
  crs(ea0, ea1, nA);
  crs(eb0, eb1, nB);
  float d00 = f3dot(ea0, ea0), d11 = f3dot(ea1, ea1), d01 = f3dot(ea0, ea1);
  float e00 = f3dot(eb0, eb0), e11 = f3dot(eb1, eb1), e01 = f3dot(eb0, eb1);
  float dena = d00 * d11 - d01 * d01, denb = e00 * e11 - e01 * e01;
  // ill-conditioned barycentric systems get no cull
  if (!(dena > 1e-4f * d00 * d11) || !(denb > 1e-4f * e00 * e11)) return 1;
  float kappa = fmaxf(d00 * d11 / dena, e00 * e11 / denb);
  float L2 = fmaxf(fmaxf(d00, d11), fmaxf(e00, e11));
  float e12[3], f12[3];
  for (int r = 0; r < 3; ++r) e12[r] = A[2][r] - A[1][r], f12[r] = B[2][r] - B[1][r];
  L2 = fmaxf(L2, fmaxf(f3dot(e12, e12), f3dot(f12, f12)));
  float L = sqrtf(L2);
  float thr = (float)thr_d;
  float margin = 1e-6f + 64.f * EPSF * kappa * L + 16.f * EPSF * coord + EPSF * thr;
  // lower bound
  float lb = 0.f;
  const float* axes[3] = {nA, nB, cd};
  for (int k = 0; k < 3; ++k) {
    const float* n = axes[k];
    float nn = f3dot(n, n);
    if (!(nn > 0.f)) continue;
    float a0 = f3dot(A[0], n), a1 = f3dot(A[1], n), a2 = f3dot(A[2], n);
    float b0 = f3dot(B[0], n), b1 = f3dot(B[1], n), b2 = f3dot(B[2], n);
    float amin = fminf(a0, fminf(a1, a2)), amax = fmaxf(a0, fmaxf(a1, a2));
    float bmin = fminf(b0, fminf(b1, b2)), bmax = fmaxf(b0, fmaxf(b1, b2));
    float g = fmaxf(bmin - amax, amin - bmax);
    if (g > 0.f) lb = fmaxf(lb, g / sqrtf(nn));
  }
  if (lb > thr + margin) return 0;
  // upper bound
  float dd[3][3];
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) {
      float t[3] = {A[x][0] - B[y][0], A[x][1] - B[y][1], A[x][2] - B[y][2]};
      dd[x][y] = f3dot(t, t);
    }
  float ub2 = 3.4e38f;
  for (int x = 0; x < 3; ++x)
    for (int y = 0; y < 3; ++y) {
      int x1 = x == 2 ? 0 : x + 1, y1 = y == 2 ? 0 : y + 1;
      ub2 = fminf(ub2, fmaxf(fmaxf(dd[x][y], dd[x][y1]), fmaxf(dd[x1][y], dd[x1][y1])));
    }
  float lim = thr - margin;
  if (lim > 0.f && ub2 < lim * lim) return 2;
  return 1;
}

// One wavefront per node pair: which triangle pairs need the FP64 test.
// FP64 mode keeps every valid pair; FP32 mode drops certain misses, and for
// node pairs with an inner node (they emit no contacts, only "any hit"
// matters) a certain hit settles the node pair without any FP64 work.
__global__ void __launch_bounds__(256) k_tri_mask(Dev D, long long n, int fp32, unsigned long long* masks, int* counts) {
  const int lane = lane_id();
  if (blockIdx.x == 0 && threadIdx.x == 0) counts[n] = 0;
  const long long wstride = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < n; w += wstride) {
    NP np = D.soup[w];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    int i = lane >> 3, j = lane & 7;
    bool valid = i < na.ntris && j < nb.ntris;
    int cls = valid ? 1 : 0;
    if (fp32 && valid) {
      const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
      const double* ba = D.buckets + (size_t)BUCKET * (pi[2] + na.bucket);
      const double* bb = D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket);
      double thr = (pr[PR_SEARCH] + ba[72 + i]) + bb[72 + j];
      cls = cull_class(pr + PR_TA1, pr + PR_TB1, load_tri<double>(ba, i), load_tri<double>(bb, j), thr);
    }
    bool inner = na.inner || nb.inner;
    unsigned long long keep = __ballot(cls == 1 || (cls == 2 && !inner));
    bool settled = inner && __ballot(cls == 2) != 0ull;
    if (settled) keep = 0ull;
    if (lane == 0) {
      masks[w] = keep;
      counts[w] = __popcll(keep);
      D.np_hit[w] = settled ? 1 : 0;
    }
  }
}

// One wavefront per node pair: kept triangle pairs to the item list, at the
// node pair's scanned offset, in triangle-pair order.
__global__ void __launch_bounds__(256) k_tri_enum_masked(Dev D, long long n, const unsigned long long* masks,
                                                          const int* offs) {
  const int lane = lane_id();
  const long long wstride = (long long)gridDim.x * (blockDim.x >> 6);
  for (long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < n; w += wstride) {
    unsigned long long m = masks[w];
    long long base = offs[w];
    if ((m >> lane) & 1ull) {
      long long slot = base + __popcll(m & lanes_below());
      if (slot < D.cap_items) D.items[slot] = Item{(uint32_t)w, (uint32_t)lane};
      else D.cnt[C_OVERFLOW] = 1;
    }
    if (lane == 0 && w == n - 1) D.cnt[C_LVL_ITEMS] = (unsigned long long)offs[n];
  }
}

// Grid-stride over a device-resident item count (no host round trip).
__global__ void __launch_bounds__(256) k_tri_dist_dev(Dev D, const int* n_items_dev) {
  long long count = (long long)*n_items_dev;
  if (count > D.cap_items) count = D.cap_items;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long long)gridDim.x * blockDim.x) {
    Item item = D.items[t];
    NP np = D.soup[item.np];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    const double* ba = D.buckets + (size_t)BUCKET * (pi[2] + na.bucket);
    const double* bb = D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket);
    int i = item.sub >> 3, j = item.sub & 7;
    tri<double> A = xform(pr + PR_TA1, load_tri<double>(ba, i));
    tri<double> B = xform(pr + PR_TB1, load_tri<double>(bb, j));
    v3<double> P, Q;
    double d = tri_tri_dist(A, B, P, Q);
    double ea = ba[72 + i], eb = bb[72 + j];
    if (!(d < (pr[PR_SEARCH] + ea) + eb)) continue;
    D.np_hit[item.np] = 1;
    if (na.inner || nb.inner) continue;
    unsigned long long slot = atomicAdd(&D.cnt[C_CONTACTS], 1ull);
    if ((long long)slot >= D.cap_contacts) {
      D.cnt[C_OVERFLOW] = 1;
      continue;
    }
    ContactRec c;
    c.key = make_key(D.level, 0, np.key, (int)item.sub);
    c.pair = (int32_t)np.pair, c.na = (int32_t)np.na, c.nb = (int32_t)np.nb, c.kind = 0;
    c.ia = i, c.ib = j, c.level = D.level, c.pad = 0;
    c.r[0] = P.x, c.r[1] = P.y, c.r[2] = P.z, c.r[3] = Q.x, c.r[4] = Q.y, c.r[5] = Q.z;
    c.r[6] = pr[PR_GEOA] + ea, c.r[7] = pr[PR_GEOB] + eb, c.r[8] = ba[80 + i], c.r[9] = bb[80 + j], c.r[10] = 0.0;
    D.contacts[slot] = c;
  }
}

// Children of hit node pairs (full_tree_comparison.h:26-62, 99-150), in two
// passes: counts per destination list, then writes at scanned offsets, which
// keeps every list in the reference's breadth-first order.
__device__ __forceinline__ bool expand_side(const NodeRec& na, const NodeRec& nb) {
  return na.inner && (!nb.inner || na.depth < nb.depth);  // true: advance a
}
__global__ void k_expand_count(Dev D, long long n, int from_soup, int* cs, int* cc) {
  const NP* list = from_soup ? D.soup : D.conn;
  for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w <= n; w += (long long)gridDim.x * blockDim.x) {
    int ns = 0, nc = 0;
    if (w < n && D.np_hit[w]) {
      NP np = list[w];
      const int32_t* pi = D.pair_info + 4 * np.pair;
      const NodeRec na = D.nodes[pi[2] + np.na];
      const NodeRec nb = D.nodes[pi[3] + np.nb];
      if (na.inner || nb.inner) {
        bool adv_a = expand_side(na, nb);
        const NodeRec& par = adv_a ? na : nb;
        for (int c = 0; c < par.nchild; ++c) {
          bool to_soup = true;
          if (!D.discrete) {
            const NodeRec x = adv_a ? D.nodes[pi[2] + par.child_first + c] : na;
            const NodeRec y = adv_a ? nb : D.nodes[pi[3] + par.child_first + c];
            to_soup = x.inner || y.inner;
          }
          ns += to_soup, nc += !to_soup;
        }
      }
    }
    cs[w] = ns;
    cc[w] = nc;
  }
}
__global__ void k_expand_write(Dev D, long long n, int from_soup, const int* os, const int* oc) {
  const NP* list = from_soup ? D.soup : D.conn;
  for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < n; w += (long long)gridDim.x * blockDim.x) {
    if (!D.np_hit[w]) continue;
    NP np = list[w];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    if (!na.inner && !nb.inner) continue;
    bool adv_a = expand_side(na, nb);
    const NodeRec& par = adv_a ? na : nb;
    long long ps = os[w], pc = oc[w];
    for (int c = 0; c < par.nchild; ++c) {
      uint32_t ca = adv_a ? (uint32_t)(par.child_first + c) : np.na;
      uint32_t cb = adv_a ? np.nb : (uint32_t)(par.child_first + c);
      NP child{np.pair, ca, cb, 0u, np.key * 8ull + (unsigned long long)c};
      bool to_soup = true;
      if (!D.discrete) to_soup = D.nodes[pi[2] + ca].inner || D.nodes[pi[3] + cb].inner;
      long long slot = to_soup ? ps++ : pc++;
      if (slot < D.cap_np) (to_soup ? D.soup_next : D.conn_next)[slot] = child;
      else D.cnt[C_OVERFLOW] = 1;
    }
  }
}
// Next-level sizes from the scanned totals.
__global__ void k_advance_scanned(unsigned long long* cnt, const int* os_total, const int* oc_total, int have_conn) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    cnt[C_SOUP] = os_total ? (unsigned long long)*os_total : 0ull;
    cnt[C_CONN] = (have_conn && oc_total) ? (unsigned long long)*oc_total : 0ull;
    cnt[C_SOUP_NEXT] = 0;
    cnt[C_CONN_NEXT] = 0;
    cnt[C_ITEMS] = 0;
    cnt[C_PT] = 0;
    cnt[C_EE] = 0;
  }
}

// -------------------------------------------------------- post-processing
__global__ void k_count(long long n, const ContactRec* c, int* pair_cnt) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    atomicAdd(&pair_cnt[c[t].pair], 1);
}
__global__ void k_scatter(long long n, const ContactRec* c, const int* pair_off, int* pair_fill, int* perm) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    int p = c[t].pair;
    int k = atomicAdd(&pair_fill[p], 1);
    perm[pair_off[p] + k] = (int)t;
  }
}

// contact.h:35-43
__device__ __forceinline__ bool contact_less(const ContactRec& x, const ContactRec& y) {
  double xe = x.r[6] + x.r[7], ye = y.r[6] + y.r[7];
  if (xe == ye) {
    double dx = vlen(sub(v3<double>{x.r[0], x.r[1], x.r[2]}, v3<double>{x.r[3], x.r[4], x.r[5]}));
    double dy = vlen(sub(v3<double>{y.r[0], y.r[1], y.r[2]}, v3<double>{y.r[3], y.r[4], y.r[5]}));
    return dx < dy;
  }
  return xe < ye;
}

// The permutation libstdc++'s std::sort produces (introsort: median-of-three
// pivot to the front, unguarded Hoare partition, depth limit 2*floor(log2 n),
// heapsort fallback, insertion sort threshold 16).  Reproducing it keeps the
// contact filter's choices among exactly tied contacts identical.
struct Sorter {
  int* a;
  const ContactRec* c;
  __device__ bool lt(int x, int y) const { return contact_less(c[x], c[y]); }
  __device__ void swp(int i, int j) {
    int t = a[i];
    a[i] = a[j];
    a[j] = t;
  }
  __device__ void median_to_first(int res, int x, int y, int z) {
    if (lt(a[x], a[y])) {
      if (lt(a[y], a[z])) swp(res, y);
      else if (lt(a[x], a[z])) swp(res, z);
      else swp(res, x);
    } else if (lt(a[x], a[z])) swp(res, x);
    else if (lt(a[y], a[z])) swp(res, z);
    else swp(res, y);
  }
  __device__ int partition(int first, int last, int pivot) {
    while (true) {
      while (lt(a[first], a[pivot])) ++first;
      --last;
      while (lt(a[pivot], a[last])) --last;
      if (!(first < last)) return first;
      swp(first, last);
      ++first;
    }
  }
  __device__ void push_heap(int first, int hole, int top, int val) {
    int parent = (hole - 1) / 2;
    while (hole > top && lt(a[first + parent], val)) {
      a[first + hole] = a[first + parent];
      hole = parent;
      parent = (hole - 1) / 2;
    }
    a[first + hole] = val;
  }
  __device__ void adjust_heap(int first, int hole, int len, int val) {
    int top = hole, child = hole;
    while (child < (len - 1) / 2) {
      child = 2 * (child + 1);
      if (lt(a[first + child], a[first + child - 1])) child--;
      a[first + hole] = a[first + child];
      hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
      child = 2 * (child + 1);
      a[first + hole] = a[first + child - 1];
      hole = child - 1;
    }
    push_heap(first, hole, top, val);
  }
  __device__ void heap_sort(int first, int last) {
    int len = last - first;
    if (len >= 2) {
      for (int parent = (len - 2) / 2;; --parent) {
        adjust_heap(first, parent, len, a[first + parent]);
        if (parent == 0) break;
      }
    }
    while (last - first > 1) {
      --last;
      int v = a[last];
      a[last] = a[first];
      adjust_heap(first, 0, last - first, v);
    }
  }
  __device__ void linear_insert(int last) {
    int v = a[last], next = last - 1;
    while (lt(v, a[next])) {
      a[last] = a[next];
      last = next;
      --next;
    }
    a[last] = v;
  }
  __device__ void insertion(int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
      if (lt(a[i], a[first])) {
        int v = a[i];
        for (int k = i; k > first; --k) a[k] = a[k - 1];
        a[first] = v;
      } else {
        linear_insert(i);
      }
    }
  }
  __device__ void sort(int first, int last) {
    if (first == last) return;
    int n = last - first;
    int depth = 2 * (63 - __clzll((long long)n));
    int st_f[64], st_l[64], st_d[64], sp = 0;
    st_f[sp] = first, st_l[sp] = last, st_d[sp] = depth, ++sp;
    while (sp) {
      --sp;
      int f = st_f[sp], l = st_l[sp], d = st_d[sp];
      while (l - f > 16) {
        if (d == 0) {
          heap_sort(f, l);
          break;
        }
        --d;
        int mid = f + (l - f) / 2;
        median_to_first(f, f + 1, mid, l - 1);
        int cut = partition(f + 1, l, f);
        st_f[sp] = cut, st_l[sp] = l, st_d[sp] = d, ++sp;  // processed independently
        l = cut;
      }
    }
    if (n > 16) {
      insertion(first, first + 16);
      for (int i = first + 16; i != last; ++i) linear_insert(i);
    } else {
      insertion(first, last);
    }
  }
};

// Per pair: restore the breadth-first order (insertion sort on the key), then
// either the contact filter (discrete) or the step rule (CCD).
__device__ void pair_finish_one(int p, const ContactRec* c, const int* pair_off, int* perm, int mode, int* keep,
                                int* kept_cnt, const double* pair_rec, double* scale) {
  {
    int b = pair_off[p], e = pair_off[p + 1];
    for (int i = b + 1; i < e; ++i) {
      int v = perm[i];
      unsigned long long k = c[v].key;
      int j = i - 1;
      while (j >= b && c[perm[j]].key > k) {
        perm[j + 1] = perm[j];
        --j;
      }
      perm[j + 1] = v;
    }
    if (mode == 1) {  // filterContacts(.., 0.0)
      Sorter s{perm, c};
      s.sort(b, e);
      int kept = 0;
      for (int i = b; i < e; ++i) {
        const ContactRec& ct = c[perm[i]];
        v3<double> m1 = dvs(add(v3<double>{ct.r[0], ct.r[1], ct.r[2]}, v3<double>{ct.r[3], ct.r[4], ct.r[5]}), 2.0);
        double search = (ct.r[6] + ct.r[7]) + 0.0;
        bool fresh = true;
        for (int h = b; h < i && fresh; ++h) {
          if (!keep[h]) continue;
          const ContactRec& hc = c[perm[h]];
          v3<double> m2 = dvs(add(v3<double>{hc.r[0], hc.r[1], hc.r[2]}, v3<double>{hc.r[3], hc.r[4], hc.r[5]}), 2.0);
          if (vlen(sub(m1, m2)) < search) fresh = false;
        }
        keep[i] = fresh;
        kept += fresh;
      }
      kept_cnt[p] = kept;
    } else if (mode == 2) {  // step rule, separate_clusters.cpp:496-564
      const double* pr = pair_rec + (size_t)PR_STRIDE * p;
      double mt = pr[PR_MAXT], ms = 1.0;
      v3<double> va{pr[PR_VA], pr[PR_VA + 1], pr[PR_VA + 2]}, wa{pr[PR_WA], pr[PR_WA + 1], pr[PR_WA + 2]};
      v3<double> xa{pr[PR_XA], pr[PR_XA + 1], pr[PR_XA + 2]};
      v3<double> vb{pr[PR_VB], pr[PR_VB + 1], pr[PR_VB + 2]}, wb{pr[PR_WB], pr[PR_WB + 1], pr[PR_WB + 2]};
      v3<double> xb{pr[PR_XB], pr[PR_XB + 1], pr[PR_XB + 2]};
      for (int i = b; i < e; ++i) {
        const ContactRec& ct = c[perm[i]];
        v3<double> A{ct.r[0], ct.r[1], ct.r[2]}, B{ct.r[3], ct.r[4], ct.r[5]};
        v3<double> nrm = sub(A, B);
        double inter = ct.r[6] + ct.r[7], toc = ct.r[10];
        v3<double> rel = sub(add(va, cross(wa, sub(A, xa))), add(vb, cross(wb, sub(B, xb))));
        double len = vlen(nrm);
        if (dot(unit(nrm), unit(rel)) <= 0.0 || len < 1e-4) {
          double s;
          if (len < 1e-4) {
            if (toc == 0.0) continue;
            double st = toc - (vlen(rel) * mt) / inter;
            st = st > 0.0 ? st : 0.0;  // std::max(0.0, x): x when x > 0 (NaN keeps 0)
            s = lerp(st, toc, 0.75);
          } else if (toc < 1.0) {
            double pv = -(dot(unit(nrm), rel));
            double depth = inter - len;
            double pd = (pv * toc) * mt;
            if (!(pd > 1e-6)) continue;
            if (depth < pd) {
              s = lerp(toc * (1.0 - depth / pd), toc, 0.5);
            } else {
              double fp = toc + ((1.0 - toc) * inter) / (pv * mt);
              s = lerp(toc, (1.0 < fp) ? 1.0 : fp, 0.8);  // std::min(fp, 1.0)
            }
          } else {
            continue;
          }
          if (s < ms) ms = s;  // std::min(ms, s): NaN never wins
        }
      }
      scale[p] = ms;
    }
  }
}
__global__ void k_pair_finish_big(const int* big, const int* n_big, const ContactRec* c, const int* pair_off, int* perm,
                                  int mode, int* keep, int* kept_cnt, const double* pair_rec, double* scale) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < *n_big; t += gridDim.x * blockDim.x)
    pair_finish_one(big[t], c, pair_off, perm, mode, keep, kept_cnt, pair_rec, scale);
}

// One wavefront per pair, contacts staged in LDS (pairs with more than
// POST_CAP contacts go to k_pair_finish, one lane per pair).
//   1. bitonic sort by emission key -> the reference's contact order
//   2a. discrete + filter: the order std::sort(compareContacts) produces.  When
//       no two contacts have equal (eps_a+eps_b, |A-B|) every correct sort
//       gives that order, so a parallel sort is used; with exact ties lane 0
//       replays libstdc++'s introsort (Sorter) on the LDS copy.  Then the
//       greedy pass of filterContacts, each candidate tested against the kept
//       set across the lanes.
//   2b. CCD: the step rule, lanes over contacts, min-reduced.
constexpr int POST_CAP = 512;
struct LdsKey {
  double eps, dist;
};
__device__ __forceinline__ bool lds_less(const LdsKey& x, const LdsKey& y) {
  if (x.eps == y.eps) return x.dist < y.dist;
  return x.eps < y.eps;
}
struct LdsSorter {  // Sorter over LDS keys (same algorithm, LDS operands)
  int* a;
  const LdsKey* k;
  __device__ bool lt(int x, int y) const { return lds_less(k[x], k[y]); }
  __device__ void swp(int i, int j) {
    int t = a[i];
    a[i] = a[j];
    a[j] = t;
  }
  __device__ void median_to_first(int res, int x, int y, int z) {
    if (lt(a[x], a[y])) {
      if (lt(a[y], a[z])) swp(res, y);
      else if (lt(a[x], a[z])) swp(res, z);
      else swp(res, x);
    } else if (lt(a[x], a[z])) swp(res, x);
    else if (lt(a[y], a[z])) swp(res, z);
    else swp(res, y);
  }
  __device__ int partition(int first, int last, int pivot) {
    while (true) {
      while (lt(a[first], a[pivot])) ++first;
      --last;
      while (lt(a[pivot], a[last])) --last;
      if (!(first < last)) return first;
      swp(first, last);
      ++first;
    }
  }
  __device__ void push_heap(int first, int hole, int top, int val) {
    int parent = (hole - 1) / 2;
    while (hole > top && lt(a[first + parent], val)) {
      a[first + hole] = a[first + parent];
      hole = parent;
      parent = (hole - 1) / 2;
    }
    a[first + hole] = val;
  }
  __device__ void adjust_heap(int first, int hole, int len, int val) {
    int top = hole, child = hole;
    while (child < (len - 1) / 2) {
      child = 2 * (child + 1);
      if (lt(a[first + child], a[first + child - 1])) child--;
      a[first + hole] = a[first + child];
      hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
      child = 2 * (child + 1);
      a[first + hole] = a[first + child - 1];
      hole = child - 1;
    }
    push_heap(first, hole, top, val);
  }
  __device__ void heap_sort(int first, int last) {
    int len = last - first;
    if (len >= 2)
      for (int parent = (len - 2) / 2;; --parent) {
        adjust_heap(first, parent, len, a[first + parent]);
        if (parent == 0) break;
      }
    while (last - first > 1) {
      --last;
      int v = a[last];
      a[last] = a[first];
      adjust_heap(first, 0, last - first, v);
    }
  }
  __device__ void linear_insert(int last) {
    int v = a[last], next = last - 1;
    while (lt(v, a[next])) {
      a[last] = a[next];
      last = next;
      --next;
    }
    a[last] = v;
  }
  __device__ void insertion(int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
      if (lt(a[i], a[first])) {
        int v = a[i];
        for (int q = i; q > first; --q) a[q] = a[q - 1];
        a[first] = v;
      } else {
        linear_insert(i);
      }
    }
  }
  __device__ void sort(int first, int last) {
    if (first == last) return;
    int n = last - first;
    int depth = 2 * (63 - __clzll((long long)n));
    int st_f[64], st_l[64], st_d[64], sp = 0;
    st_f[sp] = first, st_l[sp] = last, st_d[sp] = depth, ++sp;
    while (sp) {
      --sp;
      int f = st_f[sp], l = st_l[sp], d = st_d[sp];
      while (l - f > 16) {
        if (d == 0) {
          heap_sort(f, l);
          break;
        }
        --d;
        int mid = f + (l - f) / 2;
        median_to_first(f, f + 1, mid, l - 1);
        int cut = partition(f + 1, l, f);
        st_f[sp] = cut, st_l[sp] = l, st_d[sp] = d, ++sp;
        l = cut;
      }
    }
    if (n > 16) {
      insertion(first, first + 16);
      for (int i = first + 16; i != last; ++i) linear_insert(i);
    } else {
      insertion(first, last);
    }
  }
};

__global__ void __launch_bounds__(64) k_pair_finish_wave(int n_pairs, const ContactRec* c, const int* pair_off, int* perm,
                                                         int mode, int* keep, int* kept_cnt, const double* pair_rec,
                                                         double* scale, int* big, int* n_big) {
  __shared__ unsigned long long skey[POST_CAP];
  __shared__ int sidx[POST_CAP];
  __shared__ LdsKey sk[POST_CAP];
  __shared__ int ord[POST_CAP];
  __shared__ double smid[POST_CAP * 3];
  __shared__ int kept_list[POST_CAP];
  const int lane = threadIdx.x;
  for (int p = blockIdx.x; p < n_pairs; p += gridDim.x) {
    const int b = pair_off[p], e = pair_off[p + 1], n = e - b;
    if (n > POST_CAP) {
      if (lane == 0) big[atomicAdd(n_big, 1)] = p;
      continue;
    }
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int i = lane; i < n2; i += 64) {
      int idx = i < n ? perm[b + i] : -1;
      skey[i] = i < n ? c[idx].key : ~0ull;
      sidx[i] = idx;
    }
    __syncthreads();
    // bitonic sort of (skey, sidx) ascending (keys are unique)
    for (int kk = 2; kk <= n2; kk <<= 1)
      for (int jj = kk >> 1; jj > 0; jj >>= 1) {
        for (int i = lane; i < n2; i += 64) {
          int l = i ^ jj;
          if (l > i) {
            bool up = (i & kk) == 0;
            unsigned long long ki = skey[i], kl = skey[l];
            if ((ki > kl) == up) {
              skey[i] = kl, skey[l] = ki;
              int t = sidx[i];
              sidx[i] = sidx[l], sidx[l] = t;
            }
          }
        }
        __syncthreads();
      }
    if (mode == 2) {  // step rule, separate_clusters.cpp:496-564
      const double* pr = pair_rec + (size_t)PR_STRIDE * p;
      double mt = pr[PR_MAXT], ms = 1.0;
      v3<double> va{pr[PR_VA], pr[PR_VA + 1], pr[PR_VA + 2]}, wa{pr[PR_WA], pr[PR_WA + 1], pr[PR_WA + 2]};
      v3<double> xa{pr[PR_XA], pr[PR_XA + 1], pr[PR_XA + 2]};
      v3<double> vb{pr[PR_VB], pr[PR_VB + 1], pr[PR_VB + 2]}, wb{pr[PR_WB], pr[PR_WB + 1], pr[PR_WB + 2]};
      v3<double> xb{pr[PR_XB], pr[PR_XB + 1], pr[PR_XB + 2]};
      for (int i = lane; i < n; i += 64) {
        const ContactRec& ct = c[sidx[i]];
        v3<double> A{ct.r[0], ct.r[1], ct.r[2]}, B{ct.r[3], ct.r[4], ct.r[5]};
        v3<double> nrm = sub(A, B);
        double inter = ct.r[6] + ct.r[7], toc = ct.r[10];
        v3<double> rel = sub(add(va, cross(wa, sub(A, xa))), add(vb, cross(wb, sub(B, xb))));
        double len = vlen(nrm);
        if (!(dot(unit(nrm), unit(rel)) <= 0.0 || len < 1e-4)) continue;
        double s;
        if (len < 1e-4) {
          if (toc == 0.0) continue;
          double st = toc - (vlen(rel) * mt) / inter;
          st = st > 0.0 ? st : 0.0;
          s = lerp(st, toc, 0.75);
        } else {
          if (!(toc < 1.0)) continue;
          double pv = -(dot(unit(nrm), rel));
          double depth = inter - len;
          double pd = (pv * toc) * mt;
          if (!(pd > 1e-6)) continue;
          if (depth < pd) {
            s = lerp(toc * (1.0 - depth / pd), toc, 0.5);
          } else {
            double fp = toc + ((1.0 - toc) * inter) / (pv * mt);
            s = lerp(toc, (1.0 < fp) ? 1.0 : fp, 0.8);
          }
        }
        if (s < ms) ms = s;
      }
      for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(ms, off);
        if (o < ms) ms = o;
      }
      if (lane == 0) scale[p] = ms;
      for (int i = lane; i < n; i += 64) perm[b + i] = sidx[i];
    } else if (mode == 1) {  // filterContacts(.., 0.0)
      for (int i = lane; i < n; i += 64) {
        const ContactRec& ct = c[sidx[i]];
        sk[i].eps = ct.r[6] + ct.r[7];
        sk[i].dist = vlen(sub(v3<double>{ct.r[0], ct.r[1], ct.r[2]}, v3<double>{ct.r[3], ct.r[4], ct.r[5]}));
        v3<double> m = dvs(add(v3<double>{ct.r[0], ct.r[1], ct.r[2]}, v3<double>{ct.r[3], ct.r[4], ct.r[5]}), 2.0);
        smid[3 * i] = m.x, smid[3 * i + 1] = m.y, smid[3 * i + 2] = m.z;
      }
      for (int i = lane; i < n2; i += 64) ord[i] = i;
      __syncthreads();
      // parallel sort by (eps, dist, position); padding sorts last
      for (int kk = 2; kk <= n2; kk <<= 1)
        for (int jj = kk >> 1; jj > 0; jj >>= 1) {
          for (int i = lane; i < n2; i += 64) {
            int l = i ^ jj;
            if (l > i) {
              bool up = (i & kk) == 0;
              int oi = ord[i], ol = ord[l];
              bool gt;  // ord[i] sorts after ord[l]
              if (oi >= n || ol >= n) gt = oi > ol;
              else if (lds_less(sk[ol], sk[oi])) gt = true;
              else if (lds_less(sk[oi], sk[ol])) gt = false;
              else gt = oi > ol;
              if (gt == up) ord[i] = ol, ord[l] = oi;
            }
          }
          __syncthreads();
        }
      bool tie = false;
      for (int i = lane; i + 1 < n; i += 64) {
        const LdsKey x = sk[ord[i]], y = sk[ord[i + 1]];
        tie = tie || (x.eps == y.eps && x.dist == y.dist);
      }
      if (__ballot(tie)) {
        __syncthreads();
        for (int i = lane; i < n; i += 64) ord[i] = i;
        __syncthreads();
        if (lane == 0) {
          LdsSorter srt{ord, sk};
          srt.sort(0, n);
        }
        __syncthreads();
      }
      // greedy pass in sorted order
      int nk = 0;
      for (int t = 0; t < n; ++t) {
        int ci = ord[t];
        double mx = smid[3 * ci], my = smid[3 * ci + 1], mz = smid[3 * ci + 2];
        double search = sk[ci].eps + 0.0;
        bool near = false;
        for (int q = lane; q < nk; q += 64) {
          int h = kept_list[q];
          v3<double> dd{mx - smid[3 * h], my - smid[3 * h + 1], mz - smid[3 * h + 2]};
          near = near || (vlen(dd) < search);
        }
        bool fresh = __ballot(near) == 0ull;
        if (fresh && lane == 0) kept_list[nk] = ci;
        if (lane == 0) keep[b + t] = fresh;
        nk += fresh;
        __syncthreads();
      }
      for (int i = lane; i < n; i += 64) perm[b + i] = sidx[ord[i]];
      if (lane == 0) kept_cnt[p] = nk;
    } else {
      for (int i = lane; i < n; i += 64) perm[b + i] = sidx[i];
    }
    __syncthreads();
  }
}

// Gathers one pair per wavefront into the output arrays (CSR order).
__global__ void __launch_bounds__(64) k_gather_wave(int n_pairs, const ContactRec* c, const int* pair_off, const int* perm,
                                                    const int* keep, const int* out_off, int filtered, int32_t* out_i,
                                                    double* out_r) {
  const int lane = threadIdx.x;
  for (int p = blockIdx.x; p < n_pairs; p += gridDim.x) {
    int b = pair_off[p], e = pair_off[p + 1];
    int o = out_off[p];
    for (int base = b; base < e; base += 64) {
      int i = base + lane;
      bool take = i < e && (!filtered || keep[i]);
      unsigned long long m = __ballot(take);
      if (take) {
        int dst = o + __popcll(m & lanes_below());
        const ContactRec& ct = c[perm[i]];
        int32_t* oi = out_i + 8 * (size_t)dst;
        oi[0] = ct.pair, oi[1] = ct.na, oi[2] = ct.nb, oi[3] = ct.kind, oi[4] = ct.ia, oi[5] = ct.ib, oi[6] = ct.level,
        oi[7] = 0;
        double* orr = out_r + 11 * (size_t)dst;
        for (int k = 0; k < 11; ++k) orr[k] = ct.r[k];
      }
      o += __popcll(m);
    }
  }
}

// Body future states: future = current.extrapolate(step).
__global__ void k_project(int n, double* bodies, const double* step) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double* r = bodies + 40 * (size_t)i;
    body_state f = extrapolate(load_state(r), step[i], r + 28);
    store_state(f, r + 14);
  }
}

}  // namespace d2

// =========================================================================
// Host side: context and the C ABI.
// =========================================================================
using namespace d2;

namespace {

const char* kKernelNames[] = {"setup", "soup_cull", "soup_point", "soup_edge", "soup_reduce", "expand",
                              "leaf_enum", "leaf_point", "leaf_edge", "tri_enum", "tri_dist", "post"};
enum KId { K_SETUP, K_SOUP_CULL, K_SOUP_POINT, K_SOUP_EDGE, K_SOUP_REDUCE, K_EXPAND, K_LEAF_ENUM, K_LEAF_POINT,
           K_LEAF_EDGE, K_TRI_ENUM, K_TRI_DIST, K_POST, K_NUM };

template <class T> struct DBuf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct d2_ctx {
  int device = 0;
  int precision = D2_FP64;
  hipStream_t stream = nullptr;
  std::string err;
  long long cap = 1ll << 22;          // node pairs / items per list
  long long cap_contacts = 1ll << 22;
  int cus = 256;
  // trees / bodies
  DBuf<NodeRec> nodes;
  DBuf<double> buckets;
  DBuf<int32_t> tree_off, body_tree;
  DBuf<double> bodies;
  int n_trees = 0, n_bodies = 0;
  // pairs
  DBuf<int32_t> pairs, pair_info;
  DBuf<double> max_time, pair_rec;
  int n_pairs = 0;
  // work
  DBuf<NP> soup, conn, soup_next, conn_next;
  DBuf<Item> items, pt_items, ee_items;
  DBuf<double> res;
  DBuf<int> np_hit;
  DBuf<ContactRec> contacts;
  DBuf<unsigned long long> cnt, pass_min;
  DBuf<int> pass_flag;
  DBuf<double> min_time, scale;
  // post
  DBuf<int> pair_cnt, pair_off, pair_fill, perm, keep, kept_cnt, out_off, big, n_big;
  DBuf<int> cnt_a, off_a, cnt_s, off_s, cnt_c, off_c;  // per-level scans
  DBuf<unsigned long long> masks;
  DBuf<int32_t> out_i;
  DBuf<double> out_r;
  DBuf<unsigned char> scan_tmp;
  unsigned long long* host_cnt = nullptr;  // pinned
  long long n_out = 0;
  int last_mode = -1;  // 0 narrow, 1 ccd
  int last_filter = 1;
  // profiling
  bool profiling = false;
  long long k_launch[K_NUM] = {};
  double k_ms[K_NUM] = {};
  double k_items[K_NUM] = {};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<double> pending_items;
};

namespace {

int fail(d2_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
#define HIPCHK(ctx, expr)                                                                       \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return fail(ctx, D2_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T> int ensure(d2_ctx* c, DBuf<T>& b, size_t n) {
  if (b.n >= n && b.p) return D2_OK;
  if (b.p) HIPCHK(c, hipFree(b.p));
  b.p = nullptr;
  b.n = 0;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  HIPCHK(c, hipMalloc((void**)&b.p, bytes));
  b.n = std::max<size_t>(n, 1);
  return D2_OK;
}
template <class T> void release(DBuf<T>& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.n = 0;
}

int grid_for(d2_ctx* c, long long threads, int block = 256) {
  long long g = (threads + block - 1) / block;
  long long maxg = (long long)c->cus * 16;
  if (g > maxg) g = maxg;
  if (g < 1) g = 1;
  return (int)g;
}

struct Timer {
  d2_ctx* c;
  int id;
  double items;
  hipEvent_t a = nullptr, b = nullptr;
  Timer(d2_ctx* ctx, int k, double n) : c(ctx), id(k), items(n) {
    if (c->profiling) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~Timer() {
    if (c->profiling) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({id, {a, b}});
      c->pending_items.push_back(items);
    }
  }
};

void collect_timers(d2_ctx* c) {
  for (size_t i = 0; i < c->pending.size(); ++i) {
    auto& pe = c->pending[i];
    float ms = 0.f;
    (void)hipEventSynchronize(pe.second.second);
    (void)hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
    c->k_launch[pe.first] += 1;
    c->k_ms[pe.first] += ms;
    c->k_items[pe.first] += c->pending_items[i];
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  c->pending.clear();
  c->pending_items.clear();
}

Dev make_dev(d2_ctx* c, int level, int discrete) {
  Dev D;
  D.nodes = c->nodes.p;
  D.buckets = c->buckets.p;
  D.pair_info = c->pair_info.p;
  D.pair_rec = c->pair_rec.p;
  D.soup = c->soup.p;
  D.conn = c->conn.p;
  D.soup_next = c->soup_next.p;
  D.conn_next = c->conn_next.p;
  D.items = c->items.p;
  D.pt_items = c->pt_items.p;
  D.ee_items = c->ee_items.p;
  D.res = c->res.p;
  D.np_hit = c->np_hit.p;
  D.contacts = c->contacts.p;
  D.cnt = c->cnt.p;
  D.pass_min = c->pass_min.p;
  D.pass_flag = c->pass_flag.p;
  D.cap_np = c->cap;
  D.cap_items = c->cap;
  D.cap_contacts = c->cap_contacts;
  D.level = level;
  D.discrete = discrete;
  return D;
}

int read_counters(d2_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->host_cnt, c->cnt.p, sizeof(unsigned long long) * C_COUNT, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->host_cnt[C_OVERFLOW])
    return fail(c, D2_ERR_CAPACITY, "work list capacity exceeded; raise it with d2_set_capacity");
  return D2_OK;
}

int alloc_work(d2_ctx* c) {
  int rc;
  if ((rc = ensure(c, c->soup, c->cap)) || (rc = ensure(c, c->conn, c->cap)) || (rc = ensure(c, c->soup_next, c->cap)) ||
      (rc = ensure(c, c->conn_next, c->cap)) || (rc = ensure(c, c->items, c->cap)) || (rc = ensure(c, c->pt_items, c->cap)) ||
      (rc = ensure(c, c->ee_items, c->cap)) || (rc = ensure(c, c->np_hit, c->cap)) ||
      (rc = ensure(c, c->contacts, c->cap_contacts)) || (rc = ensure(c, c->cnt, C_COUNT)) ||
      (rc = ensure(c, c->cnt_a, c->cap + 1)) || (rc = ensure(c, c->off_a, c->cap + 1)) ||
      (rc = ensure(c, c->cnt_s, c->cap + 1)) || (rc = ensure(c, c->off_s, c->cap + 1)) ||
      (rc = ensure(c, c->cnt_c, c->cap + 1)) || (rc = ensure(c, c->off_c, c->cap + 1)) || (rc = ensure(c, c->masks, c->cap)))
    return rc;
  const long long res_items = std::min<long long>(c->cap, 1ll << 20);  // results chunk
  if ((rc = ensure(c, c->res, (size_t)res_items * RES_SLOTS * 8))) return rc;
  return D2_OK;
}

long long res_chunk(d2_ctx* c) { return (long long)(c->res.n / (RES_SLOTS * 8)); }

void swap_next(d2_ctx* c) {
  std::swap(c->soup, c->soup_next);
  std::swap(c->conn, c->conn_next);
}

// Exclusive scan of n ints on the context stream.
int scan_ints(d2_ctx* c, const int* in, int* out, long long n) {
  size_t tmp = 0;
  HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, c->stream));
  int rc = ensure(c, c->scan_tmp, tmp);
  if (rc) return rc;
  HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.p, tmp, in, out, (int)n, c->stream));
  return D2_OK;
}

// Runs the post-processing of the contacts gathered by a traversal.
int finish(d2_ctx* c, int mode, long long n_contacts) {
  int rc;
  const int np = c->n_pairs;
  if ((rc = ensure(c, c->pair_cnt, np + 1)) || (rc = ensure(c, c->pair_off, np + 1)) || (rc = ensure(c, c->pair_fill, np)) ||
      (rc = ensure(c, c->perm, n_contacts)) || (rc = ensure(c, c->keep, n_contacts)) || (rc = ensure(c, c->kept_cnt, np + 1)) ||
      (rc = ensure(c, c->out_off, np + 1)) || (rc = ensure(c, c->big, np)) || (rc = ensure(c, c->n_big, 1)))
    return rc;
  Timer tm(c, K_POST, (double)n_contacts);
  HIPCHK(c, hipMemsetAsync(c->pair_cnt.p, 0, sizeof(int) * (np + 1), c->stream));
  HIPCHK(c, hipMemsetAsync(c->pair_fill.p, 0, sizeof(int) * std::max(np, 1), c->stream));
  HIPCHK(c, hipMemsetAsync(c->n_big.p, 0, sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(c->kept_cnt.p, 0, sizeof(int) * (np + 1), c->stream));
  if (n_contacts > 0)
    hipLaunchKernelGGL(k_count, dim3(grid_for(c, n_contacts)), dim3(256), 0, c->stream, n_contacts, c->contacts.p,
                       c->pair_cnt.p);
  if ((rc = scan_ints(c, c->pair_cnt.p, c->pair_off.p, np + 1))) return rc;
  if (n_contacts > 0)
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(c, n_contacts)), dim3(256), 0, c->stream, n_contacts, c->contacts.p,
                       c->pair_off.p, c->pair_fill.p, c->perm.p);
  int fmode = mode == 0 ? (c->last_filter ? 1 : 0) : 2;
  const int wblocks = std::max(1, std::min(np, c->cus * 16));
  hipLaunchKernelGGL(k_pair_finish_wave, dim3(wblocks), dim3(64), 0, c->stream, np, c->contacts.p, c->pair_off.p,
                     c->perm.p, fmode, c->keep.p, c->kept_cnt.p, c->pair_rec.p, c->scale.p, c->big.p, c->n_big.p);
  hipLaunchKernelGGL(k_pair_finish_big, dim3(grid_for(c, std::max(np, 1), 64)), dim3(64), 0, c->stream, c->big.p,
                     c->n_big.p, c->contacts.p, c->pair_off.p, c->perm.p, fmode, c->keep.p, c->kept_cnt.p, c->pair_rec.p,
                     c->scale.p);
  const int* outoff = c->pair_off.p;
  long long n_out = n_contacts;
  int filtered = fmode == 1;
  if (filtered) {
    if ((rc = scan_ints(c, c->kept_cnt.p, c->out_off.p, np + 1))) return rc;
    int total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, c->out_off.p + np, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    n_out = total;
    outoff = c->out_off.p;
  } else {
    HIPCHK(c, hipMemcpyAsync(c->out_off.p, c->pair_off.p, sizeof(int) * (np + 1), hipMemcpyDeviceToDevice, c->stream));
  }
  if ((rc = ensure(c, c->out_i, (size_t)std::max<long long>(n_out, 1) * 8)) ||
      (rc = ensure(c, c->out_r, (size_t)std::max<long long>(n_out, 1) * 11)))
    return rc;
  hipLaunchKernelGGL(k_gather_wave, dim3(wblocks), dim3(64), 0, c->stream, np, c->contacts.p, c->pair_off.p, c->perm.p,
                     c->keep.p, outoff, filtered, c->out_i.p, c->out_r.p);
  HIPCHK(c, hipGetLastError());
  c->n_out = n_out;
  return D2_OK;
}

}  // namespace

namespace {
int traverse(d2_ctx* c, int discrete, int64_t* n_contacts_out) {
  if (c->n_pairs < 0 || c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "upload trees, bodies and pairs first");
  (void)hipSetDevice(c->device);
  int rc = alloc_work(c);
  if (rc) return rc;
  const int np = c->n_pairs;
  if (np > c->cap) return fail(c, D2_ERR_CAPACITY, "more pairs than the work-list capacity");
  {
    Timer tm(c, K_SETUP, np);
    hipLaunchKernelGGL(k_setup, dim3(grid_for(c, std::max(np, 1))), dim3(256), 0, c->stream, np, c->pairs.p,
                       c->max_time.p, c->bodies.p, c->body_tree.p, c->tree_off.p, c->pair_info.p, c->pair_rec.p,
                       discrete ? 0 : 1);
  }
  hipLaunchKernelGGL(k_roots, dim3(grid_for(c, std::max(np, 1))), dim3(256), 0, c->stream, np, c->soup.p, c->cnt.p,
                     c->pass_min.p, c->pass_flag.p, c->min_time.p, c->pair_rec.p, c->scale.p);
  HIPCHK(c, hipGetLastError());
  // The host copy of the counters must describe this run even when no level
  // executes (an empty pair list).
  if ((rc = read_counters(c))) return rc;
  long long n_soup = np, n_conn = 0;
  const int wave_block = 256;
  const int stream_grid = c->cus * 8;  // persistent-style grid for device-sized launches
  for (int level = 0; n_soup + n_conn > 0; ++level) {
    if (level > 15) return fail(c, D2_ERR_RUNTIME, "traversal deeper than 15 levels (contact keys hold 16)");
    Dev D = make_dev(c, level, discrete);
    if (n_soup > 0) {
      if (discrete) {
        {
          Timer tm(c, K_TRI_ENUM, (double)n_soup);
          hipLaunchKernelGGL(k_tri_mask, dim3(grid_for(c, n_soup * 64, wave_block)), dim3(wave_block), 0, c->stream, D,
                             n_soup, c->precision == D2_FP32 ? 1 : 0, c->masks.p, c->cnt_a.p);
          if ((rc = scan_ints(c, c->cnt_a.p, c->off_a.p, n_soup + 1))) return rc;
          hipLaunchKernelGGL(k_tri_enum_masked, dim3(grid_for(c, n_soup * 64, wave_block)), dim3(wave_block), 0,
                             c->stream, D, n_soup, c->masks.p, c->off_a.p);
        }
        {
          Timer tm(c, K_TRI_DIST, 0.0);
          hipLaunchKernelGGL(k_tri_dist_dev, dim3(stream_grid), dim3(256), 0, c->stream, D, c->off_a.p + n_soup);
        }
      } else {
        {
          Timer tm(c, K_SOUP_CULL, (double)n_soup);
          hipLaunchKernelGGL(k_soup_cull, dim3(grid_for(c, n_soup * 64, wave_block)), dim3(wave_block), 0, c->stream, D);
        }
        if ((rc = read_counters(c))) return rc;
        long long n_items = (long long)c->host_cnt[C_ITEMS];
        const long long chunk = res_chunk(c);
        for (long long f = 0; f < n_items; f += chunk) {
          long long cntc = std::min(chunk, n_items - f);
          {
            Timer tm(c, K_SOUP_POINT, (double)cntc * 6);
            hipLaunchKernelGGL(k_soup_point, dim3(grid_for(c, cntc * 6)), dim3(256), 0, c->stream, D, f, cntc);
          }
          {
            Timer tm(c, K_SOUP_EDGE, (double)cntc * 9);
            hipLaunchKernelGGL(k_soup_edge, dim3(grid_for(c, cntc * 9)), dim3(256), 0, c->stream, D, f, cntc);
          }
          {
            Timer tm(c, K_SOUP_REDUCE, (double)cntc);
            hipLaunchKernelGGL(k_soup_reduce, dim3(grid_for(c, cntc)), dim3(256), 0, c->stream, D, f, cntc);
          }
        }
      }
      {
        Timer tm(c, K_EXPAND, (double)n_soup);
        hipLaunchKernelGGL(k_expand_count, dim3(grid_for(c, n_soup + 1)), dim3(256), 0, c->stream, D, n_soup, 1,
                           c->cnt_s.p, c->cnt_c.p);
        if ((rc = scan_ints(c, c->cnt_s.p, c->off_s.p, n_soup + 1))) return rc;
        if (!discrete && (rc = scan_ints(c, c->cnt_c.p, c->off_c.p, n_soup + 1))) return rc;
        hipLaunchKernelGGL(k_expand_write, dim3(grid_for(c, n_soup)), dim3(256), 0, c->stream, D, n_soup, 1, c->off_s.p,
                           c->off_c.p);
      }
      if (!discrete)
        hipLaunchKernelGGL(k_fold_min, dim3(grid_for(c, np)), dim3(256), 0, c->stream, np, c->pass_min.p, c->pass_flag.p,
                           c->min_time.p, c->pair_rec.p);
      HIPCHK(c, hipGetLastError());
    }
    if (!discrete && n_conn > 0) {
      {
        Timer tm(c, K_LEAF_ENUM, (double)n_conn);
        hipLaunchKernelGGL(k_leaf_enum, dim3(grid_for(c, n_conn * 64, wave_block)), dim3(wave_block), 0, c->stream, D);
      }
      if ((rc = read_counters(c))) return rc;
      long long npt = (long long)c->host_cnt[C_PT], nee = (long long)c->host_cnt[C_EE];
      if (npt > 0) {
        Timer tm(c, K_LEAF_POINT, (double)npt);
        hipLaunchKernelGGL(k_leaf_point, dim3(grid_for(c, npt)), dim3(256), 0, c->stream, D, 0ll, npt);
      }
      if (nee > 0) {
        Timer tm(c, K_LEAF_EDGE, (double)nee);
        hipLaunchKernelGGL(k_leaf_edge, dim3(grid_for(c, nee)), dim3(256), 0, c->stream, D, 0ll, nee);
      }
      hipLaunchKernelGGL(k_fold_min, dim3(grid_for(c, np)), dim3(256), 0, c->stream, np, c->pass_min.p, c->pass_flag.p,
                         c->min_time.p, c->pair_rec.p);
      HIPCHK(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_advance_scanned, dim3(1), dim3(64), 0, c->stream, c->cnt.p,
                       n_soup > 0 ? c->off_s.p + n_soup : nullptr, n_soup > 0 ? c->off_c.p + n_soup : nullptr,
                       discrete ? 0 : 1);
    if ((rc = read_counters(c))) return rc;
    if (discrete && n_soup > 0 && c->profiling) c->k_items[K_TRI_DIST] += (double)c->host_cnt[C_LVL_ITEMS];
    swap_next(c);
    n_soup = (long long)c->host_cnt[C_SOUP];
    n_conn = (long long)c->host_cnt[C_CONN];
  }
  long long n_contacts = (long long)c->host_cnt[C_CONTACTS];
  if ((rc = finish(c, discrete ? 0 : 1, n_contacts))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->profiling) collect_timers(c);
  c->last_mode = discrete ? 0 : 1;
  if (n_contacts_out) *n_contacts_out = c->n_out;
  return D2_OK;
}
}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

int d2_ctx_create(int device, int precision, d2_ctx** out) {
  if (!out) return D2_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return D2_ERR_HIP;
  if (device < 0 || device >= ndev) return D2_ERR_ARG;
  if (precision != D2_FP64 && precision != D2_FP32) return D2_ERR_ARG;
  d2_ctx* c = new d2_ctx();
  c->device = device;
  c->precision = precision;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return D2_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
  if (hipHostMalloc((void**)&c->host_cnt, sizeof(unsigned long long) * C_COUNT, hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return D2_ERR_HIP;
  }
  *out = c;
  return D2_OK;
}

int d2_ctx_destroy(d2_ctx* c) {
  if (!c) return D2_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  collect_timers(c);
  release(c->nodes), release(c->buckets), release(c->tree_off), release(c->body_tree), release(c->bodies);
  release(c->pairs), release(c->pair_info), release(c->max_time), release(c->pair_rec);
  release(c->soup), release(c->conn), release(c->soup_next), release(c->conn_next);
  release(c->items), release(c->pt_items), release(c->ee_items), release(c->res), release(c->np_hit);
  release(c->contacts), release(c->cnt), release(c->pass_min), release(c->pass_flag), release(c->min_time),
      release(c->scale);
  release(c->pair_cnt), release(c->pair_off), release(c->pair_fill), release(c->perm), release(c->keep),
      release(c->kept_cnt), release(c->out_off), release(c->out_i), release(c->out_r), release(c->scan_tmp);
  release(c->big), release(c->n_big), release(c->cnt_a), release(c->off_a), release(c->cnt_s), release(c->off_s),
      release(c->cnt_c), release(c->off_c), release(c->masks);
  if (c->host_cnt) (void)hipHostFree(c->host_cnt);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return D2_OK;
}

const char* d2_last_error(const d2_ctx* c) { return c ? c->err.c_str() : "null context"; }

int d2_set_precision(d2_ctx* c, int precision) {
  if (!c) return D2_ERR_ARG;
  if (precision != D2_FP64 && precision != D2_FP32) return fail(c, D2_ERR_ARG, "precision must be D2_FP64 or D2_FP32");
  c->precision = precision;
  return D2_OK;
}

int d2_set_capacity(d2_ctx* c, int64_t entries) {
  if (!c) return D2_ERR_ARG;
  if (entries < 1024) return fail(c, D2_ERR_ARG, "capacity must be >= 1024");
  c->cap = entries;
  c->cap_contacts = entries;
  return D2_OK;
}

int d2_trees_upload(d2_ctx* c, int n_trees, const int32_t* tree_node_offset, const int32_t* nodes, const double* buckets) {
  if (!c) return D2_ERR_ARG;
  if (n_trees <= 0 || !tree_node_offset || !nodes || !buckets) return fail(c, D2_ERR_ARG, "d2_trees_upload: bad arguments");
  int total = tree_node_offset[n_trees];
  if (total <= 0 || tree_node_offset[0] != 0) return fail(c, D2_ERR_ARG, "d2_trees_upload: bad offsets");
  for (int t = 0; t < n_trees; ++t) {
    int b = tree_node_offset[t], e = tree_node_offset[t + 1];
    if (e <= b) return fail(c, D2_ERR_ARG, "d2_trees_upload: empty tree");
    for (int i = b; i < e; ++i) {
      const int32_t* r = nodes + 8 * (size_t)i;
      int cnt = e - b;
      if (r[1] < 0 || r[1] >= cnt || r[7] < 0 || r[7] > 8 || (r[2] && (r[5] < 0 || r[5] + r[6] > cnt || r[6] < 1 || r[6] > 8)))
        return fail(c, D2_ERR_ARG, "d2_trees_upload: inconsistent node record");
    }
  }
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = ensure(c, c->nodes, total)) || (rc = ensure(c, c->buckets, (size_t)total * BUCKET)) ||
      (rc = ensure(c, c->tree_off, n_trees + 1)))
    return rc;
  HIPCHK(c, hipMemcpyAsync(c->nodes.p, nodes, sizeof(NodeRec) * total, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->buckets.p, buckets, sizeof(double) * BUCKET * (size_t)total, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->tree_off.p, tree_node_offset, sizeof(int32_t) * (n_trees + 1), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_trees = n_trees;
  return D2_OK;
}

int d2_bodies_upload(d2_ctx* c, int n_bodies, const int32_t* tree_of_body, const double* records) {
  if (!c) return D2_ERR_ARG;
  if (n_bodies <= 0 || !tree_of_body || !records) return fail(c, D2_ERR_ARG, "d2_bodies_upload: bad arguments");
  if (c->n_trees <= 0) return fail(c, D2_ERR_STATE, "d2_bodies_upload: upload trees first");
  for (int i = 0; i < n_bodies; ++i)
    if (tree_of_body[i] < 0 || tree_of_body[i] >= c->n_trees) return fail(c, D2_ERR_ARG, "d2_bodies_upload: bad tree index");
  (void)hipSetDevice(c->device);
  int rc;
  if ((rc = ensure(c, c->bodies, (size_t)n_bodies * 40)) || (rc = ensure(c, c->body_tree, n_bodies))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->bodies.p, records, sizeof(double) * 40 * (size_t)n_bodies, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->body_tree.p, tree_of_body, sizeof(int32_t) * n_bodies, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_bodies = n_bodies;
  return D2_OK;
}

int d2_bodies_download(d2_ctx* c, double* records) {
  if (!c || !records) return D2_ERR_ARG;
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "no bodies uploaded");
  HIPCHK(c, hipMemcpyAsync(records, c->bodies.p, sizeof(double) * 40 * (size_t)c->n_bodies, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_project_future(d2_ctx* c, const double* step_per_body) {
  if (!c || !step_per_body) return D2_ERR_ARG;
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "no bodies uploaded");
  DBuf<double> st;
  int rc = ensure(c, st, c->n_bodies);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(st.p, step_per_body, sizeof(double) * c->n_bodies, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_project, dim3(grid_for(c, c->n_bodies)), dim3(256), 0, c->stream, c->n_bodies, c->bodies.p, st.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  release(st);
  return D2_OK;
}

int d2_pairs_upload(d2_ctx* c, int n_pairs, const int32_t* pairs, const double* max_time) {
  if (!c) return D2_ERR_ARG;
  if (n_pairs < 0 || (n_pairs > 0 && !pairs)) return fail(c, D2_ERR_ARG, "d2_pairs_upload: bad arguments");
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "d2_pairs_upload: upload bodies first");
  for (int p = 0; p < n_pairs; ++p) {
    if (pairs[2 * p] < 0 || pairs[2 * p] >= c->n_bodies || pairs[2 * p + 1] < 0 || pairs[2 * p + 1] >= c->n_bodies)
      return fail(c, D2_ERR_ARG, "d2_pairs_upload: body index out of range");
    if (max_time && !(max_time[p] >= 0.0)) return fail(c, D2_ERR_ARG, "d2_pairs_upload: max_time must be >= 0");
  }
  (void)hipSetDevice(c->device);
  int rc;
  size_t np = std::max(n_pairs, 1);
  if ((rc = ensure(c, c->pairs, 2 * np)) || (rc = ensure(c, c->max_time, np)) || (rc = ensure(c, c->pair_rec, np * PR_STRIDE)) ||
      (rc = ensure(c, c->pair_info, 4 * np)) || (rc = ensure(c, c->pass_min, np)) || (rc = ensure(c, c->pass_flag, np)) ||
      (rc = ensure(c, c->min_time, np)) || (rc = ensure(c, c->scale, np)))
    return rc;
  if (n_pairs > 0) {
    HIPCHK(c, hipMemcpyAsync(c->pairs.p, pairs, sizeof(int32_t) * 2 * n_pairs, hipMemcpyHostToDevice, c->stream));
    if (max_time)
      HIPCHK(c, hipMemcpyAsync(c->max_time.p, max_time, sizeof(double) * n_pairs, hipMemcpyHostToDevice, c->stream));
    else
      HIPCHK(c, hipMemsetAsync(c->max_time.p, 0, sizeof(double) * n_pairs, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_pairs = n_pairs;
  return D2_OK;
}


int d2_narrow_run(d2_ctx* c, int filter, int64_t* n_contacts_out) {
  if (!c) return D2_ERR_ARG;
  c->last_filter = filter ? 1 : 0;
  return traverse(c, 1, n_contacts_out);
}

int d2_ccd_run(d2_ctx* c, int64_t* n_contacts_out) {
  if (!c) return D2_ERR_ARG;
  return traverse(c, 0, n_contacts_out);
}

int d2_narrow(d2_ctx* c, int n_pairs, const int32_t* pairs, int filter, int64_t* n_contacts_out) {
  int rc = d2_pairs_upload(c, n_pairs, pairs, nullptr);
  return rc ? rc : d2_narrow_run(c, filter, n_contacts_out);
}

int d2_ccd(d2_ctx* c, int n_pairs, const int32_t* pairs, const double* max_time, int64_t* n_contacts_out) {
  if (!max_time) return fail(c, D2_ERR_ARG, "d2_ccd: max_time is required");
  int rc = d2_pairs_upload(c, n_pairs, pairs, max_time);
  return rc ? rc : d2_ccd_run(c, n_contacts_out);
}

int d2_result_contacts(d2_ctx* c, int32_t* offsets, int32_t* ints, double* reals) {
  if (!c) return D2_ERR_ARG;
  if (c->last_mode < 0) return fail(c, D2_ERR_STATE, "no result: run the path first");
  if ((ints || reals) && c->n_out > 0) {  // the CSR total and the record count must agree
    int total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, c->out_off.p + c->n_pairs, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((long long)total != c->n_out) return fail(c, D2_ERR_STATE, "result offsets and contact count disagree");
  }
  if (offsets)
    HIPCHK(c, hipMemcpyAsync(offsets, c->out_off.p, sizeof(int) * (c->n_pairs + 1), hipMemcpyDeviceToHost, c->stream));
  if (ints && c->n_out)
    HIPCHK(c, hipMemcpyAsync(ints, c->out_i.p, sizeof(int32_t) * 8 * (size_t)c->n_out, hipMemcpyDeviceToHost, c->stream));
  if (reals && c->n_out)
    HIPCHK(c, hipMemcpyAsync(reals, c->out_r.p, sizeof(double) * 11 * (size_t)c->n_out, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_result_pairs(d2_ctx* c, double* per_pair) {
  if (!c || !per_pair) return D2_ERR_ARG;
  if (c->last_mode != 1) return fail(c, D2_ERR_STATE, "per-pair results exist after d2_ccd_run");
  std::vector<double> sc(c->n_pairs), mt(c->n_pairs);
  if (c->n_pairs) {
    HIPCHK(c, hipMemcpyAsync(sc.data(), c->scale.p, sizeof(double) * c->n_pairs, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(mt.data(), c->min_time.p, sizeof(double) * c->n_pairs, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int p = 0; p < c->n_pairs; ++p) per_pair[2 * p] = sc[p], per_pair[2 * p + 1] = mt[p];
  return D2_OK;
}

// The cluster step is a minimum over pair steps and bodies' last steps:
// every update of separate_clusters.cpp:541-553 takes a minimum with the
// pair's max_time * scale, and the cluster minimum (:570-598) reads one body
// per pair, so the result does not depend on the order of the sequential
// loop.  Bodies, pairs and the per-pair scales are the resident ones.
int d2_cluster_steps(d2_ctx* c, const int32_t* cluster_of_pair, int n_clusters, float dt, double* out) {
  if (!c || !cluster_of_pair || !out || n_clusters <= 0) return D2_ERR_ARG;
  if (c->last_mode != 1) return fail(c, D2_ERR_STATE, "d2_cluster_steps needs a preceding d2_ccd_run");
  const int np = c->n_pairs;
  std::vector<double> per(2 * (size_t)np), bodies((size_t)c->n_bodies * 40);
  std::vector<int32_t> pairs(2 * (size_t)np), off(np + 1);
  int rc = d2_result_pairs(c, per.data());
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(bodies.data(), c->bodies.p, sizeof(double) * bodies.size(), hipMemcpyDeviceToHost, c->stream));
  if (np) HIPCHK(c, hipMemcpyAsync(pairs.data(), c->pairs.p, sizeof(int32_t) * 2 * np, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(off.data(), c->out_off.p, sizeof(int) * (np + 1), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> step(n_clusters, -1.0);
  std::vector<int> first_body(n_clusters, -1);
  for (int p = 0; p < np; ++p) {
    int k = cluster_of_pair[p];
    if (k < 0 || k >= n_clusters) return fail(c, D2_ERR_ARG, "d2_cluster_steps: cluster index out of range");
    int a = pairs[2 * p], b = pairs[2 * p + 1];
    const double* ra = &bodies[40 * (size_t)a];
    const double* rb = &bodies[40 * (size_t)b];
    bool sa = ra[38] != 0.0, sb = rb[38] != 0.0;
    if (first_body[k] < 0) first_body[k] = a;
    double mt = std::min(ra[39], rb[39]);
    double v = 1e300;
    if (off[p + 1] > off[p] && !(sa && sb)) v = std::min(v, mt * per[2 * p]);
    if (!(sa && sb)) {
      int id = (!sa) ? a : b;                 // separate_clusters.cpp:575-586
      double last = step[k] < 0.0 ? ra[39] : bodies[40 * (size_t)id + 39];  // first pair reads a (:589-590)
      v = std::min(v, last);
      step[k] = step[k] < 0.0 ? v : std::min(step[k], v);
    }
  }
  for (int k = 0; k < n_clusters; ++k) {
    double s = step[k];
    if (s < 0.0) s = first_body[k] >= 0 ? bodies[40 * (size_t)first_body[k] + 39] : (double)dt;
    out[k] = std::max(1e-3, std::min(s, (double)dt));
  }
  return D2_OK;
}

int d2_set_profiling(d2_ctx* c, int on) {
  if (!c) return D2_ERR_ARG;
  c->profiling = on != 0;
  return D2_OK;
}
int d2_num_kernels(void) { return K_NUM; }
const char* d2_kernel_name(int i) { return (i >= 0 && i < K_NUM) ? kKernelNames[i] : ""; }
int d2_kernel_stats(d2_ctx* c, int i, int64_t* launches, double* total_ms, double* items) {
  if (!c || i < 0 || i >= K_NUM) return D2_ERR_ARG;
  collect_timers(c);
  if (launches) *launches = c->k_launch[i];
  if (total_ms) *total_ms = c->k_ms[i];
  if (items) *items = c->k_items[i];
  return D2_OK;
}
int d2_reset_stats(d2_ctx* c) {
  if (!c) return D2_ERR_ARG;
  collect_timers(c);
  for (int i = 0; i < K_NUM; ++i) c->k_launch[i] = 0, c->k_ms[i] = 0.0, c->k_items[i] = 0.0;
  return D2_OK;
}
int d2_sync(d2_ctx* c) {
  if (!c) return D2_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

}  // extern "C"
