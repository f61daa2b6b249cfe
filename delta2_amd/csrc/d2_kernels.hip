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
//     soup_sub     one thread per (item, point-triangle sub-test)  6 per item,
//                  and (other blocks) per (item, edge-edge sub-test) 9 per item
//     soup_reduce  one thread per item: the reference's ordered selection of
//                  the 15 sub-results, hit test, leaf-pair contact emission
//     expand       one thread per node pair: children of hit node pairs
//     leaf_enum    one wavefront per leaf-leaf node pair: enumerates its
//                  vertex-triangle and edge-edge sub-tests
//     leaf_sub     one thread per sub-test (point-triangle blocks, edge-edge
//                  blocks), hit -> contact
//   Discrete level (compareTreesFull, full_tree_comparison.h:13-67)
//     tri_mask     one wavefront per node pair, lane = triangle pair: FP32
//                  cull (or keep-all in FP64 mode), kept pairs to the item
//                  list, children of node pairs that hit or may hit to the
//                  next list (dropped there if the parent did not hit)
//     tri_dist     one thread per item: FP64 _comparisonDist, hit -> contact
//                  (leaf-leaf) or the node pair's hit flag
//     disc_expand  (FP64 mode) one thread per node pair: children of hits
//   Post-processing: per-pair ordering of contacts by their breadth-first key,
//   the contact filter (contact.h:44-69, libstdc++ std::sort permutation) and
//   the TOI -> step rule (separate_clusters.cpp:496-564).
//
// Contacts are emitted with a 64-bit key (level, list, child-index path,
// sub-test) whose order is the reference's emission order, so allocation can
// use atomics and the order is restored per pair afterwards.
#include <hip/hip_runtime.h>

#include <hip/hip_ext.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <random>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/d2hip.h"
#define D2_HD __host__ __device__ __forceinline__
#include "d2_geom.hpp"
#include "d2_cull.hpp"
#include "d2_sort.hpp"
#include "d2_solver.hpp"

namespace d2 {

// ------------------------------------------------------------------ records
constexpr int NB_STRIDE = 12;  // per-node pruning bounds (d2_trees_upload)
// The FP32 form's copy of the buckets: per triangle (node * 8 + k) 12 floats,
// its corners rounded to float (a.x a.y a.z b.x ... c.z), one unused, and its
// eps's double bits in the last two (low word first); three 16-byte loads.
constexpr int FB_STRIDE = 12;
// and its planes for the CCD point-triangle culls (soup_tri: normal, offset,
// edge normals and offsets, acceptance bound) in 20 floats, five 16-byte loads.
constexpr int ST_STRIDE = 20;
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
struct DItem {  // discrete triangle-pair item: node pair, pair, global buckets, sub | inner << 8
  uint32_t np, pair, ga, gb, sub;
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
// Global counters (atomics) and per-level sizes.  Level L reads its list sizes
// from lvl[L*LV_STRIDE + ...] and the level's write kernel stores the sizes of
// level L+1; no size ever travels through the host.
enum Counter : int {
  C_CONTACTS = 0, C_OVERFLOW = 1, C_LVL_ITEMS = 2, C_OUT = 3,
  C_GATED_CCD = 4, C_CERT_CCD = 5, C_GATED_DISC = 6, C_CERT_DISC = 7,  // learned gate: proposed / certified pairs
  C_COUNT = 8
};
// LV_PACK (discrete levels): the level's item count (low 32 bits) and its
// children count (high 32 bits), taken together by k_tri_mask's allocation
// atomic and unpacked into LV_ITEMS and the next level's LV_SOUP by k_tri_dist.
// LV_SPT / LV_SEE (CCD levels, FP32 form): the soup sub-tests that survive
// the FP32 culls, point and edge (slot 5 is LV_PACK on discrete levels).
// LV_DEF (discrete levels, FP32 form): the deferred items of inner node pairs
// (k_tri_def), allocated by k_tri_mask (slot 6 is LV_SEE on CCD levels).
enum LevelSlot : int {
  LV_SOUP = 0, LV_CONN = 1, LV_ITEMS = 2, LV_PT = 3, LV_EE = 4, LV_PACK = 5, LV_SPT = 5, LV_SEE = 6, LV_DEF = 6,
  LV_STRIDE = 8
};
constexpr int MAX_LEVELS = 16;  // contact keys hold 16 base-8 path digits
// Key layout: level(5) | list(1) | path(48) | sub(10)
__device__ __forceinline__ unsigned long long make_key(int level, int list, unsigned long long path, int sub) {
  return ((unsigned long long)level << 59) | ((unsigned long long)list << 58) | (path << 10) | (unsigned long long)sub;
}

// std::min / std::max semantics (first argument on ties and NaN).
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
// Value of lane l (wave-uniform) without going through LDS.
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long bits = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(bits & 0xffffffffll), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(bits >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(v & 0xffffffffull), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}
// Orders one wavefront's LDS writes before its later LDS reads (the other
// wavefronts of the block are not waited for).
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ unsigned long long lanes_below() { return (1ull << lane_id()) - 1ull; }
// Inclusive prefix sum over the wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d);
    if (lane_id() >= d) v += o;
  }
  return v;
}

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
// Corner v (0-2) of triangle k, read from the bucket: with a runtime v this
// is an indexed global load, where selecting from a loaded triangle puts the
// triangle in scratch memory.
__device__ __forceinline__ v3<double> load_corner(const double* bucket, int k, int v) {
  return v3<double>{bucket[(3 * v) * 8 + k], bucket[(3 * v + 1) * 8 + k], bucket[(3 * v + 2) * 8 + k]};
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
  DItem* ditems;
  DItem* ditems_def;           // FP32 form: deferred items of inner node pairs (k_tri_def)
  Item* pt_items;
  Item* ee_items;
  double* res;  // [items][15][8]
  unsigned* res_mask;  // FP32 form: per soup item, the slots its FP32 culls keep
  int* np_hit;
  int* pair_cnt;               // contacts per pair, counted as they are written
  int* exp_desc;               // per discrete node pair: expansion_desc (FP64 mode)
  const int* np_hit_prev;      // the previous level's hit flags (discrete: parity buffers)
  ContactRec* contacts;
  unsigned long long* cnt;
  unsigned long long* lvl;
  // per node pair counts, masks, and per-block partial sums of the counts
  unsigned long long* masks;
  int *cnt_a, *cnt_s, *cnt_c, *cnt_pt, *cnt_ee;
  int *part_a, *part_s, *part_c, *part_pt, *part_ee;
  // per pair, min toc of the soup / connected pass (bits) and "a leaf pair
  // had hits in this pass"
  unsigned long long *pass_min_s, *pass_min_c;
  int *pass_flag_s, *pass_flag_c;
  double* min_time;
  const float* tree_eps_out;   // float-barycentric acceptance bound per tree
  const float* fbuckets;       // FP32 form: per triangle its corners rounded to float and its eps (FB_STRIDE)
  const float* stris;          // FP32 form: per triangle its soup_tri planes (ST_STRIDE)
  const float* nbound;         // per node: leaf box, slack, sphere (nodes_apart, spheres_apart)
  const uint8_t* kid_inner;    // per node: bit c set when child c is an inner node
  int prune;                   // skip node pairs nodes_apart proves contact-free
  int defer;                   // FP32 form: inner node pairs send one item to k_tri_dist, the rest to k_tri_def
  int box_test;                // FP32 form: k_tri_mask tests arriving node pairs with boxes_apart_faces
  unsigned long long* conn_mask;  // per connected node pair: surviving leaf sub-tests
  const int32_t* body_tree;
  int fp32;                    // FP32 culls of the CCD leaf sub-tests
  int n_pairs;
  long long cap_np, cap_items, cap_res, cap_contacts;  // cap_res: CCD soup items (result slots)
  int level;
  int discrete;
  __device__ long long lv(int level_, int slot) const { return (long long)lvl[level_ * LV_STRIDE + slot]; }
  __device__ long long n_soup() const { long long n = lv(level, LV_SOUP); return n < cap_np ? n : cap_np; }
  __device__ long long n_conn() const { long long n = lv(level, LV_CONN); return n < cap_np ? n : cap_np; }
  __device__ long long n_lv(int slot) const { long long n = lv(level, slot); return n < cap_items ? n : cap_items; }
  __device__ long long n_res() const { long long n = lv(level, LV_ITEMS); return n < cap_res ? n : cap_res; }
};

// ------------------------------------------------- block-level primitives
// Chunked kernels split a list of n entries into gridDim.x contiguous ranges,
// one per block; a producer stores each block's total in part[block], and its
// consumer (launched with the same grid) takes its offset as the sum of the
// partials before it.  This replaces a separate scan and keeps the lists in
// order.
__device__ __forceinline__ void chunk_of(long long n, long long& b, long long& e) {
  const long long ch = (n + gridDim.x - 1) / gridDim.x;
  b = (long long)blockIdx.x * ch;
  e = b + ch;
  if (b > n) b = n;
  if (e > n) e = n;
}
__device__ long long block_sum(long long v) {
  __shared__ long long red[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  long long s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;
}
// Exclusive scan over the block; total = sum over the block.
__device__ int block_scan(int v, int& total) {
  __shared__ int ws[16];
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  __syncthreads();
  if (lane == 63) ws[wid] = x;
  __syncthreads();
  int before = 0, tot = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
    if (i < wid) before += ws[i];
    tot += ws[i];
  }
  total = tot;
  return before + x - v;
}
__device__ void partial_prefix(const int* part, long long& before, long long& total) {
  long long b = 0, t = 0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
    long long v = part[i];
    t += v;
    if (i < (int)blockIdx.x) b += v;
  }
  before = block_sum(b);
  total = block_sum(t);
}

// ----------------------------------------------------------------- set-up
// Cluster-step accumulators (separate_clusters.cpp:541-598): per cluster the
// smallest step (as bits) and its first pair / first pair with a non-static
// body; with body_step, every body's step starts as its clamped last step (a
// body in no pair is a cluster of its own without interactions,
// separate_clusters.cpp:594-596, time_step_selection_dynamic_continuous.cpp:19).
struct ClusterAcc {
  int n_clusters, n_bodies;
  float dt;
  unsigned long long* step_bits;  // null: nothing to do
  int *first_any, *first_ns;
  int* read_later;  // per body: its step is read or changed after its cluster's first pair
  const double* bodies;
  double* body_step;
  const int32_t *pairs, *cluster_of_pair;
};
__device__ __forceinline__ void cluster_init_range(const ClusterAcc& A, long long t0, long long stride) {
  const int n = A.n_bodies > A.n_clusters ? A.n_bodies : A.n_clusters;
  for (long long k = t0; k < n; k += stride) {
    if (k < A.n_clusters) A.step_bits[k] = ~0ull, A.first_any[k] = 0x7fffffff, A.first_ns[k] = 0x7fffffff;
    if (k < A.n_bodies) A.read_later[k] = 0;
    if (A.body_step && k < A.n_bodies) {
      double s = A.bodies[40 * (size_t)k + 39];
      const double hi = (double)A.dt;
      s = (hi < s) ? hi : s;
      A.body_step[k] = (s < 1e-3) ? 1e-3 : s;
    }
  }
}
__device__ __forceinline__ void cluster_first_range(const ClusterAcc& A, int n_pairs, long long t0, long long stride) {
  for (long long p = t0; p < n_pairs; p += stride) {
    const int k = A.cluster_of_pair[p];
    atomicMin(&A.first_any[k], (int)p);
    const bool sa = A.bodies[40 * (size_t)A.pairs[2 * p] + 38] != 0.0;
    const bool sb = A.bodies[40 * (size_t)A.pairs[2 * p + 1] + 38] != 0.0;
    if (!(sa && sb)) atomicMin(&A.first_ns[k], (int)p);
  }
}

// Per pair: poses, frame changes, kinematics (CCD: future = extrapolate(max_time)).
// Four lanes per pair split the record: lane q computes the poses and the
// frame change it writes (q = 0: TA0, inv(TB0) TA0; 1: TA1, inv(TB1) TA1 and
// a's kinematics; 2: TB0, inv(TA0) TB0; 3: TB1, inv(TA1) TB1 and b's
// kinematics), so the longest chain holds two extrapolations, one inverse
// and one product instead of all of them.
// The same kernel starts the phase: per pair the root node pair, the pass
// accumulators, min_time and scale (a step's discrete phase keeps its CCD
// phase's), the post-processing counters; block 0 clears the counters (a
// step's discrete phase keeps the CCD phase's overflow flag and gate counts)
// and sets the level table to one level of n_pairs roots (none when the
// learned gate's certificate appends the phase's roots).
__global__ void __launch_bounds__(256) k_setup(Dev D, const int32_t* pairs, const double* max_time, const double* bodies,
                        const int32_t* body_tree, const int32_t* tree_off, int32_t* pair_info, double* pair_rec,
                        int ccd, int mt_from_bodies, int keep_ccd, double* scale, int* pair_fill, int* kept_cnt,
                        int* n_big, ClusterAcc CA, int gated) {
  const int n_pairs = D.n_pairs;
  if (CA.step_bits)  // a step: its cluster accumulators
    cluster_init_range(CA, (long long)blockIdx.x * blockDim.x + threadIdx.x, (long long)gridDim.x * blockDim.x);
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < C_COUNT; i += blockDim.x)
      if (!(keep_ccd && (i == C_OVERFLOW || i == C_GATED_CCD || i == C_CERT_CCD))) D.cnt[i] = 0;
    for (int i = threadIdx.x; i < (MAX_LEVELS + 1) * LV_STRIDE; i += blockDim.x)  // gated: roots come from k_certify
      D.lvl[i] = i == LV_SOUP && !gated ? (unsigned long long)n_pairs : 0ull;
    if (threadIdx.x == 0) D.pair_cnt[n_pairs] = 0, kept_cnt[n_pairs] = 0, n_big[0] = 0, n_big[1] = 0;
  }
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < 4ll * n_pairs;
       t += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(t >> 2), q = (int)(t & 3);
    int a = pairs[2 * p], b = pairs[2 * p + 1];
    const double* ra = bodies + 40 * (size_t)a;
    const double* rb = bodies + 40 * (size_t)b;
    // max_time: given per pair, or min(last steps of a, b) (separate_clusters.cpp:490)
    const double mt = !ccd ? 0.0 : mt_from_bodies ? ((rb[39] < ra[39]) ? rb[39] : ra[39]) : max_time[p];
    const bool end = q & 1;  // lanes 1, 3 work on the end poses
    const bool of_a = q < 2;  // lanes 0, 1 own a's pose, lanes 2, 3 b's
    // the own body's record and the other's (selecting pointers, not private
    // arrays, keeps the poses in registers)
    const double* ro = of_a ? ra : rb;
    const double* rx = of_a ? rb : ra;
    body_state so, sx;
    if (!end) {
      so = load_state(ro), sx = load_state(rx);
    } else if (ccd) {
      so = extrapolate(load_state(ro), mt, ro + 28);
      sx = extrapolate(load_state(rx), mt, rx + 28);
    } else {
      so = load_state(ro + 14), sx = load_state(rx + 14);
    }
    double* pr = pair_rec + (size_t)PR_STRIDE * p;
    double To[16];
    pose(so.q, so.t, To);
    double* dst = pr + (of_a ? (end ? PR_TA1 : PR_TA0) : (end ? PR_TB1 : PR_TB0));
#pragma unroll
    for (int k = 0; k < 16; ++k) dst[k] = To[k];
    if (ccd) {
      double Tx[16], inv[16];
      pose(sx.q, sx.t, Tx);
      inv44(Tx, inv);
      mul44(inv, To, pr + (of_a ? (end ? PR_MAB1 : PR_MAB0) : (end ? PR_MBA1 : PR_MBA0)));
      if (end) {
        double Iw[9];
        world_inv_inertia(so.q, ro + 28, Iw);
        const v3<double> w = mul33v(Iw, so.L);
        double* k = pr + (of_a ? PR_VA : PR_VB);
        k[0] = so.v.x, k[1] = so.v.y, k[2] = so.v.z;
        k[3] = w.x, k[4] = w.y, k[5] = w.z;  // PR_WA / PR_WB
        k[6] = so.t.x, k[7] = so.t.y, k[8] = so.t.z;  // PR_XA / PR_XB
      }
    }
    if (q == 0) {
      pr[PR_GEOA] = ra[37];
      pr[PR_GEOB] = rb[37];
      pr[PR_MAXT] = mt;
      pr[PR_SEARCH] = (double)(float)(ra[37] + rb[37]);  // a float in the reference
      pair_info[4 * p + 0] = a;
      pair_info[4 * p + 1] = b;
      pair_info[4 * p + 2] = tree_off[body_tree[a]];
      pair_info[4 * p + 3] = tree_off[body_tree[b]];
      D.soup[p] = NP{(uint32_t)p, 0u, 0u, 0u, 0ull};
      D.pass_min_s[p] = ~0ull, D.pass_min_c[p] = ~0ull;
      D.pass_flag_s[p] = 0, D.pass_flag_c[p] = 0;
      if (!keep_ccd) D.min_time[p] = mt, scale[p] = 1.0;
      D.pair_cnt[p] = 0, kept_cnt[p] = 0, pair_fill[p] = 0;
    }
  }
}


// Folds one pass: "min toc of every hit of this pass" enters min_time when a
// leaf-leaf node pair of the pair had hits (full_tree_comparison.h:106-108,167-169).
__device__ __forceinline__ void fold_pass(const Dev& D, const double* pair_rec, unsigned long long* pass_min,
                                          int* pass_flag) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < D.n_pairs; p += gridDim.x * blockDim.x) {
    if (pass_flag[p] && pass_min[p] != ~0ull) {
      double t = __longlong_as_double((long long)pass_min[p]) * pair_rec[(size_t)PR_STRIDE * p + PR_MAXT];
      if (t < D.min_time[p]) D.min_time[p] = t;
    }
    pass_min[p] = ~0ull;
    pass_flag[p] = 0;
  }
}
// Keeps the CCD phase's sizes (and its contact count in the reserved field of
// level 0) next to the discrete phase's, for sizing the capacities.
__device__ __forceinline__ void save_ccd_sizes(Dev& D) {
  const int n = (MAX_LEVELS + 1) * LV_STRIDE;
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    D.lvl[n + i] = i == LV_STRIDE - 1 ? D.cnt[C_CONTACTS] : D.lvl[i];
}
// Closing fold of the CCD phase; block 0 also keeps the phase's level sizes
// and contact count in the second half of the level table.
// In a step it also records each cluster's first pairs (accumulators
// initialised by k_setup).
__global__ void k_fold_final(Dev D, ClusterAcc CA) {
  fold_pass(D, D.pair_rec, D.pass_min_s, D.pass_flag_s);
  fold_pass(D, D.pair_rec, D.pass_min_c, D.pass_flag_c);
  if (blockIdx.x == 0) save_ccd_sizes(D);
  if (CA.step_bits)
    cluster_first_range(CA, D.n_pairs, (long long)blockIdx.x * blockDim.x + threadIdx.x,
                        (long long)gridDim.x * blockDim.x);
}

// --------------------------------------------------------- CCD soup level
// Swept box of a triangle: bbox at the start pose expanded by bbox at the end
// pose, then by geo_eps (continuous_comparison.h:804-812); lo xyz, hi xyz.
__device__ __forceinline__ void swept_box(const tri<double>& t, const double* m0, const double* m1, double g,
                                          double* o) {
  const tri<double> u = xform(m0, t), v = xform(m1, t);
  const double ux[3] = {u.a.x, u.b.x, u.c.x}, uy[3] = {u.a.y, u.b.y, u.c.y}, uz[3] = {u.a.z, u.b.z, u.c.z};
  const double vx[3] = {v.a.x, v.b.x, v.c.x}, vy[3] = {v.a.y, v.b.y, v.c.y}, vz[3] = {v.a.z, v.b.z, v.c.z};
  const double* U[3] = {ux, uy, uz};
  const double* V[3] = {vx, vy, vz};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double lo = smin(smin(U[k][0], U[k][1]), U[k][2]), hi = smax(smax(U[k][0], U[k][1]), U[k][2]);
    const double lo2 = smin(smin(V[k][0], V[k][1]), V[k][2]), hi2 = smax(smax(V[k][0], V[k][1]), V[k][2]);
    o[k] = smin(lo, lo2) - g, o[3 + k] = smax(hi, hi2) + g;
  }
}
// One wavefront per soup node pair (chunked ranges), lane = triangle pair:
// swept-AABB overlap mask, its popcount and the block partial; also folds the
// connected pass of the previous level.  Lanes 0-7 build the swept boxes of
// bucket a's triangles, lanes 8-15 bucket b's, into LDS; each triangle pair
// then compares two of them.
__global__ void __launch_bounds__(256) k_soup_cull(Dev D) {
  __shared__ double s_box[4][16][6];
  fold_pass(D, D.pair_rec, D.pass_min_c, D.pass_flag_c);
  const long long n = D.n_soup();
  const int lane = lane_id();
  long long rb, re;
  chunk_of(n, rb, re);
  long long acc = 0;
  // the wavefront index made visibly uniform: the node pair's record, nodes
  // and poses then come through scalar loads into scalar registers
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (long long w = rb + wv; w < re; w += (blockDim.x >> 6)) {
    NP np = D.soup[w];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    if (lane < 16) {
      const bool side = lane >= 8;
      const int k = lane & 7;
      if (k < (side ? nb.ntris : na.ntris)) {
        const double* bk = D.buckets + (size_t)BUCKET * (side ? pi[3] + nb.bucket : pi[2] + na.bucket);
        swept_box(load_tri<double>(bk, k), side ? pr + PR_TB0 : pr + PR_TA0, side ? pr + PR_TB1 : pr + PR_TA1,
                  side ? pr[PR_GEOB] : pr[PR_GEOA], s_box[wv][lane]);
      }
    }
    wave_sync_lds();
    int i = lane >> 3, j = lane & 7;
    bool ok = i < na.ntris && j < nb.ntris;
    bool hit = false;
    if (ok) {
      const double* A = s_box[wv][i];
      const double* B = s_box[wv][8 + j];
      hit = true;
      for (int k = 0; k < 3; ++k) {  // bbox.h:8-23, a is "this"
        bool o = (B[k] < A[k]) ? (B[3 + k] > A[k]) : (B[k] < A[3 + k]);
        hit = hit && o;
      }
    }
    wave_sync_lds();  // read before the next node pair overwrites
    const unsigned long long m = __ballot(hit);
    if (lane == 0) {
      D.masks[w] = m;
      D.cnt_a[w] = __popcll(m);
      D.np_hit[w] = 0;
      acc += __popcll(m);
    }
  }
  // The overlapping triangle pairs of the block's range to the level's item
  // list: one slot range per block (one atomic), in node-pair order within
  // it; the list's order is free (contacts carry their emission keys).
  __shared__ int toff[256];
  __shared__ unsigned long long s_base;
  const long long tot = block_sum(acc);
  if (threadIdx.x == 0) {
    unsigned long long base = 0;
    if (tot) {
      base = atomicAdd(&D.lvl[D.level * LV_STRIDE + LV_ITEMS], (unsigned long long)tot);
      atomicAdd(&D.cnt[C_LVL_ITEMS], (unsigned long long)tot);
      if ((long long)(base + tot) > D.cap_res) atomicOr(&D.cnt[C_OVERFLOW], 1ull);
    }
    s_base = base;
  }
  __syncthreads();
  if (tot == 0) return;
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  long long carry = (long long)s_base;
  for (long long t0 = rb; t0 < re; t0 += blockDim.x) {
    const long long w = t0 + threadIdx.x;
    int tile_total;
    const int ex = block_scan(w < re ? D.cnt_a[w] : 0, tile_total);
    toff[threadIdx.x] = ex;
    __syncthreads();
    const int in_tile = (int)((re - t0) < (long long)blockDim.x ? (re - t0) : (long long)blockDim.x);
    for (int j = wid; j < in_tile; j += nw) {
      const long long ww = t0 + j;
      const unsigned long long m = D.masks[ww];
      if ((m >> lane) & 1ull) {
        const long long slot = carry + toff[j] + __popcll(m & lanes_below());
        if (slot < D.cap_res) D.items[slot] = Item{(uint32_t)ww, (uint32_t)lane};
      }
    }
    carry += tile_total;
    __syncthreads();
  }
}

__device__ __forceinline__ void store_res(double* r, double d, v3<double> P, v3<double> Q, double toc) {
  r[0] = d, r[1] = toc, r[2] = P.x, r[3] = P.y, r[4] = P.z, r[5] = Q.x, r[6] = Q.y, r[7] = Q.z;
}

// The FP32 form's float copy of the buckets (FB_STRIDE floats per triangle)
// and every triangle's soup_tri planes (ST_STRIDE).
__global__ void __launch_bounds__(256) k_bucket_f32(const double* buckets, long long n_nodes, float* fb, float* st) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < 8 * n_nodes;
       t += (long long)gridDim.x * blockDim.x) {
    const long long node = t >> 3;
    const int k = (int)(t & 7);
    const double* bk = buckets + (size_t)BUCKET * node;
    const tri<double> T = load_tri<double>(bk, k);
    const double eps = bk[72 + k];
    float4* o = reinterpret_cast<float4*>(fb + (size_t)FB_STRIDE * t);
    o[0] = make_float4((float)T.a.x, (float)T.a.y, (float)T.a.z, (float)T.b.x);
    o[1] = make_float4((float)T.b.y, (float)T.b.z, (float)T.c.x, (float)T.c.y);
    o[2] = make_float4((float)T.c.z, 0.f, __int_as_float(__double2loint(eps)), __int_as_float(__double2hiint(eps)));
    const SoupTri S = soup_tri(T);
    float4* q = reinterpret_cast<float4*>(st + (size_t)ST_STRIDE * t);
    q[0] = make_float4(S.n[0], S.n[1], S.n[2], S.d);
    q[1] = make_float4(S.m[0], S.m[1], S.m[2], S.m[3]);
    q[2] = make_float4(S.m[4], S.m[5], S.m[6], S.m[7]);
    q[3] = make_float4(S.m[8], S.o[0], S.o[1], S.o[2]);
    q[4] = make_float4(S.eo, 0.f, 0.f, 0.f);
  }
}
// A triangle's float corners and its soup_tri planes from the FP32 copies
// (global triangle index = node * 8 + k).
__device__ __forceinline__ void load_corners_f(const float* fb, size_t t, float* c9) {
  const float4* f = reinterpret_cast<const float4*>(fb + (size_t)FB_STRIDE * t);
  const float4 q0 = f[0], q1 = f[1], q2 = f[2];
  c9[0] = q0.x, c9[1] = q0.y, c9[2] = q0.z, c9[3] = q0.w, c9[4] = q1.x, c9[5] = q1.y, c9[6] = q1.z, c9[7] = q1.w;
  c9[8] = q2.x;
}
__device__ __forceinline__ SoupTri load_soup_tri(const float* st, size_t t) {
  const float4* f = reinterpret_cast<const float4*>(st + (size_t)ST_STRIDE * t);
  const float4 q0 = f[0], q1 = f[1], q2 = f[2], q3 = f[3], q4 = f[4];
  SoupTri S;
  S.n[0] = q0.x, S.n[1] = q0.y, S.n[2] = q0.z, S.d = q0.w;
  S.m[0] = q1.x, S.m[1] = q1.y, S.m[2] = q1.z, S.m[3] = q1.w, S.m[4] = q2.x, S.m[5] = q2.y, S.m[6] = q2.z;
  S.m[7] = q2.w, S.m[8] = q3.x, S.o[0] = q3.y, S.o[1] = q3.z, S.o[2] = q3.w, S.eo = q4.x;
  return S;
}
// ------------------------------------------- FP32 culls of CCD soup sub-tests
// FP32 form: one thread per soup item (triangle pair) runs the FP32 bounds of
// its 15 sub-tests (k_soup_keep), keeps a 15-bit mask of the survivors per
// item and appends them to two compact lists (point and edge sub-tests, the
// leaf level's item lists, which this level's connected pass fills only
// later); the FP64 sub-tests then run over those lists only, in full
// wavefronts, and k_soup_reduce selects over the surviving slots.  A culled
// sub-test provably reports no distance below the search distance, so it
// cannot be the winner of a hit (its slot would hold +inf, toc 1, which the
// selection never takes).  The bounds (d2_cull.hpp) are those of the leaf
// level: point-triangle by the plane side and the edge half-planes of the
// triangle in its frame (soup_point_keep), edge-edge by cull_edge_edge on
// the lerped world-space edges.  Soup triangles are surrogates, so their
// plane and edge normals are computed per item in FP64 (soup_tri, rounded
// once to float) and the lateral margin adds the FP64 barycentric acceptance
// bound of that triangle.
constexpr unsigned SOUP_PT_BITS = 0xC63u;   // slots 5k and 5k+1 (point-triangle, k = 0..2)
constexpr unsigned SOUP_EE_BITS = 0x739Cu;  // slots 5k+2+l (edge k of a, edge l of b)
// FP32 form: the survivor masks of the level's soup items (soup_cull_mask's
// parts: point slots per item; edge slots by the triangle-level test per item
// and, for the items it leaves, the per-edge culls packed over the
// wavefront, nine per item, with the items' world corners staged in LDS),
// then the two compact sub-test lists (wavefront prefix sums, one atomic per
// wavefront and list; the lists' order is free, results go to fixed slots).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) k_soup_keep(Dev D) {
  __shared__ float s_w[4][64][36];        // per wavefront: world corners of the items left for the per-edge culls
  __shared__ float s_thr[4][64];          // their search distances
  __shared__ float s_ex[4][64];           // and their transforms' input magnitudes
  __shared__ unsigned char s_fl[4][64];   // their lanes
  __shared__ unsigned s_em[4][64];        // per lane: surviving edge slots
  const long long n = D.n_res();
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  unsigned long long* cpt = D.lvl + D.level * LV_STRIDE + LV_SPT;
  unsigned long long* cee = D.lvl + D.level * LV_STRIDE + LV_SEE;
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < n; b0 += (long long)gridDim.x * blockDim.x) {
    const long long it = b0 + threadIdx.x;
    bool need = false;
    const double* pr = nullptr;
    size_t ga = 0, gb = 0;  // global triangle indices (node * 8 + k) of the item's two triangles
    if (it < n) {
      const Item item = D.items[it];
      const NP np = D.soup[item.np];
      const int32_t* pi = D.pair_info + 4 * np.pair;
      pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
      ga = 8 * (size_t)(pi[2] + D.nodes[pi[2] + np.na].bucket) + (item.sub >> 3);
      gb = 8 * (size_t)(pi[3] + D.nodes[pi[3] + np.nb].bucket) + (item.sub & 7);
    }
    {  // the triangle-level edge test; the corners of the items it leaves to LDS
      float w[12][3];
      float thr = 0.f, extra = 0.f;
      if (it < n) {
        float ca[9], cb[9];
        load_corners_f(D.fbuckets, ga, ca), load_corners_f(D.fbuckets, gb, cb);
        thr = (float)pr[PR_SEARCH];
        extra = soup_world(pr + PR_TA0, pr + PR_TA1, pr + PR_TB0, pr + PR_TB1, ca, cb, w);
        need = !cull_tri_edges(w, w + 3, w + 6, w + 9, thr, extra);
      }
      const unsigned long long tm = __ballot(need);
      if (need) {
        const int r = __popcll(tm & lanes_below());
#pragma unroll
        for (int q = 0; q < 12; ++q)
#pragma unroll
          for (int c = 0; c < 3; ++c) s_w[wv][r][3 * q + c] = w[q][c];
        s_thr[wv][r] = thr;
        s_ex[wv][r] = extra;
        s_fl[wv][r] = (unsigned char)lane;
      }
    }
    s_em[wv][lane] = 0u;
    unsigned mask = 0u;  // the point slots (the corners are read again: fewer live registers)
    if (it < n) {
      float ca[9], cb[9];
      load_corners_f(D.fbuckets, ga, ca), load_corners_f(D.fbuckets, gb, cb);
      mask = soup_point_mask(pr + PR_MAB0, pr + PR_MAB1, pr + PR_MBA0, pr + PR_MBA1, ca, cb, load_soup_tri(D.stris, ga),
                             load_soup_tri(D.stris, gb), (float)pr[PR_SEARCH]);
    }
    wave_sync_lds();
    const int jobs = 9 * __popcll(__ballot(need));
    for (int q0 = 0; q0 < jobs; q0 += 64) {
      const int q = q0 + lane;
      if (q < jobs) {
        const int r = q / 9, e = q % 9, k = e / 3, l = e % 3;
        if (soup_edge_keep(reinterpret_cast<const float(*)[3]>(s_w[wv][r]), k, l, s_thr[wv][r], s_ex[wv][r]))
          atomicOr(&s_em[wv][s_fl[wv][r]], 1u << (5 * k + 2 + l));
      }
    }
    wave_sync_lds();
    mask |= s_em[wv][lane];
    if (it < n) D.res_mask[it] = mask;
    const int mp = __popc(mask & SOUP_PT_BITS), me = __popc(mask & SOUP_EE_BITS);
    const int ip = wave_incl_scan(mp), ie = wave_incl_scan(me);
    unsigned long long bp = 0ull, be = 0ull;
    if (lane == 63) {
      if (ip) bp = atomicAdd(cpt, (unsigned long long)ip);
      if (ie) be = atomicAdd(cee, (unsigned long long)ie);
      if ((long long)(bp + ip) > D.cap_items || (long long)(be + ie) > D.cap_items) atomicOr(&D.cnt[C_OVERFLOW], 1ull);
    }
    long long sp = (long long)readlane_u64(bp, 63) + (ip - mp), se = (long long)readlane_u64(be, 63) + (ie - me);
    for (unsigned m = mask; m; m &= m - 1u) {
      const int s = __ffs(m) - 1;
      const bool pt = (SOUP_PT_BITS >> s) & 1u;
      const long long slot = pt ? sp++ : se++;
      if (slot < D.cap_items) (pt ? D.pt_items : D.ee_items)[slot] = Item{(uint32_t)it, (uint32_t)s};
    }
    wave_sync_lds();  // LDS read before the next round overwrites it
  }
}

// Point-triangle sub-test of soup item it, slot 5k+dir (k: vertex index,
// dir 0: vertex of a vs triangle of b, dir 1: vertex of b vs triangle of a).
__device__ __forceinline__ void soup_point_one(Dev& D, long long it, int k, int dir) {
  Item item = D.items[it];
  NP np = D.soup[item.np];
  const int32_t* pi = D.pair_info + 4 * np.pair;
  const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
  const NodeRec na = D.nodes[pi[2] + np.na];
  const NodeRec nb = D.nodes[pi[3] + np.nb];
  int i = item.sub >> 3, j = item.sub & 7;
  const double* ba = D.buckets + (size_t)BUCKET * (pi[2] + na.bucket);
  const double* bb = D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket);
  v3<double> P, Q;
  double toc = 1.0, d;
  if (dir == 0)
    d = point_tri(load_corner(ba, i, k), pr + PR_MAB0, pr + PR_MAB1, load_tri<double>(bb, j), pr + PR_TB0, pr + PR_TB1,
                  P, Q, toc);
  else
    d = point_tri(load_corner(bb, j, k), pr + PR_MBA0, pr + PR_MBA1, load_tri<double>(ba, i), pr + PR_TA0, pr + PR_TA1,
                  P, Q, toc);
  store_res(D.res + ((size_t)it * RES_SLOTS + 5 * k + dir) * 8, d, P, Q, toc);
}

// Edge-edge sub-test of soup item it, slot 5k+2+l (edge k of a, edge l of b).
__device__ __forceinline__ void soup_edge_one(Dev& D, long long it, int k, int l) {
  Item item = D.items[it];
  NP np = D.soup[item.np];
  const int32_t* pi = D.pair_info + 4 * np.pair;
  const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
  const NodeRec na = D.nodes[pi[2] + np.na];
  const NodeRec nb = D.nodes[pi[3] + np.nb];
  int i = item.sub >> 3, j = item.sub & 7;
  const double* ba = D.buckets + (size_t)BUCKET * (pi[2] + na.bucket);
  const double* bb = D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket);
  v3<double> a0 = load_corner(ba, i, k), a1 = load_corner(ba, i, (k + 1) % 3);
  v3<double> b0 = load_corner(bb, j, l), b1 = load_corner(bb, j, (l + 1) % 3);
  v3<double> P, Q;
  double toc = 1.0;
  double d = edge_edge(xform(pr + PR_TA0, a0), xform(pr + PR_TA0, a1), xform(pr + PR_TA1, a0), xform(pr + PR_TA1, a1),
                       xform(pr + PR_TB0, b0), xform(pr + PR_TB0, b1), xform(pr + PR_TB1, b0), xform(pr + PR_TB1, b1),
                       P, Q, toc);
  store_res(D.res + ((size_t)it * RES_SLOTS + 5 * k + 2 + l) * 8, d, P, Q, toc);
}

// The point sub-tests of the level: every item's six (FP64 form), or the
// FP32 form's surviving list.
__device__ __forceinline__ void soup_point_body(Dev& D, long long blk, long long nblk) {
  if (D.fp32) {
    const long long n = D.n_lv(LV_SPT);
    for (long long t = blk * blockDim.x + threadIdx.x; t < n; t += nblk * blockDim.x) {
      const Item e = D.pt_items[t];
      soup_point_one(D, e.np, (int)e.sub / 5, (int)e.sub % 5);
    }
    return;
  }
  const long long total = D.n_res() * 6;
  for (long long t = blk * blockDim.x + threadIdx.x; t < total; t += nblk * blockDim.x) {
    const int s = (int)(t % 6);
    soup_point_one(D, t / 6, s >> 1, s & 1);
  }
}

// The edge sub-tests of the level: every item's nine (FP64 form), or the
// FP32 form's surviving list.
__device__ __forceinline__ void soup_edge_body(Dev& D, long long blk, long long nblk) {
  if (D.fp32) {
    const long long n = D.n_lv(LV_SEE);
    for (long long t = blk * blockDim.x + threadIdx.x; t < n; t += nblk * blockDim.x) {
      const Item e = D.ee_items[t];
      soup_edge_one(D, e.np, (int)e.sub / 5, (int)e.sub % 5 - 2);
    }
    return;
  }
  const long long total = D.n_res() * 9;
  for (long long t = blk * blockDim.x + threadIdx.x; t < total; t += nblk * blockDim.x) {
    const int s = (int)(t % 9);
    soup_edge_one(D, t / 9, s / 3, s % 3);
  }
}

// Both kinds of sub-tests of the level's soup items in one launch: blocks
// [0, g/2) take the point-triangle tests, [g/2, g) the edge-edge tests, so no
// wavefront mixes the two code paths.
__global__ void __launch_bounds__(256) k_soup_sub(Dev D) {
  const int h = (int)(gridDim.x >> 1);
  if ((int)blockIdx.x < h) soup_point_body(D, blockIdx.x, h);
  else soup_edge_body(D, blockIdx.x - h, gridDim.x - h);
}

// triTriCCDLinear's selection over its 15 sub-results in the reference order
// (continuous_comparison.h:680-779), hit test (continuous_comparison.h:817),
// contact emission for leaf-leaf node pairs.
__global__ void __launch_bounds__(256) k_soup_reduce(Dev D) {
  const long long count = D.n_res();
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long long)gridDim.x * blockDim.x) {
    Item item = D.items[t];
    const double* r = D.res + (size_t)t * RES_SLOTS * 8;
    double best = 1.0 / 0.0, toc = 1.0;
    int win = -1;
    const unsigned live = D.fp32 ? D.res_mask[t] : (1u << RES_SLOTS) - 1u;  // FP32 form: culled slots hold nothing
    for (int s = 0; s < RES_SLOTS; ++s) {
      if (!((live >> s) & 1u)) continue;
      double d = r[8 * s], tc = r[8 * s + 1];
      if (d < best || (d == best && tc < toc)) best = d, toc = tc, win = s;
    }
    NP np = D.soup[item.np];
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    bool hit = best < pr[PR_SEARCH];
    if (!hit) continue;
    atomicOr(&D.np_hit[item.np], 1);
    atomic_min_pos(&D.pass_min_s[np.pair], toc);
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    if (na.inner || nb.inner) continue;
    D.pass_flag_s[np.pair] = 1;
    unsigned long long slot = atomicAdd(&D.cnt[C_CONTACTS], 1ull);
    if ((long long)slot >= D.cap_contacts) {
      atomicOr(&D.cnt[C_OVERFLOW], D.discrete ? 2ull : 1ull);
      continue;
    }
    atomicAdd(&D.pair_cnt[np.pair], 1);
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
// The sub-test items are written by k_level_write.
__device__ __forceinline__ void emit_leaf(Dev& D, const NP& np, const double* pr, int kind, int ia, int ib, int sub,
                                          v3<double> A, v3<double> B, double toc) {
  unsigned long long slot = atomicAdd(&D.cnt[C_CONTACTS], 1ull);
  if ((long long)slot >= D.cap_contacts) {
    atomicOr(&D.cnt[C_OVERFLOW], D.discrete ? 2ull : 1ull);
    return;
  }
  atomicAdd(&D.pair_cnt[np.pair], 1);
  ContactRec c;
  c.key = make_key(2 * D.level + 1, 0, np.key, sub);
  c.pair = (int32_t)np.pair, c.na = (int32_t)np.na, c.nb = (int32_t)np.nb, c.kind = kind;
  c.ia = ia, c.ib = ib, c.level = 2 * D.level + 1, c.pad = 0;
  c.r[0] = A.x, c.r[1] = A.y, c.r[2] = A.z, c.r[3] = B.x, c.r[4] = B.y, c.r[5] = B.z;
  c.r[6] = pr[PR_GEOA], c.r[7] = pr[PR_GEOB], c.r[8] = 0.0, c.r[9] = 0.0, c.r[10] = toc;
  D.contacts[slot] = c;
}

__device__ __forceinline__ void leaf_point_body(Dev& D, long long blk, long long nblk) {
  const long long count = D.n_lv(LV_PT);
  for (long long t = blk * blockDim.x + threadIdx.x; t < count; t += nblk * blockDim.x) {
    Item item = D.pt_items[t];
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
      d = point_tri(load_corner(ba, v / 3, v % 3), pr + PR_MAB0, pr + PR_MAB1, load_tri<double>(bb, j), pr + PR_TB0,
                    pr + PR_TB1, P, Q, toc);
      kind = 1, ia = v, ib = j;
    } else {
      int s2 = s - 3 * n * m, v = s2 / n, i = s2 % n;
      d = point_tri(load_corner(bb, v / 3, v % 3), pr + PR_MBA0, pr + PR_MBA1, load_tri<double>(ba, i), pr + PR_TA0,
                    pr + PR_TA1, P, Q, toc);
      kind = 2, ia = i, ib = v;
    }
    if (d < pr[PR_SEARCH]) {
      D.pass_flag_c[np.pair] = 1;
      atomic_min_pos(&D.pass_min_c[np.pair], toc);
      if (kind == 1) emit_leaf(D, np, pr, kind, ia, ib, s, P, Q, toc);
      else emit_leaf(D, np, pr, kind, ia, ib, s, Q, P, toc);  // continuous_comparison.h:876
    }
  }
}

__device__ __forceinline__ void leaf_edge_body(Dev& D, long long blk, long long nblk) {
  const long long count = D.n_lv(LV_EE);
  for (long long t = blk * blockDim.x + threadIdx.x; t < count; t += nblk * blockDim.x) {
    Item item = D.ee_items[t];
    NP np = D.conn[item.np];
    const int32_t* pi = D.pair_info + 4 * np.pair;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
    const NodeRec na = D.nodes[pi[2] + np.na];
    const NodeRec nb = D.nodes[pi[3] + np.nb];
    int s = (int)item.sub, n = na.ntris, m = nb.ntris;
    int ea = s / (3 * m), eb = s % (3 * m);
    const double* ba = D.buckets + (size_t)BUCKET * (pi[2] + na.bucket);
    const double* bb = D.buckets + (size_t)BUCKET * (pi[3] + nb.bucket);
    v3<double> a0 = load_corner(ba, ea / 3, ea % 3), a1 = load_corner(ba, ea / 3, (ea % 3 + 1) % 3);
    v3<double> b0 = load_corner(bb, eb / 3, eb % 3), b1 = load_corner(bb, eb / 3, (eb % 3 + 1) % 3);
    v3<double> P, Q;
    double toc = 1.0;
    double d = edge_edge(xform(pr + PR_TA0, a0), xform(pr + PR_TA0, a1), xform(pr + PR_TA1, a0), xform(pr + PR_TA1, a1),
                         xform(pr + PR_TB0, b0), xform(pr + PR_TB0, b1), xform(pr + PR_TB1, b0), xform(pr + PR_TB1, b1),
                         P, Q, toc);
    if (d < pr[PR_SEARCH]) {
      D.pass_flag_c[np.pair] = 1;
      atomic_min_pos(&D.pass_min_c[np.pair], toc);
      emit_leaf(D, np, pr, 3, ea, eb, 6 * n * m + s, Q, P, toc);  // continuous_comparison.h:903
    }
  }
}
// The leaf-leaf sub-tests of the level in one launch, split by block as in
// k_soup_sub.
__global__ void __launch_bounds__(256) k_leaf_sub(Dev D) {
  const int h = (int)(gridDim.x >> 1);
  if ((int)blockIdx.x < h) leaf_point_body(D, blockIdx.x, h);
  else leaf_edge_body(D, blockIdx.x - h, gridDim.x - h);
}

// ------------------------------------------- ordered enumeration/expansion
// ------------------------------------------------ FP32 conservative cull
// The discrete phase's FP32 classification of triangle pairs (certain miss,
// FP64 test, certain hit) is in d2_cull.hpp: cull_prep, cull_margin,
// cull_pair.

// One wavefront per node pair (chunked ranges): which triangle pairs need the
// FP64 test.  FP64 mode keeps every valid pair; FP32 mode drops certain
// misses, and for node pairs with an inner node (they emit no contacts, only
// "any hit" matters) a certain hit settles the node pair without FP64 work.
__device__ __forceinline__ bool expand_side(const NodeRec& na, const NodeRec& nb) {
  return na.inner && (!nb.inner || na.depth < nb.depth);  // true: advance a
}
// Expansion of a node pair of the discrete list, packed: bit 0 advance a,
// bits 1-4 child count, bits 5-31 first child (0: two leaves, no children).
__device__ __forceinline__ int expansion_desc(const NodeRec& na, const NodeRec& nb) {
  if (!(na.inner || nb.inner)) return 0;
  const bool adv_a = expand_side(na, nb);
  const NodeRec& par = adv_a ? na : nb;
  return (adv_a ? 1 : 0) | (par.nchild << 1) | (par.child_first << 5);
}
__device__ __forceinline__ int rl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
// Inclusive prefix sum over the wavefront (every lane active).
// Discrete lists: the children of a node pair that hits or may hit, at slot,
// slot + 1, ... of the next list, each naming its parent: parent 0 for a
// certain hit, list index + 1 when the parent's FP64 tests (k_tri_dist, its
// hit flag) decide; the next level drops the children of parents that did
// not hit.  The discrete phase's contacts carry their emission keys (child
// index paths) and are ordered by them afterwards, so its lists need no
// breadth-first order and children take slots by atomics.
// Only the children in cmask are written (the others are provably apart:
// child_mask); the emission keys keep every child's own index.
__device__ __forceinline__ void write_children(const Dev& D, long long slot, const NP& np, int desc, uint32_t parent,
                                               unsigned cmask) {
  const bool adv_a = desc & 1;
  const int nc = (desc >> 1) & 15, first = desc >> 5;
  for (int c = 0; c < nc; ++c) {
    if (!((cmask >> c) & 1u)) continue;
    const uint32_t ca = adv_a ? (uint32_t)(first + c) : np.na;
    const uint32_t cb = adv_a ? np.nb : (uint32_t)(first + c);
    if (slot < D.cap_np) D.soup_next[slot] = NP{np.pair, ca, cb, parent, np.key * 8ull + (unsigned long long)c};
    ++slot;
  }
}
// The children of a node pair worth writing: all of them, or (pruning) those
// nodes_apart does not prove contact-free.  offa/offb: the two trees' first
// node records.
template <bool F32>
__device__ __forceinline__ unsigned child_mask(const Dev& D, const NP& np, int desc, int offa, int offb) {
  const int nc = (desc >> 1) & 15;
  unsigned m = (1u << nc) - 1u;
  if (!D.prune) return m;
  const bool adv_a = desc & 1;
  const int first = desc >> 5;
  const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
  for (int c = 0; c < nc; ++c) {
    const size_t ga = (size_t)offa + (adv_a ? (uint32_t)(first + c) : np.na);
    const size_t gb = (size_t)offb + (adv_a ? np.nb : (uint32_t)(first + c));
    const float* ba = D.nbound + NB_STRIDE * ga;
    const float* bb = D.nbound + NB_STRIDE * gb;
    // the FP32 form's mask kernel runs at five wavefronts per SIMD: the FP32
    // sphere test only (the boxes' test, or FP64 poses, would spill there);
    // elsewhere both FP64 tests
    const bool apart = F32 ? spheres_apart_f32(pr + PR_TA1, pr + PR_TB1, ba, bb, pr[PR_SEARCH])
                           : spheres_apart(pr + PR_TA1, pr + PR_TB1, ba, bb, pr[PR_SEARCH]) ||
                                 nodes_apart(pr + PR_TA1, pr + PR_TB1, ba, bb, pr[PR_SEARCH]);
    if (apart) m &= ~(1u << c);
  }
  return m;
}
static_assert(sizeof(CullTri) == 112, "prepared triangle: 28 dwords, a conflict-free LDS stride");
// Five wavefronts per SIMD (at most 96 VGPRs, no spills) instead of the four
// its natural 101 VGPRs allow: the kernel waits on dependent loads, and more
// resident wavefronts hide more of them.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) k_tri_mask(Dev D) {
  __shared__ int s_tot[2][4], s_ch[2][4], s_df[2][4];
  __shared__ CullTri s_tri[4][4][16];  // per wavefront: 4 node pairs' 8 + 8 prepared triangles
  __shared__ unsigned long long s_base[2], s_chb[2], s_dfb[2];
  const int lane = lane_id(), nw = blockDim.x >> 6, wid = threadIdx.x >> 6;
  const long long n = D.n_soup();
  long long rb, re;
  chunk_of(n, rb, re);
  // The block walks its range 64 * nw node pairs at a time, wavefront wid
  // taking rb + wid + nw * k; lane l fetches the record, node offsets and node
  // records of the wave's l-th node pair in one go, so the per-node-pair
  // dependent loads reduce to its pose and buckets.
  int it = 0;
  for (long long b0 = rb; b0 < re; b0 += 64ll * nw, ++it) {
    const long long w0 = b0 + wid;
    const long long wl = w0 + (long long)nw * lane;
    int l_pair = 0, l_offa = 0, l_offb = 0, l_bka = 0, l_bkb = 0, l_nta = 0, l_ntb = 0, l_inner = 0, l_desc = 0;
    double l_search = 0.0;
    NP l_np{};
    if (wl < re) {
      const NP np = D.soup[wl];
      const int32_t* pi = D.pair_info + 4 * np.pair;
      l_pair = (int)np.pair, l_offa = pi[2], l_offb = pi[3];
      // a child of a parent that did not hit, or a node pair whose leaves are
      // provably apart: no work, no children.  Spheres are tested when a
      // child is written (roots here); the boxes' face-axis test runs here,
      // one node pair per lane, before the node records are loaded (few
      // registers are live).
      bool live = np.pad == 0u || D.np_hit_prev[np.pad - 1u] != 0;
      if (live && D.prune) {
        const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
        const float* bA = D.nbound + NB_STRIDE * ((size_t)l_offa + np.na);
        const float* bB = D.nbound + NB_STRIDE * ((size_t)l_offb + np.nb);
        live = !((D.level == 0 && spheres_apart_f32(pr + PR_TA1, pr + PR_TB1, bA, bB, pr[PR_SEARCH])) ||
                 (D.box_test && boxes_apart_faces(pr + PR_TA1, pr + PR_TB1, bA, bB, pr[PR_SEARCH])));
      }
      if (live) {
        const NodeRec na = D.nodes[l_offa + np.na];
        const NodeRec nb = D.nodes[l_offb + np.nb];
        l_np = np;
        l_bka = na.bucket, l_bkb = nb.bucket, l_nta = na.ntris, l_ntb = nb.ntris;
        l_inner = (na.inner || nb.inner) ? 1 : 0;
        l_desc = expansion_desc(na, nb);
        l_search = D.pair_rec[(size_t)PR_STRIDE * np.pair + PR_SEARCH];
      }
    }
    const long long left = w0 < re ? (re - w0 + nw - 1) / nw : 0;
    const int batch = left < 64 ? (int)left : 64;
    unsigned long long my_keep = 0;
    int my_settled = 0;
    const int i = lane >> 3, j = lane & 7;
    for (int k0 = 0; k0 < batch; k0 += 4) {
      {
        // the triangles of four node pairs at once: lane l prepares triangle
        // (l & 15) of node pair k0 + (l >> 4) (lanes 0-7 of a group: bucket a,
        // 8-15: bucket b), so no lane idles in the preparation
        const int kp = k0 + (lane >> 4), slot = lane & 15, side = slot >> 3, kk = slot & 7;
        const int src = kp < batch ? kp : k0;
        // every lane takes part in every shuffle (a shuffle inside divergent
        // code may read an inactive source lane)
        const int pair = __shfl(l_pair, src), nta = __shfl(l_nta, src), ntb = __shfl(l_ntb, src);
        const int ga = __shfl(l_offa, src) + __shfl(l_bka, src), gb = __shfl(l_offb, src) + __shfl(l_bkb, src);
        const int gbk = side ? gb : ga;
        CullTri mine{};
        if (kp < batch && kk < (side ? ntb : nta)) {
          const double* pr = D.pair_rec + (size_t)PR_STRIDE * pair;
          const double* pm = side ? pr + PR_TB1 : pr + PR_TA1;
          float rot[9], tr[3];
          for (int r = 0; r < 3; ++r) {
            for (int q = 0; q < 3; ++q) rot[3 * r + q] = (float)pm[4 * r + q];
            tr[r] = side ? 0.f : (float)(pr[PR_TA1 + 4 * r + 3] - pr[PR_TB1 + 4 * r + 3]);
          }
          const float4* fb = reinterpret_cast<const float4*>(D.fbuckets + (size_t)FB_STRIDE * (8 * (size_t)gbk + kk));
          const float4 q0 = fb[0], q1 = fb[1], q2 = fb[2];
          const float c9[9] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x};
          cull_prep_f(rot, tr, c9, mine);
          mine.eps = __hiloint2double(__float_as_int(q2.w), __float_as_int(q2.z));
        }
        s_tri[wid][lane >> 4][slot] = mine;
        wave_sync_lds();
      }
      const int kend = k0 + 4 < batch ? k0 + 4 : batch;
      for (int k = k0; k < kend; ++k) {
        const int nta = rl(l_nta, k), ntb = rl(l_ntb, k);
        const bool inner = rl(l_inner, k) != 0;
        int cls = 0;
        if (i < nta && j < ntb) {  // operands read from LDS where they are used (fewer live registers)
          // no global load on the node pair's chain: the search distance came
          // with the node pair, the eps values with the prepared triangles
          const CullTri& A = s_tri[wid][k - k0][i];
          const CullTri& B = s_tri[wid][k - k0][8 + j];
          const double search = __longlong_as_double((long long)readlane_u64(
              (unsigned long long)__double_as_longlong(l_search), k));
          cls = cull_pair(A, B, (float)((search + A.eps) + B.eps), inner);
        }
        unsigned long long keep = __ballot(cls == 1 || (cls == 2 && !inner));
        const bool settled = inner && __ballot(cls == 2) != 0ull;
        if (settled) keep = 0ull;
        if (lane == k) my_keep = keep, my_settled = settled ? 1 : 0;
      }
      wave_sync_lds();  // read before the next group overwrites
    }
    if (lane < batch) D.np_hit[wl] = my_settled;
    // The kept triangle pairs go straight to the level's item list, and the
    // children of node pairs that hit (settled) or may hit (kept pairs) to
    // the next list: one slot range each per block and round, one atomic
    // each (order does not matter: contacts carry their emission keys).  An
    // inner node pair's items only decide whether it expands (any FP64 hit):
    // its first kept item goes to the item list, the others (deferred) to
    // k_tri_def, which tests them only while the node pair has no hit.
    const int kept = lane < batch ? __popcll(my_keep) : 0;
    const bool defer = D.defer && l_inner != 0;
    const int mine = defer && kept > 1 ? 1 : kept, mdef = kept - mine;
    const unsigned cmask =
        lane < batch && (my_settled || my_keep != 0ull) ? child_mask<true>(D, l_np, l_desc, l_offa, l_offb) : 0u;
    const int nch = __popc(cmask);
    const int incl = wave_incl_scan(mine), ich = wave_incl_scan(nch), idf = wave_incl_scan(mdef);
    const int par = it & 1;
    if (lane == 63) s_tot[par][wid] = incl, s_ch[par][wid] = ich, s_df[par][wid] = idf;
    __syncthreads();
    if (threadIdx.x == 0) {  // one atomic for both ranges (same-address atomics from every block are costly)
      int sum = 0, sch = 0, sdf = 0;
      for (int v = 0; v < nw; ++v) sum += s_tot[par][v], sch += s_ch[par][v], sdf += s_df[par][v];
      unsigned long long base = 0, chb = 0, dfb = 0;
      if (sum || sch) {
        const unsigned long long old =
            atomicAdd(&D.lvl[D.level * LV_STRIDE + LV_PACK], ((unsigned long long)sch << 32) | (unsigned)sum);
        base = old & 0xffffffffull, chb = old >> 32;
        if ((long long)(base + sum) > D.cap_items || (long long)(chb + sch) > D.cap_np)
          atomicOr(&D.cnt[C_OVERFLOW], 2ull);
      }
      if (sdf) {  // deferred items: their own range (another address)
        dfb = atomicAdd(&D.lvl[D.level * LV_STRIDE + LV_DEF], (unsigned long long)sdf);
        if ((long long)(dfb + sdf) > D.cap_items) atomicOr(&D.cnt[C_OVERFLOW], 2ull);
      }
      s_base[par] = base, s_chb[par] = chb, s_dfb[par] = dfb;
    }
    __syncthreads();
    if (nch > 0) {
      long long slot = (long long)s_chb[par] + ich - nch;
      for (int v = 0; v < wid; ++v) slot += s_ch[par][v];
      write_children(D, slot, l_np, l_desc, my_settled ? 0u : (uint32_t)(wl + 1), cmask);
    }
    long long base = (long long)s_base[par], dbase = (long long)s_dfb[par];
    for (int v = 0; v < wid; ++v) base += s_tot[par][v], dbase += s_df[par][v];
    const int excl = incl - mine, dexcl = idf - mdef;
    const int defer_i = defer ? 1 : 0;
    for (int k = 0; k < batch; ++k) {
      const unsigned long long m = readlane_u64(my_keep, k);
      if (m == 0ull) continue;
      // deferring node pair: its lowest kept lane to the item list, the rest
      // to the deferred list
      const unsigned long long first = rl(defer_i, k) ? (m & (~m + 1ull)) : m;
      const unsigned long long rest = m & ~first;
      const DItem item{(uint32_t)(w0 + (long long)nw * k), (uint32_t)rl(l_pair, k),
                       (uint32_t)(rl(l_offa, k) + rl(l_bka, k)), (uint32_t)(rl(l_offb, k) + rl(l_bkb, k)),
                       (uint32_t)lane | (rl(l_inner, k) ? 256u : 0u)};
      if ((first >> lane) & 1ull) {
        const long long slot = base + rl(excl, k) + __popcll(first & lanes_below());
        if (slot < D.cap_items) D.ditems[slot] = item;
      } else if ((rest >> lane) & 1ull) {
        const long long slot = dbase + rl(dexcl, k) + __popcll(rest & lanes_below());
        if (slot < D.cap_items) D.ditems_def[slot] = item;
      }
    }
  }
}


// The FP64 test of one item (comparison_dist.h:205-240's leaf test): the
// reference's triangle distance at the phase's poses against search + eps_a
// + eps_b.
struct ItemTest {
  v3<double> P, Q;
  double ea, eb;
  int i, j;
  bool hit;
};
__device__ __forceinline__ ItemTest item_test(const Dev& D, const DItem& item) {
  ItemTest r;
  const double* pr = D.pair_rec + (size_t)PR_STRIDE * item.pair;
  const double* ba = D.buckets + (size_t)BUCKET * item.ga;
  const double* bb = D.buckets + (size_t)BUCKET * item.gb;
  const int sub = (int)(item.sub & 63u);
  r.i = sub >> 3, r.j = sub & 7;
  const tri<double> A = xform(pr + PR_TA1, load_tri<double>(ba, r.i));
  const tri<double> B = xform(pr + PR_TB1, load_tri<double>(bb, r.j));
  const double d = tri_tri_dist(A, B, r.P, r.Q);
  r.ea = ba[72 + r.i], r.eb = bb[72 + r.j];
  r.hit = d < (pr[PR_SEARCH] + r.ea) + r.eb;
  return r;
}
// FP64 triangle-pair distances of the level's items (grid-stride).
__global__ void __launch_bounds__(256) k_tri_dist(Dev D) {
  const unsigned long long pack = D.lv(D.level, LV_PACK);
  const long long n_items = (long long)(pack & 0xffffffffull);
  const long long count = n_items < D.cap_items ? n_items : D.cap_items;
  if (blockIdx.x == 0 && threadIdx.x == 0) {  // the plain sizes, for the next level and the host
    D.lvl[D.level * LV_STRIDE + LV_ITEMS] = (unsigned long long)n_items;
    D.lvl[(D.level + 1) * LV_STRIDE + LV_SOUP] = pack >> 32;
    atomicAdd(&D.cnt[C_LVL_ITEMS], (unsigned long long)n_items);
  }
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (long long)gridDim.x * blockDim.x) {
    const DItem item = D.ditems[t];
    const ItemTest r = item_test(D, item);
    if (!r.hit) continue;
    D.np_hit[item.np] = 1;
    if (item.sub & 256u) continue;  // an inner node: the hit only expands the node pair
    unsigned long long slot = atomicAdd(&D.cnt[C_CONTACTS], 1ull);
    if ((long long)slot >= D.cap_contacts) {
      atomicOr(&D.cnt[C_OVERFLOW], D.discrete ? 2ull : 1ull);
      continue;
    }
    atomicAdd(&D.pair_cnt[item.pair], 1);
    const NP np = D.soup[item.np];
    ContactRec c;
    const double* pr = D.pair_rec + (size_t)PR_STRIDE * item.pair;
    c.key = make_key(D.level, 0, np.key, (int)(item.sub & 63u));
    c.pair = (int32_t)np.pair, c.na = (int32_t)np.na, c.nb = (int32_t)np.nb, c.kind = 0;
    c.ia = r.i, c.ib = r.j, c.level = D.level, c.pad = 0;
    c.r[0] = r.P.x, c.r[1] = r.P.y, c.r[2] = r.P.z, c.r[3] = r.Q.x, c.r[4] = r.Q.y, c.r[5] = r.Q.z;
    c.r[6] = pr[PR_GEOA] + r.ea, c.r[7] = pr[PR_GEOB] + r.eb;
    c.r[8] = D.buckets[(size_t)BUCKET * item.ga + 80 + r.i], c.r[9] = D.buckets[(size_t)BUCKET * item.gb + 80 + r.j];
    c.r[10] = 0.0;
    D.contacts[slot] = c;
  }
}

// The deferred items of inner node pairs (FP32 form, after k_tri_dist): an
// inner node pair's FP64 tests only decide whether it expands, so a deferred
// item is tested only while its node pair has no hit.  The items still
// needed are queued per wavefront in LDS and tested 64 at a time, so the
// FP64 test runs on full wavefronts; a hit sets the node pair's flag (inner
// node pairs emit no contacts).  Every wavefront walks its 64-item chunks to
// the end of the list, then drains its queue.
__global__ void __launch_bounds__(256) k_tri_def(Dev D) {
  __shared__ uint32_t s_q[4][128];
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  const long long n_def = D.lv(D.level, LV_DEF);
  const long long n = n_def < D.cap_items ? n_def : D.cap_items;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&D.cnt[C_LVL_ITEMS], (unsigned long long)n_def);
  const long long waves = (long long)gridDim.x * (blockDim.x >> 6);
  int qn = 0;  // queue length (wavefront-uniform), below 64 between chunks
  for (long long c0 = ((long long)blockIdx.x * (blockDim.x >> 6) + wid) * 64;; c0 += waves * 64) {
    const bool more = c0 < n;
    if (more) {
      const long long t = c0 + lane;
      const bool need = t < n && D.np_hit[D.ditems_def[t].np] == 0;
      const unsigned long long b = __ballot(need);
      if (need) s_q[wid][qn + __popcll(b & lanes_below())] = (uint32_t)t;
      qn += __popcll(b);
      wave_sync_lds();
    }
    while (qn >= 64 || (!more && qn > 0)) {
      const int take = qn < 64 ? qn : 64;
      if (lane < take) {
        const DItem item = D.ditems_def[s_q[wid][lane]];
        if (D.np_hit[item.np] == 0 && item_test(D, item).hit) D.np_hit[item.np] = 1;
      }
      const int rest = qn - take;  // below 64: move it to the queue's front
      const uint32_t v = lane < rest ? s_q[wid][take + lane] : 0u;
      wave_sync_lds();
      if (lane < rest) s_q[wid][lane] = v;
      wave_sync_lds();
      qn = rest;
    }
    if (!more) break;
  }
}

// FP64 discrete levels in one pass, with no item list (the reference's own
// form: every triangle pair of a node pair tested in place,
// full_tree_comparison.h:26-62 with bucketCompare, discrete_comparison.h).
// A wavefront takes 64 node pairs at a time and packs their valid triangle
// pairs (n_a * n_b each) densely over its lanes: a prefix sum over the counts,
// then each lane finds the node pair of its triangle pair by binary search
// over the sums.  The FP64 test, the contact record and its emission key are
// k_tri_dist's, so the contacts are the same bit for bit; hits set the node
// pair's flag and expansion descriptor for k_disc_expand.
struct FullSlot {
  int pair, ga, gb, ntb, excl, inner;
  long long w;
};
__global__ void __launch_bounds__(256) k_tri_full(Dev D) {
  __shared__ FullSlot s_np[4][64];
  __shared__ int s_hit[4][64];
  __shared__ unsigned long long s_tested;
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  const long long n = D.n_soup();
  const long long n_groups = (n + 63) / 64;
  const long long nwaves = (long long)gridDim.x * (blockDim.x >> 6);
  if (threadIdx.x == 0) s_tested = 0ull;
  __syncthreads();
  unsigned long long tested = 0ull;  // kept by lane 0
  for (long long g = (long long)blockIdx.x * (blockDim.x >> 6) + wid; g < n_groups; g += nwaves) {
    const long long w = g * 64 + lane;
    int cnt = 0, desc = 0;
    FullSlot f{0, 0, 0, 1, 0, 0, w};
    if (w < n) {
      const NP np = D.soup[w];
      // a child of a parent that did not hit: no work, no children
      const int32_t* pi = D.pair_info + 4 * np.pair;
      const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
      if ((np.pad == 0u || D.np_hit_prev[np.pad - 1u] != 0) &&
          !(D.prune && D.level == 0 &&
            nodes_apart(pr + PR_TA1, pr + PR_TB1, D.nbound + NB_STRIDE * ((size_t)pi[2] + np.na),
                        D.nbound + NB_STRIDE * ((size_t)pi[3] + np.nb), pr[PR_SEARCH]))) {
        const NodeRec na = D.nodes[pi[2] + np.na];
        const NodeRec nb = D.nodes[pi[3] + np.nb];
        f.pair = (int)np.pair, f.ga = pi[2] + na.bucket, f.gb = pi[3] + nb.bucket;
        f.ntb = nb.ntris > 0 ? nb.ntris : 1;
        f.inner = (na.inner || nb.inner) ? 1 : 0;
        cnt = na.ntris * nb.ntris;
        desc = expansion_desc(na, nb);
      }
    }
    const int incl = wave_incl_scan(cnt);
    f.excl = incl - cnt;
    s_np[wid][lane] = f;
    s_hit[wid][lane] = 0;
    const int total = rl(incl, 63);
    tested += (unsigned long long)total;
    wave_sync_lds();
    for (int t0 = 0; t0 < total; t0 += 64) {
      const int t = t0 + lane;
      // the owning node pair: how many inclusive sums are <= t (every lane
      // takes part in the shuffles)
      int k = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(incl, k + step - 1) <= t) k += step;
      bool hit = false, emit = false;
      int i = 0, j = 0;
      v3<double> P{}, Q{};
      double ea = 0.0, eb = 0.0;
      const FullSlot& o = s_np[wid][k];
      if (t < total) {
        const int s = t - o.excl;
        i = s / o.ntb, j = s - i * o.ntb;
        const double* pr = D.pair_rec + (size_t)PR_STRIDE * o.pair;
        const double* ba = D.buckets + (size_t)BUCKET * o.ga;
        const double* bb = D.buckets + (size_t)BUCKET * o.gb;
        const tri<double> A = xform(pr + PR_TA1, load_tri<double>(ba, i));
        const tri<double> B = xform(pr + PR_TB1, load_tri<double>(bb, j));
        const double d = tri_tri_dist(A, B, P, Q);
        ea = ba[72 + i], eb = bb[72 + j];
        hit = d < (pr[PR_SEARCH] + ea) + eb;
        if (hit) s_hit[wid][k] = 1;
        emit = hit && !o.inner;  // an inner node: the hit only expands the node pair
      }
      const long long slot = wave_alloc(emit, &D.cnt[C_CONTACTS]);
      if (!emit) continue;
      if (slot >= D.cap_contacts) {
        atomicOr(&D.cnt[C_OVERFLOW], 2ull);
        continue;
      }
      atomicAdd(&D.pair_cnt[o.pair], 1);
      const NP np = D.soup[o.w];
      const double* pr = D.pair_rec + (size_t)PR_STRIDE * o.pair;
      const double* ba = D.buckets + (size_t)BUCKET * o.ga;
      const double* bb = D.buckets + (size_t)BUCKET * o.gb;
      ContactRec c;
      c.key = make_key(D.level, 0, np.key, 8 * i + j);
      c.pair = (int32_t)np.pair, c.na = (int32_t)np.na, c.nb = (int32_t)np.nb, c.kind = 0;
      c.ia = i, c.ib = j, c.level = D.level, c.pad = 0;
      c.r[0] = P.x, c.r[1] = P.y, c.r[2] = P.z, c.r[3] = Q.x, c.r[4] = Q.y, c.r[5] = Q.z;
      c.r[6] = pr[PR_GEOA] + ea, c.r[7] = pr[PR_GEOB] + eb, c.r[8] = ba[80 + i], c.r[9] = bb[80 + j], c.r[10] = 0.0;
      D.contacts[slot] = c;
    }
    wave_sync_lds();
    if (w < n) {
      const int h = s_hit[wid][lane];
      D.np_hit[w] = h;
      if (h) D.exp_desc[w] = desc;
    }
    wave_sync_lds();  // s_np and s_hit are rewritten by the next group
  }
  if (lane == 0 && tested) atomicAdd(&s_tested, tested);
  __syncthreads();
  if (threadIdx.x == 0 && s_tested) {
    atomicAdd(&D.lvl[D.level * LV_STRIDE + LV_ITEMS], s_tested);
    atomicAdd(&D.cnt[C_LVL_ITEMS], s_tested);
  }
}

// CCD levels: children of hit node pairs (full_tree_comparison.h:26-62,
// 99-150) and the sub-tests of the connected (leaf-leaf) list, in two chunked
// passes: counts + block partials, then writes at scanned offsets, which
// keeps every list in the reference's breadth-first order (discrete levels
// are expanded by k_tri_mask, or by k_disc_expand in FP64 mode).
// ------------------------------------------- FP32 culls of CCD leaf sub-tests
// The bounds of d2_cull.hpp (soup_point_keep: plane side and edge
// half-planes with each triangle's FP64-derived normals and its
// float-barycentric acceptance bound; cull_edge_edge: separating axes).
// Per connected node pair, shared across its wavefront: lane v < 24 holds
// vertex v of a's leaf (triangle v/3, corner v%3), lanes 24-47 those of b:
// the vertex at both ends in world coordinates (relative to b's start
// position) and in the other body's frame; lanes 48-55 / 56-63 hold
// triangle k of a / b in its own frame: unit normal, plane offset, outward
// unit edge normals and offsets.  A sub-test then reads what it needs by
// shuffles.
struct LeafShare {
  float w0[3], w1[3], f0[3], f1[3];  // vertex lanes
  SoupTri t;                         // triangle lanes (soup_tri: FP64 normals, float-barycentric bound)
};
struct LeafInfo {
  int n, m;
  float thr, margin, mag;  // mag: every magnitude of the node pair's FP32 transforms
};
__device__ LeafInfo leaf_share(const Dev& D, const NP& np, LeafShare& sh) {
  const int lane = lane_id();
  const int32_t* pi = D.pair_info + 4 * np.pair;
  const double* pr = D.pair_rec + (size_t)PR_STRIDE * np.pair;
  const NodeRec na = D.nodes[pi[2] + np.na];
  const NodeRec nb = D.nodes[pi[3] + np.nb];
  const size_t ta0 = 8 * (size_t)(pi[2] + na.bucket), tb0 = 8 * (size_t)(pi[3] + nb.bucket);  // global triangle 0
  LeafInfo I;
  I.n = na.ntris, I.m = nb.ntris;
  I.thr = (float)pr[PR_SEARCH];
  float mag = 0.f;
  for (int k = 0; k < 3; ++k) sh.w0[k] = sh.w1[k] = sh.f0[k] = sh.f1[k] = 0.f;
  if (lane < 48) {
    const bool of_b = lane >= 24;
    const int v = of_b ? lane - 24 : lane;
    if (v / 3 < (of_b ? I.m : I.n)) {
      const float* pf = D.fbuckets + (size_t)FB_STRIDE * ((of_b ? tb0 : ta0) + v / 3) + 3 * (v % 3);
      const float p[3] = {pf[0], pf[1], pf[2]};  // the corner rounded to float
      const M34f W0 = to_m34f(of_b ? pr + PR_TB0 : pr + PR_TA0, pr + PR_TB0);
      const M34f W1 = to_m34f(of_b ? pr + PR_TB1 : pr + PR_TA1, pr + PR_TB0);
      const M34f F0 = to_m34f(of_b ? pr + PR_MBA0 : pr + PR_MAB0, nullptr);
      const M34f F1 = to_m34f(of_b ? pr + PR_MBA1 : pr + PR_MAB1, nullptr);
      xf_f(W0, p, sh.w0), xf_f(W1, p, sh.w1), xf_f(F0, p, sh.f0), xf_f(F1, p, sh.f1);
      mag = fmaxf(fmaxf(amax3(sh.w0), amax3(sh.w1)), fmaxf(amax3(sh.f0), amax3(sh.f1)));
      mag = fmaxf(mag, fmaxf(fmaxf(fmaxf(m34_mag(W0), m34_mag(W1)), fmaxf(m34_mag(F0), m34_mag(F1))), amax3(p)));
    }
    sh.t = SoupTri{};  // (vertex lanes: unused)
  } else {
    const bool of_b = lane >= 56;
    const int k = lane & 7;
    if (k < (of_b ? I.m : I.n)) {
      const size_t tk = (of_b ? tb0 : ta0) + k;
      float c9[9];
      load_corners_f(D.fbuckets, tk, c9);
      mag = fmaxf(amax3(c9), fmaxf(amax3(c9 + 3), amax3(c9 + 6)));
      sh.t = load_soup_tri(D.stris, tk);
    } else {
      sh.t = SoupTri{};  // n = 0: never culled
    }
  }
  for (int off = 32; off > 0; off >>= 1) mag = fmaxf(mag, __shfl_xor(mag, off));
  I.mag = mag;
  I.margin = 1e-6f + 64.f * 5.9604645e-08f * mag;
  return I;
}
__device__ __forceinline__ void sh3(const float* v, int src, float* o) {
  o[0] = __shfl(v[0], src), o[1] = __shfl(v[1], src), o[2] = __shfl(v[2], src);
}
// Point-triangle sub-test s survives (the bounds above: plane side, edge half-plane).
// Every lane of the wavefront calls this (the shuffles need all lanes).
__device__ bool leaf_point_keep(const LeafShare& sh, const LeafInfo& I, int s, bool active) {
  const int n = I.n, m = I.m;
  int vl, tl;
  if (s < 3 * n * m) {
    const int v = s / m, j = s % m;
    vl = v, tl = 56 + j;  // vertex of a in b's frame, triangle of b
  } else {
    const int s2 = s - 3 * n * m, v = s2 / n, i = s2 % n;
    vl = 24 + v, tl = 48 + i;
  }
  if (!active) vl = 0, tl = 48;
  float p0[3], p1[3];
  SoupTri t;
  sh3(sh.f0, vl, p0), sh3(sh.f1, vl, p1), sh3(sh.t.n, tl, t.n);
  t.d = __shfl(sh.t.d, tl), t.eo = __shfl(sh.t.eo, tl);
  for (int q = 0; q < 3; ++q) sh3(sh.t.m + 3 * q, tl, t.m + 3 * q), t.o[q] = __shfl(sh.t.o[q], tl);
  if (!active) return false;
  return soup_point_keep(p0, p1, t, I.thr, I.margin);
}
__device__ bool leaf_edge_keep(const LeafShare& sh, const LeafInfo& I, int s, bool active) {
  const int ea = active ? s / (3 * I.m) : 0, eb = active ? s % (3 * I.m) : 0;
  const int a0 = 3 * (ea / 3) + ea % 3, a1 = 3 * (ea / 3) + (ea % 3 + 1) % 3;
  const int b0 = 24 + 3 * (eb / 3) + eb % 3, b1 = 24 + 3 * (eb / 3) + (eb % 3 + 1) % 3;
  float as0[3], as1[3], ae0[3], ae1[3], bs0[3], bs1[3], be0[3], be1[3];
  sh3(sh.w0, a0, as0), sh3(sh.w0, a1, as1), sh3(sh.w1, a0, ae0), sh3(sh.w1, a1, ae1);
  sh3(sh.w0, b0, bs0), sh3(sh.w0, b1, bs1), sh3(sh.w1, b0, be0), sh3(sh.w1, b1, be1);
  if (!active) return false;
  return !cull_edge_edge(as0, as1, ae0, ae1, bs0, bs1, be0, be1, I.thr, I.mag);
}
constexpr int CONN_MASK_WORDS = 16;  // 6 words of point sub-tests (<= 384), 9 of edge sub-tests (<= 576)

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_level_count(Dev D) {
  {  // soup: children per destination list
    const long long n = D.n_soup();
    long long rb, re;
    chunk_of(n, rb, re);
    long long as = 0, ac = 0;
    for (long long w = rb + threadIdx.x; w < re; w += blockDim.x) {
      int ns = 0, nc = 0;
      if (D.np_hit[w]) {
        NP np = D.soup[w];
        const int32_t* pi = D.pair_info + 4 * np.pair;
        const NodeRec na = D.nodes[pi[2] + np.na];
        const NodeRec nb = D.nodes[pi[3] + np.nb];
        if (na.inner || nb.inner) {
          // a child pair goes to the soup list when either node is inner:
          // all of them when the node that stays is inner, else the inner
          // children of the one that advances
          const bool adv_a = expand_side(na, nb);
          const NodeRec& par = adv_a ? na : nb;
          const bool stay_inner = adv_a ? nb.inner : na.inner;
          ns = stay_inner ? par.nchild : __popc((unsigned)D.kid_inner[(adv_a ? pi[2] : pi[3]) + (adv_a ? np.na : np.nb)]);
          nc = par.nchild - ns;
        }
      }
      D.cnt_s[w] = ns;
      D.cnt_c[w] = nc;
      as += ns, ac += nc;
    }
    const long long ts = block_sum(as), tc = block_sum(ac);
    if (threadIdx.x == 0) D.part_s[blockIdx.x] = (int)ts, D.part_c[blockIdx.x] = (int)tc;
  }
  {
    {  // connected list: sub-test counts
      const long long n = D.n_conn();
      long long rb, re;
      chunk_of(n, rb, re);
      long long ap = 0, ae = 0;  // lane 0 of each wavefront
      const int lane = lane_id();
      const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // uniform: scalar loads
      for (long long w = rb + wv; w < re; w += (blockDim.x >> 6)) {  // one wavefront per node pair
        const NP np = D.conn[w];
        LeafShare sh;
        LeafInfo L;
        if (D.fp32) {
          L = leaf_share(D, np, sh);  // reads the FP32 copies (FP32 contexts only)
        } else {  // the FP64 form keeps every sub-test: only the sizes
          const int32_t* pi = D.pair_info + 4 * np.pair;
          L.n = D.nodes[pi[2] + np.na].ntris, L.m = D.nodes[pi[3] + np.nb].ntris;
        }
        const int npt = 6 * L.n * L.m, nee = 9 * L.n * L.m;
        int kp = 0, ke = 0;
        unsigned long long* mk = D.conn_mask + (size_t)CONN_MASK_WORDS * w;
        for (int b0 = 0, wd = 0; b0 < npt; b0 += 64, ++wd) {
          const int sb = b0 + lane;
          const bool keep = D.fp32 ? leaf_point_keep(sh, L, sb, sb < npt) : sb < npt;
          const unsigned long long bits = __ballot(keep);
          if (lane == 0) mk[wd] = bits;
          kp += __popcll(bits);
        }
        for (int b0 = 0, wd = 6; b0 < nee; b0 += 64, ++wd) {
          const int sb = b0 + lane;
          const bool keep = D.fp32 ? leaf_edge_keep(sh, L, sb, sb < nee) : sb < nee;
          const unsigned long long bits = __ballot(keep);
          if (lane == 0) mk[wd] = bits;
          ke += __popcll(bits);
        }
        if (lane == 0) {
          D.cnt_pt[w] = kp;
          D.cnt_ee[w] = ke;
          ap += kp, ae += ke;
        }
      }
      const long long tp = block_sum(ap), te = block_sum(ae);
      if (threadIdx.x == 0) D.part_pt[blockIdx.x] = (int)tp, D.part_ee[blockIdx.x] = (int)te;
    }
    fold_pass(D, D.pair_rec, D.pass_min_s, D.pass_flag_s);  // the soup pass of this level is complete
  }
}

__global__ void __launch_bounds__(256) k_level_write(Dev D) {
  __shared__ int toff_s[256], toff_c[256];
  const int lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  {
    const long long n = D.n_soup();
    long long rb, re, bs, ts, bc, tc;
    chunk_of(n, rb, re);
    partial_prefix(D.part_s, bs, ts);
    partial_prefix(D.part_c, bc, tc);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      D.lvl[(D.level + 1) * LV_STRIDE + LV_SOUP] = (unsigned long long)ts;
      D.lvl[(D.level + 1) * LV_STRIDE + LV_CONN] = (unsigned long long)tc;
      if (ts > D.cap_np || tc > D.cap_np) atomicOr(&D.cnt[C_OVERFLOW], 1ull);
    }
    for (long long t0 = rb; t0 < re; t0 += blockDim.x) {
      const long long w = t0 + threadIdx.x;
      int tts, ttc;
      const int es = block_scan(w < re ? D.cnt_s[w] : 0, tts);
      const int ec = block_scan(w < re ? D.cnt_c[w] : 0, ttc);
      if (w < re && D.cnt_s[w] + D.cnt_c[w] > 0) {
        NP np = D.soup[w];
        const int32_t* pi = D.pair_info + 4 * np.pair;
        const NodeRec na = D.nodes[pi[2] + np.na];
        const NodeRec nb = D.nodes[pi[3] + np.nb];
        bool adv_a = expand_side(na, nb);
        const NodeRec& par = adv_a ? na : nb;
        const bool stay_inner = adv_a ? nb.inner : na.inner;
        const unsigned kid = D.kid_inner[(adv_a ? pi[2] : pi[3]) + (adv_a ? np.na : np.nb)];
        long long ps = bs + es, pc = bc + ec;
        for (int c = 0; c < par.nchild; ++c) {
          uint32_t ca = adv_a ? (uint32_t)(par.child_first + c) : np.na;
          uint32_t cb = adv_a ? np.nb : (uint32_t)(par.child_first + c);
          NP child{np.pair, ca, cb, 0u, np.key * 8ull + (unsigned long long)c};
          const bool to_soup = stay_inner || ((kid >> c) & 1u);
          long long slot = to_soup ? ps++ : pc++;
          if (slot < D.cap_np) (to_soup ? D.soup_next : D.conn_next)[slot] = child;
        }
      }
      bs += tts, bc += ttc;
    }
  }
  {
    const long long n = D.n_conn();
    long long rb, re, bp, tp, be, te;
    chunk_of(n, rb, re);
    partial_prefix(D.part_pt, bp, tp);
    partial_prefix(D.part_ee, be, te);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      D.lvl[D.level * LV_STRIDE + LV_PT] = (unsigned long long)tp;
      D.lvl[D.level * LV_STRIDE + LV_EE] = (unsigned long long)te;
      if (tp > D.cap_items || te > D.cap_items) atomicOr(&D.cnt[C_OVERFLOW], 1ull);
    }
    for (long long t0 = rb; t0 < re; t0 += blockDim.x) {
      const long long w = t0 + threadIdx.x;
      int ttp, tte;
      const int ep = block_scan(w < re ? D.cnt_pt[w] : 0, ttp);
      const int ee = block_scan(w < re ? D.cnt_ee[w] : 0, tte);
      toff_s[threadIdx.x] = ep;
      toff_c[threadIdx.x] = ee;
      __syncthreads();
      const int in_tile = (int)((re - t0) < (long long)blockDim.x ? (re - t0) : (long long)blockDim.x);
      for (int j = wid; j < in_tile; j += nw) {  // one wavefront per connected node pair
        const long long ww = t0 + j;
        const NP cnp = D.conn[ww];
        const int32_t* cpi = D.pair_info + 4 * cnp.pair;
        const int nn = D.nodes[cpi[2] + cnp.na].ntris, mm = D.nodes[cpi[3] + cnp.nb].ntris;
        const int npt = 6 * nn * mm, nee = 9 * nn * mm;
        const unsigned long long* mkw = D.conn_mask + (size_t)CONN_MASK_WORDS * ww;
        long long op = bp + toff_s[j], oe = be + toff_c[j];
        for (int b0 = 0, wd = 0; b0 < npt; b0 += 64, ++wd) {  // surviving sub-tests in sub-test order
          const int q = b0 + lane;
          const unsigned long long mk = mkw[wd];
          const bool keep = (mk >> lane) & 1ull;
          const long long slot = op + __popcll(mk & lanes_below());
          if (keep && slot < D.cap_items) D.pt_items[slot] = Item{(uint32_t)ww, (uint32_t)q};
          op += __popcll(mk);
        }
        for (int b0 = 0, wd = 6; b0 < nee; b0 += 64, ++wd) {
          const int q = b0 + lane;
          const unsigned long long mk = mkw[wd];
          const bool keep = (mk >> lane) & 1ull;
          const long long slot = oe + __popcll(mk & lanes_below());
          if (keep && slot < D.cap_items) D.ee_items[slot] = Item{(uint32_t)ww, (uint32_t)q};
          oe += __popcll(mk);
        }
      }
      bp += ttp, be += tte;
      __syncthreads();
    }
  }
}

// FP64-mode discrete levels: the children of every hit node pair (from its
// expansion descriptor) to the next list, one atomic per block and round.
__global__ void __launch_bounds__(256) k_disc_expand(Dev D) {
  __shared__ int s_tot[4];
  __shared__ unsigned long long s_base;
  const long long n = D.n_soup();
  const int lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < n; b0 += stride) {  // block-uniform
    const long long w = b0 + threadIdx.x;
    int desc = 0, nc = 0;
    unsigned cmask = 0u;
    NP np{};
    if (w < n && D.np_hit[w]) {
      desc = D.exp_desc[w];
      np = D.soup[w];
      const int32_t* pi = D.pair_info + 4 * np.pair;
      cmask = child_mask<false>(D, np, desc, pi[2], pi[3]);
      nc = __popc(cmask);
    }
    const int incl = wave_incl_scan(nc);
    if (lane == 63) s_tot[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int v = 0; v < nw; ++v) tot += s_tot[v];
      unsigned long long base = 0;
      if (tot) {
        base = atomicAdd(&D.lvl[(D.level + 1) * LV_STRIDE + LV_SOUP], (unsigned long long)tot);
        if ((long long)(base + tot) > D.cap_np) atomicOr(&D.cnt[C_OVERFLOW], 2ull);
      }
      s_base = base;
    }
    __syncthreads();
    long long slot = (long long)s_base + incl - nc;
    for (int v = 0; v < wid; ++v) slot += s_tot[v];
    __syncthreads();  // s_tot and s_base are rewritten by the next round
    if (nc > 0) write_children(D, slot, np, desc, 0u, cmask);
  }
}

// ------------------------------------------- algorithmic-byte accounting
// Instrumentation only (d2_set_touch_stats): which (body, node) records a
// level's traversal tests.  Bit node_base[body] + node of a bitmap stands for
// node `node` of body `body`; one bitmap per level and one per run (the union
// over the run's levels and phases), and one bit per body for each.
struct TouchArgs {
  const NP* soup;
  const NP* conn;
  const unsigned long long* lvl;
  const int* np_hit_prev;
  const int32_t* pair_info;
  const long long* node_base;
  unsigned *lv_nodes, *run_nodes, *lv_bodies, *run_bodies;
  long long cap_np;
  int level, discrete;
};
__device__ __forceinline__ void touch_bit(unsigned* bits, long long idx) {
  atomicOr(bits + (idx >> 5), 1u << (unsigned)(idx & 31));
}
// Marks the nodes of every node pair the level tests: a discrete node pair
// whose parent missed is dropped unread by k_tri_mask (same test here); every
// CCD soup and connected node pair is tested.
__global__ void __launch_bounds__(256) k_touch(TouchArgs T) {
  const unsigned long long* lv = T.lvl + (size_t)T.level * LV_STRIDE;
  const long long ns = (long long)lv[LV_SOUP] < T.cap_np ? (long long)lv[LV_SOUP] : T.cap_np;
  const long long nc = T.discrete ? 0 : ((long long)lv[LV_CONN] < T.cap_np ? (long long)lv[LV_CONN] : T.cap_np);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ns + nc; t += (long long)gridDim.x * blockDim.x) {
    const NP np = t < ns ? T.soup[t] : T.conn[t - ns];
    if (T.discrete && np.pad != 0u && T.np_hit_prev[np.pad - 1u] == 0) continue;
    const int a = T.pair_info[4 * np.pair], b = T.pair_info[4 * np.pair + 1];
    const long long ia = T.node_base[a] + np.na, ib = T.node_base[b] + np.nb;
    touch_bit(T.lv_nodes, ia), touch_bit(T.lv_nodes, ib);
    touch_bit(T.run_nodes, ia), touch_bit(T.run_nodes, ib);
    touch_bit(T.lv_bodies, a), touch_bit(T.lv_bodies, b);
    touch_bit(T.run_bodies, a), touch_bit(T.run_bodies, b);
  }
}
// out[0..3] += inner nodes, leaf nodes, triangles of those leaves, bodies
// marked in the bitmaps.
__global__ void __launch_bounds__(256) k_touch_count(const unsigned* nodes_bits, long long n_words, const unsigned* body_bits,
                                                     long long n_body_words, const long long* node_base, int n_bodies,
                                                     const int32_t* body_tree, const int32_t* tree_off,
                                                     const NodeRec* nodes, unsigned long long* out) {
  unsigned long long inner = 0, leaf = 0, tris = 0, bodies = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < n_words; w += stride) {
    unsigned m = nodes_bits[w];
    while (m) {
      const int k = __ffs(m) - 1;
      m &= m - 1u;
      const long long idx = 32 * w + k;
      int lo = 0, hi = n_bodies - 1;  // the body: the last node_base <= idx
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (node_base[mid] <= idx) lo = mid;
        else hi = mid - 1;
      }
      const NodeRec r = nodes[tree_off[body_tree[lo]] + (idx - node_base[lo])];
      if (r.inner) ++inner;
      else ++leaf, tris += (unsigned long long)r.ntris;
    }
  }
  for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < n_body_words; w += stride)
    bodies += (unsigned long long)__popc(body_bits[w]);
  for (int o = 32; o > 0; o >>= 1) {
    inner += __shfl_down(inner, o), leaf += __shfl_down(leaf, o);
    tris += __shfl_down(tris, o), bodies += __shfl_down(bodies, o);
  }
  if (lane_id() == 0) {
    if (inner) atomicAdd(out + 0, inner);
    if (leaf) atomicAdd(out + 1, leaf);
    if (tris) atomicAdd(out + 2, tris);
    if (bodies) atomicAdd(out + 3, bodies);
  }
}

// ---------------------------------------------------- bucket-level entry
// The plugin point's unit of work (strategy::ContactDetectionStrategy::
// findContactsBucketComparison, contact_detection_strategy.h:11, and its
// implementation comparison_dist.h:205-240): a list of node pairs of two
// bodies compared at the bodies' future states.  One wavefront per node pair,
// lane 8 i + j tests triangle i of a against triangle j of b in FP64
// (tri_tri_dist); a node pair's hits take one slot range (one atomic), and
// k_bucket_scan / k_bucket_place put them in the reference's order: node-pair
// order, then i, then j -- the lane order.
struct BucketArgs {
  const int32_t* np_list;  // [n][2] node indices within the two trees
  int n;
  const NodeRec* nodes;
  const double* buckets;
  int off_a, off_b;        // first node / bucket record of each tree
  const double* body_a;    // body records (future state at [14..27])
  const double* body_b;
  ContactRec* hits;        // hit records at provisional slots
  long long cap;
  unsigned long long* counter;  // slots taken
  long long* base;         // per node pair, its first slot
  int* count;              // per node pair, its hits
};
__global__ void __launch_bounds__(256) k_bucket_pairs(BucketArgs A) {
  __shared__ double sT[32];  // future poses of a and b (basic_utils.h:79-86)
  if (threadIdx.x == 0) {
    const double* sa = A.body_a + 14;
    const double* sb = A.body_b + 14;
    pose(quat{sa[0], sa[1], sa[2], sa[3]}, v3<double>{sa[4], sa[5], sa[6]}, sT);
    pose(quat{sb[0], sb[1], sb[2], sb[3]}, v3<double>{sb[4], sb[5], sb[6]}, sT + 16);
  }
  __syncthreads();
  const int lane = lane_id();
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long n_waves = ((long long)gridDim.x * blockDim.x) >> 6;
  const double geo_a = A.body_a[37], geo_b = A.body_b[37];
  const double search = (double)(float)(geo_a + geo_b);  // float search distance (comparison_dist.h:207)
  const int i = lane >> 3, j = lane & 7;
  for (long long k = wave; k < A.n; k += n_waves) {
    const int la = A.np_list[2 * k], lb = A.np_list[2 * k + 1];
    const NodeRec na = A.nodes[A.off_a + la];
    const NodeRec nb = A.nodes[A.off_b + lb];
    const double* ba = A.buckets + (size_t)BUCKET * (A.off_a + na.bucket);
    const double* bb = A.buckets + (size_t)BUCKET * (A.off_b + nb.bucket);
    bool hit = false;
    v3<double> P{0, 0, 0}, Q{0, 0, 0};
    if (i < na.ntris && j < nb.ntris) {
      const tri<double> TA = xform(sT, load_tri<double>(ba, i));
      const tri<double> TB = xform(sT + 16, load_tri<double>(bb, j));
      const double d = tri_tri_dist(TA, TB, P, Q);
      hit = d < (search + ba[72 + i]) + bb[72 + j];
    }
    const unsigned long long mask = __ballot(hit);
    const int cnt = __popcll(mask);
    unsigned long long b0 = 0;
    if (lane == 0 && cnt) b0 = atomicAdd(A.counter, (unsigned long long)cnt);
    b0 = readlane_u64(b0, 0);
    if (lane == 0) A.base[k] = (long long)b0, A.count[k] = cnt;
    if (hit) {
      const long long slot = (long long)b0 + __popcll(mask & lanes_below());
      if (slot < A.cap) {
        ContactRec c;
        c.key = 0ull;
        c.pair = 0, c.na = la, c.nb = lb, c.kind = 0, c.ia = i, c.ib = j, c.level = 0, c.pad = 0;
        c.r[0] = P.x, c.r[1] = P.y, c.r[2] = P.z, c.r[3] = Q.x, c.r[4] = Q.y, c.r[5] = Q.z;
        c.r[6] = geo_a + ba[72 + i], c.r[7] = geo_b + bb[72 + j];  // Contact(P, Q, geo + eps, ..., inner eps)
        c.r[8] = ba[80 + i], c.r[9] = bb[80 + j], c.r[10] = 0.0;
        A.hits[slot] = c;
      }
    }
  }
}

// Scratch of the allocation-free bucket entry (d2_find_contacts_bucket_noalloc):
// at most NOALLOC_NP node pairs and NOALLOC_HITS hits per call, laid out as
// node pairs | base | count | off | used | hits | out ints | out reals |
// counter.  Each device holds its own copy until the process ends; its
// address comes from hipGetSymbolAddress (the name is a host-side stand-in).
constexpr long long NOALLOC_NP = 1ll << 15;
constexpr long long NOALLOC_HITS = 1ll << 17;
constexpr size_t NOALLOC_BYTES = (size_t)NOALLOC_NP * (8 + 8 + 4 + 8 + 4) +
                                 (size_t)NOALLOC_HITS * (sizeof(ContactRec) + 8 * 4 + 11 * 8) + 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// The standalone filter's input as one pair: contact i keeps input position i
// as its emission key, so the per-pair finishing kernels see the caller's
// order (the unstable sort's result depends on it).
__global__ void __launch_bounds__(256) k_filter_load(int n, const double* reals, ContactRec* c, int* perm, int* pair_off,
                                                     int* n_big) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    ContactRec r;
    r.key = (unsigned long long)i;
    r.pair = 0, r.na = 0, r.nb = 0, r.kind = 0, r.ia = i, r.ib = 0, r.level = 0, r.pad = 0;
    for (int q = 0; q < 11; ++q) r.r[q] = reals[11 * (size_t)i + q];
    c[i] = r;
    perm[i] = i;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) pair_off[0] = 0, pair_off[1] = n, n_big[0] = 0, n_big[1] = 0;
}

// Exclusive offsets of the node pairs' hits (node-pair order) and the
// reference's cumulative counts (pair_used_out[k] = hits up to and including
// node pair k, comparison_dist.h:238); one block.
__global__ void __launch_bounds__(1024) k_bucket_scan(int n, const int* count, long long* off, int32_t* used) {
  __shared__ long long s[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = b + per < n ? b + per : n;
  long long sum = 0;
  for (int k = b; k < e; ++k) sum += count[k];
  s[t] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan of the chunk sums
    const long long v = t >= d ? s[t - d] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  long long run = t ? s[t - 1] : 0;
  for (int k = b; k < e; ++k) {
    off[k] = run;
    run += count[k];
    used[k] = (int32_t)run;
  }
}

// Each node pair's hits from their provisional slots to their final rows, in
// the flat contact layout (8 int32 + 11 doubles per hit).
__global__ void __launch_bounds__(256) k_bucket_place(int n, const long long* base, const int* count,
                                                      const long long* off, const ContactRec* hits, int32_t* out_i,
                                                      double* out_r) {
  const int lane = lane_id();
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long long n_waves = ((long long)gridDim.x * blockDim.x) >> 6;
  for (long long k = wave; k < n; k += n_waves) {
    if (lane >= count[k]) continue;
    const ContactRec& c = hits[base[k] + lane];
    const long long row = off[k] + lane;
    int32_t* oi = out_i + 8 * row;
    oi[0] = c.pair, oi[1] = c.na, oi[2] = c.nb, oi[3] = c.kind, oi[4] = c.ia, oi[5] = c.ib, oi[6] = c.level, oi[7] = 0;
    for (int q = 0; q < 11; ++q) out_r[11 * row + q] = c.r[q];
  }
}

// -------------------------------------------------------- post-processing
__device__ __forceinline__ long long n_contacts(const unsigned long long* cnt, long long cap) {
  long long n = (long long)cnt[C_CONTACTS];
  return n < cap ? n : cap;
}
// Pair lists up to this size are scanned inside k_scatter.
constexpr int SCAN_FUSE = 4096;
// Contacts to their pair's segment of perm.  With pair_cnt (n_pairs <=
// SCAN_FUSE) every block first scans the per-pair counts into LDS itself and
// block 0 stores the offsets and their total, so no scan launch precedes it;
// without, pair_off comes from a scan launch.
__global__ void __launch_bounds__(256) k_scatter(const unsigned long long* cnt, long long cap, const ContactRec* c,
                                                 int* pair_off, int* pair_fill, int* perm, const int* pair_cnt,
                                                 int n_pairs, unsigned long long* total_dst) {
  __shared__ int s_off[SCAN_FUSE + 1];
  const int* off = pair_off;
  if (pair_cnt) {
    const int per = (n_pairs + (int)blockDim.x - 1) / (int)blockDim.x;
    const int b = (int)threadIdx.x * per, e = b + per < n_pairs ? b + per : n_pairs;
    int acc = 0;
    for (int i = b; i < e; ++i) acc += pair_cnt[i];
    int total;
    int run = block_scan(acc, total);
    for (int i = b; i < e; ++i) {
      const int v = pair_cnt[i];
      s_off[i] = run;
      if (blockIdx.x == 0) pair_off[i] = run;
      run += v;
    }
    if (threadIdx.x == 0) {
      s_off[n_pairs] = total;
      if (blockIdx.x == 0) {
        pair_off[n_pairs] = total;
        if (total_dst) *total_dst = (unsigned long long)total;
      }
    }
    __syncthreads();
    off = s_off;
  }
  const long long n = n_contacts(cnt, cap);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    int p = c[t].pair;
    int k = atomicAdd(&pair_fill[p], 1);
    perm[off[p] + k] = (int)t;
  }
}

// The step rule's scale for one contact (separate_clusters.cpp:496-564):
// false when the contact sets none.  k: the pair's future-state kinematics.
struct StepKin {
  double mt;
  v3<double> va, wa, xa, vb, wb, xb;
};
__device__ __forceinline__ StepKin step_kin(const double* pr) {
  StepKin k;
  k.mt = pr[PR_MAXT];
  k.va = {pr[PR_VA], pr[PR_VA + 1], pr[PR_VA + 2]}, k.wa = {pr[PR_WA], pr[PR_WA + 1], pr[PR_WA + 2]};
  k.xa = {pr[PR_XA], pr[PR_XA + 1], pr[PR_XA + 2]};
  k.vb = {pr[PR_VB], pr[PR_VB + 1], pr[PR_VB + 2]}, k.wb = {pr[PR_WB], pr[PR_WB + 1], pr[PR_WB + 2]};
  k.xb = {pr[PR_XB], pr[PR_XB + 1], pr[PR_XB + 2]};
  return k;
}
__device__ __forceinline__ bool step_rule_scale(const ContactRec& ct, const StepKin& k, double& s) {
  const double mt = k.mt;
  v3<double> A{ct.r[0], ct.r[1], ct.r[2]}, B{ct.r[3], ct.r[4], ct.r[5]};
  v3<double> nrm = sub(A, B);
  double inter = ct.r[6] + ct.r[7], toc = ct.r[10];
  v3<double> rel = sub(add(k.va, cross(k.wa, sub(A, k.xa))), add(k.vb, cross(k.wb, sub(B, k.xb))));
  double len = vlen(nrm);
  if (!(dot(unit(nrm), unit(rel)) <= 0.0 || len < 1e-4)) return false;
  if (len < 1e-4) {
    if (toc == 0.0) return false;
    double st = toc - (vlen(rel) * mt) / inter;
    st = st > 0.0 ? st : 0.0;  // std::max(0.0, x): x when x > 0 (NaN keeps 0)
    s = lerp(st, toc, 0.75);
    return true;
  }
  if (!(toc < 1.0)) return false;
  double pv = -(dot(unit(nrm), rel));
  double depth = inter - len;
  double pd = (pv * toc) * mt;
  if (!(pd > 1e-6)) return false;
  if (depth < pd) {
    s = lerp(toc * (1.0 - depth / pd), toc, 0.5);
  } else {
    double fp = toc + ((1.0 - toc) * inter) / (pv * mt);
    s = lerp(toc, (1.0 < fp) ? 1.0 : fp, 0.8);  // std::min(fp, 1.0)
  }
  return true;
}

// The step rule of a step's CCD phase (separate_clusters.cpp:496-564): a
// pair's scale is the minimum of 1 and its contacts' scales, which no order
// of the contacts changes, and a step never reads its CCD contacts, so they
// stay where the emitting kernels put them (no scatter, no emission order):
// one thread per contact, an atomic minimum on the scale's bits (k_setup sets
// 1.0).  Scales are non-negative (toc >= 0, st >= 0, and the last two forms
// need pv > 0 and depth < pd), so their bit patterns order as their values;
// a NaN never passes s < 1, as std::min(min_scaling, NaN) keeps min_scaling.
__global__ void __launch_bounds__(256) k_step_rule(const unsigned long long* cnt, long long cap, const ContactRec* c,
                                                   const double* pair_rec, double* scale) {
  const long long n = n_contacts(cnt, cap);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const ContactRec& ct = c[i];
    double s;
    if (!step_rule_scale(ct, step_kin(pair_rec + (size_t)PR_STRIDE * ct.pair), s) || !(s < 1.0)) continue;
    atomicMin(reinterpret_cast<unsigned long long*>(scale + ct.pair), (unsigned long long)__double_as_longlong(s));
  }
}

// Contacts a pair may have for the LDS passes of the post-processing: the
// second pass (one 256-thread block per CU, 142 KB of LDS) takes pairs up to
// POST_CAP_LARGE; above it a pair takes one block with global scratch.
constexpr int POST_CAP_LARGE = 2048;
struct LdsKey {
  double eps, dist;
};
__device__ __forceinline__ bool lds_less(const LdsKey& x, const LdsKey& y) {
  if (x.eps == y.eps) return x.dist < y.dist;
  return x.eps < y.eps;
}
// The same permutation computed by the whole wavefront.  Each Hoare partition
// of [lo, hi) around the pivot at lo-1 is derived from ranks: with A_k the k-th
// position from the left holding an element not less than the pivot and B_k
// the k-th from the right holding one not greater (the pivot's own slot ends
// that list), the scalar loop swaps (A_k, B_k) for every k < K,
// K = #{k : A_k < B_k}, and returns min(A_K, B_{K-1}).  The closing insertion
// sort never moves an element across a leaf segment's border, so it equals a
// stable sort of every leaf segment; heapsort (depth exhausted) stays on lane 0.
__device__ int wave_partition(int lo, int hi, int* ord, const int* rk, int* Apos, int* Bpos) {
  const int lane = lane_id();
  const int P = rk[ord[lo - 1]];
  int nA = 0, nB = 0;
  for (int base = lo; base < hi; base += 64) {
    const int i = base + lane;
    const bool ge = i < hi && !(rk[ord[i]] < P);
    const unsigned long long m = __ballot(ge);
    if (ge) Apos[nA + __popcll(m & lanes_below())] = i;
    nA += __popcll(m);
  }
  for (int top = hi - 1; top >= lo - 1; top -= 64) {
    const int i = top - lane;
    const bool le = i >= lo - 1 && !(P < rk[ord[i]]);
    const unsigned long long m = __ballot(le);
    if (le) Bpos[nB + __popcll(m & lanes_below())] = i;
    nB += __popcll(m);
  }
  wave_sync_lds();
  const int lim = nA < nB ? nA : nB;
  int K = 0;
  for (int base = 0; base < lim; base += 64) {  // A_k < B_k holds for a prefix of k
    const int q = base + lane;
    const unsigned long long m = __ballot(q < lim && Apos[q] < Bpos[q]);
    K += __popcll(m);
    if (m != ~0ull) break;
  }
  const int ak = K < nA ? Apos[K] : hi;
  const int bk1 = K > 0 ? Bpos[K - 1] : hi;
  for (int q = lane; q < K; q += 64) {  // the swapped positions are pairwise distinct
    const int x = ord[Apos[q]], y = ord[Bpos[q]];
    ord[Apos[q]] = y;
    ord[Bpos[q]] = x;
  }
  wave_sync_lds();
  return ak < bk1 ? ak : bk1;
}



// ---- one block (4 wavefronts) per pair: pairs with at most CAP contacts
// (BLK_CAP for every pair, then POST_CAP_LARGE for the larger ones)
constexpr int BLK_CAP = 512;
// queued segments per round (> 16 elements each: at most CAP / 17), 3 ints each
constexpr int blk_segs(int cap) { return 3 * (cap / 17 + 1); }

// ord[0, n) := the order std::sort gives, with the
// partitions of one round's segments spread over the block's wavefronts:
// partitions of disjoint segments commute, and every segment keeps its own
// depth budget, so the order equals the sequential introsort's.  Apos/Bpos
// are indexed by position (a segment's lists stay inside its range).
__device__ void block_std_sort(int n, int* ord, const int* rk, int* Apos, int* Bpos, int* mark, int* segs,
                               int* s_cnt, int seg_stride) {
  const int tid = threadIdx.x, w = tid >> 6, lane = lane_id();
  for (int i = tid; i < n; i += blockDim.x) mark[i] = 0;
  if (tid == 0) {  // thread 0 also zeroed mark[0] above
    s_cnt[0] = 0, s_cnt[1] = 0;
    if (n > 16) segs[0] = 0, segs[1] = n, segs[2] = 2 * (63 - __clzll((long long)n)), s_cnt[0] = 1;
    else if (n > 0) mark[0] = 1;
  }
  __syncthreads();
  if (n <= 1) return;
  int cur = 0;
  for (;;) {
    const int m = s_cnt[cur];
    if (m == 0) break;
    const int* L = segs + cur * seg_stride;
    int* N = segs + (cur ^ 1) * seg_stride;
    for (int q = w; q < m; q += (int)(blockDim.x >> 6)) {
      const int f = L[3 * q], l = L[3 * q + 1];
      int d = L[3 * q + 2];
      if (d == 0) {  // depth exhausted: heapsort (libstdc++ __partial_sort) on one lane
        if (lane == 0) {
          StdSort<RankLess> srt{ord, RankLess{rk}};
          srt.heap_sort(f, l);
          mark[f] = 2;
        }
        wave_sync_lds();
        continue;
      }
      --d;
      {  // median of three to the front
        const int x = f + 1, y = f + (l - f) / 2, z = l - 1;
        const int kx = rk[ord[x]], ky = rk[ord[y]], kz = rk[ord[z]];
        int sel;
        if (kx < ky) sel = ky < kz ? y : (kx < kz ? z : x);
        else sel = kx < kz ? x : (ky < kz ? z : y);
        wave_sync_lds();
        if (lane == 0) {
          const int t = ord[f];
          ord[f] = ord[sel];
          ord[sel] = t;
        }
        wave_sync_lds();
      }
      const int cut = wave_partition(f + 1, l, ord, rk, Apos + f, Bpos + f);
      if (lane == 0) {
        // left part [f, cut) and right part [cut, l), both with depth d
        const int cf[2] = {f, cut}, cl[2] = {cut, l};
        for (int k = 0; k < 2; ++k) {
          if (cl[k] - cf[k] > 16) {
            const int slot = atomicAdd(&s_cnt[cur ^ 1], 1);
            N[3 * slot] = cf[k], N[3 * slot + 1] = cl[k], N[3 * slot + 2] = d;
          } else if (cl[k] > cf[k]) {
            mark[cf[k]] = 1;
          }
        }
      }
      wave_sync_lds();
    }
    __syncthreads();
    if (tid == 0) s_cnt[cur] = 0;
    cur ^= 1;
    __syncthreads();
  }
  // leaf segments tile [0, n): stable-sort each (<= 16 elements)
  for (int j = tid; j < n; j += blockDim.x) {
    int f = j;
    while (mark[f] == 0) --f;
    int dst = j;
    if (mark[f] == 1) {
      int l = j + 1;
      while (l < n && mark[l] == 0) ++l;
      const int kj = rk[ord[j]];
      int r = 0;
      for (int i = f; i < l; ++i) {
        const int ki = rk[ord[i]];
        r += ki < kj || (i < j && !(kj < ki));
      }
      dst = f + r;
    }
    Bpos[dst] = ord[j];
  }
  __syncthreads();
  for (int j = tid; j < n; j += blockDim.x) ord[j] = Bpos[j];
  __syncthreads();
}

// Post-processing of one pair per 256-thread block: orders by rank counting instead of sorting
// networks (emission keys are unique; (eps, dist) ranks equal exactly for tied
// keys), the tie replay over the block's wavefronts, and the greedy pass of
// filterContacts 64 candidates at a time with the distance tests spread over
// the wavefronts and the decisions taken in order by the first one.
// A pair above every LDS capacity without the filter (mode 0 or 2), one block:
// the emission order by rank counting (the keys are unique; each thread ranks
// up to 8 x 256 contacts at a time against key tiles of CAP staged in LDS),
// written to perm from a copy of the pair's contact indices in keep (unused
// by these modes); in mode 2 the step rule's minimum over the block.
template <int CAP>
__device__ void pair_finish_huge(int p, const ContactRec* c, const int* pair_off, int* perm, int mode, int* keep,
                                 const double* pair_rec, double* scale, unsigned long long* skey, double* s_min) {
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  const int b = pair_off[p], n = pair_off[p + 1] - b;
  for (int i = tid; i < n; i += blockDim.x) keep[b + i] = perm[b + i];
  __syncthreads();
  constexpr int K = 8;
  for (int q0 = 0; q0 < n; q0 += K * (int)blockDim.x) {
    unsigned long long mk[K];
    int rk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = q0 + k * (int)blockDim.x + tid;
      mk[k] = i < n ? c[keep[b + i]].key : ~0ull;
      rk[k] = 0;
    }
    for (int t0 = 0; t0 < n; t0 += CAP) {
      const int m = n - t0 < CAP ? n - t0 : CAP;
      for (int j = tid; j < m; j += blockDim.x) skey[j] = c[keep[b + t0 + j]].key;
      __syncthreads();
      for (int j = 0; j < m; ++j) {
        const unsigned long long kj = skey[j];
#pragma unroll
        for (int k = 0; k < K; ++k) rk[k] += kj < mk[k] ? 1 : 0;
      }
      __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = q0 + k * (int)blockDim.x + tid;
      if (i < n) perm[b + rk[k]] = keep[b + i];
    }
  }
  if (mode == 2) {  // step rule, separate_clusters.cpp:496-564: the minimum over the pair's contacts
    const StepKin kin = step_kin(pair_rec + (size_t)PR_STRIDE * p);
    double ms = 1.0;
    for (int i = tid; i < n; i += blockDim.x) {
      double sc;
      if (step_rule_scale(c[keep[b + i]], kin, sc) && sc < ms) ms = sc;  // NaN never wins
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(ms, off);
      if (o < ms) ms = o;
    }
    if (lane == 0) s_min[w] = ms;
    __syncthreads();
    if (tid == 0) {
      double m = s_min[0];
      for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = (s_min[k] < m) ? s_min[k] : m;
      scale[p] = m;
    }
  }
}

// Global scratch of the filter above every LDS capacity: six int arrays of
// the run's contact capacity, a pair using positions [b, b + n) of each
// (its contacts' range; the sort's segment queue, at most 6 (n / 17 + 1)
// ints, fits there for n > 16).
struct HugeScratch {
  int *rk, *ord, *apos, *bpos, *mark, *segs;
};

// filterContacts(.., margin) of one pair above every LDS capacity, one block:
// the emission order by rank counting (pair_finish_huge), the (eps, dist)
// ranks by rank counting over LDS key tiles, the std::sort replay
// (block_std_sort) on global scratch when keys tie, then the greedy pass 64
// candidates at a time as in the LDS path, with the kept list in scratch.
// Returns the kept count (every thread).
template <int CAP>
__device__ int pair_filter_huge(int p, const ContactRec* c, const int* pair_off, int* perm, int* keep, int* kept_cnt,
                                double margin, const HugeScratch& H, unsigned long long* skey, LdsKey* sk,
                                int* s_misc, unsigned long long (*conf_w)[64], int (*near_w)[64]) {
  const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
  const int b = pair_off[p], n = pair_off[p + 1] - b;
  pair_finish_huge<CAP>(p, c, pair_off, perm, 0, keep, nullptr, nullptr, skey, nullptr);  // perm: emission order
  if (tid == 0) s_misc[0] = 0;
  __syncthreads();
  int* rk = H.rk + b;
  int* ord = H.ord + b;
  int* eo = H.apos + b;  // the emission order, kept while perm is rewritten
  int* kl = H.mark + b;  // the kept list (emission positions), after the sort
  constexpr int K = 8;
  bool tie = false;
  for (int q0 = 0; q0 < n; q0 += K * (int)blockDim.x) {  // ranks of the (eps, dist) keys; equal keys share one
    LdsKey mk[K];
    int r[K], eq[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = q0 + k * (int)blockDim.x + tid;
      const ContactRec& ct = c[perm[b + (i < n ? i : 0)]];
      mk[k].eps = ct.r[6] + ct.r[7];
      mk[k].dist = vlen(sub(v3<double>{ct.r[0], ct.r[1], ct.r[2]}, v3<double>{ct.r[3], ct.r[4], ct.r[5]}));
      r[k] = 0, eq[k] = 0;
    }
    for (int t0 = 0; t0 < n; t0 += CAP) {
      const int m = n - t0 < CAP ? n - t0 : CAP;
      for (int j = tid; j < m; j += blockDim.x) {
        const ContactRec& ct = c[perm[b + t0 + j]];
        sk[j].eps = ct.r[6] + ct.r[7];
        sk[j].dist = vlen(sub(v3<double>{ct.r[0], ct.r[1], ct.r[2]}, v3<double>{ct.r[3], ct.r[4], ct.r[5]}));
      }
      __syncthreads();
      for (int j = 0; j < m; ++j) {
        const LdsKey y = sk[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (lds_less(y, mk[k])) ++r[k];
          else if (!lds_less(mk[k], y)) ++eq[k];
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = q0 + k * (int)blockDim.x + tid;
      if (i < n) rk[i] = r[k], tie = tie || eq[k] > 1;
    }
  }
  if (tie) s_misc[0] = 1;
  __syncthreads();
  if (s_misc[0] == 0) {
    for (int i = tid; i < n; i += blockDim.x) ord[rk[i]] = i;
  } else {
    for (int i = tid; i < n; i += blockDim.x) ord[i] = i;
    __syncthreads();
    block_std_sort(n, ord, rk, H.apos + b, H.bpos + b, H.mark + b, H.segs + b, s_misc + 2, blk_segs(n));
  }
  __syncthreads();
  for (int i = tid; i < n; i += blockDim.x) eo[i] = perm[b + i];
  __syncthreads();
  // greedy pass in sorted order
  int nk = 0;
  for (int t0 = 0; t0 < n; t0 += 64) {
    const int m = n - t0 < 64 ? n - t0 : 64;
    const int t = t0 + lane;
    const bool valid = t < n;
    const ContactRec& ct = c[eo[ord[valid ? t : t0]]];
    const v3<double> mid = dvs(add(v3<double>{ct.r[0], ct.r[1], ct.r[2]}, v3<double>{ct.r[3], ct.r[4], ct.r[5]}), 2.0);
    const double mx = mid.x, my = mid.y, mz = mid.z;
    const double search = (ct.r[6] + ct.r[7]) + margin;
    bool near = false;
    for (int q0 = 0; q0 < nk; q0 += 64) {  // contacts kept before this group
      const ContactRec& hc = c[eo[kl[q0 + lane < nk ? q0 + lane : q0]]];
      const v3<double> hm =
          dvs(add(v3<double>{hc.r[0], hc.r[1], hc.r[2]}, v3<double>{hc.r[3], hc.r[4], hc.r[5]}), 2.0);
      const int mm = nk - q0 < 64 ? nk - q0 : 64;
      for (int u = w; u < mm; u += 4) {
        const v3<double> dd{mx - readlane_d(hm.x, u), my - readlane_d(hm.y, u), mz - readlane_d(hm.z, u)};
        near = near || near_mid(dd, search);
      }
    }
    unsigned long long conf = 0;  // bit u: candidate u (< this one) of the group is near
    for (int u = w; u < m; u += 4) {
      const v3<double> dd{mx - readlane_d(mx, u), my - readlane_d(my, u), mz - readlane_d(mz, u)};
      if (u < lane && near_mid(dd, search)) conf |= 1ull << u;
    }
    near_w[w][lane] = near ? 1 : 0;
    conf_w[w][lane] = conf;
    __syncthreads();
    if (w == 0) {
      const bool nr = (near_w[0][lane] | near_w[1][lane] | near_w[2][lane] | near_w[3][lane]) != 0;
      const unsigned long long cf = conf_w[0][lane] | conf_w[1][lane] | conf_w[2][lane] | conf_w[3][lane];
      const unsigned long long blocked = __ballot(nr);
      unsigned long long kept = 0;
      for (int i = 0; i < m; ++i) {
        const unsigned long long ci = readlane_u64(cf, i);
        if (!((blocked >> i) & 1ull) && !(ci & kept)) kept |= 1ull << i;
      }
      const bool fresh = (kept >> lane) & 1ull;
      if (valid) keep[b + t] = fresh ? 1 : 0;
      if (fresh) kl[nk + __popcll(kept & lanes_below())] = ord[t];
      if (lane == 0) s_misc[1] = nk + __popcll(kept);
    }
    __syncthreads();
    nk = s_misc[1];
  }
  for (int t = tid; t < n; t += blockDim.x) perm[b + t] = eo[ord[t]];
  if (tid == 0) kept_cnt[p] = nk;
  __syncthreads();
  return nk;
}

// CAP = BLK_CAP: every pair (list null); pairs above CAP go to `over` (up to
// POST_CAP_LARGE contacts) or `huge`.  CAP = POST_CAP_LARGE: the pairs of
// `list`, then the pairs of `tail` one block each (above every LDS capacity).
template <int CAP>
__global__ void __launch_bounds__(256) k_pair_finish_block(int n_pairs, const int* list, const int* n_list,
                                                           const ContactRec* c, const int* pair_off, int* perm,
                                                           int mode, int* keep, int* kept_cnt,
                                                           const double* pair_rec, double* scale, int* over,
                                                           int* n_over, int* huge, int* n_huge, const int* tail,
                                                           const int* n_tail, unsigned long long* out_total,
                                                           double margin, HugeScratch H) {
  constexpr int SEGS = blk_segs(CAP);
  __shared__ unsigned long long skey[CAP];  // emission keys, then the partition lists of the tie replay
  __shared__ int sidx[CAP];
  __shared__ int eo[CAP];  // contact at each emission position
  __shared__ LdsKey sk[CAP];
  __shared__ double smid[3 * CAP];
  __shared__ int rk[CAP];
  __shared__ int ord[CAP];
  __shared__ int mark[CAP];  // leaf marks of the tie replay, then the kept list
  __shared__ int segs[2 * SEGS];
  __shared__ unsigned long long conf_w[4][64];
  __shared__ int near_w[4][64];
  __shared__ double s_min[4];
  __shared__ int s_misc[4];
  const int tid = threadIdx.x, w = tid >> 6, lane = lane_id();
  const int n_items = list ? *n_list : n_pairs;
  unsigned long long kept_sum = 0;  // (thread 0) the filter's kept contacts of this block's pairs
  for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
    const int p = list ? list[it] : it;
    const int b = pair_off[p], e = pair_off[p + 1], n = e - b;
    if (n > CAP) {  // to the second launch: its LDS pass, or (above its capacity) one lane per pair
      if (tid == 0) {
        if (n > POST_CAP_LARGE) huge[atomicAdd(n_huge, 1)] = p;
        else over[atomicAdd(n_over, 1)] = p;
      }
      continue;
    }
    if (n == 0) {  // what the general path writes for a pair without contacts, without its barriers
      if (tid == 0) {
        if (mode == 2) scale[p] = 1.0;
        if (mode == 1) kept_cnt[p] = 0;
      }
      continue;
    }
    for (int i = tid; i < n; i += 256) {
      const int idx = perm[b + i];
      sidx[i] = idx;
      skey[i] = c[idx].key;
    }
    __syncthreads();
    for (int i = tid; i < n; i += 256) {  // emission position = rank of the (unique) key
      const unsigned long long k = skey[i];
      int r = 0, j = 0;
      for (; j + 4 <= n; j += 4) {  // four independent LDS reads in flight
        const unsigned long long k0 = skey[j], k1 = skey[j + 1], k2 = skey[j + 2], k3 = skey[j + 3];
        r += (k0 < k ? 1 : 0) + (k1 < k ? 1 : 0) + (k2 < k ? 1 : 0) + (k3 < k ? 1 : 0);
      }
      for (; j < n; ++j) r += skey[j] < k ? 1 : 0;
      eo[r] = sidx[i];
    }
    __syncthreads();
    if (mode == 2) {  // step rule, separate_clusters.cpp:496-564
      const double* pr = pair_rec + (size_t)PR_STRIDE * p;
      double mt = pr[PR_MAXT], ms = 1.0;
      v3<double> va{pr[PR_VA], pr[PR_VA + 1], pr[PR_VA + 2]}, wa{pr[PR_WA], pr[PR_WA + 1], pr[PR_WA + 2]};
      v3<double> xa{pr[PR_XA], pr[PR_XA + 1], pr[PR_XA + 2]};
      v3<double> vb{pr[PR_VB], pr[PR_VB + 1], pr[PR_VB + 2]}, wb{pr[PR_WB], pr[PR_WB + 1], pr[PR_WB + 2]};
      v3<double> xb{pr[PR_XB], pr[PR_XB + 1], pr[PR_XB + 2]};
      for (int i = tid; i < n; i += 256) {
        const ContactRec& ct = c[eo[i]];
        v3<double> A{ct.r[0], ct.r[1], ct.r[2]}, B{ct.r[3], ct.r[4], ct.r[5]};
        v3<double> nrm = sub(A, B);
        double inter = ct.r[6] + ct.r[7], toc = ct.r[10];
        v3<double> rel = sub(add(va, cross(wa, sub(A, xa))), add(vb, cross(wb, sub(B, xb))));
        double len = vlen(nrm);
        if (!(dot(unit(nrm), unit(rel)) <= 0.0 || len < 1e-4)) continue;
        double sc;
        if (len < 1e-4) {
          if (toc == 0.0) continue;
          double st = toc - (vlen(rel) * mt) / inter;
          st = st > 0.0 ? st : 0.0;
          sc = lerp(st, toc, 0.75);
        } else {
          if (!(toc < 1.0)) continue;
          double pv = -(dot(unit(nrm), rel));
          double depth = inter - len;
          double pd = (pv * toc) * mt;
          if (!(pd > 1e-6)) continue;
          if (depth < pd) {
            sc = lerp(toc * (1.0 - depth / pd), toc, 0.5);
          } else {
            double fp = toc + ((1.0 - toc) * inter) / (pv * mt);
            sc = lerp(toc, (1.0 < fp) ? 1.0 : fp, 0.8);
          }
        }
        if (sc < ms) ms = sc;
      }
      for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(ms, off);
        if (o < ms) ms = o;
      }
      if (lane == 0) s_min[w] = ms;
      __syncthreads();
      if (tid == 0) {
        double m = s_min[0];
        for (int k = 1; k < 4; ++k) m = (s_min[k] < m) ? s_min[k] : m;
        scale[p] = m;
      }
      for (int i = tid; i < n; i += 256) perm[b + i] = eo[i];
    } else if (mode == 1) {  // filterContacts(.., margin)
      for (int i = tid; i < n; i += 256) {
        const ContactRec& ct = c[eo[i]];
        const v3<double> A{ct.r[0], ct.r[1], ct.r[2]}, B{ct.r[3], ct.r[4], ct.r[5]};
        sk[i].eps = ct.r[6] + ct.r[7];
        sk[i].dist = vlen(sub(A, B));
        const v3<double> m = dvs(add(A, B), 2.0);
        smid[3 * i] = m.x, smid[3 * i + 1] = m.y, smid[3 * i + 2] = m.z;
      }
      if (tid == 0) s_misc[0] = 0;
      __syncthreads();
      bool tie = false;
      for (int i = tid; i < n; i += 256) {  // rank of the (eps, dist) key; equal keys share it
        const LdsKey x = sk[i];
        int r = 0, eq = 0, j = 0;
        for (; j + 4 <= n; j += 4) {
          const LdsKey y0 = sk[j], y1 = sk[j + 1], y2 = sk[j + 2], y3 = sk[j + 3];
          r += (lds_less(y0, x) ? 1 : 0) + (lds_less(y1, x) ? 1 : 0) + (lds_less(y2, x) ? 1 : 0) +
               (lds_less(y3, x) ? 1 : 0);
          eq += (!lds_less(y0, x) && !lds_less(x, y0) ? 1 : 0) + (!lds_less(y1, x) && !lds_less(x, y1) ? 1 : 0) +
                (!lds_less(y2, x) && !lds_less(x, y2) ? 1 : 0) + (!lds_less(y3, x) && !lds_less(x, y3) ? 1 : 0);
        }
        for (; j < n; ++j) {
          const LdsKey y = sk[j];
          if (lds_less(y, x)) ++r;
          else if (!lds_less(x, y)) ++eq;
        }
        rk[i] = r;
        tie = tie || eq > 1;
      }
      if (tie) s_misc[0] = 1;
      __syncthreads();
      if (s_misc[0] == 0) {
        for (int i = tid; i < n; i += 256) ord[rk[i]] = i;
        __syncthreads();
      } else {
        for (int i = tid; i < n; i += 256) ord[i] = i;
        __syncthreads();
        int* lists = reinterpret_cast<int*>(skey);
        block_std_sort(n, ord, rk, lists, lists + CAP, mark, segs, s_misc + 2, SEGS);
      }
      // greedy pass in sorted order
      int nk = 0;
      for (int t0 = 0; t0 < n; t0 += 64) {
        const int m = n - t0 < 64 ? n - t0 : 64;
        const int t = t0 + lane;
        const bool valid = t < n;
        const int pi = valid ? ord[t] : 0;
        const double mx = smid[3 * pi], my = smid[3 * pi + 1], mz = smid[3 * pi + 2];
        const double search = sk[pi].eps + margin;
        bool near = false;
        for (int q0 = 0; q0 < nk; q0 += 64) {  // contacts kept before the block
          const int h = q0 + lane < nk ? mark[q0 + lane] : 0;
          const double hx = smid[3 * h], hy = smid[3 * h + 1], hz = smid[3 * h + 2];
          const int mm = nk - q0 < 64 ? nk - q0 : 64;
          for (int u = w; u < mm; u += 4) {
            const v3<double> dd{mx - readlane_d(hx, u), my - readlane_d(hy, u), mz - readlane_d(hz, u)};
            near = near || near_mid(dd, search);
          }
        }
        unsigned long long conf = 0;  // bit u: candidate u (< this one) of the block is near
        for (int u = w; u < m; u += 4) {
          const v3<double> dd{mx - readlane_d(mx, u), my - readlane_d(my, u), mz - readlane_d(mz, u)};
          if (u < lane && near_mid(dd, search)) conf |= 1ull << u;
        }
        near_w[w][lane] = near ? 1 : 0;
        conf_w[w][lane] = conf;
        __syncthreads();
        if (w == 0) {
          const bool nr = (near_w[0][lane] | near_w[1][lane] | near_w[2][lane] | near_w[3][lane]) != 0;
          const unsigned long long cf = conf_w[0][lane] | conf_w[1][lane] | conf_w[2][lane] | conf_w[3][lane];
          const unsigned long long blocked = __ballot(nr);
          unsigned long long kept = 0;
          for (int i = 0; i < m; ++i) {
            const unsigned long long ci = readlane_u64(cf, i);
            if (!((blocked >> i) & 1ull) && !(ci & kept)) kept |= 1ull << i;
          }
          const bool fresh = (kept >> lane) & 1ull;
          if (valid) keep[b + t] = fresh ? 1 : 0;
          if (fresh) mark[nk + __popcll(kept & lanes_below())] = pi;
          if (lane == 0) s_misc[1] = nk + __popcll(kept);
        }
        __syncthreads();
        nk = s_misc[1];
      }
      for (int t = tid; t < n; t += 256) perm[b + t] = eo[ord[t]];
      if (tid == 0) kept_cnt[p] = nk, kept_sum += (unsigned long long)nk;
    } else {
      for (int i = tid; i < n; i += 256) perm[b + i] = eo[i];
    }
    __syncthreads();
  }
  if constexpr (CAP == POST_CAP_LARGE) {  // (kept out of the first launch's code: registers, scratch)
    if (tail) {  // pairs above every LDS capacity (listed by the first launch), one block each
      if (mode == 1) {
        for (int t = blockIdx.x; t < *n_tail; t += gridDim.x) {
          const int nk = pair_filter_huge<CAP>(tail[t], c, pair_off, perm, keep, kept_cnt, margin, H, skey, sk,
                                               s_misc, conf_w, near_w);
          if (tid == 0) kept_sum += (unsigned long long)nk;
        }
      } else {
        for (int t = blockIdx.x; t < *n_tail; t += gridDim.x) {
          pair_finish_huge<CAP>(tail[t], c, pair_off, perm, mode, keep, pair_rec, scale, skey, s_min);
          __syncthreads();
        }
      }
    }
  }
  if (tid == 0 && out_total && kept_sum) atomicAdd(out_total, kept_sum);  // one atomic per block
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

// Exclusive scan of in[0, n) into out[0, n], out[n] = total (also stored to
// *total_dst when given), in two chunked kernels of the same grid.
__global__ void __launch_bounds__(256) k_scan_part(const int* in, long long n, int* part) {
  long long rb, re;
  chunk_of(n, rb, re);
  long long acc = 0;
  for (long long i = rb + threadIdx.x; i < re; i += blockDim.x) acc += in[i];
  const long long t = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = (int)t;
}
__global__ void __launch_bounds__(256) k_scan_apply(const int* in, int* out, long long n, const int* part,
                                                    unsigned long long* total_dst) {
  long long rb, re, before, total;
  chunk_of(n, rb, re);
  partial_prefix(part, before, total);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out[n] = (int)total;
    if (total_dst) *total_dst = (unsigned long long)total;
  }
  for (long long t0 = rb; t0 < re; t0 += blockDim.x) {
    const long long i = t0 + threadIdx.x;
    int tt;
    const int ex = block_scan(i < re ? in[i] : 0, tt);
    if (i < re) out[i] = (int)(before + ex);
    before += tt;
  }
}

// The same scan for a short array (n <= SCAN_SMALL) in one block of 1024
// threads: thread t sums its run of ceil(n / 1024) elements, the block scans
// the sums, and each thread writes its run.
constexpr long long SCAN_SMALL = 16384;
__global__ void __launch_bounds__(1024) k_scan_small(const int* in, int* out, long long n,
                                                     unsigned long long* total_dst) {
  const long long per = (n + blockDim.x - 1) / blockDim.x;
  const long long b = (long long)threadIdx.x * per;
  const long long e = b + per < n ? b + per : n;
  int acc = 0;
  for (long long i = b; i < e; ++i) acc += in[i];
  int total;
  int run = block_scan(acc, total);
  for (long long i = b; i < e; ++i) {
    const int v = in[i];
    out[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) {
    out[n] = total;
    if (total_dst) *total_dst = (unsigned long long)total;
  }
}

// ------------------------------------------------------ cluster step sizes
// The cluster step (separate_clusters.cpp:555-598,
// time_step_selection_dynamic_continuous.cpp:17-21).  The reference runs two
// loops over the cluster's pairs in order: the first lowers both dynamic
// bodies' steps to use = min(max_time * scale, step[a], step[b]) for a pair
// with CCD contacts; the second starts from step[a] of the first pair that is
// not static-static (:589-590) and takes min(., step[id]) over the others (id
// = a unless a is static).  Since max_time * scale <= max_time <= both last
// steps, every pair's use reaches the step its own pair reads in the second
// loop, so the result is the minimum over pairs of max_time * scale (with
// contacts) and last[id] -- order does not matter -- except for the first
// pair when its body a is static: the second loop reads a's unchanged last
// step there, and that pair's use only counts when its body b is read or
// lowered again later (a later pair whose id is b, or a later pair with
// contacts touching b: every other pair comes later).  read_later marks
// those bodies.  Steps are non-negative, so they order like their bit
// patterns.
__global__ void k_cluster_init(ClusterAcc CA) {
  cluster_init_range(CA, (long long)blockIdx.x * blockDim.x + threadIdx.x, (long long)gridDim.x * blockDim.x);
}
__global__ void k_cluster_first(int n_pairs, ClusterAcc CA) {
  cluster_first_range(CA, n_pairs, (long long)blockIdx.x * blockDim.x + threadIdx.x, (long long)gridDim.x * blockDim.x);
}
__global__ void k_cluster_min(int n_pairs, const int32_t* pairs, const int32_t* cluster_of_pair, const double* bodies,
                              const int* pair_cnt, const double* scale, const int* first_ns, int* read_later,
                              unsigned long long* step_bits) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    const int k = cluster_of_pair[p], a = pairs[2 * p], b = pairs[2 * p + 1];
    const double* ra = bodies + 40 * (size_t)a;
    const double* rb = bodies + 40 * (size_t)b;
    const bool sa = ra[38] != 0.0, sb = rb[38] != 0.0;
    if (sa && sb) continue;
    const bool first = p == first_ns[k];
    const int id = !sa ? a : b;
    if (!first) {  // the second loop reads step[id]; a pair with contacts lowers a's and b's steps
      read_later[id] = 1;
      if (pair_cnt[p] > 0) read_later[a] = 1, read_later[b] = 1;
    }
    const double mt = (rb[39] < ra[39]) ? rb[39] : ra[39];
    double v = 1e300;
    if (pair_cnt[p] > 0 && !(first && sa)) {  // a static first pair's own use: see cluster_step_of
      const double s = mt * scale[p];
      v = (s < v) ? s : v;
    }
    const double last = first ? ra[39] : bodies[40 * (size_t)id + 39];
    v = (last < v) ? last : v;
    atomicMin(&step_bits[k], (unsigned long long)__double_as_longlong(v));
  }
}
struct ClusterFinal {
  const int32_t* pairs;
  const double* bodies;
  const unsigned long long* step_bits;
  const int *first_any, *first_ns, *first_member, *pair_cnt, *read_later;
  const double* scale;
};
// A cluster without pairs keeps its first particle's last step (:594-596);
// first_member (clusters built on the device) names it, otherwise dt.  A first
// pair with a static body a adds its use only when its body b is read later.
__device__ __forceinline__ double cluster_step_of(int k, float dt, const ClusterFinal& F) {
  double s;
  if (F.step_bits[k] != ~0ull) s = __longlong_as_double((long long)F.step_bits[k]);
  else if (F.first_any[k] != 0x7fffffff) s = F.bodies[40 * (size_t)F.pairs[2 * F.first_any[k]] + 39];
  else if (F.first_member) s = F.bodies[40 * (size_t)F.first_member[k] + 39];
  else s = (double)dt;
  const int p0 = F.first_ns[k];
  if (p0 != 0x7fffffff && F.pair_cnt[p0] > 0) {
    const double* ra = F.bodies + 40 * (size_t)F.pairs[2 * p0];
    const double* rb = F.bodies + 40 * (size_t)F.pairs[2 * p0 + 1];
    if (ra[38] != 0.0 && F.read_later[F.pairs[2 * p0 + 1]]) {
      const double mt = (rb[39] < ra[39]) ? rb[39] : ra[39];
      const double u = mt * F.scale[p0];
      s = (u < s) ? u : s;
    }
  }
  const double hi = (double)dt;
  s = (hi < s) ? hi : s;         // std::min(s, dt)
  return (s < 1e-3) ? 1e-3 : s;  // std::max(1e-3, .)
}
// Cluster steps, and with body_step the per-body steps of bodies in pairs
// (their cluster's step, computed per pair from the same accumulators).
__global__ void k_cluster_final(int n_clusters, float dt, ClusterFinal F, double* out, int n_pairs,
                                const int32_t* cluster_of_pair, double* body_step) {
  const int n = body_step && n_pairs > n_clusters ? n_pairs : n_clusters;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    if (k < n_clusters) out[k] = cluster_step_of(k, dt, F);
    if (body_step && k < n_pairs) {
      const double s = cluster_step_of(cluster_of_pair[k], dt, F);
      body_step[F.pairs[2 * k]] = s;
      body_step[F.pairs[2 * k + 1]] = s;
    }
  }
}

#include "d2_learned.hpp"
#include "d2_broad.hpp"

// Body future states: future = current.extrapolate(step).
// skip_static: static bodies keep their future state (PDE_explicit.cpp:49-56).
__global__ void k_project(int n, double* bodies, const double* step, int skip_static) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double* r = bodies + 40 * (size_t)i;
    if (skip_static && r[38] != 0.0) continue;
    body_state f = extrapolate(load_state(r), step[i], r + 28);
    store_state(f, r + 14);
  }
}

// ------------------------------------------- sequential impulses (the solver)
// After a filtered step: the clusters' pair lists and hit counts, then one
// thread per cluster walks SequentialImpulses::solve (d2_solver.hpp).
struct SolveArgs {
  int n_pairs, n_clusters, n_bodies;
  const int32_t* pairs;
  const int32_t* cluster_of_pair;
  const int* kept_cnt;    // per pair: filtered contacts
  const int* pair_off;    // per pair: first contact position
  const int* out_off;     // per pair: first output row (filtered CSR)
  const int* perm;        // contact at each position
  const int* keep;        // kept flag at each position
  const ContactRec* contacts;
  const double* cluster_step;
  const double* body_step;
  double* bodies;
  const double* mass;
  const double* fric;
  double gx, gy, gz;
  int* ccount;  // [n_clusters + 1] pairs per cluster, then their offsets (poff)
  int* chits;   // [n_clusters + 1] hits per cluster, then offsets (hoff)
  int* cfill;   // [n_clusters]
  int* cpairs;  // [n_pairs]
  int* inpair;  // [n_bodies] 1 + the cluster of the body's pairs, 0: in no pair
  int* loc;     // [n_bodies] local particle index of a dynamic body
  int* flag;    // set when a dynamic body's pairs lie in two clusters
  int* status;  // [n_clusters]
  int* hl;      // [hits][2]
  int* ord;     // [hits]
  int* hrow;    // [hits] output row
  int* hpc;     // [hits] contacts of the hit's pair
  int* pbody;   // [2 n_pairs]
  double* hA;   // [hits][3]
  double* hB;   // [hits][3]
  double* heps; // [hits]
  double* hkey; // [hits] |A - B| / (eps_a + eps_b)
  double* hs;   // [hits][SH]
  double* ps;   // [2 n_pairs][SP]
  double* out;  // [hits][11], cluster-gathered order
  WarmTable warm;
};
__global__ void k_fill_i32(int* p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_sv_count(SolveArgs A) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < A.n_pairs; p += gridDim.x * blockDim.x) {
    const int k = A.cluster_of_pair[p];
    atomicAdd(&A.ccount[k], 1);
    atomicAdd(&A.chits[k], A.kept_cnt[p]);
    for (int e = 0; e < 2; ++e) {  // a dynamic body belongs to the one cluster of its pairs
      const int b = A.pairs[2 * p + e];
      const int old = atomicCAS(&A.inpair[b], 0, k + 1);
      if (old != 0 && old != k + 1 && A.bodies[40 * (size_t)b + 38] == 0.0) atomicOr(A.flag, 1);
    }
  }
}
// ccount / chits hold the exclusive offsets here (poff / hoff)
__global__ void k_sv_place(SolveArgs A) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < A.n_pairs; p += gridDim.x * blockDim.x) {
    const int k = A.cluster_of_pair[p];
    A.cpairs[A.ccount[k] + atomicAdd(&A.cfill[k], 1)] = p;
  }
}
struct IdLess {
  D2_HD bool operator()(int x, int y) const { return x < y; }
};
struct KeyLess {
  const double* k;
  D2_HD bool operator()(int x, int y) const { return k[x] < k[y]; }
};
// One thread per cluster.  The cluster's pairs in index order (the order its
// interactions were gathered in, separate_clusters.cpp:405-434), its hits per
// pair in the filtered order (PDE_explicit.cpp:77-95), the std::sort replay
// of the hits by Contact::operator< (contact.h:14-18), then the solve.
__global__ void __launch_bounds__(64) k_sv_solve(SolveArgs A) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A.n_clusters; k += gridDim.x * blockDim.x) {
    const int p0 = A.ccount[k], np_k = A.ccount[k + 1] - p0;
    const int h0 = A.chits[k];
    int* cp = A.cpairs + p0;
    StdSort<IdLess>{cp, IdLess{}}.sort(0, np_k);
    int* pb = A.pbody + 2 * (size_t)p0;
    int n_part = 0, nh = 0;
    auto local_of = [&](int b) {
      if (A.bodies[40 * (size_t)b + 38] == 0.0) {  // a dynamic body belongs to one cluster
        if (A.loc[b] < 0) A.loc[b] = n_part, pb[n_part++] = b;
        return A.loc[b];
      }
      for (int q = 0; q < n_part; ++q)  // a static body: one particle per cluster it touches
        if (pb[q] == b) return q;
      pb[n_part] = b;
      return n_part++;
    };
    for (int i = 0; i < np_k; ++i) {
      const int p = cp[i];
      const int la = local_of(A.pairs[2 * p]), lb = local_of(A.pairs[2 * p + 1]);
      int row = A.out_off[p];
      for (int t = A.pair_off[p]; t < A.pair_off[p + 1]; ++t) {
        if (!A.keep[t]) continue;
        const ContactRec& c = A.contacts[A.perm[t]];
        const size_t H = (size_t)h0 + nh;
        for (int j = 0; j < 3; ++j) A.hA[3 * H + j] = c.r[j], A.hB[3 * H + j] = c.r[3 + j];
        A.heps[H] = c.r[6] + c.r[7];
        A.hkey[H] = vlen(sub(v3<double>{c.r[0], c.r[1], c.r[2]}, v3<double>{c.r[3], c.r[4], c.r[5]})) / A.heps[H];
        A.hl[2 * H] = la, A.hl[2 * H + 1] = lb;
        A.hrow[H] = row++;
        A.hpc[H] = A.kept_cnt[p];
        A.ord[H] = nh;
        ++nh;
      }
    }
    StdSort<KeyLess>{A.ord + h0, KeyLess{A.hkey + h0}}.sort(0, nh);
    ClusterSolve S;
    S.nh = nh, S.np = n_part;
    S.hA = A.hA + 3 * (size_t)h0, S.hB = A.hB + 3 * (size_t)h0, S.heps = A.heps + h0, S.hl = A.hl + 2 * (size_t)h0;
    S.hpc = A.hpc + h0;
    S.ord = A.ord + h0, S.pbody = pb, S.bodies = A.bodies, S.mass = A.mass, S.fric = A.fric;
    S.t = A.cluster_step[k], S.g = v3<double>{A.gx, A.gy, A.gz}, S.warm = A.warm;
    S.hs = A.hs + (size_t)SH * h0, S.ps = A.ps + (size_t)SP * 2 * p0;
    S.out_h = A.out + 11 * (size_t)h0;
    A.status[k] = solve_cluster(S) ? 1 : 0;
  }
}
// Bodies in no pair: a cluster of their own without contacts (gravity only).
__global__ void k_sv_free(SolveArgs A) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < A.n_bodies; b += gridDim.x * blockDim.x) {
    double* r = A.bodies + 40 * (size_t)b;
    if (A.inpair[b] || r[38] != 0.0) continue;
    const double t = A.body_step[b];
    const v3<double> g{A.gx, A.gy, A.gz}, z{0.0, 0.0, 0.0};
    const SolverState u = sv_update(sv_load(r + 14), t, scl(sclL(A.mass[b], g), t), z, z, z, A.mass[b], r + 28);
    r[14] = u.q.w, r[15] = u.q.x, r[16] = u.q.y, r[17] = u.q.z;
    r[18] = u.x.x, r[19] = u.x.y, r[20] = u.x.z, r[21] = u.v.x, r[22] = u.v.y, r[23] = u.v.z;
    r[24] = u.L.x, r[25] = u.L.y, r[26] = u.L.z;
  }
}
// Per-hit results to the filtered CSR rows.
__global__ void k_sv_scatter(long long n, const int* hrow, const double* in, double* out) {
  for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < n; h += (long long)gridDim.x * blockDim.x)
    for (int j = 0; j < 11; ++j) out[11 * (size_t)hrow[h] + j] = in[11 * (size_t)h + j];
}
// PDEExplicit::step's tail (PDE_explicit.cpp:117-131) for the dynamic bodies:
// current <- future, last step <- the body's step, future <- extrapolate.
__global__ void k_commit(int n, double* bodies, const double* step) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double* r = bodies + 40 * (size_t)i;
    if (r[38] != 0.0) continue;
    for (int j = 0; j < 14; ++j) r[j] = r[14 + j];
    r[39] = step[i];
    store_state(extrapolate(load_state(r), step[i], r + 28), r + 14);
  }
}
// Re-inserts the keys of an old warm table into a new one.
__global__ void k_warm_rehash(const unsigned long long* keys, const double* vals, long long n, WarmTable W) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    if (keys[i] != WARM_EMPTY) warm_put(W, keys[i], vals[i]);
}

}  // namespace d2

// =========================================================================
// Host side: context and the C ABI.
// =========================================================================
using namespace d2;

namespace {

const char* kKernelNames[] = {"setup", "soup_cull", "soup_point", "soup_edge", "soup_reduce", "expand",
                              "leaf_point", "leaf_edge", "tri_mask", "tri_enum", "tri_dist", "post", "post_pairs",
                              "cluster", "gate", "tri_full", "soup_keep", "tri_def"};
// K_TRI_ENUM launches nothing (the enumeration is part of k_tri_mask); its
// item count is the discrete node pairs received, kept for instrumentation.
// FP64 contexts run K_TRI_FULL (k_tri_full) in place of K_TRI_MASK + K_TRI_DIST.
enum KId { K_SETUP, K_SOUP_CULL, K_SOUP_POINT, K_SOUP_EDGE, K_SOUP_REDUCE, K_EXPAND, K_LEAF_POINT, K_LEAF_EDGE,
           K_TRI_MASK, K_TRI_ENUM, K_TRI_DIST, K_POST, K_POST_PAIRS, K_CLUSTER, K_GATE, K_TRI_FULL, K_SOUP_KEEP, K_TRI_DEF,
           K_NUM };

template <class T> struct DBuf {
  T* p = nullptr;
  size_t n = 0;
};

enum RunKind { RUN_NARROW = 0, RUN_CCD = 1, RUN_STEP = 2, RUN_KINDS = 3 };
constexpr int MAX_SEGMENTS = 3;

// Everything a captured run bakes in besides buffer addresses (a
// reallocation drops every graph).
struct Sig {
  int n_pairs = -1, n_bodies = -1, n_clusters = -1, max_depth = -1, precision = -1, filter = -1;
  float dt = -1.f;
  d2_tune tune{0, 0, 0, 0};
  int levels[2] = {-1, -1};  // graph levels of each phase
  int learned = -1;
  float gate_t = 0.f, gate_s = 0.f;
  int cl_dev = -1;  // clusters built on the device (first members known)
  unsigned long long grids = 0;  // hash of the per-level launch grids
  bool operator==(const Sig& o) const {
    return grids == o.grids && cl_dev == o.cl_dev && levels[0] == o.levels[0] && levels[1] == o.levels[1] && learned == o.learned && gate_t == o.gate_t &&
           gate_s == o.gate_s && n_pairs == o.n_pairs && n_bodies == o.n_bodies && n_clusters == o.n_clusters && max_depth == o.max_depth &&
           precision == o.precision && filter == o.filter && dt == o.dt &&
           tune.chunk_blocks == o.tune.chunk_blocks && tune.stream_blocks == o.tune.stream_blocks &&
           tune.post_blocks == o.tune.post_blocks;
  }
};
struct Captured {  // one instantiated graph per run kind
  hipGraphExec_t exec = nullptr;
  bool valid = false;
  Sig sig;
};

}  // namespace

struct d2_ctx {
  int device = 0;
  int precision = D2_FP64;
  hipStream_t stream = nullptr;
  std::string err;
  // capacities: node pairs per list, triangle-pair / sub-test items per
  // level, contacts per run, CCD soup items with result slots (15 x 64 B)
  long long cap = 1ll << 22;
  long long cap_items = 1ll << 24;
  long long cap_contacts = 1ll << 22;
  long long cap_res = 1ll << 20;
  bool auto_grow = true;  // grow the capacities and re-run on overflow
  int cus = 256;
  d2_tune tune{1024, 2048, 4096, 3};  // 4, 8, 16 per CU of a 256-CU device; reset at creation
  // trees / bodies
  DBuf<NodeRec> nodes;
  DBuf<double> buckets;
  DBuf<float> fbuckets;  // FP32 form: float corners + eps per triangle (FB_STRIDE), built on first use
  DBuf<float> stris;     // FP32 form: soup_tri planes per triangle (ST_STRIDE), built with fbuckets
  bool fb_ready = false;
  long long n_nodes_total = 0;  // nodes of the uploaded trees
  long long graph_captures = 0, graph_launches = 0;  // graph segments captured / launched (d2_graph_stats)
  DBuf<int32_t> tree_off, body_tree;
  DBuf<double> bodies;
  int n_trees = 0, n_bodies = 0;
  int pair_depth = 0;                       // max over pairs of depth(tree a) + depth(tree b)
  std::vector<int> tree_depth, host_body_tree;
  // pairs and clusters
  DBuf<int32_t> pairs, pair_info, cluster_of_pair;
  DBuf<double> max_time, pair_rec;
  int n_pairs = 0, n_clusters = 0;
  // a pair list describes the uploaded bodies: a tree upload or a body upload
  // of another count invalidates it (and its clusters) until pairs come again
  bool pairs_valid = false;
  // work
  DBuf<NP> soup_list[2], conn_list[2];  // level l reads [l & 1] and writes [(l + 1) & 1]
  DBuf<Item> items, pt_items, ee_items;
  DBuf<DItem> ditems;
  DBuf<DItem> ditems_def;  // FP32 form: deferred items of inner node pairs
  DBuf<double> res;
  DBuf<unsigned> res_mask;  // FP32 form: per soup item, the slots its FP32 culls keep
  DBuf<int> np_hit, np_hit2, exp_desc;
  DBuf<ContactRec> contacts;
  DBuf<unsigned long long> cnt, lvl, masks, conn_mask;
  DBuf<int> cnt_a, cnt_s, cnt_c, cnt_pt, cnt_ee, parts;
  DBuf<unsigned long long> pass_min_s, pass_min_c;
  DBuf<int> pass_flag_s, pass_flag_c;
  DBuf<double> min_time, scale;
  // post
  DBuf<int> pair_cnt, pair_off, pair_fill, perm, keep, kept_cnt, out_off, big, huge, n_big, scan_part;
  DBuf<int> hscratch;  // 6 x contact capacity: the filter of pairs above every LDS capacity
  DBuf<float> nbound;  // per node: leaf box lo, hi (body frame), largest leaf eps + slack, sphere r, centre, 0
  DBuf<uint8_t> kid_inner;  // per node: which children are inner nodes (bit c: child c)
  bool prune_on = true;  // skip node pairs whose leaf triangles are provably apart
  bool defer_on = true;  // FP32 form: an inner node pair's items after its first go to k_tri_def
  // the solver (d2_solve): per-body mass and friction, per-cluster and per-hit scratch, the warm store
  DBuf<double> sv_mass, sv_fric;
  DBuf<int> sv_ccount, sv_chits, sv_poff, sv_hoff, sv_cfill, sv_cpairs, sv_inpair, sv_loc, sv_status, sv_flag;
  DBuf<int> sv_hl, sv_ord, sv_hrow, sv_hpc, sv_pbody;
  DBuf<double> sv_hA, sv_hB, sv_heps, sv_hkey, sv_hs, sv_ps, sv_out, sv_out_csr;
  DBuf<unsigned long long> warm_keys, warm_used;
  DBuf<double> warm_vals;
  int last_kind = -1;         // kind of the last finished run
  long long run_serial = 0;   // finished runs
  long long solved_serial = -1;  // run_serial the last d2_solve worked on
  long long solved_hits = 0;
  DBuf<int32_t> out_i;
  DBuf<double> out_r;
  // clusters
  DBuf<unsigned long long> step_bits;
  DBuf<int> first_any, first_ns, read_later;
  DBuf<double> cluster_step, body_step;
  DBuf<int32_t> tmp_cluster;
  // broad phase (per tree: float AABB of the mesh and bounding radius)
  DBuf<float> tb_box;
  DBuf<double> tb_rad, last_state, bp_R;
  DBuf<int32_t> tree_depth_d, group;
  bool has_groups = false;
  std::vector<int> host_static;
  int n_items = 0;
  DBuf<int> bp_statics, bp_item_body, bp_item_seg, bp_item_cnt, bp_item_sorted, bp_item_off, bp_meta, bp_slot,
      bp_rank, bp_counts, bp_start, bp_members, bp_red;
  DBuf<long long> bp_item_pos;
  DBuf<unsigned long long> bp_keys, bp_fill;
  DBuf<Box8> bp_box;
  DBuf<int2> bp_chunk;
  long long bp_cap = 0;
  int* host_red = nullptr;  // pinned: BR_N reductions + pair total
  // clusters built on the device
  DBuf<int> cc_parent, cc_label, cc_root, cc_root_off, cc_of_body, cc_first;
  DBuf<long long> cc_tmin;
  DBuf<unsigned long long> cc_smin;
  bool clusters_on_device = false;
  // learned gate
  DBuf<float> tree_radius, tree_eps_out, hv, feat, pred;
  DBuf<int32_t> hv_off;
  DBuf<int> cert;  // per pair: 1 = traverse (not certified)
  int gate_phases = 3;       // bit 0: CCD phase, bit 1: discrete phase
  DBuf<float> mlp_w, mlp_m, mlp_v, mlp_wf, mlp_grad;
  DBuf<short> mlp_wb;
  DBuf<float> tr_x, tr_t, tr_act, tr_dlt, tr_loss;
  DBuf<int> tr_idx;
  bool mlp_ready = false, learned_on = false, has_max_time = false;
  float gate_t = 1.1f, gate_s = 0.99f;
  long long mlp_step = 0;
  // host mirrors (pinned)
  unsigned long long* host_cnt = nullptr;
  unsigned long long* host_lvl = nullptr;
  long long n_out = 0;
  int last_mode = -1;  // 0 narrow, 1 ccd (per-pair CCD values valid in both after a step)
  bool ccd_valid = false;
  int last_filter = 1;
  bool out_filtered = false;
  // graphs
  bool graph_on = false;
  bool capturing = false;
  Captured graphs[RUN_KINDS][MAX_SEGMENTS];
  int phase_levels[4] = {-1, -1, -1, -1};  // deepest level count seen per phase (narrow, ccd, step ccd, step narrow)
  int graph_levels[2] = {0, 0};             // level counts of the last graph run's phases
  // level sizes of the last run of each phase (0 discrete, 1 CCD): launch
  // grids of the next runs' levels are sized from them
  long long prev_lv[2][MAX_LEVELS + 1][LV_STRIDE] = {};
  bool prev_ok[2] = {false, false};
  long long prev_contacts[2] = {-1, -1};  // contacts of the last run's discrete / CCD phase (-1: none yet)
  // algorithmic-byte accounting (d2_set_touch_stats): bitmaps over (body, node)
  // and bodies, per level and per run; counts per phase and level, then per run
  bool touch_on = false;
  std::vector<int32_t> host_tree_off;
  DBuf<long long> touch_base;
  DBuf<unsigned> touch_lv, touch_run, touch_lvb, touch_runb;
  DBuf<unsigned long long> touch_acc;
  long long touch_words = 0, touch_body_words = 0;
  std::vector<unsigned long long> touch_host;
  // bucket-level entry and standalone filter
  unsigned char* noalloc = nullptr;  // this device's copy of `scratch` (resolved at creation)
  DBuf<int32_t> bk_np, bk_used, bk_out_i;
  DBuf<long long> bk_base, bk_off;
  DBuf<int> bk_count;
  DBuf<ContactRec> bk_hits;
  DBuf<double> bk_out_r;
  DBuf<unsigned long long> bk_counter;
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
#define RC(expr)            \
  do {                      \
    int rc_ = (expr);       \
    if (rc_) return rc_;    \
  } while (0)

void drop_graphs(d2_ctx* c) {
  for (auto& row : c->graphs)
    for (auto& g : row) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = Captured{};
    }
}

// The pair list, its clusters and every result refer to the uploaded bodies:
// forget them when the bodies they index change (runs then fail with
// D2_ERR_STATE until a new pair list is uploaded or built).
void invalidate_pairs(d2_ctx* c) {
  c->pairs_valid = false;
  c->n_pairs = 0;
  c->n_clusters = 0;
  c->clusters_on_device = false;
  c->last_mode = -1;
  c->ccd_valid = false;
}

// (Re)allocation invalidates captured graphs: they hold buffer addresses.
template <class T> int ensure(d2_ctx* c, DBuf<T>& b, size_t n) {
  if (b.n >= n && b.p) return D2_OK;
  if (c->capturing) return fail(c, D2_ERR_STATE, "internal: allocation during graph capture");
  drop_graphs(c);
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
int chunk_grid(const d2_ctx* c) { return c->tune.chunk_blocks; }
int stream_grid(const d2_ctx* c) { return c->tune.stream_blocks; }

struct Timer {
  d2_ctx* c;
  int id;
  double items;
  hipEvent_t a = nullptr, b = nullptr;
  bool on;
  Timer(d2_ctx* ctx, int k, double n) : c(ctx), id(k), items(n), on(ctx->profiling && !ctx->capturing) {
    if (on) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~Timer() {
    if (on) {
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

// Launch of one kernel that, when profiling, is timed by events the GPU
// stamps at the kernel's own start and end (no host submission gaps).
template <typename F, typename... Args>
void timed_launch(d2_ctx* c, int kid, F kernel, dim3 grid, dim3 block, Args... args) {
  if (c->profiling && !c->capturing) {
    hipEvent_t a = nullptr, b = nullptr;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    hipExtLaunchKernelGGL(kernel, grid, block, 0, c->stream, a, b, 0, args...);
    c->pending.push_back({kid, {a, b}});
    c->pending_items.push_back(0.0);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, c->stream, args...);
  }
}

Dev make_dev(d2_ctx* c, int level, int discrete) {
  Dev D;
  D.nodes = c->nodes.p;
  D.buckets = c->buckets.p;
  D.pair_info = c->pair_info.p;
  D.pair_rec = c->pair_rec.p;
  D.soup = c->soup_list[level & 1].p;
  D.conn = c->conn_list[level & 1].p;
  D.soup_next = c->soup_list[(level + 1) & 1].p;
  D.conn_next = c->conn_list[(level + 1) & 1].p;
  D.items = c->items.p;
  D.ditems = c->ditems.p;
  D.ditems_def = c->ditems_def.p;
  D.pt_items = c->pt_items.p;
  D.ee_items = c->ee_items.p;
  D.res = c->res.p;
  D.res_mask = c->res_mask.p;
  D.np_hit = (level & 1) ? c->np_hit2.p : c->np_hit.p;
  D.np_hit_prev = (level & 1) ? c->np_hit.p : c->np_hit2.p;
  D.exp_desc = c->exp_desc.p;
  D.pair_cnt = c->pair_cnt.p;
  D.contacts = c->contacts.p;
  D.cnt = c->cnt.p;
  D.lvl = c->lvl.p;
  D.masks = c->masks.p;
  D.cnt_a = c->cnt_a.p, D.cnt_s = c->cnt_s.p, D.cnt_c = c->cnt_c.p, D.cnt_pt = c->cnt_pt.p, D.cnt_ee = c->cnt_ee.p;
  const int g = chunk_grid(c);
  D.part_a = c->parts.p, D.part_s = c->parts.p + g, D.part_c = c->parts.p + 2 * g, D.part_pt = c->parts.p + 3 * g,
  D.part_ee = c->parts.p + 4 * g;
  D.pass_min_s = c->pass_min_s.p, D.pass_min_c = c->pass_min_c.p;
  D.pass_flag_s = c->pass_flag_s.p, D.pass_flag_c = c->pass_flag_c.p;
  D.min_time = c->min_time.p;
  D.tree_eps_out = c->tree_eps_out.p;
  D.fbuckets = c->fbuckets.p;
  D.stris = c->stris.p;
  D.nbound = c->nbound.p;
  D.kid_inner = c->kid_inner.p;
  D.prune = c->prune_on && !c->touch_on && c->nbound.p ? 1 : 0;  // accounting runs walk the reference's node pairs
  D.defer = c->defer_on && c->precision == D2_FP32 && c->ditems_def.p ? 1 : 0;  // per level: enqueue_level
  D.box_test = 1;
  D.conn_mask = c->conn_mask.p;
  D.body_tree = c->body_tree.p;
  D.fp32 = c->precision == D2_FP32 ? 1 : 0;
  D.n_pairs = c->n_pairs;
  D.cap_np = c->cap;
  D.cap_items = c->cap_items;
  D.cap_res = c->cap_res;
  D.cap_contacts = c->cap_contacts;
  D.level = level;
  D.discrete = discrete;
  return D;
}

// A node pair's level is the sum of its nodes' depths, so a pair list needs
// (deepest depth sum) + 1 levels.
int level_budget(const d2_ctx* c) { return std::min(MAX_LEVELS, c->pair_depth + 1); }

// Every buffer a run touches, sized before any launch (and before capture).
int prepare(d2_ctx* c) {
  const size_t np = (size_t)std::max(c->n_pairs, 1);
  const size_t g = (size_t)chunk_grid(c);
  for (int k = 0; k < 2; ++k) {
    RC(ensure(c, c->soup_list[k], c->cap));
    RC(ensure(c, c->conn_list[k], c->cap));
  }
  RC(ensure(c, c->items, c->cap_items)); RC(ensure(c, c->pt_items, c->cap_items));
  RC(ensure(c, c->ee_items, c->cap_items));
  if (c->precision == D2_FP32) {  // FP64: k_tri_full keeps no item list
    RC(ensure(c, c->ditems, c->cap_items));
    RC(ensure(c, c->ditems_def, c->cap_items));
  }
  RC(ensure(c, c->np_hit, c->cap)); RC(ensure(c, c->np_hit2, c->cap)); RC(ensure(c, c->exp_desc, c->cap)); RC(ensure(c, c->masks, c->cap));
  RC(ensure(c, c->cnt_a, c->cap)); RC(ensure(c, c->cnt_s, c->cap)); RC(ensure(c, c->cnt_c, c->cap));
  RC(ensure(c, c->cnt_pt, c->cap)); RC(ensure(c, c->cnt_ee, c->cap)); RC(ensure(c, c->parts, 5 * g));
  RC(ensure(c, c->conn_mask, (size_t)c->cap * CONN_MASK_WORDS));
  RC(ensure(c, c->contacts, c->cap_contacts)); RC(ensure(c, c->cnt, C_COUNT));
  RC(ensure(c, c->hscratch, 6 * (size_t)c->cap_contacts));
  RC(ensure(c, c->lvl, (size_t)2 * (MAX_LEVELS + 1) * LV_STRIDE));
  RC(ensure(c, c->res, (size_t)c->cap_res * RES_SLOTS * 8));
  if (c->precision == D2_FP32) RC(ensure(c, c->res_mask, (size_t)c->cap_res));
  if (c->precision == D2_FP32 && !c->fb_ready) {  // the float copy of the buckets, once per tree upload
    const size_t nodes = (size_t)c->n_nodes_total;
    RC(ensure(c, c->fbuckets, nodes * 8 * FB_STRIDE));
    RC(ensure(c, c->stris, nodes * 8 * ST_STRIDE));
    hipLaunchKernelGGL(k_bucket_f32, dim3(grid_for(c, (long long)nodes * 8)), dim3(256), 0, c->stream, c->buckets.p,
                       (long long)nodes, c->fbuckets.p, c->stris.p);
    HIPCHK(c, hipGetLastError());
    c->fb_ready = true;
  }
  RC(ensure(c, c->pass_min_s, np)); RC(ensure(c, c->pass_min_c, np)); RC(ensure(c, c->pass_flag_s, np));
  RC(ensure(c, c->pass_flag_c, np)); RC(ensure(c, c->min_time, np)); RC(ensure(c, c->scale, np));
  RC(ensure(c, c->pair_cnt, np + 1)); RC(ensure(c, c->pair_off, np + 1)); RC(ensure(c, c->pair_fill, np));
  RC(ensure(c, c->kept_cnt, np + 1)); RC(ensure(c, c->out_off, np + 1)); RC(ensure(c, c->big, np));
  RC(ensure(c, c->n_big, 2)); RC(ensure(c, c->huge, np)); RC(ensure(c, c->perm, c->cap_contacts)); RC(ensure(c, c->keep, c->cap_contacts));
  RC(ensure(c, c->scan_part, g));
  if (c->learned_on) {
    RC(ensure(c, c->feat, np * MLP_IN)); RC(ensure(c, c->pred, np * MLP_OUT)); RC(ensure(c, c->cert, np));
  }
  return D2_OK;
}

// Exclusive scan of n (host-known) ints on the context stream.
void scan_host_n(d2_ctx* c, const int* in, int* out, long long n, unsigned long long* total_dst) {
  if (n <= SCAN_SMALL) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, c->stream, in, out, n, total_dst);
    return;
  }
  const int g = chunk_grid(c);
  hipLaunchKernelGGL(k_scan_part, dim3(g), dim3(256), 0, c->stream, in, n, c->scan_part.p);
  hipLaunchKernelGGL(k_scan_apply, dim3(g), dim3(256), 0, c->stream, in, out, n, c->scan_part.p, total_dst);
}

// One breadth-first level of a phase.
// Launch grids of one level: the tuned block counts are the caps; a level
// that was small in the last run of its phase gets fewer blocks (at least 16,
// in powers of two, with 25% headroom): the culling kernels keep about one
// node pair per wavefront (they are latency-bound, so as many wavefronts as
// there is work), the count/write pair one soup node pair per thread or one
// connected node pair per wavefront, the stream kernels one unit per thread.
// Any grid gives the same results; only the time changes.
// def: k_tri_def's grid, 0 where the level defers nothing; box: whether
// k_tri_mask runs the face-axis box test.  Both pay on large levels only (on
// small ones the extra launch and the extra dependent load cost more than
// they save), judged by the last run's sizes of the level: at least 64 K
// items, at least 256 K node pairs.
struct LevelGrids {
  int cull, expand, stream, def, box;
};
LevelGrids level_grids(const d2_ctx* c, int discrete, int level) {
  LevelGrids g{chunk_grid(c), chunk_grid(c), stream_grid(c), stream_grid(c), 1};
  const int ph = discrete ? 0 : 1;
  if (!c->prev_ok[ph] || level > MAX_LEVELS) return g;
  const long long* v = c->prev_lv[ph][level];
  const long long ns = v[LV_SOUP], nc = v[LV_CONN];
  const long long units = discrete ? v[LV_ITEMS] : std::max(std::max(9 * v[LV_ITEMS], v[LV_PT]), v[LV_EE]);
  auto fit = [](long long n, long long per, int hi) {
    long long b = 16;
    while (b < hi && b * per < n + n / 4) b <<= 1;
    return (int)std::min<long long>(b, hi);
  };
  g.cull = fit(ns, 4, g.cull);
  g.expand = fit(std::max((ns + 63) / 64, nc), 4, g.expand);
  g.stream = fit(units, 256, g.stream);
  if (discrete) {
    g.def = v[LV_ITEMS] + v[LV_DEF] >= (1ll << 16) ? fit(v[LV_DEF], 256, g.def) : 0;
    g.box = ns >= (1ll << 18) ? 1 : 0;
  }
  return g;
}
unsigned long long grids_hash(const d2_ctx* c, int discrete, int levels) {
  unsigned long long h = 1469598103934665603ull;
  for (int l = 0; l < levels; ++l) {
    const LevelGrids g = level_grids(c, discrete, l);
    h = (h ^ (unsigned long long)((((g.cull * 131071ll + g.expand) * 131071ll + g.stream) * 131071ll + g.def) * 2 +
                                  g.box)) *
        1099511628211ull;
  }
  return h;
}
// Accounting runs (eager only): counts the (body, node) records this level
// tests into touch_acc[phase][level], and marks them in the run's bitmaps.
constexpr int TOUCH_FIELDS = 4;  // inner nodes, leaf nodes, leaf triangles, bodies
constexpr int TOUCH_RUN = 2 * (MAX_LEVELS + 1) * TOUCH_FIELDS;  // offset of the run's union
void touch_level(d2_ctx* c, int level, int discrete) {
  (void)hipMemsetAsync(c->touch_lv.p, 0, sizeof(unsigned) * c->touch_words, c->stream);
  (void)hipMemsetAsync(c->touch_lvb.p, 0, sizeof(unsigned) * c->touch_body_words, c->stream);
  TouchArgs T;
  T.soup = c->soup_list[level & 1].p, T.conn = c->conn_list[level & 1].p, T.lvl = c->lvl.p;
  T.np_hit_prev = (level & 1) ? c->np_hit.p : c->np_hit2.p;
  T.pair_info = c->pair_info.p, T.node_base = c->touch_base.p;
  T.lv_nodes = c->touch_lv.p, T.run_nodes = c->touch_run.p, T.lv_bodies = c->touch_lvb.p, T.run_bodies = c->touch_runb.p;
  T.cap_np = c->cap, T.level = level, T.discrete = discrete;
  hipLaunchKernelGGL(k_touch, dim3(grid_for(c, c->cap)), dim3(256), 0, c->stream, T);
  unsigned long long* out = c->touch_acc.p + ((discrete ? 0 : 1) * (MAX_LEVELS + 1) + level) * TOUCH_FIELDS;
  hipLaunchKernelGGL(k_touch_count, dim3(grid_for(c, c->touch_words)), dim3(256), 0, c->stream, c->touch_lv.p,
                     c->touch_words, c->touch_lvb.p, c->touch_body_words, c->touch_base.p, c->n_bodies,
                     c->body_tree.p, c->tree_off.p, c->nodes.p, out);
}

// Start of an accounting run: the (body, node) numbering of the current
// bodies, cleared bitmaps and counts.
int touch_begin(d2_ctx* c) {
  std::vector<long long> base(c->n_bodies + 1, 0);
  for (int b = 0; b < c->n_bodies; ++b) {
    const int t = c->host_body_tree[b];
    base[b + 1] = base[b] + (c->host_tree_off[t + 1] - c->host_tree_off[t]);
  }
  c->touch_words = std::max<long long>((base[c->n_bodies] + 31) / 32, 1);
  c->touch_body_words = std::max<long long>((c->n_bodies + 31) / 32, 1);
  RC(ensure(c, c->touch_base, base.size()));
  RC(ensure(c, c->touch_lv, c->touch_words)); RC(ensure(c, c->touch_run, c->touch_words));
  RC(ensure(c, c->touch_lvb, c->touch_body_words)); RC(ensure(c, c->touch_runb, c->touch_body_words));
  RC(ensure(c, c->touch_acc, TOUCH_RUN + TOUCH_FIELDS));
  HIPCHK(c, hipMemcpyAsync(c->touch_base.p, base.data(), sizeof(long long) * base.size(), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemsetAsync(c->touch_run.p, 0, sizeof(unsigned) * c->touch_words, c->stream));
  HIPCHK(c, hipMemsetAsync(c->touch_runb.p, 0, sizeof(unsigned) * c->touch_body_words, c->stream));
  HIPCHK(c, hipMemsetAsync(c->touch_acc.p, 0, sizeof(unsigned long long) * (TOUCH_RUN + TOUCH_FIELDS), c->stream));
  return D2_OK;
}

// End of an accounting run: the union's counts, and every count to the host.
int touch_end(d2_ctx* c) {
  hipLaunchKernelGGL(k_touch_count, dim3(grid_for(c, c->touch_words)), dim3(256), 0, c->stream, c->touch_run.p,
                     c->touch_words, c->touch_runb.p, c->touch_body_words, c->touch_base.p, c->n_bodies,
                     c->body_tree.p, c->tree_off.p, c->nodes.p, c->touch_acc.p + TOUCH_RUN);
  HIPCHK(c, hipGetLastError());
  c->touch_host.assign(TOUCH_RUN + TOUCH_FIELDS, 0ull);
  HIPCHK(c, hipMemcpyAsync(c->touch_host.data(), c->touch_acc.p, sizeof(unsigned long long) * c->touch_host.size(),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

void enqueue_level(d2_ctx* c, int level, int discrete) {
  if (c->touch_on && !c->capturing) touch_level(c, level, discrete);
  Dev D = make_dev(c, level, discrete);
  const LevelGrids G = level_grids(c, discrete, level);
  const int cg = G.cull, eg = G.expand, sg = G.stream;
  D.defer = D.defer && G.def > 0 ? 1 : 0;
  D.box_test = G.box;
  if (discrete && c->precision == D2_FP32) {
    timed_launch(c, K_TRI_MASK, k_tri_mask, dim3(cg), dim3(256), D);
    timed_launch(c, K_TRI_DIST, k_tri_dist, dim3(sg), dim3(256), D);
    if (D.defer) timed_launch(c, K_TRI_DEF, k_tri_def, dim3(G.def), dim3(256), D);
  } else if (discrete) {  // one wavefront per 64 node pairs
    timed_launch(c, K_TRI_FULL, k_tri_full, dim3(eg), dim3(256), D);
  } else {
    timed_launch(c, K_SOUP_CULL, k_soup_cull, dim3(cg), dim3(256), D);
    if (D.fp32) timed_launch(c, K_SOUP_KEEP, k_soup_keep, dim3(sg), dim3(256), D);  // FP32 culls + survivor lists
    timed_launch(c, K_SOUP_EDGE, k_soup_sub, dim3(2 * sg), dim3(256), D);  // point and edge sub-tests
    timed_launch(c, K_SOUP_REDUCE, k_soup_reduce, dim3(sg), dim3(256), D);
  }
  if (discrete && c->precision != D2_FP32) {
    Timer tm(c, K_EXPAND, 0.0);
    hipLaunchKernelGGL(k_disc_expand, dim3(eg), dim3(256), 0, c->stream, D);
  } else if (!discrete) {  // FP32 discrete levels are expanded by k_tri_mask
    Timer tm(c, K_EXPAND, 0.0);
    hipLaunchKernelGGL(k_level_count, dim3(eg), dim3(256), 0, c->stream, D);
    hipLaunchKernelGGL(k_level_write, dim3(eg), dim3(256), 0, c->stream, D);
  }
  if (!discrete) {
    // the connected list of level 0 is empty (roots go to the soup list)
    if (level > 0) timed_launch(c, K_LEAF_EDGE, k_leaf_sub, dim3(2 * sg), dim3(256), D);
  }
}

// Grid of the per-contact kernels of the post-processing: one contact per
// thread at the last run's count of the phase (mode 2: CCD) with 25%
// headroom, at least 16 blocks, at most the tuned stream grid.
int post_grid(const d2_ctx* c, int mode) {
  const int hi = stream_grid(c);
  const long long n = c->prev_contacts[mode >= 2 ? 1 : 0];
  if (n < 0) return hi;
  long long b = 16;
  while (b < hi && b * 256 < n + n / 4) b <<= 1;
  return (int)std::min<long long>(b, hi);
}

// The per-pair finishing launches over contacts, pair_off and perm: pairs up to
// 512 contacts in the first (wblocks blocks), those up to 2048 in the second
// launch's LDS pass, larger ones by the second launch's per-pair forms (with
// the filter, one lane per pair; n_big[0..1] count the two lists, zeroed
// before).  mode 1 applies filterContacts(.., margin).
HugeScratch huge_scratch(d2_ctx* c) {
  const size_t m = c->hscratch.n / 6;
  int* h = c->hscratch.p;
  return HugeScratch{h, h + m, h + 2 * m, h + 3 * m, h + 4 * m, h + 5 * m};
}
void launch_finish(d2_ctx* c, int np, int wblocks, int mode, unsigned long long* kept_total, double margin) {
  hipLaunchKernelGGL(k_pair_finish_block<BLK_CAP>, dim3(wblocks), dim3(256), 0, c->stream, np, (const int*)nullptr,
                     (const int*)nullptr, c->contacts.p, c->pair_off.p, c->perm.p, mode, c->keep.p, c->kept_cnt.p,
                     c->pair_rec.p, c->scale.p, c->big.p, c->n_big.p, c->huge.p, c->n_big.p + 1,
                     (const int*)nullptr, (const int*)nullptr, kept_total, margin, huge_scratch(c));
  hipLaunchKernelGGL(k_pair_finish_block<POST_CAP_LARGE>, dim3(std::max(1, std::min(np, c->cus))), dim3(256), 0,
                     c->stream, np, (const int*)c->big.p, (const int*)c->n_big.p, c->contacts.p, c->pair_off.p,
                     c->perm.p, mode, c->keep.p, c->kept_cnt.p, c->pair_rec.p, c->scale.p, c->huge.p,
                     c->n_big.p + 1, c->huge.p, c->n_big.p + 1, (const int*)c->huge.p,
                     (const int*)(c->n_big.p + 1), kept_total, margin, huge_scratch(c));
}

// Post-processing of the contacts of the last phase: per-pair order, then the
// contact filter (mode 1), the step rule (mode 2) or nothing (mode 0); mode 3
// (a step's CCD phase) runs the step rule alone, on the unordered contacts.
int enqueue_post(d2_ctx* c, int mode) {
  const int np = c->n_pairs;
  if (mode == 3) {
    Timer tm(c, K_POST_PAIRS, (double)np);
    hipLaunchKernelGGL(k_step_rule, dim3(post_grid(c, mode)), dim3(256), 0, c->stream, c->cnt.p, c->cap_contacts,
                       c->contacts.p, c->pair_rec.p, c->scale.p);
    HIPCHK(c, hipGetLastError());
    return D2_OK;
  }
  const int sg = post_grid(c, mode);
  const bool filtered = mode == 1;
  {
    Timer tm(c, K_POST, 0.0);  // per-pair counts come from the emitting kernels, zeroed by k_setup
    unsigned long long* total = filtered ? nullptr : c->cnt.p + C_OUT;
    const bool fuse = np <= SCAN_FUSE;  // the scan inside k_scatter
    if (!fuse) scan_host_n(c, c->pair_cnt.p, c->pair_off.p, np, total);
    hipLaunchKernelGGL(k_scatter, dim3(sg), dim3(256), 0, c->stream, c->cnt.p, c->cap_contacts, c->contacts.p,
                       c->pair_off.p, c->pair_fill.p, c->perm.p, fuse ? (const int*)c->pair_cnt.p : nullptr, np,
                       fuse ? total : nullptr);
  }
  {
    Timer tm(c, K_POST_PAIRS, (double)np);
    const int wblocks = std::max(1, std::min(np, c->tune.post_blocks));
    // pairs above the first launch's capacity (512 contacts): the second
    // launch's 2048-contact LDS pass (list big); above that, one block per
    // pair, or one lane per pair with the filter (list huge, filled by the
    // first launch only)
    // the filter's kept total accumulates into C_OUT (zeroed by k_setup); the
    // output offsets are scanned when the contacts are read (d2_result_contacts)
    unsigned long long* kept_total = filtered ? c->cnt.p + C_OUT : nullptr;
    launch_finish(c, np, wblocks, mode, kept_total, 0.0);  // filterContacts(contacts, 0.0), PDE_explicit.cpp:88
  }
  HIPCHK(c, hipGetLastError());
  return D2_OK;
}

ClusterAcc cluster_acc(d2_ctx* c, const int32_t* cluster_of_pair, int n_clusters, float dt, bool project) {
  ClusterAcc A{};
  A.n_clusters = n_clusters, A.n_bodies = c->n_bodies, A.dt = dt;
  A.step_bits = c->step_bits.p, A.first_any = c->first_any.p, A.first_ns = c->first_ns.p;
  A.read_later = c->read_later.p;
  A.bodies = c->bodies.p, A.body_step = project ? c->body_step.p : nullptr;
  A.pairs = c->pairs.p, A.cluster_of_pair = cluster_of_pair;
  return A;
}
// started: the accumulators were initialised and the first pairs recorded
// by the phase's k_setup / k_fold_final.
int enqueue_clusters(d2_ctx* c, const int32_t* cluster_of_pair, int n_clusters, float dt, bool project,
                     bool first_member, bool started) {
  const int np = c->n_pairs;
  Timer tm(c, K_CLUSTER, np);
  double* body_step = project ? c->body_step.p : nullptr;
  const ClusterAcc A = cluster_acc(c, cluster_of_pair, n_clusters, dt, project);
  if (!started)
    hipLaunchKernelGGL(k_cluster_init, dim3(grid_for(c, std::max(n_clusters, c->n_bodies))), dim3(256), 0, c->stream, A);
  if (np > 0) {
    if (!started) hipLaunchKernelGGL(k_cluster_first, dim3(grid_for(c, np)), dim3(256), 0, c->stream, np, A);
    hipLaunchKernelGGL(k_cluster_min, dim3(grid_for(c, np)), dim3(256), 0, c->stream, np, c->pairs.p, cluster_of_pair,
                       c->bodies.p, c->pair_cnt.p, c->scale.p, c->first_ns.p, c->read_later.p, c->step_bits.p);
  }
  ClusterFinal F;
  F.pairs = c->pairs.p, F.bodies = c->bodies.p, F.step_bits = c->step_bits.p, F.first_any = c->first_any.p;
  F.first_ns = c->first_ns.p, F.first_member = first_member ? c->cc_first.p : (const int*)nullptr;
  F.pair_cnt = c->pair_cnt.p, F.read_later = c->read_later.p, F.scale = c->scale.p;
  hipLaunchKernelGGL(k_cluster_final, dim3(grid_for(c, project ? std::max(n_clusters, np) : n_clusters)), dim3(256), 0,
                     c->stream, n_clusters, dt, F, c->cluster_step.p, np, cluster_of_pair, body_step);
  if (project)
    hipLaunchKernelGGL(k_project, dim3(grid_for(c, c->n_bodies)), dim3(256), 0, c->stream, c->n_bodies, c->bodies.p,
                       c->body_step.p, 1);
  HIPCHK(c, hipGetLastError());
  return D2_OK;
}

// ---------------------------------------------------------------- phases
// A run is a list of traversal phases, each: set-up + roots, breadth-first
// levels, closing fold + post-processing (a step adds the cluster steps and
// projections between its CCD and discrete phases).
struct PhaseSpec {
  int discrete, keep_ccd, mt_from_bodies, post_mode, slot;
};
int phases_of(int kind, int filter, PhaseSpec* out) {
  const int fmode = filter ? 1 : 0;
  if (kind == RUN_NARROW) {
    out[0] = {1, 0, 0, fmode, 0};
    return 1;
  }
  if (kind == RUN_CCD) {
    out[0] = {0, 0, 0, 2, 1};
    return 1;
  }
  out[0] = {0, 0, 1, 3, 2};  // the step never reads its CCD contacts: step rule only
  out[1] = {1, 1, 0, fmode, 3};
  return 2;
}

// Whether the learned gate runs in this phase.
bool gate_on(const d2_ctx* c, const PhaseSpec& ph) {
  return c->learned_on && c->mlp_ready && c->n_pairs > 0 && (c->gate_phases & (ph.discrete ? 2 : 1));
}

// The cluster-step accumulators of a step run, set up by its CCD phase's
// k_setup (initialisation) and k_fold_final (first pairs); none otherwise.
ClusterAcc step_acc(d2_ctx* c, int kind, const PhaseSpec& ph, float dt) {
  if (kind != RUN_STEP || ph.discrete) return ClusterAcc{};
  return cluster_acc(c, c->cluster_of_pair.p, c->n_clusters, dt, true);
}

void enqueue_begin(d2_ctx* c, int kind, const PhaseSpec& ph, float dt) {
  const int np = c->n_pairs;
  Timer tm(c, K_SETUP, np);
  Dev D = make_dev(c, 0, ph.discrete);
  hipLaunchKernelGGL(k_setup, dim3(grid_for(c, 4ll * std::max(np, 1))), dim3(256), 0, c->stream, D, c->pairs.p,
                     c->max_time.p, c->bodies.p, c->body_tree.p, c->tree_off.p, c->pair_info.p, c->pair_rec.p,
                     ph.discrete ? 0 : 1, ph.mt_from_bodies, ph.keep_ccd, c->scale.p, c->pair_fill.p, c->kept_cnt.p,
                     c->n_big.p, step_acc(c, kind, ph, dt), gate_on(c, ph) ? 1 : 0);
}

// The learned gate of a phase: proposals (once per run), certificates, and
// the phase's root list without the certified pairs.
void enqueue_gate(d2_ctx* c, const PhaseSpec& ph) {
  const int np = c->n_pairs;
  if (!gate_on(c, ph)) return;
  {
    Timer tg(c, K_GATE, np);
    Dev D = make_dev(c, 0, ph.discrete);
    if (!ph.keep_ccd || !(c->gate_phases & 1)) {  // a step's discrete phase reuses its CCD phase's proposals
      const double* mt = (ph.mt_from_bodies || !c->has_max_time) ? nullptr : c->max_time.p;
      const int tiles = (np + 15) / 16;  // features computed inside the inference kernel
      hipLaunchKernelGGL(k_mlp_infer<true>, dim3(std::max(1, std::min((tiles + 3) / 4, c->cus * 4))), dim3(256), 0,
                         c->stream, np, (const float*)nullptr, c->mlp_wb.p, c->pred.p,
                         FeatIn{c->pairs.p, c->bodies.p, c->body_tree.p, c->tree_radius.p, mt});
    }
    // at most CERT_LIST pairs per block; the kept pairs are the phase's roots
    const int grid = std::max(std::max(1, std::min(np, c->cus * 8)), (np + CERT_LIST - 1) / CERT_LIST);
    hipLaunchKernelGGL(k_certify, dim3(grid), dim3(256), 0, c->stream, np, c->pair_info.p, c->body_tree.p,
                       c->pair_rec.p, c->pred.p, c->gate_t, c->gate_s, c->hv_off.p, c->hv.p, c->tree_eps_out.p,
                       ph.discrete ? 1 : 0, c->cert.p, c->cnt.p + (ph.discrete ? C_GATED_DISC : C_GATED_CCD), D.soup,
                       c->lvl.p + LV_SOUP);
  }
}

// Items of level l per kernel, from its sizes (profiling).
void add_level_items(d2_ctx* c, int level, int discrete) {
  const unsigned long long* s = c->host_lvl + level * LV_STRIDE;
  if (discrete && c->precision != D2_FP32) {
    c->k_items[K_TRI_FULL] += (double)s[LV_ITEMS];
  } else if (discrete) {
    c->k_items[K_TRI_MASK] += (double)s[LV_SOUP];
    c->k_items[K_TRI_ENUM] += (double)s[LV_SOUP];
    c->k_items[K_TRI_DIST] += (double)s[LV_ITEMS];
    c->k_items[K_TRI_DEF] += (double)s[LV_DEF];
  } else {
    c->k_items[K_SOUP_CULL] += (double)s[LV_SOUP];
    const bool fp32 = c->precision == D2_FP32;  // FP32 form: the sub-tests its culls keep
    if (fp32) c->k_items[K_SOUP_KEEP] += (double)s[LV_ITEMS];
    c->k_items[K_SOUP_POINT] += fp32 ? (double)s[LV_SPT] : 6.0 * (double)s[LV_ITEMS];
    c->k_items[K_SOUP_EDGE] += fp32 ? (double)s[LV_SPT] + (double)s[LV_SEE] : 15.0 * (double)s[LV_ITEMS];
    c->k_items[K_SOUP_REDUCE] += (double)s[LV_ITEMS];
    c->k_items[K_LEAF_POINT] += (double)s[LV_PT];
    c->k_items[K_LEAF_EDGE] += (double)s[LV_EE];
  }
  c->k_items[K_EXPAND] += (double)s[LV_SOUP];
}

int read_slots(d2_ctx* c, int first, int count) {
  HIPCHK(c, hipMemcpyAsync(c->host_lvl + first * LV_STRIDE, c->lvl.p + first * LV_STRIDE,
                           sizeof(unsigned long long) * LV_STRIDE * count, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

// Eager levels from `level` on while the lists are not empty (sizes read back
// after every level).  *used receives the index one past the last level run.
int eager_levels(d2_ctx* c, const PhaseSpec& ph, int level, long long n_soup, long long n_conn, int* used) {
  const int budget = level_budget(c);
  for (; level < budget && n_soup + n_conn > 0; ++level) {
    enqueue_level(c, level, ph.discrete);
    HIPCHK(c, hipGetLastError());
    RC(read_slots(c, level, 2));
    if (c->profiling) add_level_items(c, level, ph.discrete);
    n_soup = (long long)c->host_lvl[(level + 1) * LV_STRIDE + LV_SOUP];
    n_conn = (long long)c->host_lvl[(level + 1) * LV_STRIDE + LV_CONN];
  }
  if (n_soup + n_conn > 0) return fail(c, D2_ERR_RUNTIME, "traversal deeper than the level budget of the uploaded trees");
  *used = level;
  return D2_OK;
}

void enqueue_levels(d2_ctx* c, const PhaseSpec& ph, int from, int to) {
  for (int level = from; level < to; ++level) enqueue_level(c, level, ph.discrete);
}

// Closing fold (CCD) and post-processing, then (a step's CCD phase) the
// cluster steps and future states.
int enqueue_end(d2_ctx* c, int kind, const PhaseSpec& ph, int levels, float dt) {
  if (!ph.discrete) {
    Dev D = make_dev(c, levels, 0);
    hipLaunchKernelGGL(k_fold_final, dim3(grid_for(c, std::max(c->n_pairs, 1))), dim3(256), 0, c->stream, D,
                       step_acc(c, kind, ph, dt));
  }
  RC(enqueue_post(c, ph.post_mode));
  if (kind == RUN_STEP && !ph.discrete)
    RC(enqueue_clusters(c, c->cluster_of_pair.p, c->n_clusters, dt, true, c->clusters_on_device, true));
  HIPCHK(c, hipGetLastError());
  return D2_OK;
}

int run_eager(d2_ctx* c, int kind, int filter, float dt) {
  PhaseSpec ph[2];
  const int n = phases_of(kind, filter, ph);
  for (int i = 0; i < n; ++i) {
    enqueue_begin(c, kind, ph[i], dt);
    enqueue_gate(c, ph[i]);
    HIPCHK(c, hipGetLastError());
    int used = 0;
    RC(eager_levels(c, ph[i], 0, c->n_pairs, 0, &used));
    c->phase_levels[ph[i].slot] = std::max(c->phase_levels[ph[i].slot], used);
    RC(enqueue_end(c, kind, ph[i], used, dt));
  }
  return D2_OK;
}

// Graph runs: each phase runs its observed level count inside the graph; when
// that is below the level budget, the graph ends there, the host reads the
// next level's sizes and continues eagerly if a deeper level is needed (and
// the phase's level count grows for the next capture).
int launch_segment(d2_ctx* c, int kind, int seg, const Sig& sig, const std::vector<std::function<int()>>& ops) {
  if (seg >= MAX_SEGMENTS) return fail(c, D2_ERR_RUNTIME, "internal: too many graph segments");
  Captured& g = c->graphs[kind][seg];
  if (!g.valid || !(g.sig == sig)) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = Captured{};
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = D2_OK;
    for (auto& op : ops)
      if ((rc = op())) break;
    c->capturing = false;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(c, D2_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(c, D2_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    g.valid = true, g.sig = sig;
    ++c->graph_captures;
  }
  HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
  ++c->graph_launches;
  return D2_OK;
}

// Graph runs need every phase's level count observed by an eager run first.
bool graph_ready(const d2_ctx* c, int kind, int filter) {
  PhaseSpec ph[2];
  const int n = phases_of(kind, filter, ph);
  for (int i = 0; i < n; ++i)
    if (c->phase_levels[ph[i].slot] < 0) return false;
  return true;
}

int run_graph(d2_ctx* c, int kind, int filter, float dt) {
  PhaseSpec ph[2];
  const int n = phases_of(kind, filter, ph);
  const int budget = level_budget(c);
  Sig sig;
  sig.n_pairs = c->n_pairs, sig.n_bodies = c->n_bodies, sig.n_clusters = kind == RUN_STEP ? c->n_clusters : 0;
  sig.max_depth = c->pair_depth, sig.precision = c->precision, sig.filter = filter, sig.dt = dt;
  sig.tune = c->tune;
  sig.learned = (c->learned_on && c->mlp_ready) ? c->gate_phases : 0, sig.gate_t = c->gate_t, sig.gate_s = c->gate_s;
  sig.cl_dev = kind == RUN_STEP && c->clusters_on_device ? 1 : 0;
  for (int i = 0; i < n; ++i) sig.levels[i] = std::min(c->phase_levels[ph[i].slot], budget);
  for (int i = 0; i < n; ++i)
    sig.grids = (sig.grids * 31 + grids_hash(c, ph[i].discrete, sig.levels[i])) * 31 + post_grid(c, ph[i].post_mode);
  std::vector<std::function<int()>> ops;
  int seg = 0;
  for (int i = 0; i < n; ++i) {
    const PhaseSpec p = ph[i];
    const int L = sig.levels[i];
    ops.push_back([c, kind, p, dt] {
      enqueue_begin(c, kind, p, dt);
      enqueue_gate(c, p);
      return D2_OK;
    });
    ops.push_back([c, p, L] { enqueue_levels(c, p, 0, L); return D2_OK; });
    ops.push_back([c, kind, p, L, dt] { return enqueue_end(c, kind, p, L, dt); });
    c->graph_levels[i] = L;
  }
  RC(launch_segment(c, kind, seg, sig, ops));
  return D2_OK;
}

// After a graph run: whether a phase's traversal went on past the levels in
// the graph (its next level is not empty).  The graph runs every phase to the
// deepest level count seen so far without stopping for the host; when a run
// needs more, it is repeated eagerly (the inputs are unchanged by a run) and
// the next capture holds more levels.
bool graph_cut_short(const d2_ctx* c, int kind, int filter) {
  PhaseSpec ph[2];
  const int n = phases_of(kind, filter, ph);
  const int half = (MAX_LEVELS + 1) * LV_STRIDE;
  for (int i = 0; i < n; ++i) {
    const int L = c->graph_levels[i];
    if (L >= level_budget(c)) continue;  // the budget check covers it
    // a step's CCD phase: its sizes as saved by k_fold_final
    const unsigned long long* t = c->host_lvl + (kind == RUN_STEP && !ph[i].discrete ? half : 0) + L * LV_STRIDE;
    if (t[LV_SOUP] || t[LV_CONN]) return true;
  }
  return false;
}

int ensure_clusters(d2_ctx* c, int n_clusters) {
  RC(ensure(c, c->step_bits, n_clusters)); RC(ensure(c, c->first_any, n_clusters)); RC(ensure(c, c->first_ns, n_clusters));
  RC(ensure(c, c->cluster_step, n_clusters)); RC(ensure(c, c->body_step, std::max(c->n_bodies, 1)));
  RC(ensure(c, c->read_later, std::max(c->n_bodies, 1)));
  return D2_OK;
}

// Waits for the run, reads the counters, checks capacity.
int finish_run(d2_ctx* c, int kind, int filter) {
  HIPCHK(c, hipMemcpyAsync(c->host_cnt, c->cnt.p, sizeof(unsigned long long) * C_COUNT, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->host_lvl, c->lvl.p, sizeof(unsigned long long) * 2 * (MAX_LEVELS + 1) * LV_STRIDE,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipGetLastError());
  if (c->profiling) collect_timers(c);
  const int budget = level_budget(c);
  const unsigned long long* last = c->host_lvl + budget * LV_STRIDE;
  if (c->host_cnt[C_OVERFLOW] || (long long)c->host_cnt[C_CONTACTS] > c->cap_contacts)
    return fail(c, D2_ERR_CAPACITY, "work list capacity exceeded; raise it with d2_set_capacity");
  if (last[LV_SOUP] || last[LV_CONN])
    return fail(c, D2_ERR_RUNTIME, "traversal deeper than the level budget of the uploaded trees");
  {
    const int half = (MAX_LEVELS + 1) * LV_STRIDE;
    auto keep = [&](int ph, const unsigned long long* src) {
      for (int l = 0; l <= MAX_LEVELS; ++l)
        for (int k = 0; k < LV_STRIDE; ++k) c->prev_lv[ph][l][k] = (long long)src[l * LV_STRIDE + k];
      c->prev_ok[ph] = true;
    };
    if (kind == RUN_CCD) keep(1, c->host_lvl);
    if (kind == RUN_NARROW) keep(0, c->host_lvl);
    if (kind == RUN_STEP) keep(0, c->host_lvl), keep(1, c->host_lvl + half);
    if (kind == RUN_CCD) c->prev_contacts[1] = (long long)c->host_cnt[C_CONTACTS];
    else c->prev_contacts[0] = (long long)c->host_cnt[C_CONTACTS];
    if (kind == RUN_STEP) c->prev_contacts[1] = (long long)c->host_lvl[half + LV_STRIDE - 1];
  }
  c->n_out = (long long)c->host_cnt[C_OUT];
  c->last_kind = kind;
  ++c->run_serial;
  c->last_mode = kind == RUN_CCD ? 1 : 0;
  c->ccd_valid = kind != RUN_NARROW;
  c->out_filtered = kind != RUN_CCD && filter != 0;
  c->last_filter = filter ? 1 : 0;
  return D2_OK;
}

// Capacities large enough for the sizes the failed run recorded (exact for
// the levels it reached; a deeper level may need another round).  The
// discrete phase's sizes are in the first half of the level table, the CCD
// phase's in the second; the overflow counter says which phase overflowed.
void grow_capacities(d2_ctx* c, int kind) {
  const int half = (MAX_LEVELS + 1) * LV_STRIDE;
  long long np = c->n_pairs, items = 0, res = 0, contacts = 0;
  auto scan = [&](const unsigned long long* t, bool ccd) {
    for (int l = 0; l <= MAX_LEVELS; ++l) {
      const unsigned long long* v = t + l * LV_STRIDE;
      np = std::max<long long>(np, (long long)std::max(v[LV_SOUP], v[LV_CONN]));
      if (ccd) {
        res = std::max<long long>(res, (long long)v[LV_ITEMS]);
        items = std::max<long long>(items, (long long)std::max(v[LV_PT], v[LV_EE]));
        items = std::max<long long>(items, (long long)std::max(v[LV_SPT], v[LV_SEE]));  // FP32 form's soup survivors
      } else if (c->precision == D2_FP32) {  // FP64 discrete levels keep no item list (k_tri_full)
        items = std::max<long long>(items, (long long)std::max(v[LV_ITEMS], v[LV_DEF]));
      }
    }
  };
  if (kind != RUN_CCD) scan(c->host_lvl, false), contacts = (long long)c->host_cnt[C_CONTACTS];
  if (kind != RUN_NARROW) scan(c->host_lvl + half, true),
      contacts = std::max<long long>(contacts, (long long)c->host_lvl[half + LV_STRIDE - 1]);
  auto grown = [](long long cur, long long need) { return need > cur ? std::max(2 * cur, need + need / 4) : cur; };
  bool any = false;
  long long v;
  if ((v = grown(c->cap, np)) != c->cap) c->cap = v, any = true;
  if ((v = grown(c->cap_items, items)) != c->cap_items) c->cap_items = v, any = true;
  if ((v = grown(c->cap_res, res)) != c->cap_res) c->cap_res = v, any = true;
  if ((v = grown(c->cap_contacts, contacts)) != c->cap_contacts) c->cap_contacts = v, any = true;
  if (!any) {  // sizes beyond the recorded ones (an overflow cut a level short): double the phase's lists
    const unsigned long long ov = c->host_cnt[C_OVERFLOW];
    c->cap *= 2, c->cap_contacts *= 2;
    if ((ov & 2ull) && c->precision == D2_FP32) c->cap_items *= 2;
    if (ov & 1ull) c->cap_res *= 2, c->cap_items *= 2;
  }
}

int run_once(d2_ctx* c, int kind, int filter, float dt, int64_t* n_contacts_out);

// This device's copy of `scratch`, resolved once per device with the device
// current (the first context created on it) and kept until the process ends;
// nullptr when the runtime cannot resolve it.
unsigned char* noalloc_scratch(int device) {
  static std::mutex mu;
  static unsigned char* addr[256] = {};
  static bool tried[256] = {};
  if (device < 0 || device >= 256) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!tried[device]) {
    tried[device] = true;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) addr[device] = static_cast<unsigned char*>(p);
  }
  return addr[device];
}

// ---- bucket-level entry: scratch and one call
struct BucketScratch {
  int32_t* np;       // [cap_np][2] staged node pairs
  long long* base;   // [cap_np] first provisional slot of each node pair
  int* count;        // [cap_np] hits of each node pair
  long long* off;    // [cap_np] first output row of each node pair
  int32_t* used;     // [cap_np] cumulative hit counts
  ContactRec* hits;  // [cap_hits] provisional hit records
  int32_t* out_i;    // [cap_hits][8]
  double* out_r;     // [cap_hits][11]
  unsigned long long* counter;
  long long cap_np, cap_hits;
};

// Validates a bucket call against the resident trees and bodies.
int bucket_check(d2_ctx* c, int a, int b, int n, const int32_t* node_pairs, const char* who) {
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, std::string(who) + ": upload trees and bodies first");
  if (a < 0 || a >= c->n_bodies || b < 0 || b >= c->n_bodies || n < 0 || (n > 0 && !node_pairs))
    return fail(c, D2_ERR_ARG, std::string(who) + ": bad arguments");
  const int ta = c->host_body_tree[a], tb = c->host_body_tree[b];
  const int na = c->host_tree_off[ta + 1] - c->host_tree_off[ta], nb = c->host_tree_off[tb + 1] - c->host_tree_off[tb];
  for (int k = 0; k < n; ++k)
    if (node_pairs[2 * k] < 0 || node_pairs[2 * k] >= na || node_pairs[2 * k + 1] < 0 || node_pairs[2 * k + 1] >= nb)
      return fail(c, D2_ERR_ARG, std::string(who) + ": node index out of range");
  return D2_OK;
}

// One bucket call on the given scratch.  *total receives the hit count; the
// cumulative counts are always written, the hits when they fit both the
// scratch and max_hits.
int bucket_run(d2_ctx* c, const BucketScratch& S, int a, int b, int n, const int32_t* node_pairs,
               int32_t* pair_used_out, int64_t max_hits, int32_t* ints_out, double* reals_out, long long* total) {
  *total = 0;
  if (n == 0) return D2_OK;
  HIPCHK(c, hipMemcpyAsync(S.np, node_pairs, sizeof(int32_t) * 2 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, S.counter, 1ll, 0ull);
  BucketArgs A;
  A.np_list = S.np, A.n = n, A.nodes = c->nodes.p, A.buckets = c->buckets.p;
  A.off_a = c->host_tree_off[c->host_body_tree[a]], A.off_b = c->host_tree_off[c->host_body_tree[b]];
  A.body_a = c->bodies.p + 40 * (size_t)a, A.body_b = c->bodies.p + 40 * (size_t)b;
  A.hits = S.hits, A.cap = S.cap_hits, A.counter = S.counter, A.base = S.base, A.count = S.count;
  const int blocks = (int)std::min<long long>((n + 3) / 4, (long long)c->cus * 8);
  hipLaunchKernelGGL(k_bucket_pairs, dim3(blocks), dim3(256), 0, c->stream, A);
  hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, c->stream, n, (const int*)S.count, S.off, S.used);
  HIPCHK(c, hipGetLastError());
  unsigned long long hits = 0;
  HIPCHK(c, hipMemcpyAsync(&hits, S.counter, sizeof(hits), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(pair_used_out, S.used, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *total = (long long)hits;
  if ((long long)hits > S.cap_hits || (long long)hits > max_hits || hits == 0) return D2_OK;
  hipLaunchKernelGGL(k_bucket_place, dim3(blocks), dim3(256), 0, c->stream, n, (const long long*)S.base,
                     (const int*)S.count, (const long long*)S.off, (const ContactRec*)S.hits, S.out_i, S.out_r);
  HIPCHK(c, hipGetLastError());
  if (ints_out)
    HIPCHK(c, hipMemcpyAsync(ints_out, S.out_i, sizeof(int32_t) * 8 * hits, hipMemcpyDeviceToHost, c->stream));
  if (reals_out)
    HIPCHK(c, hipMemcpyAsync(reals_out, S.out_r, sizeof(double) * 11 * hits, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int run(d2_ctx* c, int kind, int filter, float dt, int64_t* n_contacts_out) {
  int rc = run_once(c, kind, filter, dt, n_contacts_out);
  for (int attempt = 0; rc == D2_ERR_CAPACITY && c->auto_grow && attempt < 8; ++attempt) {
    grow_capacities(c, kind);
    rc = run_once(c, kind, filter, dt, n_contacts_out);
  }
  return rc;
}

int run_once(d2_ctx* c, int kind, int filter, float dt, int64_t* n_contacts_out) {
  if (!c->pairs_valid || c->n_bodies <= 0 || c->n_trees <= 0)
    return fail(c, D2_ERR_STATE, "upload trees, bodies and pairs first");
  if (c->pair_depth + 1 > MAX_LEVELS)
    return fail(c, D2_ERR_ARG, "a pair's trees are too deep: depth(a) + depth(b) must stay below 16");
  if (kind == RUN_STEP && c->n_clusters <= 0) return fail(c, D2_ERR_STATE, "upload the pairs' clusters first");
  (void)hipSetDevice(c->device);
  if (c->n_pairs > c->cap) return fail(c, D2_ERR_CAPACITY, "more pairs than the work-list capacity");
  RC(prepare(c));
  if (kind == RUN_STEP) RC(ensure_clusters(c, c->n_clusters));
  c->ccd_valid = false;
  c->last_mode = -1;
  // accounting runs are eager: their extra kernels sit between the levels
  const bool graph = c->graph_on && !c->touch_on && graph_ready(c, kind, filter);
  if (c->touch_on) RC(touch_begin(c));
  if (graph) RC(run_graph(c, kind, filter, dt));
  else RC(run_eager(c, kind, filter, dt));
  RC(finish_run(c, kind, filter));
  if (c->touch_on) RC(touch_end(c));
  if (graph && graph_cut_short(c, kind, filter)) {
    RC(run_eager(c, kind, filter, dt));
    RC(finish_run(c, kind, filter));
  }
  if (n_contacts_out) *n_contacts_out = c->n_out;
  return D2_OK;
}

// ---- the solver's warm store
constexpr size_t WARM_MIN = 1u << 20;
WarmTable warm_table(d2_ctx* c) {
  return WarmTable{c->warm_keys.p, c->warm_vals.p, (unsigned long long)c->warm_keys.n - 1, c->warm_used.p};
}
// Room for `extra` more keys at a load factor of at most one half: the table
// doubles (re-inserting its keys) until it fits.  Synchronises the stream.
int warm_reserve(d2_ctx* c, long long extra) {
  if (!c->warm_keys.p) {
    RC(ensure(c, c->warm_keys, WARM_MIN)); RC(ensure(c, c->warm_vals, WARM_MIN)); RC(ensure(c, c->warm_used, 1));
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(c, WARM_MIN)), dim3(256), 0, c->stream, c->warm_keys.p,
                       (long long)WARM_MIN, WARM_EMPTY);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->warm_used.p, 1ll, 0ull);
  }
  unsigned long long used = 0;
  HIPCHK(c, hipMemcpyAsync(&used, c->warm_used.p, sizeof(used), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  size_t cap = c->warm_keys.n;
  while ((long long)used + extra > (long long)(cap / 2)) cap *= 2;
  if (cap == c->warm_keys.n) return D2_OK;
  DBuf<unsigned long long> old_k = c->warm_keys;
  DBuf<double> old_v = c->warm_vals;
  c->warm_keys = DBuf<unsigned long long>{}, c->warm_vals = DBuf<double>{};
  RC(ensure(c, c->warm_keys, cap)); RC(ensure(c, c->warm_vals, cap));
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(c, (long long)cap)), dim3(256), 0, c->stream, c->warm_keys.p,
                     (long long)cap, WARM_EMPTY);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->warm_used.p, 1ll, 0ull);
  hipLaunchKernelGGL(k_warm_rehash, dim3(grid_for(c, (long long)old_k.n)), dim3(256), 0, c->stream,
                     (const unsigned long long*)old_k.p, (const double*)old_v.p, (long long)old_k.n, warm_table(c));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  release(old_k), release(old_v);
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
  c->tune = d2_tune{4 * c->cus, 8 * c->cus, 16 * c->cus, 3};
  if (hipHostMalloc((void**)&c->host_cnt, sizeof(unsigned long long) * C_COUNT, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->host_lvl, sizeof(unsigned long long) * 2 * (MAX_LEVELS + 1) * LV_STRIDE,
                    hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->host_red, sizeof(int) * (BR_N + 2), hipHostMallocDefault) != hipSuccess) {
    if (c->host_cnt) (void)hipHostFree(c->host_cnt);
    if (c->host_lvl) (void)hipHostFree(c->host_lvl);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return D2_ERR_HIP;
  }
  std::memset(c->host_lvl, 0, sizeof(unsigned long long) * 2 * (MAX_LEVELS + 1) * LV_STRIDE);
  c->noalloc = noalloc_scratch(device);
  *out = c;
  return D2_OK;
}

int d2_ctx_destroy(d2_ctx* c) {
  if (!c) return D2_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  collect_timers(c);
  drop_graphs(c);
  release(c->nodes), release(c->buckets), release(c->fbuckets), release(c->stris), release(c->tree_off), release(c->body_tree), release(c->bodies);
  release(c->pairs), release(c->pair_info), release(c->cluster_of_pair), release(c->max_time), release(c->pair_rec);
  for (int k = 0; k < 2; ++k) release(c->soup_list[k]), release(c->conn_list[k]);
  release(c->items), release(c->ditems), release(c->ditems_def), release(c->pt_items), release(c->ee_items), release(c->res), release(c->res_mask), release(c->np_hit), release(c->np_hit2), release(c->exp_desc);
  release(c->contacts), release(c->cnt), release(c->lvl), release(c->masks), release(c->conn_mask);
  release(c->cnt_a), release(c->cnt_s), release(c->cnt_c), release(c->cnt_pt), release(c->cnt_ee), release(c->parts);
  release(c->pass_min_s), release(c->pass_min_c), release(c->pass_flag_s), release(c->pass_flag_c);
  release(c->min_time), release(c->scale);
  release(c->pair_cnt), release(c->pair_off), release(c->pair_fill), release(c->perm), release(c->keep),
      release(c->kept_cnt), release(c->out_off), release(c->out_i), release(c->out_r), release(c->scan_part);
  release(c->big), release(c->huge), release(c->n_big), release(c->hscratch);
  release(c->nbound), release(c->kid_inner);
  release(c->sv_mass), release(c->sv_fric), release(c->sv_ccount), release(c->sv_chits), release(c->sv_poff),
      release(c->sv_hoff), release(c->sv_cfill), release(c->sv_cpairs), release(c->sv_inpair), release(c->sv_loc),
      release(c->sv_status), release(c->sv_flag), release(c->sv_hl), release(c->sv_ord), release(c->sv_hrow),
      release(c->sv_hpc), release(c->sv_pbody), release(c->sv_hA), release(c->sv_hB), release(c->sv_heps),
      release(c->sv_hkey), release(c->sv_hs), release(c->sv_ps), release(c->sv_out), release(c->sv_out_csr),
      release(c->warm_keys), release(c->warm_used), release(c->warm_vals);
  release(c->step_bits), release(c->first_any), release(c->first_ns), release(c->cluster_step), release(c->body_step),
      release(c->tmp_cluster), release(c->read_later);
  release(c->touch_base), release(c->touch_lv), release(c->touch_run), release(c->touch_lvb), release(c->touch_runb),
      release(c->touch_acc);
  release(c->bk_np), release(c->bk_used), release(c->bk_out_i), release(c->bk_base), release(c->bk_off),
      release(c->bk_count), release(c->bk_hits), release(c->bk_out_r), release(c->bk_counter);
  release(c->tree_radius), release(c->tree_eps_out), release(c->hv), release(c->feat), release(c->pred), release(c->hv_off);
  release(c->cert), release(c->mlp_w), release(c->mlp_m), release(c->mlp_v), release(c->mlp_wf), release(c->mlp_grad);
  release(c->mlp_wb), release(c->tr_x), release(c->tr_t), release(c->tr_act), release(c->tr_dlt), release(c->tr_loss);
  release(c->tr_idx);
  release(c->tb_box), release(c->tb_rad), release(c->last_state), release(c->bp_R), release(c->tree_depth_d),
      release(c->group);
  release(c->bp_statics), release(c->bp_item_body), release(c->bp_item_seg), release(c->bp_item_cnt),
      release(c->bp_item_sorted), release(c->bp_item_off), release(c->bp_meta), release(c->bp_slot), release(c->bp_rank),
      release(c->bp_counts), release(c->bp_start), release(c->bp_members), release(c->bp_red);
  release(c->bp_item_pos), release(c->bp_keys), release(c->bp_fill), release(c->bp_box), release(c->bp_chunk);
  release(c->cc_parent), release(c->cc_label), release(c->cc_root), release(c->cc_root_off), release(c->cc_of_body),
      release(c->cc_first), release(c->cc_tmin), release(c->cc_smin);
  if (c->host_red) (void)hipHostFree(c->host_red);
  if (c->host_cnt) (void)hipHostFree(c->host_cnt);
  if (c->host_lvl) (void)hipHostFree(c->host_lvl);
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
  drop_graphs(c);
  c->cap = entries;
  c->cap_items = entries;
  c->cap_contacts = entries;
  c->cap_res = std::min<long long>(entries, 1ll << 20);
  c->auto_grow = false;
  return D2_OK;
}

int d2_set_auto_grow(d2_ctx* c, int on) {
  if (!c) return D2_ERR_ARG;
  c->auto_grow = on != 0;
  return D2_OK;
}

int d2_set_tune(d2_ctx* c, const d2_tune* t) {
  if (!c || !t) return D2_ERR_ARG;
  if (t->chunk_blocks < 1 || t->chunk_blocks > 4096 || t->stream_blocks < 1 || t->stream_blocks > 16384 ||
      t->post_blocks < 1 || t->post_blocks > 16384 || t->gate_phases < 0 || t->gate_phases > 3)
    return fail(c, D2_ERR_ARG, "d2_set_tune: value out of range");
  c->tune = *t;
  if (c->learned_on) c->gate_phases = t->gate_phases;
  return D2_OK;
}
int d2_get_tune(const d2_ctx* c, d2_tune* t) {
  if (!c || !t) return D2_ERR_ARG;
  *t = c->tune;
  t->gate_phases = c->gate_phases;
  return D2_OK;
}

// Coordinate descent over the launch shapes, timing whole runs of the given
// kind (graph or eager, as the context is set) on the resident inputs.
int d2_autotune(d2_ctx* c, int kind, float dt, int filter, int reps, d2_tune* best_out, double* best_ms_out) {
  if (!c || kind < 0 || kind >= RUN_KINDS || reps < 1) return D2_ERR_ARG;
  d2_tune start = c->tune;
  start.gate_phases = c->gate_phases;
  auto time_it = [&](const d2_tune& t, double* ms) -> int {
    c->tune = t;
    if (c->learned_on) c->gate_phases = t.gate_phases;
    RC(run(c, kind, filter, dt, nullptr));  // warm-up (and capture)
    std::vector<double> v;
    for (int r = 0; r < reps; ++r) {
      auto t0 = std::chrono::steady_clock::now();
      RC(run(c, kind, filter, dt, nullptr));
      v.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(v.begin(), v.end());
    *ms = v[v.size() / 2];
    return D2_OK;
  };
  const int u = c->cus;
  const std::vector<int> chunk = {u, 2 * u, 4 * u, 8 * u, 16 * u};
  const std::vector<int> stream = {2 * u, 4 * u, 8 * u, 16 * u, 32 * u};
  const std::vector<int> post = {4 * u, 16 * u, 64 * u};
  d2_tune best = start;
  double best_ms = 0.0;
  int rc = time_it(best, &best_ms);
  const std::vector<int> gate = {0, 1, 2, 3};  // 0: the gate gated in no phase (it may not pay for itself)
  const int dims = c->learned_on ? 4 : 3;
  for (int sweep = 0; sweep < 2 && !rc; ++sweep)
    for (int dim = 0; dim < dims && !rc; ++dim) {
      const std::vector<int>& vals = dim == 0 ? chunk : dim == 1 ? stream : dim == 2 ? post : gate;
      for (int v : vals) {
        d2_tune t = best;
        int* f = dim == 0 ? &t.chunk_blocks : dim == 1 ? &t.stream_blocks : dim == 2 ? &t.post_blocks : &t.gate_phases;
        if (*f == v || v > (dim == 0 ? 4096 : 16384)) continue;
        *f = v;
        double ms = 0.0;
        if ((rc = time_it(t, &ms))) break;
        if (ms < best_ms) best_ms = ms, best = t;
      }
    }
  c->tune = rc ? start : best;
  if (c->learned_on) c->gate_phases = c->tune.gate_phases;
  if (rc) return rc;
  if (best_out) *best_out = best;
  if (best_ms_out) *best_ms_out = best_ms;
  return D2_OK;
}

int d2_set_graph(d2_ctx* c, int on) {
  if (!c) return D2_ERR_ARG;
  if (!on) drop_graphs(c);
  c->graph_on = on != 0;
  return D2_OK;
}

int d2_trees_upload(d2_ctx* c, int n_trees, const int32_t* tree_node_offset, const int32_t* nodes, const double* buckets) {
  if (!c) return D2_ERR_ARG;
  if (n_trees <= 0 || !tree_node_offset || !nodes || !buckets) return fail(c, D2_ERR_ARG, "d2_trees_upload: bad arguments");
  int total = tree_node_offset[n_trees];
  if (total <= 0 || tree_node_offset[0] != 0) return fail(c, D2_ERR_ARG, "d2_trees_upload: bad offsets");
  std::vector<int> depth(n_trees, 0);
  for (int t = 0; t < n_trees; ++t) {
    int b = tree_node_offset[t], e = tree_node_offset[t + 1];
    if (e <= b) return fail(c, D2_ERR_ARG, "d2_trees_upload: empty tree");
    for (int i = b; i < e; ++i) {
      const int32_t* r = nodes + 8 * (size_t)i;
      int cnt = e - b;
      if (r[1] < 0 || r[1] >= cnt || r[7] < 0 || r[7] > 8 || r[3] < 0 ||
          (r[2] && (r[5] < 0 || r[5] + r[6] > cnt || r[6] < 1 || r[6] > 8)))
        return fail(c, D2_ERR_ARG, "d2_trees_upload: inconsistent node record");
      depth[t] = std::max(depth[t], (int)r[3]);
    }
  }
  (void)hipSetDevice(c->device);
  c->n_trees = 0;  // no trees until every per-node table of this upload is in place
  RC(ensure(c, c->nodes, total)); RC(ensure(c, c->buckets, (size_t)total * BUCKET)); RC(ensure(c, c->tree_off, n_trees + 1));
  c->n_nodes_total = total, c->fb_ready = false;  // the float copy follows on the next FP32 run
  HIPCHK(c, hipMemcpyAsync(c->nodes.p, nodes, sizeof(NodeRec) * total, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->buckets.p, buckets, sizeof(double) * BUCKET * (size_t)total, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->tree_off.p, tree_node_offset, sizeof(int32_t) * (n_trees + 1), hipMemcpyHostToDevice, c->stream));
  // Per tree, for the learned gate's certificate: the distinct vertices of the
  // leaf triangles (body coordinates), the bounding radius, and how far beyond
  // the search distance a leaf contact can be accepted: the largest distance a
  // float-barycentric "inside" test can accept a point outside a triangle,
  // 64 eps_f kappa L (kappa = d00 d11 / (d00 d11 - d01^2), L the longest edge;
  // 4x the 16 eps_f kappa L rounding bound), plus the largest eps of the leaf
  // buckets (the discrete hit test adds them, comparison_dist.h:233).
  std::vector<float> hv, radius(n_trees), eps_out(n_trees);
  std::vector<int32_t> hv_off(n_trees + 1, 0);
  for (int t = 0; t < n_trees; ++t) {
    std::vector<std::array<double, 3>> vs;
    double rmax = 0.0, eo = 0.0, leaf_eps = 0.0;
    for (int i = tree_node_offset[t]; i < tree_node_offset[t + 1]; ++i) {
      const int32_t* r = nodes + 8 * (size_t)i;
      if (r[2]) continue;
      const double* bk = buckets + (size_t)BUCKET * (tree_node_offset[t] + r[1]);
      for (int k = 0; k < r[7]; ++k) {
        leaf_eps = std::max(leaf_eps, bk[72 + k]);
        std::array<double, 3> v[3];
        for (int q = 0; q < 3; ++q) v[q] = {bk[(3 * q) * 8 + k], bk[(3 * q + 1) * 8 + k], bk[(3 * q + 2) * 8 + k]};
        double e0[3], e1[3], e2[3];
        for (int d = 0; d < 3; ++d) e0[d] = v[1][d] - v[0][d], e1[d] = v[2][d] - v[0][d], e2[d] = v[2][d] - v[1][d];
        const double d00 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
        const double d11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
        const double d01 = e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2];
        const double d22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
        const double den = d00 * d11 - d01 * d01;
        const double kappa = den > 0.0 ? d00 * d11 / den : 1e30;
        eo = std::max(eo, 64.0 * 5.9604645e-08 * kappa * std::sqrt(std::max(d00, std::max(d11, d22))));
        for (int q = 0; q < 3; ++q) {
          vs.push_back(v[q]);
          rmax = std::max(rmax, std::sqrt(v[q][0] * v[q][0] + v[q][1] * v[q][1] + v[q][2] * v[q][2]));
        }
      }
    }
    std::sort(vs.begin(), vs.end());
    vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
    for (auto& v : vs) hv.push_back((float)v[0]), hv.push_back((float)v[1]), hv.push_back((float)v[2]);
    hv_off[t + 1] = hv_off[t] + (int32_t)vs.size();
    radius[t] = (float)rmax * (1.0f + 1e-6f);
    // rounded up so the float margin is never below the double one
    eps_out[t] = (float)std::min((eo + leaf_eps) * (1.0 + 1e-6), 1e30);
  }
  // Per node, for pruning (NB_STRIDE floats): the body-frame box of every
  // leaf-triangle vertex under the node, the largest (leaf eps +
  // float-barycentric acceptance bound 64 eps_f kappa L) of those triangles,
  // and a sphere around the same vertices (their centroid, the largest
  // distance from it), all rounded outward; children follow their parent, so
  // one pass from the back merges the boxes and slacks, and a second pass
  // per node over its leaves' vertices gives the spheres.
  std::vector<float> nb((size_t)total * NB_STRIDE);
  for (int t = 0; t < n_trees; ++t) {
    const int b0 = tree_node_offset[t];
    for (int i = tree_node_offset[t + 1] - 1; i >= b0; --i) {
      const int32_t* r = nodes + 8 * (size_t)i;
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, E = 0.0;
      if (!r[2]) {
        const double* bk = buckets + (size_t)BUCKET * (b0 + r[1]);
        for (int k = 0; k < r[7]; ++k) {
          double v[3][3];
          for (int q = 0; q < 3; ++q)
            for (int d = 0; d < 3; ++d) v[q][d] = bk[(3 * q + d) * 8 + k];
          E = std::max(E, leaf_slack(v[0], v[1], v[2], bk[72 + k]));
          for (int q = 0; q < 3; ++q)
            for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], v[q][d]), hi[d] = std::max(hi[d], v[q][d]);
        }
      } else {
        for (int ch = 0; ch < r[6]; ++ch) {
          const float* cb = nb.data() + NB_STRIDE * ((size_t)b0 + r[5] + ch);
          for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], (double)cb[d]), hi[d] = std::max(hi[d], (double)cb[3 + d]);
          E = std::max(E, (double)cb[6]);
        }
      }
      float* o = nb.data() + NB_STRIDE * (size_t)i;
      const bool empty = !(lo[0] <= hi[0]);
      for (int d = 0; d < 3; ++d) {  // an empty node never prunes
        float fl = empty ? -FLT_MAX : (float)lo[d], fh = empty ? FLT_MAX : (float)hi[d];
        if ((double)fl > lo[d]) fl = std::nextafter(fl, -INFINITY);
        if ((double)fh < hi[d]) fh = std::nextafter(fh, INFINITY);
        o[d] = fl, o[3 + d] = fh;
      }
      float fe = (float)std::min(E, 1e30);
      if ((double)fe < E) fe = std::nextafter(fe, INFINITY);
      o[6] = fe, o[7] = FLT_MAX, o[8] = o[9] = o[10] = o[11] = 0.0f;
    }
    // spheres: the leaves under a node are a contiguous run of the tree's
    // leaves in depth-first order, so each node's vertex set is walked once
    // per level of the tree (leaf vertices gathered bottom-up by child lists)
    std::vector<std::vector<int>> leaves_of(tree_node_offset[t + 1] - b0);
    for (int i = tree_node_offset[t + 1] - 1; i >= b0; --i) {
      const int32_t* r = nodes + 8 * (size_t)i;
      std::vector<int>& L = leaves_of[i - b0];
      if (!r[2]) L.push_back(i);
      else
        for (int ch = 0; ch < r[6]; ++ch) {
          std::vector<int>& C = leaves_of[r[5] + ch];
          L.insert(L.end(), C.begin(), C.end());
        }
      double cen[3] = {0.0, 0.0, 0.0};
      long long nv = 0;
      for (int leaf : L) {
        const int32_t* lr = nodes + 8 * (size_t)leaf;
        const double* bk = buckets + (size_t)BUCKET * (b0 + lr[1]);
        for (int k = 0; k < lr[7]; ++k)
          for (int q = 0; q < 3; ++q, ++nv)
            for (int d = 0; d < 3; ++d) cen[d] += bk[(3 * q + d) * 8 + k];
      }
      if (nv == 0) continue;  // no vertices: the sphere never prunes
      float fc[3];
      for (int d = 0; d < 3; ++d) fc[d] = (float)(cen[d] / (double)nv);
      double rad = 0.0;
      for (int leaf : L) {
        const int32_t* lr = nodes + 8 * (size_t)leaf;
        const double* bk = buckets + (size_t)BUCKET * (b0 + lr[1]);
        for (int k = 0; k < lr[7]; ++k)
          for (int q = 0; q < 3; ++q) {
            double s2 = 0.0;
            for (int d = 0; d < 3; ++d) {
              const double x = bk[(3 * q + d) * 8 + k] - (double)fc[d];
              s2 += x * x;
            }
            rad = std::max(rad, std::sqrt(s2));
          }
      }
      float* o = nb.data() + NB_STRIDE * (size_t)i;
      float fr = (float)(rad * (1.0 + 1e-12));
      if ((double)fr < rad) fr = std::nextafter(fr, INFINITY);
      o[7] = fr, o[8] = fc[0], o[9] = fc[1], o[10] = fc[2];
      if (!r[2]) continue;
      for (int ch = 0; ch < r[6]; ++ch) std::vector<int>().swap(leaves_of[r[5] + ch]);  // children are done
    }
  }
  RC(ensure(c, c->nbound, nb.size()));
  HIPCHK(c, hipMemcpyAsync(c->nbound.p, nb.data(), sizeof(float) * nb.size(), hipMemcpyHostToDevice, c->stream));
  // Per node, which of its (at most 8) children are inner nodes: the CCD
  // expansion routes a node pair's children to the soup or the connected
  // list by it instead of reading every child's record.
  std::vector<uint8_t> kid((size_t)total, 0);
  for (int t = 0; t < n_trees; ++t) {
    const int b0 = tree_node_offset[t];
    for (int i = b0; i < tree_node_offset[t + 1]; ++i) {
      const int32_t* r = nodes + 8 * (size_t)i;
      if (!r[2]) continue;
      unsigned m = 0;
      for (int ch = 0; ch < r[6]; ++ch)
        if (nodes[8 * ((size_t)b0 + r[5] + ch) + 2]) m |= 1u << ch;
      kid[i] = (uint8_t)m;
    }
  }
  RC(ensure(c, c->kid_inner, kid.size()));
  HIPCHK(c, hipMemcpyAsync(c->kid_inner.p, kid.data(), kid.size(), hipMemcpyHostToDevice, c->stream));
  // Per tree for the broad phase, as MeshData::initBoudingData (model/mesh.cpp:44-51)
  // computes it over the mesh vertices (here: the leaf triangles' vertices,
  // which are every vertex of a mesh whose vertices all belong to a face):
  // the float AABB and the largest vertex norm.
  std::vector<float> tbox(6 * (size_t)n_trees);
  std::vector<double> trad(n_trees, 0.0);
  for (int t = 0; t < n_trees; ++t) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = tree_node_offset[t]; i < tree_node_offset[t + 1]; ++i) {
      const int32_t* r = nodes + 8 * (size_t)i;
      if (r[2]) continue;
      const double* bk = buckets + (size_t)BUCKET * (tree_node_offset[t] + r[1]);
      for (int k = 0; k < r[7]; ++k)
        for (int q = 0; q < 3; ++q) {
          const double x = bk[(3 * q) * 8 + k], y = bk[(3 * q + 1) * 8 + k], z = bk[(3 * q + 2) * 8 + k];
          const float f[3] = {(float)x, (float)y, (float)z};
          for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], f[d]), hi[d] = std::max(hi[d], f[d]);
          trad[t] = std::max(trad[t], std::sqrt((x * x + y * y) + z * z));
        }
    }
    for (int d = 0; d < 3; ++d) tbox[6 * t + d] = lo[d], tbox[6 * t + 3 + d] = hi[d];
  }
  RC(ensure(c, c->tb_box, 6 * (size_t)n_trees)); RC(ensure(c, c->tb_rad, n_trees)); RC(ensure(c, c->tree_depth_d, n_trees));
  HIPCHK(c, hipMemcpyAsync(c->tb_box.p, tbox.data(), sizeof(float) * tbox.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->tb_rad.p, trad.data(), sizeof(double) * n_trees, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->tree_depth_d.p, depth.data(), sizeof(int32_t) * n_trees, hipMemcpyHostToDevice, c->stream));
  RC(ensure(c, c->hv, std::max<size_t>(hv.size(), 3))); RC(ensure(c, c->hv_off, n_trees + 1));
  RC(ensure(c, c->tree_radius, n_trees)); RC(ensure(c, c->tree_eps_out, n_trees));
  HIPCHK(c, hipMemcpyAsync(c->hv.p, hv.data(), sizeof(float) * hv.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->hv_off.p, hv_off.data(), sizeof(int32_t) * (n_trees + 1), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->tree_radius.p, radius.data(), sizeof(float) * n_trees, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->tree_eps_out.p, eps_out.data(), sizeof(float) * n_trees, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_trees = n_trees;
  c->tree_depth = depth;
  c->host_tree_off.assign(tree_node_offset, tree_node_offset + n_trees + 1);
  c->pair_depth = std::min(MAX_LEVELS - 1, 2 * *std::max_element(depth.begin(), depth.end()));  // until pairs come
  // body records name trees by index: bodies, pairs and clusters must come again
  c->host_body_tree.clear();
  c->n_bodies = 0;
  c->has_groups = false;
  invalidate_pairs(c);
  return D2_OK;
}

int d2_bodies_upload(d2_ctx* c, int n_bodies, const int32_t* tree_of_body, const double* records) {
  if (!c) return D2_ERR_ARG;
  if (n_bodies <= 0 || !tree_of_body || !records) return fail(c, D2_ERR_ARG, "d2_bodies_upload: bad arguments");
  if (c->n_trees <= 0) return fail(c, D2_ERR_STATE, "d2_bodies_upload: upload trees first");
  for (int i = 0; i < n_bodies; ++i)
    if (tree_of_body[i] < 0 || tree_of_body[i] >= c->n_trees) return fail(c, D2_ERR_ARG, "d2_bodies_upload: bad tree index");
  (void)hipSetDevice(c->device);
  RC(ensure(c, c->bodies, (size_t)n_bodies * 40)); RC(ensure(c, c->body_tree, n_bodies));
  HIPCHK(c, hipMemcpyAsync(c->bodies.p, records, sizeof(double) * 40 * (size_t)n_bodies, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->body_tree.p, tree_of_body, sizeof(int32_t) * n_bodies, hipMemcpyHostToDevice, c->stream));
  // last states start equal to the current ones (d2_bodies_last_upload sets them)
  std::vector<double> last(14 * (size_t)n_bodies);
  for (int i = 0; i < n_bodies; ++i) std::memcpy(&last[14 * (size_t)i], records + 40 * (size_t)i, 14 * sizeof(double));
  RC(ensure(c, c->last_state, 14 * (size_t)n_bodies));
  HIPCHK(c, hipMemcpyAsync(c->last_state.p, last.data(), sizeof(double) * last.size(), hipMemcpyHostToDevice, c->stream));
  // broad-phase work items: one per non-static body, one per BP_SEG later
  // bodies for a static one (it sweeps all of them)
  c->host_static.clear();
  std::vector<int> ib, is;
  for (int i = 0; i < n_bodies; ++i) {
    if (records[40 * (size_t)i + 38] != 0.0) {
      c->host_static.push_back(i);
      const int segs = std::max(1, (n_bodies - 1 - i + BP_SEG - 1) / BP_SEG);
      for (int q = 0; q < segs; ++q) ib.push_back(i), is.push_back(q);
    } else {
      ib.push_back(i), is.push_back(0);
    }
  }
  c->n_items = (int)ib.size();
  RC(ensure(c, c->bp_item_body, ib.size())); RC(ensure(c, c->bp_item_seg, is.size()));
  RC(ensure(c, c->bp_statics, std::max<size_t>(c->host_static.size(), 1)));
  HIPCHK(c, hipMemcpyAsync(c->bp_item_body.p, ib.data(), sizeof(int) * ib.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->bp_item_seg.p, is.data(), sizeof(int) * is.size(), hipMemcpyHostToDevice, c->stream));
  if (!c->host_static.empty())
    HIPCHK(c, hipMemcpyAsync(c->bp_statics.p, c->host_static.data(), sizeof(int) * c->host_static.size(),
                             hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n_bodies != c->n_bodies) {
    // pairs, clusters and scene groups index bodies: a new count voids them
    invalidate_pairs(c);
    c->has_groups = false;
  } else if (c->pairs_valid && !std::equal(tree_of_body, tree_of_body + n_bodies, c->host_body_tree.begin())) {
    // same bodies, other trees: the kept pairs' level budget becomes the
    // deepest any pair of these trees can need
    c->pair_depth = std::min(MAX_LEVELS - 1, 2 * *std::max_element(c->tree_depth.begin(), c->tree_depth.end()));
  }
  c->n_bodies = n_bodies;
  c->host_body_tree.assign(tree_of_body, tree_of_body + n_bodies);
  return D2_OK;
}

int d2_bodies_download(d2_ctx* c, double* records) {
  if (!c || !records) return D2_ERR_ARG;
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "no bodies uploaded");
  HIPCHK(c, hipMemcpyAsync(records, c->bodies.p, sizeof(double) * 40 * (size_t)c->n_bodies, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_solver_bodies(d2_ctx* c, int n_bodies, const double* mass, const double* friction) {
  if (!c) return D2_ERR_ARG;
  if (c->n_bodies <= 0 || n_bodies != c->n_bodies) return fail(c, D2_ERR_ARG, "d2_solver_bodies: one value per uploaded body");
  if (!mass || !friction) return fail(c, D2_ERR_ARG, "d2_solver_bodies: bad arguments");
  for (int i = 0; i < n_bodies; ++i)
    if (!(mass[i] > 0.0) || !(friction[i] >= 0.0))
      return fail(c, D2_ERR_ARG, "d2_solver_bodies: masses must be > 0 and friction coefficients >= 0");
  (void)hipSetDevice(c->device);
  RC(ensure(c, c->sv_mass, n_bodies)); RC(ensure(c, c->sv_fric, n_bodies));
  HIPCHK(c, hipMemcpyAsync(c->sv_mass.p, mass, sizeof(double) * n_bodies, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->sv_fric.p, friction, sizeof(double) * n_bodies, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_solve(d2_ctx* c, const double* gravity, int32_t* cluster_status, int* n_failed) {
  if (!c) return D2_ERR_ARG;
  if (!gravity) return fail(c, D2_ERR_ARG, "d2_solve: gravity is NULL");
  if (c->last_mode != 0 || c->last_kind != RUN_STEP || !c->out_filtered)
    return fail(c, D2_ERR_STATE, "d2_solve needs a preceding d2_step_run with the contact filter");
  if (c->sv_mass.n < (size_t)c->n_bodies || !c->sv_mass.p)
    return fail(c, D2_ERR_STATE, "d2_solve: set the bodies' masses and friction first (d2_solver_bodies)");
  (void)hipSetDevice(c->device);
  const int np = c->n_pairs, nk = c->n_clusters, nb = c->n_bodies;
  const long long nh = c->n_out;
  RC(warm_reserve(c, nh));
  RC(ensure(c, c->sv_ccount, nk + 1)); RC(ensure(c, c->sv_chits, nk + 1)); RC(ensure(c, c->sv_poff, nk + 1));
  RC(ensure(c, c->sv_hoff, nk + 1)); RC(ensure(c, c->sv_cfill, nk)); RC(ensure(c, c->sv_status, nk));
  RC(ensure(c, c->sv_cpairs, std::max(np, 1))); RC(ensure(c, c->sv_pbody, 2 * (size_t)std::max(np, 1)));
  RC(ensure(c, c->sv_inpair, nb)); RC(ensure(c, c->sv_loc, nb)); RC(ensure(c, c->sv_flag, 1));
  RC(ensure(c, c->out_off, (size_t)np + 1));
  const size_t H = (size_t)std::max<long long>(nh, 1);
  RC(ensure(c, c->sv_hl, 2 * H)); RC(ensure(c, c->sv_ord, H)); RC(ensure(c, c->sv_hrow, H)); RC(ensure(c, c->sv_hpc, H));
  RC(ensure(c, c->sv_hA, 3 * H)); RC(ensure(c, c->sv_hB, 3 * H)); RC(ensure(c, c->sv_heps, H)); RC(ensure(c, c->sv_hkey, H));
  RC(ensure(c, c->sv_hs, (size_t)SH * H)); RC(ensure(c, c->sv_out, 11 * H));
  RC(ensure(c, c->sv_ps, (size_t)SP * 2 * std::max(np, 1)));
  SolveArgs A;
  A.n_pairs = np, A.n_clusters = nk, A.n_bodies = nb;
  A.pairs = c->pairs.p, A.cluster_of_pair = c->cluster_of_pair.p, A.kept_cnt = c->kept_cnt.p;
  A.pair_off = c->pair_off.p, A.out_off = c->out_off.p, A.perm = c->perm.p, A.keep = c->keep.p;
  A.contacts = c->contacts.p, A.cluster_step = c->cluster_step.p, A.body_step = c->body_step.p;
  A.bodies = c->bodies.p, A.mass = c->sv_mass.p, A.fric = c->sv_fric.p;
  A.gx = gravity[0], A.gy = gravity[1], A.gz = gravity[2];
  A.ccount = c->sv_ccount.p, A.chits = c->sv_chits.p, A.cfill = c->sv_cfill.p, A.cpairs = c->sv_cpairs.p;
  A.inpair = c->sv_inpair.p, A.loc = c->sv_loc.p, A.status = c->sv_status.p;
  A.hl = c->sv_hl.p, A.ord = c->sv_ord.p, A.hrow = c->sv_hrow.p, A.hpc = c->sv_hpc.p, A.pbody = c->sv_pbody.p;
  A.hA = c->sv_hA.p, A.hB = c->sv_hB.p, A.heps = c->sv_heps.p, A.hkey = c->sv_hkey.p, A.hs = c->sv_hs.p;
  A.ps = c->sv_ps.p, A.out = c->sv_out.p, A.warm = warm_table(c), A.flag = c->sv_flag.p;
  auto fill = [&](int* p, long long n, int v) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c, n)), dim3(256), 0, c->stream, p, n, v);
  };
  fill(A.ccount, nk + 1, 0), fill(A.chits, nk + 1, 0), fill(A.cfill, nk, 0);
  fill(A.inpair, nb, 0), fill(A.loc, nb, -1), fill(c->sv_flag.p, 1, 0);
  hipLaunchKernelGGL(k_sv_count, dim3(grid_for(c, np)), dim3(256), 0, c->stream, A);
  scan_host_n(c, c->sv_ccount.p, c->sv_poff.p, nk, nullptr);
  scan_host_n(c, c->sv_chits.p, c->sv_hoff.p, nk, nullptr);
  scan_host_n(c, c->kept_cnt.p, c->out_off.p, np, nullptr);
  SolveArgs B = A;
  B.ccount = c->sv_poff.p, B.chits = c->sv_hoff.p;  // the offsets from here on
  hipLaunchKernelGGL(k_sv_place, dim3(grid_for(c, np)), dim3(256), 0, c->stream, B);
  int bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, c->sv_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (bad) return fail(c, D2_ERR_ARG, "d2_solve: a dynamic body's pairs lie in more than one cluster");
  hipLaunchKernelGGL(k_sv_solve, dim3((nk + 63) / 64 > 0 ? (nk + 63) / 64 : 1), dim3(64), 0, c->stream, B);
  hipLaunchKernelGGL(k_sv_free, dim3(grid_for(c, nb)), dim3(256), 0, c->stream, B);
  HIPCHK(c, hipGetLastError());
  std::vector<int32_t> st(std::max(nk, 1));
  if (nk > 0) HIPCHK(c, hipMemcpyAsync(st.data(), c->sv_status.p, sizeof(int32_t) * nk, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int failed = 0;
  for (int k = 0; k < nk; ++k) failed += st[k] == 0;
  if (cluster_status && nk > 0) std::memcpy(cluster_status, st.data(), sizeof(int32_t) * nk);
  if (n_failed) *n_failed = failed;
  c->solved_serial = c->run_serial;
  c->solved_hits = nh;
  return D2_OK;
}

int d2_solve_results(d2_ctx* c, double* per_contact) {
  if (!c || !per_contact) return D2_ERR_ARG;
  if (c->last_mode != 0 || c->solved_serial != c->run_serial)
    return fail(c, D2_ERR_STATE, "d2_solve_results: no solve of the last step run");
  const long long nh = c->solved_hits;
  if (nh == 0) return D2_OK;
  (void)hipSetDevice(c->device);
  RC(ensure(c, c->sv_out_csr, 11 * (size_t)nh));
  hipLaunchKernelGGL(k_sv_scatter, dim3(grid_for(c, nh)), dim3(256), 0, c->stream, nh, (const int*)c->sv_hrow.p,
                     (const double*)c->sv_out.p, c->sv_out_csr.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(per_contact, c->sv_out_csr.p, sizeof(double) * 11 * (size_t)nh, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_commit_step(d2_ctx* c) {
  if (!c) return D2_ERR_ARG;
  if (c->last_kind != RUN_STEP || c->last_mode != 0 || c->body_step.n < (size_t)c->n_bodies)
    return fail(c, D2_ERR_STATE, "d2_commit_step needs a preceding d2_step_run");
  (void)hipSetDevice(c->device);
  hipLaunchKernelGGL(k_commit, dim3(grid_for(c, c->n_bodies)), dim3(256), 0, c->stream, c->n_bodies, c->bodies.p,
                     (const double*)c->body_step.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->last_mode = -1;  // the step's results no longer describe the bodies
  c->ccd_valid = false;
  return D2_OK;
}

int d2_solver_reset(d2_ctx* c) {
  if (!c) return D2_ERR_ARG;
  if (!c->warm_keys.p) return D2_OK;
  (void)hipSetDevice(c->device);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(c, (long long)c->warm_keys.n)), dim3(256), 0, c->stream,
                     c->warm_keys.p, (long long)c->warm_keys.n, WARM_EMPTY);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->warm_used.p, 1ll, 0ull);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_project_future(d2_ctx* c, const double* step_per_body) {
  if (!c || !step_per_body) return D2_ERR_ARG;
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "no bodies uploaded");
  RC(ensure(c, c->body_step, c->n_bodies));
  HIPCHK(c, hipMemcpyAsync(c->body_step.p, step_per_body, sizeof(double) * c->n_bodies, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_project, dim3(grid_for(c, c->n_bodies)), dim3(256), 0, c->stream, c->n_bodies, c->bodies.p,
                     c->body_step.p, 0);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
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
  size_t np = std::max(n_pairs, 1);
  RC(ensure(c, c->pairs, 2 * np)); RC(ensure(c, c->max_time, np)); RC(ensure(c, c->pair_rec, np * PR_STRIDE));
  RC(ensure(c, c->pair_info, 4 * np));
  if (n_pairs > 0) {
    HIPCHK(c, hipMemcpyAsync(c->pairs.p, pairs, sizeof(int32_t) * 2 * n_pairs, hipMemcpyHostToDevice, c->stream));
    if (max_time)
      HIPCHK(c, hipMemcpyAsync(c->max_time.p, max_time, sizeof(double) * n_pairs, hipMemcpyHostToDevice, c->stream));
    else
      HIPCHK(c, hipMemsetAsync(c->max_time.p, 0, sizeof(double) * n_pairs, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_pairs = n_pairs;
  int deepest = 0;
  for (int p = 0; p < n_pairs; ++p)
    deepest = std::max(deepest, c->tree_depth[c->host_body_tree[pairs[2 * p]]] +
                                    c->tree_depth[c->host_body_tree[pairs[2 * p + 1]]]);
  c->pair_depth = deepest;
  c->n_clusters = 0;  // clusters describe a pair list
  c->clusters_on_device = false;
  c->has_max_time = max_time != nullptr;
  c->last_mode = -1;
  c->ccd_valid = false;
  c->pairs_valid = true;
  return D2_OK;
}

int d2_clusters_upload(d2_ctx* c, const int32_t* cluster_of_pair, int n_clusters) {
  if (!c) return D2_ERR_ARG;
  if (!c->pairs_valid) return fail(c, D2_ERR_STATE, "d2_clusters_upload: upload or build the pair list first");
  if (n_clusters <= 0 || (c->n_pairs > 0 && !cluster_of_pair)) return fail(c, D2_ERR_ARG, "d2_clusters_upload: bad arguments");
  for (int p = 0; p < c->n_pairs; ++p)
    if (cluster_of_pair[p] < 0 || cluster_of_pair[p] >= n_clusters)
      return fail(c, D2_ERR_ARG, "d2_clusters_upload: cluster index out of range");
  (void)hipSetDevice(c->device);
  RC(ensure(c, c->cluster_of_pair, std::max(c->n_pairs, 1)));
  if (c->n_pairs > 0)
    HIPCHK(c, hipMemcpyAsync(c->cluster_of_pair.p, cluster_of_pair, sizeof(int32_t) * c->n_pairs, hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_clusters = n_clusters;
  c->clusters_on_device = false;
  return D2_OK;
}

int d2_narrow_run(d2_ctx* c, int filter, int64_t* n_contacts_out) {
  if (!c) return D2_ERR_ARG;
  return run(c, RUN_NARROW, filter ? 1 : 0, 0.f, n_contacts_out);
}

int d2_ccd_run(d2_ctx* c, int64_t* n_contacts_out) {
  if (!c) return D2_ERR_ARG;
  return run(c, RUN_CCD, 0, 0.f, n_contacts_out);
}

int d2_step_run(d2_ctx* c, float dt, int filter, double* cluster_step_out, int64_t* n_contacts_out) {
  if (!c) return D2_ERR_ARG;
  if (!(dt > 0.f)) return fail(c, D2_ERR_ARG, "d2_step_run: dt must be > 0");
  RC(run(c, RUN_STEP, filter ? 1 : 0, dt, n_contacts_out));
  if (cluster_step_out) {
    HIPCHK(c, hipMemcpyAsync(cluster_step_out, c->cluster_step.p, sizeof(double) * c->n_clusters, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return D2_OK;
}

int d2_find_contacts_bucket(d2_ctx* c, int body_a, int body_b, int n_node_pairs, const int32_t* node_pairs,
                            int32_t* pair_used_out, int64_t max_hits, int32_t* ints_out, double* reals_out,
                            int64_t* n_hits_out) {
  if (!c) return D2_ERR_ARG;
  RC(bucket_check(c, body_a, body_b, n_node_pairs, node_pairs, "d2_find_contacts_bucket"));
  if (n_node_pairs > 0 && !pair_used_out) return fail(c, D2_ERR_ARG, "d2_find_contacts_bucket: pair_used_out is NULL");
  (void)hipSetDevice(c->device);
  const size_t n = std::max(n_node_pairs, 1);
  RC(ensure(c, c->bk_np, 2 * n)); RC(ensure(c, c->bk_base, n)); RC(ensure(c, c->bk_count, n));
  RC(ensure(c, c->bk_off, n)); RC(ensure(c, c->bk_used, n)); RC(ensure(c, c->bk_counter, 1));
  long long cap = std::max<long long>((long long)c->bk_hits.n, 8ll * (long long)n);
  long long total = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    RC(ensure(c, c->bk_hits, cap)); RC(ensure(c, c->bk_out_i, 8 * cap)); RC(ensure(c, c->bk_out_r, 11 * cap));
    const BucketScratch S{c->bk_np.p, c->bk_base.p, c->bk_count.p, c->bk_off.p, c->bk_used.p, c->bk_hits.p,
                          c->bk_out_i.p, c->bk_out_r.p, c->bk_counter.p, (long long)n, cap};
    RC(bucket_run(c, S, body_a, body_b, n_node_pairs, node_pairs, pair_used_out, max_hits, ints_out, reals_out, &total));
    if (total <= cap) break;
    cap = total;  // the scratch grows to the recorded count and the call repeats
  }
  if (n_hits_out) *n_hits_out = total;
  if (total > max_hits) return fail(c, D2_ERR_CAPACITY, "d2_find_contacts_bucket: more hits than max_hits");
  return D2_OK;
}

int d2_find_contacts_bucket_noalloc(d2_ctx* c, int body_a, int body_b, int n_node_pairs, const int32_t* node_pairs,
                                    int32_t* pair_used_out, int64_t max_hits, int32_t* ints_out, double* reals_out,
                                    int64_t* n_hits_out) {
  if (!c) return D2_ERR_ARG;
  RC(bucket_check(c, body_a, body_b, n_node_pairs, node_pairs, "d2_find_contacts_bucket_noalloc"));
  if (n_node_pairs > 0 && !pair_used_out)
    return fail(c, D2_ERR_ARG, "d2_find_contacts_bucket_noalloc: pair_used_out is NULL");
  if (n_node_pairs > NOALLOC_NP)
    return fail(c, D2_ERR_CAPACITY, "d2_find_contacts_bucket_noalloc: at most 32768 node pairs per call");
  if (!c->noalloc) return fail(c, D2_ERR_HIP, "d2_find_contacts_bucket_noalloc: the scratch array has no address");
  (void)hipSetDevice(c->device);
  // the scratch array carved into the bucket call's buffers (8-byte fields first)
  unsigned char* p = c->noalloc;
  BucketScratch S;
  S.base = reinterpret_cast<long long*>(p), p += 8 * NOALLOC_NP;
  S.off = reinterpret_cast<long long*>(p), p += 8 * NOALLOC_NP;
  S.np = reinterpret_cast<int32_t*>(p), p += 8 * NOALLOC_NP;
  S.count = reinterpret_cast<int*>(p), p += 4 * NOALLOC_NP;
  S.used = reinterpret_cast<int32_t*>(p), p += 4 * NOALLOC_NP;
  S.hits = reinterpret_cast<ContactRec*>(p), p += sizeof(ContactRec) * NOALLOC_HITS;
  S.out_r = reinterpret_cast<double*>(p), p += 11 * 8 * NOALLOC_HITS;
  S.out_i = reinterpret_cast<int32_t*>(p), p += 8 * 4 * NOALLOC_HITS;
  S.counter = reinterpret_cast<unsigned long long*>(p);
  S.cap_np = NOALLOC_NP, S.cap_hits = NOALLOC_HITS;
  long long total = 0;
  RC(bucket_run(c, S, body_a, body_b, n_node_pairs, node_pairs, pair_used_out, max_hits, ints_out, reals_out, &total));
  if (n_hits_out) *n_hits_out = total;
  if (total > NOALLOC_HITS)
    return fail(c, D2_ERR_CAPACITY, "d2_find_contacts_bucket_noalloc: at most 131072 hits per call");
  if (total > max_hits) return fail(c, D2_ERR_CAPACITY, "d2_find_contacts_bucket_noalloc: more hits than max_hits");
  return D2_OK;
}

int d2_filter_contacts(d2_ctx* c, int n, const double* reals, double margin, int32_t* sorted_out, int32_t* kept_out,
                       int* n_kept_out) {
  if (!c) return D2_ERR_ARG;
  if (n < 0 || (n > 0 && (!reals || !sorted_out || !kept_out)) || !(margin >= 0.0 || margin < 0.0))
    return fail(c, D2_ERR_ARG, "d2_filter_contacts: bad arguments");
  if (n_kept_out) *n_kept_out = 0;
  if (n == 0) return D2_OK;
  (void)hipSetDevice(c->device);
  const size_t m = (size_t)n;
  RC(ensure(c, c->contacts, m)); RC(ensure(c, c->perm, m)); RC(ensure(c, c->keep, m));
  RC(ensure(c, c->kept_cnt, 2)); RC(ensure(c, c->pair_off, 2)); RC(ensure(c, c->n_big, 2));
  RC(ensure(c, c->big, 1)); RC(ensure(c, c->huge, 1)); RC(ensure(c, c->pair_rec, PR_STRIDE)); RC(ensure(c, c->scale, 1));
  RC(ensure(c, c->bk_out_r, 11 * m)); RC(ensure(c, c->hscratch, 6 * m));
  // the run buffers now hold this call's contacts: no earlier result stays readable
  c->last_mode = -1;
  c->ccd_valid = false;
  HIPCHK(c, hipMemcpyAsync(c->bk_out_r.p, reals, sizeof(double) * 11 * m, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_filter_load, dim3(grid_for(c, n)), dim3(256), 0, c->stream, n, (const double*)c->bk_out_r.p,
                     c->contacts.p, c->perm.p, c->pair_off.p, c->n_big.p);
  launch_finish(c, 1, 1, 1, nullptr, margin);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(sorted_out, c->perm.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(kept_out, c->keep.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int k = 0;  // keep flags per sorted position -> the kept inputs in sorted order
  for (int t = 0; t < n; ++t)
    if (kept_out[t]) kept_out[k++] = sorted_out[t];
  if (n_kept_out) *n_kept_out = k;
  return D2_OK;
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
  const int np = c->n_pairs;
  if (c->out_filtered) {  // the filtered output's offsets (the run only counted its total)
    scan_host_n(c, c->kept_cnt.p, c->out_off.p, np, nullptr);
    HIPCHK(c, hipGetLastError());
  }
  const int* outoff = c->out_filtered ? c->out_off.p : c->pair_off.p;
  if ((ints || reals) && c->n_out > 0) {  // the CSR total and the record count must agree
    int total = 0;
    HIPCHK(c, hipMemcpyAsync(&total, outoff + np, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((long long)total != c->n_out) return fail(c, D2_ERR_STATE, "result offsets and contact count disagree");
    RC(ensure(c, c->out_i, (size_t)c->n_out * 8)); RC(ensure(c, c->out_r, (size_t)c->n_out * 11));
    const int wblocks = std::max(1, std::min(np, c->cus * 16));
    hipLaunchKernelGGL(k_gather_wave, dim3(wblocks), dim3(64), 0, c->stream, np, c->contacts.p, c->pair_off.p,
                       c->perm.p, c->keep.p, outoff, c->out_filtered ? 1 : 0, c->out_i.p, c->out_r.p);
    HIPCHK(c, hipGetLastError());
  }
  if (offsets) HIPCHK(c, hipMemcpyAsync(offsets, outoff, sizeof(int) * (np + 1), hipMemcpyDeviceToHost, c->stream));
  if (ints && c->n_out)
    HIPCHK(c, hipMemcpyAsync(ints, c->out_i.p, sizeof(int32_t) * 8 * (size_t)c->n_out, hipMemcpyDeviceToHost, c->stream));
  if (reals && c->n_out)
    HIPCHK(c, hipMemcpyAsync(reals, c->out_r.p, sizeof(double) * 11 * (size_t)c->n_out, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_result_pairs(d2_ctx* c, double* per_pair) {
  if (!c || !per_pair) return D2_ERR_ARG;
  if (!c->ccd_valid) return fail(c, D2_ERR_STATE, "per-pair results exist after d2_ccd_run or d2_step_run");
  std::vector<double> sc(c->n_pairs), mt(c->n_pairs);
  if (c->n_pairs) {
    HIPCHK(c, hipMemcpyAsync(sc.data(), c->scale.p, sizeof(double) * c->n_pairs, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(mt.data(), c->min_time.p, sizeof(double) * c->n_pairs, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int p = 0; p < c->n_pairs; ++p) per_pair[2 * p] = sc[p], per_pair[2 * p + 1] = mt[p];
  return D2_OK;
}

int d2_cluster_steps(d2_ctx* c, const int32_t* cluster_of_pair, int n_clusters, float dt, double* out) {
  if (!c || !cluster_of_pair || !out || n_clusters <= 0) return D2_ERR_ARG;
  if (c->last_mode != 1) return fail(c, D2_ERR_STATE, "d2_cluster_steps needs a preceding d2_ccd_run");
  const int np = c->n_pairs;
  for (int p = 0; p < np; ++p)
    if (cluster_of_pair[p] < 0 || cluster_of_pair[p] >= n_clusters)
      return fail(c, D2_ERR_ARG, "d2_cluster_steps: cluster index out of range");
  RC(ensure_clusters(c, n_clusters)); RC(ensure(c, c->tmp_cluster, std::max(np, 1)));
  if (np) HIPCHK(c, hipMemcpyAsync(c->tmp_cluster.p, cluster_of_pair, sizeof(int32_t) * np, hipMemcpyHostToDevice, c->stream));
  RC(enqueue_clusters(c, c->tmp_cluster.p, n_clusters, dt, false, false, false));
  HIPCHK(c, hipMemcpyAsync(out, c->cluster_step.p, sizeof(double) * n_clusters, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

// ------------------------------------------------ broad phase, clusters
int d2_tree_bounds_upload(d2_ctx* c, const float* aabb6, const double* radius) {
  if (!c || !aabb6 || !radius) return D2_ERR_ARG;
  if (c->n_trees <= 0) return fail(c, D2_ERR_STATE, "d2_tree_bounds_upload: upload trees first");
  for (int t = 0; t < c->n_trees; ++t)
    if (!(radius[t] >= 0.0)) return fail(c, D2_ERR_ARG, "d2_tree_bounds_upload: radius must be >= 0");
  HIPCHK(c, hipMemcpyAsync(c->tb_box.p, aabb6, sizeof(float) * 6 * c->n_trees, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->tb_rad.p, radius, sizeof(double) * c->n_trees, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_tree_bounds_download(d2_ctx* c, float* aabb6, double* radius) {
  if (!c || !aabb6 || !radius) return D2_ERR_ARG;
  if (c->n_trees <= 0) return fail(c, D2_ERR_STATE, "no trees uploaded");
  HIPCHK(c, hipMemcpyAsync(aabb6, c->tb_box.p, sizeof(float) * 6 * c->n_trees, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(radius, c->tb_rad.p, sizeof(double) * c->n_trees, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_bodies_last_upload(d2_ctx* c, const double* last14) {
  if (!c || !last14) return D2_ERR_ARG;
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "d2_bodies_last_upload: upload bodies first");
  HIPCHK(c, hipMemcpyAsync(c->last_state.p, last14, sizeof(double) * 14 * (size_t)c->n_bodies, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_bodies_groups_upload(d2_ctx* c, const int32_t* group_of_body) {
  if (!c) return D2_ERR_ARG;
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "d2_bodies_groups_upload: upload bodies first");
  if (!group_of_body) {
    c->has_groups = false;
    return D2_OK;
  }
  for (int i = 0; i < c->n_bodies; ++i)
    if (group_of_body[i] < 0 || group_of_body[i] > (1 << 29))
      return fail(c, D2_ERR_ARG, "d2_bodies_groups_upload: group out of range");
  RC(ensure(c, c->group, c->n_bodies));
  HIPCHK(c, hipMemcpyAsync(c->group.p, group_of_body, sizeof(int32_t) * c->n_bodies, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->has_groups = true;
  return D2_OK;
}

int d2_broad_phase(d2_ctx* c, int mode, double t, int64_t* n_pairs_out) {
  if (!c) return D2_ERR_ARG;
  if (mode != 0 && mode != 1) return fail(c, D2_ERR_ARG, "d2_broad_phase: mode must be 0 or 1");
  if (mode == 0 && !(t >= 0.0 && t < 1e300)) return fail(c, D2_ERR_ARG, "d2_broad_phase: t must be finite and >= 0");
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "d2_broad_phase: upload bodies first");
  (void)hipSetDevice(c->device);
  // the resident pair list is replaced (and invalid until this call succeeds)
  invalidate_pairs(c);
  const int n = c->n_bodies;
  long long T = 64;
  while (T < 2ll * n) T <<= 1;
  RC(ensure(c, c->bp_box, n)); RC(ensure(c, c->bp_R, n)); RC(ensure(c, c->bp_meta, n));
  RC(ensure(c, c->bp_slot, n)); RC(ensure(c, c->bp_rank, n)); RC(ensure(c, c->bp_members, n));
  RC(ensure(c, c->bp_keys, T)); RC(ensure(c, c->bp_counts, T)); RC(ensure(c, c->bp_start, T + 1));
  RC(ensure(c, c->bp_red, BR_N + 2)); RC(ensure(c, c->bp_fill, 2));
  RC(ensure(c, c->bp_item_cnt, c->n_items)); RC(ensure(c, c->bp_item_pos, c->n_items));
  RC(ensure(c, c->bp_item_sorted, c->n_items)); RC(ensure(c, c->bp_item_off, c->n_items + 1));
  RC(ensure(c, c->scan_part, chunk_grid(c)));
  if (c->bp_cap <= 0) c->bp_cap = std::max<long long>(1 << 16, 8ll * n);
  const int* grp = c->has_groups ? c->group.p : nullptr;
  const int g = grid_for(c, n);
  for (int attempt = 0; attempt < 4; ++attempt) {
    RC(ensure(c, c->bp_chunk, c->bp_cap)); RC(ensure(c, c->pairs, 2 * c->bp_cap)); RC(ensure(c, c->max_time, c->bp_cap));
    hipLaunchKernelGGL(k_bp_clear, dim3(grid_for(c, T)), dim3(256), 0, c->stream, T, c->bp_keys.p, c->bp_counts.p,
                       c->bp_red.p);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->bp_fill.p, 2ll, 0ull);
    hipLaunchKernelGGL(k_bp_bounds, dim3(g), dim3(256), 0, c->stream, n, c->bodies.p, c->last_state.p, c->body_tree.p,
                       c->tb_box.p, c->tb_rad.p, grp, mode, t, c->bp_box.p, c->bp_R.p, c->bp_meta.p, c->bp_red.p);
    hipLaunchKernelGGL(k_bp_insert, dim3(g), dim3(256), 0, c->stream, n, c->bp_box.p, c->bp_meta.p, c->bp_red.p, T,
                       c->bp_keys.p, c->bp_counts.p, c->bp_slot.p, c->bp_rank.p);
    scan_host_n(c, c->bp_counts.p, c->bp_start.p, T, nullptr);
    hipLaunchKernelGGL(k_bp_scatter, dim3(g), dim3(256), 0, c->stream, n, c->bp_meta.p, c->bp_slot.p, c->bp_rank.p,
                       c->bp_start.p, c->bp_members.p);
    BpIn in;
    in.n = n, in.mode = mode, in.t = t, in.bodies = c->bodies.p, in.last = c->last_state.p, in.box = c->bp_box.p;
    in.R = c->bp_R.p, in.meta = c->bp_meta.p, in.statics = c->bp_statics.p, in.n_static = (int)c->host_static.size();
    in.red = c->bp_red.p, in.T = T, in.keys = c->bp_keys.p, in.counts = c->bp_counts.p, in.start = c->bp_start.p;
    in.members = c->bp_members.p;
    const int gi = (int)std::min<long long>((c->n_items + 3) / 4, (long long)c->cus * 16);
    hipLaunchKernelGGL(k_bp_pairs, dim3(std::max(gi, 1)), dim3(256), 0, c->stream, in, c->n_items, c->bp_item_body.p,
                       c->bp_item_seg.p, c->bp_cap, c->bp_chunk.p, c->bp_fill.p, c->bp_item_cnt.p, c->bp_item_pos.p,
                       c->bp_item_sorted.p, c->bp_red.p);
    scan_host_n(c, c->bp_item_cnt.p, c->bp_item_off.p, c->n_items, c->bp_fill.p + 1);
    hipLaunchKernelGGL(k_bp_gather, dim3(std::max(gi, 1)), dim3(256), 0, c->stream, c->n_items, c->bp_item_cnt.p,
                       c->bp_item_pos.p, c->bp_item_sorted.p, c->bp_item_off.p, c->bp_chunk.p, c->bodies.p,
                       c->body_tree.p, c->tree_depth_d.p, c->pairs.p, c->max_time.p, c->bp_red.p);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->host_red, c->bp_red.p, sizeof(int) * BR_N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->host_red + BR_N, c->bp_item_off.p + c->n_items, sizeof(int), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->host_red[BR_ERR] & 1)
      return fail(c, D2_ERR_RUNTIME, "d2_broad_phase: Invalid args for interpolation (state.cpp:56-58)");
    const long long total = c->host_red[BR_N];
    if (c->host_red[BR_OVF]) {
      c->bp_cap = std::max(c->bp_cap * 2, total + total / 4 + 1024);
      continue;
    }
    const int np = (int)total;
    RC(ensure(c, c->pair_rec, (size_t)std::max(np, 1) * PR_STRIDE)); RC(ensure(c, c->pair_info, 4 * (size_t)std::max(np, 1)));
    c->n_pairs = np;
    c->pair_depth = c->host_red[BR_DEPTH];
    c->has_max_time = true;
    c->pairs_valid = true;
    if (n_pairs_out) *n_pairs_out = np;
    return D2_OK;
  }
  return fail(c, D2_ERR_CAPACITY, "d2_broad_phase: pair list kept overflowing");
}

int d2_pairs_download(d2_ctx* c, int32_t* pairs, double* max_time) {
  if (!c) return D2_ERR_ARG;
  if (c->n_pairs > 0 && pairs)
    HIPCHK(c, hipMemcpyAsync(pairs, c->pairs.p, sizeof(int32_t) * 2 * c->n_pairs, hipMemcpyDeviceToHost, c->stream));
  if (c->n_pairs > 0 && max_time)
    HIPCHK(c, hipMemcpyAsync(max_time, c->max_time.p, sizeof(double) * c->n_pairs, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_clusters_build(d2_ctx* c, int* n_clusters_out) {
  if (!c) return D2_ERR_ARG;
  if (c->n_bodies <= 0) return fail(c, D2_ERR_STATE, "d2_clusters_build: upload bodies first");
  if (!c->pairs_valid) return fail(c, D2_ERR_STATE, "d2_clusters_build: upload or build the pair list first");
  (void)hipSetDevice(c->device);
  const int n = c->n_bodies, np = c->n_pairs;
  c->n_clusters = 0;  // invalid until this call succeeds
  c->clusters_on_device = false;
  RC(ensure(c, c->cc_parent, n)); RC(ensure(c, c->cc_label, n)); RC(ensure(c, c->cc_root, n));
  RC(ensure(c, c->cc_root_off, n + 1)); RC(ensure(c, c->cc_of_body, n)); RC(ensure(c, c->cc_first, n));
  RC(ensure(c, c->cc_tmin, n)); RC(ensure(c, c->cc_smin, n)); RC(ensure(c, c->bp_red, BR_N + 2));
  RC(ensure(c, c->cluster_of_pair, std::max(np, 1))); RC(ensure(c, c->scan_part, chunk_grid(c)));
  RC(ensure(c, c->bp_fill, 2));
  const int g = grid_for(c, n), gp = grid_for(c, std::max(np, 1));
  hipLaunchKernelGGL(k_bp_clear, dim3(1), dim3(64), 0, c->stream, 0ll, c->bp_keys.p, c->bp_counts.p, c->bp_red.p);
  hipLaunchKernelGGL(k_cc_init, dim3(g), dim3(256), 0, c->stream, n, c->cc_parent.p);
  if (np > 0)
    hipLaunchKernelGGL(k_cc_union, dim3(gp), dim3(256), 0, c->stream, np, c->pairs.p, c->bodies.p, c->cc_parent.p,
                       c->bp_red.p);
  hipLaunchKernelGGL(k_cc_roots, dim3(g), dim3(256), 0, c->stream, n, c->bodies.p, c->cc_parent.p, c->cc_label.p,
                     c->cc_root.p);
  scan_host_n(c, c->cc_root.p, c->cc_root_off.p, n, nullptr);
  hipLaunchKernelGGL(k_cc_init_cl, dim3(g), dim3(256), 0, c->stream, n, c->cc_tmin.p, c->cc_smin.p);
  hipLaunchKernelGGL(k_cc_assign, dim3(g), dim3(256), 0, c->stream, n, c->bodies.p, c->cc_label.p, c->cc_root_off.p,
                     c->cc_of_body.p, c->cc_first.p, c->cc_tmin.p, c->cc_smin.p);
  if (np > 0)
    hipLaunchKernelGGL(k_cc_pairs, dim3(gp), dim3(256), 0, c->stream, np, c->pairs.p, c->bodies.p, c->cc_of_body.p,
                       c->cluster_of_pair.p);
  hipLaunchKernelGGL(k_cc_rollback, dim3(g), dim3(256), 0, c->stream, n, c->bodies.p, c->last_state.p, c->cc_of_body.p,
                     c->cc_tmin.p, c->bp_red.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->host_red, c->bp_red.p, sizeof(int) * BR_N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->host_red + BR_N, c->cc_root_off.p + n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->host_red[BR_ERR] & 2)
    return fail(c, D2_ERR_ARG, "d2_clusters_build: a pair of two static bodies belongs to no cluster");
  if (c->host_red[BR_ERR] & 1)
    return fail(c, D2_ERR_RUNTIME, "d2_clusters_build: Invalid args for interpolation (state.cpp:56-58)");
  const int k = c->host_red[BR_N];
  RC(ensure_clusters(c, std::max(k, 1)));
  c->n_clusters = k;
  c->clusters_on_device = true;
  c->ccd_valid = false;
  if (n_clusters_out) *n_clusters_out = k;
  return D2_OK;
}

int d2_clusters_download(d2_ctx* c, int32_t* cluster_of_pair, int32_t* cluster_of_body, int32_t* first,
                         double* min_time, double* step) {
  if (!c) return D2_ERR_ARG;
  if (!c->clusters_on_device) return fail(c, D2_ERR_STATE, "d2_clusters_download needs d2_clusters_build");
  const int n = c->n_bodies, k = c->n_clusters;
  if (cluster_of_pair && c->n_pairs > 0)
    HIPCHK(c, hipMemcpyAsync(cluster_of_pair, c->cluster_of_pair.p, sizeof(int32_t) * c->n_pairs, hipMemcpyDeviceToHost,
                             c->stream));
  if (cluster_of_body)
    HIPCHK(c, hipMemcpyAsync(cluster_of_body, c->cc_of_body.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (first && k) HIPCHK(c, hipMemcpyAsync(first, c->cc_first.p, sizeof(int32_t) * k, hipMemcpyDeviceToHost, c->stream));
  std::vector<long long> tm(min_time ? k : 0);
  std::vector<unsigned long long> sm(step ? k : 0);
  if (min_time && k) HIPCHK(c, hipMemcpyAsync(tm.data(), c->cc_tmin.p, sizeof(long long) * k, hipMemcpyDeviceToHost, c->stream));
  if (step && k) HIPCHK(c, hipMemcpyAsync(sm.data(), c->cc_smin.p, sizeof(long long) * k, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < (min_time ? k : 0); ++i) {
    long long o = tm[i];
    o = o >= 0 ? o : o ^ 0x7fffffffffffffffll;
    std::memcpy(&min_time[i], &o, sizeof(double));
  }
  for (int i = 0; i < (step ? k : 0); ++i) std::memcpy(&step[i], &sm[i], sizeof(double));
  return D2_OK;
}

// ----------------------------------------------------------- learned gate
namespace {
int mlp_alloc(d2_ctx* c) {
  RC(ensure(c, c->mlp_w, MLP_PARAMS)); RC(ensure(c, c->mlp_m, MLP_PARAMS)); RC(ensure(c, c->mlp_v, MLP_PARAMS));
  RC(ensure(c, c->mlp_wf, MLP_PARAMS)); RC(ensure(c, c->mlp_grad, MLP_PARAMS)); RC(ensure(c, c->mlp_wb, MLP_PARAMS));
  return D2_OK;
}
// Minibatch forward/backward + parameter gradient (grad in c->mlp_grad,
// per-sample losses in c->tr_loss).
void mlp_batch_grad(d2_ctx* c, const int* idx, int nb, float lw0, float lw1, float delta) {
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3((nb + 255) / 256), dim3(256), 0, c->stream, nb, idx, c->tr_x.p, c->tr_t.p,
                     c->mlp_wf.p, lw0, lw1, delta, c->tr_act.p, c->tr_dlt.p, c->tr_loss.p);
  hipLaunchKernelGGL(k_mlp_grad, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, c->stream, nb, c->tr_act.p, c->tr_dlt.p,
                     c->mlp_grad.p);
}
int mlp_upload_data(d2_ctx* c, const float* X, const float* T, int n, int batch) {
  RC(ensure(c, c->tr_x, (size_t)n * MLP_IN)); RC(ensure(c, c->tr_t, (size_t)n * MLP_OUT));
  RC(ensure(c, c->tr_act, (size_t)batch * MLP_ACT)); RC(ensure(c, c->tr_dlt, (size_t)batch * MLP_DELTA));
  RC(ensure(c, c->tr_loss, batch)); RC(ensure(c, c->tr_idx, n));
  HIPCHK(c, hipMemcpyAsync(c->tr_x.p, X, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice, c->stream));
  if (T) HIPCHK(c, hipMemcpyAsync(c->tr_t.p, T, sizeof(float) * MLP_OUT * (size_t)n, hipMemcpyHostToDevice, c->stream));
  return D2_OK;
}
}  // namespace

int d2_mlp_num_params(void) { return MLP_PARAMS; }

int d2_mlp_set_params(d2_ctx* c, const float* params) {
  if (!c || !params) return D2_ERR_ARG;
  (void)hipSetDevice(c->device);
  RC(mlp_alloc(c));
  HIPCHK(c, hipMemcpyAsync(c->mlp_w.p, params, sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->mlp_m.p, 0, sizeof(float) * MLP_PARAMS, c->stream));
  HIPCHK(c, hipMemsetAsync(c->mlp_v.p, 0, sizeof(float) * MLP_PARAMS, c->stream));
  hipLaunchKernelGGL(k_mlp_round, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, c->stream, c->mlp_w.p, c->mlp_wb.p,
                     c->mlp_wf.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mlp_ready = true;
  c->mlp_step = 0;
  return D2_OK;
}

int d2_mlp_get_params(d2_ctx* c, float* params) {
  if (!c || !params) return D2_ERR_ARG;
  if (!c->mlp_ready) return fail(c, D2_ERR_STATE, "no network parameters set");
  HIPCHK(c, hipMemcpyAsync(params, c->mlp_w.p, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_mlp_infer(d2_ctx* c, const float* X, int n, float* Y) {
  if (!c || n < 0 || (n > 0 && (!X || !Y))) return D2_ERR_ARG;
  if (!c->mlp_ready) return fail(c, D2_ERR_STATE, "no network parameters set");
  if (n == 0) return D2_OK;
  RC(ensure(c, c->tr_x, (size_t)n * MLP_IN)); RC(ensure(c, c->tr_t, (size_t)n * MLP_OUT));
  HIPCHK(c, hipMemcpyAsync(c->tr_x.p, X, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice, c->stream));
  const int tiles = (n + 15) / 16;
  hipLaunchKernelGGL(k_mlp_infer<false>, dim3(std::max(1, std::min((tiles + 3) / 4, c->cus * 4))), dim3(256), 0,
                     c->stream, n, (const float*)c->tr_x.p, c->mlp_wb.p, c->tr_t.p, FeatIn{});
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(Y, c->tr_t.p, sizeof(float) * MLP_OUT * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_mlp_loss_grad(d2_ctx* c, const float* X, const float* T, int n, float lw0, float lw1, float delta, float* loss,
                     float* grad) {
  if (!c || !X || !T || n <= 0 || n > (1 << 20)) return D2_ERR_ARG;
  if (!c->mlp_ready) return fail(c, D2_ERR_STATE, "no network parameters set");
  RC(mlp_upload_data(c, X, T, n, n));
  mlp_batch_grad(c, nullptr, n, lw0, lw1, delta);
  HIPCHK(c, hipGetLastError());
  std::vector<float> l(n);
  HIPCHK(c, hipMemcpyAsync(l.data(), c->tr_loss.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  if (grad) HIPCHK(c, hipMemcpyAsync(grad, c->mlp_grad.p, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double sum = 0.0;
  for (float v : l) sum += v;
  if (loss) *loss = (float)(sum / n);
  return D2_OK;
}

int d2_mlp_train(d2_ctx* c, const float* X, const float* T, int n, int epochs, int batch, float lr, float lw0,
                 float lw1, float delta, uint64_t seed, float* epoch_loss) {
  if (!c || !X || !T || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.f)) return D2_ERR_ARG;
  if (!c->mlp_ready) return fail(c, D2_ERR_STATE, "set initial network parameters first");
  batch = std::min(batch, n);
  RC(mlp_upload_data(c, X, T, n, batch));
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::mt19937_64 rng(seed);
  std::vector<float> l(batch);
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(idx.begin(), idx.end(), rng);
    HIPCHK(c, hipMemcpyAsync(c->tr_idx.p, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    double sum = 0.0;
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int nb = std::min(batch, n - b0);
      mlp_batch_grad(c, c->tr_idx.p + b0, nb, lw0, lw1, delta);
      ++c->mlp_step;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)c->mlp_step), c2 = 1.f - std::pow(b2, (float)c->mlp_step);
      hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, c->stream, c->mlp_w.p, c->mlp_m.p,
                         c->mlp_v.p, c->mlp_grad.p, lr, b1, b2, 1e-8f, c1, c2, c->mlp_wb.p, c->mlp_wf.p);
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipMemcpyAsync(l.data(), c->tr_loss.p, sizeof(float) * nb, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (int i = 0; i < nb; ++i) sum += l[i];
    }
    if (epoch_loss) epoch_loss[e] = (float)(sum / n);
  }
  drop_graphs(c);  // captured inferences read the weights in place; shapes are unchanged, but be explicit
  return D2_OK;
}

int d2_pair_features(d2_ctx* c, float* X) {
  if (!c || !X) return D2_ERR_ARG;
  if (c->n_pairs <= 0) return D2_OK;
  RC(ensure(c, c->feat, (size_t)c->n_pairs * MLP_IN));
  hipLaunchKernelGGL(k_pair_features, dim3(grid_for(c, c->n_pairs)), dim3(256), 0, c->stream, c->n_pairs, c->pairs.p,
                     c->bodies.p, c->body_tree.p, c->tree_radius.p, c->has_max_time ? c->max_time.p : nullptr,
                     c->feat.p);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(X, c->feat.p, sizeof(float) * MLP_IN * (size_t)c->n_pairs, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

int d2_set_learned(d2_ctx* c, int on, float t_gate, float s_gate) {
  if (!c || on < 0 || on > 3) return D2_ERR_ARG;
  if (on && !c->mlp_ready) return fail(c, D2_ERR_STATE, "set network parameters before turning the gate on");
  c->learned_on = on != 0;
  c->gate_phases = on;
  c->gate_t = t_gate;
  c->gate_s = s_gate;
  return D2_OK;
}

int d2_set_pruning(d2_ctx* c, int on) {
  if (!c) return D2_ERR_ARG;
  if (c->prune_on != (on != 0)) drop_graphs(c);  // captured launches carry the flag
  c->prune_on = on != 0;
  return D2_OK;
}

int d2_set_touch_stats(d2_ctx* c, int on) {
  if (!c) return D2_ERR_ARG;
  // turning it on forgets the last counts; turning it off keeps them readable
  if (on && !c->touch_on) c->touch_host.clear();
  c->touch_on = on != 0;
  return D2_OK;
}

int d2_touch_stats(const d2_ctx* c, int64_t* out, int n) {
  if (!c) return D2_ERR_ARG;
  const int total = TOUCH_RUN + TOUCH_FIELDS;
  if (out)
    for (int i = 0; i < std::min(n, total); ++i) out[i] = i < (int)c->touch_host.size() ? (int64_t)c->touch_host[i] : 0;
  return total;
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
int d2_run_info(const d2_ctx* c, int64_t* info, int n) {
  if (!c || !info || n < 0) return D2_ERR_ARG;
  const int total = C_COUNT + 2 * (MAX_LEVELS + 1) * LV_STRIDE;
  for (int i = 0; i < n && i < total; ++i)
    info[i] = (int64_t)(i < C_COUNT ? c->host_cnt[i] : c->host_lvl[i - C_COUNT]);
  return total;
}
int d2_graph_stats(const d2_ctx* c, int64_t* captures, int64_t* launches) {
  if (!c || !captures || !launches) return D2_ERR_ARG;
  *captures = c->graph_captures;
  *launches = c->graph_launches;
  return D2_OK;
}
int d2_sync(d2_ctx* c) {
  if (!c) return D2_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return D2_OK;
}

}  // extern "C"
