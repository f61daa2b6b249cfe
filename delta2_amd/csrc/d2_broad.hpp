// d2_broad.hpp -- device side of the broad phase and the cluster separation
// (included by d2_kernels.hip inside namespace d2).
//
// Broad phase: broadPhaseEmbree (collision_detection/broad_phase_embree.cpp:
// 248-289) without Embree.  Every body gets the reference's float bounding
// box (boundsFunc :60-84 for a call with a step t, boundsFuncIndividualTime
// :86-112 for the cluster stepper); candidate pairs come from a uniform grid
// over the non-static bodies (cell side = the largest non-static box extent,
// so overlapping boxes sit in the same or adjacent cells) and a sweep of each
// static body over all later bodies; a candidate is kept when the boxes
// overlap (closed intervals, as Embree's collider tests them) and the sphere
// test of collideFunc / collideFuncIndividualTime (:143-159, :178-209) passes.
// Each body's partners with a larger index are sorted in the wave that found
// them, so the pair list comes out sorted by (a, b) without a global sort.
//
// Cluster separation: separateCollisionClusters (separate_clusters.cpp:
// 355-459): lock-free union-find over pairs of two non-static bodies (a root
// is always the smallest index of its tree), clusters numbered by their
// smallest member, roll-back of the members to the cluster's earliest time.

constexpr int BP_SEG = 4096;  // bodies one static body's wave sweeps per item
constexpr int BP_WAVE_CAP = 512;  // partners a wave sorts in LDS
enum BpRed : int {  // reductions over the non-static bodies' boxes
  BR_EXT = 0,       // largest box extent (float bits, non-negative)
  BR_LO = 1,        // smallest lower corner x, y, z (ordered ints)
  BR_HIZ = 4,       // largest upper z (ordered int)
  BR_ERR = 5,       // bit 0: State::interpolate would throw; bit 1: pair of two static bodies
  BR_OVF = 6,       // pair capacity exceeded
  BR_DEPTH = 7,     // deepest depth(a) + depth(b) over the pairs
  BR_N = 8
};
enum BpMeta : int { BM_STATIC = 1 };

__device__ __forceinline__ int f2ord(float f) {
  const int b = __float_as_int(f);
  return b >= 0 ? b : b ^ 0x7fffffff;
}
__device__ __forceinline__ float ord2f(int o) { return __int_as_float(o >= 0 ? o : o ^ 0x7fffffff); }
__device__ __forceinline__ long long d2ord(double d) {
  const long long b = __double_as_longlong(d);
  return b >= 0 ? b : b ^ 0x7fffffffffffffffll;
}
__device__ __forceinline__ double ord2d(long long o) {
  return __longlong_as_double(o >= 0 ? o : o ^ 0x7fffffffffffffffll);
}

struct Box8 {  // float lo xyz, pad, hi xyz, pad
  float v[8];
};
__device__ __forceinline__ void box_extend(float* lo, float* hi, const float* blo, const float* bhi) {
#pragma unroll
  for (int k = 0; k < 3; ++k) lo[k] = fminf(lo[k], blo[k]), hi[k] = fmaxf(hi[k], bhi[k]);
}
__device__ __forceinline__ bool box_overlap(const Box8& a, const Box8& b) {
  return !(a.v[4] < b.v[0] || b.v[4] < a.v[0] || a.v[5] < b.v[1] || b.v[5] < a.v[1] || a.v[6] < b.v[2] ||
           b.v[6] < a.v[2]);
}

// Bounding boxes of every body and the grid reductions.
// mode 0: boundsFunc with the call's step t; mode 1: boundsFuncIndividualTime.
__global__ void k_bp_bounds(int n, const double* bodies, const double* last, const int32_t* body_tree,
                            const float* tb_box, const double* tb_rad, const int32_t* group, int mode, double t,
                            Box8* box, double* R, int* meta, int* red) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double* r = bodies + 40 * (size_t)i;
    const int tr = body_tree[i];
    const bool st = r[38] != 0.0;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (st) {
      // MeshData::getAABB(transform) (model/mesh.cpp:136-153): float corners
      // through the float cast of the pose.
      double m[16];
      pose(quat{r[0], r[1], r[2], r[3]}, v3<double>{r[4], r[5], r[6]}, m);
      float mf[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) mf[k] = (float)m[k];
      const float* mb = tb_box + 6 * (size_t)tr;
      for (int xi = 0; xi < 2; ++xi)
        for (int yi = 0; yi < 2; ++yi)
          for (int zi = 0; zi < 2; ++zi) {
            const v3<float> c = xform(mf, v3<float>{xi ? mb[3] : mb[0], yi ? mb[4] : mb[1], zi ? mb[5] : mb[2]});
            const float cv[3] = {c.x, c.y, c.z};
            box_extend(lo, hi, cv, cv);
          }
    } else {
      const float br = (float)tb_rad[tr];  // MeshData::getSphereBounds (mesh.cpp:155-158)
      const float s0[3] = {(float)r[4], (float)r[5], (float)r[6]};
      float s1[3];
      if (mode == 0) {
        const float tf = (float)t;
        s1[0] = s0[0] + (float)r[7] * tf, s1[1] = s0[1] + (float)r[8] * tf, s1[2] = s0[2] + (float)r[9] * tf;
      } else {
        s1[0] = (float)r[18], s1[1] = (float)r[19], s1[2] = (float)r[20];
      }
      const float* pts[3] = {s0, s1, s1};
      float sl[3];
      if (mode == 1) {
        const double* l = last + 14 * (size_t)i;
        sl[0] = (float)l[4], sl[1] = (float)l[5], sl[2] = (float)l[6];
        pts[0] = sl, pts[1] = s0, pts[2] = s1;  // last, current, future (:98-104)
      }
      for (int q = 0; q < (mode == 1 ? 3 : 2); ++q) {
        const float blo[3] = {-br + pts[q][0], -br + pts[q][1], -br + pts[q][2]};
        const float bhi[3] = {br + pts[q][0], br + pts[q][1], br + pts[q][2]};
        box_extend(lo, hi, blo, bhi);
      }
    }
    const float e = (float)r[37];  // BBox3fa::enlarge_by
    Box8 b;
    b.v[0] = lo[0] - e, b.v[1] = lo[1] - e, b.v[2] = lo[2] - e, b.v[3] = 0.f;
    b.v[4] = hi[0] + e, b.v[5] = hi[1] + e, b.v[6] = hi[2] + e, b.v[7] = 0.f;
    box[i] = b;
    R[i] = tb_rad[tr] + r[37];  // getBoundingRadius() + geo_eps (double)
    meta[i] = (group ? group[i] : 0) * 2 + (st ? BM_STATIC : 0);
    if (!st) {
      const float ext = fmaxf(b.v[4] - b.v[0], fmaxf(b.v[5] - b.v[1], b.v[6] - b.v[2]));
      atomicMax(&red[BR_EXT], __float_as_int(ext > 0.f ? ext : 0.f));
      atomicMin(&red[BR_LO + 0], f2ord(b.v[0]));
      atomicMin(&red[BR_LO + 1], f2ord(b.v[1]));
      atomicMin(&red[BR_LO + 2], f2ord(b.v[2]));
      atomicMax(&red[BR_HIZ], f2ord(b.v[6]));
    }
  }
}

// Grid of the non-static bodies: cell side h (1% above the largest extent, so
// rounding never puts two overlapping boxes two cells apart), origin at the
// smallest lower corner; groups are stacked along z with a gap of two cells.
struct BpGrid {
  double h, o[3];
  long long zspan;
};
__device__ __forceinline__ BpGrid bp_grid(const int* red) {
  BpGrid g;
  const double ext = (double)__int_as_float(red[BR_EXT]);
  g.h = ext > 0.0 ? ext * 1.01 : 1.0;
  for (int k = 0; k < 3; ++k) g.o[k] = (double)ord2f(red[BR_LO + k]);
  g.zspan = (long long)floor(((double)ord2f(red[BR_HIZ]) - g.o[2]) / g.h) + 3;
  return g;
}
__device__ __forceinline__ void bp_cell(const BpGrid& g, const Box8& b, int grp, long long* c) {
  c[0] = (long long)floor(((double)b.v[0] - g.o[0]) / g.h);
  c[1] = (long long)floor(((double)b.v[1] - g.o[1]) / g.h);
  c[2] = (long long)floor(((double)b.v[2] - g.o[2]) / g.h) + (long long)grp * g.zspan;
}
constexpr unsigned long long BP_EMPTY = ~0ull;
__device__ __forceinline__ unsigned long long bp_key(long long x, long long y, long long z) {
  const unsigned long long m = (1ull << 21) - 1;  // coordinates wrap: aliasing only adds candidates
  return ((unsigned long long)x & m) | (((unsigned long long)y & m) << 21) | (((unsigned long long)z & m) << 42);
}
__device__ __forceinline__ unsigned bp_hash(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (unsigned)k;
}

__global__ void k_fill_u64(unsigned long long* p, long long n, unsigned long long v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_bp_clear(long long T, unsigned long long* keys, int* counts, int* red) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < T; i += (long long)gridDim.x * blockDim.x)
    keys[i] = BP_EMPTY, counts[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    red[BR_EXT] = 0;
    red[BR_LO] = red[BR_LO + 1] = red[BR_LO + 2] = 0x7fffffff;
    red[BR_HIZ] = (int)0x80000000;
    red[BR_ERR] = red[BR_OVF] = red[BR_DEPTH] = 0;
  }
}

// Open addressing with linear probing; the table holds twice as many slots as
// bodies, so every probe sequence ends.
__global__ void k_bp_insert(int n, const Box8* box, const int* meta, const int* red, long long T,
                            unsigned long long* keys, int* counts, int* slot_of, int* rank_of) {
  const BpGrid g = bp_grid(red);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (meta[i] & BM_STATIC) continue;
    long long c[3];
    bp_cell(g, box[i], meta[i] >> 1, c);
    const unsigned long long key = bp_key(c[0], c[1], c[2]);
    long long s = bp_hash(key) & (T - 1);
    for (;;) {
      const unsigned long long prev = atomicCAS(&keys[s], BP_EMPTY, key);
      if (prev == BP_EMPTY || prev == key) break;
      s = (s + 1) & (T - 1);
    }
    slot_of[i] = (int)s;
    rank_of[i] = atomicAdd(&counts[s], 1);
  }
}
__device__ __forceinline__ long long bp_lookup(unsigned long long key, long long T, const unsigned long long* keys) {
  long long s = bp_hash(key) & (T - 1);
  for (;;) {
    const unsigned long long k = keys[s];
    if (k == key) return s;
    if (k == BP_EMPTY) return -1;
    s = (s + 1) & (T - 1);
  }
}
__global__ void k_bp_scatter(int n, const int* meta, const int* slot_of, const int* rank_of, const int* start,
                             int* members) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (!(meta[i] & BM_STATIC)) members[start[slot_of[i]] + rank_of[i]] = i;
}

// State::interpolate (model/state.cpp:55-70) of translation and velocity.
__device__ __forceinline__ bool bp_interp(const double* cur, const double* last, double time, v3<double>& t,
                                          v3<double>& v) {
  const double ct = cur[13], lt = last[13];
  if (time > ct || time < lt || lt > ct) return false;
  if (lt == ct) {
    t = v3<double>{cur[4], cur[5], cur[6]};
    v = v3<double>{cur[7], cur[8], cur[9]};
    return true;
  }
  const double f = (time - lt) / (ct - lt);
  t = v3<double>{f * cur[4] + (1 - f) * last[4], f * cur[5] + (1 - f) * last[5], f * cur[6] + (1 - f) * last[6]};
  v = v3<double>{f * cur[7] + (1 - f) * last[7], f * cur[8] + (1 - f) * last[8], f * cur[9] + (1 - f) * last[9]};
  return true;
}
// sphereSphereFirstIntersect (collision_detection/continuous_comparison.h:12-36).
__device__ __forceinline__ double sphere_first(v3<double> ap, v3<double> av, double ar, v3<double> bp, v3<double> bv,
                                               double br) {
  const v3<double> ab = sub(ap, bp), rv = sub(av, bv);
  const double rr = ar + br;
  const double c = dot(ab, ab) - rr * rr;
  if (c < 0.0) return 0.0;
  const double a = dot(rv, rv);
  const double b = dot(rv, ab);
  if (b >= 0.0) return -1.0;
  const double d = b * b - a * c;
  if (d < 0.0) return -1.0;
  return (-b - sqrt(d)) / a;
}
// collideFunc / collideFuncIndividualTime for a < b (both not static).
__device__ __forceinline__ bool bp_collide(int a, int b, int mode, double t, const double* bodies, const double* last,
                                           const double* R, const int* meta, int* red) {
  const double* ra = bodies + 40 * (size_t)a;
  const double* rb = bodies + 40 * (size_t)b;
  v3<double> pa{ra[4], ra[5], ra[6]}, va{ra[7], ra[8], ra[9]};
  v3<double> pb{rb[4], rb[5], rb[6]}, vb{rb[7], rb[8], rb[9]};
  double window = t;
  if (mode == 1) {
    const bool sa = meta[a] & BM_STATIC, sb = meta[b] & BM_STATIC;
    if (!sa && !sb) {
      bool ok;
      if (ra[13] < rb[13])
        ok = bp_interp(rb, last + 14 * (size_t)b, ra[13], pb, vb);
      else
        ok = bp_interp(ra, last + 14 * (size_t)a, rb[13], pa, va);
      if (!ok) {
        atomicOr(&red[BR_ERR], 1);
        return false;
      }
    }
    window = smax(ra[39], rb[39]);
  }
  const double s = sphere_first(pa, va, R[a], pb, vb, R[b]);
  return s >= 0.0 && s < window;
}

// Items: one per non-static body, ceil((n - 1 - a) / BP_SEG) (at least one)
// per static body a, in body order; item_body / item_seg describe them.
// Each item's wave writes its partners to a chunk of `chunk` allocated with
// one atomic and records the chunk position, the count and whether the chunk
// is sorted (a static body's sweep and a list that fit in LDS are; a longer
// list is placed by rank when it is gathered).
__device__ __forceinline__ void lds_rank_sort(int* buf, int cnt) {
  // distinct body indices: a value's rank is the number of smaller values
  constexpr int PER = BP_WAVE_CAP / 64;
  const int lane = lane_id();
  int vals[PER], ranks[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int k = lane + 64 * q;
    vals[q] = k < cnt ? buf[k] : 0x7fffffff;
    ranks[q] = 0;
  }
  for (int j = 0; j < cnt; ++j) {
    const int x = buf[j];
#pragma unroll
    for (int q = 0; q < PER; ++q) ranks[q] += x < vals[q] ? 1 : 0;
  }
  wave_sync_lds();
#pragma unroll
  for (int q = 0; q < PER; ++q)
    if (lane + 64 * q < cnt) buf[ranks[q]] = vals[q];
  wave_sync_lds();
}

struct BpIn {
  int n, mode;
  double t;
  const double* bodies;
  const double* last;
  const Box8* box;
  const double* R;
  const int* meta;
  const int* statics;
  int n_static;
  const int* red;
  long long T;
  const unsigned long long* keys;
  const int* counts;
  const int* start;
  const int* members;
};
__device__ __forceinline__ bool bp_test(const BpIn& in, int a, const Box8& ba, int ma, int j, int* red_w) {
  if (j <= a || (in.meta[j] >> 1) != (ma >> 1)) return false;
  if ((in.meta[j] & BM_STATIC) && (ma & BM_STATIC)) return false;  // collideFunc: not both static
  if (!box_overlap(ba, in.box[j])) return false;
  return bp_collide(a, j, in.mode, in.t, in.bodies, in.last, in.R, in.meta, red_w);
}

__global__ void __launch_bounds__(256) k_bp_pairs(BpIn in, int n_items, const int* item_body, const int* item_seg,
                                                  long long cap, int2* chunk, unsigned long long* chunk_fill,
                                                  int* item_cnt, long long* item_pos, int* item_sorted, int* red_w) {
  __shared__ int lds[4][BP_WAVE_CAP];
  __shared__ int nb_pre[4][28], nb_st[4][27];
  const int w = threadIdx.x >> 6, lane = lane_id();
  const BpGrid g = bp_grid(in.red);
  for (int it = blockIdx.x * 4 + w; it < n_items; it += gridDim.x * 4) {
    const int a = item_body[it];
    const Box8 ba = in.box[a];
    const int ma = in.meta[a];
    int cnt = 0, sorted = 1;
    long long pos = 0;
    if (ma & BM_STATIC) {
      // sweep over the bodies after a, in index order (the chunk comes out sorted)
      const long long lo = (long long)a + 1 + (long long)item_seg[it] * BP_SEG;
      const long long hi = lo + BP_SEG < in.n ? lo + BP_SEG : in.n;
      for (int pass = 0; pass < 2; ++pass) {
        int c = 0;
        for (long long b0 = lo; b0 < hi; b0 += 64) {
          const long long j = b0 + lane;
          const bool h = j < hi && bp_test(in, a, ba, ma, (int)j, red_w);
          const unsigned long long m = __ballot(h);
          if (pass == 1 && h) chunk[pos + c + __popcll(m & lanes_below())] = make_int2(a, (int)j);
          c += __popcll(m);
        }
        if (pass == 1) break;
        cnt = c;
        if (cnt == 0) break;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(chunk_fill, (unsigned long long)cnt);
        pos = (long long)__shfl((long long)base, 0);
        if (pos + cnt > cap) {
          if (lane == 0) atomicOr(&red_w[BR_OVF], 1);
          break;
        }
      }
    } else {
      // candidates: the members of the 27 cells around a's cell, then the static bodies
      long long c0[3];
      bp_cell(g, ba, ma >> 1, c0);
      int my_cnt = 0, my_st = 0;
      long long s = -1;
      if (lane < 27)
        s = bp_lookup(bp_key(c0[0] + lane % 3 - 1, c0[1] + (lane / 3) % 3 - 1, c0[2] + lane / 9 - 1), in.T, in.keys);
      // far-apart cells can share a key (coordinates wrap): visit each slot once
      for (int d = 0; d < 27; ++d) {
        const long long o = __shfl(s, d);
        if (d < lane && o == s) s = -1;
      }
      if (s >= 0) my_cnt = in.counts[s], my_st = in.start[s];
      int incl = my_cnt;
#pragma unroll
      for (int d = 1; d < 32; d <<= 1) {
        const int o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
      }
      if (lane < 27) nb_pre[w][lane + 1] = incl, nb_st[w][lane] = my_st;
      if (lane == 0) nb_pre[w][0] = 0;
      wave_sync_lds();
      const int total = nb_pre[w][27];
      for (int pass = 0; pass < 2; ++pass) {
        int c = 0;
        for (int k0 = 0; k0 < total + in.n_static; k0 += 64) {
          const int k = k0 + lane;
          int j = -1;
          if (k < total) {
            int q = 0;
            while (nb_pre[w][q + 1] <= k) ++q;
            j = in.members[nb_st[w][q] + (k - nb_pre[w][q])];
          } else if (k < total + in.n_static) {
            j = in.statics[k - total];
          }
          const bool h = j >= 0 && bp_test(in, a, ba, ma, j, red_w);
          const unsigned long long m = __ballot(h);
          const int at = c + __popcll(m & lanes_below());
          if (h) {
            if (pass == 0 && at < BP_WAVE_CAP) lds[w][at] = j;
            if (pass == 1) chunk[pos + at] = make_int2(a, j);
          }
          c += __popcll(m);
        }
        if (pass == 1) break;
        cnt = c;
        if (cnt == 0) break;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(chunk_fill, (unsigned long long)cnt);
        pos = (long long)__shfl((long long)base, 0);
        if (pos + cnt > cap) {
          if (lane == 0) atomicOr(&red_w[BR_OVF], 1);
          break;
        }
        if (cnt <= BP_WAVE_CAP) {
          wave_sync_lds();
          lds_rank_sort(lds[w], cnt);
          for (int k = lane; k < cnt; k += 64) chunk[pos + k] = make_int2(a, lds[w][k]);
          break;
        }
        sorted = 0;  // second pass writes the list unsorted
      }
    }
    if (lane == 0) {
      item_cnt[it] = cnt;
      item_pos[it] = pos;
      item_sorted[it] = sorted;
    }
    wave_sync_lds();
  }
}

// Final pair list: item chunks at their scanned offsets (unsorted chunks by
// rank), the CCD interval min(last steps) of each pair (separate_clusters.cpp:
// 490) and the deepest tree-depth sum.
__global__ void __launch_bounds__(256) k_bp_gather(int n_items, const int* item_cnt, const long long* item_pos,
                                                   const int* item_sorted, const int* item_off, const int2* chunk,
                                                   const double* bodies, const int32_t* body_tree,
                                                   const int32_t* tree_depth, int32_t* pairs, double* max_time,
                                                   int* red_w) {
  if (red_w[BR_OVF]) return;
  const int w = threadIdx.x >> 6, lane = lane_id();
  int deepest = 0;
  for (int it = blockIdx.x * 4 + w; it < n_items; it += gridDim.x * 4) {
    const int cnt = item_cnt[it];
    if (cnt == 0) continue;
    const long long pos = item_pos[it];
    const int off = item_off[it];
    const bool sorted = item_sorted[it] != 0;
    for (int k = lane; k < cnt; k += 64) {
      const int2 pr = chunk[pos + k];
      int rk = k;
      if (!sorted) {
        rk = 0;
        for (int q = 0; q < cnt; ++q) rk += chunk[pos + q].y < pr.y ? 1 : 0;
      }
      const int p = off + rk;
      pairs[2 * p] = pr.x;
      pairs[2 * p + 1] = pr.y;
      const double la = bodies[40 * (size_t)pr.x + 39], lb = bodies[40 * (size_t)pr.y + 39];
      max_time[p] = smin(la, lb);
      const int dsum = tree_depth[body_tree[pr.x]] + tree_depth[body_tree[pr.y]];
      deepest = dsum > deepest ? dsum : deepest;
    }
  }
  if (deepest) atomicMax(&red_w[BR_DEPTH], deepest);
}

// ---------------------------------------------------------------- clusters
__device__ __forceinline__ int cc_find(const int* parent, int x) {
  int p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
  while (p != x) {  // parents only ever point to smaller indices
    x = p;
    p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
  }
  return x;
}
__global__ void k_cc_init(int n, int* parent) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) parent[i] = i;
}
// Hooks the larger root under the smaller; a failed CAS means the root was
// hooked meanwhile, so the loop retries from the new roots.
__global__ void k_cc_union(int n_pairs, const int32_t* pairs, const double* bodies, int* parent, int* red_w) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    const int a = pairs[2 * p], b = pairs[2 * p + 1];
    const bool sa = bodies[40 * (size_t)a + 38] != 0.0, sb = bodies[40 * (size_t)b + 38] != 0.0;
    if (sa && sb) atomicOr(&red_w[BR_ERR], 2);
    if (sa || sb) continue;
    int ra = cc_find(parent, a), rb = cc_find(parent, b);
    while (ra != rb) {
      if (ra < rb) {
        const int x = ra;
        ra = rb;
        rb = x;
      }
      const int old = atomicCAS(&parent[ra], ra, rb);
      if (old == ra) break;
      ra = cc_find(parent, old);
      rb = cc_find(parent, rb);
    }
  }
}
__global__ void k_cc_roots(int n, const double* bodies, const int* parent, int* label, int* is_root) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int r = cc_find(parent, i);
    label[i] = r;
    is_root[i] = (r == i && bodies[40 * (size_t)i + 38] == 0.0) ? 1 : 0;
  }
}
// Cluster of each body (-1 static), first member, earliest current time and
// smallest last step per cluster.
__global__ void k_cc_assign(int n, const double* bodies, const int* label, const int* root_off, int* of_body,
                            int* first, long long* tmin, unsigned long long* smin_bits) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double* r = bodies + 40 * (size_t)i;
    if (r[38] != 0.0) {
      of_body[i] = -1;
      continue;
    }
    const int c = root_off[label[i]];
    of_body[i] = c;
    if (label[i] == i) first[c] = i;
    atomicMin(&tmin[c], d2ord(r[13]));
    atomic_min_pos(&smin_bits[c], r[39]);
  }
}
__global__ void k_cc_init_cl(int n_clusters, long long* tmin, unsigned long long* smin_bits) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_clusters; c += gridDim.x * blockDim.x)
    tmin[c] = 0x7fffffffffffffffll, smin_bits[c] = ~0ull;
}
__global__ void k_cc_pairs(int n_pairs, const int32_t* pairs, const double* bodies, const int* of_body,
                           int32_t* of_pair) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs; p += gridDim.x * blockDim.x) {
    const int a = pairs[2 * p], b = pairs[2 * p + 1];
    of_pair[p] = of_body[bodies[40 * (size_t)a + 38] != 0.0 ? b : a];
  }
}
// Quaternion::slerp (Eigen's published algorithm: sine weights over the
// shorter arc, linear weights when the quaternions are within epsilon).
__device__ __forceinline__ quat slerp_q(const quat& a, double t, const quat& b) {
  const double one = 1.0 - 2.220446049250313e-16;
  const double d = ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
  const double absD = fabs(d);
  double s0, s1;
  if (absD >= one) {
    s0 = 1.0 - t;
    s1 = t;
  } else {
    const double th = acos(absD);
    const double st = sin(th);
    s0 = sin((1.0 - t) * th) / st;
    s1 = sin(t * th) / st;
  }
  if (d < 0.0) s1 = -s1;
  return quat{s0 * a.w + s1 * b.w, s0 * a.x + s1 * b.x, s0 * a.y + s1 * b.y, s0 * a.z + s1 * b.z};
}
// Particle::rollBackState (model/particle.cpp:96-98) to the cluster's earliest time.
__global__ void k_cc_rollback(int n, double* bodies, const double* last, const int* of_body, const long long* tmin,
                              int* red_w) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = of_body[i];
    if (c < 0) continue;
    double* r = bodies + 40 * (size_t)i;
    const double* l = last + 14 * (size_t)i;
    const double time = ord2d(tmin[c]);
    const double ct = r[13], lt = l[13];
    if (time > ct || time < lt || lt > ct) {
      atomicOr(&red_w[BR_ERR], 1);
      continue;
    }
    if (lt == ct) continue;
    const double f = (time - lt) / (ct - lt);
    double o[14];
#pragma unroll
    for (int k = 4; k < 13; ++k) o[k] = f * r[k] + (1 - f) * l[k];
    const quat q = slerp_q(quat{l[0], l[1], l[2], l[3]}, f, quat{r[0], r[1], r[2], r[3]});
    o[0] = q.w, o[1] = q.x, o[2] = q.y, o[3] = q.z;
    o[13] = time;
#pragma unroll
    for (int k = 0; k < 14; ++k) r[k] = o[k];
  }
}
