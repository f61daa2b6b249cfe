// Host check of the partition-rank formulation the post-processing kernel
// uses to replay libstdc++'s std::sort in parallel.
//
// std::sort (introsort) is: median-of-three pivot moved to the front, an
// unguarded Hoare partition, recursion on the right part and a loop on the
// left while a segment holds more than 16 elements, heapsort at depth 0, and a
// final insertion sort.  The kernel computes each partition without a scan:
// with A_k the k-th position (from the left) holding an element not less than
// the pivot and B_k the k-th position (from the right) holding one not greater,
// the partition swaps (A_k, B_k) for every k < K, K = #{k : A_k < B_k}, and
// returns min(A_K, B_{K-1}).  The final insertion sort moves no element across
// a leaf segment's border, so it equals a stable sort of every leaf segment.
//
// This program restates that formulation sequentially and compares the
// permutation with std::sort on tie-heavy random keys (the comparator of
// contact.h:35-43 on (eps sum, distance) pairs).  Prints "sort <cases> <mismatches>".
#include <algorithm>
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

struct Key {
  double eps, dist;
};
static bool less_key(const Key& x, const Key& y) {
  if (x.eps == y.eps) return x.dist < y.dist;
  return x.eps < y.eps;
}

struct Rec {
  Key k;
  int id;
};

// Reference: std::sort of the records themselves.
static std::vector<int> std_perm(const std::vector<Key>& keys) {
  std::vector<Rec> r(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) r[i] = {keys[i], (int)i};
  std::sort(r.begin(), r.end(), [](const Rec& a, const Rec& b) { return less_key(a.k, b.k); });
  std::vector<int> out(r.size());
  for (size_t i = 0; i < r.size(); ++i) out[i] = r[i].id;
  return out;
}

struct RankSort {
  std::vector<int> a;
  const std::vector<Key>& k;
  std::vector<std::pair<int, int>> leaves;  // [f, l) to stable-sort at the end
  bool used_heap = false;
  explicit RankSort(const std::vector<Key>& keys) : a(keys.size()), k(keys) {
    for (size_t i = 0; i < a.size(); ++i) a[i] = (int)i;
  }
  bool lt(int x, int y) const { return less_key(k[x], k[y]); }
  void median_to_first(int res, int x, int y, int z) {
    if (lt(a[x], a[y])) {
      if (lt(a[y], a[z])) std::swap(a[res], a[y]);
      else if (lt(a[x], a[z])) std::swap(a[res], a[z]);
      else std::swap(a[res], a[x]);
    } else if (lt(a[x], a[z])) std::swap(a[res], a[x]);
    else if (lt(a[y], a[z])) std::swap(a[res], a[z]);
    else std::swap(a[res], a[y]);
  }
  // Partition of [lo, hi) around the element at lo - 1, by ranks.
  int partition(int lo, int hi) {
    const int pv = a[lo - 1];
    std::vector<int> A, B;
    for (int i = lo; i < hi; ++i)
      if (!lt(a[i], pv)) A.push_back(i);
    for (int i = hi - 1; i >= lo - 1; --i)  // the pivot's own slot ends the right scan
      if (!lt(pv, a[i])) B.push_back(i);
    size_t K = 0;
    while (K < A.size() && K < B.size() && A[K] < B[K]) ++K;
    const int ak = K < A.size() ? A[K] : hi;
    const int bk1 = K > 0 ? B[K - 1] : hi;
    for (size_t q = 0; q < K; ++q) std::swap(a[A[q]], a[B[q]]);
    return std::min(ak, bk1);
  }
  void heap_sort(int f, int l) {  // the kernel hands this (rare) case to one lane
    used_heap = true;
    std::vector<Rec> r;
    for (int i = f; i < l; ++i) r.push_back({k[a[i]], a[i]});
    auto cmp = [](const Rec& x, const Rec& y) { return less_key(x.k, y.k); };
    std::make_heap(r.begin(), r.end(), cmp);
    std::sort_heap(r.begin(), r.end(), cmp);
    for (int i = f; i < l; ++i) a[i] = r[i - f].id;
  }
  void sort() {
    const int n = (int)a.size();
    if (n == 0) return;
    int depth = 2 * (63 - __builtin_clzll((unsigned long long)n));
    std::vector<int> sf{0}, sl{n}, sd{depth};
    while (!sf.empty()) {
      int f = sf.back(), l = sl.back(), d = sd.back();
      sf.pop_back(), sl.pop_back(), sd.pop_back();
      bool heaped = false;
      while (l - f > 16) {
        if (d == 0) {
          heap_sort(f, l);
          heaped = true;
          break;
        }
        --d;
        median_to_first(f, f + 1, f + (l - f) / 2, l - 1);
        int cut = partition(f + 1, l);
        sf.push_back(cut), sl.push_back(l), sd.push_back(d);
        l = cut;
      }
      if (!heaped) leaves.push_back({f, l});
    }
    // stable sort of each leaf segment by rank
    for (auto [f, l] : leaves) {
      std::vector<int> out(l - f);
      for (int j = f; j < l; ++j) {
        int r = 0;
        for (int i = f; i < l; ++i) r += lt(a[i], a[j]) || (i < j && !lt(a[j], a[i]));
        out[r] = a[j];
      }
      for (int j = f; j < l; ++j) a[j] = out[j - f];
    }
  }
};

int main() {
  std::mt19937_64 g(12345);
  long cases = 0, bad = 0;
  for (int rep = 0; rep < 4000; ++rep) {
    int n = (int)(g() % 600);
    int n_eps = 1 + (int)(g() % 3), n_dist = 1 + (int)(g() % (rep % 4 == 0 ? 4 : 40));
    std::vector<Key> keys(n);
    for (auto& kk : keys) {
      kk.eps = 0.02 + 0.01 * (double)(g() % n_eps);
      kk.dist = (g() % 3 == 0) ? 0.0 : 0.001 * (double)(g() % n_dist);
    }
    if (rep % 7 == 0) std::sort(keys.begin(), keys.end(), less_key);  // presorted
    if (rep % 11 == 0) std::reverse(keys.begin(), keys.end());
    RankSort rs(keys);
    rs.sort();
    ++cases;
    if (rs.a != std_perm(keys)) ++bad;
  }
  std::printf("sort %ld %ld\n", cases, bad);
  return 0;
}
