// d2_sort.hpp -- the permutation libstdc++'s std::sort produces, on one
// thread (host/device code; the includer defines D2_HD).
//
// Introsort as libstdc++ implements it: median-of-three pivot moved to the
// front, unguarded Hoare partition, depth limit 2 floor(log2 n), heapsort
// (make_heap + sort_heap) when the depth is spent, segments of at most 16
// left to the final insertion sort (guarded over the first 16 elements,
// unguarded after).  Reproducing it keeps every choice among exactly tied keys
// identical to the reference's.  a[] holds element ids; lt(x, y) compares two
// element ids.
#pragma once

namespace d2 {

template <class Less> struct StdSort {
  int* a;
  Less lt;
  D2_HD void swp(int i, int j) {
    int t = a[i];
    a[i] = a[j];
    a[j] = t;
  }
  D2_HD void median_to_first(int res, int x, int y, int z) {
    if (lt(a[x], a[y])) {
      if (lt(a[y], a[z])) swp(res, y);
      else if (lt(a[x], a[z])) swp(res, z);
      else swp(res, x);
    } else if (lt(a[x], a[z])) swp(res, x);
    else if (lt(a[y], a[z])) swp(res, z);
    else swp(res, y);
  }
  D2_HD int partition(int first, int last, int pivot) {
    while (true) {
      while (lt(a[first], a[pivot])) ++first;
      --last;
      while (lt(a[pivot], a[last])) --last;
      if (!(first < last)) return first;
      swp(first, last);
      ++first;
    }
  }
  D2_HD void push_heap(int first, int hole, int top, int val) {
    int parent = (hole - 1) / 2;
    while (hole > top && lt(a[first + parent], val)) {
      a[first + hole] = a[first + parent];
      hole = parent;
      parent = (hole - 1) / 2;
    }
    a[first + hole] = val;
  }
  D2_HD void adjust_heap(int first, int hole, int len, int val) {
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
  D2_HD void heap_sort(int first, int last) {
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
  D2_HD void linear_insert(int last) {
    int v = a[last], next = last - 1;
    while (lt(v, a[next])) {
      a[last] = a[next];
      last = next;
      --next;
    }
    a[last] = v;
  }
  D2_HD void insertion(int first, int last) {
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
  D2_HD void sort(int first, int last) {
    if (first == last) return;
    const int n = last - first;
    int lg = 0;
    while ((n >> (lg + 1)) != 0) ++lg;
    int st_f[64], st_l[64], st_d[64], sp = 0;
    st_f[sp] = first, st_l[sp] = last, st_d[sp] = 2 * lg, ++sp;
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
        st_f[sp] = cut, st_l[sp] = l, st_d[sp] = d, ++sp;  // the right part, on its own
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

// Integer keys per element id (ranks, equal exactly when the keys tie).
struct RankLess {
  const int* k;
  D2_HD bool operator()(int x, int y) const { return k[x] < k[y]; }
};

}  // namespace d2
