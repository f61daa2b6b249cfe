// CPU build of the device solver (delta2_amd/csrc/d2_solver.hpp and the
// std::sort replay of d2_sort.hpp are __host__ __device__), as a shared
// library for tests/test_solver_cpu.py, which compares it with the oracle.
// One call solves one cluster exactly as the device's per-cluster thread
// does after gathering: sort the hits by proportion with the replay, then
// solve_cluster.
#include <cmath>
#include <cstdint>
#include <vector>

#define D2_HD inline
#include <algorithm>
using std::cos;
using std::fabs;
using std::sin;
using std::sqrt;
#include "../delta2_amd/csrc/d2_geom.hpp"
#include "../delta2_amd/csrc/d2_sort.hpp"
#include "../delta2_amd/csrc/d2_solver.hpp"

struct HostKeyLess {
  const double* k;
  bool operator()(int x, int y) const { return k[x] < k[y]; }
};

extern "C" int solver_host(int nh, const double* hA, const double* hB, const double* heps, const int* hl,
                           const int* hpc, int np,
                           const int* pbody, double* bodies, const double* mass, const double* fric, double t,
                           const double* g, unsigned long long* wkeys, double* wvals, unsigned long long wcap,
                           unsigned long long* wused, double* out_h) {
  std::vector<double> key(nh > 0 ? nh : 1);
  std::vector<int> ord(nh > 0 ? nh : 1);
  for (int h = 0; h < nh; ++h) {
    d2::v3<double> A{hA[3 * h], hA[3 * h + 1], hA[3 * h + 2]}, B{hB[3 * h], hB[3 * h + 1], hB[3 * h + 2]};
    key[h] = d2::vlen(d2::sub(A, B)) / heps[h];
    ord[h] = h;
  }
  d2::StdSort<HostKeyLess>{ord.data(), HostKeyLess{key.data()}}.sort(0, nh);
  std::vector<double> hs((size_t)d2::SH * (nh > 0 ? nh : 1)), ps((size_t)d2::SP * (np > 0 ? np : 1));
  d2::ClusterSolve S;
  S.nh = nh, S.np = np, S.hA = hA, S.hB = hB, S.heps = heps, S.hl = hl, S.hpc = hpc, S.ord = ord.data();
  S.pbody = pbody;
  S.bodies = bodies, S.mass = mass, S.fric = fric, S.t = t, S.g = d2::v3<double>{g[0], g[1], g[2]};
  S.warm = d2::WarmTable{wkeys, wvals, wcap - 1, wused};
  S.hs = hs.data(), S.ps = ps.data(), S.out_h = out_h;
  return d2::solve_cluster(S) ? 1 : 0;
}
