// d2_solver.hpp -- sequential impulses of one cluster (host/device code).
//
// strategy::SequentialImpulses::solve (strategies/contact_force/
// sequential_impulses.cpp:104-638) over a cluster's filtered contacts, as one
// sequential walk (the Gauss-Seidel order is the reference's, so the result
// is its result): setup of the local contacts (:109-163), impulses from
// gravity (:171-180, contact_force_strategy.cpp:10-13), warm start
// (sequential_impulses_warm_start.cpp:12-29) and its impulses (:185-213),
// the first prediction (:215-227), the normal pass (:231-375, at most 1000
// sweeps), the friction pass (:392-571, at most 100 sweeps), the new future
// states (:624-626) and the warm-store update (warm_start.cpp:31-54).
// Not computed, because nothing observable depends on them: the friction
// pass's drift terms (:432-438, :455-503; they feed only an offset impulse
// the reference never applies), unused d1 values and contact_draws.
//
// The includer defines D2_HD and includes d2_geom.hpp.  Every array is the
// cluster's own slice (one thread per cluster on the device; the same code
// runs on the host for checks).  Evaluation orders match the oracle's
// restatement (dot left to right, M*v row by row, Hamilton products).
#pragma once

namespace d2 {

// The warm store: open addressing over 64-bit keys (lower body id << 32 |
// upper body id), EMPTY = ~0; values are the stored impulses.  Keys are
// never removed, so a lookup that meets an empty slot has missed.
constexpr unsigned long long WARM_EMPTY = ~0ull;
struct WarmTable {
  unsigned long long* keys;
  double* vals;
  unsigned long long mask;   // capacity - 1 (a power of two)
  unsigned long long* used;  // number of keys
};
D2_HD unsigned long long warm_key(int a, int b) {
  const unsigned lo = (unsigned)(a < b ? a : b), hi = (unsigned)(a < b ? b : a);
  return ((unsigned long long)lo << 32) | hi;
}
D2_HD unsigned long long warm_hash(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  return k;
}
D2_HD bool warm_find(const WarmTable& W, unsigned long long key, double& v) {
  for (unsigned long long h = warm_hash(key) & W.mask;; h = (h + 1) & W.mask) {
    const unsigned long long k = W.keys[h];
    if (k == key) {
      v = W.vals[h];
      return true;
    }
    if (k == WARM_EMPTY) return false;
  }
}
// Insert or overwrite (the caller keeps the load factor below one half).
D2_HD void warm_put(const WarmTable& W, unsigned long long key, double v) {
  for (unsigned long long h = warm_hash(key) & W.mask;; h = (h + 1) & W.mask) {
    unsigned long long k = W.keys[h];
    if (k == WARM_EMPTY) {
#if defined(__HIP_DEVICE_COMPILE__)
      k = atomicCAS(W.keys + h, WARM_EMPTY, key);
      if (k == WARM_EMPTY) atomicAdd(W.used, 1ull);
#else
      W.keys[h] = key;
      ++*W.used;
#endif
      if (k == WARM_EMPTY || k == key) {
        W.vals[h] = v;
        return;
      }
    } else if (k == key) {
      W.vals[h] = v;
      return;
    }
  }
}

struct SolverState {
  quat q;
  v3<double> x, v, L;
};
D2_HD SolverState sv_load(const double* r) {
  return SolverState{quat{r[0], r[1], r[2], r[3]}, {r[4], r[5], r[6]}, {r[7], r[8], r[9]}, {r[10], r[11], r[12]}};
}
D2_HD v3<double> sv_neg(v3<double> a) { return {-a.x, -a.y, -a.z}; }
// v^T M v with the row vector v^T M formed first (rn.transpose() * M * rn).
D2_HD double sv_quad(v3<double> v, const double* M) {
  const v3<double> u{(v.x * M[0] + v.y * M[3]) + v.z * M[6], (v.x * M[1] + v.y * M[4]) + v.z * M[7],
                     (v.x * M[2] + v.y * M[5]) + v.z * M[8]};
  return dot(u, v);
}
// model/forces.h:101-108: (-force) x (point - centre of mass)
D2_HD v3<double> sv_torque(v3<double> f, v3<double> p, v3<double> com) { return cross(sv_neg(f), sub(p, com)); }
D2_HD v3<double> sv_normalized(v3<double> a) {
  const double z = dot(a, a);
  return z > 0.0 ? dvs(a, sqrt(z)) : a;
}
// transform(p, pose(q, x)): the homogeneous transform with the divide by w.
D2_HD v3<double> sv_xform(const SolverState& s, v3<double> p) {
  double m[16];
  pose(s.q, s.x, m);
  return xform(m, p);
}
// sequential_impulses.cpp:46-70 (the future state's own L drives the turn)
D2_HD SolverState sv_update(const SolverState& f, double t, v3<double> imp, v3<double> off, v3<double> rimp,
                                   v3<double> roff, double mass, const double* ii) {
  SolverState u;
  u.v = add(f.v, dvs(imp, mass));
  u.x = add(f.x, scl(dvs(off, mass), t));
  u.L = add(f.L, rimp);
  double Iw[9];
  world_inv_inertia(f.q, ii, Iw);
  const v3<double> om = mul33v(Iw, add(f.L, sclL(t, roff)));
  const v3<double> d = scl(sclL(0.5, om), t);
  quat e{1.0, 0.0, 0.0, 0.0};
  if (!(vlen(d) == 0.0)) {
    const double n = vlen(d);
    const v3<double> s = scl(dvs(d, n), (double)sin(n));
    e = quat{(double)cos(n), s.x, s.y, s.z};
  }
  u.q = qmul(e, f.q);
  return u;
}
D2_HD double sv_max(double a, double b) { return (a < b) ? b : a; }  // std::max
D2_HD double sv_min(double a, double b) { return (b < a) ? b : a; }  // std::min

// One cluster.  Inputs per hit h (the cluster's filtered contacts in gathered
// order): A, B (3 each), eps_a + eps_b and the two bodies' local indices;
// per local particle k: its body id.  `ord` receives the hits' sorted order
// (the caller sorts by proportion |A - B| / (eps_a + eps_b) with the std::sort
// replay and passes the permutation).  Scratch per hit and per particle is
// the caller's.  Returns false (nothing written) where the reference returns
// false; on success the particles' future states are written to their body
// records and out_h[11 * h] gets (normal, impulse, impulse offset, friction
// impulse, tangential velocity) for hit h (in gathered order).
struct ClusterSolve {
  // inputs
  int nh, np;
  const double* hA;    // [nh][3]
  const double* hB;    // [nh][3]
  const double* heps;  // [nh] eps_a + eps_b
  const int* hl;       // [nh][2] local particle indices of a and b
  const int* hpc;      // [nh] contacts of the hit's body pair in the cluster
  const int* ord;      // [nh] sorted order (hit ids)
  const int* pbody;    // [np] body ids
  double* bodies;      // body records (40 doubles)
  const double* mass;  // per body
  const double* fric;  // per body
  double t;
  v3<double> g;
  WarmTable warm;
  // scratch per hit (sorted position): 35 doubles
  double* hs;  // [nh][SH]
  // scratch per particle: 28 doubles
  double* ps;  // [np][SP]
  double* out_h;  // [nh][11]
};
// scratch layout
constexpr int SH_LA = 0, SH_LB = 3, SH_N = 6, SH_RA = 9, SH_RB = 12, SH_LF = 15, SH_LRA = 18, SH_LRB = 21,
              SH_UT = 24, SH_VT = 27, SH_START = 30, SH_MEFF = 31, SH_IMP = 32, SH_OFF = 33, SH_OUTER = 34, SH = 35;
constexpr int SP_Q = 0, SP_X = 4, SP_V = 7, SP_L = 10, SP_IMP = 13, SP_ROT = 16, SP_OFF = 19, SP_ROFF = 22,
              SP_LAST = 25, SP = 28;

D2_HD v3<double> ld3(const double* p) { return {p[0], p[1], p[2]}; }
D2_HD void st3(double* p, v3<double> v) { p[0] = v.x, p[1] = v.y, p[2] = v.z; }
D2_HD SolverState ps_state(const double* p) {
  return SolverState{quat{p[SP_Q], p[SP_Q + 1], p[SP_Q + 2], p[SP_Q + 3]}, ld3(p + SP_X), ld3(p + SP_V),
                     ld3(p + SP_L)};
}
D2_HD void ps_store(double* p, const SolverState& s) {
  p[SP_Q] = s.q.w, p[SP_Q + 1] = s.q.x, p[SP_Q + 2] = s.q.y, p[SP_Q + 3] = s.q.z;
  st3(p + SP_X, s.x), st3(p + SP_V, s.v), st3(p + SP_L, s.L);
}

D2_HD bool solve_cluster(const ClusterSolve& S) {
  const int nh = S.nh, np = S.np;
  const double t = S.t;
  // local contacts, in sorted order (:109-163)
  for (int c = 0; c < nh; ++c) {
    const int h = S.ord[c];
    const int la = S.hl[2 * h], lb = S.hl[2 * h + 1];
    const double* ra = S.bodies + 40 * (size_t)S.pbody[la];
    const double* rb = S.bodies + 40 * (size_t)S.pbody[lb];
    const v3<double> A = ld3(S.hA + 3 * h), B = ld3(S.hB + 3 * h);
    double* s = S.hs + (size_t)SH * c;
    s[SH_START] = vlen(sub(A, B));
    if (s[SH_START] < 1e-6 && t > 1e-4) return false;
    const SolverState fa = sv_load(ra + 14), fb = sv_load(rb + 14);
    double m[16], mi[16];
    pose(fa.q, fa.x, m);
    inv44(m, mi);
    st3(s + SH_LA, xform(mi, A));
    pose(fb.q, fb.x, m);
    inv44(m, mi);
    st3(s + SH_LB, xform(mi, B));
    const v3<double> gn = dvs(sub(B, A), 2.0);
    st3(s + SH_N, sv_normalized(gn));
    const v3<double> r_A = sub(A, fa.x), r_B = sub(B, fb.x);
    st3(s + SH_RA, r_A), st3(s + SH_RB, r_B);
    const v3<double> rn_A = cross(r_A, gn), rn_B = cross(r_B, gn);
    double k_normal = 0.0;
    if (ra[38] == 0.0) k_normal += 1.0 / S.mass[S.pbody[la]] + sv_quad(rn_A, ra + 28);
    if (rb[38] == 0.0) k_normal += 1.0 / S.mass[S.pbody[lb]] + sv_quad(rn_B, rb + 28);
    s[SH_MEFF] = k_normal < 1e-6 ? 0.0 : 1.0 / k_normal;
    s[SH_IMP] = 0.0, s[SH_OFF] = 0.0;
    s[SH_OUTER] = S.heps[h] * 0.25;
    for (int k = 0; k < 3; ++k) s[SH_LF + k] = 0.0, s[SH_LRA + k] = 0.0, s[SH_LRB + k] = 0.0, s[SH_UT + k] = 0.0;
  }
  // per particle: state, impulses from gravity (:171-180)
  for (int k = 0; k < np; ++k) {
    const int b = S.pbody[k];
    double* p = S.ps + (size_t)SP * k;
    ps_store(p, sv_load(S.bodies + 40 * (size_t)b + 14));
    st3(p + SP_IMP, scl(sclL(S.mass[b], S.g), t));
    for (int j = 0; j < 3; ++j) p[SP_ROT + j] = 0.0, p[SP_OFF + j] = 0.0, p[SP_ROFF + j] = 0.0;
  }
  // warm start (warm_start.cpp:12-29) and its impulses (:185-213)
  for (int c = 0; c < nh; ++c) {
    const int h = S.ord[c];
    double w;
    if (warm_find(S.warm, warm_key(S.pbody[S.hl[2 * h]], S.pbody[S.hl[2 * h + 1]]), w)) S.hs[(size_t)SH * c + SH_IMP] = w * 0.75;
  }
  for (int c = 0; c < nh; ++c) {
    const int h = S.ord[c];
    const double* s = S.hs + (size_t)SH * c;
    double* pa = S.ps + (size_t)SP * S.hl[2 * h];
    double* pb = S.ps + (size_t)SP * S.hl[2 * h + 1];
    const SolverState fa = ps_state(pa), fb = ps_state(pb);
    const v3<double> n = ld3(s + SH_N);
    const v3<double> p1A = sv_xform(fa, ld3(s + SH_LA)), p1B = sv_xform(fb, ld3(s + SH_LB));
    const v3<double> u = scl(sv_neg(n), s[SH_IMP]);
    st3(pa + SP_IMP, add(ld3(pa + SP_IMP), u));
    st3(pb + SP_IMP, sub(ld3(pb + SP_IMP), u));
    st3(pa + SP_ROT, add(ld3(pa + SP_ROT), sv_torque(u, p1A, fa.x)));
    st3(pb + SP_ROT, add(ld3(pb + SP_ROT), sv_torque(sv_neg(u), p1B, fb.x)));
  }
  auto integrate = [&](int k) {
    const int b = S.pbody[k];
    const double* r = S.bodies + 40 * (size_t)b;
    double* p = S.ps + (size_t)SP * k;
    ps_store(p, sv_update(sv_load(r + 14), t, ld3(p + SP_IMP), ld3(p + SP_OFF), ld3(p + SP_ROT), ld3(p + SP_ROFF),
                          S.mass[b], r + 28));
  };
  auto is_static = [&](int k) { return S.bodies[40 * (size_t)S.pbody[k] + 38] != 0.0; };
  for (int k = 0; k < np; ++k)  // first prediction (:215-227)
    if (!is_static(k)) integrate(k);
  if (nh > 0) {
    // normal pass (:231-375)
    bool converged = false;
    for (int it = 0; !converged && it < 1000; ++it) {
      for (int k = 0; k < np; ++k) {
        double* p = S.ps + (size_t)SP * k;
        for (int j = 0; j < 3; ++j) p[SP_LAST + j] = p[SP_IMP + j];
      }
      for (int c = 0; c < nh; ++c) {
        const int h = S.ord[c];
        double* s = S.hs + (size_t)SH * c;
        const int la = S.hl[2 * h], lb = S.hl[2 * h + 1];
        double* pa = S.ps + (size_t)SP * la;
        double* pb = S.ps + (size_t)SP * lb;
        const double* ra = S.bodies + 40 * (size_t)S.pbody[la];
        const double* rb = S.bodies + 40 * (size_t)S.pbody[lb];
        const SolverState fa = ps_state(pa), fb = ps_state(pb);
        const v3<double> n = ld3(s + SH_N);
        const double outer = s[SH_OUTER];
        const v3<double> p1A = sv_xform(fa, ld3(s + SH_LA)), p1B = sv_xform(fb, ld3(s + SH_LB));
        const double d1 = dot(p1B, n) - dot(p1A, n);
        const v3<double> v1A = add(fa.v, cross(mul33v(ra + 28, fa.L), sub(p1A, fa.x)));
        const v3<double> v1B = add(fb.v, cross(mul33v(rb + 28, fb.L), sub(p1B, fb.x)));
        const double v1 = dot(v1A, n) - dot(v1B, n);
        const double m = s[SH_MEFF];
        const double damp = 0.1;
        const double dimp = d1 < outer ? (v1 * m) * damp : -s[SH_IMP];
        const double last = s[SH_IMP];
        s[SH_IMP] = sv_max(s[SH_IMP] + dimp, 0.0);
        const v3<double> u = scl(sv_neg(n), s[SH_IMP] - last);
        st3(pa + SP_IMP, add(ld3(pa + SP_IMP), u));
        st3(pb + SP_IMP, sub(ld3(pb + SP_IMP), u));
        st3(pa + SP_ROT, add(ld3(pa + SP_ROT), sv_torque(u, p1A, fa.x)));
        st3(pb + SP_ROT, add(ld3(pb + SP_ROT), sv_torque(sv_neg(u), p1B, fb.x)));
        const double lower_slop = sv_max(outer * 0.2, outer - s[SH_START]);
        const double slop = lerp(lower_slop, outer * 0.2, 0.3);
        const double pen = outer - d1;
        const double doff = d1 < outer ? ((damp * m) * sv_max(0.0, pen - slop)) / t : -s[SH_OFF];
        const double last_off = s[SH_OFF];
        s[SH_OFF] = sv_max(s[SH_OFF] + doff, 0.0);
        const v3<double> uo = scl(sv_neg(n), s[SH_OFF] - last_off);
        st3(pa + SP_OFF, add(ld3(pa + SP_OFF), uo));
        st3(pb + SP_OFF, sub(ld3(pb + SP_OFF), uo));
        if (ra[38] == 0.0) integrate(la);
        if (rb[38] == 0.0) integrate(lb);
      }
      converged = true;
      for (int k = 0; k < np && converged; ++k) {
        const double* p = S.ps + (size_t)SP * k;
        if (vlen(sub(ld3(p + SP_IMP), ld3(p + SP_LAST))) > 1e-6) converged = false;
      }
    }
    // friction pass (:392-571); contacts_per_pair is the hit's pair's count (hpc)
    converged = false;
    for (int it = 0; !converged && it < 100; ++it) {
      for (int c = 0; c < nh; ++c) {
        const int h = S.ord[c];
        double* s = S.hs + (size_t)SH * c;
        const int la = S.hl[2 * h], lb = S.hl[2 * h + 1];
        const double* ra = S.bodies + 40 * (size_t)S.pbody[la];
        const double* rb = S.bodies + 40 * (size_t)S.pbody[lb];
        const SolverState fa = ps_state(S.ps + (size_t)SP * la), fb = ps_state(S.ps + (size_t)SP * lb);
        const v3<double> n = ld3(s + SH_N);
        const v3<double> p1A = sv_xform(fa, ld3(s + SH_LA)), p1B = sv_xform(fb, ld3(s + SH_LB));
        const bool sa = ra[38] != 0.0, sb = rb[38] != 0.0;
        const v3<double> v1A = sa ? v3<double>{0.0, 0.0, 0.0} : add(fa.v, cross(mul33v(ra + 28, fa.L), sub(p1A, fa.x)));
        const v3<double> v1B = sb ? v3<double>{0.0, 0.0, 0.0} : add(fb.v, cross(mul33v(rb + 28, fb.L), sub(p1B, fb.x)));
        const v3<double> v1AB = sub(v1A, v1B);
        const v3<double> vt = sub(v1AB, sclL(dot(v1AB, n), n));
        st3(s + SH_VT, vt);
        const double mu = sv_min(S.fric[S.pbody[la]], S.fric[S.pbody[lb]]);
        const double max_force = (s[SH_IMP] * mu) * 10;
        const v3<double> tan = sv_normalized(vt);
        const v3<double> rtA = cross(ld3(s + SH_RA), tan), rtB = cross(ld3(s + SH_RB), tan);
        double kt = 0.0;
        if (!sa) kt += 1.0 / S.mass[S.pbody[la]] + sv_quad(rtA, ra + 28);
        if (!sb) kt += 1.0 / S.mass[S.pbody[lb]] + sv_quad(rtB, rb + 28);
        const double mt = kt < 1e-6 ? 0.0 : 1.0 / kt;
        const v3<double> delta = scl(scl(sv_neg(vt), mt), 0.9);
        v3<double> fit = add(ld3(s + SH_LF), delta);
        const double len = vlen(fit);
        const double cl = (len < 0.0) ? 0.0 : (max_force < len) ? max_force : len;  // std::clamp(len, 0, max)
        fit = scl(sv_normalized(fit), cl);
        st3(s + SH_UT, fit);
      }
      converged = true;
      for (int c = 0; c < nh; ++c) {
        const int h = S.ord[c];
        double* s = S.hs + (size_t)SH * c;
        const int la = S.hl[2 * h], lb = S.hl[2 * h + 1];
        double* pa = S.ps + (size_t)SP * la;
        double* pb = S.ps + (size_t)SP * lb;
        const SolverState fa = ps_state(pa), fb = ps_state(pb);
        const double cpp = (double)S.hpc[h];
        const v3<double> p1A = sv_xform(fa, ld3(s + SH_LA)), p1B = sv_xform(fb, ld3(s + SH_LB));
        const v3<double> u = dvs(sub(ld3(s + SH_UT), ld3(s + SH_LF)), cpp);
        st3(s + SH_LF, ld3(s + SH_UT));
        st3(pa + SP_IMP, add(ld3(pa + SP_IMP), u));
        st3(pb + SP_IMP, sub(ld3(pb + SP_IMP), u));
        if (vlen(u) > 1e-4) converged = false;
        const v3<double> lf = ld3(s + SH_LF);
        const v3<double> rA = sv_torque(lf, p1A, fa.x), rB = sv_torque(sv_neg(lf), p1B, fb.x);
        st3(pa + SP_ROT, sub(ld3(pa + SP_ROT), ld3(s + SH_LRA)));
        st3(pb + SP_ROT, sub(ld3(pb + SP_ROT), ld3(s + SH_LRB)));
        st3(s + SH_LRA, rA), st3(s + SH_LRB, rB);
        st3(pa + SP_ROT, add(ld3(pa + SP_ROT), rA));
        st3(pb + SP_ROT, add(ld3(pb + SP_ROT), rB));
        // the offset friction totals stay zero: (0 - 0) / count
        const v3<double> uo = dvs(sub(v3<double>{0.0, 0.0, 0.0}, v3<double>{0.0, 0.0, 0.0}), cpp);
        st3(pa + SP_OFF, add(ld3(pa + SP_OFF), uo));
        st3(pb + SP_OFF, sub(ld3(pb + SP_OFF), uo));
        st3(pa + SP_ROFF, add(ld3(pa + SP_ROFF), sv_torque(uo, p1A, fa.x)));
        st3(pb + SP_ROFF, add(ld3(pb + SP_ROFF), sv_torque(sv_neg(uo), p1B, fb.x)));
      }
      for (int k = 0; k < np; ++k)
        if (!is_static(k)) integrate(k);
    }
  }
  // the new future states (:624-626; a static body's stays as it was)
  for (int k = 0; k < np; ++k) {
    if (is_static(k)) continue;
    const double* p = S.ps + (size_t)SP * k;
    double* r = S.bodies + 40 * (size_t)S.pbody[k] + 14;
    for (int j = 0; j < 13; ++j) r[j] = p[j];
  }
  // warm store: each body pair keeps its last contact's impulse
  for (int c = 0; c < nh; ++c) {
    const int h = S.ord[c];
    warm_put(S.warm, warm_key(S.pbody[S.hl[2 * h]], S.pbody[S.hl[2 * h + 1]]), S.hs[(size_t)SH * c + SH_IMP]);
  }
  for (int c = 0; c < nh; ++c) {
    const int h = S.ord[c];
    const double* s = S.hs + (size_t)SH * c;
    double* o = S.out_h + 11 * (size_t)h;
    o[0] = s[SH_N], o[1] = s[SH_N + 1], o[2] = s[SH_N + 2], o[3] = s[SH_IMP], o[4] = s[SH_OFF];
    for (int j = 0; j < 3; ++j) o[5 + j] = s[SH_LF + j], o[8 + j] = s[SH_VT + j];
  }
  return true;
}

}  // namespace d2
