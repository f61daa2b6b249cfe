// d2o_solver.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Delta2's sequential-impulses contact solver of one cluster, restated
// without Eigen: strategy::SequentialImpulses::solve
// (strategies/contact_force/sequential_impulses.cpp:104-638) with its state
// updates (:46-96), the warm start (sequential_impulses_warm_start.cpp:
// 12-54), the contact order (collision_detection/contact.h:14-18), the point
// velocity (model/state.cpp:81-84) and the torque (model/forces.h:101-108).
// Evaluation orders follow d2o_math.hpp.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load the oracle.
//
// Left out because nothing observable depends on them: the "drift" terms of
// the friction pass (:432-438, :455-503; FStates_friction, updateStateImmediate
// :72-96) feed only delta_friction_impulse_offset, which the reference never
// applies (updated_friction_total_offset stays zero), the unused d1 of the
// warm-start pass, the tangent mass of the setup pass, and contact_draws.
#pragma once
#include <algorithm>
#include <map>
#include <vector>

#include "d2o_tree.hpp"

namespace d2o {

// strategies/contact_force/LocalContact.h:6-23
struct LocalContact {
  V local_A, local_B, global_centre, global_normal, global_tangent;
  double mass_eff_normal = 0, mass_eff_tangent = 0, start_dist = 0, impulse = 0, impulse_offset = 0;
  V last_friction_impulse, last_friction_rotational_impulse_A, last_friction_rotational_impulse_B;
};

// A hit as the solver sees it: the contact and its two bodies (scene ids).
struct SolverHit {
  Contact ct;
  int a = 0, b = 0;
  int row = 0;  // position in the caller's list (results are reported there)
};

// What the solver reads of a particle besides its states (model/particle.h).
struct SolverBody {
  double mass = 1.0, friction = 0.0;
};

// Per-hit results: the contact normal n = normalized((B - A) / 2), the normal
// impulse and its position offset, the friction impulse and the tangential
// relative velocity v - (v.n) n of the last friction pass.
struct SolverOut {
  V normal, friction, v_tangent;
  double impulse = 0, impulse_offset = 0;
};

// Vector3d(0, 0, 0).cross-free helpers in the reference's expression order.
inline V neg(const V& a) { return V(-a.x, -a.y, -a.z); }
// v^T M v as Eigen evaluates rn.transpose() * M * rn: the row vector first.
inline double quad(const V& v, const M3<double>& M) {
  V u((v.x * M(0, 0) + v.y * M(1, 0)) + v.z * M(2, 0), (v.x * M(0, 1) + v.y * M(1, 1)) + v.z * M(2, 1),
      (v.x * M(0, 2) + v.y * M(1, 2)) + v.z * M(2, 2));
  return dot(u, v);
}
// model/forces.h:101-108
inline V calcTorque(const V& force, const V& point, const V& com) { return cross(neg(force), point - com); }
// model/state.cpp:81-84
inline V pointVelocity(const State& s, const V& pt, const M3<double>& inv_inertia) {
  V w = mul(inv_inertia, s.L);
  return s.vel + cross(w, pt - s.trans);
}
// sequential_impulses.cpp:46-70
inline State updateState(const State& future, double t, const V& impulse, const V& impulse_offset, const V& rot_impulse,
                         const V& rot_offset, double mass, const M3<double>& ii) {
  State u = future;
  u.vel = future.vel + impulse / mass;
  u.trans = future.trans + (impulse_offset / mass) * t;
  u.L = future.L + rot_impulse;
  M3<double> R = toRotation(future.rot);
  M3<double> Iinv = mul(mul(R, ii), transpose(R));
  V omega = mul(Iinv, future.L + t * rot_offset);
  Quat<double> d = qexp((0.5 * omega) * t);
  u.rot = qmul(d, future.rot);
  return u;
}
// std::max / std::min / std::clamp as libstdc++ defines them (NaN handling).
inline double smax(double a, double b) { return (a < b) ? b : a; }
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double sclamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// The warm-start store (sequential_impulses_warm_start.h): impulse per body
// pair (lower id, upper id), kept across solves.
using WarmStore = std::map<std::pair<int, int>, double>;

// SequentialImpulses::solve for one cluster.  `cluster` lists the cluster's
// bodies (any order: every per-body quantity is independent of it); `hits`
// are the cluster's filtered contacts in the order the caller gathered them
// (PDE_explicit.cpp:77-95).  Returns false (and changes nothing) where the
// reference returns false.  On success the bodies' future states are
// replaced, the store updated, and out[h.row] filled for every hit.
inline bool solveCluster(std::vector<Body>& bodies, const std::vector<SolverBody>& props,
                         const std::vector<int>& cluster, std::vector<SolverHit> hits, double t, const V& g,
                         WarmStore& store, std::vector<SolverOut>& out) {
  // std::sort(hits) by Contact::operator< (contact.h:14-18)
  std::sort(hits.begin(), hits.end(), [](const SolverHit& x, const SolverHit& y) {
    double px = norm(x.ct.A - x.ct.B) / (x.ct.eps_a + x.ct.eps_b);
    double py = norm(y.ct.A - y.ct.B) / (y.ct.eps_a + y.ct.eps_b);
    return px < py;
  });
  const int nh = (int)hits.size(), np = (int)cluster.size();
  std::map<int, int> local;  // cluster.particles.getLocalID
  for (int i = 0; i < np; ++i) local[cluster[i]] = i;
  std::vector<LocalContact> contacts(nh);
  for (int c = 0; c < nh; ++c) {  // :109-163
    const Contact& h = hits[c].ct;
    const Body& pa = bodies[hits[c].a];
    const Body& pb = bodies[hits[c].b];
    LocalContact& lc = contacts[c];
    lc.start_dist = norm(h.A - h.B);
    if (lc.start_dist < 1e-6 && t > 1e-4) return false;
    lc.local_A = transform(h.A, inverse(pa.future.transformation()));
    lc.local_B = transform(h.B, inverse(pb.future.transformation()));
    lc.global_centre = (h.A + h.B) / 2.0;
    lc.global_normal = (h.B - h.A) / 2.0;
    lc.global_tangent = V(0, 0, 0);
    V r_A = h.A - pa.future.trans, r_B = h.B - pb.future.trans;
    V rn_A = cross(r_A, lc.global_normal), rn_B = cross(r_B, lc.global_normal);
    double im_A = 1.0 / props[hits[c].a].mass, im_B = 1.0 / props[hits[c].b].mass;
    double k_normal = 0.0;
    if (!pa.is_static) k_normal += im_A + quad(rn_A, pa.inv_inertia);
    if (!pb.is_static) k_normal += im_B + quad(rn_B, pb.inv_inertia);
    lc.mass_eff_normal = k_normal < 1e-6 ? 0.0 : 1.0 / k_normal;
  }
  std::vector<V> impulses(np), rot(np), offs(np), rotoffs(np);
  std::vector<State> FS(np);
  for (int p = 0; p < np; ++p) {  // :171-180, external_force = m g (contact_force_strategy.cpp:10-13)
    const Body& b = bodies[cluster[p]];
    impulses[p] = (props[cluster[p]].mass * g) * t;
    FS[p] = b.future;
  }
  auto key_of = [](int a, int b) { return std::make_pair(std::min(a, b), std::max(a, b)); };
  for (int c = 0; c < nh; ++c) {  // _warm_start.apply (warm_start.cpp:12-29)
    auto it = store.find(key_of(hits[c].a, hits[c].b));
    if (it != store.end()) contacts[c].impulse = it->second * 0.75;
  }
  for (int c = 0; c < nh; ++c) {  // :185-213
    int a = local[hits[c].a], b = local[hits[c].b];
    V n = normalized(contacts[c].global_normal);
    V p1_A = transform(contacts[c].local_A, FS[a].transformation());
    V p1_B = transform(contacts[c].local_B, FS[b].transformation());
    V u = neg(n) * contacts[c].impulse;
    impulses[a] = impulses[a] + u;
    impulses[b] = impulses[b] - u;
    rot[a] = rot[a] + calcTorque(u, p1_A, FS[a].trans);
    rot[b] = rot[b] + calcTorque(neg(u), p1_B, FS[b].trans);
  }
  auto integrate = [&](int p) {
    const Body& b = bodies[cluster[p]];
    double m = props[cluster[p]].mass;
    FS[p] = updateState(b.future, t, impulses[p], offs[p], rot[p], rotoffs[p], m, b.inv_inertia);
  };
  for (int p = 0; p < np; ++p)  // :215-227
    if (!bodies[cluster[p]].is_static) integrate(p);
  std::vector<V> v_tangents(nh);
  if (nh > 0) {
    bool converged = false;
    int it = 0;
    while (!converged && it < 1000) {  // :231-375
      std::vector<V> last_impulse = impulses;
      for (int c = 0; c < nh; ++c) {
        int a = local[hits[c].a], b = local[hits[c].b];
        const Body& pa = bodies[hits[c].a];
        const Body& pb = bodies[hits[c].b];
        LocalContact& lc = contacts[c];
        V n = normalized(lc.global_normal);
        const double outer = (hits[c].ct.eps_a + hits[c].ct.eps_b) * 0.25;
        V p1_A = transform(lc.local_A, FS[a].transformation());
        V p1_B = transform(lc.local_B, FS[b].transformation());
        const double d1 = dot(p1_B, n) - dot(p1_A, n);
        V v1_A = pointVelocity(FS[a], p1_A, pa.inv_inertia);
        V v1_B = pointVelocity(FS[b], p1_B, pb.inv_inertia);
        const double v1 = dot(v1_A, n) - dot(v1_B, n);
        const double m = lc.mass_eff_normal;
        const double delta_damp = 0.1;
        const double delta_impulse_mag = d1 < outer ? (v1 * m) * delta_damp : -lc.impulse;
        double last_mag = lc.impulse;
        lc.impulse = smax(lc.impulse + delta_impulse_mag, 0.0);
        V u = neg(n) * (lc.impulse - last_mag);
        impulses[a] = impulses[a] + u;
        impulses[b] = impulses[b] - u;
        rot[a] = rot[a] + calcTorque(u, p1_A, FS[a].trans);
        rot[b] = rot[b] + calcTorque(neg(u), p1_B, FS[b].trans);
        const double min_slop_factor = 0.2;
        const double lower_slop = smax(outer * min_slop_factor, outer - lc.start_dist);
        const double slop = lerp(lower_slop, outer * min_slop_factor, 0.3);
        const double penetration_depth = outer - d1;
        const double delta_off_mag =
            d1 < outer ? ((delta_damp * m) * smax(0.0, penetration_depth - slop)) / t : -lc.impulse_offset;
        double last_off = lc.impulse_offset;
        lc.impulse_offset = smax(lc.impulse_offset + delta_off_mag, 0.0);
        V uo = neg(n) * (lc.impulse_offset - last_off);
        offs[a] = offs[a] + uo;
        offs[b] = offs[b] - uo;
        if (!bodies[cluster[a]].is_static) integrate(a);
        if (!bodies[cluster[b]].is_static) integrate(b);
      }
      converged = true;
      for (int p = 0; p < np; ++p)
        if (norm(impulses[p] - last_impulse[p]) > 1e-6) {
          converged = false;
          break;
        }
      ++it;
    }
    std::vector<V> updated_total(nh), updated_total_last_offset(nh), updated_total_offset(nh);
    int it2 = 0;
    converged = false;
    while (!converged && it2 < 100) {  // :392-571
      std::map<std::pair<int, int>, int> contacts_per_pair;
      for (int c = 0; c < nh; ++c) {
        int a = local[hits[c].a], b = local[hits[c].b];
        const Body& pa = bodies[hits[c].a];
        const Body& pb = bodies[hits[c].b];
        LocalContact& lc = contacts[c];
        V n = normalized(lc.global_normal);
        V p1_A = transform(lc.local_A, FS[a].transformation());
        V p1_B = transform(lc.local_B, FS[b].transformation());
        V v1_A = pa.is_static ? V(0, 0, 0) : pointVelocity(FS[a], p1_A, pa.inv_inertia);
        V v1_B = pb.is_static ? V(0, 0, 0) : pointVelocity(FS[b], p1_B, pb.inv_inertia);
        V v1_AB = v1_A - v1_B;
        V v_tangent = v1_AB - dot(v1_AB, n) * n;
        v_tangents[c] = v_tangent;
        double mu = smin(props[hits[c].a].friction, props[hits[c].b].friction);
        double max_force = (lc.impulse * mu) * 10;
        V r_A = hits[c].ct.A - pa.future.trans, r_B = hits[c].ct.B - pb.future.trans;
        V tangent = normalized(v_tangent);
        V rt_A = cross(r_A, tangent), rt_B = cross(r_B, tangent);
        double im_A = 1.0 / props[hits[c].a].mass, im_B = 1.0 / props[hits[c].b].mass;
        double k_tangent = 0.0;
        if (!pa.is_static) k_tangent += im_A + quad(rt_A, pa.inv_inertia);
        if (!pb.is_static) k_tangent += im_B + quad(rt_B, pb.inv_inertia);
        double m_inv_t = k_tangent < 1e-6 ? 0.0 : 1.0 / k_tangent;
        V delta = (neg(v_tangent) * m_inv_t) * 0.9;
        V fit = lc.last_friction_impulse + delta;
        fit = normalized(fit) * sclamp(norm(fit), 0.0, max_force);
        updated_total[c] = fit;
        contacts_per_pair[std::make_pair(std::min(a, b), std::max(a, b))]++;
      }
      converged = true;
      for (int c = 0; c < nh; ++c) {
        int a = local[hits[c].a], b = local[hits[c].b];
        LocalContact& lc = contacts[c];
        const double cpp = (double)contacts_per_pair[std::make_pair(std::min(a, b), std::max(a, b))];
        V p1_A = transform(lc.local_A, FS[a].transformation());
        V p1_B = transform(lc.local_B, FS[b].transformation());
        V u = (updated_total[c] - lc.last_friction_impulse) / cpp;
        lc.last_friction_impulse = updated_total[c];
        impulses[a] = impulses[a] + u;
        impulses[b] = impulses[b] - u;
        if (norm(u) > 1e-4) converged = false;
        V rA = calcTorque(lc.last_friction_impulse, p1_A, FS[a].trans);
        V rB = calcTorque(neg(lc.last_friction_impulse), p1_B, FS[b].trans);
        rot[a] = rot[a] - lc.last_friction_rotational_impulse_A;
        rot[b] = rot[b] - lc.last_friction_rotational_impulse_B;
        lc.last_friction_rotational_impulse_A = rA;
        lc.last_friction_rotational_impulse_B = rB;
        rot[a] = rot[a] + rA;
        rot[b] = rot[b] + rB;
        V uo = (updated_total_offset[c] - updated_total_last_offset[c]) / cpp;
        updated_total_last_offset[c] = updated_total_offset[c];
        offs[a] = offs[a] + uo;
        offs[b] = offs[b] - uo;
        if (norm(uo) > 1e-4) converged = false;
        rotoffs[a] = rotoffs[a] + calcTorque(uo, p1_A, FS[a].trans);
        rotoffs[b] = rotoffs[b] + calcTorque(neg(uo), p1_B, FS[b].trans);
      }
      for (int p = 0; p < np; ++p)
        if (!bodies[cluster[p]].is_static) integrate(p);
      ++it2;
    }
  }
  for (int p = 0; p < np; ++p) bodies[cluster[p]].future = FS[p];  // :624-626
  // _warm_start.update (warm_start.cpp:31-54): the found/not-found branches
  // both leave the key's last contact impulse with count 1
  std::map<std::pair<int, int>, std::pair<double, int>> acc;
  for (int c = 0; c < nh; ++c) {
    auto key = key_of(hits[c].a, hits[c].b);
    if (acc.find(key) != acc.end()) {
      acc[key] = std::make_pair(contacts[c].impulse, 1);
    } else {
      acc[key].first += contacts[c].impulse;
      acc[key].second++;
    }
  }
  for (auto& kv : acc) store[kv.first] = kv.second.first / kv.second.second;
  for (int c = 0; c < nh; ++c) {
    SolverOut& o = out[hits[c].row];
    o.normal = normalized(contacts[c].global_normal);
    o.impulse = contacts[c].impulse;
    o.impulse_offset = contacts[c].impulse_offset;
    o.friction = contacts[c].last_friction_impulse;
    o.v_tangent = v_tangents[c];
  }
  return true;
}

}  // namespace d2o
