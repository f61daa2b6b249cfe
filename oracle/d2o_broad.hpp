// d2o_broad.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// The broad phase and the cluster separation that feed the contact-detection
// path, restated without Embree, Eigen or libigl, by brute force over every
// body pair:
//   * broadPhaseEmbree(particles, t)  (collision_detection/broad_phase_embree.cpp:248-268,
//     boundsFunc :60-84, collideFunc :143-176)
//   * broadPhaseEmbree(particles)     (:270-289, boundsFuncIndividualTime :86-112,
//     collideFuncIndividualTime :178-226)
//   * separateCollisionClusters       (collision_detection/separate_clusters.cpp:355-459)
// Third-party pieces, restated from their published behaviour:
//   * Embree 3.13.2 rtcCollide (src/Makefile:30-32): reports every pair of
//     user primitives whose float bounding boxes overlap, touching boxes
//     included (closed intervals); the order of its reports is not
//     deterministic, so the pair list here is sorted by (a, b).
//   * Embree BBox3fa: extend = componentwise min/max, enlarge_by(a) =
//     [lower - a, upper + a], all in float.
//   * libigl connected_components (separate_clusters.cpp:375): a breadth-first
//     search from every unvisited vertex in index order, components numbered
//     in that order.
//   * Eigen Quaternion::slerp (state.cpp:68): scale0 = sin((1-t) theta) /
//     sin(theta), scale1 = sin(t theta) / sin(theta) with theta = acos|d|,
//     linear weights when |d| >= 1 - epsilon, scale1 negated when d < 0.
// Groups (several independent scenes in one call) have no reference
// counterpart: bodies of different groups never pair.
#pragma once
#include <algorithm>
#include <cmath>
#include <limits>
#include <queue>
#include <stdexcept>
#include <vector>

#include "d2o_tree.hpp"

namespace d2o {

// embree::BBox3fa over its three used components.
struct Box3f {
  float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                 std::numeric_limits<float>::infinity()};
  float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                 -std::numeric_limits<float>::infinity()};
  void extend(const Box3f& b) {
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], b.lo[k]), hi[k] = std::max(hi[k], b.hi[k]);
  }
  void extend(const float v[3]) {
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], v[k]), hi[k] = std::max(hi[k], v[k]);
  }
  Box3f translated(const float v[3]) const {  // broad_phase_embree.cpp:53-58
    Box3f r = *this;
    for (int k = 0; k < 3; ++k) r.lo[k] += v[k], r.hi[k] += v[k];
    return r;
  }
  Box3f enlarged(float a) const {
    Box3f r = *this;
    for (int k = 0; k < 3; ++k) r.lo[k] = lo[k] - a, r.hi[k] = hi[k] + a;
    return r;
  }
};
inline bool overlap(const Box3f& a, const Box3f& b) {
  for (int k = 0; k < 3; ++k)
    if (a.hi[k] < b.lo[k] || b.hi[k] < a.lo[k]) return false;
  return true;
}

// MeshData::initBoudingData (model/mesh.cpp:44-51): float AABB of the
// vertices and the largest vertex norm (double).
struct MeshBounds {
  Box3f aabb;
  double radius = 0.0;
};
inline MeshBounds meshBounds(const std::vector<V>& verts) {
  MeshBounds m;
  for (const V& v : verts) {
    float f[3] = {(float)v.x, (float)v.y, (float)v.z};
    m.aabb.extend(f);
    m.radius = std::max(m.radius, norm(v));
  }
  return m;
}
// MeshData::getAABB(transform) (model/mesh.cpp:136-153): the 8 corners of
// the float AABB through the float cast of the transform.
inline Box3f transformedAabb(const MeshBounds& m, const M& T) {
  M4<float> Tf = T.cast<float>();
  Box3f r;
  for (int xi = 0; xi < 2; ++xi)
    for (int yi = 0; yi < 2; ++yi)
      for (int zi = 0; zi < 2; ++zi) {
        V3<float> c(xi == 0 ? m.aabb.lo[0] : m.aabb.hi[0], yi == 0 ? m.aabb.lo[1] : m.aabb.hi[1],
                    zi == 0 ? m.aabb.lo[2] : m.aabb.hi[2]);
        V3<float> v = transform(c, Tf);
        float f[3] = {v.x, v.y, v.z};
        r.extend(f);
      }
  return r;
}
// MeshData::getSphereBounds (model/mesh.cpp:155-158).
inline Box3f sphereBox(double radius) {
  float br = (float)radius;
  Box3f b;
  for (int k = 0; k < 3; ++k) b.lo[k] = -br, b.hi[k] = br;
  return b;
}
inline void f3(const V& v, float o[3]) { o[0] = (float)v.x, o[1] = (float)v.y, o[2] = (float)v.z; }

// State::interpolate (model/state.cpp:55-70): throws on a time outside
// [last.time, this.time].
inline Quat<double> slerp(const Quat<double>& a, double t, const Quat<double>& b) {
  const double one = 1.0 - std::numeric_limits<double>::epsilon();
  const double d = ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
  const double absD = std::fabs(d);
  double s0, s1;
  if (absD >= one) {
    s0 = 1.0 - t;
    s1 = t;
  } else {
    const double theta = std::acos(absD);
    const double st = std::sin(theta);
    s0 = std::sin((1.0 - t) * theta) / st;
    s1 = std::sin(t * theta) / st;
  }
  if (d < 0.0) s1 = -s1;
  return {s0 * a.w + s1 * b.w, s0 * a.x + s1 * b.x, s0 * a.y + s1 * b.y, s0 * a.z + s1 * b.z};
}
inline State interpolate(const State& cur, const State& last, double time) {
  if (time > cur.time || time < last.time || last.time > cur.time)
    throw std::runtime_error("Invalid args for interpolation");
  if (last.time == cur.time) return cur;
  const double t = (time - last.time) / (cur.time - last.time);
  State r;
  r.time = time;
  r.trans = t * cur.trans + (1 - t) * last.trans;
  r.vel = t * cur.vel + (1 - t) * last.vel;
  r.L = t * cur.L + (1 - t) * last.L;
  r.rot = slerp(last.rot, t, cur.rot);
  return r;
}

// One body as the broad phase reads it.
struct BpBody {
  State current, future, last;
  MeshBounds mesh;
  double geo_eps = 0.01;
  bool is_static = false;
  double last_time_step_size = 1.0;
  int group = 0;
};

// boundsFunc (broad_phase_embree.cpp:60-84): t is the call's step.
inline Box3f boundsFixed(const BpBody& p, double t) {
  Box3f b;
  if (p.is_static) {
    b.extend(transformedAabb(p.mesh, p.current.transformation()));
  } else {
    const Box3f s = sphereBox(p.mesh.radius);
    float start[3], vel[3];
    f3(p.current.trans, start);
    f3(p.current.vel, vel);
    b.extend(s.translated(start));
    const float tf = (float)t;
    float end[3] = {start[0] + vel[0] * tf, start[1] + vel[1] * tf, start[2] + vel[2] * tf};
    b.extend(s.translated(end));
  }
  return b.enlarged((float)p.geo_eps);
}
// boundsFuncIndividualTime (broad_phase_embree.cpp:86-112).
inline Box3f boundsIndividual(const BpBody& p) {
  Box3f b;
  if (p.is_static) {
    b.extend(transformedAabb(p.mesh, p.current.transformation()));
  } else {
    const Box3f s = sphereBox(p.mesh.radius);
    float l[3], c[3], f[3];
    f3(p.last.trans, l);
    f3(p.current.trans, c);
    f3(p.future.trans, f);
    b.extend(s.translated(l));
    b.extend(s.translated(c));
    b.extend(s.translated(f));
  }
  return b.enlarged((float)p.geo_eps);
}

// collideFunc (:143-159) for geomID0 = a < geomID1 = b.
inline bool collideFixed(const BpBody& a, const BpBody& b, double t) {
  if (a.is_static && b.is_static) return false;
  const double s = sphereSphereFirstIntersect(a.current.trans, a.current.vel, a.mesh.radius + a.geo_eps, b.current.trans,
                                              b.current.vel, b.mesh.radius + b.geo_eps);
  return s >= 0.0 && s < t;
}
// collideFuncIndividualTime (:178-209): the later body is rolled back to the
// earlier one's time before the sphere test.
inline bool collideIndividual(const BpBody& a, const BpBody& b) {
  if (a.is_static && b.is_static) return false;
  State as = a.current, bs = b.current;
  if (!a.is_static && !b.is_static) {
    if (a.current.time < b.current.time)
      bs = interpolate(b.current, b.last, a.current.time);
    else
      as = interpolate(a.current, a.last, b.current.time);
  }
  const double s = sphereSphereFirstIntersect(as.trans, as.vel, a.mesh.radius + a.geo_eps, bs.trans, bs.vel,
                                              b.mesh.radius + b.geo_eps);
  return s >= 0.0 && s < std::max(a.last_time_step_size, b.last_time_step_size);
}

// mode 0: broadPhaseEmbree(particles, t); mode 1: broadPhaseEmbree(particles).
inline std::vector<std::pair<int, int>> broadPhase(const std::vector<BpBody>& bodies, int mode, double t) {
  const int n = (int)bodies.size();
  std::vector<Box3f> box(n);
  for (int i = 0; i < n; ++i) box[i] = mode == 0 ? boundsFixed(bodies[i], t) : boundsIndividual(bodies[i]);
  std::vector<std::pair<int, int>> out;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (bodies[i].group != bodies[j].group || !overlap(box[i], box[j])) continue;
      const bool hit = mode == 0 ? collideFixed(bodies[i], bodies[j], t) : collideIndividual(bodies[i], bodies[j]);
      if (hit) out.push_back({i, j});
    }
  return out;
}

// separateCollisionClusters (separate_clusters.cpp:355-459) without sleeping
// bodies.  Components of the graph of pairs between two non-static bodies
// over every body (libigl numbering); components without a non-static body
// are dropped and the rest keep their order.  A pair belongs to the cluster
// of its non-static body (a static body joins every cluster it touches).
// Non-static members are rolled back to the cluster's smallest current time.
struct Clusters {
  int n = 0;
  std::vector<int> of_body;  // -1 for static bodies
  std::vector<int> of_pair;  // -1 for a pair of two static bodies
  std::vector<int> first;    // cluster.particles[0]: the smallest member index
  std::vector<double> min_current_time, step_size;
};
inline Clusters separateClusters(std::vector<BpBody>& bodies, const std::vector<std::pair<int, int>>& pairs) {
  const int n = (int)bodies.size();
  std::vector<std::vector<int>> adj(n);
  for (auto& p : pairs)
    if (!bodies[p.first].is_static && !bodies[p.second].is_static) {
      adj[p.first].push_back(p.second);
      adj[p.second].push_back(p.first);
    }
  std::vector<int> comp(n, -1);
  int n_comp = 0;
  for (int f = 0; f < n; ++f) {
    if (comp[f] >= 0) continue;
    std::queue<int> q;
    q.push(f);
    comp[f] = n_comp;
    while (!q.empty()) {
      int g = q.front();
      q.pop();
      for (int h : adj[g])
        if (comp[h] < 0) comp[h] = n_comp, q.push(h);
    }
    ++n_comp;
  }
  // a component is kept when it has a non-static member (statics are alone)
  std::vector<int> keep(n_comp, -1);
  Clusters C;
  for (int i = 0; i < n; ++i)
    if (!bodies[i].is_static && keep[comp[i]] < 0) keep[comp[i]] = 0;
  for (int k = 0; k < n_comp; ++k)
    if (keep[k] == 0) keep[k] = C.n++;
  C.of_body.assign(n, -1);
  C.first.assign(C.n, -1);
  C.min_current_time.assign(C.n, std::numeric_limits<double>::infinity());
  C.step_size.assign(C.n, std::numeric_limits<double>::infinity());
  for (int i = 0; i < n; ++i) {
    if (bodies[i].is_static) continue;
    const int c = keep[comp[i]];
    C.of_body[i] = c;
    if (C.first[c] < 0) C.first[c] = i;
    C.min_current_time[c] = std::min(C.min_current_time[c], bodies[i].current.time);
    C.step_size[c] = std::min(C.step_size[c], bodies[i].last_time_step_size);
  }
  for (auto& p : pairs) {
    const BpBody &a = bodies[p.first], &b = bodies[p.second];
    C.of_pair.push_back(a.is_static && b.is_static ? -1 : C.of_body[a.is_static ? p.second : p.first]);
  }
  for (int i = 0; i < n; ++i)
    if (!bodies[i].is_static)  // Particle::rollBackState (particle.cpp:96-98)
      bodies[i].current = interpolate(bodies[i].current, bodies[i].last, C.min_current_time[C.of_body[i]]);
  return C;
}

}  // namespace d2o
