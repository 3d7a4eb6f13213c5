// advect.cpp -- advection schemes (advect.cpp:11-216, advect_factory.cpp of
// the reference) for host-held sheets; stage velocities are GPU N-body sums.
#include "advect.h"

#include <cstdio>
#include <cstdlib>
#include <map>

#include "../../include/uvlm_gpu.h"
#include "device.h"
#include "vortex.h"

namespace UVLM {
namespace advect {

namespace {
std::vector<Vector3d> flow(const std::vector<Vector3d>& x,
                           const MultipleSheet<Vector3d>& wing_pos,
                           const MultipleSheet<double>& wing_gamma,
                           const MultipleSheet<Vector3d>& wake_pos,
                           const MultipleSheet<double>& wake_gamma,
                           const Vector3d& ff) {
  const auto b = InducedVelocity(x, wing_pos, wing_gamma);
  const auto w = InducedVelocity(x, wake_pos, wake_gamma);
  std::vector<Vector3d> r(x.size());
  for (std::size_t k = 0; k < x.size(); ++k) r[k] = (-ff + b[k]) + w[k];
  return r;
}
}  // namespace

Vector3d Velocity(const Vector3d& x, const MultipleSheet<Vector3d>& wing_pos,
                  const MultipleSheet<double>& wing_gamma,
                  const MultipleSheet<Vector3d>& wake_pos,
                  const MultipleSheet<double>& wake_gamma,
                  const Vector3d& forward_flight) {
  return flow({x}, wing_pos, wing_gamma, wake_pos, wake_gamma, forward_flight)[0];
}

void Velocity(MultipleSheet<Vector3d>* result, const MultipleSheet<Vector3d>& wing_pos,
              const MultipleSheet<double>& wing_gamma,
              const MultipleSheet<Vector3d>& wake_pos,
              const MultipleSheet<double>& wake_gamma,
              const Vector3d& forward_flight) {
  result->resize(wake_pos);
  std::vector<Vector3d> x(wake_pos.begin(), wake_pos.end());
  const auto v = flow(x, wing_pos, wing_gamma, wake_pos, wake_gamma, forward_flight);
  std::copy(v.begin(), v.end(), result->begin());
  for (std::size_t n = 0; n < result->num(); n++)
    std::fill(result->iterator_at(n, 0, 0), result->iterator_at(n, 1, 0), -forward_flight);
}

void Euler::Advect(MultipleSheet<Vector3d>* next, const MultipleSheet<Vector3d>& wing_pos,
                   const MultipleSheet<double>& wing_gamma,
                   const MultipleSheet<Vector3d>& wake_pos,
                   const MultipleSheet<double>& wake_gamma, const Vector3d& ff,
                   const double dt) const {
  MultipleSheet<Vector3d> vel;
  Velocity(&vel, wing_pos, wing_gamma, wake_pos, wake_gamma, ff);
  MultipleSheet<Vector3d> src(wake_pos);
  next->resize(src);
  for (std::size_t K = 0; K < src.size(); K++) (*next)[K] = src[K] + vel[K] * dt;
}
int Euler::scheme() const { return UVLM_ADVECT_EULER; }

void RungeKutta2::Advect(MultipleSheet<Vector3d>* next,
                         const MultipleSheet<Vector3d>& wing_pos,
                         const MultipleSheet<double>& wing_gamma,
                         const MultipleSheet<Vector3d>& wake_pos,
                         const MultipleSheet<double>& wake_gamma, const Vector3d& ff,
                         const double dt) const {
  const MultipleSheet<Vector3d> w(wake_pos);
  MultipleSheet<Vector3d> k1, k2, p2(w);
  Velocity(&k1, wing_pos, wing_gamma, w, wake_gamma, ff);
  for (std::size_t K = 0; K < w.size(); K++) p2[K] = w[K] + k1[K] * dt;
  Velocity(&k2, wing_pos, wing_gamma, p2, wake_gamma, ff);
  next->resize(w);
  for (const auto& idx : w.list_index()) {
    const std::size_t K = std::get<0>(idx), i = std::get<2>(idx);
    if (i > ignore_steps_)
      (*next)[K] = w[K] + (k1[K] + k2[K]) * dt / 2;
    else
      (*next)[K] = w[K] + k1[K] * dt;
  }
}
int RungeKutta2::scheme() const { return UVLM_ADVECT_RK2; }

namespace {
void rk4(MultipleSheet<Vector3d>* next, const MultipleSheet<Vector3d>& wing_pos,
         const MultipleSheet<double>& wing_gamma, const MultipleSheet<Vector3d>& wake_pos,
         const MultipleSheet<double>& wake_gamma, const Vector3d& ff, const double dt,
         std::size_t ign, bool classic) {
  const MultipleSheet<Vector3d> w(wake_pos);
  MultipleSheet<Vector3d> k1, k2, k3, k4, p(w);
  Velocity(&k1, wing_pos, wing_gamma, w, wake_gamma, ff);
  for (std::size_t K = 0; K < w.size(); K++) p[K] = w[K] + k1[K] * dt / 2;
  Velocity(&k2, wing_pos, wing_gamma, p, wake_gamma, ff);
  for (std::size_t K = 0; K < w.size(); K++) p[K] = w[K] + k2[K] * dt / 2;
  Velocity(&k3, wing_pos, wing_gamma, p, wake_gamma, ff);
  if (classic) {
    for (std::size_t K = 0; K < w.size(); K++) p[K] = w[K] + k3[K] * dt;
    Velocity(&k4, wing_pos, wing_gamma, p, wake_gamma, ff);
  }
  next->resize(w);
  for (const auto& idx : w.list_index()) {
    const std::size_t K = std::get<0>(idx), i = std::get<2>(idx);
    if (i > ign) {
      const Vector3d last = classic ? k4[K] : k3[K];
      (*next)[K] = w[K] + (k1[K] + k2[K] * 2 + k3[K] * 2 + last) * dt / 6;
    } else {
      (*next)[K] = w[K] + k1[K] * dt;
    }
  }
}
}  // namespace

void RungeKutta4::Advect(MultipleSheet<Vector3d>* next,
                         const MultipleSheet<Vector3d>& wing_pos,
                         const MultipleSheet<double>& wing_gamma,
                         const MultipleSheet<Vector3d>& wake_pos,
                         const MultipleSheet<double>& wake_gamma, const Vector3d& ff,
                         const double dt) const {
  rk4(next, wing_pos, wing_gamma, wake_pos, wake_gamma, ff, dt, ignore_steps_, false);
}
int RungeKutta4::scheme() const { return UVLM_ADVECT_RK4; }

void RungeKutta4Classic::Advect(MultipleSheet<Vector3d>* next,
                                const MultipleSheet<Vector3d>& wing_pos,
                                const MultipleSheet<double>& wing_gamma,
                                const MultipleSheet<Vector3d>& wake_pos,
                                const MultipleSheet<double>& wake_gamma,
                                const Vector3d& ff, const double dt) const {
  rk4(next, wing_pos, wing_gamma, wake_pos, wake_gamma, ff, dt, ignore_steps_, true);
}
int RungeKutta4Classic::scheme() const { return UVLM_ADVECT_RK4_CLASSIC; }

void AdamsBashforth2::Advect(MultipleSheet<Vector3d>* next,
                             const MultipleSheet<Vector3d>& wing_pos,
                             const MultipleSheet<double>& wing_gamma,
                             const MultipleSheet<Vector3d>& wake_pos,
                             const MultipleSheet<double>& wake_gamma, const Vector3d& ff,
                             const double dt) const {
  const MultipleSheet<Vector3d> w(wake_pos);
  MultipleSheet<Vector3d> v;
  Velocity(&v, wing_pos, wing_gamma, w, wake_gamma, ff);
  next->resize(w);
  if (v_prev.size() == 0) {
    for (std::size_t K = 0; K < w.size(); K++) (*next)[K] = w[K] + v[K] * dt;
  } else {
    for (const auto& idx : w.list_index()) {
      std::size_t K, n, i, j;
      std::tie(K, n, i, j) = idx;
      if (i >= 1)
        next->at(n, i, j) = w.at(n, i, j) + (v.at(n, i, j) * 3 - v_prev.at(n, i - 1, j)) * dt / 2;
      else
        (*next)[K] = w[K] + v[K] * dt;
    }
  }
  v_prev = v;
}
int AdamsBashforth2::scheme() const { return UVLM_ADVECT_AB2; }

namespace {
enum Kind { EULER, RK2, RK4, AB2, RK4C };
const std::map<std::string, Kind>& names() {
  static const std::map<std::string, Kind> m{
      {"AB2", AB2}, {"RK2", RK2}, {"RK4", RK4}, {"RK4classic", RK4C}, {"euler", EULER}};
  return m;
}
}  // namespace

std::string GetNames() {
  std::string s;
  for (const auto& kv : names()) {
    if (!s.empty()) s += ", ";
    s += kv.first;
  }
  return s;
}

Advection* AdvectFactory(const std::string& name) {
  auto it = names().find(name);
  if (it == names().end()) {
    std::fprintf(stderr, "Invalid advection name! Please choose from following: [%s]\n",
                 GetNames().c_str());
    std::abort();
  }
  switch (it->second) {
    case EULER: return new Euler;
    case RK2: return new RungeKutta2;
    case RK4: return new RungeKutta4;
    case AB2: return new AdamsBashforth2;
    case RK4C: return new RungeKutta4Classic;
  }
  return nullptr;
}

}  // namespace advect
}  // namespace UVLM
