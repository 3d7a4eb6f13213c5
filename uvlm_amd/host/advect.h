// advect.h -- wake advection schemes of the host API (advect.h:13-125 of
// the reference) and the plugin seam AdvectFactory (advect_factory.h).
//
// Each scheme carries the C-ABI scheme id; SimpleSimulator hands it to
// uvlm_sim_advect so the whole update runs on the device-resident wake.  The
// stand-alone Advect() keeps the reference signature for sheets held on the
// host: the stage velocities are GPU N-body sums, the stage arithmetic is the
// reference's.
#pragma once

#include <string>

#include "multiple_sheet.h"
#include "vector3d.h"

namespace UVLM {
namespace advect {

// advect.h:16-24: ((-ff + bound) + wake) at one point.
Vector3d Velocity(const Vector3d& x, const MultipleSheet<Vector3d>& wing_pos,
                  const MultipleSheet<double>& wing_gamma,
                  const MultipleSheet<Vector3d>& wake_pos,
                  const MultipleSheet<double>& wake_gamma,
                  const Vector3d& forward_flight);

// advect.h:26-46: every node of wake_pos; trailing-edge rows get -ff.
void Velocity(MultipleSheet<Vector3d>* result,
              const MultipleSheet<Vector3d>& wing_pos,
              const MultipleSheet<double>& wing_gamma,
              const MultipleSheet<Vector3d>& wake_pos,
              const MultipleSheet<double>& wake_gamma,
              const Vector3d& forward_flight);

class Advection {
 public:
  Advection() = default;
  virtual ~Advection() = default;
  virtual void Advect(MultipleSheet<Vector3d>* next,
                      const MultipleSheet<Vector3d>& wing_pos,
                      const MultipleSheet<double>& wing_gamma,
                      const MultipleSheet<Vector3d>& wake_pos,
                      const MultipleSheet<double>& wake_gamma,
                      const Vector3d& forward_flight, const double dt) const = 0;
  // UVLM_ADVECT_* id and the Euler-row count of the device-resident update.
  virtual int scheme() const = 0;
  virtual std::size_t ignore_steps() const { return 0; }
};

class Euler : public Advection {
 public:
  void Advect(MultipleSheet<Vector3d>* next, const MultipleSheet<Vector3d>& wing_pos,
              const MultipleSheet<double>& wing_gamma,
              const MultipleSheet<Vector3d>& wake_pos,
              const MultipleSheet<double>& wake_gamma,
              const Vector3d& forward_flight, const double dt) const override;
  int scheme() const override;
};

class RungeKutta2 : public Advection {
 public:
  explicit RungeKutta2(std::size_t steps = 5) : ignore_steps_(steps) {}
  void Advect(MultipleSheet<Vector3d>* next, const MultipleSheet<Vector3d>& wing_pos,
              const MultipleSheet<double>& wing_gamma,
              const MultipleSheet<Vector3d>& wake_pos,
              const MultipleSheet<double>& wake_gamma,
              const Vector3d& forward_flight, const double dt) const override;
  int scheme() const override;
  std::size_t ignore_steps() const override { return ignore_steps_; }

 private:
  const std::size_t ignore_steps_;
};

// The reference's "RK4": the update uses k1 + 2 k2 + 2 k3 + k3
// (advect.cpp:172-173); its k4 stage never reaches the result and is not
// evaluated here.
class RungeKutta4 : public Advection {
 public:
  explicit RungeKutta4(std::size_t steps = 5) : ignore_steps_(steps) {}
  void Advect(MultipleSheet<Vector3d>* next, const MultipleSheet<Vector3d>& wing_pos,
              const MultipleSheet<double>& wing_gamma,
              const MultipleSheet<Vector3d>& wake_pos,
              const MultipleSheet<double>& wake_gamma,
              const Vector3d& forward_flight, const double dt) const override;
  int scheme() const override;
  std::size_t ignore_steps() const override { return ignore_steps_; }

 protected:
  const std::size_t ignore_steps_;
};

// Textbook RK4 (k1 + 2 k2 + 2 k3 + k4); no reference counterpart, factory
// name "RK4classic".
class RungeKutta4Classic : public RungeKutta4 {
 public:
  explicit RungeKutta4Classic(std::size_t steps = 5) : RungeKutta4(steps) {}
  void Advect(MultipleSheet<Vector3d>* next, const MultipleSheet<Vector3d>& wing_pos,
              const MultipleSheet<double>& wing_gamma,
              const MultipleSheet<Vector3d>& wake_pos,
              const MultipleSheet<double>& wake_gamma,
              const Vector3d& forward_flight, const double dt) const override;
  int scheme() const override;
};

class AdamsBashforth2 : public Advection {
 public:
  void Advect(MultipleSheet<Vector3d>* next, const MultipleSheet<Vector3d>& wing_pos,
              const MultipleSheet<double>& wing_gamma,
              const MultipleSheet<Vector3d>& wake_pos,
              const MultipleSheet<double>& wake_gamma,
              const Vector3d& forward_flight, const double dt) const override;
  int scheme() const override;

 private:
  mutable MultipleSheet<Vector3d> v_prev;
};

// advect_factory.cpp:43-60: "euler", "RK2", "RK4", "AB2" (+ "RK4classic").
// Unknown names abort with the list of valid names.  The caller owns the
// returned object.
Advection* AdvectFactory(const std::string& name);
std::string GetNames();

}  // namespace advect
}  // namespace UVLM
