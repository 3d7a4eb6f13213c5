// morphing.h -- wing kinematics of the host API (morphing.h:15-110 of the
// reference): plunge (plug), flap about x, pitch alpha, twist(x0, t) and
// bend(x0, t); surface velocity by central difference.
#pragma once

#include <functional>

#include "vector3d.h"

namespace UVLM {

class Morphing {
 public:
  Morphing();
  Morphing(const Morphing&) = default;
  Morphing& operator=(const Morphing&) = default;

  // morphing.cpp:23-43
  void Perfome(Vector3d* x, const Vector3d& x0, const double t) const;
  Vector3d Perfome(const Vector3d& x0, const double t) const {
    Vector3d r;
    Perfome(&r, x0, t);
    return r;
  }
  // morphing.cpp:45-52
  void Velocity(Vector3d* v, const Vector3d& x0, const double t,
                const double dt = 1e-6) const;
  Vector3d Velocity(const Vector3d& x0, const double t, const double dt = 1e-6) const {
    Vector3d r;
    Velocity(&r, x0, t, dt);
    return r;
  }
  // morphing.cpp:54-76: attack * flap * twist
  void PrepareMatrix(Matrix3d* m, const Vector3d& x0, double t) const;

  Morphing& set_alpha(double alpha) {
    alpha_ = alpha;
    return *this;
  }
  Morphing& set_plug(std::function<double(double)> f) {
    plug_ = f;
    return *this;
  }
  Morphing& set_flap(std::function<double(double)> f) {
    flap_ = f;
    return *this;
  }
  Morphing& set_twist(std::function<double(const Vector3d&, double)> f) {
    twist_ = f;
    return *this;
  }
  Morphing& set_bend(std::function<double(const Vector3d&, double)> f) {
    bend_ = f;
    return *this;
  }
  void set_origin(const Vector3d& origin) { origin_ = origin; }
  const Vector3d& origin() const { return origin_; }
  void Clear();

 private:
  double alpha_;
  std::function<double(double)> plug_;
  std::function<double(double)> flap_;
  std::function<double(const Vector3d&, double)> twist_;
  std::function<double(const Vector3d&, double)> bend_;
  Vector3d origin_;
};

}  // namespace UVLM
