// morphing.cpp -- kinematics (morphing.cpp:23-84 of the reference).  Host
// code: it produces the node positions and surface velocities the GPU path
// consumes.  Built with -ffp-contract=off so the finite difference of
// Velocity (which amplifies rounding by 1/dt) matches plain IEEE evaluation.
#include "morphing.h"

#include <cmath>

namespace UVLM {

namespace {
double Zero1(double) { return 0; }
double Zero2(const Vector3d&, double) { return 0; }
}  // namespace

Morphing::Morphing() : alpha_(0), origin_(0, 0, 0) { Clear(); }

void Morphing::Clear() {
  plug_ = Zero1;
  flap_ = Zero1;
  twist_ = Zero2;
  bend_ = Zero2;
}

void Morphing::Perfome(Vector3d* x, const Vector3d& x0, const double t) const {
  Vector3d x_ref = x0 - origin_;
  const bool is_negative = x_ref.y() < 0;
  if (is_negative) x_ref.y() *= -1;
  Matrix3d T;
  PrepareMatrix(&T, x_ref, t);
  const double gamma_z = plug_(t);
  const double gamma_b = bend_(x_ref, t);
  *x = Vector3d::UnitZ() * gamma_z + T * (x_ref + Vector3d::UnitZ() * gamma_b);
  if (x_ref.y() == 0) x->y() = 0;
  if (is_negative) x->y() *= -1;
  *x += origin_;
}

void Morphing::Velocity(Vector3d* v, const Vector3d& x0, const double t,
                        const double dt) const {
  Vector3d x1, x2;
  Perfome(&x1, x0, t - dt);
  Perfome(&x2, x0, t + dt);
  *v = (x2 - x1) / (2 * dt);
}

void Morphing::PrepareMatrix(Matrix3d* m, const Vector3d& x0, double t) const {
  const double phi = flap_(t);
  const Matrix3d flap(1, 0, 0, 0, std::cos(phi), -std::sin(phi), 0, std::sin(phi),
                      std::cos(phi));
  const double alpha = alpha_;
  const Matrix3d attack(std::cos(alpha), 0, std::sin(alpha), 0, 1, 0, -std::sin(alpha),
                        0, std::cos(alpha));
  const double beta = twist_(x0, t);
  const Matrix3d twist(std::cos(beta), 0, std::sin(beta), 0, 1, 0, -std::sin(beta), 0,
                       std::cos(beta));
  *m = (attack * flap) * twist;
}

}  // namespace UVLM
