// vector3d.h -- UVLM::Vector3d, the 3-double value type of the host API.
//
// The reference uses Eigen::Vector3d (3 packed doubles).  This library does
// not depend on Eigen, so the host API defines a 24-byte POD with the members the
// reference's hot path uses.  Arithmetic conventions: dot and squaredNorm are
// (x*x' + y*y') + z*z'; normalize divides each component by the norm when the
// squared norm is positive; scalar division divides each component.
#pragma once

#include <cmath>
#include <cstddef>

namespace UVLM {

struct Vector3d {
  double v[3];

  Vector3d() : v{0.0, 0.0, 0.0} {}
  Vector3d(double x, double y, double z) : v{x, y, z} {}

  static Vector3d Zero() { return Vector3d(0, 0, 0); }
  static Vector3d UnitX() { return Vector3d(1, 0, 0); }
  static Vector3d UnitY() { return Vector3d(0, 1, 0); }
  static Vector3d UnitZ() { return Vector3d(0, 0, 1); }

  double& x() { return v[0]; }
  double& y() { return v[1]; }
  double& z() { return v[2]; }
  double x() const { return v[0]; }
  double y() const { return v[1]; }
  double z() const { return v[2]; }
  double& operator[](std::size_t i) { return v[i]; }
  double operator[](std::size_t i) const { return v[i]; }
  double& operator()(std::size_t i) { return v[i]; }
  double operator()(std::size_t i) const { return v[i]; }
  double* data() { return v; }
  const double* data() const { return v; }

  Vector3d operator+(const Vector3d& o) const {
    return Vector3d(v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]);
  }
  Vector3d operator-(const Vector3d& o) const {
    return Vector3d(v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]);
  }
  Vector3d operator-() const { return Vector3d(-v[0], -v[1], -v[2]); }
  Vector3d operator*(double s) const { return Vector3d(v[0] * s, v[1] * s, v[2] * s); }
  Vector3d operator/(double s) const { return Vector3d(v[0] / s, v[1] / s, v[2] / s); }
  Vector3d& operator+=(const Vector3d& o) {
    v[0] += o.v[0];
    v[1] += o.v[1];
    v[2] += o.v[2];
    return *this;
  }
  Vector3d& operator-=(const Vector3d& o) {
    v[0] -= o.v[0];
    v[1] -= o.v[1];
    v[2] -= o.v[2];
    return *this;
  }
  Vector3d& operator*=(double s) {
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
    return *this;
  }
  Vector3d& operator/=(double s) {
    v[0] /= s;
    v[1] /= s;
    v[2] /= s;
    return *this;
  }
  bool operator==(const Vector3d& o) const {
    return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2];
  }
  bool operator!=(const Vector3d& o) const { return !(*this == o); }

  double dot(const Vector3d& o) const {
    return (v[0] * o.v[0] + v[1] * o.v[1]) + v[2] * o.v[2];
  }
  Vector3d cross(const Vector3d& o) const {
    return Vector3d(v[1] * o.v[2] - v[2] * o.v[1], v[2] * o.v[0] - v[0] * o.v[2],
                    v[0] * o.v[1] - v[1] * o.v[0]);
  }
  double squaredNorm() const { return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]; }
  double norm() const { return std::sqrt(squaredNorm()); }
  void normalize() {
    const double z = squaredNorm();
    if (z > 0) *this /= std::sqrt(z);
  }
  Vector3d normalized() const {
    Vector3d r(*this);
    r.normalize();
    return r;
  }
};

inline Vector3d operator*(double s, const Vector3d& a) { return a * s; }

static_assert(sizeof(Vector3d) == 24, "Vector3d must be 3 packed doubles");

// 3x3 matrix, row-major, with the product used by Morphing (attack * flap *
// twist, morphing.cpp:75): (A*B)(i,j) = (A(i,0)B(0,j) + A(i,1)B(1,j)) + A(i,2)B(2,j).
struct Matrix3d {
  double a[3][3];
  Matrix3d() : a{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}} {}
  Matrix3d(double a00, double a01, double a02, double a10, double a11, double a12,
           double a20, double a21, double a22)
      : a{{a00, a01, a02}, {a10, a11, a12}, {a20, a21, a22}} {}
  static Matrix3d Identity() { return Matrix3d(1, 0, 0, 0, 1, 0, 0, 0, 1); }
  double& operator()(int i, int j) { return a[i][j]; }
  double operator()(int i, int j) const { return a[i][j]; }
  Matrix3d operator*(const Matrix3d& B) const {
    Matrix3d r;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        r.a[i][j] = (a[i][0] * B.a[0][j] + a[i][1] * B.a[1][j]) + a[i][2] * B.a[2][j];
    return r;
  }
  Vector3d operator*(const Vector3d& x) const {
    return Vector3d((a[0][0] * x[0] + a[0][1] * x[1]) + a[0][2] * x[2],
                    (a[1][0] * x[0] + a[1][1] * x[1]) + a[1][2] * x[2],
                    (a[2][0] * x[0] + a[2][1] * x[1]) + a[2][2] * x[2]);
  }
};

}  // namespace UVLM
