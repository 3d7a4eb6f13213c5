// wing.h -- wing lattice generators of the host API (wing/*.h, wing.cpp of
// the reference).  The reference fills a protobuf UVLM::proto::Wing; this
// build has no protobuf, so proto::Wing below is a plain struct with the
// accessors the generators use.
#pragma once

#include <memory>
#include <vector>

#include "vector3d.h"

namespace UVLM {
namespace proto {

struct Point {
  double x_ = 0, y_ = 0, z_ = 0;
  double x() const { return x_; }
  double y() const { return y_; }
  double z() const { return z_; }
  void set_x(double v) { x_ = v; }
  void set_y(double v) { y_ = v; }
  void set_z(double v) { z_ = v; }
};

struct Wing {
  double chord_ = 0, span_ = 0;
  std::size_t rows_ = 0, cols_ = 0;
  Point origin_;
  std::vector<Point> points_;
  void Clear() { *this = Wing(); }
  double chord() const { return chord_; }
  double span() const { return span_; }
  std::size_t rows() const { return rows_; }
  std::size_t cols() const { return cols_; }
  void set_chord(double v) { chord_ = v; }
  void set_span(double v) { span_ = v; }
  void set_rows(std::size_t v) { rows_ = v; }
  void set_cols(std::size_t v) { cols_ = v; }
  const Point& origin() const { return origin_; }
  Point* mutable_origin() { return &origin_; }
  const std::vector<Point>& points() const { return points_; }
  Point* add_points() {
    points_.emplace_back();
    return &points_.back();
  }
  Point* mutable_points(int i) { return &points_[i]; }
  int points_size() const { return (int)points_.size(); }
};

}  // namespace proto

// util.h:36-43
std::vector<double> linspace(double a, double b, std::size_t N);

std::vector<Vector3d> PointsToVector(const std::vector<proto::Point>& points);

namespace wing {

// rect.h:17-26
void TransformChebyshev(std::vector<double>* result, std::size_t N, double xmin,
                        double xmax);
// naca4digit.cpp:12-25 (mean camber of a NACA 4-digit airfoil)
double NACA4digit(double x, double c, int xx);
// naca00XX.cpp:15-20 (thickness used as camber)
double NACA00XX(double x, double c, int xx);

// wing_generator.h:14-30
class WingGenerator {
 public:
  virtual ~WingGenerator() = default;
  virtual void Generate(proto::Wing* wing, const double chord, const double span,
                        const std::size_t rows, const std::size_t cols) const {
    wing->Clear();
    wing->set_chord(chord);
    wing->set_span(span);
    wing->set_cols(cols);
    wing->set_rows(rows);
  }
  void operator()(proto::Wing* wing, const double chord, const double span,
                  const std::size_t rows, const std::size_t cols) const {
    Generate(wing, chord, span, rows, cols);
  }
};

// rect.cpp:18-44; chebyshev_chord / chebyshev_span replace the reference's
// --chebyshev_chord / --chebyshev_span flags.
class RectGenerator : public WingGenerator {
 public:
  RectGenerator(bool chebyshev_chord = false, bool chebyshev_span = false)
      : cheb_chord_(chebyshev_chord), cheb_span_(chebyshev_span) {}
  void Generate(proto::Wing* wing, const double chord, const double span,
                const std::size_t rows, const std::size_t cols) const override;

 private:
  bool cheb_chord_, cheb_span_;
};

class NACA4digitGenerator : public RectGenerator {
 public:
  explicit NACA4digitGenerator(int digit) : digit_(digit) {}
  void Generate(proto::Wing* wing, const double chord, const double span,
                const std::size_t rows, const std::size_t cols) const override;

 private:
  int digit_;
};

class NACA00XXGenerator : public RectGenerator {
 public:
  explicit NACA00XXGenerator(int digit) : digit_(digit) {}
  void Generate(proto::Wing* wing, const double chord, const double span,
                const std::size_t rows, const std::size_t cols) const override;

 private:
  int digit_;
};

// wing.h:19-25
void SetOrigin(proto::Wing* wing, const Vector3d& origin);
// wing.cpp:12-50: mirror the half wing (y >= 0) into a whole wing.
void WholeWing(proto::Wing* wing, const proto::Wing& half);

}  // namespace wing
}  // namespace UVLM
