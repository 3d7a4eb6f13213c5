// wing.cpp -- wing lattice generators (wing/rect.cpp, naca4digit.cpp,
// naca00XX.cpp, wing.cpp of the reference).  Host-side input preparation.
#include "wing.h"

#include <algorithm>
#include <cmath>

namespace UVLM {

std::vector<double> linspace(double a, double b, std::size_t N) {
  const double dx = (b - a) / (N - 1);
  std::vector<double> res;
  for (std::size_t i = 0; i < N; i++) res.push_back(a + i * dx);
  return res;
}

std::vector<Vector3d> PointsToVector(const std::vector<proto::Point>& points) {
  std::vector<Vector3d> res;
  for (const auto& p : points) res.emplace_back(p.x(), p.y(), p.z());
  return res;
}

namespace wing {

void TransformChebyshev(std::vector<double>* result, std::size_t N, double xmin,
                        double xmax) {
  auto theta = linspace(-M_PI_2, 0, N);
  result->clear();
  for (double t : theta) result->push_back(xmin + (xmax - xmin) * std::cos(t));
  *result->begin() = xmin;
  *result->rbegin() = xmax;
}

double NACA4digit(double x, double c, int xx) {
  const int first = xx / 10;
  const int second = xx % 10;
  const double m = first / 100.0;
  const double p = second / 10.0;
  double res = 0;
  if (0 <= x && x <= p * c) {
    res = m * x / p / p * (2 * p - x / c);
  } else if (p * c <= x && x <= c) {
    res = m * (c - x) / (1 - p) / (1 - p) * (1 + x / c - 2 * p);
  }
  return res;
}

double NACA00XX(double x, double c, int xx) {
  const double t = xx * 0.01;
  const double z = x / c;
  return 5 * t * c *
         (0.2969 * std::sqrt(z) + (-0.1260) * (z) + (-0.3516) * z * z +
          0.2843 * z * z * z + (-0.1015) * z * z * z * z);
}

void RectGenerator::Generate(proto::Wing* wing, const double chord,
                             const double span, const std::size_t rows,
                             const std::size_t cols) const {
  WingGenerator::Generate(wing, chord, span, rows, cols);
  std::vector<double> xs, ys;
  if (cheb_chord_)
    TransformChebyshev(&xs, rows + 1, 0, chord);
  else
    xs = linspace(0, chord, rows + 1);
  if (cheb_span_)
    TransformChebyshev(&ys, cols + 1, 0, span);
  else
    ys = linspace(0, span, cols + 1);
  for (const double x : xs)
    for (const double y : ys) {
      auto* p = wing->add_points();
      p->set_x(x);
      p->set_y(y);
      p->set_z(0);
    }
}

void NACA4digitGenerator::Generate(proto::Wing* wing, const double chord,
                                   const double span, const std::size_t rows,
                                   const std::size_t cols) const {
  RectGenerator::Generate(wing, chord, span, rows, cols);
  for (int i = 0; i < wing->points_size(); i++) {
    auto* p = wing->mutable_points(i);
    p->set_z(NACA4digit(p->x(), chord, digit_));
  }
}

void NACA00XXGenerator::Generate(proto::Wing* wing, const double chord,
                                 const double span, const std::size_t rows,
                                 const std::size_t cols) const {
  RectGenerator::Generate(wing, chord, span, rows, cols);
  for (int i = 0; i < wing->points_size(); i++) {
    auto* p = wing->mutable_points(i);
    p->set_z(NACA00XX(p->x(), chord, digit_));
  }
}

void SetOrigin(proto::Wing* wing, const Vector3d& origin) {
  wing->mutable_origin()->set_x(origin.x());
  wing->mutable_origin()->set_y(origin.y());
  wing->mutable_origin()->set_z(origin.z());
}

void WholeWing(proto::Wing* wing, const proto::Wing& half) {
  wing->Clear();
  wing->set_cols(half.cols() * 2);
  wing->set_rows(half.rows());
  *wing->mutable_origin() = half.origin();
  wing->set_chord(half.chord());
  wing->set_span(half.span());
  const std::vector<Vector3d> hp = PointsToVector(half.points());
  const std::size_t COLS = half.cols() + 1;
  const std::size_t ROWS = hp.size() / COLS;
  const Vector3d origin(half.origin().x(), half.origin().y(), half.origin().z());
  std::vector<Vector3d> whole;
  for (std::size_t i = 0; i < ROWS; ++i) {
    for (std::size_t k = COLS; k-- > 0;) {  // mirrored row, reversed
      Vector3d p = hp[i * COLS + k];
      p.y() *= -1;
      whole.push_back(p);
    }
    for (std::size_t k = 1; k < COLS; ++k) whole.push_back(hp[i * COLS + k]);
  }
  for (auto& p : whole) {
    p += origin;
    auto* q = wing->add_points();
    q->set_x(p.x());
    q->set_y(p.y());
    q->set_z(p.z());
  }
}

}  // namespace wing
}  // namespace UVLM
