// calc_load.cpp -- legacy loads (declarations in calc_load.h).
#include "calc_load.h"

#include <cmath>
#include <vector>

namespace UVLM {
namespace calc_load {
namespace internal {

Matrix3 CalcProjectionOperator(const Vector3d& Um) {
  Vector3d u = Um;
  u.normalize();
  Matrix3 P;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P.a[i][j] = (i == j ? 1.0 : 0.0) - u[i] * u[j];
  return P;
}

Vector3d CalcUm(const Morphing& morphing, const Vector3d& x0, const Vector3d& freestream,
                double t) {
  Vector3d Uls;
  morphing.Velocity(&Uls, x0, t);
  return freestream - Uls;
}

double CalcLocalLiftSt(const Vector3d& Um, const Vector3d& Uw, const Vector3d& grad, double b,
                       double c, double alpha, double rho) {
  return rho * (Um + Uw).dot(grad) * b * c * std::cos(alpha);
}

double CalcLocalDragSt(const Vector3d& Ubc, const Vector3d& Uw, const Matrix3& P,
                       const Vector3d& n, double dg, double b, double rho) {
  return rho * (-(Ubc + Uw).dot(P * n) * dg * b);
}

}  // namespace internal

AerodynamicLoad CalcLoadKatzPlotkin(const VortexContainer& c, const VortexContainer& cp,
                                    const Morphing& morphing, const UVLMVortexRing& rings,
                                    const Vector3d& freestream, double rho, double t,
                                    double dt) {
  const std::size_t R = c.rows(), C = c.cols();
  // the ring sums at every centroid, one batched call per ring set
  std::vector<Vector3d> cent(R * C);
  for (std::size_t i = 0; i < R; ++i)
    for (std::size_t j = 0; j < C; ++j) cent[i * C + j] = c.at(i, j).Centroid();
  const auto Uw = UVLM::InducedVelocity(cent, rings.wake_vortices().cbegin(),
                                        rings.wake_vortices().cend(), false);
  const auto Ubc = UVLM::InducedVelocity(cent, rings.bound_vortices().cbegin(),
                                         rings.bound_vortices().cend(), true);
  Vector3d F(0, 0, 0);
  double Pin = 0, Pout = 0;
  for (std::size_t i = 0; i < R; ++i)
    for (std::size_t j = 0; j < C; ++j) {
      const VortexRing& v = c.at(i, j);
      const Vector3d& centroid = cent[i * C + j];
      const double db = v.CalcB();
      const double dc = v.CalcC();
      const Vector3d Um = internal::CalcUm(morphing, centroid, freestream, t);
      const internal::Matrix3 P = internal::CalcProjectionOperator(Um);
      const Vector3d n = v.Normal();
      const double alpha = v.AngleOfAttack(Um);
      // steady part: the reference passes (Um, Ubc) as (Ubc, Uw) to the drag
      const double Lst =
          internal::CalcLocalLiftSt(Um, Uw[i * C + j], c.Grad(i, j), db, dc, alpha, rho);
      const double Dst =
          internal::CalcLocalDragSt(Um, Ubc[i * C + j], P, n, c.DeltaGamma(i, j), db, rho);
      // unsteady part: dGamma/dt averaged with the ring ahead (row 0 halved)
      double dg_dt = (c.at(i, j).gamma() - cp.at(i, j).gamma()) / dt;
      if (i > 0) dg_dt += (c.at(i - 1, j).gamma() - cp.at(i - 1, j).gamma()) / dt;
      dg_dt /= 2;
      const double Lunst = rho * dg_dt * db * dc * std::cos(alpha);
      const double Dunst = rho * dg_dt * db * dc * std::sin(alpha);
      const Vector3d e_drag = Um / Um.norm();
      const Vector3d e_lift = P * n;
      const Vector3d dF = e_drag * (Dst + Dunst) + e_lift * (Lst + Lunst);
      F += dF;
      Pin += (-1) * Um.dot(n) * dF.dot(n);
    }
  return AerodynamicLoad{F, Pin, Pout};
}

AerodynamicLoad CalcLoadJoukowski(const VortexContainer& vb, const VortexContainer& vb_prev,
                                  const Morphing& morphing, const Vector3d& freestream,
                                  double rho, double t, double dt) {
  const auto& vortices = *vb.vortices();
  const std::size_t R = vb.rows(), C = vb.cols();
  // segment midpoints of every ring (the last row without its trailing edge)
  std::vector<Vector3d> mids, mids0, dls;
  std::vector<std::size_t> first(R * C + 1, 0);
  for (std::size_t i = 0; i < R; ++i)
    for (std::size_t j = 0; j < C; ++j) {
      const VortexRing& v = vb.at(i, j);
      const bool edge = i + 1 == R;
      const std::size_t m = v.nodes().size();
      for (std::size_t k = 0; k < m; ++k) {
        if (edge && k == 1) continue;
        const Vector3d& start = v.nodes()[(k + 1) % m];
        const Vector3d& end = v.nodes()[k];
        const Vector3d& start0 = v.nodes0()[(k + 1) % m];
        const Vector3d& end0 = v.nodes0()[k];
        mids.push_back((start + end) / 2);
        mids0.push_back((start0 + end0) / 2);
        dls.push_back(end - start);
      }
      first[i * C + j + 1] = mids.size();
    }
  const auto Uind = UVLM::InducedVelocity(mids, vortices.cbegin(), vortices.cend(), false);
  double Fx = 0, Fy = 0, Fz = 0, Pin = 0, Pout = 0;
  for (std::size_t i = 0; i < R; ++i)
    for (std::size_t j = 0; j < C; ++j) {
      const VortexRing& v = vb.at(i, j);
      Vector3d dF_st(0, 0, 0);
      for (std::size_t q = first[i * C + j]; q < first[i * C + j + 1]; ++q) {
        const Vector3d Um = internal::CalcUm(morphing, mids0[q], freestream, t);
        const Vector3d U = Uind[q] + Um;
        dF_st += U.cross(dls[q]) * rho * v.gamma();
      }
      const double dGdt = (v.gamma() - vb_prev.at(i, j).gamma()) / dt;
      const Vector3d dF_unst = v.Normal() * rho * dGdt * v.CalcB() * v.CalcC();
      const Vector3d dF = dF_st + dF_unst;
      Fx += dF.x();
      Fy += dF.y();
      Fz += dF.z();
      Vector3d Vls;
      morphing.Velocity(&Vls, v.ReferenceCentroid(), t);
      Pin += dF.dot(Vls);
    }
  const Vector3d F(Fx, Fy, Fz);
  Pout = Fx * freestream.norm() * (-1);
  return AerodynamicLoad{F, Pin, Pout};
}

}  // namespace calc_load
}  // namespace UVLM
