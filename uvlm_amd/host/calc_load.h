// calc_load.h -- aerodynamic loads of the legacy ring loop (calc_load/*.h,
// katz_plotkin.cpp of the reference): the Katz-Plotkin panel forces and the
// container form of the Joukowski load.  The ring sums they need go to the
// GPU in one batched call per load.
#pragma once

#include "legacy.h"
#include "morphing.h"
#include "vector3d.h"

namespace UVLM {
namespace calc_load {

// functions.h:33-36
struct AerodynamicLoad {
  Vector3d F;
  double Pin, Pout;
};

namespace internal {
// Row-major 3 x 3 matrix for the projection operator.
struct Matrix3 {
  double a[3][3];
  Vector3d operator*(const Vector3d& v) const {
    return Vector3d((a[0][0] * v.x() + a[0][1] * v.y()) + a[0][2] * v.z(),
                    (a[1][0] * v.x() + a[1][1] * v.y()) + a[1][2] * v.z(),
                    (a[2][0] * v.x() + a[2][1] * v.y()) + a[2][2] * v.z());
  }
};
// functions.h:14-18: I - u u^T with u = Um / |Um|.
Matrix3 CalcProjectionOperator(const Vector3d& Um);
// functions.h:23-30: freestream minus the surface velocity at x0.
Vector3d CalcUm(const Morphing& morphing, const Vector3d& x0, const Vector3d& freestream,
                double t);
// katz_plotkin.h:20-35
double CalcLocalLiftSt(const Vector3d& Um, const Vector3d& Uw, const Vector3d& grad, double b,
                       double c, double alpha, double rho);
double CalcLocalDragSt(const Vector3d& Ubc, const Vector3d& Uw, const Matrix3& P,
                       const Vector3d& n, double dg, double b, double rho);
}  // namespace internal

// katz_plotkin.cpp:16-68.  `rings` supplies the wake and chordwise bound
// velocities; the legacy loop passes its never-initialised global ring set
// (simulator.cpp:37), so both are zero there.  F starts from zero (the
// reference leaves it uninitialised).
AerodynamicLoad CalcLoadKatzPlotkin(const VortexContainer& c, const VortexContainer& cp,
                                    const Morphing& morphing, const UVLMVortexRing& rings,
                                    const Vector3d& freestream, double rho, double t, double dt);

// joukowski.h:63-99: per ring, (U x dl) rho Gamma over its segments
// node[(k+1)%4] -> node[k] (the trailing-edge segment k = 1 of the last row
// skipped) with U = velocity of every ring of the shared vector at the
// segment midpoint + freestream - surface velocity at the reference
// midpoint, plus normal rho dGamma/dt CalcB CalcC; Pin = sum dF . Vls
// (surface velocity at the reference centroid), Pout = -F.x |freestream|.
AerodynamicLoad CalcLoadJoukowski(const VortexContainer& vb, const VortexContainer& vb_prev,
                                  const Morphing& morphing, const Vector3d& freestream,
                                  double rho, double t, double dt);

}  // namespace calc_load
}  // namespace UVLM
