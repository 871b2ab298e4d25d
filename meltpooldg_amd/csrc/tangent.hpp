// Launchers of the full Newton tangent kernels (tangent.hip), Cartesian cells.
#pragma once

#include "kernels.hpp"

namespace mpdg {

// coefficients per quadrature point in the tangent data
constexpr int kTangentCoeffs = 8;

// td: kTangentCoeffs n_cells (k+1)^3 values, [cell][coefficient][q]. vfield: null (the constant
// vel) or a velocity DoF field, 3 n_dofs values (x, y, z blocks) on the temperature's nodes.
int launch_tangent_data(const OperatorData<double> &op, double dt, const mpdg_heat_material &mat, int solid,
                        const double *vel, const double *vfield, const double *T, const double *T_old, double *td,
                        hipStream_t s);
// dst = J src (dst zeroed first; constrained rows stay 0, constrained src entries read as 0)
// (conv: the data carries the convection coefficients; solid: as given to launch_tangent_data)
int launch_tangent_apply(const OperatorData<double> &op, const double *td, int conv, int solid,
                         const double *src, double *dst, hipStream_t s);
// dst += -R(T) (cell part; q may be null)
int launch_tangent_residual(const OperatorData<double> &op, double dt, const mpdg_heat_material &mat,
                            const double *vel, const double *vfield, const double *T, const double *T_old,
                            const double *q, double *dst, hipStream_t s);
// diag = diagonal of the cell part of J (constrained rows: 1)
int launch_tangent_diagonal(const OperatorData<double> &op, const double *td, int conv, int solid,
                            double *diag, hipStream_t s);
// whether the tangent has a convection term: a velocity field, or a non-zero constant velocity
inline int tangent_convects(const double *vel, const double *vfield) {
  return (vfield || vel[0] != 0.0 || vel[1] != 0.0 || vel[2] != 0.0) ? 1 : 0;
}

}  // namespace mpdg
