// Shared declarations of the HIP kernels and their host launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/mpdg_heat.h"
#include "fe1d.hpp"

namespace mpdg {

enum KappaMode { KAPPA_CONST = MPDG_KAPPA_CONST, KAPPA_QP = MPDG_KAPPA_QP, KAPPA_CELL = MPDG_KAPPA_CELL, KAPPA_LAW = MPDG_KAPPA_LAW };
enum LawId { LAW_TANH = MPDG_LAW_TANH, LAW_MUSHY = MPDG_LAW_MUSHY };

// Everything a level operator needs on the device (pointers are device memory).
template <typename Number>
struct OperatorData {
  int degree = 1;
  int nc[3] = {1, 1, 1};
  int nd[3] = {2, 2, 2};
  double L[3] = {1, 1, 1};
  double h[3] = {1, 1, 1};
  long n_dofs = 0, n_cells = 0;
  double c_mass = 1.0;
  int kappa_mode = KAPPA_CONST;
  double kappa = 1.0;
  const Number *kappa_arr = nullptr;
  int law = 0;
  double law_p[4] = {0, 0, 0, 1};
  const Number *T = nullptr;
  int dmask = 0;
  int robin_mask = 0;
  double alpha = 0, emissivity = 0, T_inf_conv = 0, T_inf_rad = 0;
  const double *T_robin = nullptr;  // temperature for the radiation linearisation (FP64)
  // mapped geometry: per quadrature point [cell][7][q]: JxW, then G = JxW J^-1 J^-T as (00, 01, 11, 02, 12, 22)
  const Number *metric = nullptr;
  // metric with the conductivity folded in (kappa_q G), used for per-qp / k(T) conductivity on mapped meshes
  const Number *metric_k = nullptr;
  const double *coords = nullptr;   // mapping node coordinates (X block, Y block, Z block), FP64
  // apply scratch: contributions of the upper z-layer of columns to the DoF rows where four
  // columns meet, summed into dst after the cell kernel so that vmult is bitwise reproducible
  // (apply_edge_elems Number entries, all zero between applies)
  Number *edge = nullptr;
  size_t edge_elems = 0;
  void set_mesh(int k, const int *ncells, const double *ext) {
    degree = k;
    for (int d = 0; d < 3; ++d) {
      nc[d] = ncells[d];
      nd[d] = k * ncells[d] + 1;
      L[d] = ext[d];
      h[d] = ext[d] / ncells[d];
    }
    n_dofs = (long)nd[0] * nd[1] * nd[2];
    n_cells = (long)nc[0] * nc[1] * nc[2];
  }
};

// A 1D N1 x N1 table stored by its symmetry about the cell centre (Gauss-Lobatto nodes and
// Gauss points are symmetric): T[N1-1-q][N1-1-i] = T[q][i] (ANTI = false, shape values) or
// -T[q][i] (ANTI = true, derivatives); only rows q < (N1+1)/2 are kept. Halving the tables
// halves the scalar registers the kernels hold them in. Indices are compile-time constants
// after unrolling, so the mirrored half costs at most a sign modifier.
template <typename Number, int N1, bool ANTI>
struct Half1D {
  Number h[(N1 + 1) / 2][N1];
  __host__ __device__ __forceinline__ Number operator()(int q, int i) const {
    return (2 * q < N1) ? h[q][i] : (ANTI ? -h[N1 - 1 - q][N1 - 1 - i] : h[N1 - 1 - q][N1 - 1 - i]);
  }
  void set(const double (&full)[kMaxN1][kMaxN1]) {
    for (int q = 0; 2 * q < N1; ++q)
      for (int i = 0; i < N1; ++i) h[q][i] = (Number)full[q][i];
  }
};

template <typename Number, int N1>
struct ApplyArgs {
  Half1D<Number, N1, false> S;  // shape values at the Gauss points
  Half1D<Number, N1, true> D;   // collocation derivative on the Gauss points
  Number w[N1];      // Gauss weights
  int nc[3], nd[3], ncol[2];
  int dmask, xcd_cols, law;
  Number cm;         // c_mass * hx hy hz
  Number ks[3];      // hx hy hz / h_d^2
  Number kappa;
  Number law_p[4];
  const Number *kappa_arr;
  const Number *T;
  const Number *metric;
  Number *edge;      // rows shared by four columns: the upper layer's contributions
};

template <typename Number, int N1>
void fill_apply_args(const OperatorData<Number> &op, ApplyArgs<Number, N1> &A) {
  const FE1D &fe = fe1d(N1 - 1);
  for (int a = 0; a < N1; ++a) A.w[a] = (Number)fe.qw[a];
  A.S.set(fe.S);
  A.D.set(fe.D);
  for (int d = 0; d < 3; ++d) {
    A.nc[d] = op.nc[d];
    A.nd[d] = op.nd[d];
  }
  const double vol = op.metric ? 1.0 : op.h[0] * op.h[1] * op.h[2];
  A.metric = op.metric;
  A.cm = (Number)(op.c_mass * vol);
  for (int d = 0; d < 3; ++d) A.ks[d] = (Number)(vol / (op.h[d] * op.h[d]));
  A.dmask = op.dmask;
  A.kappa = (Number)op.kappa;
  A.kappa_arr = op.kappa_arr;
  A.T = op.T;
  A.law = op.law;
  for (int i = 0; i < 4; ++i) A.law_p[i] = (Number)op.law_p[i];
  A.xcd_cols = 0;
  A.ncol[0] = A.ncol[1] = 1;
  A.edge = op.edge;
}

// heat_apply.hip; `kernel_start` and `kernel_end`, when given, are recorded on the stream
// right before and after the cell kernel (between the face-row pre-pass and the edge fix-up)
// so that the event pair times that kernel alone
template <typename Number>
int launch_apply(const OperatorData<Number> &op, const Number *src, Number *dst, hipStream_t stream,
                 hipEvent_t kernel_start = nullptr, hipEvent_t kernel_end = nullptr);
// entries of OperatorData::edge the apply of a mesh of degree k and n_cells cells needs with
// the Cartesian (mapped = 0) or mapped (1) chunk shape, or the larger of the two (mapped = -1);
// 0 when no DoF row is shared by four columns
size_t apply_edge_elems(int degree, const int *n_cells, bool fp32, int mapped = -1);

// heat_aux.hip: diagonal, boundary faces, residual, norms, conductivity
// diag = the operator's diagonal (cells only; constrained rows 1); scratch: 2 n_cells (k+1)^3 values
template <typename Number>
int launch_diagonal(const OperatorData<Number> &op, Number *diag, Number *scratch, hipStream_t stream);
// boundary faces: mode 0 Robin tangent on src, 1 its diagonal, 2 boundary residual (+ surface flux)
int launch_boundary(const OperatorData<double> &op, int mode, const double *flux, double *dst,
                    const double *src, const double *T, hipStream_t stream);
struct FluxNodes {
  double nodes[kMaxN1];
  double p[6];
};
// The operator the kernels see: on mapped meshes with per-qp or k(T) conductivity the
// conductivity sits in metric_k and the kernels run with a constant unit conductivity.
template <typename Number>
OperatorData<Number> kernel_view(const OperatorData<Number> &op) {
  OperatorData<Number> v = op;
  if (op.metric && (op.kappa_mode == KAPPA_QP || op.kappa_mode == KAPPA_LAW)) {
    v.metric = op.metric_k;
    v.kappa_mode = KAPPA_CONST;
    v.kappa = 1.0;
  }
  return v;
}
template <typename Number>
int launch_scale_metric(const OperatorData<Number> &op, const Number *kq, Number *metric_k, hipStream_t s);
template <typename Number>
int launch_metric(const OperatorData<Number> &op, const double *coords, Number *metric, hipStream_t s);
int launch_surface_flux_field(const OperatorData<double> &op, const double *p6, double *out, hipStream_t s);
template <typename Number>
int launch_kappa_qp(const OperatorData<Number> &op, Number *kq, hipStream_t stream);
// quadrature-weighted cell averages sum_q w_q k_q of the conductivity (any mode)
template <typename Number>
int launch_cell_average(const OperatorData<Number> &op, Number *kcell, hipStream_t stream);
template <typename Number>
int launch_coarsen_cells(const int *nc_fine, const Number *kf, Number *kc, hipStream_t stream);

}  // namespace mpdg
