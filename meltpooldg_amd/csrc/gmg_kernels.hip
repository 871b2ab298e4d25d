// Setup kernels of the multigrid preconditioner: conductivity cell averages
// straight from an operator's data, and an in-place Gauss-Jordan inverse of the
// (small, symmetric positive definite) coarse-grid matrix.
#include "blas.hpp"

namespace mpdg {

// quadrature-weighted cell average of the conductivity of an FP64 operator
__global__ __launch_bounds__(256) void op_cell_average_kernel(int n1, int nc0, int nc1, long ncells, int nd0, int nd1,
                                                              int mode, double kappa, const double *__restrict__ karr,
                                                              int law, double p0, double p1, double p2, double p3,
                                                              const double *__restrict__ T, double S00, double *Sg,
                                                              double *wg, double *__restrict__ kc) {
  (void)S00;
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= ncells) return;
  const int nq = n1 * n1 * n1, K = n1 - 1;
  double s = 0.0;
  if (mode == KAPPA_CONST) s = kappa;
  else if (mode == KAPPA_CELL) s = karr[c];
  else if (mode == KAPPA_QP) {
    for (int q = 0; q < nq; ++q) s += wg[q % n1] * wg[(q / n1) % n1] * wg[q / (n1 * n1)] * karr[c * nq + q];
  } else {
    const int cx = c % nc0, cy = (c / nc0) % nc1, cz = c / ((long)nc0 * nc1);
    for (int q = 0; q < nq; ++q) {
      const int qx = q % n1, qy = (q / n1) % n1, qz = q / (n1 * n1);
      double Tq = 0.0;
      for (int l = 0; l < n1; ++l)
        for (int j = 0; j < n1; ++j) {
          const long row = (long)K * cx + (long)nd0 * ((K * cy + j) + (long)nd1 * (K * cz + l));
          double r = 0.0;
          for (int i = 0; i < n1; ++i) r += Sg[qx * kMaxN1 + i] * T[row + i];
          Tq += Sg[qy * kMaxN1 + j] * Sg[qz * kMaxN1 + l] * r;
        }
      double k;
      if (law == LAW_TANH) k = p0 * (1.0 + p1 * tanh((Tq - p2) / p3));
      else {
        double fs = (p3 - Tq) / (p3 - p2);
        fs = fs < 0.0 ? 0.0 : (fs > 1.0 ? 1.0 : fs);
        k = p0 + (p1 - p0) * (3.0 - 2.0 * fs) * fs * fs;
      }
      s += wg[qx] * wg[qy] * wg[qz] * k;
    }
  }
  kc[c] = s;
}

int launch_op_cell_average(const OperatorData<double> &op, double *kc, double *table_scratch, hipStream_t stream) {
  FE1D fe(op.degree);
  double host[kMaxN1 * kMaxN1 + kMaxN1] = {};
  for (int a = 0; a < fe.n1; ++a) {
    host[kMaxN1 * kMaxN1 + a] = fe.qw[a];
    for (int b = 0; b < fe.n1; ++b) host[a * kMaxN1 + b] = fe.S[a][b];
  }
  hipMemcpyAsync(table_scratch, host, sizeof(host), hipMemcpyHostToDevice, stream);
  hipLaunchKernelGGL(op_cell_average_kernel, dim3((op.n_cells + 255) / 256), dim3(256), 0, stream, op.degree + 1,
                     op.nc[0], op.nc[1], op.n_cells, op.nd[0], op.nd[1], op.kappa_mode, op.kappa, op.kappa_arr, op.law,
                     op.law_p[0], op.law_p[1], op.law_p[2], op.law_p[3], op.T, 0.0, table_scratch,
                     table_scratch + kMaxN1 * kMaxN1, kc);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// In-place Gauss-Jordan inversion without pivoting (SPD matrix), step k in three kernels.
__global__ void gj_row_kernel(double *A, int n, int k) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n || j == k) return;
  A[(long)k * n + j] /= A[(long)k * n + k];
}
__global__ void gj_update_kernel(double *A, int n, int k) {
  const long g = (long)blockIdx.x * 256 + threadIdx.x;
  if (g >= (long)n * n) return;
  const int i = g / n, j = g % n;
  if (i == k || j == k) return;
  A[(long)i * n + j] -= A[(long)i * n + k] * A[(long)k * n + j];
}
__global__ void gj_col_kernel(double *A, int n, int k) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double piv = 1.0 / A[(long)k * n + k];
  if (i == k) return;
  A[(long)i * n + k] = -A[(long)i * n + k] * piv;
}
__global__ void gj_pivot_kernel(double *A, int n, int k) { A[(long)k * n + k] = 1.0 / A[(long)k * n + k]; }

int launch_gauss_jordan_inverse(double *A, int n, hipStream_t s) {
  const int b1 = (n + 255) / 256;
  const long b2 = ((long)n * n + 255) / 256;
  for (int k = 0; k < n; ++k) {
    hipLaunchKernelGGL(gj_row_kernel, dim3(b1), dim3(256), 0, s, A, n, k);
    hipLaunchKernelGGL(gj_update_kernel, dim3(b2), dim3(256), 0, s, A, n, k);
    hipLaunchKernelGGL(gj_col_kernel, dim3(b1), dim3(256), 0, s, A, n, k);
    hipLaunchKernelGGL(gj_pivot_kernel, dim3(1), dim3(1), 0, s, A, n, k);
  }
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

}  // namespace mpdg
