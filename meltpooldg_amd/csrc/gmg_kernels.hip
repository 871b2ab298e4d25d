// Setup kernels of the multigrid preconditioner: the dense coarse-grid matrix
// assembled on the device from element matrices, and its in-place Gauss-Jordan
// inverse (the matrix is symmetric positive definite: no pivoting), blocked so
// that most of the work is a rank-32 update with the panels staged in LDS.
#include "blas.hpp"

namespace mpdg {

struct CoarseTables {
  double S[kMaxN1][kMaxN1], G[kMaxN1][kMaxN1], w[kMaxN1];
};

__device__ __forceinline__ bool constrained(long g, int nd0, int nd1, int nd2, int m) {
  const int ix = g % nd0, iy = (g / nd0) % nd1, iz = g / ((long)nd0 * nd1);
  return ((m & 1) && ix == 0) || ((m & 2) && ix == nd0 - 1) || ((m & 4) && iy == 0) ||
         ((m & 8) && iy == nd1 - 1) || ((m & 16) && iz == 0) || ((m & 32) && iz == nd2 - 1);
}

// One thread per (cell, I, J) entry of the element matrix
//   A_e[I][J] = sum_q m_q phi_I phi_J + k_cell grad phi_I . G_q grad phi_J,
// Cartesian cells: m_q = c_m vol W_q, G_q = diag(vol/h_d^2) W_q; mapped cells: m_q = c_m JxW_q,
// G_q = JxW_q B B^T, stored per cell in E[cell][I][J]; coarse_gather_kernel sums them into the
// dense row-major matrix A in a fixed order (no atomics); rows and columns of constrained DoFs
// are left out (their identity diagonal is set separately).
template <typename T>
__global__ __launch_bounds__(256) void coarse_matrix_kernel(CoarseTables t, int n1, int nc0, int nc1, long ncells,
                                                            int nd0, int nd1, int nd2, int dmask, double cm,
                                                            double ks0, double ks1, double ks2, double c_mass,
                                                            const double *__restrict__ kc, double kconst,
                                                            const T *__restrict__ metric, double *__restrict__ E) {
  const int nloc = n1 * n1 * n1, K = n1 - 1;
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= ncells * nloc * nloc) return;
  const long cell = gid / ((long)nloc * nloc);
  const int e = (int)(gid - cell * nloc * nloc), I = e / nloc, J = e - I * nloc;
  const int cx = cell % nc0, cy = (cell / nc0) % nc1, cz = cell / ((long)nc0 * nc1);
  const int i0 = I % n1, i1 = (I / n1) % n1, i2 = I / (n1 * n1);
  const int j0 = J % n1, j1 = (J / n1) % n1, j2 = J / (n1 * n1);
  const double kap = kc ? kc[cell] : kconst;
  double acc = 0.0;
  for (int qz = 0; qz < n1; ++qz)
    for (int qy = 0; qy < n1; ++qy)
      for (int qx = 0; qx < n1; ++qx) {
        const double pI = t.S[qx][i0] * t.S[qy][i1] * t.S[qz][i2], pJ = t.S[qx][j0] * t.S[qy][j1] * t.S[qz][j2];
        const double a[3] = {t.G[qx][i0] * t.S[qy][i1] * t.S[qz][i2], t.S[qx][i0] * t.G[qy][i1] * t.S[qz][i2],
                             t.S[qx][i0] * t.S[qy][i1] * t.G[qz][i2]};
        const double b[3] = {t.G[qx][j0] * t.S[qy][j1] * t.S[qz][j2], t.S[qx][j0] * t.G[qy][j1] * t.S[qz][j2],
                             t.S[qx][j0] * t.S[qy][j1] * t.G[qz][j2]};
        if (metric) {
          const int q = qx + n1 * (qy + n1 * qz);
          const T *M = metric + cell * 7 * nloc + q;
          const double g00 = M[nloc], g01 = M[2 * nloc], g11 = M[3 * nloc], g02 = M[4 * nloc], g12 = M[5 * nloc],
                       g22 = M[6 * nloc];
          const double gg = a[0] * (g00 * b[0] + g01 * b[1] + g02 * b[2]) +
                            a[1] * (g01 * b[0] + g11 * b[1] + g12 * b[2]) +
                            a[2] * (g02 * b[0] + g12 * b[1] + g22 * b[2]);
          acc += c_mass * (double)M[0] * pI * pJ + kap * gg;
        } else {
          const double W = t.w[qx] * t.w[qy] * t.w[qz];
          acc += W * (cm * pI * pJ + kap * (ks0 * a[0] * b[0] + ks1 * a[1] * b[1] + ks2 * a[2] * b[2]));
        }
      }
  E[gid] = acc;
}

// A[I][J] = sum over the cells holding both DoFs (z, y, x order) of E[cell][I_loc][J_loc];
// along each direction the cells holding node index x are c with K c <= x <= K c + K
__global__ __launch_bounds__(256) void coarse_gather_kernel(const double *__restrict__ E, int n1, int nc0, int nc1,
                                                            int nc2, int nd0, int nd1, int nd2, int dmask,
                                                            double *__restrict__ A) {
  const int K = n1 - 1, nloc = n1 * n1 * n1;
  const long n = (long)nd0 * nd1 * nd2;
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= n * n) return;
  const long gI = gid / n, gJ = gid - gI * n;
  double acc = 0.0;
  if (!(dmask && (constrained(gI, nd0, nd1, nd2, dmask) || constrained(gJ, nd0, nd1, nd2, dmask)))) {
    const int ii[3] = {(int)(gI % nd0), (int)((gI / nd0) % nd1), (int)(gI / ((long)nd0 * nd1))};
    const int jj[3] = {(int)(gJ % nd0), (int)((gJ / nd0) % nd1), (int)(gJ / ((long)nd0 * nd1))};
    const int nc[3] = {nc0, nc1, nc2};
    int lo[3], hi[3];
    bool empty = false;
    for (int d = 0; d < 3; ++d) {
      const int mx = max(ii[d], jj[d]), mn = min(ii[d], jj[d]);
      lo[d] = max(0, (mx + K - 1) / K - 1);
      hi[d] = min(nc[d] - 1, mn / K);
      empty = empty || lo[d] > hi[d];
    }
    if (!empty)
      for (int cz = lo[2]; cz <= hi[2]; ++cz)
        for (int cy = lo[1]; cy <= hi[1]; ++cy)
          for (int cx = lo[0]; cx <= hi[0]; ++cx) {
            const long cell = cx + (long)nc0 * (cy + (long)nc1 * cz);
            const int I = (ii[0] - K * cx) + n1 * ((ii[1] - K * cy) + n1 * (ii[2] - K * cz));
            const int J = (jj[0] - K * cx) + n1 * ((jj[1] - K * cy) + n1 * (jj[2] - K * cz));
            acc += E[(cell * nloc + I) * nloc + J];
          }
  }
  A[gid] = acc;
}

__global__ __launch_bounds__(256) void coarse_dirichlet_kernel(double *A, long n, int nd0, int nd1, int nd2,
                                                               int dmask) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n && constrained(i, nd0, nd1, nd2, dmask)) A[i * n + i] = 1.0;
}

template <typename T>
int launch_coarse_matrix(const OperatorData<T> &od, double c_mass, const double *kc, double kconst, double *A,
                         double *E, hipStream_t s) {
  const FE1D &fe = fe1d(od.degree);
  CoarseTables t{};
  for (int a = 0; a < fe.n1; ++a) {
    t.w[a] = fe.qw[a];
    for (int b = 0; b < fe.n1; ++b) {
      t.S[a][b] = fe.S[a][b];
      t.G[a][b] = fe.G[a][b];
    }
  }
  const long n = od.n_dofs, nloc = (long)fe.n1 * fe.n1 * fe.n1, work = od.n_cells * nloc * nloc;
  const double vol = od.h[0] * od.h[1] * od.h[2];
  hipLaunchKernelGGL(coarse_matrix_kernel<T>, dim3((work + 255) / 256), dim3(256), 0, s, t, fe.n1, od.nc[0],
                     od.nc[1], od.n_cells, od.nd[0], od.nd[1], od.nd[2], od.dmask, c_mass * vol,
                     vol / (od.h[0] * od.h[0]), vol / (od.h[1] * od.h[1]), vol / (od.h[2] * od.h[2]), c_mass, kc,
                     kconst, od.metric, E);
  hipLaunchKernelGGL(coarse_gather_kernel, dim3((n * n + 255) / 256), dim3(256), 0, s, E, fe.n1, od.nc[0], od.nc[1],
                     od.nc[2], od.nd[0], od.nd[1], od.nd[2], od.dmask, A);
  if (od.dmask)
    hipLaunchKernelGGL(coarse_dirichlet_kernel, dim3((n + 255) / 256), dim3(256), 0, s, A, n, od.nd[0], od.nd[1],
                       od.nd[2], od.dmask);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// Blocked in-place Gauss-Jordan inversion. For the diagonal block [k0, k0 + nb) of the current
// matrix [[A11, A12], [A21, A22]] one step writes
//   A11 <- A11^-1,  A12 <- A11^-1 A12,  A22 <- A22 - A21 A11^-1 A12,  A21 <- -A21 A11^-1
// (the scalar step with 1x1 blocks is the textbook in-place Gauss-Jordan); after the last block
// the array holds A^-1. Four launches per block: diagonal block, row panel, rank-nb update of
// everything outside the block's rows and columns (which reads the old column panel and the new
// row panel), column panel.
constexpr int kGJB = 32;    // block width
constexpr int kGJTile = 64;  // update tile (64 x 64 per workgroup, 4 x 4 per thread)

__global__ __launch_bounds__(256) void gjb_diag_kernel(double *A, int n, int k0, int nb) {
  __shared__ double D[kGJB][kGJB + 1];
  for (int e = threadIdx.x; e < nb * nb; e += 256) {
    const int i = e / nb, j = e - i * nb;
    D[i][j] = A[(long)(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  constexpr int kPer = kGJB * kGJB / 256;
  for (int k = 0; k < nb; ++k) {
    const double piv = 1.0 / D[k][k];
    double v[kPer];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
      const int e = threadIdx.x + 256 * m, i = e / kGJB, j = e - i * kGJB;
      double x = 0.0;
      if (i < nb && j < nb) {
        if (i == k && j == k) x = piv;
        else if (i == k) x = D[k][j] * piv;
        else if (j == k) x = -D[i][k] * piv;
        else x = D[i][j] - D[i][k] * D[k][j] * piv;
      }
      v[m] = x;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
      const int e = threadIdx.x + 256 * m, i = e / kGJB, j = e - i * kGJB;
      if (i < nb && j < nb) D[i][j] = v[m];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < nb * nb; e += 256) {
    const int i = e / nb, j = e - i * nb;
    A[(long)(k0 + i) * n + k0 + j] = D[i][j];
  }
}

// A12 <- A11^-1 A12, one thread per column outside the block
__global__ __launch_bounds__(256) void gjb_row_kernel(double *A, int n, int k0, int nb) {
  __shared__ double D[kGJB][kGJB + 1];
  for (int e = threadIdx.x; e < nb * nb; e += 256) {
    const int i = e / nb, j = e - i * nb;
    D[i][j] = A[(long)(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n || (j >= k0 && j < k0 + nb)) return;
  double c[kGJB];
#pragma unroll
  for (int r = 0; r < kGJB; ++r) c[r] = r < nb ? A[(long)(k0 + r) * n + j] : 0.0;
#pragma unroll
  for (int r = 0; r < kGJB; ++r) {
    if (r >= nb) break;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kGJB; ++q) s += (q < nb ? D[r][q] : 0.0) * c[q];
    A[(long)(k0 + r) * n + j] = s;
  }
}

// A22 <- A22 - A21 A12 for rows and columns outside the block (A12 already updated)
__global__ __launch_bounds__(256) void gjb_update_kernel(double *A, int n, int k0, int nb) {
  __shared__ double Ct[kGJB][kGJTile + 1], Rt[kGJB][kGJTile + 1];
  const int i0 = blockIdx.y * kGJTile, j0 = blockIdx.x * kGJTile;
  for (int e = threadIdx.x; e < kGJTile * kGJB; e += 256) {
    const int ii = e / kGJB, c = e - ii * kGJB;
    Ct[c][ii] = (i0 + ii < n && c < nb) ? A[(long)(i0 + ii) * n + k0 + c] : 0.0;
  }
  for (int e = threadIdx.x; e < kGJTile * kGJB; e += 256) {
    const int c = e / kGJTile, jj = e - c * kGJTile;
    Rt[c][jj] = (j0 + jj < n && c < nb) ? A[(long)(k0 + c) * n + j0 + jj] : 0.0;
  }
  __syncthreads();
  const int ty = threadIdx.x / 16, tx = threadIdx.x - ty * 16;
  double acc[4][4] = {};
#pragma unroll 8
  for (int c = 0; c < kGJB; ++c) {
    double a[4], b[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      a[m] = Ct[c][ty + 16 * m];
      b[m] = Rt[c][tx + 16 * m];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int p = 0; p < 4; ++p) acc[m][p] = fma(a[m], b[p], acc[m][p]);
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int i = i0 + ty + 16 * m;
    if (i >= n || (i >= k0 && i < k0 + nb)) continue;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int j = j0 + tx + 16 * p;
      if (j < n && !(j >= k0 && j < k0 + nb)) A[(long)i * n + j] -= acc[m][p];
    }
  }
}

// A21 <- -A21 A11^-1, one thread per row outside the block
__global__ __launch_bounds__(256) void gjb_col_kernel(double *A, int n, int k0, int nb) {
  __shared__ double D[kGJB][kGJB + 1];
  for (int e = threadIdx.x; e < nb * nb; e += 256) {
    const int i = e / nb, j = e - i * nb;
    D[i][j] = A[(long)(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || (i >= k0 && i < k0 + nb)) return;
  double r[kGJB];
#pragma unroll
  for (int q = 0; q < kGJB; ++q) r[q] = q < nb ? A[(long)i * n + k0 + q] : 0.0;
#pragma unroll
  for (int c = 0; c < kGJB; ++c) {
    if (c >= nb) break;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kGJB; ++q) s += r[q] * (q < nb ? D[q][c] : 0.0);
    A[(long)i * n + k0 + c] = -s;
  }
}

int launch_gauss_jordan_inverse(double *A, int n, hipStream_t s) {
  const int b1 = (n + 255) / 256, bt = (n + kGJTile - 1) / kGJTile;
  for (int k0 = 0; k0 < n; k0 += kGJB) {
    const int nb = std::min(kGJB, n - k0);
    hipLaunchKernelGGL(gjb_diag_kernel, dim3(1), dim3(256), 0, s, A, n, k0, nb);
    hipLaunchKernelGGL(gjb_row_kernel, dim3(b1), dim3(256), 0, s, A, n, k0, nb);
    hipLaunchKernelGGL(gjb_update_kernel, dim3(bt, bt), dim3(256), 0, s, A, n, k0, nb);
    hipLaunchKernelGGL(gjb_col_kernel, dim3(b1), dim3(256), 0, s, A, n, k0, nb);
  }
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

template int launch_coarse_matrix<double>(const OperatorData<double> &, double, const double *, double, double *,
                                          double *, hipStream_t);
template int launch_coarse_matrix<float>(const OperatorData<float> &, double, const double *, double, double *,
                                         double *, hipStream_t);

}  // namespace mpdg
