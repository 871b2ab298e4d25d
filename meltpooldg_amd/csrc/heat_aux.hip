// Auxiliary kernels of the heat operator: diagonal, conductivity at quadrature
// points, cell averages for the multigrid levels, and the boundary-face terms.
//
// These run once per time step or once per Newton iteration, not once per
// Krylov iteration. The per-cell tensor work (k(T) at the quadrature points, the
// diagonal, cell averages) is sum-factorized through LDS; the rest favours simple
// indexing over tiling.
#include "blas.hpp"

namespace mpdg {

constexpr double kStefanBoltzmann = 5.670374419e-8;  // include/meltpooldg/utilities/physical_constants.hpp:6

template <typename Number>
struct AuxTables {
  Number S[kMaxN1][kMaxN1], G[kMaxN1][kMaxN1], w[kMaxN1];
};
template <typename Number>
AuxTables<Number> make_aux_tables(int k) {
  const FE1D &fe = fe1d(k);
  AuxTables<Number> t{};
  for (int a = 0; a < fe.n1; ++a) {
    t.w[a] = (Number)fe.qw[a];
    for (int b = 0; b < fe.n1; ++b) {
      t.S[a][b] = (Number)fe.S[a][b];
      t.G[a][b] = (Number)fe.G[a][b];
    }
  }
  return t;
}

template <typename Number>
__device__ __forceinline__ Number law_value(int law, const Number *p, Number T) {
  if (law == LAW_TANH) return p[0] * (Number(1) + p[1] * tanh((T - p[2]) / p[3]));
  const Number inv = Number(1) / (p[3] - p[2]);
  Number fs = (p[3] - T) * inv;
  fs = fs < Number(0) ? Number(0) : (fs > Number(1) ? Number(1) : fs);
  return p[0] + (p[1] - p[0]) * (Number(3) - Number(2) * fs) * fs * fs;
}

// Cell-local tensor kernels. A 256-thread workgroup holds kCellsPerBlock<N1> cells with N1^2
// threads per cell, one per line of the cell's N1^3 tensor; a contraction with a 1D matrix
// along each direction is one sweep in registers and one transpose through LDS, so a cell
// costs 3 N1^4 operations instead of the N1^6 of an element-by-element sum.
template <int N1>
constexpr int kCellsPerBlock = 256 / (N1 * N1);

template <int N1, typename Number>
__device__ __forceinline__ void cell_of(long cell, int nc0, int nc1, int *c) {
  c[0] = (int)(cell % nc0);
  c[1] = (int)((cell / nc0) % nc1);
  c[2] = (int)(cell / ((long)nc0 * nc1));
}

// k(T) at the quadrature points of every cell: T interpolated by sum factorization (S along x,
// y, z), the law evaluated per point. kq != nullptr: k per point, cell-major [cell][qz][qy][qx];
// kc != nullptr: the quadrature-weighted cell average sum_q w_q k_q instead.
template <typename Number, int N1>
__global__ __launch_bounds__(256) void kappa_law_kernel(AuxTables<Number> t, int nc0, int nc1, long ncells, int nd0,
                                                        int nd1, int law, Number p0, Number p1, Number p2, Number p3,
                                                        const Number *__restrict__ T, Number *__restrict__ kq,
                                                        Number *__restrict__ kc) {
  constexpr int L2 = N1 * N1, NQ = N1 * L2, CPB = kCellsPerBlock<N1>, K = N1 - 1;
  __shared__ Number buf[2][CPB * NQ];
  const int lc = threadIdx.x / L2, r = threadIdx.x - lc * L2, a = r / N1, b = r - a * N1;
  const long cell = (long)blockIdx.x * CPB + lc;
  const bool in_block = lc < CPB, active = in_block && cell < ncells;
  Number *B0 = buf[0] + (in_block ? lc : 0) * NQ, *B1 = buf[1] + (in_block ? lc : 0) * NQ;
  Number u[N1];
  // x: thread (l = a, j = b) interpolates its row of DoFs
  if (active) {
    int c[3];
    cell_of<N1, Number>(cell, nc0, nc1, c);
    const long row = (long)K * c[0] + (long)nd0 * ((K * c[1] + b) + (long)nd1 * (K * c[2] + a));
#pragma unroll
    for (int i = 0; i < N1; ++i) u[i] = T[row + i];
  } else {
#pragma unroll
    for (int i = 0; i < N1; ++i) u[i] = Number(0);
  }
  if (in_block) {
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      Number s = Number(0);
#pragma unroll
      for (int i = 0; i < N1; ++i) s += t.S[q][i] * u[i];
      B0[r * N1 + q] = s;  // [l][j][qx]
    }
  }
  __syncthreads();
  // y: thread (l = a, qx = b)
  if (in_block) {
#pragma unroll
    for (int j = 0; j < N1; ++j) u[j] = B0[(a * N1 + j) * N1 + b];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      Number s = Number(0);
#pragma unroll
      for (int j = 0; j < N1; ++j) s += t.S[q][j] * u[j];
      B1[(a * N1 + q) * N1 + b] = s;  // [l][qy][qx]
    }
  }
  __syncthreads();
  // z: thread (qy = a, qx = b), then the law
  const Number p[4] = {p0, p1, p2, p3};
  Number part = Number(0);
  if (in_block) {
#pragma unroll
    for (int l = 0; l < N1; ++l) u[l] = B1[(l * N1 + a) * N1 + b];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      Number s = Number(0);
#pragma unroll
      for (int l = 0; l < N1; ++l) s += t.S[q][l] * u[l];
      const Number k = law_value<Number>(law, p, s);
      if (kq && active) kq[cell * NQ + q * L2 + r] = k;
      part += t.w[q] * k;
    }
  }
  if (!kc) return;  // uniform across the grid: no thread waits below
  // cell average: the L2 partial sums of a cell through LDS (B0 is free again)
  if (in_block) B0[r] = part * t.w[a] * t.w[b];
  __syncthreads();
  if (active && r == 0) {
    Number s = Number(0);
    for (int i = 0; i < L2; ++i) s += B0[i];
    kc[cell] = s;
  }
}

// quadrature-weighted cell averages of per-point values, cell-major [cell][q]
template <typename Number, int N1>
__global__ __launch_bounds__(256) void cell_average_qp_kernel(AuxTables<Number> t, long ncells,
                                                              const Number *__restrict__ kq, Number *__restrict__ kc) {
  constexpr int L2 = N1 * N1, NQ = N1 * L2, CPB = kCellsPerBlock<N1>;
  __shared__ Number buf[CPB * L2];
  const int lc = threadIdx.x / L2, r = threadIdx.x - lc * L2, a = r / N1, b = r - a * N1;
  const long cell = (long)blockIdx.x * CPB + lc;
  const bool in_block = lc < CPB, active = in_block && cell < ncells;
  Number part = Number(0);
  if (active)
#pragma unroll
    for (int q = 0; q < N1; ++q) part += t.w[q] * kq[cell * NQ + q * L2 + r];
  if (in_block) buf[lc * L2 + r] = part * t.w[a] * t.w[b];
  __syncthreads();
  if (active && r == 0) {
    Number s = Number(0);
    for (int i = 0; i < L2; ++i) s += buf[lc * L2 + i];
    kc[cell] = s;
  }
}

// conductivity at every quadrature point of every cell, non-law modes
template <typename Number>
__global__ __launch_bounds__(256) void kappa_fill_kernel(long nq, long n, int mode, Number kappa,
                                                         const Number *__restrict__ karr, Number *__restrict__ kq) {
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= n) return;
  Number k = kappa;
  if (mode == KAPPA_QP) k = karr[gid];
  else if (mode == KAPPA_CELL) k = karr[gid / nq];
  kq[gid] = k;
}

template <typename Number, int N1>
static void launch_kappa_law_n1(const OperatorData<Number> &op, Number *kq, Number *kc, hipStream_t stream) {
  constexpr int CPB = kCellsPerBlock<N1>;
  hipLaunchKernelGGL((kappa_law_kernel<Number, N1>), dim3((op.n_cells + CPB - 1) / CPB), dim3(256), 0, stream,
                     make_aux_tables<Number>(op.degree), op.nc[0], op.nc[1], op.n_cells, op.nd[0], op.nd[1], op.law,
                     (Number)op.law_p[0], (Number)op.law_p[1], (Number)op.law_p[2], (Number)op.law_p[3], op.T, kq, kc);
}
template <typename Number>
static int launch_kappa_law(const OperatorData<Number> &op, Number *kq, Number *kc, hipStream_t stream) {
  switch (op.degree) {
    case 1: launch_kappa_law_n1<Number, 2>(op, kq, kc, stream); break;
    case 2: launch_kappa_law_n1<Number, 3>(op, kq, kc, stream); break;
    case 3: launch_kappa_law_n1<Number, 4>(op, kq, kc, stream); break;
    case 4: launch_kappa_law_n1<Number, 5>(op, kq, kc, stream); break;
    default: return MPDG_ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

template <typename Number>
int launch_kappa_qp(const OperatorData<Number> &op, Number *kq, hipStream_t stream) {
  if (op.kappa_mode == KAPPA_LAW) return launch_kappa_law<Number>(op, kq, nullptr, stream);
  const int n1 = op.degree + 1;
  const long nq = (long)n1 * n1 * n1, n = op.n_cells * nq;
  hipLaunchKernelGGL(kappa_fill_kernel<Number>, dim3((n + 255) / 256), dim3(256), 0, stream, nq, n, op.kappa_mode,
                     (Number)op.kappa, op.kappa_arr, kq);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

template <typename Number>
int launch_cell_average(const OperatorData<Number> &op, Number *kc, hipStream_t stream) {
  switch (op.kappa_mode) {
    case KAPPA_CONST: return launch_fill<Number>(kc, (Number)op.kappa, op.n_cells, stream);
    case KAPPA_CELL:
      return hipMemcpyAsync(kc, op.kappa_arr, sizeof(Number) * op.n_cells, hipMemcpyDeviceToDevice, stream) ==
                     hipSuccess
                 ? MPDG_OK
                 : MPDG_ERR_DEVICE;
    case KAPPA_LAW: return launch_kappa_law<Number>(op, nullptr, kc, stream);
    default: break;
  }
  const AuxTables<Number> t = make_aux_tables<Number>(op.degree);
  switch (op.degree) {
#define MPDG_CELL_AVG(N1)                                                                                       \
  case N1 - 1:                                                                                                  \
    hipLaunchKernelGGL((cell_average_qp_kernel<Number, N1>), dim3((op.n_cells + kCellsPerBlock<N1> - 1) /      \
                                                                  kCellsPerBlock<N1>),                          \
                       dim3(256), 0, stream, t, op.n_cells, op.kappa_arr, kc);                                 \
    break;
    MPDG_CELL_AVG(2)
    MPDG_CELL_AVG(3)
    MPDG_CELL_AVG(4)
    MPDG_CELL_AVG(5)
#undef MPDG_CELL_AVG
    default: return MPDG_ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// Diagonal of the cell matrices, summed into the global vector:
//   diag_I = sum_q c_q^(0) phi_I^2 + sum_(a,b) c_q^(ab) d_a phi_I d_b phi_I.
// Every term is a tensor product of 1D factors S^2, G^2 or G S, so each is three sum-factorized
// sweeps of its per-point coefficient. Cartesian cells: c^(0) = c_m vol W, c^(aa) = k vol/h_a^2 W;
// mapped cells: c^(0) = c_m JxW, c^(ab) = k (JxW B B^T)_ab (cross terms twice).
// diagonal_kernel stores each cell's N1^3 entries (cell-major, lexicographic inside the cell);
// diag_gather_kernel then sums, for every DoF, the entries of the (up to 8) cells that hold it in
// a fixed order: no atomics, and the result does not depend on the schedule.
template <typename Number>
__device__ __forceinline__ Number diag_factor(const AuxTables<Number> &t, int code, int q, int i) {
  const Number s = t.S[q][i], g = t.G[q][i];
  return code == 0 ? s * s : (code == 1 ? g * g : g * s);
}

template <typename Number, int N1, bool MAPPED>
__global__ __launch_bounds__(256) void diagonal_kernel(AuxTables<Number> t, int nc0, int nc1, long ncells, int nd0,
                                                       int nd1, Number cm, Number ks0, Number ks1, Number ks2,
                                                       int kmode, Number kappa, const Number *__restrict__ karr,
                                                       const Number *__restrict__ metric, Number *__restrict__ cellwise) {
  constexpr int L2 = N1 * N1, NQ = N1 * L2, CPB = kCellsPerBlock<N1>, K = N1 - 1, NT = MAPPED ? 7 : 4;
  // term: metric component, 1D factor codes along x, y, z (0: S^2, 1: G^2, 2: G S)
  constexpr int comp[7] = {0, 1, 3, 6, 2, 4, 5};
  constexpr int fx[7] = {0, 1, 0, 0, 2, 2, 0}, fy[7] = {0, 0, 1, 0, 2, 0, 2}, fz[7] = {0, 0, 0, 1, 0, 2, 2};
  __shared__ Number buf[2][CPB * NQ];
  const int lc = threadIdx.x / L2, r = threadIdx.x - lc * L2, a = r / N1, b = r - a * N1;
  const long cell = (long)blockIdx.x * CPB + lc;
  const bool in_block = lc < CPB, active = in_block && cell < ncells;
  Number *B0 = buf[0] + (in_block ? lc : 0) * NQ, *B1 = buf[1] + (in_block ? lc : 0) * NQ;
  const Number ks[3] = {ks0, ks1, ks2};
  // conductivity along this thread's qz line: constant, per point, or per cell
  Number kk[N1];
#pragma unroll
  for (int qz = 0; qz < N1; ++qz)
    kk[qz] = !active ? Number(0) : (kmode == 0 ? kappa : (kmode == 1 ? karr[cell * NQ + qz * L2 + r] : karr[cell]));
  Number d[N1], u[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) d[i] = Number(0);
#pragma unroll
  for (int term = 0; term < NT; ++term) {
    // z: thread (qy = a, qx = b) forms the coefficient along qz
    if (in_block) {
#pragma unroll
      for (int qz = 0; qz < N1; ++qz) {
        Number c = Number(0);
        if (active) {
          if (MAPPED) {
            const Number m = metric[(cell * 7 + comp[term]) * NQ + qz * L2 + r];
            c = term == 0 ? cm * m : (term >= 4 ? Number(2) : Number(1)) * kk[qz] * m;
          } else {
            const Number W = t.w[qz] * t.w[a] * t.w[b];
            c = term == 0 ? cm * W : ks[term - 1] * kk[qz] * W;
          }
        }
        u[qz] = c;
      }
#pragma unroll
      for (int l = 0; l < N1; ++l) {
        Number s = Number(0);
#pragma unroll
        for (int qz = 0; qz < N1; ++qz) s += diag_factor(t, fz[term], qz, l) * u[qz];
        B0[(l * N1 + a) * N1 + b] = s;  // [l][qy][qx]
      }
    }
    __syncthreads();
    // y: thread (l = a, qx = b)
    if (in_block) {
#pragma unroll
      for (int qy = 0; qy < N1; ++qy) u[qy] = B0[(a * N1 + qy) * N1 + b];
#pragma unroll
      for (int j = 0; j < N1; ++j) {
        Number s = Number(0);
#pragma unroll
        for (int qy = 0; qy < N1; ++qy) s += diag_factor(t, fy[term], qy, j) * u[qy];
        B1[(a * N1 + j) * N1 + b] = s;  // [l][j][qx]
      }
    }
    __syncthreads();
    // x: thread (l = a, j = b) accumulates its row of DoFs
    if (in_block) {
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) u[qx] = B1[r * N1 + qx];
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        Number s = Number(0);
#pragma unroll
        for (int qx = 0; qx < N1; ++qx) s += diag_factor(t, fx[term], qx, i) * u[qx];
        d[i] += s;
      }
    }
  }
  if (!active) return;
  Number *out = cellwise + cell * NQ + r * N1;
#pragma unroll
  for (int i = 0; i < N1; ++i) out[i] = d[i];
}

// diag[I] = sum over the cells holding DoF I of their entry for I; along each direction a DoF on
// a cell boundary belongs to the cells on both sides (lower cell first)
template <typename Number, int N1>
__global__ __launch_bounds__(256) void diag_gather_kernel(const Number *__restrict__ cellwise, int nc0, int nc1,
                                                          int nc2, Number *__restrict__ diag) {
  constexpr int K = N1 - 1, NQ = N1 * N1 * N1;
  const int nd0 = K * nc0 + 1, nd1 = K * nc1 + 1, nd2 = K * nc2 + 1;
  const long n = (long)nd0 * nd1 * nd2;
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= n) return;
  const int ix = gid % nd0, iy = (gid / nd0) % nd1, iz = gid / ((long)nd0 * nd1);
  // per direction: first cell, its local index, and whether the previous cell also holds the DoF
  int c[3], l[3], two[3];
  const int idx[3] = {ix, iy, iz}, nc[3] = {nc0, nc1, nc2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = min(idx[d] / K, nc[d] - 1);
    l[d] = idx[d] - K * c[d];
    two[d] = (l[d] == 0 && c[d] > 0) ? 1 : 0;
  }
  Number acc = Number(0);
  for (int sz = two[2]; sz >= 0; --sz)
    for (int sy = two[1]; sy >= 0; --sy)
      for (int sx = two[0]; sx >= 0; --sx) {
        const long cell = (c[0] - sx) + (long)nc0 * ((c[1] - sy) + (long)nc1 * (c[2] - sz));
        const int loc = (sx ? K : l[0]) + N1 * ((sy ? K : l[1]) + N1 * (sz ? K : l[2]));
        acc += cellwise[cell * NQ + loc];
      }
  diag[gid] = acc;
}

template <typename Number>
__global__ __launch_bounds__(256) void diag_dirichlet_kernel(Number *__restrict__ d, int nd0, int nd1, int nd2,
                                                             int mask) {
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= (long)nd0 * nd1 * nd2) return;
  const int ix = gid % nd0, iy = (gid / nd0) % nd1, iz = gid / ((long)nd0 * nd1);
  const bool con = ((mask & 1) && ix == 0) || ((mask & 2) && ix == nd0 - 1) || ((mask & 4) && iy == 0) ||
                   ((mask & 8) && iy == nd1 - 1) || ((mask & 16) && iz == 0) || ((mask & 32) && iz == nd2 - 1);
  if (con) d[gid] = Number(1);
}

// Boundary faces (Robin terms, surface flux); the sums over faces and cells per DoF are below.
// mode 0: Robin tangent applied to src: dst_I += sum_q phi_I (alpha + 4 eps sigma T_q^3) u_q JxW_q
//         (heat_diffuse_operator.cpp:1090-1129)
// mode 1: its diagonal
// mode 2: boundary residual -alpha (T - T_inf) - eps sigma (T^4 - T_inf^4) on Robin faces plus the
//         Gaussian surface flux on z+ (heat_diffuse_operator.cpp:721-781)
// On mapped meshes (coords set) the face geometry comes from the mapping's node coordinates on
// the face: JxW_q = |dx/dxi_a x dx/dxi_b| w_qa w_qb and the laser sees the mapped (x, y) of the
// point (FEFaceValues with MappingQ, which MatrixFree's face data hold).
struct FaceArgs {
  AuxTables<double> t;
  double qp[kMaxN1];
  int n1, nc[3], nd[3], robin_mask, face_mask, dmask, mode, flux_on;
  double h[3], alpha, emissivity, T_inf_conv, T_inf_rad;
  double flux[6];  // peak, sigma, x0, y0, vx, t
  const double *coords;  // null: the affine box
  long n_dofs;
};

__device__ __forceinline__ bool face_dof_constrained(const FaceArgs &A, long g) {
  const int ix = g % A.nd[0], iy = (g / A.nd[0]) % A.nd[1], iz = g / ((long)A.nd[0] * A.nd[1]);
  const int m = A.dmask;
  return ((m & 1) && ix == 0) || ((m & 2) && ix == A.nd[0] - 1) || ((m & 4) && iy == 0) ||
         ((m & 8) && iy == A.nd[1] - 1) || ((m & 16) && iz == 0) || ((m & 32) && iz == A.nd[2] - 1);
}

// contribution of face f of cell c to the face DoF (ia, ib) of that cell face
__device__ double face_contribution(const FaceArgs &A, const double *__restrict__ src, const double *__restrict__ T,
                                    int f, const int *c, int ia, int ib) {
  const int n1 = A.n1, K = n1 - 1;
  const int d = f / 2, s = f % 2;
  const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
  auto dof = [&](int xa, int xb) {
    int idx[3];
    idx[d] = K * c[d] + s * K;
    idx[ta] = K * c[ta] + xa;
    idx[tb] = K * c[tb] + xb;
    return (long)idx[0] + (long)A.nd[0] * (idx[1] + (long)A.nd[1] * idx[2]);
  };
  const bool robin = (A.robin_mask >> f) & 1;
  double acc = 0.0;
  for (int qb = 0; qb < n1; ++qb)
    for (int qa = 0; qa < n1; ++qa) {
      double Tq = 0.0, uq = 0.0;
      double dxa[3] = {0.0, 0.0, 0.0}, dxb[3] = {0.0, 0.0, 0.0}, xq[3] = {0.0, 0.0, 0.0};
      for (int xb = 0; xb < n1; ++xb)
        for (int xa = 0; xa < n1; ++xa) {
          const double phi = A.t.S[qa][xa] * A.t.S[qb][xb];
          const long g = dof(xa, xb);
          if (T) Tq += phi * T[g];
          if (A.mode == 0 && !face_dof_constrained(A, g)) uq += phi * src[g];
          if (A.coords) {
            const double ga = A.t.G[qa][xa] * A.t.S[qb][xb], gb = A.t.S[qa][xa] * A.t.G[qb][xb];
            for (int e = 0; e < 3; ++e) {
              const double X = A.coords[e * A.n_dofs + g];
              dxa[e] += ga * X;
              dxb[e] += gb * X;
              xq[e] += phi * X;
            }
          }
        }
      double JxW = A.h[ta] * A.h[tb] * A.t.w[qa] * A.t.w[qb];
      if (A.coords) {
        const double c0 = dxa[1] * dxb[2] - dxa[2] * dxb[1], c1 = dxa[2] * dxb[0] - dxa[0] * dxb[2],
                     c2 = dxa[0] * dxb[1] - dxa[1] * dxb[0];
        JxW = sqrt(c0 * c0 + c1 * c1 + c2 * c2) * A.t.w[qa] * A.t.w[qb];
      }
      const double phiI = A.t.S[qa][ia] * A.t.S[qb][ib];
      double val = 0.0;
      if (A.mode == 2) {
        if (robin && A.emissivity > 0) {
          const double T4 = Tq * Tq * Tq * Tq, Ti = A.T_inf_rad;
          val -= A.emissivity * kStefanBoltzmann * (T4 - Ti * Ti * Ti * Ti);
        }
        if (robin && A.alpha > 0) val -= A.alpha * (Tq - A.T_inf_conv);
        if (A.flux_on && f == 5) {
          double pos[3];
          pos[d] = (c[d] + s) * A.h[d];
          pos[ta] = (c[ta] + A.qp[qa]) * A.h[ta];
          pos[tb] = (c[tb] + A.qp[qb]) * A.h[tb];
          if (A.coords)
            for (int e = 0; e < 3; ++e) pos[e] = xq[e];
          const double dx = pos[0] - A.flux[2] - A.flux[4] * A.flux[5], dy = pos[1] - A.flux[3];
          val += A.flux[0] * exp(-(dx * dx + dy * dy) / (2.0 * A.flux[1] * A.flux[1]));
        }
      } else {
        double coef = 0.0;
        if (A.alpha > 0) coef += A.alpha;
        if (A.emissivity > 0) coef += 4.0 * A.emissivity * kStefanBoltzmann * Tq * Tq * Tq;
        val = coef * (A.mode == 0 ? uq : phiI);
      }
      acc += phiI * val * JxW;
    }
  return acc;
}

// One thread per boundary DoF: the DoFs of the six face grids, each handled by the first face
// (in the order x-, x+, y-, y+, z-, z+) that holds it. The thread sums the contributions of every
// selected face holding the DoF and, on each, of the (up to four) cells whose face holds it, in
// that fixed order, and adds the total to dst once: no atomics, so the result is bitwise
// repeatable.
__global__ __launch_bounds__(256) void boundary_face_kernel(FaceArgs A, const double *__restrict__ src,
                                                            const double *__restrict__ T, double *__restrict__ dst) {
  const int K = A.n1 - 1;
  long gid = (long)blockIdx.x * 256 + threadIdx.x;
  int idx[3];
  int owner = -1;
  for (int f = 0; f < 6 && owner < 0; ++f) {
    const int d = f / 2, s = f % 2, ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const long count = (long)A.nd[ta] * A.nd[tb];
    if (gid >= count) {
      gid -= count;
      continue;
    }
    idx[d] = s ? A.nd[d] - 1 : 0;
    idx[ta] = (int)(gid % A.nd[ta]);
    idx[tb] = (int)(gid / A.nd[ta]);
    owner = f;
  }
  if (owner < 0) return;
  auto on_face = [&](int f) { return idx[f / 2] == ((f % 2) ? A.nd[f / 2] - 1 : 0); };
  for (int f = 0; f < owner; ++f)
    if (on_face(f)) return;  // an earlier face owns this DoF
  const long gI = (long)idx[0] + (long)A.nd[0] * (idx[1] + (long)A.nd[1] * idx[2]);
  if (face_dof_constrained(A, gI)) return;
  double total = 0.0;
  bool any = false;
  for (int f = 0; f < 6; ++f) {
    if (!((A.face_mask >> f) & 1) || !on_face(f)) continue;
    const int d = f / 2, s = f % 2, ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    // cells along ta and tb whose face holds the DoF: the one at or below it, and the one
    // before when the DoF sits on their shared edge
    const int ca = min(idx[ta] / K, A.nc[ta] - 1), cb = min(idx[tb] / K, A.nc[tb] - 1);
    const int la = idx[ta] - K * ca, lb = idx[tb] - K * cb;
    const int two_a = (la == 0 && ca > 0) ? 1 : 0, two_b = (lb == 0 && cb > 0) ? 1 : 0;
    for (int sb = two_b; sb >= 0; --sb)
      for (int sa = two_a; sa >= 0; --sa) {
        int c[3];
        c[d] = s ? A.nc[d] - 1 : 0;
        c[ta] = ca - sa;
        c[tb] = cb - sb;
        total += face_contribution(A, src, T, f, c, sa ? K : la, sb ? K : lb);
        any = true;
      }
  }
  if (any) dst[gI] += total;
}

int launch_boundary(const OperatorData<double> &op, int mode, const double *flux, double *dst,
                    const double *src, const double *T, hipStream_t stream) {
  FaceArgs A{};
  A.t = make_aux_tables<double>(op.degree);
  const FE1D &fe = fe1d(op.degree);
  for (int i = 0; i < fe.n1; ++i) A.qp[i] = fe.qp[i];
  A.n1 = op.degree + 1;
  for (int d = 0; d < 3; ++d) {
    A.nc[d] = op.nc[d];
    A.nd[d] = op.nd[d];
    A.h[d] = op.h[d];
  }
  A.robin_mask = op.robin_mask;
  A.flux_on = (mode == 2 && flux && flux[0] != 0.0) ? 1 : 0;
  A.face_mask = op.robin_mask | (A.flux_on ? 32 : 0);
  if (!A.face_mask) return MPDG_OK;
  // a Neumann flux and a radiation/convection term on the same face are refused, as the
  // reference's rhs_boundary_loop does (heat_diffuse_operator.cpp:739-743)
  if (A.flux_on && (op.robin_mask & 32)) return MPDG_ERR_INVALID;
  A.coords = op.coords;
  A.n_dofs = op.n_dofs;
  A.dmask = op.dmask;
  A.mode = mode;
  A.alpha = op.alpha;
  A.emissivity = op.emissivity;
  A.T_inf_conv = op.T_inf_conv;
  A.T_inf_rad = op.T_inf_rad;
  if (flux)
    for (int i = 0; i < 6; ++i) A.flux[i] = flux[i];
  long total = 0;  // the DoFs of all six face grids (each boundary DoF has one owner among them)
  for (int f = 0; f < 6; ++f) {
    const int d = f / 2, ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    total += (long)op.nd[ta] * op.nd[tb];
  }
  hipLaunchKernelGGL(boundary_face_kernel, dim3((total + 255) / 256), dim3(256), 0, stream, A, src, T, dst);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

template <typename Number, int N1>
static void launch_diagonal_n1(const OperatorData<Number> &op, int kmode, const Number *karr, Number *cellwise,
                               Number *diag, hipStream_t stream) {
  constexpr int CPB = kCellsPerBlock<N1>;
  const double vol = op.metric ? 1.0 : op.h[0] * op.h[1] * op.h[2];
  const dim3 grid((op.n_cells + CPB - 1) / CPB);
  const AuxTables<Number> t = make_aux_tables<Number>(op.degree);
  const Number cm = (Number)(op.c_mass * vol), ks0 = (Number)(vol / (op.h[0] * op.h[0])),
               ks1 = (Number)(vol / (op.h[1] * op.h[1])), ks2 = (Number)(vol / (op.h[2] * op.h[2]));
  if (op.metric)
    hipLaunchKernelGGL((diagonal_kernel<Number, N1, true>), grid, dim3(256), 0, stream, t, op.nc[0], op.nc[1],
                       op.n_cells, op.nd[0], op.nd[1], cm, ks0, ks1, ks2, kmode, (Number)op.kappa, karr, op.metric,
                       cellwise);
  else
    hipLaunchKernelGGL((diagonal_kernel<Number, N1, false>), grid, dim3(256), 0, stream, t, op.nc[0], op.nc[1],
                       op.n_cells, op.nd[0], op.nd[1], cm, ks0, ks1, ks2, kmode, (Number)op.kappa, karr, op.metric,
                       cellwise);
  hipLaunchKernelGGL((diag_gather_kernel<Number, N1>), dim3((op.n_dofs + 255) / 256), dim3(256), 0, stream, cellwise,
                     op.nc[0], op.nc[1], op.nc[2], diag);
}

// scratch: 2 n_cells (k+1)^3 values (the cells' entries, and k at the points for a conductivity law)
template <typename Number>
int launch_diagonal(const OperatorData<Number> &op_in, Number *diag, Number *scratch, hipStream_t stream) {
  const long ncq = op_in.n_cells * (long)((op_in.degree + 1) * (op_in.degree + 1) * (op_in.degree + 1));
  Number *cellwise = scratch, *kq = scratch + ncq;
  const OperatorData<Number> op = kernel_view(op_in);
  if (op.metric == nullptr && op_in.metric != nullptr) return MPDG_ERR_INVALID;
  int kmode = 0;
  const Number *karr = nullptr;
  if (op.kappa_mode == KAPPA_QP || op.kappa_mode == KAPPA_CELL) {
    kmode = op.kappa_mode == KAPPA_QP ? 1 : 2;
    karr = op.kappa_arr;
  } else if (op.kappa_mode == KAPPA_LAW) {
    const int rc = launch_kappa_law<Number>(op, kq, nullptr, stream);
    if (rc) return rc;
    kmode = 1;
    karr = kq;
  }
  switch (op.degree) {
    case 1: launch_diagonal_n1<Number, 2>(op, kmode, karr, cellwise, diag, stream); break;
    case 2: launch_diagonal_n1<Number, 3>(op, kmode, karr, cellwise, diag, stream); break;
    case 3: launch_diagonal_n1<Number, 4>(op, kmode, karr, cellwise, diag, stream); break;
    case 4: launch_diagonal_n1<Number, 5>(op, kmode, karr, cellwise, diag, stream); break;
    default: return MPDG_ERR_INVALID;
  }
  if (op.dmask)
    hipLaunchKernelGGL(diag_dirichlet_kernel<Number>, dim3((op.n_dofs + 255) / 256), dim3(256), 0, stream, diag,
                       op.nd[0], op.nd[1], op.nd[2], op.dmask);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// average of the 8 children of every coarse cell
template <typename Number>
__global__ __launch_bounds__(256) void coarsen_cells_kernel(int f0, int f1, int f2, const Number *__restrict__ kf,
                                                            Number *__restrict__ kc) {
  const int c0 = f0 / 2, c1 = f1 / 2, c2 = f2 / 2;
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= (long)c0 * c1 * c2) return;
  const int x = c % c0, y = (c / c0) % c1, z = c / ((long)c0 * c1);
  Number s = 0;
  for (int i = 0; i < 8; ++i)
    s += kf[(2 * x + (i & 1)) + (long)f0 * ((2 * y + ((i >> 1) & 1)) + (long)f1 * (2 * z + (i >> 2)))];
  kc[c] = Number(0.125) * s;
}
template <typename Number>
int launch_coarsen_cells(const int *ncf, const Number *kf, Number *kc, hipStream_t stream) {
  const long n = (long)(ncf[0] / 2) * (ncf[1] / 2) * (ncf[2] / 2);
  hipLaunchKernelGGL(coarsen_cells_kernel<Number>, dim3((n + 255) / 256), dim3(256), 0, stream, ncf[0], ncf[1],
                     ncf[2], kf, kc);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// Geometry of mapped cells at every quadrature point from the mapping node coordinates
// (MappingQ(k) with the FE's support points, source/utilities/fe_util.cpp:18-28):
// J[a][b] = d x_a / d xi_b, B = J^-1, JxW = |det J| w, G = JxW B B^T.
template <typename Number>
__global__ __launch_bounds__(256) void metric_kernel(AuxTables<double> t, int n1, int nc0, int nc1, long ncells,
                                                     int nd0, int nd1, long ndofs, const double *__restrict__ X,
                                                     Number *__restrict__ metric) {
  const int nq = n1 * n1 * n1, K = n1 - 1;
  const long gid = (long)blockIdx.x * 256 + threadIdx.x;
  if (gid >= ncells * nq) return;
  const long cell = gid / nq;
  const int q = gid - cell * nq;
  const int qx = q % n1, qy = (q / n1) % n1, qz = q / (n1 * n1);
  const int cx = cell % nc0, cy = (cell / nc0) % nc1, cz = cell / ((long)nc0 * nc1);
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int l = 0; l < n1; ++l)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i) {
        const long g = (long)(K * cx + i) + (long)nd0 * ((K * cy + j) + (long)nd1 * (K * cz + l));
        const double d0 = t.G[qx][i] * t.S[qy][j] * t.S[qz][l], d1 = t.S[qx][i] * t.G[qy][j] * t.S[qz][l],
                     d2 = t.S[qx][i] * t.S[qy][j] * t.G[qz][l];
        for (int a = 0; a < 3; ++a) {
          const double x = X[a * ndofs + g];
          J[a][0] += x * d0;
          J[a][1] += x * d1;
          J[a][2] += x * d2;
        }
      }
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double id = 1.0 / det;
  double B[3][3];
  B[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  B[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  B[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  B[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  B[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  B[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  B[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  B[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  B[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  const double jxw = fabs(det) * t.w[qx] * t.w[qy] * t.w[qz];
  auto g = [&](int a, int b) { return jxw * (B[a][0] * B[b][0] + B[a][1] * B[b][1] + B[a][2] * B[b][2]); };
  Number *M = metric + cell * 7 * nq + q;
  M[0] = (Number)jxw;
  M[nq] = (Number)g(0, 0);
  M[2 * nq] = (Number)g(0, 1);
  M[3 * nq] = (Number)g(1, 1);
  M[4 * nq] = (Number)g(0, 2);
  M[5 * nq] = (Number)g(1, 2);
  M[6 * nq] = (Number)g(2, 2);
}

template <typename Number>
int launch_metric(const OperatorData<Number> &op, const double *coords, Number *metric, hipStream_t s) {
  const int n1 = op.degree + 1;
  const long n = op.n_cells * n1 * n1 * n1;
  hipLaunchKernelGGL(metric_kernel<Number>, dim3((n + 255) / 256), dim3(256), 0, s, make_aux_tables<double>(op.degree),
                     n1, op.nc[0], op.nc[1], op.n_cells, op.nd[0], op.nd[1], op.n_dofs, coords, metric);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
// metric_k[c][0][q] = metric[c][0][q] (JxW), metric_k[c][i][q] = kq[c][q] metric[c][i][q] (i = 1..6)
template <typename Number>
__global__ __launch_bounds__(256) void scale_metric_kernel(const Number *__restrict__ metric, const Number *__restrict__ kq,
                                                           Number *__restrict__ out, long ncells, int nq) {
  const long n = ncells * 7 * nq;
  for (long g = (long)blockIdx.x * 256 + threadIdx.x; g < n; g += (long)gridDim.x * 256) {
    const long c = g / (7 * nq);
    const int r = g - c * 7 * nq, comp = r / nq, q = r - comp * nq;
    out[g] = comp == 0 ? metric[g] : kq[c * nq + q] * metric[g];
  }
}
template <typename Number>
int launch_scale_metric(const OperatorData<Number> &op, const Number *kq, Number *metric_k, hipStream_t s) {
  const int n1 = op.degree + 1, nq = n1 * n1 * n1;
  const long n = op.n_cells * 7 * nq;
  hipLaunchKernelGGL(scale_metric_kernel<Number>, dim3(std::min<long>((n + 255) / 256, 16384)), dim3(256), 0, s,
                     op.metric, kq, metric_k, op.n_cells, nq);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_scale_metric<double>(const OperatorData<double> &, const double *, double *, hipStream_t);
template int launch_scale_metric<float>(const OperatorData<float> &, const float *, float *, hipStream_t);

template int launch_metric<double>(const OperatorData<double> &, const double *, double *, hipStream_t);
template int launch_metric<float>(const OperatorData<float> &, const double *, float *, hipStream_t);

// nodal values of the z+ surface flux q = peak exp(-((x - x0 - vx t)^2 + (y - y0)^2)/(2 sigma^2)), 0 elsewhere
// the flux at the nodes of face z+ (0 elsewhere); mapped meshes: at the nodes' mapped (x, y)
__global__ __launch_bounds__(256) void surface_flux_field_kernel(double *__restrict__ out, int nd0, int nd1, int nd2,
                                                                 int k, double h0, double h1, FluxNodes fn,
                                                                 const double *__restrict__ coords) {
  const long n = (long)nd0 * nd1 * nd2;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int ix = i % nd0, iy = (i / nd0) % nd1, iz = i / ((long)nd0 * nd1);
    double v = 0.0;
    if (iz == nd2 - 1) {
      const int cx = min(ix / k, (nd0 - 1) / k - 1), cy = min(iy / k, (nd1 - 1) / k - 1);
      double x = (cx + fn.nodes[ix - k * cx]) * h0, y = (cy + fn.nodes[iy - k * cy]) * h1;
      if (coords) {
        x = coords[i];
        y = coords[n + i];
      }
      const double dx = x - fn.p[2] - fn.p[4] * fn.p[5], dy = y - fn.p[3];
      v = fn.p[0] * exp(-(dx * dx + dy * dy) / (2.0 * fn.p[1] * fn.p[1]));
    }
    out[i] = v;
  }
}
int launch_surface_flux_field(const OperatorData<double> &op, const double *p6, double *out, hipStream_t s) {
  FluxNodes fn{};
  const FE1D &fe = fe1d(op.degree);
  for (int i = 0; i < fe.n1; ++i) fn.nodes[i] = fe.nodes[i];
  for (int i = 0; i < 6; ++i) fn.p[i] = p6[i];
  hipLaunchKernelGGL(surface_flux_field_kernel, dim3(std::min<long>((op.n_dofs + 255) / 256, 8192)), dim3(256), 0, s,
                     out, op.nd[0], op.nd[1], op.nd[2], op.degree, op.h[0], op.h[1], fn, op.coords);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

template int launch_kappa_qp<double>(const OperatorData<double> &, double *, hipStream_t);
template int launch_kappa_qp<float>(const OperatorData<float> &, float *, hipStream_t);
template int launch_diagonal<double>(const OperatorData<double> &, double *, double *, hipStream_t);
template int launch_diagonal<float>(const OperatorData<float> &, float *, float *, hipStream_t);
template int launch_cell_average<double>(const OperatorData<double> &, double *, hipStream_t);
template int launch_cell_average<float>(const OperatorData<float> &, float *, hipStream_t);
template int launch_coarsen_cells<double>(const int *, const double *, double *, hipStream_t);
template int launch_coarsen_cells<float>(const int *, const float *, float *, hipStream_t);

}  // namespace mpdg
