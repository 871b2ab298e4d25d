// Full Newton tangent of the heat residual on Cartesian cells: temperature-dependent
// material (liquid/solid mushy zone) and a prescribed velocity, constant or a velocity DoF field
// on the temperature's nodes (HeatDiffuseMultiPhaseOperator::tangent_local_cell_operation with
// do_solidification and a velocity, source/heat/heat_diffuse_operator.cpp:975-978, :1028-1056;
// residual :449-452, :467-527).
//
// Per quadrature point the tangent is val = a u + c . grad u, flux = k grad u + b u. The
// coefficients depend on T, T_old and the velocity only, so they are evaluated once per Newton
// iteration (tangent_data_kernel) into kTangentCoeffs values per point, with the quadrature
// weight, the cell volume and the 1/h factors folded in (v = the velocity at the point):
//   t0 = JxW (rho_cp/dt [+ d rho_cp/dT ((T - T_old)/dt + grad T . v)])
//   t1 = JxW k                                           (times 1/h_d^2 d u / d xi_d)
//   t2..t4 = JxW rho_cp v_d / h_d          [velocity]     (times d u / d xi_d)
//   t5..t7 = JxW dk/dT (grad T)_d / h_d  [solidification] (gradient slot, times u)
// Coefficients a problem does not have (no velocity, no solidification) are neither written
// nor read. The apply, residual and diagonal kernels are cell-local sum factorization with one
// thread per line of the cell's N1^3 tensor (N1^2 threads per cell) and LDS transposes between
// the directions. A block holds a segment of an x-row of cells; it loads and stores its DoFs as
// N1^2 lines along x with consecutive lanes on consecutive addresses, sums the shared x-faces in
// LDS, stores the DoFs only it touches and adds the rest to dst. The kernels that add run in
// eight launches, one per colour (parity of the segment, of the row's y and of its z index):
// blocks of one colour share no DoF, so the adds need no atomics, and the colours follow each
// other in a fixed order, which makes the result bitwise repeatable.
#include "blas.hpp"
#include "tangent.hpp"

namespace mpdg {

namespace {

struct Tables1D {
  double S[kMaxN1][kMaxN1], G[kMaxN1][kMaxN1], w[kMaxN1];
};
Tables1D make_tables(int k) {
  const FE1D &fe = fe1d(k);
  Tables1D t{};
  for (int a = 0; a < fe.n1; ++a) {
    t.w[a] = fe.qw[a];
    for (int b = 0; b < fe.n1; ++b) {
      t.S[a][b] = fe.S[a][b];
      t.G[a][b] = fe.G[a][b];
    }
  }
  return t;
}

template <int N1>
constexpr int kCells = 256 / (N1 * N1);

struct Geometry {
  int nc0, nc1, nc2;
  int color;  // 0..7: the blocks of one colour (segment, y, z parities); -1: all blocks
  long ncells;
  int nd[3];
  int dmask;
  double h[3], ih2[3], vol, dt;
  int conv, solid;  // coefficients t2..t4 / t5..t7 present
};

__device__ __forceinline__ bool is_constrained(int m, const int *nd, int ix, int iy, int iz) {
  return ((m & 1) && ix == 0) || ((m & 2) && ix == nd[0] - 1) || ((m & 4) && iy == 0) ||
         ((m & 8) && iy == nd[1] - 1) || ((m & 16) && iz == 0) || ((m & 32) && iz == nd[2] - 1);
}

// Blocks walk the x-rows of cells: block = (row of cells, segment of kCells cells along x; the
// last segment of a row may be shorter). The block's DoFs form N1^2 lines (a, b) along x of
// n_cells_in_block K + 1 DoFs each.
template <int N1>
constexpr int kLine = kCells<N1> * (N1 - 1) + 1;

inline long tangent_blocks(int n1, const int *nc, int color = -1) {
  const int cpb = 256 / (n1 * n1);
  const long nseg = (nc[0] + cpb - 1) / cpb;
  if (color < 0) return (long)nc[1] * nc[2] * nseg;
  const int px = color & 1, py = (color >> 1) & 1, pz = (color >> 2) & 1;
  return ((nseg - px + 1) / 2) * (long)((nc[1] - py + 1) / 2) * ((nc[2] - pz + 1) / 2);
}

// The thread's place: cell `cell` (valid when `active`), line (a, b) of the cell's tensor;
// block-uniform: ncb cells in this block, first DoF (ix0, iy0, iz0) of its lines.
template <int N1>
struct Place {
  int lc, r, a, b, ncb, ix0, iy0, iz0;
  long cell;
  bool in_block, active;
  __device__ Place(const Geometry &g) {
    constexpr int L2 = N1 * N1, K = N1 - 1, CPB = kCells<N1>;
    lc = threadIdx.x / L2;
    r = threadIdx.x - lc * L2;
    a = r / N1;
    b = r - a * N1;
    const int nseg = (g.nc0 + CPB - 1) / CPB;
    int seg, cy, cz;
    if (g.color < 0) {
      const long row = blockIdx.x / nseg;
      seg = (int)(blockIdx.x - row * nseg);
      cy = (int)(row % g.nc1);
      cz = (int)(row / g.nc1);
    } else {
      const int px = g.color & 1, py = (g.color >> 1) & 1, pz = (g.color >> 2) & 1;
      const int nseg_c = (nseg - px + 1) / 2, ny_c = (g.nc1 - py + 1) / 2;
      const long rest = blockIdx.x / nseg_c;
      seg = 2 * (int)(blockIdx.x - rest * nseg_c) + px;
      cy = 2 * (int)(rest % ny_c) + py;
      cz = 2 * (int)(rest / ny_c) + pz;
    }
    const long row = cy + (long)g.nc1 * cz;
    const int x0 = seg * CPB;
    ncb = min(CPB, g.nc0 - x0);
    ix0 = K * x0;
    iy0 = K * cy;
    iz0 = K * cz;
    in_block = lc < CPB;
    active = lc < ncb;
    cell = row * g.nc0 + x0 + lc;
  }
  // DoF indices of entry e = r' kLine + x of the block's lines (r' = a' N1 + b': z-line a',
  // y-line b')
  __device__ void at(const Geometry &g, int e, int &la, int &lb, int &x, int &ix, int &iy, int &iz, long &gi) const {
    const int rr = e / kLine<N1>;
    x = e - rr * kLine<N1>;
    la = rr / N1;
    lb = rr - la * N1;
    ix = ix0 + x;
    iy = iy0 + lb;
    iz = iz0 + la;
    gi = ix + (long)g.nd[0] * (iy + (long)g.nd[1] * iz);
  }
};

// Values and reference gradients at the quadrature points (qz = 0..N1-1 of this thread's
// (qy = a, qx = b) column) of the field whose DoF row (l = a, j = b) the thread holds in u.
// X0, X1: 2 N1^3 LDS values of this cell; Y: 3 N1^3. Ends with the Y arrays still being read
// (a following evaluate() may start at once: its first barrier precedes its Y writes).
template <int N1, bool GRAD>
__device__ void evaluate(const Tables1D &t, const Place<N1> &p, const double *u, double *X0, double *X1,
                         double *Y0, double *Y1, double *Y2, double *val, double *dx, double *dy, double *dz) {
  if (p.in_block) {
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      double s = 0.0, d = 0.0;
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        s += t.S[q][i] * u[i];
        if (GRAD) d += t.G[q][i] * u[i];
      }
      X0[p.r * N1 + q] = s;  // [l][j][qx]
      if (GRAD) X1[p.r * N1 + q] = d;
    }
  }
  __syncthreads();
  if (p.in_block) {  // thread (l = a, qx = b)
    double c0[N1], c1[N1];
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      c0[j] = X0[(p.a * N1 + j) * N1 + p.b];
      if (GRAD) c1[j] = X1[(p.a * N1 + j) * N1 + p.b];
    }
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      double ss = 0.0, gs = 0.0, sg = 0.0;
#pragma unroll
      for (int j = 0; j < N1; ++j) {
        ss += t.S[q][j] * c0[j];
        if (GRAD) {
          gs += t.G[q][j] * c0[j];
          sg += t.S[q][j] * c1[j];
        }
      }
      const int o = (p.a * N1 + q) * N1 + p.b;  // [l][qy][qx]
      Y0[o] = ss;
      if (GRAD) {
        Y1[o] = gs;
        Y2[o] = sg;
      }
    }
  }
  __syncthreads();
  if (p.in_block) {  // thread (qy = a, qx = b)
    double c0[N1], c1[N1], c2[N1];
#pragma unroll
    for (int l = 0; l < N1; ++l) {
      const int o = (l * N1 + p.a) * N1 + p.b;
      c0[l] = Y0[o];
      if (GRAD) {
        c1[l] = Y1[o];
        c2[l] = Y2[o];
      }
    }
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      double v = 0.0, gz = 0.0, gy = 0.0, gx = 0.0;
#pragma unroll
      for (int l = 0; l < N1; ++l) {
        v += t.S[q][l] * c0[l];
        if (GRAD) {
          gz += t.G[q][l] * c0[l];
          gy += t.S[q][l] * c1[l];
          gx += t.S[q][l] * c2[l];
        }
      }
      val[q] = v;
      if (GRAD) {
        dx[q] = gx;
        dy[q] = gy;
        dz[q] = gz;
      }
    }
  }
}

// out[i] (this thread's DoF row) = sum_q v_q phi_i + sum_d f_d,q d phi_i / d xi_d, from the
// per-point values of this thread's (qy, qx) column. Starts with a barrier, so any LDS array
// may be reused after an evaluate().
template <int N1>
__device__ void integrate(const Tables1D &t, const Place<N1> &p, const double *v, const double *fx,
                          const double *fy, const double *fz, double *X0, double *X1, double *X2, double *Y0,
                          double *Y1, double *out) {
  __syncthreads();
  if (p.in_block) {  // z^T, thread (qy = a, qx = b)
#pragma unroll
    for (int l = 0; l < N1; ++l) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
      for (int q = 0; q < N1; ++q) {
        s0 += t.S[q][l] * v[q] + t.G[q][l] * fz[q];
        s1 += t.S[q][l] * fy[q];
        s2 += t.S[q][l] * fx[q];
      }
      const int o = (l * N1 + p.a) * N1 + p.b;  // [l][qy][qx]
      X0[o] = s0;
      X1[o] = s1;
      X2[o] = s2;
    }
  }
  __syncthreads();
  if (p.in_block) {  // y^T, thread (l = a, qx = b)
    double c0[N1], c1[N1], c2[N1];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      const int o = (p.a * N1 + q) * N1 + p.b;
      c0[q] = X0[o];
      c1[q] = X1[o];
      c2[q] = X2[o];
    }
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int q = 0; q < N1; ++q) {
        s0 += t.S[q][j] * c0[q] + t.G[q][j] * c1[q];
        s1 += t.S[q][j] * c2[q];
      }
      const int o = (p.a * N1 + j) * N1 + p.b;  // [l][j][qx]
      Y0[o] = s0;
      Y1[o] = s1;
    }
  }
  __syncthreads();
  if (p.in_block) {  // x^T, thread (l = a, j = b)
    double c0[N1], c1[N1];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      c0[q] = Y0[p.r * N1 + q];
      c1[q] = Y1[p.r * N1 + q];
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < N1; ++q) s += t.S[q][i] * c0[q] + t.G[q][i] * c1[q];
      out[i] = s;
    }
  }
}

// The thread's DoF row (l = a, j = b) of a global vector in u (constrained entries read as 0
// when `zero_con`; 0 for threads without a cell): the block loads its lines into `line`
// (kCells N1^3 LDS values; any array: starts with a barrier) with consecutive lanes on
// consecutive addresses, then each thread takes its row.
template <int N1>
__device__ void gather(const Geometry &g, const Place<N1> &p, const double *src, bool zero_con, double *line,
                       double *u) {
  constexpr int K = N1 - 1, NE = N1 * N1 * kLine<N1>;
  static_assert(NE <= kCells<N1> * N1 * N1 * N1, "block lines exceed the LDS array");
  const int len = p.ncb * K + 1;
  __syncthreads();
  for (int e = threadIdx.x; e < NE; e += blockDim.x) {
    int la, lb, x, ix, iy, iz;
    long gi;
    p.at(g, e, la, lb, x, ix, iy, iz, gi);
    if (x < len) line[e] = (zero_con && g.dmask && is_constrained(g.dmask, g.nd, ix, iy, iz)) ? 0.0 : src[gi];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N1; ++i) u[i] = p.active ? line[p.r * kLine<N1> + p.lc * K + i] : 0.0;
}

// dst += the threads' DoF rows (l = a, j = b) in out: the rows are summed into the block's
// lines in LDS (shared x-faces of neighbouring cells added there), then written with
// consecutive lanes on consecutive addresses. With ZEROED (dst cleared before the launch) the
// DoFs that only this block touches -- y, z interior to the cells, x interior to the block --
// are stored, the others added (no other block of the launch's colour touches them).
// Constrained rows are left alone.
template <int N1, bool ZEROED>
__device__ void scatter(const Geometry &g, const Place<N1> &p, double *dst, const double *out, double *line) {
  constexpr int K = N1 - 1, NE = N1 * N1 * kLine<N1>;
  const int len = p.ncb * K + 1;
  double *my = line + (p.active ? p.r * kLine<N1> + p.lc * K : 0);
  __syncthreads();
  if (p.active) {
#pragma unroll
    for (int i = 0; i < K; ++i) my[i] = out[i];
    if (p.lc == p.ncb - 1) my[K] = out[K];
  }
  __syncthreads();
  if (p.active && p.lc < p.ncb - 1) my[K] += out[K];  // first DoF of the next cell's row
  __syncthreads();
  for (int e = threadIdx.x; e < NE; e += blockDim.x) {
    int la, lb, x, ix, iy, iz;
    long gi;
    p.at(g, e, la, lb, x, ix, iy, iz, gi);
    if (x >= len || (g.dmask && is_constrained(g.dmask, g.nd, ix, iy, iz))) continue;
    const bool own = ZEROED && la > 0 && la < K && lb > 0 && lb < K && x > 0 && x < len - 1;
    if (own)
      dst[gi] = line[e];
    else
      dst[gi] += line[e];
  }
}

// material of the tangent (the oracle's `material`, material.templates.hpp:296-380, :546-598, :765-778)
__device__ __forceinline__ void material(const mpdg_heat_material &m, double T, double &rc, double &k, double &drc,
                                         double &dk) {
  if (m.transition == MPDG_TRANSITION_NONE) {
    rc = m.density[0] * m.specific_heat[0];
    k = m.conductivity[0];
    drc = dk = 0.0;
    return;
  }
  const double inv = 1.0 / (m.liquidus - m.solidus);
  const double fs = fmin(1.0, fmax(0.0, (m.liquidus - T) * inv));
  const bool pure = fs == 0.0 || fs == 1.0;
  const double c = 3.0 * fs * fs - 2.0 * fs * fs * fs, dc = pure ? 0.0 : -inv * (6.0 * fs - 6.0 * fs * fs);
  auto blend = [&](const double *x) { return fs == 0.0 ? x[0] : (fs == 1.0 ? x[1] : x[0] + (x[1] - x[0]) * c); };
  const double rho = blend(m.density), cp = blend(m.specific_heat);
  k = blend(m.conductivity);
  dk = (m.conductivity[1] - m.conductivity[0]) * dc;
  rc = rho * cp;
  drc = (m.specific_heat[1] - m.specific_heat[0]) * dc * rho + (m.density[1] - m.density[0]) * dc * cp;
}

struct TangentArgs {
  Geometry g;
  mpdg_heat_material mat;
  double vel[3];
  const double *vf;  // velocity DoF field (3 blocks of n_dofs) or null: vel
  int solid;
};

// velocity at the quadrature points (qz = 0..N1-1 of this thread's (qy, qx) column): the
// constant vector, or the field's values (gathered into lds[0], evaluated through lds[1], lds[3])
template <int N1>
__device__ void velocity_at_points(const Tables1D &t, const TangentArgs &A, const Place<N1> &p, double (*lds)[kCells<N1> * N1 * N1 * N1],
                                   int o, double (&w)[3][N1]) {
  if (!A.vf) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < N1; ++q) w[c][q] = A.vel[c];
    return;
  }
  const long n = (long)A.g.nd[0] * A.g.nd[1] * A.g.nd[2];
  double u[N1];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    gather<N1>(A.g, p, A.vf + c * n, false, lds[0], u);
    evaluate<N1, false>(t, p, u, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, lds[0] + o, w[c], nullptr, nullptr,
                        nullptr);
  }
}

template <int N1>
__global__ __launch_bounds__(256) void tangent_data_kernel(Tables1D t, TangentArgs A, const double *__restrict__ T,
                                                           const double *__restrict__ T_old, double *__restrict__ td) {
  constexpr int NQ = N1 * N1 * N1, CPB = kCells<N1>;
  __shared__ double lds[5][CPB * NQ];
  const Place<N1> p(A.g);
  const int o = (p.in_block ? p.lc : 0) * NQ;
  double u[N1], tq[N1], gx[N1], gy[N1], gz[N1], to[N1];
  gather<N1>(A.g, p, T, false, lds[4], u);
  evaluate<N1, true>(t, p, u, lds[0] + o, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, tq, gx, gy, gz);
  gather<N1>(A.g, p, T_old, false, lds[0], u);
  evaluate<N1, false>(t, p, u, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, lds[0] + o, to, nullptr, nullptr,
                      nullptr);
  double w[3][N1];
  velocity_at_points<N1>(t, A, p, lds, o, w);
  if (!p.active) return;
  const Geometry &g = A.g;
#pragma unroll
  for (int qz = 0; qz < N1; ++qz) {
    const double JxW = g.vol * t.w[qz] * t.w[p.a] * t.w[p.b];
    double rc, k, drc, dk;
    material(A.mat, tq[qz], rc, k, drc, dk);
    const double gT[3] = {gx[qz] / g.h[0], gy[qz] / g.h[1], gz[qz] / g.h[2]};
    double m = rc / g.dt;
    if (A.solid) m += drc * ((tq[qz] - to[qz]) / g.dt + gT[0] * w[0][qz] + gT[1] * w[1][qz] + gT[2] * w[2][qz]);
    double *c = td + p.cell * kTangentCoeffs * NQ + qz * N1 * N1 + p.r;
    c[0] = JxW * m;
    c[NQ] = JxW * k;
    if (g.conv)
      for (int d = 0; d < 3; ++d) c[(2 + d) * NQ] = JxW * rc * w[d][qz] / g.h[d];
    if (g.solid)
      for (int d = 0; d < 3; ++d) c[(5 + d) * NQ] = JxW * dk * gT[d] / g.h[d];
  }
}

template <int N1>
__global__ __launch_bounds__(256) void tangent_apply_kernel(Tables1D t, Geometry g, const double *__restrict__ td,
                                                            const double *__restrict__ src, double *__restrict__ dst) {
  constexpr int NQ = N1 * N1 * N1, CPB = kCells<N1>;
  __shared__ double lds[5][CPB * NQ];
  const Place<N1> p(g);
  const int o = (p.in_block ? p.lc : 0) * NQ;
  double u[N1], v[N1], gx[N1], gy[N1], gz[N1];
  gather<N1>(g, p, src, true, lds[4], u);
  evaluate<N1, true>(t, p, u, lds[0] + o, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, v, gx, gy, gz);
  if (p.active) {
#pragma unroll
    for (int qz = 0; qz < N1; ++qz) {
      const double *c = td + p.cell * kTangentCoeffs * NQ + qz * N1 * N1 + p.r;
      const double uq = v[qz], ax = gx[qz], ay = gy[qz], az = gz[qz], k = c[NQ];
      double val = c[0] * uq;
      if (g.conv) val += c[2 * NQ] * ax + c[3 * NQ] * ay + c[4 * NQ] * az;
      v[qz] = val;
      gx[qz] = k * g.ih2[0] * ax;
      gy[qz] = k * g.ih2[1] * ay;
      gz[qz] = k * g.ih2[2] * az;
      if (g.solid) {
        gx[qz] += c[5 * NQ] * uq;
        gy[qz] += c[6 * NQ] * uq;
        gz[qz] += c[7 * NQ] * uq;
      }
    }
  }
  integrate<N1>(t, p, v, gx, gy, gz, lds[0] + o, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, u);
  scatter<N1, true>(g, p, dst, u, lds[0]);
}

// dst += -R(T): -[rho_cp/dt (T - T_old) - q + rho_cp v . grad T] N - k grad T . grad N
template <int N1>
__global__ __launch_bounds__(256) void tangent_residual_kernel(Tables1D t, TangentArgs A, const double *__restrict__ T,
                                                               const double *__restrict__ T_old,
                                                               const double *__restrict__ qv, double *__restrict__ dst) {
  constexpr int NQ = N1 * N1 * N1, CPB = kCells<N1>;
  __shared__ double lds[5][CPB * NQ];
  const Place<N1> p(A.g);
  const Geometry &g = A.g;
  const int o = (p.in_block ? p.lc : 0) * NQ;
  double u[N1], tq[N1], gx[N1], gy[N1], gz[N1], to[N1], qq[N1];
  gather<N1>(g, p, T, false, lds[4], u);
  evaluate<N1, true>(t, p, u, lds[0] + o, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, tq, gx, gy, gz);
  gather<N1>(g, p, T_old, false, lds[0], u);
  evaluate<N1, false>(t, p, u, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, lds[0] + o, to, nullptr, nullptr,
                      nullptr);
#pragma unroll
  for (int q = 0; q < N1; ++q) qq[q] = 0.0;
  if (qv) {
    gather<N1>(g, p, qv, false, lds[0], u);
    evaluate<N1, false>(t, p, u, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, lds[0] + o, qq, nullptr, nullptr,
                        nullptr);
  }
  double w[3][N1];
  velocity_at_points<N1>(t, A, p, lds, o, w);
  if (p.in_block) {
#pragma unroll
    for (int qz = 0; qz < N1; ++qz) {
      const double JxW = g.vol * t.w[qz] * t.w[p.a] * t.w[p.b];
      double rc, k, drc, dk;
      material(A.mat, tq[qz], rc, k, drc, dk);
      const double gT[3] = {gx[qz] / g.h[0], gy[qz] / g.h[1], gz[qz] / g.h[2]};
      const double val = rc / g.dt * (tq[qz] - to[qz]) - qq[qz] +
                         rc * (gT[0] * w[0][qz] + gT[1] * w[1][qz] + gT[2] * w[2][qz]);
      tq[qz] = -JxW * val;
      gx[qz] = -JxW * k * gT[0] / g.h[0];
      gy[qz] = -JxW * k * gT[1] / g.h[1];
      gz[qz] = -JxW * k * gT[2] / g.h[2];
    }
  }
  integrate<N1>(t, p, tq, gx, gy, gz, lds[0] + o, lds[1] + o, lds[2] + o, lds[3] + o, lds[4] + o, u);
  scatter<N1, false>(g, p, dst, u, lds[0]);
}

__device__ __forceinline__ double factor(const Tables1D &t, int code, int q, int i) {
  const double s = t.S[q][i], d = t.G[q][i];
  return code == 0 ? s * s : (code == 1 ? d * d : d * s);
}

// diag_I = sum_q t0 phi^2 + sum_d (t_{2+d} + t_{5+d}) phi d_d phi + t1/h_d^2 (d_d phi)^2:
// seven tensor-product terms with 1D factors S^2 (0), G^2 (1), G S (2), three sweeps each
template <int N1>
__global__ __launch_bounds__(256) void tangent_diagonal_kernel(Tables1D t, Geometry g, const double *__restrict__ td,
                                                               double *__restrict__ diag) {
  constexpr int NQ = N1 * N1 * N1, CPB = kCells<N1>;
  constexpr int fx[7] = {0, 2, 0, 0, 1, 0, 0}, fy[7] = {0, 0, 2, 0, 0, 1, 0}, fz[7] = {0, 0, 0, 2, 0, 0, 1};
  __shared__ double lds[2][CPB * NQ];
  const Place<N1> p(g);
  double *B0 = lds[0] + (p.in_block ? p.lc : 0) * NQ, *B1 = lds[1] + (p.in_block ? p.lc : 0) * NQ;
  double d[N1], u[N1];
#pragma unroll
  for (int i = 0; i < N1; ++i) d[i] = 0.0;
#pragma unroll
  for (int term = 0; term < 7; ++term) {
    if (p.in_block) {  // z: thread (qy = a, qx = b)
#pragma unroll
      for (int qz = 0; qz < N1; ++qz) {
        double c = 0.0;
        if (p.active) {
          const double *cc = td + p.cell * kTangentCoeffs * NQ + qz * N1 * N1 + p.r;
          if (term == 0) {
            c = cc[0];
          } else if (term <= 3) {
            if (g.conv) c += cc[(term + 1) * NQ];
            if (g.solid) c += cc[(term + 4) * NQ];
          } else {
            c = cc[NQ] * g.ih2[term - 4];
          }
        }
        u[qz] = c;
      }
#pragma unroll
      for (int l = 0; l < N1; ++l) {
        double s = 0.0;
#pragma unroll
        for (int qz = 0; qz < N1; ++qz) s += factor(t, fz[term], qz, l) * u[qz];
        B0[(l * N1 + p.a) * N1 + p.b] = s;
      }
    }
    __syncthreads();
    if (p.in_block) {  // y: thread (l = a, qx = b)
#pragma unroll
      for (int qy = 0; qy < N1; ++qy) u[qy] = B0[(p.a * N1 + qy) * N1 + p.b];
#pragma unroll
      for (int j = 0; j < N1; ++j) {
        double s = 0.0;
#pragma unroll
        for (int qy = 0; qy < N1; ++qy) s += factor(t, fy[term], qy, j) * u[qy];
        B1[(p.a * N1 + j) * N1 + p.b] = s;
      }
    }
    __syncthreads();
    if (p.in_block) {  // x: thread (l = a, j = b)
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) u[qx] = B1[p.r * N1 + qx];
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        double s = 0.0;
#pragma unroll
        for (int qx = 0; qx < N1; ++qx) s += factor(t, fx[term], qx, i) * u[qx];
        d[i] += s;
      }
    }
  }
  scatter<N1, true>(g, p, diag, d, lds[0]);
}

__global__ __launch_bounds__(256) void unit_constrained_kernel(double *v, Geometry g, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int ix = i % g.nd[0], iy = (i / g.nd[0]) % g.nd[1], iz = i / ((long)g.nd[0] * g.nd[1]);
  if (is_constrained(g.dmask, g.nd, ix, iy, iz)) v[i] = 1.0;
}

Geometry geometry(const OperatorData<double> &op, double dt, int conv = 0, int solid = 0) {
  Geometry g{};
  g.conv = conv ? 1 : 0;
  g.solid = solid ? 1 : 0;
  g.nc0 = op.nc[0];
  g.nc1 = op.nc[1];
  g.nc2 = op.nc[2];
  g.color = -1;
  g.ncells = op.n_cells;
  for (int d = 0; d < 3; ++d) {
    g.nd[d] = op.nd[d];
    g.h[d] = op.h[d];
    g.ih2[d] = 1.0 / (op.h[d] * op.h[d]);
  }
  g.dmask = op.dmask;
  g.vol = op.h[0] * op.h[1] * op.h[2];
  g.dt = dt;
  return g;
}

int finish(hipStream_t) { return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE; }

// one launch over all blocks (GEO.color = -1), or eight, one per colour (COLORS = 8)
#define MPDG_TANGENT_DISPATCH(KERNEL, GEO, COLORS, ...)                                                           \
  if (op.degree < 1 || op.degree > 4) return MPDG_ERR_INVALID;                                                   \
  for (int color_ = (COLORS) > 1 ? 0 : -1; color_ < ((COLORS) > 1 ? (COLORS) : 0); ++color_) {                  \
    GEO.color = color_;                                                                                          \
    const long nb_ = tangent_blocks(op.degree + 1, op.nc, color_);                                               \
    if (nb_ == 0) continue;                                                                                      \
    switch (op.degree) {                                                                                         \
      case 1: hipLaunchKernelGGL(KERNEL<2>, dim3(nb_), dim3(256), 0, s, __VA_ARGS__); break;                     \
      case 2: hipLaunchKernelGGL(KERNEL<3>, dim3(nb_), dim3(256), 0, s, __VA_ARGS__); break;                     \
      case 3: hipLaunchKernelGGL(KERNEL<4>, dim3(nb_), dim3(256), 0, s, __VA_ARGS__); break;                     \
      default: hipLaunchKernelGGL(KERNEL<5>, dim3(nb_), dim3(256), 0, s, __VA_ARGS__); break;                    \
    }                                                                                                            \
  }

}  // namespace

int launch_tangent_data(const OperatorData<double> &op, double dt, const mpdg_heat_material &mat, int solid,
                        const double *vel, const double *vfield, const double *T, const double *T_old, double *td,
                        hipStream_t s) {
  TangentArgs A{};
  A.g = geometry(op, dt, tangent_convects(vel, vfield), solid);
  A.mat = mat;
  for (int d = 0; d < 3; ++d) A.vel[d] = vel[d];
  A.vf = vfield;
  A.solid = solid;
  const Tables1D t = make_tables(op.degree);
  MPDG_TANGENT_DISPATCH(tangent_data_kernel, A.g, 1, t, A, T, T_old, td)
  return finish(s);
}

int launch_tangent_apply(const OperatorData<double> &op, const double *td, int conv, int solid,
                         const double *src, double *dst, hipStream_t s) {
  Geometry g = geometry(op, 1.0, conv, solid);
  const Tables1D t = make_tables(op.degree);
  if (hipMemsetAsync(dst, 0, sizeof(double) * op.n_dofs, s) != hipSuccess) return MPDG_ERR_DEVICE;
  MPDG_TANGENT_DISPATCH(tangent_apply_kernel, g, 8, t, g, td, src, dst)
  return finish(s);
}

int launch_tangent_residual(const OperatorData<double> &op, double dt, const mpdg_heat_material &mat,
                            const double *vel, const double *vfield, const double *T, const double *T_old,
                            const double *q, double *dst, hipStream_t s) {
  TangentArgs A{};
  A.g = geometry(op, dt);
  A.mat = mat;
  for (int d = 0; d < 3; ++d) A.vel[d] = vel[d];
  A.vf = vfield;
  const Tables1D t = make_tables(op.degree);
  MPDG_TANGENT_DISPATCH(tangent_residual_kernel, A.g, 8, t, A, T, T_old, q, dst)
  return finish(s);
}

int launch_tangent_diagonal(const OperatorData<double> &op, const double *td, int conv, int solid,
                            double *diag, hipStream_t s) {
  Geometry g = geometry(op, 1.0, conv, solid);
  const Tables1D t = make_tables(op.degree);
  if (hipMemsetAsync(diag, 0, sizeof(double) * op.n_dofs, s) != hipSuccess) return MPDG_ERR_DEVICE;
  MPDG_TANGENT_DISPATCH(tangent_diagonal_kernel, g, 8, t, g, td, diag)
  if (op.dmask)
    hipLaunchKernelGGL(unit_constrained_kernel, dim3((op.n_dofs + 255) / 256), dim3(256), 0, s, diag, g, op.n_dofs);
  return finish(s);
}

}  // namespace mpdg
