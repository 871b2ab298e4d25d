// Vector kernels of the Krylov solvers and the multigrid smoother.
//
// Reductions are deterministic: every block writes one partial per quantity,
// the last block to finish (atomic ticket) sums the partials in block order and
// updates the solver scalars on the device, so the host never has to read a
// dot product to launch the next kernel. Kernels of an iteration return early
// once the device-side "done" flag is set, which lets the host enqueue several
// iterations before it looks at the convergence state.
#include "blas.hpp"

namespace mpdg {

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T *x, T v, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) x[i] = v;
}
template <typename T>
int launch_fill(T *x, T v, long n, hipStream_t s) {
  const long nb = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_kernel<T>, dim3(std::max<long>(nb, 1)), dim3(256), 0, s, x, v, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_fill<double>(double *, double, long, hipStream_t);
template int launch_fill<float>(float *, float, long, hipStream_t);

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void convert_kernel(const TA *a, TB *b, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) b[i] = (TB)a[i];
}
template <typename TA, typename TB>
int launch_convert(const TA *a, TB *b, long n, hipStream_t s) {
  const long nb = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL((convert_kernel<TA, TB>), dim3(std::max<long>(nb, 1)), dim3(256), 0, s, a, b, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_convert<double, float>(const double *, float *, long, hipStream_t);
template int launch_convert<float, double>(const float *, double *, long, hipStream_t);
template int launch_convert<double, double>(const double *, double *, long, hipStream_t);

__global__ __launch_bounds__(256) void axpby_kernel(double *y, double a, const double *x, double b, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) y[i] = a * x[i] + b * y[i];
}
int launch_axpby(double *y, double a, const double *x, double b, long n, hipStream_t s) {
  const long nb = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(axpby_kernel, dim3(std::max<long>(nb, 1)), dim3(256), 0, s, y, a, x, b, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

constexpr int kMdotBlocks = 512;
long mdot_scratch() { return (long)kMdotBlocks * kMdotMax; }

// partial sums of the J dot products against w: block (i, b) sums row i of V times w over the
// b-th share of the index range; blockIdx.x = i runs fastest, so the rows' blocks over one
// range of w are dispatched together and share its reads in the caches. Fixed summation order
// (grid-stride, two interleaved accumulators, then a wavefront and block tree).
__global__ __launch_bounds__(256) void mdot_partial_kernel(const double *__restrict__ V, long n,
                                                           const double *__restrict__ w, double *__restrict__ part) {
  const int i = blockIdx.x;
  const double *__restrict__ v = V + (long)i * n;
  const long stride = (long)gridDim.y * 256;
  long k = (long)blockIdx.y * 256 + threadIdx.x;
  double a0 = 0.0, a1 = 0.0;
  for (; k + stride < n; k += 2 * stride) {
    a0 = fma(v[k], w[k], a0);
    a1 = fma(v[k + stride], w[k + stride], a1);
  }
  if (k < n) a0 = fma(v[k], w[k], a0);
  double t = a0 + a1;
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) part[(long)i * kMdotBlocks + blockIdx.y] = (red[0] + red[1]) + (red[2] + red[3]);
}
// h[i] (+)= sum over the nb partials of row i: one block per row
__global__ __launch_bounds__(256) void mdot_final_kernel(const double *__restrict__ part, int nb,
                                                         double *__restrict__ h, int accumulate) {
  const int i = blockIdx.x;
  double t = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) t += part[(long)i * kMdotBlocks + b];
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double sum = (red[0] + red[1]) + (red[2] + red[3]);
    h[i] = accumulate ? h[i] + sum : sum;
  }
}
int launch_mdot(const double *V, long n, int J, const double *w, double *part, double *h, bool accumulate,
                hipStream_t s) {
  if (J < 1 || J > kMdotMax) return MPDG_ERR_INVALID;
  const int nb = (int)std::max<long>(1, std::min<long>(kMdotBlocks, (n + 255) / 256));
  hipLaunchKernelGGL(mdot_partial_kernel, dim3(J, nb), dim3(256), 0, s, V, n, w, part);
  hipLaunchKernelGGL(mdot_final_kernel, dim3(J), dim3(256), 0, s, part, nb, h, accumulate ? 1 : 0);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

__global__ __launch_bounds__(256) void maxpy_kernel(double *__restrict__ w, const double *__restrict__ V, long n,
                                                    int J, const double *__restrict__ c, double sign) {
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long)gridDim.x * 256) {
    double s = 0.0;
    for (int i = 0; i < J; ++i) s = fma(c[i], V[i * n + k], s);
    w[k] += sign * s;
  }
}
int launch_maxpy(double *w, const double *V, long n, int J, const double *c, double sign, hipStream_t s) {
  if (J < 1) return MPDG_OK;
  const long nb = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(maxpy_kernel, dim3(std::max<long>(nb, 1)), dim3(256), 0, s, w, V, n, J, c, sign);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

__global__ __launch_bounds__(256) void scale_kernel(double *__restrict__ y, const double *__restrict__ x, double a,
                                                    long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) y[i] = a * x[i];
}
int launch_scale(double *y, const double *x, double a, long n, hipStream_t s) {
  const long nb = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(scale_kernel, dim3(std::max<long>(nb, 1)), dim3(256), 0, s, y, x, a, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

struct CombineArgs {
  const double *x[kCombineMax];
  double w[kCombineMax];
};
template <int M>
__global__ __launch_bounds__(256) void combine_kernel(double *__restrict__ y, CombineArgs a, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j) s = fma(a.w[j], a.x[j][i], s);
    y[i] = s;
  }
}
int launch_combine(double *y, const double *const *x, const double *w, int m, long n, hipStream_t s) {
  if (m < 1 || m > kCombineMax) return MPDG_ERR_INVALID;
  CombineArgs a{};
  for (int j = 0; j < m; ++j) {
    a.x[j] = x[j];
    a.w[j] = w[j];
  }
  const dim3 g(std::max<long>(1, std::min<long>((n + 255) / 256, 8192)));
  switch (m) {
    case 1: hipLaunchKernelGGL(combine_kernel<1>, g, dim3(256), 0, s, y, a, n); break;
    case 2: hipLaunchKernelGGL(combine_kernel<2>, g, dim3(256), 0, s, y, a, n); break;
    case 3: hipLaunchKernelGGL(combine_kernel<3>, g, dim3(256), 0, s, y, a, n); break;
    case 4: hipLaunchKernelGGL(combine_kernel<4>, g, dim3(256), 0, s, y, a, n); break;
    case 5: hipLaunchKernelGGL(combine_kernel<5>, g, dim3(256), 0, s, y, a, n); break;
    case 6: hipLaunchKernelGGL(combine_kernel<6>, g, dim3(256), 0, s, y, a, n); break;
    case 7: hipLaunchKernelGGL(combine_kernel<7>, g, dim3(256), 0, s, y, a, n); break;
    default: hipLaunchKernelGGL(combine_kernel<8>, g, dim3(256), 0, s, y, a, n); break;
  }
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

template <typename T>
__global__ __launch_bounds__(256) void invert_diag_kernel(T *d, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const T v = d[i];
    d[i] = fabs((double)v) > 1.0e-10 ? T(1) / v : T(1);  // heat_diffuse_operator.cpp:282-283
  }
}
template <typename T>
int launch_invert_diagonal(T *d, long n, hipStream_t s) {
  const long nb = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(invert_diag_kernel<T>, dim3(std::max<long>(nb, 1)), dim3(256), 0, s, d, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_invert_diagonal<double>(double *, long, hipStream_t);
template int launch_invert_diagonal<float>(float *, long, hipStream_t);

// ---- block reduction helpers ----
template <int NQ>
__device__ __forceinline__ void block_reduce(double (&v)[NQ], double *sh) {
  // sh: 256*NQ doubles
#pragma unroll
  for (int q = 0; q < NQ; ++q) sh[q * 256 + threadIdx.x] = v[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int q = 0; q < NQ; ++q) sh[q * 256 + threadIdx.x] += sh[q * 256 + threadIdx.x + s];
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = sh[q * 256];
}

// Write the block partials; returns true in the last block, whose v then holds the totals.
template <int NQ>
__device__ __forceinline__ bool reduce_to_last_block(double (&v)[NQ], double *partials, unsigned *counter) {
  __shared__ double sh[256 * NQ];
  __shared__ bool is_last;
  block_reduce<NQ>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) partials[q * kRedBlocks + blockIdx.x] = v[q];
    __threadfence();
    const unsigned t = atomicAdd(counter, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  // thread t adds the partials t, t + 256, ... in that order; all its loads are issued first
  constexpr int PER = kRedBlocks / 256;
  double part[NQ][PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int b = threadIdx.x + 256 * r;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      part[q][r] = b < (int)gridDim.x ? __builtin_nontemporal_load(partials + q * kRedBlocks + b) : 0.0;
  }
  double w[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    w[q] = 0.0;
#pragma unroll
    for (int r = 0; r < PER; ++r) w[q] += part[q][r];
  }
  // fixed-order tree over the 256 thread sums (deterministic for a fixed grid)
  block_reduce<NQ>(w, sh);
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = w[q];
  if (threadIdx.x == 0) *counter = 0;
  return true;
}

// workgroups of a reduction: one per 8192 elements, at most kRedBlocks. Every workgroup ends with a
// device-scope release (L2 write-back) and an atomic on one counter, so fewer, longer workgroups
// pay where the vector is small: a Q2 64^3 laser step (2.1 M DoFs, 256 instead of 1024
// workgroups) runs 2.8 % faster, one workgroup per 16384 elements 2-4 % slower; above 8.4 M
// elements the count stays at kRedBlocks (capping it at 512 or 256 slows the Q3 128^3 step by
// 5-20 %)
inline int red_blocks(long n) { return (int)std::max<long>(1, std::min<long>(kRedBlocks, (n + 8191) / 8192)); }

// ---- generic dot product into a device scalar ----
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void dot_kernel(const TA *a, const TB *b, long n, double *partials,
                                                  unsigned *counter, double *out) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    v[0] += (double)a[i] * (double)b[i];
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) *out = v[0];
}
template <typename TA, typename TB>
int launch_dot(const TA *a, const TB *b, long n, const Reduction &r, double *out, hipStream_t s) {
  hipLaunchKernelGGL((dot_kernel<TA, TB>), dim3(red_blocks(n)), dim3(256), 0, s, a, b, n, r.partials, r.counter,
                     out);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_dot<double, double>(const double *, const double *, long, const Reduction &, double *, hipStream_t);

// z = b - z, |z|^2 into *rr (the dot's grid and order), then z *= scale (if given): a residual,
// its norm and its diagonal scaling in one pass
__global__ __launch_bounds__(256) void residual_norm_kernel(double *z, const double *b, const double *scale, long n,
                                                            double *partials, unsigned *counter, double *rr,
                                                            double *raw) {
  // four grid strides per trip, their loads issued together; each thread adds in index order
  double v[1] = {0.0};
  const long st = (long)gridDim.x * 256;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * st < n; i += 4 * st) {
    double r[4], c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      r[u] = b[i + u * st] - z[i + u * st];
      c[u] = scale ? scale[i + u * st] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      v[0] += r[u] * r[u];
      if (raw) raw[i + u * st] = r[u];
      z[i + u * st] = scale ? r[u] * c[u] : r[u];
    }
  }
  for (; i < n; i += st) {
    const double r = b[i] - z[i];
    v[0] += r * r;
    if (raw) raw[i] = r;
    z[i] = scale ? r * scale[i] : r;
  }
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) *rr = v[0];
}
int launch_residual_norm(double *z, const double *b, const double *scale, long n, const Reduction &r, double *rr,
                         hipStream_t s, double *raw) {
  hipLaunchKernelGGL(residual_norm_kernel, dim3(red_blocks(n)), dim3(256), 0, s, z, b, scale, n, r.partials, r.counter,
                     rr, raw);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// Guess check on the device: the guess (in x, its residual in r) is kept when its squared
// residual norm sc[0] is finite and below the base guess's sc[1]; otherwise x and r are
// replaced by the base guess and its residual. The host reads sc later, without waiting here.
__global__ __launch_bounds__(256) void guess_select_kernel(const double *sc, double *x, const double *xb, double *r,
                                                           const double *rb, long n) {
  const double g = sc[0], base = sc[1];
  if (g == g && g < base) return;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    x[i] = xb[i];
    r[i] = rb[i];
  }
}
int launch_guess_select(const double *sc, double *x, const double *xb, double *r, const double *rb, long n,
                        hipStream_t s) {
  hipLaunchKernelGGL(guess_select_kernel, dim3(std::max<long>(1, std::min<long>((n + 255) / 256, 8192))), dim3(256), 0,
                     s, sc, x, xb, r, rb, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_dot<float, float>(const float *, const float *, long, const Reduction &, double *, hipStream_t);

// ================= preconditioned CG =================
// k1: alpha = rz / (p . q)
__global__ __launch_bounds__(256) void cg_alpha_kernel(const double *p, const double *q, long n, double *partials,
                                                       unsigned *counter, CGState *st) {
  if (st->done) return;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) v[0] += p[i] * q[i];
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) {
    st->pq = v[0];
    st->alpha = st->rz / v[0];
  }
}

// k2: x += alpha p, r -= alpha q, (Jacobi: z = inv r), reductions of r.r and r.z (or r.z_old)
template <bool JACOBI>
__global__ __launch_bounds__(256) void cg_update_kernel(double *x, double *r, const double *p, const double *q,
                                                        double *z, const double *inv, long n, double *partials,
                                                        unsigned *counter, CGState *st) {
  if (st->done) return;
  const double alpha = st->alpha;
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    v[0] += ri * ri;
    if (JACOBI) {
      const double zi = inv ? inv[i] * ri : ri;
      z[i] = zi;
      v[1] += ri * zi;
    } else {
      v[1] += ri * z[i];  // r_new . z_old for the flexible beta
    }
  }
  if (reduce_to_last_block<2>(v, partials, counter) && threadIdx.x == 0) {
    st->rr = v[0];
    st->iters += 1;
    const bool conv = v[0] <= st->tol2;
    const bool bad = !(v[0] == v[0]) || !(fabs(v[0]) <= 1.7e308);
    if (JACOBI) {
      st->beta = v[1] / st->rz;
      st->rz = v[1];
    } else {
      st->rzold_new = v[1];
    }
    if (conv || bad || st->iters >= st->max_it) st->done = bad ? 2 : 1;
  }
}

// k3 (external preconditioner): beta from r.z_new
__global__ __launch_bounds__(256) void cg_beta_kernel(const double *r, const double *z, long n, double *partials,
                                                      unsigned *counter, CGState *st) {
  if (st->done) return;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) v[0] += r[i] * z[i];
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) {
    const double num = st->flexible ? v[0] - st->rzold_new : v[0];
    st->beta = num / st->rz;
    st->rz = v[0];
  }
}

// k4: p = z + beta p
__global__ __launch_bounds__(256) void cg_direction_kernel(double *p, const double *z, long n, const CGState *st) {
  if (st->done) return;
  const double beta = st->beta;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = z[i] + beta * p[i];
}

// initial: r = b - Ax (Ax in q; or q is the residual itself, q_is_r), z = P r (Jacobi /
// identity), rr, rz, bb; p = z
template <bool JACOBI>
__global__ __launch_bounds__(256) void cg_init_kernel(const double *b, double *r, const double *q, double *z,
                                                      double *p, const double *inv, long n, double *partials,
                                                      unsigned *counter, CGState *st, double rel_tol,
                                                      double abs_tol, int criterion, float *r32, int q_is_r) {
  double v[3] = {0.0, 0.0, 0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const double ri = q_is_r ? q[i] : b[i] - q[i];
    r[i] = ri;
    if (r32) r32[i] = (float)ri;
    v[0] += ri * ri;
    v[2] += b[i] * b[i];
    if (JACOBI) {
      const double zi = inv ? inv[i] * ri : ri;
      z[i] = zi;
      p[i] = zi;
      v[1] += ri * zi;
    }
  }
  if (reduce_to_last_block<3>(v, partials, counter) && threadIdx.x == 0) {
    st->rr = v[0];
    st->rr0 = v[0];
    st->bb = v[2];
    st->rz = v[1];
    st->iters = 0;
    st->rzold_new = 0.0;
    double tol;
    if (criterion == MPDG_CRIT_RHS) tol = rel_tol * sqrt(v[2]);
    else tol = fmax(abs_tol, rel_tol * sqrt(v[0]));
    st->tol2 = tol * tol;
    const bool bad = !(v[0] == v[0]) || !(fabs(v[0]) <= 1.7e308);
    st->done = bad ? 2 : (v[0] <= st->tol2 ? 1 : 0);
  }
}

__global__ void cg_copy_kernel(double *p, const double *z, long n, const CGState *st) {
  if (st->done) return;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = z[i];
}
__global__ __launch_bounds__(256) void cg_rz0_kernel(const double *r, const double *z, long n, double *partials,
                                                     unsigned *counter, CGState *st) {
  if (st->done) return;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) v[0] += r[i] * z[i];
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) st->rz = v[0];
}

int cg_init(const double *b, double *r, const double *q, double *z, double *p, const double *inv, bool jacobi,
            long n, const Reduction &red, CGState *st, double rel_tol, double abs_tol, int criterion, hipStream_t s,
            float *r32, bool q_is_residual) {
  const int qr = q_is_residual ? 1 : 0;
  if (jacobi)
    hipLaunchKernelGGL(cg_init_kernel<true>, dim3(red_blocks(n)), dim3(256), 0, s, b, r, q, z, p, inv, n,
                       red.partials, red.counter, st, rel_tol, abs_tol, criterion, (float *)nullptr, qr);
  else
    hipLaunchKernelGGL(cg_init_kernel<false>, dim3(red_blocks(n)), dim3(256), 0, s, b, r, q, z, p, inv, n,
                       red.partials, red.counter, st, rel_tol, abs_tol, criterion, r32, qr);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// ---- the same iteration with an FP32 preconditioner working in place: r is also written in
//      FP32 (the preconditioner's input) and z is read in FP32 (its output), which spares the
//      two conversion passes per iteration and reads half the bytes of z ----
__global__ __launch_bounds__(256) void fcg32_start_kernel(const double *r, const float *z, double *p, long n,
                                                          double *partials, unsigned *counter, CGState *st) {
  if (st->done) return;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const double zi = z[i];
    p[i] = zi;
    v[0] += r[i] * zi;
  }
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) st->rz = v[0];
}
__global__ __launch_bounds__(256) void fcg32_update_kernel(double *x, double *r, const double *p, const double *q,
                                                           const float *z, float *r32, long n, double *partials,
                                                           unsigned *counter, CGState *st) {
  if (st->done) return;
  const double alpha = st->alpha;
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    r32[i] = (float)ri;
    v[0] += ri * ri;
    v[1] += ri * (double)z[i];  // r_new . z_old for the flexible beta
  }
  if (reduce_to_last_block<2>(v, partials, counter) && threadIdx.x == 0) {
    st->rr = v[0];
    st->iters += 1;
    const bool conv = v[0] <= st->tol2;
    const bool bad = !(v[0] == v[0]) || !(fabs(v[0]) <= 1.7e308);
    st->rzold_new = v[1];
    if (conv || bad || st->iters >= st->max_it) st->done = bad ? 2 : 1;
  }
}
__global__ __launch_bounds__(256) void fcg32_beta_kernel(const double *r, const float *z, long n, double *partials,
                                                         unsigned *counter, CGState *st) {
  if (st->done) return;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) v[0] += r[i] * (double)z[i];
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) {
    const double num = st->flexible ? v[0] - st->rzold_new : v[0];
    st->beta = num / st->rz;
    st->rz = v[0];
  }
}
__global__ void fcg32_direction_kernel(double *p, const float *z, long n, const CGState *st) {
  if (st->done) return;
  const double beta = st->beta;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = (double)z[i] + beta * p[i];
}
int fcg32_start(const double *r, const float *z, double *p, long n, const Reduction &red, CGState *st,
                hipStream_t s) {
  hipLaunchKernelGGL(fcg32_start_kernel, dim3(red_blocks(n)), dim3(256), 0, s, r, z, p, n, red.partials, red.counter,
                     st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int fcg32_update(double *x, double *r, const double *p, const double *q, const float *z, float *r32, long n,
                 const Reduction &red, CGState *st, hipStream_t s) {
  hipLaunchKernelGGL(fcg32_update_kernel, dim3(red_blocks(n)), dim3(256), 0, s, x, r, p, q, z, r32, n, red.partials,
                     red.counter, st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int fcg32_beta(const double *r, const float *z, long n, const Reduction &red, CGState *st, hipStream_t s) {
  hipLaunchKernelGGL(fcg32_beta_kernel, dim3(red_blocks(n)), dim3(256), 0, s, r, z, n, red.partials, red.counter, st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int fcg32_direction(double *p, const float *z, long n, const CGState *st, hipStream_t s) {
  hipLaunchKernelGGL(fcg32_direction_kernel, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, p, z, n,
                     st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int cg_after_external_precond_init(const double *r, const double *z, double *p, long n, const Reduction &red,
                                   CGState *st, hipStream_t s) {
  hipLaunchKernelGGL(cg_rz0_kernel, dim3(red_blocks(n)), dim3(256), 0, s, r, z, n, red.partials, red.counter, st);
  hipLaunchKernelGGL(cg_copy_kernel, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, p, z, n, st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int cg_alpha(const double *p, const double *q, long n, const Reduction &red, CGState *st, hipStream_t s) {
  hipLaunchKernelGGL(cg_alpha_kernel, dim3(red_blocks(n)), dim3(256), 0, s, p, q, n, red.partials, red.counter, st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int cg_update(double *x, double *r, const double *p, const double *q, double *z, const double *inv, bool jacobi,
              long n, const Reduction &red, CGState *st, hipStream_t s) {
  if (jacobi)
    hipLaunchKernelGGL(cg_update_kernel<true>, dim3(red_blocks(n)), dim3(256), 0, s, x, r, p, q, z, inv, n,
                       red.partials, red.counter, st);
  else
    hipLaunchKernelGGL(cg_update_kernel<false>, dim3(red_blocks(n)), dim3(256), 0, s, x, r, p, q, z, inv, n,
                       red.partials, red.counter, st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int cg_beta(const double *r, const double *z, long n, const Reduction &red, CGState *st, hipStream_t s) {
  hipLaunchKernelGGL(cg_beta_kernel, dim3(red_blocks(n)), dim3(256), 0, s, r, z, n, red.partials, red.counter, st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int cg_direction(double *p, const double *z, long n, const CGState *st, hipStream_t s) {
  hipLaunchKernelGGL(cg_direction_kernel, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, p, z, n, st);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// ================= multigrid smoother / transfer =================
// Chebyshev step: r = b - Ax; d = c1 d + c2 inv r; x += d (from x = 0: d = c2 inv b, x = d).
// d is read only by the steps after the first of a smoothing call (c1 = 0 in the first) and
// written only when a later step of the same call reads it: a degree-p call (p >= 2) from a
// nonzero x moves 6 + 7 (p - 2) + 6 vectors instead of 7 p, and a degree-1 call from x = 0
// moves 3 instead of 4.
template <typename T>
__global__ __launch_bounds__(256) void cheb_kernel(T *x, T *d, const T *b, const T *Ax, const T *inv, long n,
                                                   T c1, T c2, int zero_start, int read_d, int write_d) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const T r = zero_start ? b[i] : b[i] - Ax[i];
    const T di = read_d ? c1 * d[i] + c2 * inv[i] * r : c2 * inv[i] * r;
    if (write_d) d[i] = di;
    x[i] = zero_start ? di : x[i] + di;
  }
}
template <typename T>
int launch_chebyshev_step(T *x, T *d, const T *b, const T *Ax, const T *inv, long n, double c1, double c2,
                          bool zero_start, bool read_d, bool write_d, hipStream_t s) {
  hipLaunchKernelGGL(cheb_kernel<T>, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, x, d, b, Ax, inv,
                     n, (T)c1, (T)c2, zero_start ? 1 : 0, read_d ? 1 : 0, write_d ? 1 : 0);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_chebyshev_step<float>(float *, float *, const float *, const float *, const float *, long,
                                          double, double, bool, bool, bool, hipStream_t);
template int launch_chebyshev_step<double>(double *, double *, const double *, const double *, const double *, long,
                                           double, double, bool, bool, bool, hipStream_t);

// r = b - Ax
template <typename T>
__global__ __launch_bounds__(256) void residual_kernel(T *r, const T *b, const T *Ax, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) r[i] = b[i] - Ax[i];
}
template <typename T>
int launch_residual(T *r, const T *b, const T *Ax, long n, hipStream_t s) {
  hipLaunchKernelGGL(residual_kernel<T>, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, r, b, Ax, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_residual<float>(float *, const float *, const float *, long, hipStream_t);
template int launch_residual<double>(double *, const double *, const double *, long, hipStream_t);

// x += y
template <typename T>
__global__ __launch_bounds__(256) void add_kernel(T *x, const T *y, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) x[i] += y[i];
}
template <typename T>
int launch_add(T *x, const T *y, long n, hipStream_t s) {
  hipLaunchKernelGGL(add_kernel<T>, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, x, y, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_add<float>(float *, const float *, long, hipStream_t);
template int launch_add<double>(double *, const double *, long, hipStream_t);

// x *= y (pointwise)
__global__ __launch_bounds__(256) void mul_kernel(double *x, const double *y, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) x[i] *= y[i];
}
int launch_mul(double *x, const double *y, long n, hipStream_t s) {
  hipLaunchKernelGGL(mul_kernel, dim3(std::max<long>(1, std::min<long>((n + 255) / 256, 8192))), dim3(256), 0, s, x,
                     y, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// 1D ELL matrix applied along direction DIR of a lexicographic 3D array:
// out(i with i_DIR = row) (+)= sum_e val[row][e] (in - in2)(i with i_DIR = col[row][e]).
// in2 (may be null) fuses a residual b - Ax into the first pass of a restriction; add fuses
// x += P e into the last pass of a prolongation.
// DIR 0: one wavefront per (y, z) line; the input line is staged in LDS and the wavefront
//        forms the outputs of the line from it, one output row per lane. The table is stored
//        entry-major (col[e * o0 + row]), so that the lanes read consecutive addresses.
// DIR 1, 2: each thread walks a few output rows along DIR; the ELL rows are uniform over the
//        workgroup (scalar loads). The input rows those outputs use (a contiguous range that
//        fits the LDS stage for the multigrid transfers) are read once into a per-thread LDS
//        column and the sums are formed from there, so every input element is fetched about
//        once per workgroup instead of once per ELL entry that uses it.
constexpr int kEllLine = 1024;  // longest x line staged in LDS (DoF grids up to 1024 nodes in x)
template <typename T> struct EllStage { static constexpr int rows = 32; };  // 32 KB of LDS either way
template <> struct EllStage<double> { static constexpr int rows = 16; };
template <typename T, int MAXW>
__global__ __launch_bounds__(256) void ell_dir0_kernel(EllFields<T> F, int o0, int i0, long nlines, int width,
                                                       const int *__restrict__ col, const T *__restrict__ val,
                                                       int add) {
  __shared__ T lines[4][kEllLine];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long line_id = (long)blockIdx.x * 4 + wave;
  if (line_id >= nlines) return;  // whole wavefront; only wavefront-local LDS use below
  T *line = lines[wave];
  const long lin = (long)i0 * line_id, lout = (long)o0 * line_id;
  for (int f = 0; f < F.nf; ++f) {
    const T *in = F.in[f], *in2 = F.in2[f];
    T *out = F.out[f];
    if (f > 0) {  // every lane is done with the previous field's line
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    for (int x = lane; x < i0; x += 64) line[x] = in2 ? in[lin + x] - in2[lin + x] : in[lin + x];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int x = lane; x < o0; x += 64) {
      T acc = 0;
#pragma unroll
      for (int e = 0; e < MAXW; ++e) {
        if (e < width) {
          const int c = col[e * o0 + x];
          if (c >= 0) acc += val[e * o0 + x] * line[c];
        }
      }
      out[lout + x] = add ? out[lout + x] + acc : acc;
    }
  }
}
// DIR 1: thread index = x, blockIdx.y = z; DIR 2: thread index = position in the x-y plane
// (consecutive threads read consecutive addresses of a plane). Each thread walks `walk`
// output rows along DIR.
template <typename T, int DIR, int MAXW>
__global__ __launch_bounds__(256) void ell_dir12_kernel(EllFields<T> F, int o0, int o1, int o2, int i0, int i1,
                                                        int width, const int *__restrict__ col,
                                                        const T *__restrict__ val, int add, int walk) {
  constexpr int SR = EllStage<T>::rows;
  __shared__ T stage[SR][256];
  const int tx = threadIdx.x;
  const long t = (long)blockIdx.x * 256 + tx;
  const int r0 = blockIdx.z * walk;
  const int nrows = DIR == 1 ? o1 : o2;
  const int r1 = min(nrows, r0 + walk);
  // input rows used by the output rows r0 .. r1-1 (uniform over the workgroup)
  int cmin = 0x7fffffff, cmax = -1;
  for (int row = r0; row < r1; ++row)
    for (int e = 0; e < width; ++e) {
      const int c = col[row * width + e];
      if (c >= 0) {
        cmin = min(cmin, c);
        cmax = max(cmax, c);
      }
    }
  long base, stride, obase, ostride;
  if (DIR == 1) {
    if (t >= o0) return;  // no workgroup barrier below: each thread uses only its own LDS column
    const long z = blockIdx.y;
    base = t + (long)i0 * i1 * z;
    stride = i0;
    obase = t + (long)o0 * o1 * z;
    ostride = o0;
  } else {
    const long plane = (long)o0 * o1;
    if (t >= plane) return;
    base = t;
    stride = (long)i0 * i1;
    obase = t;
    ostride = plane;
  }
  const bool staged = cmax >= 0 && cmax - cmin < SR;
  // each thread stages and reads only its own LDS column, so the fields can reuse it in turn
  for (int f = 0; f < F.nf; ++f) {
    const T *in = F.in[f], *in2 = F.in2[f];
    T *out = F.out[f];
    if (staged) {
      for (int j = 0; j <= cmax - cmin; ++j) {
        const long src = base + stride * (cmin + j);
        stage[j][tx] = in2 ? in[src] - in2[src] : in[src];
      }
    }
    for (int row = r0; row < r1; ++row) {
      // the row's ELL entries first (uniform), then all reads, then the sum
      int cs[MAXW];
      T vs[MAXW], xs[MAXW];
#pragma unroll
      for (int e = 0; e < MAXW; ++e) {
        cs[e] = e < width ? col[row * width + e] : -1;
        vs[e] = e < width ? val[row * width + e] : T(0);
      }
#pragma unroll
      for (int e = 0; e < MAXW; ++e) {
        const int c = cs[e];
        T v = 0;
        if (c >= 0) {
          if (staged) {
            v = stage[c - cmin][tx];
          } else {
            const long src = base + stride * c;
            v = in2 ? in[src] - in2[src] : in[src];
          }
        }
        xs[e] = v;
      }
      T acc = 0;
#pragma unroll
      for (int e = 0; e < MAXW; ++e) acc += vs[e] * xs[e];
      const long g = obase + ostride * row;
      out[g] = add ? out[g] + acc : acc;
    }
  }
}
template <typename T>
int launch_ell_dir_fields(const EllFields<T> &F, const int *odims, const int *idims, int dir, int width,
                          const int *col, const T *val, hipStream_t s, bool add) {
  if (width > 32 || F.nf < 1 || F.nf > kEllMaxFields) return MPDG_ERR_INVALID;
  if (dir == 0) {
    if (idims[0] > kEllLine) return MPDG_ERR_INVALID;
    const long nlines = (long)odims[1] * odims[2];
    const dim3 grid((unsigned)((nlines + 3) / 4));
#define MPDG_ELL0(W)                                                                                             \
  hipLaunchKernelGGL((ell_dir0_kernel<T, W>), grid, dim3(256), 0, s, F, odims[0], idims[0], nlines, width, col, \
                     val, add ? 1 : 0)
    if (width <= 4) MPDG_ELL0(4);
    else if (width <= 8) MPDG_ELL0(8);
    else if (width <= 16) MPDG_ELL0(16);
    else MPDG_ELL0(32);
#undef MPDG_ELL0
    return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
  }
  // output rows per thread: as many as keep the input rows they use within the LDS stage
  // (a row of a 2:1 restriction spans about 2 walk + width input rows, of a prolongation
  // about walk / 2 + width)
  int walk = 8;
  while (walk > 1 && 2 * walk + width > EllStage<T>::rows) walk /= 2;
  dim3 grid;
  if (dir == 1) grid = dim3((odims[0] + 255) / 256, odims[2], (odims[1] + walk - 1) / walk);
  else grid = dim3((unsigned)(((long)odims[0] * odims[1] + 255) / 256), 1, (odims[2] + walk - 1) / walk);
#define MPDG_ELL12(D, W)                                                                                        \
  hipLaunchKernelGGL((ell_dir12_kernel<T, D, W>), grid, dim3(256), 0, s, F, odims[0], odims[1], odims[2],       \
                     idims[0], idims[1], width, col, val, add ? 1 : 0, walk)
  if (dir == 1) {
    if (width <= 4) MPDG_ELL12(1, 4);
    else if (width <= 8) MPDG_ELL12(1, 8);
    else if (width <= 16) MPDG_ELL12(1, 16);
    else MPDG_ELL12(1, 32);
  } else {
    if (width <= 4) MPDG_ELL12(2, 4);
    else if (width <= 8) MPDG_ELL12(2, 8);
    else if (width <= 16) MPDG_ELL12(2, 16);
    else MPDG_ELL12(2, 32);
  }
#undef MPDG_ELL12
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
// ---------------------------------------------------------------------------
// Structured 2:1 h-transfers of Q_K along one direction (the multigrid's h levels).
// P of one coarse cell, tab[m][j] = coarse basis j at fine node m of the cell (m = 0..2K);
// every coarse cell has the same table, so the kernels need no index arrays: prolongation
// reads the K+1 coarse values of a cell once and forms its 2K fine values, restriction
// (P^T) reads the 2K fine values of a cell once and forms its K coarse values, carrying the
// shared vertex's partial sum to the next cell. lo / hi: homogeneous Dirichlet end (the end
// value is dropped on input and output, as the ELL form drops its rows and columns).
// DIR 1, 2: one thread per line along DIR, kHSeg coarse cells per thread (restriction
// recomputes the carry of the cell before its segment). DIR 0: one wavefront per x line,
// the line staged in LDS.
// ---------------------------------------------------------------------------
template <typename T, int K> struct HTab { T v[2 * K + 1][K + 1]; };
constexpr int kHSeg = 4;

// element i of line (t, by): base + i * stride, base = t + by * bstride
template <typename T, int K>
__global__ __launch_bounds__(256) void hprolong_kernel(const T *__restrict__ in, T *__restrict__ out, long nl,
                                                       long ibstride, long obstride, long istride, long ostride,
                                                       int ncc, HTab<T, K> P, int lo, int hi, int add) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nl) return;
  const long ib = t + ibstride * blockIdx.y, ob = t + obstride * blockIdx.y;
  const int nfine = 2 * K * ncc + 1, ncoarse = K * ncc + 1;
  const int c0 = blockIdx.z * kHSeg, c1 = min(ncc, c0 + kHSeg);
  for (int cc = c0; cc < c1; ++cc) {
    T c[K + 1];
#pragma unroll
    for (int j = 0; j <= K; ++j) {
      const int n = K * cc + j;
      c[j] = ((lo && n == 0) || (hi && n == ncoarse - 1)) ? T(0) : in[ib + istride * n];
    }
#pragma unroll
    for (int m = 0; m < 2 * K; ++m) {
      const int g = 2 * K * cc + m;
      T a = 0;
#pragma unroll
      for (int j = 0; j <= K; ++j) a += P.v[m][j] * c[j];
      if (lo && g == 0) a = 0;
      const long o = ob + ostride * g;
      out[o] = add ? out[o] + a : a;
    }
    if (cc == ncc - 1) {
      const long o = ob + ostride * (nfine - 1);
      const T a = hi ? T(0) : c[K];
      out[o] = add ? out[o] + a : a;
    }
  }
}
// the fields of FS in turn (each with the same lines)
template <typename T, int K>
__global__ __launch_bounds__(256) void hrestrict_kernel(EllFields<T> FS, long nl, long ibstride, long obstride,
                                                        long istride, long ostride, int ncc, HTab<T, K> P, int lo,
                                                        int hi) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nl) return;
  const long ib = t + ibstride * blockIdx.y, ob = t + obstride * blockIdx.y;
  const int nfine = 2 * K * ncc + 1, ncoarse = K * ncc + 1;
  const int c0 = blockIdx.z * kHSeg, c1 = min(ncc, c0 + kHSeg);
  for (int fi = 0; fi < FS.nf; ++fi) {
    const T *__restrict__ in = FS.in[fi];
    const T *__restrict__ in2 = FS.in2[fi];
    T *__restrict__ out = FS.out[fi];
    auto F = [&](int g) -> T {
      if ((lo && g == 0) || (hi && g == nfine - 1)) return T(0);
      const long a = ib + istride * g;
      return in2 ? in[a] - in2[a] : in[a];
    };
    T carry = 0;
    if (c0 > 0) {
#pragma unroll
      for (int m = 1; m < 2 * K; ++m) carry += P.v[m][K] * F(2 * K * (c0 - 1) + m);
    }
    for (int cc = c0; cc < c1; ++cc) {
      T f[2 * K];
#pragma unroll
      for (int m = 0; m < 2 * K; ++m) f[m] = F(2 * K * cc + m);
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int n = K * cc + j;
        T a = j == 0 ? carry : T(0);
#pragma unroll
        for (int m = 0; m < 2 * K; ++m) a += P.v[m][j] * f[m];
        out[ob + ostride * n] = (lo && n == 0) ? T(0) : a;
      }
      carry = 0;
#pragma unroll
      for (int m = 1; m < 2 * K; ++m) carry += P.v[m][K] * f[m];
      if (cc == ncc - 1) out[ob + ostride * (ncoarse - 1)] = hi ? T(0) : carry + F(nfine - 1);
    }
  }
}
// x lines: one wavefront per line, the input line in LDS, the table read from global memory
template <typename T, int K, bool RESTRICT>
__global__ __launch_bounds__(256) void htransfer_x_kernel(EllFields<T> FS, long nlines, int ncc,
                                                          const T *__restrict__ tab, int lo, int hi, int add) {
  __shared__ T lines[4][kEllLine];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long line_id = (long)blockIdx.x * 4 + wave;
  if (line_id >= nlines) return;  // whole wavefront; only wavefront-local LDS use below
  const int nfine = 2 * K * ncc + 1, ncoarse = K * ncc + 1;
  const int ni = RESTRICT ? nfine : ncoarse, no = RESTRICT ? ncoarse : nfine;
  T *line = lines[wave];
  const long lin = (long)ni * line_id, lout = (long)no * line_id;
  for (int fi = 0; fi < FS.nf; ++fi) {
    const T *__restrict__ in = FS.in[fi];
    const T *__restrict__ in2 = FS.in2[fi];
    T *__restrict__ out = FS.out[fi];
    if (fi > 0) {  // every lane is done with the previous field's line
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    for (int x = lane; x < ni; x += 64) {
      const bool drop = (lo && x == 0) || (hi && x == ni - 1);
      line[x] = drop ? T(0) : (in2 ? in[lin + x] - in2[lin + x] : in[lin + x]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    constexpr int W = K + 1;
    for (int x = lane; x < no; x += 64) {
      T a = 0;
      if (!RESTRICT) {
        // fine node x: cell cc, row m (the last node is the last cell's right vertex)
        const int cc = min(x / (2 * K), ncc - 1), m = x - 2 * K * cc;
#pragma unroll
        for (int j = 0; j <= K; ++j) a += tab[m * W + j] * line[K * cc + j];
        if ((lo && x == 0) || (hi && x == no - 1)) a = 0;
        out[lout + x] = add ? out[lout + x] + a : a;
      } else {
        // coarse node x = K cc + j: this cell's rows 0..2K-1, the previous cell's for a vertex,
        // and the last fine node for the last vertex
        const int cc = min(x / K, ncc - 1), j = x - K * cc;
        if (j < K) {
#pragma unroll
          for (int m = 0; m < 2 * K; ++m) a += tab[m * W + j] * line[2 * K * cc + m];
        }
        const int cp = j == K ? cc : cc - 1;  // cell whose right vertex x is
        if ((j == 0 && cc > 0) || j == K) {
#pragma unroll
          for (int m = 1; m < 2 * K; ++m) a += tab[m * W + K] * line[2 * K * cp + m];
        }
        if (j == K) a += line[nfine - 1];
        if ((lo && x == 0) || (hi && x == no - 1)) a = 0;
        out[lout + x] = a;
      }
    }
  }
}
template <typename T, int K>
static int htransfer_k(bool restrict_mode, int dir, const EllFields<T> &FS, const int *idims, const int *odims,
                       const double *tab_host, const T *tab_dev, int lo, int hi, bool add, hipStream_t s) {
  if (FS.nf < 1 || FS.nf > kEllMaxFields || (!restrict_mode && FS.nf != 1)) return MPDG_ERR_INVALID;
  const T *in = FS.in[0];
  T *out = FS.out[0];
  HTab<T, K> P;
  for (int m = 0; m <= 2 * K; ++m)
    for (int j = 0; j <= K; ++j) P.v[m][j] = (T)tab_host[m * (K + 1) + j];
  const int ncoarse = restrict_mode ? odims[dir] : idims[dir];
  if ((ncoarse - 1) % K != 0) return MPDG_ERR_INVALID;
  const int ncc = (ncoarse - 1) / K;
  if ((restrict_mode ? idims[dir] : odims[dir]) != 2 * K * ncc + 1) return MPDG_ERR_INVALID;
  if (dir == 0) {
    if (idims[0] > kEllLine) return MPDG_ERR_INVALID;
    const long nlines = (long)odims[1] * odims[2];
    const dim3 grid((unsigned)((nlines + 3) / 4));
    if (restrict_mode)
      hipLaunchKernelGGL((htransfer_x_kernel<T, K, true>), grid, dim3(256), 0, s, FS, nlines, ncc, tab_dev, lo, hi, 0);
    else
      hipLaunchKernelGGL((htransfer_x_kernel<T, K, false>), grid, dim3(256), 0, s, FS, nlines, ncc, tab_dev, lo, hi,
                         add ? 1 : 0);
    return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
  }
  long nl, ibs, obs, is, os;
  unsigned gy;
  if (dir == 2) {  // lines over the x-y plane (unchanged by a z pass)
    nl = (long)odims[0] * odims[1];
    ibs = obs = 0;
    is = os = nl;
    gy = 1;
  } else {  // lines over x, one block row per z
    nl = odims[0];
    ibs = (long)idims[0] * idims[1];
    obs = (long)odims[0] * odims[1];
    is = idims[0];
    os = odims[0];
    gy = (unsigned)odims[2];
  }
  const dim3 grid((unsigned)((nl + 255) / 256), gy, (ncc + kHSeg - 1) / kHSeg);
  if (restrict_mode)
    hipLaunchKernelGGL((hrestrict_kernel<T, K>), grid, dim3(256), 0, s, FS, nl, ibs, obs, is, os, ncc, P, lo, hi);
  else
    hipLaunchKernelGGL((hprolong_kernel<T, K>), grid, dim3(256), 0, s, in, out, nl, ibs, obs, is, os, ncc, P, lo, hi,
                       add ? 1 : 0);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template <typename T>
int launch_htransfer_fields(bool restrict_mode, int dir, const EllFields<T> &FS, const int *idims, const int *odims,
                            int k, const double *tab_host, const T *tab_dev, int lo, int hi, bool add, hipStream_t s) {
  switch (k) {
    case 1: return htransfer_k<T, 1>(restrict_mode, dir, FS, idims, odims, tab_host, tab_dev, lo, hi, add, s);
    case 2: return htransfer_k<T, 2>(restrict_mode, dir, FS, idims, odims, tab_host, tab_dev, lo, hi, add, s);
    case 3: return htransfer_k<T, 3>(restrict_mode, dir, FS, idims, odims, tab_host, tab_dev, lo, hi, add, s);
    case 4: return htransfer_k<T, 4>(restrict_mode, dir, FS, idims, odims, tab_host, tab_dev, lo, hi, add, s);
    default: return MPDG_ERR_INVALID;
  }
}
template <typename T>
int launch_htransfer(bool restrict_mode, int dir, const T *in, const T *in2, T *out, const int *idims,
                     const int *odims, int k, const double *tab_host, const T *tab_dev, int lo, int hi, bool add,
                     hipStream_t s) {
  EllFields<T> FS{};
  FS.in[0] = in;
  FS.in2[0] = in2;
  FS.out[0] = out;
  FS.nf = 1;
  return launch_htransfer_fields<T>(restrict_mode, dir, FS, idims, odims, k, tab_host, tab_dev, lo, hi, add, s);
}
template int launch_htransfer_fields<double>(bool, int, const EllFields<double> &, const int *, const int *, int,
                                             const double *, const double *, int, int, bool, hipStream_t);
template int launch_htransfer<float>(bool, int, const float *, const float *, float *, const int *, const int *, int,
                                     const double *, const float *, int, int, bool, hipStream_t);
template int launch_htransfer<double>(bool, int, const double *, const double *, double *, const int *, const int *,
                                      int, const double *, const double *, int, int, bool, hipStream_t);

template <typename T>
int launch_ell_dir(const T *in, T *out, const int *odims, const int *idims, int dir, int width, const int *col,
                   const T *val, hipStream_t s, const T *in2, bool add) {
  EllFields<T> F{};
  F.in[0] = in;
  F.in2[0] = in2;
  F.out[0] = out;
  F.nf = 1;
  return launch_ell_dir_fields<T>(F, odims, idims, dir, width, col, val, s, add);
}
template int launch_ell_dir<float>(const float *, float *, const int *, const int *, int, int, const int *,
                                   const float *, hipStream_t, const float *, bool);
template int launch_ell_dir<double>(const double *, double *, const int *, const int *, int, int, const int *,
                                    const double *, hipStream_t, const double *, bool);
template int launch_ell_dir_fields<double>(const EllFields<double> &, const int *, const int *, int, int,
                                           const int *, const double *, hipStream_t, bool);

// Coarse-grid solve with the stored inverse (dense, column-major, FP64):
// y_i = sum_j A[j n + i] x_j in two deterministic stages: segment k of the j range
// (kGemvSeg segments) gives part[k][i], then the segments are summed in order.
constexpr int kGemvSeg = 32;
// x and y in the level precision TX / TY (read and written as doubles and rounded on the way
// out, the same values a conversion pass before and after would give)
template <typename TX>
__global__ __launch_bounds__(256) void dense_gemv_part_kernel(const double *__restrict__ A,
                                                              const TX *__restrict__ x,
                                                              double *__restrict__ part, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= n) return;
  const int len = (n + kGemvSeg - 1) / kGemvSeg;
  const int j0 = k * len, j1 = min(n, j0 + len);
  double acc = 0.0;
  for (int j = j0; j < j1; ++j) acc += A[(long)j * n + i] * (double)x[j];
  part[(long)k * n + i] = acc;
}
template <typename TY>
__global__ __launch_bounds__(256) void dense_gemv_sum_kernel(const double *__restrict__ part, TY *__restrict__ y,
                                                             int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int k = 0; k < kGemvSeg; ++k) acc += part[(long)k * n + i];
  y[i] = (TY)acc;
}
long dense_gemv_scratch(int n) { return (long)kGemvSeg * n; }
template <typename T>
int launch_dense_gemv(const double *A, const T *x, T *y, double *part, int n, hipStream_t s) {
  hipLaunchKernelGGL(dense_gemv_part_kernel<T>, dim3((n + 255) / 256, kGemvSeg), dim3(256), 0, s, A, x, part, n);
  hipLaunchKernelGGL(dense_gemv_sum_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, s, part, y, n);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_dense_gemv<double>(const double *, const double *, double *, double *, int, hipStream_t);
template int launch_dense_gemv<float>(const double *, const float *, float *, double *, int, hipStream_t);

// v = 0 on the DoFs of homogeneous Dirichlet faces
template <typename T>
__global__ __launch_bounds__(256) void zero_constrained_kernel(T *v, int nd0, int nd1, int nd2, int mask) {
  const long n = (long)nd0 * nd1 * nd2;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int ix = i % nd0, iy = (i / nd0) % nd1, iz = i / ((long)nd0 * nd1);
    const bool con = ((mask & 1) && ix == 0) || ((mask & 2) && ix == nd0 - 1) || ((mask & 4) && iy == 0) ||
                     ((mask & 8) && iy == nd1 - 1) || ((mask & 16) && iz == 0) || ((mask & 32) && iz == nd2 - 1);
    if (con) v[i] = T(0);
  }
}
template <typename T>
int launch_zero_constrained(T *v, const int *nd, int mask, hipStream_t s) {
  if (!mask) return MPDG_OK;
  const long n = (long)nd[0] * nd[1] * nd[2];
  hipLaunchKernelGGL(zero_constrained_kernel<T>, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, v,
                     nd[0], nd[1], nd[2], mask);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_zero_constrained<float>(float *, const int *, int, hipStream_t);
template int launch_zero_constrained<double>(double *, const int *, int, hipStream_t);

// deterministic start vector of the eigenvalue estimate (constrained entries 0)
template <typename T>
__global__ __launch_bounds__(256) void eig_start_kernel(T *v, int nd0, int nd1, int nd2, int mask) {
  const long n = (long)nd0 * nd1 * nd2;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const int ix = i % nd0, iy = (i / nd0) % nd1, iz = i / ((long)nd0 * nd1);
    const bool con = ((mask & 1) && ix == 0) || ((mask & 2) && ix == nd0 - 1) || ((mask & 4) && iy == 0) ||
                     ((mask & 8) && iy == nd1 - 1) || ((mask & 16) && iz == 0) || ((mask & 32) && iz == nd2 - 1);
    v[i] = con ? T(0) : (T)(1.0 + 0.5 * sin(0.9 * (double)(i % 1000003)));
  }
}
template <typename T>
int launch_eig_start(T *v, const int *nd, int mask, hipStream_t s) {
  const long n = (long)nd[0] * nd[1] * nd[2];
  hipLaunchKernelGGL(eig_start_kernel<T>, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, v, nd[0],
                     nd[1], nd[2], mask);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_eig_start<float>(float *, const int *, int, hipStream_t);
template int launch_eig_start<double>(double *, const int *, int, hipStream_t);

// CG step kernels of the eigenvalue estimate in the level precision, scalars in FP64:
// alpha_k = rz/(p.Ap) recorded in hist[2k]; x += alpha p; r -= alpha Ap; z = inv r;
// beta_k = rz_new/rz recorded in hist[2k+1]; p = z + beta p.
// In FP32 the two reductions take four grid strides per trip, their loads issued together (a
// 4-byte load per lane leaves too few bytes in flight otherwise; in FP64 the one-stride loop is
// faster); each thread still adds its elements in index order, so the sums are those of the
// one-stride loop.
template <typename T> constexpr int kEigStrides = sizeof(T) == 4 ? 4 : 1;
template <typename T>
__global__ __launch_bounds__(256) void eig_alpha_kernel(const T *p, const T *q, long n, double *partials,
                                                        unsigned *counter, double *sc, double *hist, int k) {
  constexpr int U = kEigStrides<T>;
  double v[1] = {0.0};
  const long st = (long)gridDim.x * 256;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  for (; i + (U - 1) * st < n; i += U * st) {
    T a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = p[i + u * st];
      b[u] = q[i + u * st];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) v[0] += (double)a[u] * b[u];
  }
  for (; i < n; i += st) v[0] += (double)p[i] * q[i];
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) {
    sc[1] = sc[0] / v[0];
    hist[2 * k] = sc[1];
  }
}
template <typename T>
__global__ __launch_bounds__(256) void eig_update_kernel(T *r, T *z, const T *q, const T *inv, long n,
                                                         double *partials, unsigned *counter, double *sc,
                                                         double *hist, int k) {
  constexpr int U = kEigStrides<T>;
  const T alpha = (T)sc[1];
  double v[1] = {0.0};
  const long st = (long)gridDim.x * 256;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  for (; i + (U - 1) * st < n; i += U * st) {
    T a[U], b[U], c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = r[i + u * st];
      b[u] = q[i + u * st];
      c[u] = inv[i + u * st];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T ri = a[u] - alpha * b[u];
      const T zi = c[u] * ri;
      r[i + u * st] = ri;
      z[i + u * st] = zi;
      v[0] += (double)ri * zi;
    }
  }
  for (; i < n; i += st) {
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = inv[i] * ri;
    z[i] = zi;
    v[0] += (double)ri * zi;
  }
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) {
    hist[2 * k + 1] = v[0] / sc[0];
    sc[2] = v[0] / sc[0];
    sc[0] = v[0];
  }
}
template <typename T>
__global__ __launch_bounds__(256) void eig_dir_kernel(T *p, const T *z, long n, const double *sc) {
  const T beta = (T)sc[2];
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = z[i] + beta * p[i];
}
template <typename T>
__global__ __launch_bounds__(256) void eig_init_kernel(T *r, T *z, T *p, const T *inv, long n, double *partials,
                                                       unsigned *counter, double *sc) {
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const T zi = inv[i] * r[i];
    z[i] = zi;
    p[i] = zi;
    v[0] += (double)r[i] * zi;
  }
  if (reduce_to_last_block<1>(v, partials, counter) && threadIdx.x == 0) sc[0] = v[0];
}
template <typename T>
int launch_eig_step(int phase, T *r, T *z, T *p, const T *q, const T *inv, long n, const Reduction &red, double *sc,
                    double *hist, int k, hipStream_t s) {
  const dim3 g(red_blocks(n)), b(256);
  if (phase == 0) hipLaunchKernelGGL(eig_init_kernel<T>, g, b, 0, s, r, z, p, inv, n, red.partials, red.counter, sc);
  if (phase == 1) hipLaunchKernelGGL(eig_alpha_kernel<T>, g, b, 0, s, p, q, n, red.partials, red.counter, sc, hist, k);
  if (phase == 2)
    hipLaunchKernelGGL(eig_update_kernel<T>, g, b, 0, s, r, z, q, inv, n, red.partials, red.counter, sc, hist, k);
  if (phase == 3) hipLaunchKernelGGL(eig_dir_kernel<T>, dim3(std::min<long>((n + 255) / 256, 8192)), b, 0, s, p, z, n, sc);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
template int launch_eig_step<float>(int, float *, float *, float *, const float *, const float *, long,
                                    const Reduction &, double *, double *, int, hipStream_t);
template int launch_eig_step<double>(int, double *, double *, double *, const double *, const double *, long,
                                     const Reduction &, double *, double *, int, hipStream_t);

// ---- bandwidth probes (benchmark helpers): copy, read-only and write-only streams over FP64
//      vectors, 16-byte accesses, grid-stride over n/2 pairs (n even) ----
__global__ __launch_bounds__(256) void stream_copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                          long m) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long)gridDim.x * 256) dst[i] = src[i];
}
__global__ __launch_bounds__(256) void stream_read_kernel(const double2 *__restrict__ src, double *__restrict__ part,
                                                          long m) {
  __shared__ double sh[256];
  double a = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long)gridDim.x * 256) {
    const double2 v = src[i];
    a += v.x + v.y;
  }
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(256) void stream_write_kernel(double2 *__restrict__ dst, long m, double v) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long)gridDim.x * 256) dst[i] = make_double2(v, v);
}
int launch_stream(int kind, double *dst, const double *src, long n, hipStream_t s) {
  const long m = n / 2;
  const int blocks = 256 * 16;  // 16 workgroups per CU on 256 CUs
  switch (kind) {
    case 0:
      hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const double2 *>(src),
                         reinterpret_cast<double2 *>(dst), m);
      break;
    case 1:
      hipLaunchKernelGGL(stream_read_kernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<const double2 *>(src),
                         dst, m);
      break;
    case 2:
      hipLaunchKernelGGL(stream_write_kernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<double2 *>(dst), m, 1.0);
      break;
    default: return MPDG_ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

}  // namespace mpdg
