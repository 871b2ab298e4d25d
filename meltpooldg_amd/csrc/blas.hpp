// Vector kernels, deterministic reductions and the device-side CG state.
#pragma once

#include "kernels.hpp"

namespace mpdg {

constexpr int kRedBlocks = 1024;

// scratch of the two-stage reductions
struct Reduction {
  double *partials = nullptr;  // 4 * kRedBlocks doubles
  unsigned *counter = nullptr; // zero-initialised ticket
};

// Krylov scalars kept on the device between kernels
struct CGState {
  double rz, pq, alpha, beta, rr, rr0, bb, tol2, rzold_new;
  int done;  // 0 running, 1 converged or out of iterations, 2 non-finite
  int iters;
  int max_it;
  int flexible;
};

template <typename T> int launch_fill(T *x, T v, long n, hipStream_t s);
template <typename TA, typename TB> int launch_convert(const TA *a, TB *b, long n, hipStream_t s);
int launch_axpby(double *y, double a, const double *x, double b, long n, hipStream_t s);
template <typename T> int launch_invert_diagonal(T *d, long n, hipStream_t s);
template <typename TA, typename TB>
int launch_dot(const TA *a, const TB *b, long n, const Reduction &r, double *out, hipStream_t s);

// q_is_residual: q already holds b - A x (the initial apply is skipped by the caller)
int cg_init(const double *b, double *r, const double *q, double *z, double *p, const double *inv, bool jacobi,
            long n, const Reduction &red, CGState *st, double rel_tol, double abs_tol, int criterion, hipStream_t s,
            float *r32 = nullptr, bool q_is_residual = false);
// CG / flexible CG with an FP32 preconditioner run in place: r32 its input, z its output
int fcg32_start(const double *r, const float *z, double *p, long n, const Reduction &red, CGState *st,
                hipStream_t s);
int fcg32_update(double *x, double *r, const double *p, const double *q, const float *z, float *r32, long n,
                 const Reduction &red, CGState *st, hipStream_t s);
int fcg32_beta(const double *r, const float *z, long n, const Reduction &red, CGState *st, hipStream_t s);
int fcg32_direction(double *p, const float *z, long n, const CGState *st, hipStream_t s);
int cg_after_external_precond_init(const double *r, const double *z, double *p, long n, const Reduction &red,
                                   CGState *st, hipStream_t s);
int cg_alpha(const double *p, const double *q, long n, const Reduction &red, CGState *st, hipStream_t s);
int cg_update(double *x, double *r, const double *p, const double *q, double *z, const double *inv, bool jacobi,
              long n, const Reduction &red, CGState *st, hipStream_t s);
int cg_beta(const double *r, const double *z, long n, const Reduction &red, CGState *st, hipStream_t s);
int cg_direction(double *p, const double *z, long n, const CGState *st, hipStream_t s);

// GMRES building blocks over a basis V of J rows of length n (row i at V + i n), J <= kMdotMax:
// h[i] (+)= V_i . w (deterministic two-stage reduction; part: mdot_scratch() doubles),
// w += sign * sum_i c_i V_i with c in device memory, y = a x.
constexpr int kMdotMax = 32;
long mdot_scratch();
int launch_mdot(const double *V, long n, int J, const double *w, double *part, double *h, bool accumulate,
                hipStream_t s);
int launch_maxpy(double *w, const double *V, long n, int J, const double *c, double sign, hipStream_t s);
int launch_scale(double *y, const double *x, double a, long n, hipStream_t s);
// y = sum_j w_j x_j over m <= kCombineMax vectors in one pass (y distinct from the x_j)
constexpr int kCombineMax = 8;
int launch_combine(double *y, const double *const *x, const double *w, int m, long n, hipStream_t s);

template <typename T>
int launch_chebyshev_step(T *x, T *d, const T *b, const T *Ax, const T *inv, long n, double c1, double c2,
                          bool zero_start, bool read_d, bool write_d, hipStream_t s);
template <typename T> int launch_residual(T *r, const T *b, const T *Ax, long n, hipStream_t s);
template <typename T> int launch_add(T *x, const T *y, long n, hipStream_t s);
int launch_mul(double *x, const double *y, long n, hipStream_t s);
// z = b - z; *rr = |z|^2 (device scalar); raw (if given) = z; then z *= scale pointwise when
// scale is given
int launch_residual_norm(double *z, const double *b, const double *scale, long n, const Reduction &r, double *rr,
                         hipStream_t s, double *raw = nullptr);
// x, r <- xb, rb unless sc[0] (the guess's squared residual norm) is finite and below sc[1]
int launch_guess_select(const double *sc, double *x, const double *xb, double *r, const double *rb, long n,
                        hipStream_t s);
// 1D ELL matrix (rows = odims[dir], `width` entries each, col < 0 marks an empty entry) along
// direction dir; the tables of dir 1 and 2 are row-major (entry e of row r at r * width + e),
// those of dir 0 entry-major (at e * odims[0] + r, see ell_entry_major)
template <typename T>
int launch_ell_dir(const T *in, T *out, const int *odims, const int *idims, int dir, int width, const int *col,
                   const T *val, hipStream_t s, const T *in2 = nullptr, bool add = false);
// the same 1D operator applied to up to kEllMaxFields fields of equal shape in one launch
// (out[f] = A in[f], or A (in[f] - in2[f]) where in2[f] is set)
constexpr int kEllMaxFields = 4;
template <typename T> struct EllFields {
  const T *in[kEllMaxFields];
  const T *in2[kEllMaxFields];
  T *out[kEllMaxFields];
  int nf;
};
template <typename T>
int launch_ell_dir_fields(const EllFields<T> &F, const int *odims, const int *idims, int dir, int width,
                          const int *col, const T *val, hipStream_t s, bool add = false);
// structured 2:1 h-transfer of Q_k along `dir` (see blas.hip): restrict_mode false: out (+)= P in
// (fine <- coarse, accumulating when add); true: out = P^T (in - in2) (in2 may be null).
// tab_host / tab_dev: the (2k+1) x (k+1) table of one coarse cell; lo / hi: Dirichlet ends
template <typename T>
int launch_htransfer(bool restrict_mode, int dir, const T *in, const T *in2, T *out, const int *idims,
                     const int *odims, int k, const double *tab_host, const T *tab_dev, int lo, int hi, bool add,
                     hipStream_t s);
// restriction of up to kEllMaxFields fields on the same grids in one launch (in2[f] as above)
template <typename T>
int launch_htransfer_fields(bool restrict_mode, int dir, const EllFields<T> &FS, const int *idims, const int *odims,
                            int k, const double *tab_host, const T *tab_dev, int lo, int hi, bool add, hipStream_t s);
// y = A^T x for a dense column-major n x n matrix (coarse-grid inverse), accumulated in FP64 with
// x and y in precision T; part: dense_gemv_scratch(n) doubles
long dense_gemv_scratch(int n);
template <typename T>
int launch_dense_gemv(const double *A, const T *x, T *y, double *part, int n, hipStream_t s);
template <typename T> int launch_zero_constrained(T *v, const int *nd, int mask, hipStream_t s);
template <typename T> int launch_eig_start(T *v, const int *nd, int mask, hipStream_t s);
template <typename T>
int launch_eig_step(int phase, T *r, T *z, T *p, const T *q, const T *inv, long n, const Reduction &red, double *sc,
                    double *hist, int k, hipStream_t s);

// bandwidth probe: kind 0 dst = src, 1 read src (4096 block partial sums into dst), 2 write
// dst; n even, 16-byte aligned vectors
int launch_stream(int kind, double *dst, const double *src, long n, hipStream_t s);

}  // namespace mpdg
