// Host-side objects behind the C ABI: the operator, the preconditioners
// (inverse diagonal, geometric multigrid), the Krylov solver and the implicit
// heat time step. They mirror MeltPoolDG's OperatorMatrixFree,
// Preconditioner, LinearSolver, NewtonRaphsonSolver, Predictor and
// HeatDiffuseOperation (see the citations on each class).
#pragma once

#include <algorithm>
#include <memory>
#include <utility>
#include <vector>

#include "blas.hpp"
#include "fe1d.hpp"

namespace mpdg {

// dense coarse-grid matrix (row-major, n_dofs^2) of the level `od` with cell conductivities kc
// (nullptr: kconst everywhere); constrained DoFs get identity rows and columns. E: scratch of
// n_cells (k+1)^6 element-matrix entries
template <typename T>
int launch_coarse_matrix(const OperatorData<T> &od, double c_mass, const double *kc, double kconst, double *A,
                         double *E, hipStream_t s);
int launch_gauss_jordan_inverse(double *A, int n, hipStream_t s);

#define MPDG_TRY_INNER(...)         \
  do {                              \
    const int rc_i__ = (__VA_ARGS__); \
    if (rc_i__ != MPDG_OK) return rc_i__; \
  } while (0)

// device buffer owned by the library
template <typename T>
struct DeviceArray {
  T *ptr = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  int resize(size_t count) {
    if (count == n && ptr) return MPDG_OK;
    release();
    if (count == 0) return MPDG_OK;
    if (hipMalloc((void **)&ptr, count * sizeof(T)) != hipSuccess) {
      ptr = nullptr;
      return MPDG_ERR_NOMEM;
    }
    n = count;
    return MPDG_OK;
  }
  // resize and zero; keeps the contents when the size is unchanged
  int resize_zeroed(size_t count) {
    if (count == n && ptr) return MPDG_OK;
    MPDG_TRY_INNER(resize(count));
    if (count && hipMemset(ptr, 0, count * sizeof(T)) != hipSuccess) return MPDG_ERR_DEVICE;
    return MPDG_OK;
  }
  T *get() const { return ptr; }
};

struct ReductionScratch {
  DeviceArray<double> partials;
  DeviceArray<unsigned> counter;
  int init() {
    int rc = partials.resize(4 * kRedBlocks);
    if (rc) return rc;
    rc = counter.resize(1);
    if (rc) return rc;
    return hipMemset(counter.get(), 0, sizeof(unsigned)) == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
  }
  Reduction red() const {
    Reduction r;
    r.partials = partials.get();
    r.counter = counter.get();
    return r;
  }
};

#define MPDG_TRY(...)            \
  do {                           \
    const int rc__ = (__VA_ARGS__); \
    if (rc__ != MPDG_OK) return rc__; \
  } while (0)

// 1D transfer between (kc, ncc) and (kf, ncf), ncf a multiple of ncc: prolongation P (fine
// node g <- coarse basis functions at g, ELL width kc+1) and its transpose R (rows = coarse
// nodes); Dirichlet end points drop their rows and columns.
struct Transfer1DHost {
  int nf = 0, nc = 0, pw = 0, rw = 0;
  std::vector<int> pcol, rcol;
  std::vector<double> pval, rval;
  Transfer1DHost(int kc, int ncc, int kf, int ncf, bool dir_lo, bool dir_hi) {
    FE1D fc(kc), ff(kf);
    nf = kf * ncf + 1;
    nc = kc * ncc + 1;
    pw = kc + 1;
    pcol.assign((size_t)nf * pw, -1);
    pval.assign((size_t)nf * pw, 0.0);
    std::vector<std::vector<std::pair<int, double>>> cols(nc);
    for (int g = 0; g < nf; ++g) {
      const int cf = std::min(g / kf, ncf - 1), i = g - kf * cf;
      const int cc = (int)((long)cf * ncc / ncf);
      const double x = (cf + ff.nodes[i]) * (double)ncc / (double)ncf - cc;
      double v[kMaxN1];
      FE1D::lagrange(kc + 1, fc.nodes, x, v, nullptr);
      const bool drop_row = (dir_lo && g == 0) || (dir_hi && g == nf - 1);
      for (int j = 0; j < pw; ++j) {
        const int c = kc * cc + j;
        const bool drop_col = (dir_lo && c == 0) || (dir_hi && c == nc - 1);
        if (drop_row || drop_col) continue;
        pcol[(size_t)g * pw + j] = c;
        pval[(size_t)g * pw + j] = v[j];
        cols[c].push_back({g, v[j]});
      }
    }
    rw = 1;
    for (auto &c : cols) rw = std::max<int>(rw, (int)c.size());
    rcol.assign((size_t)nc * rw, -1);
    rval.assign((size_t)nc * rw, 0.0);
    for (int c = 0; c < nc; ++c)
      for (size_t e = 0; e < cols[c].size(); ++e) {
        rcol[(size_t)c * rw + e] = cols[c][e].first;
        rval[(size_t)c * rw + e] = cols[c][e].second;
      }
  }
};

// A row-major ELL table (entry e of row r at r * width + e) in the entry-major order of the
// direction-0 transfers (at e * rows + r), whose lanes each form one output row
template <typename V>
std::vector<V> ell_entry_major(const std::vector<V> &t, int width) {
  const size_t rows = t.size() / width;
  std::vector<V> o(t.size());
  for (size_t r = 0; r < rows; ++r)
    for (int e = 0; e < width; ++e) o[(size_t)e * rows + r] = t[r * width + e];
  return o;
}

}  // namespace mpdg

// ---------------------------------------------------------------------------
// HeatDiffuseMultiPhaseOperator (source/heat/heat_diffuse_operator.cpp) as an
// OperatorMatrixFree (include/meltpooldg/core/operator_base.hpp:189-282).
// ---------------------------------------------------------------------------
struct mpdg_heat_op {
  mpdg_heat_desc desc{};
  int precision = MPDG_FP64;
  double dt = 0.0;
  mutable mpdg::OperatorData<double> d64;
  mutable mpdg::OperatorData<float> d32;
  double flux[6] = {0, 0, 0, 0, 0, 0};
  const double *T_robin = nullptr;
  mutable mpdg::DeviceArray<double> kq, tmp1, tmp2, scal, tables, metric64, metric64_k;
  mutable mpdg::DeviceArray<double> diag_scratch;  // launch_diagonal: the cells' entries, k at the points
  mpdg::DeviceArray<float> metric32;
  mutable mpdg::DeviceArray<float> metric32_k, kq32;
  const double *coords = nullptr;
  mutable mpdg::ReductionScratch red;
  // apply scratch for the rows where four columns meet (OperatorData::edge), zero between applies
  mutable mpdg::DeviceArray<double> edge64;
  mutable mpdg::DeviceArray<float> edge32;
  // full Newton tangent (mpdg_heat_op_set_tangent): material, velocity, per-qp data at the bound T
  bool tangent = false;
  mpdg_heat_material mat{};
  int solid = 0;
  double vel[3] = {0, 0, 0};
  const double *vfield = nullptr;  // velocity DoF field (mpdg_heat_op_set_velocity_field), caller-owned
  mpdg::DeviceArray<double> td;
  bool td_bound = false;
  // optional HIP-event timing of the cell-apply kernel (bench / profiling)
  mutable std::vector<std::pair<hipEvent_t, hipEvent_t>> timing_events;
  mutable size_t timing_used = 0;
  bool timing_on = false;
  ~mpdg_heat_op() {
    for (auto &e : timing_events) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  }

  // mapped meshes with per-qp / k(T) conductivity: fold the conductivity into metric_k
  int refresh_metric_k(hipStream_t s) const;
  // allocates (zeroed) and binds the apply's edge scratch on first use
  int ensure_edge() const;
  long n_dofs() const { return d64.n_dofs; }
  long n_cells() const { return d64.n_cells; }
  int vmult(double *dst, const double *src, hipStream_t s) const;
  int vmult_cells(double *dst, const double *src, hipStream_t s) const;  // without boundary faces
  int vmult_f32(float *dst, const float *src, hipStream_t s) const;
  int diagonal64(double *diag, hipStream_t s) const;
  int inverse_diagonal(double *inv, hipStream_t s) const;
  // T_zero: T is the zero vector (the right-hand side of a linear step), so K T is not applied
  int create_rhs(double *dst, const double *T, const double *T_old, const double *q, hipStream_t s,
                 bool T_zero = false) const;
  int l2_norm(const double *u, double *norm, hipStream_t s) const;
};

// ---------------------------------------------------------------------------
// Preconditioner concept update()/vmult()/reinit()
// (include/meltpooldg/linear_algebra/preconditioner.hpp:19-37)
// ---------------------------------------------------------------------------
struct mpdg_precond {
  virtual ~mpdg_precond() = default;
  int type = MPDG_PRECOND_IDENTITY;
  mpdg_heat_op *op = nullptr;
  virtual int update(hipStream_t s) = 0;
  virtual int vmult(double *dst, const double *src, hipStream_t s) = 0;
  virtual const double *inverse_diagonal() const { return nullptr; }
  // dst = an approximate solution of J dst = src by `degree` smoothing steps from zero on the
  // finest level (multigrid: its Chebyshev smoother; others: the preconditioner itself)
  virtual int smooth(double *dst, const double *src, int degree, hipStream_t s) {
    (void)degree;
    return vmult(dst, src, s);
  }
  // an FP32 preconditioner that runs in place: the Krylov solver writes its input and reads its
  // output in FP32 (nullptr: not available)
  virtual float *f32_input() { return nullptr; }
  virtual const float *f32_output() const { return nullptr; }
  virtual int vmult_f32_in_place(hipStream_t) { return MPDG_ERR_INVALID; }
  virtual int n_levels() const { return 0; }
  virtual int level_info(int, int *, int *, double *) const { return MPDG_ERR_INVALID; }
};

struct mpdg_heat_operation;
