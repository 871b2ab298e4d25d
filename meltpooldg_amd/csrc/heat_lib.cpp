// Host side of the MI355X heat operator library and its C ABI (include/mpdg_heat.h).
#include "heat_lib.hpp"
#include "tangent.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>

using namespace mpdg;

namespace {

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int dev_ok() { return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE; }

}  // namespace

// ===========================================================================
// operator
// ===========================================================================
int mpdg_heat_op::ensure_edge() const {
  const size_t n = apply_edge_elems(d64.degree, d64.nc, precision == MPDG_FP32);
  if (precision == MPDG_FP64) {
    MPDG_TRY(edge64.resize_zeroed(n));
    d64.edge = edge64.get();
    d64.edge_elems = n;
  } else {
    MPDG_TRY(edge32.resize_zeroed(n));
    d32.edge = edge32.get();
    d32.edge_elems = n;
  }
  return MPDG_OK;
}

int mpdg_heat_op::vmult_cells(double *dst, const double *src, hipStream_t s) const {
  if (precision != MPDG_FP64) return MPDG_ERR_INVALID;
  MPDG_TRY(ensure_edge());
  if (d64.metric && d64.kappa_mode == KAPPA_LAW) MPDG_TRY(refresh_metric_k(s));  // k(T) with the current T
  if (tangent) {
    if (!td_bound) return MPDG_ERR_INVALID;
    return launch_tangent_apply(d64, td.get(), tangent_convects(vel, vfield), solid, src, dst, s);
  }
  if (!timing_on) return launch_apply<double>(d64, src, dst, s);
  if (timing_used == timing_events.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return MPDG_ERR_DEVICE;
    timing_events.push_back({a, b});
  }
  auto &ev = timing_events[timing_used++];
  return launch_apply<double>(d64, src, dst, s, ev.first, ev.second);
}

int mpdg_heat_op::refresh_metric_k(hipStream_t s) const {
  const bool needs = (precision == MPDG_FP64 ? d64.metric : (const double *)d32.metric) &&
                     (d64.kappa_mode == KAPPA_QP || d64.kappa_mode == KAPPA_LAW);
  if (!needs) return MPDG_OK;
  const int n1 = d64.degree + 1;
  const size_t nq = (size_t)d64.n_cells * n1 * n1 * n1;
  if (precision == MPDG_FP64) {
    MPDG_TRY(kq.resize(nq));
    MPDG_TRY(metric64_k.resize(nq * 7));
    MPDG_TRY(launch_kappa_qp<double>(d64, kq.get(), s));
    MPDG_TRY(launch_scale_metric<double>(d64, kq.get(), metric64_k.get(), s));
    d64.metric_k = metric64_k.get();
  } else {
    MPDG_TRY(kq32.resize(nq));
    MPDG_TRY(metric32_k.resize(nq * 7));
    MPDG_TRY(launch_kappa_qp<float>(d32, kq32.get(), s));
    MPDG_TRY(launch_scale_metric<float>(d32, kq32.get(), metric32_k.get(), s));
    d32.metric_k = metric32_k.get();
  }
  return MPDG_OK;
}

// HeatDiffuseMultiPhaseOperator::vmult: cell loop + boundary loop, dst zeroed
// (heat_diffuse_operator.cpp:181-197)
int mpdg_heat_op::vmult(double *dst, const double *src, hipStream_t s) const {
  MPDG_TRY(vmult_cells(dst, src, s));
  if (d64.robin_mask) MPDG_TRY(launch_boundary(d64, 0, nullptr, dst, src, T_robin, s));
  return MPDG_OK;
}

int mpdg_heat_op::vmult_f32(float *dst, const float *src, hipStream_t s) const {
  if (precision != MPDG_FP32 || tangent) return MPDG_ERR_INVALID;
  MPDG_TRY(ensure_edge());
  if (d32.metric && d32.kappa_mode == KAPPA_LAW) MPDG_TRY(refresh_metric_k(s));
  return launch_apply<float>(d32, src, dst, s);
}

int mpdg_heat_op::diagonal64(double *diag, hipStream_t s) const {
  if (tangent) {
    if (!td_bound) return MPDG_ERR_INVALID;
    MPDG_TRY(launch_tangent_diagonal(d64, td.get(), tangent_convects(vel, vfield), solid, diag, s));
    if (d64.robin_mask) MPDG_TRY(launch_boundary(d64, 1, nullptr, diag, nullptr, T_robin, s));
    return MPDG_OK;
  }
  const int n1 = d64.degree + 1;
  MPDG_TRY(refresh_metric_k(s));
  MPDG_TRY(diag_scratch.resize(2 * (size_t)d64.n_cells * n1 * n1 * n1));
  MPDG_TRY(launch_diagonal<double>(d64, diag, diag_scratch.get(), s));
  if (d64.robin_mask) MPDG_TRY(launch_boundary(d64, 1, nullptr, diag, nullptr, T_robin, s));
  return MPDG_OK;
}

// compute_inverse_diagonal_from_matrixfree (heat_diffuse_operator.cpp:273-284)
int mpdg_heat_op::inverse_diagonal(double *inv, hipStream_t s) const {
  MPDG_TRY(diagonal64(inv, s));
  return launch_invert_diagonal<double>(inv, d64.n_dofs, s);
}

// create_rhs (heat_diffuse_operator.cpp:785-806, cell part :395-529, faces :721-781):
//   dst += -[(c_m (T - T_old) - q) N + k(T) grad T . grad N] + boundary terms.
// As in the reference the difference T - T_old is formed before the mass is applied
// (:493-497), which keeps the residual free of the cancellation between c_m M T and
// c_m M T_old:  dst -= M_{c_m} (T - T_old - q/c_m) + K_{k(T)} T, then the boundary residual.
int mpdg_heat_op::create_rhs(double *dst, const double *T, const double *T_old, const double *q, hipStream_t s,
                             bool T_zero) const {
  if (precision != MPDG_FP64) return MPDG_ERR_INVALID;
  if (dt == 0.0) return MPDG_ERR_ZERO_TIME_INCREMENT;
  if (tangent) {
    // residual with the tangent's material and velocity (heat_diffuse_operator.cpp:467-527)
    MPDG_TRY(launch_tangent_residual(d64, dt, mat, vel, vfield, T, T_old, q, dst, s));
    return launch_boundary(d64, 2, flux, dst, nullptr, T, s);
  }
  const long n = d64.n_dofs;
  MPDG_TRY(tmp1.resize(n));
  MPDG_TRY(tmp2.resize(n));
  MPDG_TRY(ensure_edge());
  if (T_zero && !tangent) {
    // dst += c_m M (T_old + q / c_m) + boundary terms at T = 0 (K T vanishes)
    OperatorData<double> mass = d64;
    mass.kappa_mode = KAPPA_CONST;
    mass.kappa = 0.0;
    mass.robin_mask = 0;
    const double *src = T_old;
    if (q) {
      MPDG_TRY(launch_scale(tmp2.get(), T_old, 1.0, n, s));
      MPDG_TRY(launch_axpby(tmp2.get(), 1.0 / d64.c_mass, q, 1.0, n, s));
      src = tmp2.get();
    }
    MPDG_TRY(launch_apply<double>(mass, src, tmp1.get(), s));
    MPDG_TRY(launch_axpby(dst, 1.0, tmp1.get(), 1.0, n, s));
    return launch_boundary(d64, 2, flux, dst, nullptr, T, s);
  }
  // tmp2 = T - T_old - q / c_m ; tmp1 = c_m M tmp2
  MPDG_TRY(launch_convert<double, double>(T, tmp2.get(), n, s));
  MPDG_TRY(launch_axpby(tmp2.get(), -1.0, T_old, 1.0, n, s));
  if (q) MPDG_TRY(launch_axpby(tmp2.get(), -1.0 / d64.c_mass, q, 1.0, n, s));
  OperatorData<double> mass = d64;
  mass.kappa_mode = KAPPA_CONST;
  mass.kappa = 0.0;
  mass.robin_mask = 0;
  MPDG_TRY(launch_apply<double>(mass, tmp2.get(), tmp1.get(), s));
  MPDG_TRY(launch_axpby(dst, -1.0, tmp1.get(), 1.0, n, s));
  // tmp1 = K_{k(T)} T, conductivity law evaluated at T
  OperatorData<double> jt = d64;
  jt.c_mass = 0.0;
  if (jt.kappa_mode == KAPPA_LAW) {
    jt.T = T;
    if (jt.metric) {
      const int n1 = d64.degree + 1;
      const size_t nq = (size_t)d64.n_cells * n1 * n1 * n1;
      MPDG_TRY(kq.resize(nq));
      MPDG_TRY(metric64_k.resize(nq * 7));
      MPDG_TRY(launch_kappa_qp<double>(jt, kq.get(), s));
      MPDG_TRY(launch_scale_metric<double>(jt, kq.get(), metric64_k.get(), s));
      jt.metric_k = metric64_k.get();
    }
  }
  MPDG_TRY(launch_apply<double>(jt, T, tmp1.get(), s));
  MPDG_TRY(launch_axpby(dst, -1.0, tmp1.get(), 1.0, n, s));
  // boundary residuals and surface flux
  MPDG_TRY(launch_boundary(d64, 2, flux, dst, nullptr, T, s));
  return MPDG_OK;
}

// ||u||_L2^2 = u . (M u) with the operator's quadrature (VectorTools::compute_norm)
int mpdg_heat_op::l2_norm(const double *u, double *norm, hipStream_t s) const {
  if (precision != MPDG_FP64) return MPDG_ERR_INVALID;
  const long n = d64.n_dofs;
  MPDG_TRY(tmp1.resize(n));
  MPDG_TRY(scal.resize(4));
  MPDG_TRY(red.init());
  MPDG_TRY(ensure_edge());
  OperatorData<double> mass = d64;
  mass.c_mass = 1.0;
  mass.kappa_mode = KAPPA_CONST;
  mass.kappa = 0.0;
  mass.dmask = 0;
  MPDG_TRY(launch_apply<double>(mass, u, tmp1.get(), s));
  MPDG_TRY(launch_dot<double, double>(u, tmp1.get(), n, red.red(), scal.get(), s));
  double h = 0.0;
  if (hipMemcpyAsync(&h, scal.get(), sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return MPDG_ERR_DEVICE;
  if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
  *norm = std::sqrt(std::max(h, 0.0));
  return MPDG_OK;
}

// ===========================================================================
// preconditioners
// ===========================================================================
namespace {

struct IdentityPrecond : mpdg_precond {
  int update(hipStream_t) override { return MPDG_OK; }
  int vmult(double *dst, const double *src, hipStream_t s) override {
    return launch_convert<double, double>(src, dst, op->n_dofs(), s);
  }
};

// JacobiPreconditioner (include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:17-89)
struct DiagonalPrecond : mpdg_precond {
  DeviceArray<double> inv;
  int update(hipStream_t s) override {
    MPDG_TRY(inv.resize(op->n_dofs()));
    return op->inverse_diagonal(inv.get(), s);
  }
  int vmult(double *dst, const double *src, hipStream_t s) override;
  const double *inverse_diagonal() const override { return inv.get(); }
};

__global__ void scale_kernel(double *dst, const double *a, const double *b, long n) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = a[i] * b[i];
}
int DiagonalPrecond::vmult(double *dst, const double *src, hipStream_t s) {
  const long n = op->n_dofs();
  hipLaunchKernelGGL(scale_kernel, dim3(std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, dst, src,
                     inv.get(), n);
  return dev_ok();
}

// Largest eigenvalue of a symmetric tridiagonal matrix (cyclic Jacobi rotations).
double tridiag_max_eig(const std::vector<double> &d, const std::vector<double> &e) {
  const int n = (int)d.size();
  if (n == 0) return 0.0;
  std::vector<double> A(n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    A[i * n + i] = d[i];
    if (i + 1 < n) A[i * n + i + 1] = A[(i + 1) * n + i] = e[i];
  }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < n; ++k) {
          const double a = A[k * n + p], b = A[k * n + q];
          A[k * n + p] = c * a - sn * b;
          A[k * n + q] = sn * a + c * b;
        }
        for (int k = 0; k < n; ++k) {
          const double a = A[p * n + k], b = A[q * n + k];
          A[p * n + k] = c * a - sn * b;
          A[q * n + k] = sn * a + c * b;
        }
      }
  }
  double mx = A[0];
  for (int i = 1; i < n; ++i) mx = std::max(mx, A[i * n + i]);
  return mx;
}

// 1D evaluation of a fine Q_kf function at the nodes of (kc, ncc): carries the mapping
// node coordinates from a level to the next coarser one.
struct Interp1DHost {
  int nf = 0, nc = 0, w = 0;
  std::vector<int> col;
  std::vector<double> val;
  Interp1DHost(int kf, int ncf, int kc, int ncc) {
    FE1D fc(kc), ff(kf);
    nf = kf * ncf + 1;
    nc = kc * ncc + 1;
    w = kf + 1;
    col.assign((size_t)nc * w, -1);
    val.assign((size_t)nc * w, 0.0);
    for (int g = 0; g < nc; ++g) {
      const int cc = std::min(g / kc, ncc - 1), i = g - kc * cc;
      const double xi = (cc + fc.nodes[i]) * (double)ncf / (double)ncc;
      const int cf = std::min((int)std::floor(xi), ncf - 1);
      double v[kMaxN1];
      FE1D::lagrange(kf + 1, ff.nodes, xi - cf, v, nullptr);
      for (int j = 0; j < w; ++j) {
        col[(size_t)g * w + j] = kf * cf + j;
        val[(size_t)g * w + j] = v[j];
      }
    }
  }
};

template <typename T>
struct Level {
  OperatorData<T> od;
  DeviceArray<T> x, b, r, d, Ax, inv, kcell, kqp, Tlev, kq_scratch, diag_scratch;
  double lambda_max = 0.0;
  // transfer from the next coarser level (unused on the coarsest)
  DeviceArray<int> pcol[3], rcol[3];
  DeviceArray<T> pval[3], rval[3];
  int pw[3] = {0, 0, 0}, rw[3] = {0, 0, 0};
  // 2:1 h-transfer of the same degree: the table of one coarse cell (structured kernels)
  bool hstruct = false;
  std::vector<double> htab;
  DeviceArray<T> htab_dev;
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  DeviceArray<T> t1, t2;
  std::vector<double> cheb_c1, cheb_c2;
  DeviceArray<double> coords;  // mapping node coordinates of this level (mapped meshes)
  DeviceArray<T> metric, metric_k;
  // the level's own apply scratch and reduction scratch: update() runs the levels' diagonals
  // and eigenvalue estimates concurrently, one stream per level
  DeviceArray<T> edge;
  ReductionScratch red;
  DeviceArray<double> eig_sc;
};

template <typename T>
int upload(DeviceArray<T> &dst, const std::vector<double> &src) {
  MPDG_TRY(dst.resize(src.size()));
  std::vector<T> h(src.begin(), src.end());
  return hipMemcpy(dst.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess
             ? MPDG_OK
             : MPDG_ERR_DEVICE;
}
int upload_int(DeviceArray<int> &dst, const std::vector<int> &src) {
  MPDG_TRY(dst.resize(src.size()));
  return hipMemcpy(dst.get(), src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess
             ? MPDG_OK
             : MPDG_ERR_DEVICE;
}

template <typename T>
int level_apply(const OperatorData<T> &od, const T *src, T *dst, hipStream_t s) {
  return launch_apply<T>(od, src, dst, s);
}

// Geometric multigrid V-cycle: recipe of unit_tests/gmg.cc:343-383, :447-635
// (global coarsening, Chebyshev smoother with the inverse diagonal, eigenvalue
// estimate by CG), with level operators in precision T, an optional p-coarsening
// on the coarsest mesh and an FP64 direct coarse solve.
template <typename T>
struct GMGPrecond : mpdg_precond {
  mpdg_gmg_params prm{};
  std::vector<std::unique_ptr<Level<T>>> lev;  // 0 = coarsest
  DeviceArray<double> coarse_inv, coarse_elem, cvec_part, eig_hist, kcell64, kcell_pp[2];
  ReductionScratch red;
  int nc_coarse_dofs = 0;
  // update(): the coarse matrix and the levels below the finest run on streams of their own
  // (default priority), joined to the caller's stream by events; the finest level stays on it
  std::vector<hipStream_t> lstreams;
  std::vector<hipEvent_t> levents;
  hipEvent_t ev_setup = nullptr;
  // V-cycle replay (vcycle_graph): the kernels of one V-cycle read fixed level buffers, and their
  // arguments (level data, smoother coefficients) change only in update(); so the cycle is
  // captured on a stream of its own at the first vmult after an update and then replayed on the
  // caller's stream as one graph launch. update() drops the graph.
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t vgraph = nullptr;
  ~GMGPrecond() override {
    drop_vcycle_graph();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (hipStream_t st : lstreams) (void)hipStreamDestroy(st);
    for (hipEvent_t e : levents) (void)hipEventDestroy(e);
    if (ev_setup) (void)hipEventDestroy(ev_setup);
  }
  void drop_vcycle_graph() {
    if (vgraph) (void)hipGraphExecDestroy(vgraph);
    vgraph = nullptr;
  }
  // one V-cycle from the finest level's b into its x, on stream s
  int run_vcycle(hipStream_t s) {
    const int top = (int)lev.size() - 1;
    if (!prm.vcycle_graph) return vcycle(top, s);
    if (!vgraph) {
      if (!cap_stream && hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking) != hipSuccess)
        return MPDG_ERR_DEVICE;
      if (hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return MPDG_ERR_DEVICE;
      const int rc = vcycle(top, cap_stream);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(cap_stream, &g);
      hipError_t ei = hipErrorUnknown;
      if (rc == MPDG_OK && ec == hipSuccess && g) ei = hipGraphInstantiate(&vgraph, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      if (rc != MPDG_OK) return rc;
      if (ei != hipSuccess) {
        vgraph = nullptr;
        return MPDG_ERR_DEVICE;
      }
    }
    return hipGraphLaunch(vgraph, s) == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
  }

  int build() {
    const auto &f = op->d64;
    std::vector<OperatorData<T>> ods;
    OperatorData<T> top;
    top.set_mesh(f.degree, f.nc, f.L);
    top.c_mass = f.c_mass;
    top.kappa_mode = f.kappa_mode;
    top.kappa = f.kappa;
    top.law = f.law;
    for (int i = 0; i < 4; ++i) top.law_p[i] = f.law_p[i];
    top.dmask = f.dmask;
    ods.push_back(top);
    while (true) {
      const OperatorData<T> &p = ods.back();
      if (p.nc[0] <= prm.n_cells_coarse || p.nc[0] % 2 || p.nc[1] % 2 || p.nc[2] % 2) break;
      OperatorData<T> c = p;
      const int nc[3] = {p.nc[0] / 2, p.nc[1] / 2, p.nc[2] / 2};
      c.set_mesh(p.degree, nc, p.L);
      c.kappa_mode = KAPPA_CELL;
      ods.push_back(c);
    }
    if (prm.coarse_degree > 0 && prm.coarse_degree < f.degree) {
      OperatorData<T> c = ods.back();
      c.set_mesh(prm.coarse_degree, ods.back().nc, ods.back().L);
      c.kappa_mode = KAPPA_CELL;
      ods.push_back(c);
    }
    lev.clear();
    for (auto it = ods.rbegin(); it != ods.rend(); ++it) {
      auto L = std::make_unique<Level<T>>();
      L->od = *it;
      lev.push_back(std::move(L));
    }
    for (size_t l = 0; l < lev.size(); ++l) {
      Level<T> &L = *lev[l];
      const long n = L.od.n_dofs;
      MPDG_TRY(L.x.resize(n));
      MPDG_TRY(L.b.resize(n));
      MPDG_TRY(L.r.resize(n));
      MPDG_TRY(L.d.resize(n));
      MPDG_TRY(L.Ax.resize(n));
      MPDG_TRY(L.inv.resize(n));
      MPDG_TRY(L.kcell.resize(L.od.n_cells));
      if (L.od.kappa_mode == KAPPA_CELL) L.od.kappa_arr = L.kcell.get();
      if (l == 0) continue;
      const Level<T> &C = *lev[l - 1];
      long t1 = 0, t2 = 0;
      for (int d = 0; d < 3; ++d) {
        Transfer1DHost tr(C.od.degree, C.od.nc[d], L.od.degree, L.od.nc[d], (f.dmask >> (2 * d)) & 1,
                          (f.dmask >> (2 * d + 1)) & 1);
        L.pw[d] = tr.pw;
        L.rw[d] = tr.rw;
        if (d == 0) {  // entry-major along x
          MPDG_TRY(upload_int(L.pcol[d], ell_entry_major(tr.pcol, tr.pw)));
          MPDG_TRY(upload_int(L.rcol[d], ell_entry_major(tr.rcol, tr.rw)));
          MPDG_TRY(upload<T>(L.pval[d], ell_entry_major(tr.pval, tr.pw)));
          MPDG_TRY(upload<T>(L.rval[d], ell_entry_major(tr.rval, tr.rw)));
        } else {
          MPDG_TRY(upload_int(L.pcol[d], tr.pcol));
          MPDG_TRY(upload_int(L.rcol[d], tr.rcol));
          MPDG_TRY(upload<T>(L.pval[d], tr.pval));
          MPDG_TRY(upload<T>(L.rval[d], tr.rval));
        }
        L.lo[d] = (f.dmask >> (2 * d)) & 1;
        L.hi[d] = (f.dmask >> (2 * d + 1)) & 1;
      }
      L.hstruct = C.od.degree == L.od.degree && L.od.nd[0] <= 1024;
      for (int d = 0; d < 3; ++d) L.hstruct = L.hstruct && L.od.nc[d] == 2 * C.od.nc[d];
      if (L.hstruct) {
        const int k = L.od.degree;
        Transfer1DHost cell(k, 1, k, 2, false, false);  // one coarse cell: rows m = 0..2k, cols j = 0..k
        L.htab.assign((size_t)(2 * k + 1) * (k + 1), 0.0);
        for (int m = 0; m <= 2 * k; ++m)
          for (int e = 0; e < cell.pw; ++e) {
            const int j = cell.pcol[(size_t)m * cell.pw + e];
            if (j >= 0) L.htab[(size_t)m * (k + 1) + j] = cell.pval[(size_t)m * cell.pw + e];
          }
        MPDG_TRY(upload<T>(L.htab_dev, L.htab));
      }
      t1 = (long)L.od.nd[0] * C.od.nd[1] * C.od.nd[2];
      t2 = (long)L.od.nd[0] * L.od.nd[1] * C.od.nd[2];
      MPDG_TRY(L.t1.resize(t1));
      MPDG_TRY(L.t2.resize(t2));
    }
    for (auto &L : lev) {
      const size_t ne = apply_edge_elems(L->od.degree, L->od.nc, std::is_same<T, float>::value);
      MPDG_TRY(L->edge.resize_zeroed(ne));
      L->od.edge = L->edge.get();
      L->od.edge_elems = ne;
      MPDG_TRY(L->red.init());
      MPDG_TRY(L->eig_sc.resize(4));
    }
    if (hipEventCreateWithFlags(&ev_setup, hipEventDisableTiming) != hipSuccess) return MPDG_ERR_DEVICE;
    for (size_t l = 0; l + 1 < lev.size(); ++l) {
      hipStream_t st;
      hipEvent_t e;
      if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return MPDG_ERR_DEVICE;
      lstreams.push_back(st);
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return MPDG_ERR_DEVICE;
      levents.push_back(e);
    }
    if (op->coords) MPDG_TRY(build_geometry());
    MPDG_TRY(red.init());
    MPDG_TRY(eig_hist.resize(std::max<size_t>(lev.size(), 2) * hist_stride()));
    nc_coarse_dofs = (int)lev[0]->od.n_dofs;
    MPDG_TRY(coarse_inv.resize((size_t)nc_coarse_dofs * nc_coarse_dofs));
    MPDG_TRY(cvec_part.resize(dense_gemv_scratch((int)nc_coarse_dofs)));
    MPDG_TRY(kcell64.resize(f.n_cells));
    for (auto &k : kcell_pp) MPDG_TRY(k.resize(std::max<long>(1, f.n_cells / 8)));
    return MPDG_OK;
  }

  // mapped meshes: every level's node coordinates are the next finer level's mapping
  // evaluated at its nodes; each level keeps its metric in precision T
  int build_geometry() {
    const auto &f = op->d64;
    Level<T> &top = *lev.back();
    MPDG_TRY(top.coords.resize(3 * (size_t)f.n_dofs));
    if (hipMemcpy(top.coords.get(), op->coords, 3 * f.n_dofs * sizeof(double), hipMemcpyDeviceToDevice) != hipSuccess)
      return MPDG_ERR_DEVICE;
    for (int l = (int)lev.size() - 2; l >= 0; --l) {
      Level<T> &P = *lev[l + 1];
      Level<T> &L = *lev[l];
      DeviceArray<int> col[3];
      DeviceArray<double> val[3];
      int w[3];
      for (int d = 0; d < 3; ++d) {
        Interp1DHost e(P.od.degree, P.od.nc[d], L.od.degree, L.od.nc[d]);
        w[d] = e.w;
        MPDG_TRY(upload_int(col[d], d == 0 ? ell_entry_major(e.col, e.w) : e.col));
        MPDG_TRY(upload<double>(val[d], d == 0 ? ell_entry_major(e.val, e.w) : e.val));
      }
      MPDG_TRY(L.coords.resize(3 * (size_t)L.od.n_dofs));
      DeviceArray<double> a, b;
      const int f3[3] = {P.od.nd[0], P.od.nd[1], P.od.nd[2]};
      const int o1[3] = {L.od.nd[0], P.od.nd[1], P.od.nd[2]}, o2[3] = {L.od.nd[0], L.od.nd[1], P.od.nd[2]};
      const int o3[3] = {L.od.nd[0], L.od.nd[1], L.od.nd[2]};
      MPDG_TRY(a.resize((size_t)o1[0] * o1[1] * o1[2]));
      MPDG_TRY(b.resize((size_t)o2[0] * o2[1] * o2[2]));
      for (int c = 0; c < 3; ++c) {
        MPDG_TRY(launch_ell_dir<double>(P.coords.get() + c * P.od.n_dofs, a.get(), o1, f3, 0, w[0], col[0].get(),
                                        val[0].get(), nullptr));
        MPDG_TRY(launch_ell_dir<double>(a.get(), b.get(), o2, o1, 1, w[1], col[1].get(), val[1].get(), nullptr));
        MPDG_TRY(launch_ell_dir<double>(b.get(), L.coords.get() + c * L.od.n_dofs, o3, o2, 2, w[2], col[2].get(),
                                        val[2].get(), nullptr));
      }
      if (hipDeviceSynchronize() != hipSuccess) return MPDG_ERR_DEVICE;
    }
    for (auto &L : lev) {
      const int n1 = L->od.degree + 1;
      MPDG_TRY(L->metric.resize((size_t)L->od.n_cells * 7 * n1 * n1 * n1));
      MPDG_TRY(launch_metric<T>(L->od, L->coords.get(), L->metric.get(), nullptr));
      L->od.metric = L->metric.get();
      L->od.coords = L->coords.get();
    }
    return hipDeviceSynchronize() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
  }

  // level data that depends on the fine operator's coefficient and time increment
  int update(hipStream_t s) override {
    if (op->tangent) return MPDG_ERR_INVALID;
    drop_vcycle_graph();
    const auto &f = op->d64;
    Level<T> &top = *lev.back();
    for (auto &L : lev) L->od.c_mass = f.c_mass;
    // finest level: the fine operator's conductivity in precision T
    if (op->coords && !top.od.metric) MPDG_TRY(build_geometry());
    top.od.kappa_mode = f.kappa_mode;
    top.od.kappa = f.kappa;
    if (f.kappa_mode == KAPPA_QP) {
      const long nq = f.n_cells * (long)((f.degree + 1) * (f.degree + 1) * (f.degree + 1));
      if (std::is_same<T, double>::value) top.od.kappa_arr = (const T *)f.kappa_arr;
      else {
        MPDG_TRY(top.kqp.resize(nq));
        MPDG_TRY(launch_convert<double, T>(f.kappa_arr, top.kqp.get(), nq, s));
        top.od.kappa_arr = top.kqp.get();
      }
    } else if (f.kappa_mode == KAPPA_CELL) {
      MPDG_TRY(launch_convert<double, T>(f.kappa_arr, top.kcell.get(), f.n_cells, s));
      top.od.kappa_arr = top.kcell.get();
    } else if (f.kappa_mode == KAPPA_LAW) {
      if (std::is_same<T, double>::value) top.od.T = (const T *)f.T;
      else {
        MPDG_TRY(top.Tlev.resize(f.n_dofs));
        MPDG_TRY(launch_convert<double, T>(f.T, top.Tlev.get(), f.n_dofs, s));
        top.od.T = top.Tlev.get();
      }
    }
    // mapped mesh with per-qp / k(T) conductivity: fold it into the level metric
    if (top.od.metric && (top.od.kappa_mode == KAPPA_QP || top.od.kappa_mode == KAPPA_LAW)) {
      const int n1 = top.od.degree + 1;
      const size_t nq = (size_t)top.od.n_cells * n1 * n1 * n1;
      MPDG_TRY(top.kq_scratch.resize(nq));
      MPDG_TRY(top.metric_k.resize(nq * 7));
      MPDG_TRY(launch_kappa_qp<T>(top.od, top.kq_scratch.get(), s));
      MPDG_TRY(launch_scale_metric<T>(top.od, top.kq_scratch.get(), top.metric_k.get(), s));
      top.od.metric_k = top.metric_k.get();
    }
    // quadrature-weighted cell averages of the fine conductivity, coarsened level by level
    if (lev.size() > 1) {
      MPDG_TRY(launch_cell_average<double>(f, kcell64.get(), s));
      const double *kf = kcell64.get();
      int nc[3] = {f.nc[0], f.nc[1], f.nc[2]}, pp = 0;
      for (int l = (int)lev.size() - 2; l >= 0; --l) {
        Level<T> &L = *lev[l];
        if (L.od.nc[0] != nc[0]) {
          double *kc = kcell_pp[pp].get();
          pp ^= 1;
          MPDG_TRY(launch_coarsen_cells<double>(nc, kf, kc, s));
          kf = kc;
          for (int d = 0; d < 3; ++d) nc[d] = L.od.nc[d];
        }
        MPDG_TRY(launch_convert<double, T>(kf, L.kcell.get(), L.od.n_cells, s));
        L.od.kappa_arr = L.kcell.get();
      }
      // the coarse matrix and its inverse, concurrently with the levels' set-up below (kf: the
      // coarsest cells' conductivities, not written again in this update)
      if (hipEventRecord(ev_setup, s) != hipSuccess) return MPDG_ERR_DEVICE;
      if (hipStreamWaitEvent(lstreams[0], ev_setup, 0) != hipSuccess) return MPDG_ERR_DEVICE;
      MPDG_TRY(assemble_coarse(kf, lstreams[0]));
      if (hipEventRecord(levents[0], lstreams[0]) != hipSuccess) return MPDG_ERR_DEVICE;
    } else {
      MPDG_TRY(assemble_coarse(nullptr, s));
    }
    // inverse diagonals and eigenvalue estimates, level l > 0 on its own stream (the finest on
    // the caller's): the levels share no data here, and the coarse ones are short, launch-bound
    // chains that run beside the finest level's
    for (size_t l = 1; l < lev.size(); ++l) {
      Level<T> &L = *lev[l];
      const bool own = l + 1 < lev.size();
      hipStream_t sl = own ? lstreams[l] : s;
      if (own && hipStreamWaitEvent(sl, ev_setup, 0) != hipSuccess) return MPDG_ERR_DEVICE;
      const int n1 = L.od.degree + 1;
      MPDG_TRY(L.diag_scratch.resize(2 * (size_t)L.od.n_cells * n1 * n1 * n1));
      MPDG_TRY(launch_diagonal<T>(L.od, L.inv.get(), L.diag_scratch.get(), sl));
      MPDG_TRY(launch_invert_diagonal<T>(L.inv.get(), L.od.n_dofs, sl));
      // smoothing leaves the constrained DoFs at zero
      MPDG_TRY(launch_zero_constrained<T>(L.inv.get(), L.od.nd, L.od.dmask, sl));
      MPDG_TRY(run_eigenvalue_cg(L, eig_hist.get() + (l - 1) * hist_stride(), sl));
      if (own && hipEventRecord(levents[l], sl) != hipSuccess) return MPDG_ERR_DEVICE;
    }
    if (lev.size() > 1)
      for (size_t l = 0; l + 1 < lev.size(); ++l)
        if (hipStreamWaitEvent(s, levents[l], 0) != hipSuccess) return MPDG_ERR_DEVICE;
    // one copy of every level's CG coefficients, then lambda_max and the Chebyshev coefficients
    const size_t nh = (lev.size() - 1) * hist_stride();
    std::vector<double> h(nh);
    if (nh) {
      if (hipMemcpyAsync(h.data(), eig_hist.get(), nh * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
        return MPDG_ERR_DEVICE;
      if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
    }
    for (size_t l = 1; l < lev.size(); ++l) {
      lev[l]->lambda_max = lambda_max_from(h.data() + (l - 1) * hist_stride());
      set_chebyshev(*lev[l]);
    }
    return MPDG_OK;
  }

  size_t hist_stride() const { return 2 * (size_t)std::max(prm.eig_cg_n_iterations, 1) + 2; }

  // eig_cg_n_iterations CG steps on L x = start vector with the inverse diagonal; the step
  // coefficients (alpha_k, beta_k) go to hist[2k], hist[2k + 1]
  int run_eigenvalue_cg(Level<T> &L, double *hist, hipStream_t s) {
    const long n = L.od.n_dofs;
    const int it = prm.eig_cg_n_iterations;
    // r = start vector (b, x0 = 0), z = inv r, p = z
    MPDG_TRY(launch_eig_start<T>(L.r.get(), L.od.nd, L.od.dmask, s));
    MPDG_TRY(launch_eig_step<T>(0, L.r.get(), L.b.get(), L.d.get(), nullptr, L.inv.get(), n, L.red.red(),
                                L.eig_sc.get(), hist, 0, s));
    for (int k = 0; k < it; ++k) {
      MPDG_TRY(level_apply<T>(L.od, L.d.get(), L.Ax.get(), s));
      MPDG_TRY(launch_eig_step<T>(1, L.r.get(), L.b.get(), L.d.get(), L.Ax.get(), L.inv.get(), n, L.red.red(),
                                  L.eig_sc.get(), hist, k, s));
      MPDG_TRY(launch_eig_step<T>(2, L.r.get(), L.b.get(), L.d.get(), L.Ax.get(), L.inv.get(), n, L.red.red(),
                                  L.eig_sc.get(), hist, k, s));
      MPDG_TRY(launch_eig_step<T>(3, L.r.get(), L.b.get(), L.d.get(), L.Ax.get(), L.inv.get(), n, L.red.red(),
                                  L.eig_sc.get(), hist, k, s));
    }
    return MPDG_OK;
  }

  // 1.2 x the largest eigenvalue of the Lanczos tridiagonal matrix of the CG coefficients
  // (deal.II's PreconditionChebyshev estimate)
  double lambda_max_from(const double *h) const {
    const int it = prm.eig_cg_n_iterations;
    std::vector<double> alphas, betas;
    for (int k = 0; k < it; ++k) {
      const double a = h[2 * k], b = h[2 * k + 1];
      if (!(a == a) || !(b == b)) break;
      alphas.push_back(a);
      betas.push_back(b);
      if (b <= 0.0) break;
    }
    const int m = (int)alphas.size();
    std::vector<double> dd(m), ee(m > 0 ? m - 1 : 0);
    for (int j = 0; j < m; ++j) {
      dd[j] = 1.0 / alphas[j] + (j > 0 ? betas[j - 1] / alphas[j - 1] : 0.0);
      if (j + 1 < m) ee[j] = std::sqrt(betas[j]) / alphas[j];
    }
    return 1.2 * tridiag_max_eig(dd, ee);
  }

  // PreconditionChebyshev coefficients on [lambda_max/range, lambda_max]
  void set_chebyshev(Level<T> &L) {
    const double lmax = L.lambda_max, lmin = lmax / prm.smoothing_range;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    double rho = 1.0 / sigma;
    L.cheb_c1.assign(prm.smoother_degree, 0.0);
    L.cheb_c2.assign(prm.smoother_degree, 0.0);
    L.cheb_c2[0] = 1.0 / theta;
    for (int k = 1; k < prm.smoother_degree; ++k) {
      const double rn = 1.0 / (2.0 * sigma - rho);
      L.cheb_c1[k] = rn * rho;
      L.cheb_c2[k] = 2.0 * rn / delta;
      rho = rn;
    }
  }

  // dense coarse matrix from element matrices (FP64) with the cell-averaged conductivities kc
  // (device), Dirichlet rows/columns identity, inverted on the device
  int assemble_coarse(const double *kc, hipStream_t s) {
    const OperatorData<T> &od = lev[0]->od;
    const auto &f = op->d64;
    if (!kc && f.kappa_mode != KAPPA_CONST) {
      // single-level hierarchy: the fine operator itself is the coarse operator
      MPDG_TRY(launch_cell_average<double>(f, kcell64.get(), s));
      kc = kcell64.get();
    }
    const size_t n1 = od.degree + 1;
    MPDG_TRY(coarse_elem.resize((size_t)od.n_cells * n1 * n1 * n1 * n1 * n1 * n1));
    MPDG_TRY(launch_coarse_matrix<T>(od, f.c_mass, kc, f.kappa, coarse_inv.get(), coarse_elem.get(), s));
    // the inverse of a symmetric matrix is symmetric: row-major == column-major for the gemv
    return launch_gauss_jordan_inverse(coarse_inv.get(), (int)od.n_dofs, s);
  }

  int prolongate_add(Level<T> &F, const Level<T> &C, hipStream_t s) {
    // F.x += P C.x via x, y, z passes, the last one adding into F.x
    int o[3] = {F.od.nd[0], C.od.nd[1], C.od.nd[2]}, i[3] = {C.od.nd[0], C.od.nd[1], C.od.nd[2]};
    if (F.hstruct) {
      int o2[3] = {F.od.nd[0], F.od.nd[1], C.od.nd[2]}, o3[3] = {F.od.nd[0], F.od.nd[1], F.od.nd[2]};
      const int k = F.od.degree;
      MPDG_TRY(launch_htransfer<T>(false, 0, C.x.get(), nullptr, F.t1.get(), i, o, k, F.htab.data(),
                                   F.htab_dev.get(), F.lo[0], F.hi[0], false, s));
      MPDG_TRY(launch_htransfer<T>(false, 1, F.t1.get(), nullptr, F.t2.get(), o, o2, k, F.htab.data(),
                                   F.htab_dev.get(), F.lo[1], F.hi[1], false, s));
      return launch_htransfer<T>(false, 2, F.t2.get(), nullptr, F.x.get(), o2, o3, k, F.htab.data(),
                                 F.htab_dev.get(), F.lo[2], F.hi[2], true, s);
    }
    MPDG_TRY(launch_ell_dir<T>(C.x.get(), F.t1.get(), o, i, 0, F.pw[0], F.pcol[0].get(), F.pval[0].get(), s));
    int o2[3] = {F.od.nd[0], F.od.nd[1], C.od.nd[2]};
    MPDG_TRY(launch_ell_dir<T>(F.t1.get(), F.t2.get(), o2, o, 1, F.pw[1], F.pcol[1].get(), F.pval[1].get(), s));
    int o3[3] = {F.od.nd[0], F.od.nd[1], F.od.nd[2]};
    return launch_ell_dir<T>(F.t2.get(), F.x.get(), o3, o2, 2, F.pw[2], F.pcol[2].get(), F.pval[2].get(), s,
                             nullptr, true);
  }
  int restrict_residual(const Level<T> &F, Level<T> &C, hipStream_t s) {
    // C.b = R (F.b - F.Ax) via z, y, x passes, the residual formed in the first
    int i3[3] = {F.od.nd[0], F.od.nd[1], F.od.nd[2]}, o2[3] = {F.od.nd[0], F.od.nd[1], C.od.nd[2]};
    if (F.hstruct) {
      int o1[3] = {F.od.nd[0], C.od.nd[1], C.od.nd[2]}, o0[3] = {C.od.nd[0], C.od.nd[1], C.od.nd[2]};
      const int k = F.od.degree;
      MPDG_TRY(launch_htransfer<T>(true, 2, F.b.get(), F.Ax.get(), F.t2.get(), i3, o2, k, F.htab.data(),
                                   F.htab_dev.get(), F.lo[2], F.hi[2], false, s));
      MPDG_TRY(launch_htransfer<T>(true, 1, F.t2.get(), nullptr, F.t1.get(), o2, o1, k, F.htab.data(),
                                   F.htab_dev.get(), F.lo[1], F.hi[1], false, s));
      return launch_htransfer<T>(true, 0, F.t1.get(), nullptr, C.b.get(), o1, o0, k, F.htab.data(),
                                 F.htab_dev.get(), F.lo[0], F.hi[0], false, s);
    }
    MPDG_TRY(launch_ell_dir<T>(F.b.get(), F.t2.get(), o2, i3, 2, F.rw[2], F.rcol[2].get(), F.rval[2].get(), s,
                               F.Ax.get()));
    int o1[3] = {F.od.nd[0], C.od.nd[1], C.od.nd[2]};
    MPDG_TRY(launch_ell_dir<T>(F.t2.get(), F.t1.get(), o1, o2, 1, F.rw[1], F.rcol[1].get(), F.rval[1].get(), s));
    int o0[3] = {C.od.nd[0], C.od.nd[1], C.od.nd[2]};
    return launch_ell_dir<T>(F.t1.get(), C.b.get(), o0, o1, 0, F.rw[0], F.rcol[0].get(), F.rval[0].get(), s);
  }

  int smooth(Level<T> &L, bool x_zero, hipStream_t s, int degree_override = 0) {
    const long n = L.od.n_dofs;
    int degree = (x_zero && prm.pre_smoother_degree > 0) ? prm.pre_smoother_degree : prm.smoother_degree;
    if (degree_override > 0) degree = std::min(degree_override, prm.smoother_degree);
    for (int k = 0; k < degree; ++k) {
      const bool first = (k == 0), last = (k == degree - 1);  // cheb_c1[0] = 0: the first step ignores d
      if (first && x_zero) {
        MPDG_TRY(launch_chebyshev_step<T>(L.x.get(), L.d.get(), L.b.get(), L.Ax.get(), L.inv.get(), n, 0.0,
                                          L.cheb_c2[0], true, false, !last, s));
        continue;
      }
      MPDG_TRY(level_apply<T>(L.od, L.x.get(), L.Ax.get(), s));
      MPDG_TRY(launch_chebyshev_step<T>(L.x.get(), L.d.get(), L.b.get(), L.Ax.get(), L.inv.get(), n, L.cheb_c1[k],
                                        L.cheb_c2[k], false, !first, !last, s));
    }
    return MPDG_OK;
  }

  int vcycle(int l, hipStream_t s) {
    Level<T> &L = *lev[l];
    if (l == 0) {
      // the level's b and x directly (FP64 accumulation, one rounding to T on the way out)
      return launch_dense_gemv<T>(coarse_inv.get(), L.b.get(), L.x.get(), cvec_part.get(), (int)L.od.n_dofs, s);
    }
    MPDG_TRY(smooth(L, true, s));
    MPDG_TRY(level_apply<T>(L.od, L.x.get(), L.Ax.get(), s));
    Level<T> &C = *lev[l - 1];
    MPDG_TRY(restrict_residual(L, C, s));
    MPDG_TRY(vcycle(l - 1, s));
    MPDG_TRY(prolongate_add(L, C, s));
    return smooth(L, false, s);
  }

  int vmult(double *dst, const double *src, hipStream_t s) override {
    Level<T> &top = *lev.back();
    const long n = top.od.n_dofs;
    MPDG_TRY(launch_convert<double, T>(src, top.b.get(), n, s));
    MPDG_TRY(run_vcycle(s));
    return launch_convert<T, double>(top.x.get(), dst, n, s);
  }
  float *f32_input() override {
    if constexpr (std::is_same<T, float>::value) return lev.back()->b.get();
    return nullptr;
  }
  const float *f32_output() const override {
    if constexpr (std::is_same<T, float>::value) return lev.back()->x.get();
    return nullptr;
  }
  int vmult_f32_in_place(hipStream_t s) override {
    if constexpr (!std::is_same<T, float>::value) return MPDG_ERR_INVALID;
    return run_vcycle(s);
  }
  int smooth(double *dst, const double *src, int degree, hipStream_t s) override {
    Level<T> &top = *lev.back();
    const long n = top.od.n_dofs;
    MPDG_TRY(launch_convert<double, T>(src, top.b.get(), n, s));
    MPDG_TRY(smooth(top, true, s, degree));
    return launch_convert<T, double>(top.x.get(), dst, n, s);
  }

  int n_levels() const override { return (int)lev.size(); }
  int level_info(int l, int *degree, int *nc, double *lambda) const override {
    if (l < 0 || l >= (int)lev.size()) return MPDG_ERR_INVALID;
    *degree = lev[l]->od.degree;
    for (int d = 0; d < 3; ++d) nc[d] = lev[l]->od.nc[d];
    *lambda = lev[l]->lambda_max;
    return MPDG_OK;
  }
};

}  // namespace

// ===========================================================================
// Krylov solver: SolverCG driven as LinearSolver::solve (linear_solver.hpp:21-134)
// ===========================================================================
namespace {

struct GmresWork {
  DeviceArray<double> V, w, z, r, hd, part;
  int m = 0;
  int init(long n, int basis) {
    if (m != basis || V.n != (size_t)(basis + 1) * n) {
      MPDG_TRY(V.resize((size_t)(basis + 1) * n));
      m = basis;
    }
    MPDG_TRY(w.resize(n));
    MPDG_TRY(z.resize(n));
    MPDG_TRY(r.resize(n));
    MPDG_TRY(hd.resize(3 * kMdotMax));
    return part.resize(mdot_scratch());
  }
};

struct KrylovWork {
  DeviceArray<double> r, z, p, q, scal;
  DeviceArray<CGState> st;
  ReductionScratch red;
  GmresWork gmres;
  int init(long n) {
    MPDG_TRY(r.resize(n));
    MPDG_TRY(z.resize(n));
    MPDG_TRY(p.resize(n));
    MPDG_TRY(q.resize(n));
    MPDG_TRY(st.resize(1));
    return red.init();
  }
};

// SolverGMRES with right preconditioning (linear_solver.hpp:112-118): restarted Arnoldi with
// classical Gram-Schmidt applied twice (fused multi-dot / multi-axpy kernels, one host read of
// the Hessenberg column per iteration), Givens rotations and the small triangular solve on
// the host; x += M^-1 V y at the end of each cycle. Stops when the rotated residual estimate
// meets the tolerance (after which the true residual is recomputed) or after max_iterations.
int gmres_solve(mpdg_heat_op *op, mpdg_precond *pc, const mpdg_solver_params &prm, const double *b, double *x,
                GmresWork &g, ReductionScratch &red, double *scal, int *iterations, double *res_out,
                hipStream_t s) {
  const long n = op->n_dofs();
  const int m = prm.gmres_restart > 0 ? prm.gmres_restart : 28;
  if (m >= kMdotMax) return MPDG_ERR_INVALID;
  MPDG_TRY(g.init(n, m));
  double *V = g.V.get();
  auto precondition = [&](double *dst, const double *src) -> int {
    if (pc) return pc->vmult(dst, src, s);
    return launch_convert<double, double>(src, dst, n, s);
  };
  auto norm = [&](const double *v, double &out) -> int {
    MPDG_TRY(launch_dot<double, double>(v, v, n, red.red(), scal, s));
    double h2 = 0.0;
    if (hipMemcpyAsync(&h2, scal, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return MPDG_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
    out = std::sqrt(std::max(h2, 0.0));
    return (h2 == h2) ? MPDG_OK : MPDG_ERR_NONFINITE;
  };
  // r = b - A x
  auto residual = [&](double &out) -> int {
    MPDG_TRY(op->vmult(g.w.get(), x, s));
    MPDG_TRY(launch_convert<double, double>(b, g.r.get(), n, s));
    MPDG_TRY(launch_axpby(g.r.get(), -1.0, g.w.get(), 1.0, n, s));
    return norm(g.r.get(), out);
  };
  double res = 0.0, bnorm = 0.0;
  MPDG_TRY(residual(res));
  if (prm.criterion == MPDG_CRIT_RHS) MPDG_TRY(norm(b, bnorm));
  const double tol = prm.criterion == MPDG_CRIT_RHS ? prm.rel_tolerance * bnorm
                                                    : std::max(prm.abs_tolerance, prm.rel_tolerance * res);
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m), hh(2 * kMdotMax + 1);
  auto h = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
  int it = 0;
  while (res > tol && it < prm.max_iterations) {
    MPDG_TRY(launch_scale(V, g.r.get(), 1.0 / res, n, s));
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = res;
    int k = 0;
    while (k < m && it < prm.max_iterations) {
      ++it;
      double *vk1 = V + (size_t)(k + 1) * n;
      MPDG_TRY(precondition(g.z.get(), V + (size_t)k * n));
      MPDG_TRY(op->vmult(g.w.get(), g.z.get(), s));
      double *h1 = g.hd.get(), *h2 = g.hd.get() + kMdotMax;
      MPDG_TRY(launch_mdot(V, n, k + 1, g.w.get(), g.part.get(), h1, false, s));
      MPDG_TRY(launch_maxpy(g.w.get(), V, n, k + 1, h1, -1.0, s));
      MPDG_TRY(launch_mdot(V, n, k + 1, g.w.get(), g.part.get(), h2, false, s));
      MPDG_TRY(launch_maxpy(g.w.get(), V, n, k + 1, h2, -1.0, s));
      MPDG_TRY(launch_dot<double, double>(g.w.get(), g.w.get(), n, red.red(), g.hd.get() + 2 * kMdotMax, s));
      if (hipMemcpyAsync(hh.data(), g.hd.get(), (2 * kMdotMax + 1) * sizeof(double), hipMemcpyDeviceToHost, s) !=
          hipSuccess)
        return MPDG_ERR_DEVICE;
      if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
      for (int i = 0; i <= k; ++i) h(i, k) = hh[i] + hh[kMdotMax + i];
      const double hn = std::sqrt(std::max(hh[2 * kMdotMax], 0.0));
      if (!(hn == hn)) return MPDG_ERR_NONFINITE;
      h(k + 1, k) = hn;
      if (hn > 0.0) MPDG_TRY(launch_scale(vk1, g.w.get(), 1.0 / hn, n, s));
      for (int i = 0; i < k; ++i) {  // earlier rotations on the new column
        const double a = h(i, k), c = h(i + 1, k);
        h(i, k) = cs[i] * a + sn[i] * c;
        h(i + 1, k) = -sn[i] * a + cs[i] * c;
      }
      const double a = h(k, k), c = h(k + 1, k), d = std::sqrt(a * a + c * c);
      if (!(d == d)) return MPDG_ERR_NONFINITE;
      cs[k] = d > 0.0 ? a / d : 1.0;
      sn[k] = d > 0.0 ? c / d : 0.0;
      h(k, k) = d;
      h(k + 1, k) = 0.0;
      gv[k + 1] = -sn[k] * gv[k];
      gv[k] = cs[k] * gv[k];
      ++k;
      if (std::fabs(gv[k]) <= tol || hn == 0.0) break;
    }
    // x += M^-1 V y, H y = g
    for (int i = k - 1; i >= 0; --i) {
      double t = gv[i];
      for (int j = i + 1; j < k; ++j) t -= h(i, j) * y[j];
      y[i] = h(i, i) != 0.0 ? t / h(i, i) : 0.0;
    }
    if (hipMemcpyAsync(g.hd.get(), y.data(), k * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
      return MPDG_ERR_DEVICE;
    MPDG_TRY(launch_fill<double>(g.w.get(), 0.0, n, s));
    MPDG_TRY(launch_maxpy(g.w.get(), V, n, k, g.hd.get(), 1.0, s));
    MPDG_TRY(precondition(g.z.get(), g.w.get()));
    MPDG_TRY(launch_axpby(x, 1.0, g.z.get(), 1.0, n, s));
    MPDG_TRY(residual(res));
  }
  *iterations = it;
  *res_out = res;
  return res <= tol ? MPDG_OK : MPDG_ERR_NO_CONVERGENCE;
}

// (flexible) CG with an FP32 preconditioner run in place: the iteration of krylov_solve below
// with r also written in FP32 (the preconditioner's input) and z read in FP32 (its output)
int fcg32_solve(mpdg_heat_op *op, mpdg_precond *pc, const mpdg_solver_params &prm, const double *b, double *x,
                KrylovWork &w, int *iterations, double *res, float *r32, const float *z32, const double *r_init,
                hipStream_t s, int first_batch) {
  const long n = op->n_dofs();
  CGState h{};
  h.max_it = prm.max_iterations;
  h.flexible = prm.solver_type == MPDG_SOLVER_FCG ? 1 : 0;
  if (hipMemcpyAsync(w.st.get(), &h, sizeof(CGState), hipMemcpyHostToDevice, s) != hipSuccess) return MPDG_ERR_DEVICE;
  if (!r_init) MPDG_TRY(op->vmult(w.q.get(), x, s));
  MPDG_TRY(cg_init(b, w.r.get(), r_init ? r_init : w.q.get(), nullptr, nullptr, nullptr, false, n, w.red.red(),
                   w.st.get(), prm.rel_tolerance, prm.abs_tolerance, prm.criterion, s, r32, r_init != nullptr));
  MPDG_TRY(pc->vmult_f32_in_place(s));
  MPDG_TRY(fcg32_start(w.r.get(), z32, w.p.get(), n, w.red.red(), w.st.get(), s));
  const int batch = n > 2000000 ? 1 : 8;
  int it = 0, k_batch = std::max(batch, first_batch);
  while (true) {
    for (int k = 0; k < k_batch; ++k) {
      MPDG_TRY(op->vmult(w.q.get(), w.p.get(), s));
      MPDG_TRY(cg_alpha(w.p.get(), w.q.get(), n, w.red.red(), w.st.get(), s));
      MPDG_TRY(fcg32_update(x, w.r.get(), w.p.get(), w.q.get(), z32, r32, n, w.red.red(), w.st.get(), s));
      MPDG_TRY(pc->vmult_f32_in_place(s));
      MPDG_TRY(fcg32_beta(w.r.get(), z32, n, w.red.red(), w.st.get(), s));
      MPDG_TRY(fcg32_direction(w.p.get(), z32, n, w.st.get(), s));
    }
    it += k_batch;
    k_batch = batch;
    if (hipMemcpyAsync(&h, w.st.get(), sizeof(CGState), hipMemcpyDeviceToHost, s) != hipSuccess) return MPDG_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
    if (h.done || it > prm.max_iterations + batch) break;
  }
  *iterations = h.iters;
  *res = std::sqrt(std::max(h.rr, 0.0));
  if (h.done == 2) return MPDG_ERR_NONFINITE;
  if (h.rr > h.tol2) return MPDG_ERR_NO_CONVERGENCE;
  return MPDG_OK;
}

// r_init (may be nullptr): b - A x, already computed by the caller (CG / FCG only)
// first_batch: iterations enqueued before the first host check (at least the usual batch); a
// caller that expects the solve to take about as many iterations as its previous one passes one
// fewer than those, which saves that many host round trips
int krylov_solve(mpdg_heat_op *op, mpdg_precond *pc, const mpdg_solver_params &prm, const double *b, double *x,
                 KrylovWork &w, int *iterations, double *res, hipStream_t s, const double *r_init = nullptr,
                 int first_batch = 0) {
  const long n = op->n_dofs();
  MPDG_TRY(w.init(n));
  if (prm.solver_type == MPDG_SOLVER_GMRES) {
    MPDG_TRY(w.scal.resize(1));
    return gmres_solve(op, pc, prm, b, x, w.gmres, w.red, w.scal.get(), iterations, res, s);
  }
  const bool external = pc && pc->type == MPDG_PRECOND_GMG;
  const double *inv = (pc && pc->type == MPDG_PRECOND_DIAGONAL) ? pc->inverse_diagonal() : nullptr;
  const bool jacobi = !external;
  CGState h{};
  h.max_it = prm.max_iterations;
  h.flexible = prm.solver_type == MPDG_SOLVER_FCG ? 1 : 0;
  if (hipMemcpyAsync(w.st.get(), &h, sizeof(CGState), hipMemcpyHostToDevice, s) != hipSuccess) return MPDG_ERR_DEVICE;
  // FP32 multigrid: its input and output are used in place, in FP32
  float *r32 = external ? pc->f32_input() : nullptr;
  const float *z32 = external ? pc->f32_output() : nullptr;
  if (r32 && z32) return fcg32_solve(op, pc, prm, b, x, w, iterations, res, r32, z32, r_init, s, first_batch);
  // r = b - A x
  if (!r_init) MPDG_TRY(op->vmult(w.q.get(), x, s));
  MPDG_TRY(cg_init(b, w.r.get(), r_init ? r_init : w.q.get(), w.z.get(), w.p.get(), inv, jacobi, n, w.red.red(),
                   w.st.get(), prm.rel_tolerance, prm.abs_tolerance, prm.criterion, s, nullptr, r_init != nullptr));
  if (external) {
    MPDG_TRY(pc->vmult(w.z.get(), w.r.get(), s));
    MPDG_TRY(cg_after_external_precond_init(w.r.get(), w.z.get(), w.p.get(), n, w.red.red(), w.st.get(), s));
  }
  // The done flag makes every kernel of a finished solve a no-op, so a batch of
  // iterations can be enqueued before the host reads the state back.
  // large systems: one iteration per host check (an iteration far outweighs the check; a
  // batch would run whole iterations after convergence); small ones: batches hide launch latency
  const int batch = n > 2000000 ? 1 : 8;
  int it = 0, k_batch = std::max(batch, first_batch);
  while (true) {
    for (int k = 0; k < k_batch; ++k) {
      MPDG_TRY(op->vmult(w.q.get(), w.p.get(), s));
      MPDG_TRY(cg_alpha(w.p.get(), w.q.get(), n, w.red.red(), w.st.get(), s));
      MPDG_TRY(cg_update(x, w.r.get(), w.p.get(), w.q.get(), w.z.get(), inv, jacobi, n, w.red.red(), w.st.get(), s));
      if (external) {
        MPDG_TRY(pc->vmult(w.z.get(), w.r.get(), s));
        MPDG_TRY(cg_beta(w.r.get(), w.z.get(), n, w.red.red(), w.st.get(), s));
      }
      MPDG_TRY(cg_direction(w.p.get(), w.z.get(), n, w.st.get(), s));
    }
    it += k_batch;
    k_batch = batch;
    if (hipMemcpyAsync(&h, w.st.get(), sizeof(CGState), hipMemcpyDeviceToHost, s) != hipSuccess) return MPDG_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
    if (h.done || it > prm.max_iterations + batch) break;
  }
  *iterations = h.iters;
  *res = std::sqrt(std::max(h.rr, 0.0));
  if (h.done == 2) return MPDG_ERR_NONFINITE;
  if (h.rr > h.tol2) return MPDG_ERR_NO_CONVERGENCE;
  return MPDG_OK;
}

}  // namespace

// ===========================================================================
// HeatDiffuseOperation (source/heat/heat_diffuse_operation.cpp:215-298)
// ===========================================================================
struct mpdg_heat_operation {
  mpdg_heat_op *op = nullptr;
  mpdg_precond *pc = nullptr;
  mpdg_heat_operation_params prm{};
  static constexpr int kMaxOlder = 5;  // solutions kept before T_{n-1} (extrapolation order <= 6,
                                       // least-squares projection over <= 7 solutions)
  DeviceArray<double> buf[3 + kMaxOlder], rhs, du;
  int cur = 0, old = 1, spare = 2;
  int older[kMaxOlder] = {3, 4, 5, 6, 7};  // buffers of T_{n-2}, T_{n-3}, ...
  int depth = 0;                     // older buffers in use (extrapolation order - 1)
  int n_older = 0;                   // of those, how many hold past solutions
  double dt_old = 0.0;
  double dt_hist[kMaxOlder + 1] = {};  // dt_hist[j] = t_{n-j} - t_{n-j-1} (dt_hist[0] = dt_old)
  double dt_precond = 0.0;
  KrylovWork work;
  ReductionScratch red;
  DeviceArray<double> scal, bvec, zero, alt;
  mpdg_learned *net = nullptr;
  const double *q_field = nullptr;
  double *z_record = nullptr;        // caller's vector for the residual input of each step
  double *base_record = nullptr;     // caller's vector for the base guess of each step
  DeviceArray<double> zbuf, invd;    // residual input and inverse diagonal (learned predictor)
  DeviceArray<double> tbase;         // base guess the network corrected (the fallback)
  DeviceArray<double> rbase;         // its residual b - J T_b (the solve's start when it is kept)
  double *scal_host = nullptr;       // pinned copy of the guess check's two norms, read after the solve
  ~mpdg_heat_operation() {
    if (scal_host) (void)hipHostFree(scal_host);
  }
  bool learned_applied = false;      // this step's guess came from the network
  double invd_dt = 0.0;              // time step the cached inverse diagonal was computed for
  DeviceArray<double> kq_step;       // k(T_n) per quadrature point, frozen for a linear step
  DeviceArray<double> lsq_x, lsq_q, lsq_part, lsq_h;  // least-squares projection: bases and scalars
  bool law_frozen = false;
  int fell_back = 0;            // guess status of the last step (0 used, 1 rejected, 2 paused)
  int guess_rejections = 0;     // consecutive steps whose learned guess the check rejected
  int guess_paused = 0;         // steps left on which the network is not evaluated
  int steps_since_update = 0;  // linear steps since the preconditioner was rebuilt
  int last_lin_iters = 0;       // iterations of the previous linear step's solve
  int n_calls = 0;             // predictor calls since set_initial (Predictor::n_calls, predictor.hpp:97)
  double rel_res0 = 0.0, rel_res = 0.0;
};

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int mpdg_heat_op_create(const mpdg_heat_desc *desc, mpdg_heat_op **out) {
  if (!desc || !out) return MPDG_ERR_INVALID;
  if (desc->degree < 1 || desc->degree > 4) return MPDG_ERR_INVALID;
  for (int d = 0; d < 3; ++d)
    if (desc->n_cells[d] < 1 || !(desc->extent[d] > 0.0)) return MPDG_ERR_INVALID;
  if (desc->precision != MPDG_FP64 && desc->precision != MPDG_FP32) return MPDG_ERR_INVALID;
  auto *op = new (std::nothrow) mpdg_heat_op;
  if (!op) return MPDG_ERR_NOMEM;
  op->desc = *desc;
  op->precision = desc->precision;
  op->d64.set_mesh(desc->degree, desc->n_cells, desc->extent);
  op->d32.set_mesh(desc->degree, desc->n_cells, desc->extent);
  op->d64.dmask = op->d32.dmask = desc->dirichlet_faces & 63;
  op->d64.c_mass = op->d32.c_mass = desc->rho_cp;  // dt = 1 until reset_time_increment
  *out = op;
  return MPDG_OK;
}

int mpdg_heat_op_destroy(mpdg_heat_op *op) {
  delete op;
  return MPDG_OK;
}
long mpdg_heat_op_n_dofs(const mpdg_heat_op *op) { return op ? op->n_dofs() : -1; }
long mpdg_heat_op_n_cells(const mpdg_heat_op *op) { return op ? op->n_cells() : -1; }

int mpdg_heat_op_reset_time_increment(mpdg_heat_op *op, double dt) {
  if (!op) return MPDG_ERR_INVALID;
  if (dt == 0.0) return MPDG_ERR_ZERO_TIME_INCREMENT;
  op->dt = dt;
  op->d64.c_mass = op->d32.c_mass = op->desc.rho_cp / dt;
  return MPDG_OK;
}

int mpdg_heat_op_set_conductivity_constant(mpdg_heat_op *op, double k) {
  if (!op) return MPDG_ERR_INVALID;
  op->d64.kappa_mode = op->d32.kappa_mode = KAPPA_CONST;
  op->d64.kappa = op->d32.kappa = k;
  return MPDG_OK;
}
int mpdg_heat_op_set_conductivity_qp(mpdg_heat_op *op, const void *k) {
  if (!op || !k) return MPDG_ERR_INVALID;
  if (reinterpret_cast<uintptr_t>(k) % 16 != 0) return MPDG_ERR_INVALID;  // read with 16-byte loads
  op->d64.kappa_mode = op->d32.kappa_mode = KAPPA_QP;
  op->d64.kappa_arr = (const double *)k;
  op->d32.kappa_arr = (const float *)k;
  MPDG_TRY(op->refresh_metric_k(nullptr));
  return hipDeviceSynchronize() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int mpdg_heat_op_set_conductivity_cell(mpdg_heat_op *op, const void *k) {
  if (!op || !k) return MPDG_ERR_INVALID;
  op->d64.kappa_mode = op->d32.kappa_mode = KAPPA_CELL;
  op->d64.kappa_arr = (const double *)k;
  op->d32.kappa_arr = (const float *)k;
  return MPDG_OK;
}
int mpdg_heat_op_set_conductivity_law(mpdg_heat_op *op, int law, const double *p, const void *T) {
  if (!op || !p || !T || (law != MPDG_LAW_TANH && law != MPDG_LAW_MUSHY)) return MPDG_ERR_INVALID;
  op->d64.kappa_mode = op->d32.kappa_mode = KAPPA_LAW;
  op->d64.law = op->d32.law = law;
  for (int i = 0; i < 4; ++i) op->d64.law_p[i] = op->d32.law_p[i] = p[i];
  op->d64.T = (const double *)T;
  op->d32.T = (const float *)T;
  return MPDG_OK;
}
int mpdg_heat_op_set_robin(mpdg_heat_op *op, int mask, double alpha, double T_inf_conv, double emissivity,
                           double T_inf_rad, const double *T) {
  if (!op || op->precision != MPDG_FP64) return MPDG_ERR_INVALID;
  if (emissivity > 0.0 && !T) return MPDG_ERR_INVALID;
  // Neumann flux and Robin terms on one face (z+) are refused (heat_diffuse_operator.cpp:739-743)
  if ((mask & 32) && op->flux[0] != 0.0) return MPDG_ERR_INVALID;
  op->d64.robin_mask = mask & 63;
  op->d64.alpha = alpha;
  op->d64.T_inf_conv = T_inf_conv;
  op->d64.emissivity = emissivity;
  op->d64.T_inf_rad = T_inf_rad;
  op->T_robin = T;
  return MPDG_OK;
}
int mpdg_heat_op_set_mapping(mpdg_heat_op *op, const double *coords) {
  if (!op || !coords) return MPDG_ERR_INVALID;
  const int n1 = op->d64.degree + 1;
  const size_t nm = (size_t)op->d64.n_cells * 7 * n1 * n1 * n1;
  op->coords = coords;
  op->d64.coords = op->d32.coords = coords;
  if (op->precision == MPDG_FP64) {
    MPDG_TRY(op->metric64.resize(nm));
    MPDG_TRY(launch_metric<double>(op->d64, coords, op->metric64.get(), nullptr));
    op->d64.metric = op->metric64.get();
  } else {
    MPDG_TRY(op->metric32.resize(nm));
    MPDG_TRY(launch_metric<float>(op->d32, coords, op->metric32.get(), nullptr));
    op->d32.metric = op->metric32.get();
  }
  MPDG_TRY(op->refresh_metric_k(nullptr));
  return hipDeviceSynchronize() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int mpdg_heat_op_set_surface_flux(mpdg_heat_op *op, const double *p) {
  if (!op || !p) return MPDG_ERR_INVALID;
  if (p[0] != 0.0 && (p[1] <= 0.0 || (op->d64.robin_mask & 32))) return MPDG_ERR_INVALID;
  for (int i = 0; i < 6; ++i) op->flux[i] = p[i];
  return MPDG_OK;
}

int mpdg_heat_op_set_tangent(mpdg_heat_op *op, const mpdg_heat_material *mat, int do_solidification,
                             const double *velocity3) {
  if (!op || !mat || !velocity3) return MPDG_ERR_INVALID;
  if (op->precision != MPDG_FP64 || op->coords) return MPDG_ERR_INVALID;
  if (mat->transition != MPDG_TRANSITION_NONE && mat->transition != MPDG_TRANSITION_MUSHY) return MPDG_ERR_INVALID;
  if (mat->transition == MPDG_TRANSITION_MUSHY && !(mat->liquidus > mat->solidus)) return MPDG_ERR_INVALID;
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(velocity3[i])) return MPDG_ERR_INVALID;
  op->mat = *mat;
  op->solid = do_solidification ? 1 : 0;
  for (int i = 0; i < 3; ++i) op->vel[i] = velocity3[i];
  const int n1 = op->d64.degree + 1;
  MPDG_TRY(op->td.resize((size_t)kTangentCoeffs * op->d64.n_cells * n1 * n1 * n1));
  op->tangent = true;
  op->td_bound = false;
  return MPDG_OK;
}

int mpdg_heat_op_set_velocity_field(mpdg_heat_op *op, const double *velocity) {
  if (!op || !op->tangent) return MPDG_ERR_INVALID;
  op->vfield = velocity;
  op->td_bound = false;
  return MPDG_OK;
}

int mpdg_heat_op_clear_tangent(mpdg_heat_op *op) {
  if (!op) return MPDG_ERR_INVALID;
  op->tangent = false;
  op->vfield = nullptr;
  op->td_bound = false;
  op->td.release();
  return MPDG_OK;
}

int mpdg_heat_op_bind_temperature(mpdg_heat_op *op, const double *T, const double *T_old, void *stream) {
  if (!op || !T || !op->tangent) return MPDG_ERR_INVALID;
  if (op->dt == 0.0) return MPDG_ERR_ZERO_TIME_INCREMENT;
  MPDG_TRY(launch_tangent_data(op->d64, op->dt, op->mat, op->solid, op->vel, op->vfield, T, T_old ? T_old : T, op->td.get(),
                               S(stream)));
  op->td_bound = true;
  return MPDG_OK;
}

int mpdg_heat_op_set_timing(mpdg_heat_op *op, int enable) {
  if (!op) return MPDG_ERR_INVALID;
  op->timing_on = enable != 0;
  op->timing_used = 0;
  return MPDG_OK;
}
int mpdg_heat_op_get_timing(mpdg_heat_op *op, double *total_ms, int *count) {
  if (!op || !total_ms || !count) return MPDG_ERR_INVALID;
  double t = 0.0;
  for (size_t i = 0; i < op->timing_used; ++i) {
    if (hipEventSynchronize(op->timing_events[i].second) != hipSuccess) return MPDG_ERR_DEVICE;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, op->timing_events[i].first, op->timing_events[i].second) != hipSuccess)
      return MPDG_ERR_DEVICE;
    t += ms;
  }
  *total_ms = t;
  *count = (int)op->timing_used;
  op->timing_used = 0;
  return MPDG_OK;
}

int mpdg_heat_op_vmult(const mpdg_heat_op *op, double *dst, const double *src, void *stream) {
  if (!op || !dst || !src) return MPDG_ERR_INVALID;
  return op->vmult(dst, src, S(stream));
}
// ---------------------------------------------------------------------------
// vmult without allocation (mpdg_heat_op_vmult_noalloc)
// ---------------------------------------------------------------------------
// The apply's only scratch is the edge buffer (rows where four columns meet). Here it lives
// in a fixed device array of kNoAllocScratchDoubles doubles (8 MiB), one copy per GPU, held
// until the process ends. Every call on a GPU uses that GPU's copy: calls that overlap on one
// GPU (other streams, other threads) race.
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxScratchDevices] = {};

// device address of the current device's copy of `scratch`, looked up once per device with
// that device current (the array's name is a host-side stand-in, not a device address)
int scratch_address(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return MPDG_ERR_DEVICE;
  if (dev < 0 || dev >= kMaxScratchDevices) return MPDG_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return MPDG_ERR_DEVICE;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return MPDG_OK;
}
}  // namespace

int mpdg_heat_op_vmult_noalloc(const mpdg_heat_op *op, double *dst, const double *src, void *stream) {
  if (!op || !dst || !src || op->precision != MPDG_FP64) return MPDG_ERR_INVALID;
  hipStream_t s = S(stream);
  if (op->tangent) {
    // the full tangent's kernels use no scratch
    if (!op->td_bound) return MPDG_ERR_INVALID;
    MPDG_TRY(launch_tangent_apply(op->d64, op->td.get(), tangent_convects(op->vel, op->vfield), op->solid, src, dst, s));
  } else {
    if (op->d64.metric && op->d64.kappa_mode == KAPPA_LAW) {
      // k(T) on a mapped mesh rescales the metric into operator-owned arrays: only when they
      // already exist at their size (an earlier vmult made them), so that nothing is allocated
      const int n1 = op->d64.degree + 1;
      const size_t nq = (size_t)op->d64.n_cells * n1 * n1 * n1;
      if (op->kq.n != nq || op->metric64_k.n != nq * 7) return MPDG_ERR_INVALID;
      MPDG_TRY(op->refresh_metric_k(s));
    }
    const size_t need = apply_edge_elems(op->d64.degree, op->d64.nc, false, op->d64.metric ? 1 : 0);
    if (need > kNoAllocScratchDoubles) return MPDG_ERR_TOO_LARGE;
    double *scr = nullptr;
    MPDG_TRY(scratch_address(&scr));
    OperatorData<double> od = op->d64;
    od.edge = scr;
    od.edge_elems = kNoAllocScratchDoubles;
    // the previous call's contents stay in the array: clear the part this apply uses (a fill
    // kernel, which a captured graph replays correctly)
    if (need) MPDG_TRY(launch_fill<double>(scr, 0.0, (long)need, s));
    MPDG_TRY(launch_apply<double>(od, src, dst, s));
  }
  if (op->d64.robin_mask) MPDG_TRY(launch_boundary(op->d64, 0, nullptr, dst, src, op->T_robin, s));
  return MPDG_OK;
}

int mpdg_heat_op_vmult_f32(const mpdg_heat_op *op, float *dst, const float *src, void *stream) {
  if (!op || !dst || !src) return MPDG_ERR_INVALID;
  return op->vmult_f32(dst, src, S(stream));
}
int mpdg_heat_op_compute_inverse_diagonal(const mpdg_heat_op *op, double *inv, void *stream) {
  if (!op || !inv || op->precision != MPDG_FP64) return MPDG_ERR_INVALID;
  return op->inverse_diagonal(inv, S(stream));
}
int mpdg_heat_op_compute_diagonal(const mpdg_heat_op *op, void *diag, void *stream) {
  if (!op || !diag) return MPDG_ERR_INVALID;
  if (op->precision == MPDG_FP64) return op->diagonal64((double *)diag, S(stream));
  const int n1 = op->d32.degree + 1;
  DeviceArray<float> scratch;
  MPDG_TRY(scratch.resize(2 * (size_t)op->d32.n_cells * n1 * n1 * n1));
  MPDG_TRY(launch_diagonal<float>(op->d32, (float *)diag, scratch.get(), S(stream)));
  return hipStreamSynchronize(S(stream)) == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int mpdg_heat_op_create_rhs(const mpdg_heat_op *op, double *dst, const double *T, const double *T_old,
                            const double *q, void *stream) {
  if (!op || !dst || !T || !T_old) return MPDG_ERR_INVALID;
  return op->create_rhs(dst, T, T_old, q, S(stream));
}
int mpdg_heat_op_l2_norm(const mpdg_heat_op *op, const double *u, double *norm, void *stream) {
  if (!op || !u || !norm) return MPDG_ERR_INVALID;
  return op->l2_norm(u, norm, S(stream));
}

int mpdg_precond_create(mpdg_heat_op *op, int type, const mpdg_gmg_params *gmg, mpdg_precond **out) {
  if (!op || !out || op->precision != MPDG_FP64) return MPDG_ERR_INVALID;
  mpdg_precond *p = nullptr;
  if (type == MPDG_PRECOND_IDENTITY) p = new (std::nothrow) IdentityPrecond;
  else if (type == MPDG_PRECOND_DIAGONAL) p = new (std::nothrow) DiagonalPrecond;
  else if (type == MPDG_PRECOND_GMG) {
    if (op->tangent) return MPDG_ERR_INVALID;  // the level operators are the symmetric ones
    if (!gmg || gmg->smoother_degree < 1 || gmg->smoothing_range <= 1.0 || gmg->eig_cg_n_iterations < 1 ||
        gmg->pre_smoother_degree < 0 || gmg->pre_smoother_degree > gmg->smoother_degree)
      return MPDG_ERR_INVALID;
    if (gmg->precision == MPDG_FP32) {
      auto *g = new (std::nothrow) GMGPrecond<float>;
      if (!g) return MPDG_ERR_NOMEM;
      g->prm = *gmg;
      g->op = op;
      g->type = type;
      const int rc = g->build();
      if (rc) { delete g; return rc; }
      p = g;
    } else if (gmg->precision == MPDG_FP64) {
      auto *g = new (std::nothrow) GMGPrecond<double>;
      if (!g) return MPDG_ERR_NOMEM;
      g->prm = *gmg;
      g->op = op;
      g->type = type;
      const int rc = g->build();
      if (rc) { delete g; return rc; }
      p = g;
    } else {
      return MPDG_ERR_INVALID;
    }
  } else {
    return MPDG_ERR_INVALID;
  }
  if (!p) return MPDG_ERR_NOMEM;
  p->op = op;
  p->type = type;
  *out = p;
  return MPDG_OK;
}
int mpdg_precond_update(mpdg_precond *p, void *stream) {
  if (!p) return MPDG_ERR_INVALID;
  MPDG_TRY(p->update(S(stream)));
  return hipStreamSynchronize(S(stream)) == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
int mpdg_precond_vmult(mpdg_precond *p, double *dst, const double *src, void *stream) {
  if (!p || !dst || !src) return MPDG_ERR_INVALID;
  return p->vmult(dst, src, S(stream));
}
int mpdg_precond_destroy(mpdg_precond *p) {
  delete p;
  return MPDG_OK;
}
int mpdg_gmg_n_levels(const mpdg_precond *p) { return p ? p->n_levels() : 0; }
int mpdg_gmg_level_info(const mpdg_precond *p, int level, int *degree, int *nc, double *lambda) {
  if (!p || !degree || !nc || !lambda) return MPDG_ERR_INVALID;
  return p->level_info(level, degree, nc, lambda);
}

static bool valid_solver(const mpdg_solver_params &p) {
  if (p.solver_type != MPDG_SOLVER_CG && p.solver_type != MPDG_SOLVER_FCG && p.solver_type != MPDG_SOLVER_GMRES)
    return false;
  return p.gmres_restart >= 0 && p.gmres_restart < kMdotMax;
}

int mpdg_linear_solve(mpdg_heat_op *op, mpdg_precond *p, const mpdg_solver_params *prm, const double *b, double *x,
                      int *iterations, double *res, void *stream) {
  if (!op || !prm || !b || !x || !iterations || !res || op->precision != MPDG_FP64) return MPDG_ERR_INVALID;
  if (!valid_solver(*prm)) return MPDG_ERR_INVALID;
  KrylovWork w;
  return krylov_solve(op, p, *prm, b, x, w, iterations, res, S(stream));
}

int mpdg_heat_operation_create(mpdg_heat_op *op, mpdg_precond *p, const mpdg_heat_operation_params *prm,
                               mpdg_heat_operation **out) {
  if (!op || !prm || !out || op->precision != MPDG_FP64) return MPDG_ERR_INVALID;
  if (prm->predictor < MPDG_PREDICTOR_NONE || prm->predictor > MPDG_PREDICTOR_LSQ_SOLUTION) return MPDG_ERR_INVALID;
  const bool lsq = prm->predictor == MPDG_PREDICTOR_LSQ || prm->predictor == MPDG_PREDICTOR_LSQ_SOLUTION;
  if (lsq &&
      (prm->n_old_solution_vectors < 1 || prm->n_old_solution_vectors > mpdg_heat_operation::kMaxOlder + 2))
    return MPDG_ERR_INVALID;
  if (prm->step_mode != MPDG_STEP_NEWTON && prm->step_mode != MPDG_STEP_LINEAR) return MPDG_ERR_INVALID;
  if (!valid_solver(prm->linear)) return MPDG_ERR_INVALID;
  if (prm->guess_pause_after < 0 || prm->guess_pause_steps < 0) return MPDG_ERR_INVALID;
  if ((prm->predictor == MPDG_PREDICTOR_EXTRAPOLATION || prm->predictor == MPDG_PREDICTOR_LEARNED) &&
      (prm->extrapolation_order < (prm->predictor == MPDG_PREDICTOR_LEARNED ? 0 : 1) ||
       prm->extrapolation_order > mpdg_heat_operation::kMaxOlder + 1))
    return MPDG_ERR_INVALID;
  auto *h = new (std::nothrow) mpdg_heat_operation;
  if (!h) return MPDG_ERR_NOMEM;
  h->op = op;
  h->pc = p;
  h->prm = *prm;
  const long n = op->n_dofs();
  int rc = MPDG_OK;
  // the learned predictor corrects the extrapolation of order extrapolation_order (0 or 1: linear)
  if (prm->predictor == MPDG_PREDICTOR_LEARNED) h->prm.extrapolation_order = std::max(1, prm->extrapolation_order);
  const bool extrap = prm->predictor == MPDG_PREDICTOR_EXTRAPOLATION || prm->predictor == MPDG_PREDICTOR_LEARNED;
  h->depth = extrap ? h->prm.extrapolation_order - 1
             : lsq  ? std::max(0, prm->n_old_solution_vectors - 2)
                    : 0;
  for (int b = 0; b < 3 + h->depth; ++b) rc = rc ? rc : h->buf[b].resize(n);
  rc = rc ? rc : h->rhs.resize(n);
  rc = rc ? rc : h->du.resize(n);
  rc = rc ? rc : h->red.init();
  rc = rc ? rc : h->scal.resize(4);
  rc = rc ? rc : h->bvec.resize(n);
  rc = rc ? rc : h->zero.resize(n);
  rc = rc ? rc : h->alt.resize(n);
  if (!rc) rc = hipMemset(h->zero.get(), 0, n * sizeof(double)) == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
  if (rc) {
    delete h;
    return rc;
  }
  *out = h;
  return MPDG_OK;
}
int mpdg_heat_operation_destroy(mpdg_heat_operation *h) {
  delete h;
  return MPDG_OK;
}
int mpdg_heat_operation_set_initial(mpdg_heat_operation *h, const double *T0, void *stream) {
  if (!h || !T0) return MPDG_ERR_INVALID;
  const long n = h->op->n_dofs();
  MPDG_TRY(launch_convert<double, double>(T0, h->buf[h->cur].get(), n, S(stream)));
  MPDG_TRY(launch_convert<double, double>(T0, h->buf[h->old].get(), n, S(stream)));
  h->dt_old = 0.0;
  h->n_older = 0;
  h->n_calls = 0;
  h->guess_rejections = 0;
  h->guess_paused = 0;
  h->last_lin_iters = 0;
  return hipStreamSynchronize(S(stream)) == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

// Residual input of the learned guess for the predicted T (the base guess, an extrapolation):
// z = D^-1 (b - J T) with b given (linear step), or z = D^-1 r_Newton(T, T_old) (b == nullptr).
// rr (device scalar, may be nullptr): |b - J T|^2 before the scaling.
static int guess_feature(mpdg_heat_operation *h, const double *b, const double *T, const double *T_old, double *z,
                         double *rr, hipStream_t s, double *raw = nullptr) {
  mpdg_heat_op *op = h->op;
  const long n = op->n_dofs();
  // the inverse diagonal is kept while dt is unchanged and the conductivity does not follow T
  const bool t_dependent = op->d64.kappa_mode == KAPPA_LAW || h->law_frozen || op->d64.robin_mask != 0;
  if (h->invd.n != (size_t)n || t_dependent || h->invd_dt != op->dt) {
    MPDG_TRY(h->invd.resize(n));
    MPDG_TRY(op->inverse_diagonal(h->invd.get(), s));
    h->invd_dt = t_dependent ? 0.0 : op->dt;
  }
  if (b && rr) {
    // residual, its norm and the diagonal scaling in one pass
    MPDG_TRY(op->vmult(z, T, s));
    return launch_residual_norm(z, b, h->invd.get(), n, h->red.red(), rr, s, raw);
  }
  if (b) {
    MPDG_TRY(op->vmult(z, T, s));
    MPDG_TRY(launch_axpby(z, 1.0, b, -1.0, n, s));
  } else {
    MPDG_TRY(launch_fill<double>(z, 0.0, n, s));
    MPDG_TRY(op->create_rhs(z, T, T_old, nullptr, s));
  }
  if (rr) MPDG_TRY(launch_dot<double, double>(z, z, n, h->red.red(), rr, s));
  return launch_mul(z, h->invd.get(), n, s);
}

// Learned predictor after the base guess T (the extrapolation) is in place: records the
// residual input (and the base) and, with equal steps and a network, replaces T by the
// network's correction of it; the base stays in tbase as the fallback, and the base's squared
// residual norm |b - J T|^2 in scal[1] for the guess check (linear step). While the predictor
// is paused (after repeated rejections by the check) the base guess stays and the network is
// not evaluated.
static int learned_guess(mpdg_heat_operation *h, double dt, const double *b, double *T, const double *T_old,
                         hipStream_t s) {
  bool learned = h->prm.predictor == MPDG_PREDICTOR_LEARNED;
  h->learned_applied = false;
  if (learned && b && h->prm.guess_check && h->guess_paused > 0) {
    --h->guess_paused;
    h->fell_back = 2;
    learned = false;
  }
  if (!learned && !h->z_record) return MPDG_OK;
  const long n = h->op->n_dofs();
  double *z = h->z_record;
  if (!z) {
    MPDG_TRY(h->zbuf.resize(n));
    z = h->zbuf.get();
  }
  if (h->base_record) MPDG_TRY(launch_convert<double, double>(T, h->base_record, n, s));
  const bool check = learned && b && h->prm.guess_check;
  if (check) {  // allocated with the network (no-ops then)
    MPDG_TRY(h->rbase.resize(n));
    MPDG_TRY(h->tbase.resize(n));
  }
  MPDG_TRY(guess_feature(h, b, T, T_old, z, check ? h->scal.get() + 1 : nullptr, s, check ? h->rbase.get() : nullptr));
  // the network extrapolates with equal steps (as in its training trajectories)
  const bool equal_steps = h->dt_old > 0.0 && std::fabs(dt - h->dt_old) <= 1e-12 * dt;
  if (learned && equal_steps) {
    const double *Tnm1 = h->buf[h->depth > 0 ? h->older[0] : h->spare].get();  // T_{n-1} of this step
    MPDG_TRY(h->tbase.resize(n));
    MPDG_TRY(launch_convert<double, double>(T, h->tbase.get(), n, s));
    MPDG_TRY(mpdg_learned_predict(h->net, Tnm1, T_old, h->q_field, z, h->tbase.get(), T, s));
    h->learned_applied = true;
  }
  return MPDG_OK;
}

static void commit_time_increment(mpdg_heat_operation *h, double dt) {
  for (int j = mpdg_heat_operation::kMaxOlder; j > 0; --j) h->dt_hist[j] = h->dt_hist[j - 1];
  h->dt_hist[0] = dt;
  h->dt_old = dt;
}

// A conductivity law k(T) frozen at T_n for one linear step: evaluated once per quadrature
// point and applied as per-point data, so the step's many applies read k instead of
// interpolating T and evaluating the law at every point of every apply. Restores the law on
// destruction (every return path of the step).
struct FrozenLaw {
  mpdg_heat_operation *h;
  int mode = -1;
  const double *arr = nullptr;
  FrozenLaw(mpdg_heat_operation *h_) : h(h_) {}
  int freeze(hipStream_t s) {
    OperatorData<double> &od = h->op->d64;
    if (od.kappa_mode != KAPPA_LAW || od.metric) return MPDG_OK;  // mapped cells fold k into the metric
    MPDG_TRY(h->kq_step.resize((size_t)od.n_cells * (od.degree + 1) * (od.degree + 1) * (od.degree + 1)));
    MPDG_TRY(launch_kappa_qp<double>(od, h->kq_step.get(), s));
    mode = od.kappa_mode;
    arr = od.kappa_arr;
    od.kappa_mode = KAPPA_QP;
    od.kappa_arr = h->kq_step.get();
    h->law_frozen = true;
    return MPDG_OK;
  }
  ~FrozenLaw() {
    if (mode < 0) return;
    h->op->d64.kappa_mode = mode;
    h->op->d64.kappa_arr = arr;
    h->law_frozen = false;
  }
};

// One linear implicit step with the coefficient frozen at T_n:
// J(T_n) T_{n+1} = b,  b = M (c_m T_n + q) + surface flux, from the predicted T.
static int linear_step(mpdg_heat_operation *h, double dt, double *T, const double *T_old, int *lin_iters,
                       int *nl_iters, hipStream_t s) {
  mpdg_heat_op *op = h->op;
  const long n = op->n_dofs();
  if (op->d64.emissivity > 0.0 || op->tangent) return MPDG_ERR_INVALID;
  if (op->d64.kappa_mode == KAPPA_LAW) op->d64.T = T_old;
  if (op->d64.robin_mask) op->T_robin = T_old;
  const bool coeff_changes = op->d64.kappa_mode == KAPPA_LAW;
  FrozenLaw frozen(h);
  MPDG_TRY(frozen.freeze(s));
  // preconditioner.update() (heat_diffuse_operation.cpp:275); with a conductivity law it may be
  // kept for a few steps (any fixed SPD preconditioner keeps CG exact; the FP64 test decides)
  const int every = std::max(1, h->prm.preconditioner_update_interval);
  if (h->pc && (dt != h->dt_precond || (coeff_changes && h->steps_since_update + 1 >= every))) {
    MPDG_TRY(h->pc->update(s));
    h->dt_precond = dt;
    h->steps_since_update = 0;
  } else {
    ++h->steps_since_update;
  }
  MPDG_TRY(launch_fill<double>(h->bvec.get(), 0.0, n, s));
  MPDG_TRY(op->create_rhs(h->bvec.get(), h->zero.get(), T_old, nullptr, s, true));
  MPDG_TRY(learned_guess(h, dt, h->bvec.get(), T, T_old, s));
  const double *r_init = nullptr;
  if (h->learned_applied && h->prm.guess_check) {
    // the learned guess (T) against the base guess it corrects (its residual norm came with the
    // network's input); the residual of T is kept in alt. With guess_smoothing the network's
    // guess is first smoothed (the correction is prolonged from its working grid and carries
    // high-frequency error, which the residual and CG weigh most), then checked.
    MPDG_TRY(op->vmult(h->alt.get(), T, s));
    if (h->prm.guess_smoothing > 0 && h->pc) {
      MPDG_TRY(launch_axpby(h->alt.get(), 1.0, h->bvec.get(), -1.0, n, s));
      MPDG_TRY(h->pc->smooth(h->du.get(), h->alt.get(), h->prm.guess_smoothing, s));
      MPDG_TRY(launch_axpby(T, 1.0, h->du.get(), 1.0, n, s));
      MPDG_TRY(op->vmult(h->alt.get(), T, s));
    }
    // alt = b - J T and its squared norm; on the device, the guess and its residual or the base
    // guess and its residual (kept from the network's input) become the solve's start, so the
    // host does not wait for the decision
    MPDG_TRY(launch_residual_norm(h->alt.get(), h->bvec.get(), nullptr, n, h->red.red(), h->scal.get(), s));
    MPDG_TRY(launch_guess_select(h->scal.get(), T, h->tbase.get(), h->alt.get(), h->rbase.get(), n, s));
    if (!h->scal_host) return MPDG_ERR_INVALID;  // allocated with the network (set_learned)
    if (hipMemcpyAsync(h->scal_host, h->scal.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
      return MPDG_ERR_DEVICE;
    r_init = h->alt.get();
  }
  mpdg_solver_params lp = h->prm.linear;
  lp.criterion = MPDG_CRIT_RHS;
  int li = 0;
  double res = 0.0;
  // consecutive linear steps take similar iteration counts: all but one of the previous step's
  // iterations are enqueued before the first host check
  const int rc = krylov_solve(op, h->pc, lp, h->bvec.get(), T, h->work, &li, &res, s,
                              lp.solver_type == MPDG_SOLVER_GMRES ? nullptr : r_init, h->last_lin_iters - 1);
  h->last_lin_iters = li;
  CGState st{};
  if (hipMemcpyAsync(&st, h->work.st.get(), sizeof(CGState), hipMemcpyDeviceToHost, s) != hipSuccess)
    return MPDG_ERR_DEVICE;
  if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
  if (r_init) {
    // the check's outcome (copied before the solve): a run of rejections pauses the network (its
    // evaluation and check cost about three operator applies and a smoothing per step for
    // nothing), then it is tried again
    const double g = h->scal_host[0], base = h->scal_host[1];
    if (!(g == g) || !(g < base)) {
      h->fell_back = 1;
      if (h->prm.guess_pause_after > 0 && ++h->guess_rejections >= h->prm.guess_pause_after) {
        h->guess_paused = h->prm.guess_pause_steps;
        h->guess_rejections = 0;
      }
    } else {
      h->guess_rejections = 0;
    }
  }
  h->rel_res0 = std::sqrt(st.rr0 / st.bb);
  h->rel_res = std::sqrt(st.rr / st.bb);
  commit_time_increment(h, dt);
  if (lin_iters) *lin_iters = li;
  if (nl_iters) *nl_iters = 1;
  return rc;
}

// Polynomial extrapolation of order p through T_n, T_{n-1}, ..., T_{n-p} at their times to
// t_{n+1} = t_n + dt (Lagrange weights; p = 1 is compute_linear_predictor,
// source/linear_algebra/predictor_linear.cpp:76-97). The order drops to what the history holds.
static int extrapolate(mpdg_heat_operation *h, double dt, double *pred, hipStream_t s) {
  const long n = h->op->n_dofs();
  const double *pts[mpdg_heat_operation::kMaxOlder + 2];
  double tau[mpdg_heat_operation::kMaxOlder + 2];
  int m = 1;
  pts[0] = h->buf[h->cur].get();
  tau[0] = 0.0;
  if (h->dt_old > 0.0) {
    pts[m] = h->buf[h->old].get();
    tau[m] = -h->dt_hist[0];
    ++m;
    for (int j = 0; j < h->n_older && m <= h->prm.extrapolation_order; ++j, ++m) {
      pts[m] = h->buf[h->older[j]].get();
      tau[m] = tau[m - 1] - h->dt_hist[j + 1];
    }
  }
  m = std::min(m, h->prm.extrapolation_order + 1);
  double w[mpdg_heat_operation::kMaxOlder + 2];
  for (int j = 0; j < m; ++j) {
    w[j] = 1.0;
    for (int i = 0; i < m; ++i)
      if (i != j) w[j] *= (dt - tau[i]) / (tau[j] - tau[i]);
  }
  return launch_combine(pred, pts, w, m, n, s);
}

// MPDG_PREDICTOR_LSQ_SOLUTION, a variant of LeastSquaresProjection::vmult (predictor.hpp:36-80)
// with the step's right-hand side: x = sum_s c_s T_{n-s} (s < L) minimising ||b - J x||. The
// reference orthogonalises the images J T_{n-s} by modified Gram-Schmidt; here
// the basis is first changed to backward differences (T_n, T_n - T_{n-1}, second difference,
// ...), whose images differ by orders of magnitude instead of being nearly parallel, and the
// images are orthonormalised by classical Gram-Schmidt applied twice (fused multi-dot /
// multi-axpy) with the same operations on the solution basis; a direction whose image
// vanishes against the previous ones is dropped (the reference stops there, :58-59).
static int lsq_project(mpdg_heat_operation *h, const double *b, double *pred, hipStream_t s) {
  mpdg_heat_op *op = h->op;
  const long n = op->n_dofs();
  const double *pts[mpdg_heat_operation::kMaxOlder + 2];
  int L = 0;
  pts[L++] = h->buf[h->cur].get();
  if (h->dt_old > 0.0) {
    pts[L++] = h->buf[h->old].get();
    for (int j = 0; j < h->n_older; ++j) pts[L++] = h->buf[h->older[j]].get();
  }
  if (L < h->prm.n_old_solution_vectors) return launch_convert<double, double>(pts[0], pred, n, s);
  L = h->prm.n_old_solution_vectors;
  MPDG_TRY(h->lsq_x.resize((size_t)L * n));
  MPDG_TRY(h->lsq_q.resize((size_t)L * n));
  MPDG_TRY(h->lsq_part.resize(mdot_scratch()));
  MPDG_TRY(h->lsq_h.resize(3 * kMdotMax));
  double *X = h->lsq_x.get(), *Q = h->lsq_q.get(), *hd = h->lsq_h.get();
  // backward differences: D_j = sum_i (-1)^i C(j, i) T_{n-i}
  for (int j = 0; j < L; ++j) {
    double *D = X + (size_t)j * n;
    double c = 1.0;
    for (int i = 0; i <= j; ++i) {
      if (i == 0) MPDG_TRY(launch_scale(D, pts[0], 1.0, n, s));
      else MPDG_TRY(launch_axpby(D, c, pts[i], 1.0, n, s));
      c = -c * (double)(j - i) / (double)(i + 1);
    }
  }
  int m = 0;  // accepted directions
  for (int j = 0; j < L; ++j) {
    double *xj = X + (size_t)m * n, *qj = Q + (size_t)m * n;
    if (m != j) MPDG_TRY(launch_scale(xj, X + (size_t)j * n, 1.0, n, s));
    MPDG_TRY(op->vmult(qj, xj, s));
    double nrm0 = 0.0, hh[2 * kMdotMax + 1];
    MPDG_TRY(launch_dot<double, double>(qj, qj, n, h->red.red(), hd + 2 * kMdotMax, s));
    for (int pass = 0; pass < 2 && m > 0; ++pass) {
      MPDG_TRY(launch_mdot(Q, n, m, qj, h->lsq_part.get(), hd + pass * kMdotMax, false, s));
      MPDG_TRY(launch_maxpy(qj, Q, n, m, hd + pass * kMdotMax, -1.0, s));
      MPDG_TRY(launch_maxpy(xj, X, n, m, hd + pass * kMdotMax, -1.0, s));
    }
    if (hipMemcpyAsync(&nrm0, hd + 2 * kMdotMax, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
      return MPDG_ERR_DEVICE;
    MPDG_TRY(launch_dot<double, double>(qj, qj, n, h->red.red(), hd + 2 * kMdotMax, s));
    if (hipMemcpyAsync(hh, hd + 2 * kMdotMax, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
      return MPDG_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
    const double nrm = std::sqrt(std::max(hh[0], 0.0));
    if (!(nrm == nrm)) return MPDG_ERR_NONFINITE;
    if (!(nrm > 1e-12 * std::sqrt(std::max(nrm0, 0.0)))) continue;  // dependent on the earlier ones
    MPDG_TRY(launch_axpby(qj, 0.0, qj, 1.0 / nrm, n, s));
    MPDG_TRY(launch_axpby(xj, 0.0, xj, 1.0 / nrm, n, s));
    ++m;
  }
  if (m == 0) return launch_convert<double, double>(pts[0], pred, n, s);
  // x = sum_j (q_j . b) x_j
  MPDG_TRY(launch_mdot(Q, n, m, b, h->lsq_part.get(), hd, false, s));
  MPDG_TRY(launch_fill<double>(pred, 0.0, n, s));
  return launch_maxpy(pred, X, n, m, hd, 1.0, s);
}

// LeastSquaresProjection::vmult as the reference computes it (predictor.hpp:36-80): images
// v_s = J T_{n-s} orthogonalised in order by modified Gram-Schmidt without normalisation
// (P(j, s) = v_s . v_j after the first j projections, P(s, s) = |v_s|^2), stopped at the first
// image with P(s, s) < 1e-16 P(0, 0); coefficients by back substitution of
// P(s, s) c_s = v_s . rhs - sum_{j > s} P(s, j) c_j; x = sum_s c_s T_{n-s}. Until the predictor
// has been called n_old_solution_vectors times the previous solution is used (:147-151).
static int lsq_reference(mpdg_heat_operation *h, const double *rhs, double *pred, hipStream_t s) {
  mpdg_heat_op *op = h->op;
  const long n = op->n_dofs();
  const int L = h->prm.n_old_solution_vectors;
  const double *pts[mpdg_heat_operation::kMaxOlder + 2];
  int avail = 0;
  pts[avail++] = h->buf[h->cur].get();
  if (h->dt_old > 0.0) {
    pts[avail++] = h->buf[h->old].get();
    for (int j = 0; j < h->n_older; ++j) pts[avail++] = h->buf[h->older[j]].get();
  }
  if (h->n_calls < L || avail < L) return launch_convert<double, double>(pts[0], pred, n, s);
  MPDG_TRY(h->lsq_q.resize((size_t)L * n));
  MPDG_TRY(h->lsq_h.resize(4));
  double *V = h->lsq_q.get(), *hd = h->lsq_h.get();
  auto dot = [&](const double *a, const double *b, double *out) -> int {
    MPDG_TRY(launch_dot<double, double>(a, b, n, h->red.red(), hd, s));
    if (hipMemcpyAsync(out, hd, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return MPDG_ERR_DEVICE;
    return hipStreamSynchronize(s) == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
  };
  double P[mpdg_heat_operation::kMaxOlder + 2][mpdg_heat_operation::kMaxOlder + 2] = {};
  int step = 0;
  for (; step < L; ++step) {
    double *vs = V + (size_t)step * n;
    MPDG_TRY(op->vmult(vs, pts[step], s));
    MPDG_TRY(dot(vs, V, &P[0][step]));
    for (int j = 0; j < step; ++j) {
      MPDG_TRY(launch_axpby(vs, -P[j][step] / P[j][j], V + (size_t)j * n, 1.0, n, s));
      MPDG_TRY(dot(vs, V + (size_t)(j + 1) * n, &P[j + 1][step]));
    }
    if (!(P[step][step] == P[step][step])) return MPDG_ERR_NONFINITE;
    if (P[step][step] < 1e-16 * P[0][0]) break;
  }
  if (step == 0) return launch_convert<double, double>(pts[0], pred, n, s);
  double c[mpdg_heat_operation::kMaxOlder + 2];
  for (int k = 0; k < step; ++k) MPDG_TRY(dot(V + (size_t)k * n, rhs, &c[k]));
  for (int k = step - 1; k >= 0; --k) {
    double sum = c[k];
    for (int j = k + 1; j < step; ++j) sum -= c[j] * P[k][j];
    c[k] = sum / P[k][k];
  }
  return launch_combine(pred, pts, c, step, n, s);
}

// HeatDiffuseOperation::init_time_advance + solve (heat_diffuse_operation.cpp:215-289)
int mpdg_heat_operation_solve_step(mpdg_heat_operation *h, double dt, int *lin_iters, int *nl_iters, void *stream) {
  if (!h) return MPDG_ERR_INVALID;
  hipStream_t s = S(stream);
  mpdg_heat_op *op = h->op;
  const long n = op->n_dofs();
  MPDG_TRY(mpdg_heat_op_reset_time_increment(op, dt));
  // Predictor::vmult (predictor.hpp:118-160): pred from T_n (cur) and T_{n-1} (old)
  double *pred = h->buf[h->spare].get();
  const double *Tn = h->buf[h->cur].get(), *Tnm1 = h->buf[h->old].get();
  // the predictor's operator and residual see the conductivity law and the radiation
  // linearisation at T_n (the step rebinds them to its own iterate afterwards)
  if (op->d64.kappa_mode == KAPPA_LAW) op->d64.T = Tn;
  if (op->d64.robin_mask) op->T_robin = Tn;
  switch (h->prm.predictor) {
    case MPDG_PREDICTOR_NONE: MPDG_TRY(launch_convert<double, double>(Tn, pred, n, s)); break;
    case MPDG_PREDICTOR_ZERO: MPDG_TRY(launch_fill<double>(pred, 0.0, n, s)); break;
    case MPDG_PREDICTOR_LEARNED:
      // the network corrects the extrapolation of order extrapolation_order once b is known
      if (!h->net) return MPDG_ERR_INVALID;
      if (h->prm.extrapolation_order > 1) {
        MPDG_TRY(extrapolate(h, dt, pred, s));
        break;
      }
      [[fallthrough]];
    case MPDG_PREDICTOR_LINEAR: {
      // compute_linear_predictor (source/linear_algebra/predictor_linear.cpp:76-97)
      if (std::fabs(h->dt_old) < 1e-12) MPDG_TRY(launch_convert<double, double>(Tn, pred, n, s));
      else {
        const double f = dt / h->dt_old;
        MPDG_TRY(launch_convert<double, double>(Tn, pred, n, s));
        MPDG_TRY(launch_axpby(pred, -f, Tnm1, 1.0 + f, n, s));
      }
      break;
    }
    case MPDG_PREDICTOR_EXTRAPOLATION: MPDG_TRY(extrapolate(h, dt, pred, s)); break;
    case MPDG_PREDICTOR_LSQ: {
      // predictor rhs (heat_diffuse_operation.cpp:233-244): user_rhs (zero) + create_rhs with the
      // current solution as T and as T_old
      MPDG_TRY(launch_fill<double>(h->bvec.get(), 0.0, n, s));
      MPDG_TRY(op->create_rhs(h->bvec.get(), Tn, Tn, nullptr, s));
      MPDG_TRY(lsq_reference(h, h->bvec.get(), pred, s));
      break;
    }
    case MPDG_PREDICTOR_LSQ_SOLUTION: {
      // the step's b with T_old = T_n, for which the projection is a solution guess
      MPDG_TRY(launch_fill<double>(h->bvec.get(), 0.0, n, s));
      MPDG_TRY(op->create_rhs(h->bvec.get(), h->zero.get(), Tn, nullptr, s, true));
      MPDG_TRY(lsq_project(h, h->bvec.get(), pred, s));
      break;
    }
    default: return MPDG_ERR_INVALID;
  }
  // commit_old_solutions + swap: old <- T_n, cur <- prediction (older <- T_{n-1}, ...)
  int new_spare = h->old;
  if (h->depth > 0) {
    new_spare = h->older[h->depth - 1];
    for (int j = h->depth - 1; j > 0; --j) h->older[j] = h->older[j - 1];
    h->older[0] = h->old;
    h->n_older = h->dt_old > 0.0 ? std::min(h->n_older + 1, h->depth) : 0;
  }
  ++h->n_calls;
  const int new_old = h->cur, new_cur = h->spare;
  h->old = new_old;
  h->cur = new_cur;
  h->spare = new_spare;
  double *T = h->buf[h->cur].get();
  const double *T_old = h->buf[h->old].get();
  if (op->d64.kappa_mode == KAPPA_LAW) op->d64.T = T;
  if (op->d64.robin_mask) op->T_robin = T;
  h->fell_back = 0;
  if (h->prm.step_mode == MPDG_STEP_LINEAR) return linear_step(h, dt, T, T_old, lin_iters, nl_iters, s);
  MPDG_TRY(learned_guess(h, dt, nullptr, T, T_old, s));
  // the full tangent is evaluated at the current iterate (and T_old)
  if (op->tangent) MPDG_TRY(mpdg_heat_op_bind_temperature(op, T, T_old, s));
  // preconditioner.update() at the predicted state (heat_diffuse_operation.cpp:276)
  if (h->pc) MPDG_TRY(h->pc->update(s));
  // NewtonRaphsonSolver::solve (newton_raphson_solver.cpp:24-66)
  int acc = 0, it = 0;
  bool converged = false;
  const int max_it = std::max(1, h->prm.max_nonlinear_iterations);
  for (it = 0; it < max_it; ++it) {
    if (op->tangent && it > 0) MPDG_TRY(mpdg_heat_op_bind_temperature(op, T, T_old, s));
    MPDG_TRY(launch_fill<double>(h->rhs.get(), 0.0, n, s));
    MPDG_TRY(op->create_rhs(h->rhs.get(), T, T_old, nullptr, s));
    MPDG_TRY(launch_fill<double>(h->du.get(), 0.0, n, s));
    int li = 0;
    double res = 0.0;
    const int rc = krylov_solve(op, h->pc, h->prm.linear, h->rhs.get(), h->du.get(), h->work, &li, &res, s);
    acc += li;
    if (rc == MPDG_ERR_NONFINITE) return rc;
    if (rc != MPDG_OK && rc != MPDG_ERR_NO_CONVERGENCE) return rc;
    if (rc == MPDG_ERR_NO_CONVERGENCE) return rc;  // SolverControl::NoConvergence is rethrown (linear_solver.hpp:124-128)
    // is_converged (newton_raphson_solver.cpp:104-127)
    MPDG_TRY(launch_dot<double, double>(h->rhs.get(), h->rhs.get(), n, h->red.red(), h->scal.get(), s));
    MPDG_TRY(launch_dot<double, double>(h->du.get(), h->du.get(), n, h->red.red(), h->scal.get() + 1, s));
    double nr[2];
    if (hipMemcpyAsync(nr, h->scal.get(), 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
      return MPDG_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
    if (!(nr[0] == nr[0]) || !(nr[1] == nr[1])) return MPDG_ERR_NONFINITE;
    if (std::sqrt(nr[0]) < h->prm.residual_tolerance && std::sqrt(nr[1]) < h->prm.field_correction_tolerance) {
      converged = true;
      break;
    }
    MPDG_TRY(launch_axpby(T, 1.0, h->du.get(), 1.0, n, s));
  }
  commit_time_increment(h, dt);
  if (lin_iters) *lin_iters = acc;
  if (nl_iters) *nl_iters = it + 1;
  if (!converged) return MPDG_ERR_NO_CONVERGENCE;  // ExcNewtonDidNotConverge
  return hipStreamSynchronize(s) == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int mpdg_heat_operation_set_learned(mpdg_heat_operation *h, mpdg_learned *net, const double *q_field) {
  if (!h) return MPDG_ERR_INVALID;
  if (net) {
    int fd[3];
    MPDG_TRY(mpdg_learned_grid(net, fd, nullptr));
    for (int d = 0; d < 3; ++d)
      if (fd[d] != h->op->d64.nd[d]) return MPDG_ERR_INVALID;  // network made for another DoF grid
  }
  // pinned host copy of the guess check's norms, read after each step's solve
  if (net && !h->scal_host && hipHostMalloc((void **)&h->scal_host, 2 * sizeof(double), 0) != hipSuccess) {
    h->scal_host = nullptr;
    return MPDG_ERR_NOMEM;
  }
  if (net) {
    // the learned step's vectors (residual input, inverse diagonal, base guess and its
    // residual), so that no step allocates
    const size_t n = (size_t)h->op->n_dofs();
    MPDG_TRY(h->zbuf.resize(n));
    MPDG_TRY(h->invd.resize(n));
    MPDG_TRY(h->rbase.resize(n));
    MPDG_TRY(h->tbase.resize(n));
  }
  h->net = net;
  h->q_field = q_field;
  return MPDG_OK;
}
int mpdg_heat_operation_record_guess_feature(mpdg_heat_operation *h, double *z, double *base) {
  if (!h) return MPDG_ERR_INVALID;
  h->z_record = z;
  h->base_record = z ? base : nullptr;
  return MPDG_OK;
}
int mpdg_heat_operation_last_step_info(const mpdg_heat_operation *h, int *fell_back, double *r0, double *r) {
  if (!h) return MPDG_ERR_INVALID;
  if (fell_back) *fell_back = h->fell_back;
  if (r0) *r0 = h->rel_res0;
  if (r) *r = h->rel_res;
  return MPDG_OK;
}
int mpdg_stream_probe(int kind, double *dst, const double *src, long n, void *stream) {
  if (!dst || n < 2 || n % 2 || (kind != 2 && !src)) return MPDG_ERR_INVALID;
  if (reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16) return MPDG_ERR_INVALID;
  if (kind == 1 && n < 4096) return MPDG_ERR_INVALID;  // dst holds the 4096 partial sums
  return launch_stream(kind, dst, src, n, S(stream));
}
int mpdg_heat_op_surface_flux_field(const mpdg_heat_op *op, double *out, void *stream) {
  if (!op || !out) return MPDG_ERR_INVALID;
  return launch_surface_flux_field(op->d64, op->flux, out, S(stream));
}

const double *mpdg_heat_operation_solution(const mpdg_heat_operation *h) { return h ? h->buf[h->cur].get() : nullptr; }
const double *mpdg_heat_operation_old_solution(const mpdg_heat_operation *h) {
  return h ? h->buf[h->old].get() : nullptr;
}

}  // extern "C"
