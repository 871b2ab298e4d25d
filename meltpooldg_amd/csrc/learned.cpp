// Learned initial guess for the implicit heat step: a 4-layer 3D convolutional
// network (4 -> 16 -> 16 -> 16 -> 1 channels, 3x3x3 taps, ReLU) on the nodal
// grid predicts T_{n+1} from (T_{n-1}, T_n, heat source). It plugs into the
// reference's predictor slot (include/meltpooldg/linear_algebra/predictor.hpp:118-160)
// as a new predictor type; training and inference run in the BF16 MFMA kernels
// of cnn.hip with FP32 master weights and Adam.
//
// Features (per sample, s = rms(T_n - T_{n-1})): c0 = (T_n - T_{n-1})/s,
// c1 = q * qscale / s, c2 = (T_n - mean T_n)/rms, c3 = z / sz, where z is the
// residual of the base guess T_b for the step (a polynomial extrapolation of the
// history; by default the linear one, T_lin = 2 T_n - T_{n-1}), scaled by the inverse
// operator diagonal (z = D^-1 (b - J T_b)), and sz = rms(z). The network output y is
// the correction of the base guess in units of sz: guess = T_b + sz y. The exact
// correction is J^-1 (b - J T_b);
// where the operator is mass-dominated (dt k / (rho c_p h^2) of order one or
// less) that is a short-range smoothing of z, which the network learns from the
// solver's own (T_{n-1}, T_n, q_{n+1}, z_n) -> T_{n+1} pairs.
//
// Working grid: the network runs on the DoF grid of Q_k on (n_cells / rho) cells, rho in
// {1, 2, 4, 8}. The fields are carried there by full weighting W = diag(P^T 1)^-1 P^T and
// the correction back by the FE prolongation P (Q_k on n_cells/rho -> Q_k on n_cells), both
// applied direction by direction as log2(rho) structured 2:1 transfers. The correction of the extrapolation is smooth on the
// diffusion length sqrt(k dt / rho c_p); when that spans more nodes than the 9^3 receptive
// field (stiff steps, dt k / h^2 >> 1) a coarser working grid lets the network see it, at
// rho^3 less work.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "cnn.hpp"
#include "heat_lib.hpp"

using namespace mpdg;

namespace {
constexpr int kLayers = 4;
constexpr int kCin[kLayers] = {4, 16, 16, 16};
constexpr int kCout[kLayers] = {16, 16, 16, 1};
constexpr int kWgradBlocks = 256;
inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

// deterministic Gaussian numbers (splitmix64 + Box-Muller)
struct Rng {
  unsigned long long x;
  unsigned long long next() {
    unsigned long long z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};
}  // namespace

struct mpdg_learned {
  int fdims[3] = {0, 0, 0};  // DoF grid of the fields passed in
  int dims[3] = {0, 0, 0};   // network grid
  int rho = 1, degree = 1;
  long nf = 0, n = 0;
  double qscale = 1.0;
  // output scale: the correction is oscale * rms(z) * y. Set from the first training sample
  // (so that its target has rms 1) unless given; where the step is stiff the correction is
  // much larger than the diagonal-scaled residual and an O(1) output could not reach it.
  double oscale = 1.0;
  bool oscale_set = false;
  int step = 0;
  DeviceArray<float> w[kLayers], b[kLayers], mw[kLayers], vw[kLayers], mb[kLayers], vb[kLayers];
  DeviceArray<unsigned short> W[kLayers], Wt[kLayers], act[kLayers + 1], g0, g1;
  DeviceArray<float> dW[kLayers], db[kLayers], part;
  DeviceArray<double> stats, spart, lpart;
  DeviceArray<unsigned> scount;  // workgroups done in the statistics pass (back at zero after it)
  // rho > 1: the fields on the network grid and the intermediate arrays of the direction passes
  DeviceArray<double> cf[6], tmpa, tmpb, corr, corr_f;
  long tmpa_n = 0, tmpb_n = 0;  // per field: the transfers carry up to kEllMaxFields fields at once
  // rho = 2, 4 or 8: the structured 2:1 transfers of the multigrid (one coarse cell's table, no
  // index arrays), log2(rho) times through the DoF grids of n_cells / 2 and n_cells / 4 (the Q_k
  // spaces are nested, so P of rho is the product of the 2:1 prolongations); the full weighting
  // as passes of P^T and the inverse row sums of the whole transfer per direction
  int hlev = 0;                  // 2:1 levels
  int mdims[2][3] = {};          // DoF grids between them
  long mid_n[2] = {0, 0};
  std::vector<double> htab;
  DeviceArray<double> htab_dev, isum[3], mid[2];
  // DoF grid l of the 2:1 levels: 0 the fields', hlev the network's
  const int *hgrid(int l) const { return l == 0 ? fdims : l == hlev ? dims : mdims[l - 1]; }

  int init(const int *cells, int k, int coarsening, double qs, unsigned long long seed) {
    degree = k;
    rho = coarsening;
    for (int i = 0; i < 3; ++i) {
      fdims[i] = k * cells[i] + 1;
      dims[i] = k * (cells[i] / rho) + 1;
    }
    nf = (long)fdims[0] * fdims[1] * fdims[2];
    n = (long)dims[0] * dims[1] * dims[2];
    qscale = qs;
    Rng rng{seed};
    for (int l = 0; l < kLayers; ++l) {
      const int nw = kCout[l] * kCin[l] * 27;
      std::vector<float> hw(nw), hb(kCout[l], 0.0f);
      const double sd = std::sqrt(2.0 / (kCin[l] * 27.0)) * (l == kLayers - 1 ? 0.1 : 1.0);
      for (auto &v : hw) v = (float)(sd * rng.normal());
      MPDG_TRY(w[l].resize(nw));
      MPDG_TRY(b[l].resize(kCout[l]));
      MPDG_TRY(mw[l].resize(nw));
      MPDG_TRY(vw[l].resize(nw));
      MPDG_TRY(mb[l].resize(kCout[l]));
      MPDG_TRY(vb[l].resize(kCout[l]));
      MPDG_TRY(W[l].resize(27 * 256));
      MPDG_TRY(Wt[l].resize(27 * 256));
      MPDG_TRY(dW[l].resize(27 * 256));
      MPDG_TRY(db[l].resize(16));
      hipMemcpy(w[l].get(), hw.data(), nw * sizeof(float), hipMemcpyHostToDevice);
      hipMemcpy(b[l].get(), hb.data(), hb.size() * sizeof(float), hipMemcpyHostToDevice);
      hipMemset(mw[l].get(), 0, nw * sizeof(float));
      hipMemset(vw[l].get(), 0, nw * sizeof(float));
      hipMemset(mb[l].get(), 0, kCout[l] * sizeof(float));
      hipMemset(vb[l].get(), 0, kCout[l] * sizeof(float));
    }
    for (int l = 0; l <= kLayers; ++l) MPDG_TRY(act[l].resize((size_t)n * 16));
    MPDG_TRY(g0.resize((size_t)n * 16));
    MPDG_TRY(g1.resize((size_t)n * 16));
    MPDG_TRY(part.resize((size_t)kWgradBlocks * 4 * (27 * 256 + 16)));  // one slot per wavefront
    MPDG_TRY(stats.resize(4));
    MPDG_TRY(spart.resize(4 * 1024));
    MPDG_TRY(scount.resize(1));
    if (hipMemset(scount.get(), 0, sizeof(unsigned)) != hipSuccess) return MPDG_ERR_DEVICE;
    MPDG_TRY(lpart.resize(1024));
    MPDG_TRY(corr.resize(n));
    if (rho > 1) {
      // rho = 2^hlev; the transfers stage x lines of at most 1024 nodes (create checks both)
      hlev = rho == 2 ? 1 : rho == 4 ? 2 : 3;
      Transfer1DHost cell(k, 1, k, 2, false, false);  // one coarse cell: rows m = 0..2k, cols j = 0..k
      htab.assign((size_t)(2 * k + 1) * (k + 1), 0.0);
      for (int m = 0; m <= 2 * k; ++m)
        for (int e = 0; e < cell.pw; ++e) {
          const int j = cell.pcol[(size_t)m * cell.pw + e];
          if (j >= 0) htab[(size_t)m * (k + 1) + j] = cell.pval[(size_t)m * cell.pw + e];
        }
      MPDG_TRY(htab_dev.resize(htab.size()));
      hipMemcpy(htab_dev.get(), htab.data(), htab.size() * sizeof(double), hipMemcpyHostToDevice);
      // full weighting: the rows of P^T (of the whole transfer) scaled to sum 1
      for (int d = 0; d < 3; ++d) {
        Transfer1DHost t(k, cells[d] / rho, k, cells[d], false, false);
        std::vector<double> inv(t.nc);
        for (int c = 0; c < t.nc; ++c) {
          double sum = 0.0;
          for (int e = 0; e < t.rw; ++e) sum += t.rval[(size_t)c * t.rw + e];
          inv[c] = 1.0 / sum;
        }
        MPDG_TRY(isum[d].resize(inv.size()));
        hipMemcpy(isum[d].get(), inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice);
      }
      for (auto &a : cf) MPDG_TRY(a.resize(n));
      for (int l = 0; l + 1 < hlev; ++l) {
        for (int d = 0; d < 3; ++d) mdims[l][d] = k * (cells[d] >> (l + 1)) + 1;
        mid_n[l] = (long)mdims[l][0] * mdims[l][1] * mdims[l][2];
        MPDG_TRY(mid[l].resize((size_t)kEllMaxFields * mid_n[l]));
      }
      // x then y then z, every pass into tmpa, tmpb, out (restriction) or tmpb, tmpa, out
      // (prolongation): per field the largest intermediate of either, over the levels
      long a1 = 0, a2 = 0;
      for (int l = 0; l < hlev; ++l) {
        const int *f = hgrid(l), *c = hgrid(l + 1);
        a1 = std::max({a1, (long)c[0] * f[1] * f[2], (long)f[0] * f[1] * c[2]});
        a2 = std::max({a2, (long)c[0] * c[1] * f[2], (long)f[0] * c[1] * c[2]});
      }
      tmpa_n = a1;
      tmpb_n = a2;
      MPDG_TRY(tmpa.resize((size_t)kEllMaxFields * a1));
      MPDG_TRY(tmpb.resize((size_t)kEllMaxFields * a2));
      MPDG_TRY(corr_f.resize(nf));
    }
    return pack(nullptr);
  }
  int pack(hipStream_t s) {
    for (int l = 0; l < kLayers; ++l) MPDG_TRY(cnn_pack(w[l].get(), kCin[l], kCout[l], W[l].get(), Wt[l].get(), s));
    return MPDG_OK;
  }
  // DoF-grid fields f[0..count) on the network grid, into the network-grid slots slot0, slot0+1,
  // ... (the fields themselves when rho = 1; nullptr stays nullptr): every 1D pass of the full
  // weighting carries all the fields in one launch
  // with_stats (forward's inputs T_{n-1}, T_n, q, z): the input statistics come with the
  // transfer (done is set), otherwise they are left to cnn_stats
  int to_net(const double *const *f, int slot0, int count, const double **net, hipStream_t s,
             bool with_stats = false, bool *done = nullptr) {
    if (done) *done = false;
    if (rho > 1) {
      // per 2:1 level three passes of P^T (x, y, z), each carrying all the fields, then the row
      // scaling of all of them at once
      EllFields<double> F[9] = {};
      CnnFields C{};
      for (int i = 0; i < count; ++i) {
        net[i] = f[i];
        if (!f[i]) continue;
        const int k = F[0].nf;
        const double *src = f[i];
        for (int l = 0; l < hlev; ++l) {
          EllFields<double> *G = F + 3 * l;
          double *dst = l + 1 < hlev ? mid[l].get() + k * mid_n[l] : cf[slot0 + i].get();
          G[0].in[k] = src;
          G[0].out[k] = tmpa.get() + k * tmpa_n;
          G[1].in[k] = G[0].out[k];
          G[1].out[k] = tmpb.get() + k * tmpb_n;
          G[2].in[k] = G[1].out[k];
          G[2].out[k] = dst;
          src = dst;
        }
        for (auto &g : F) g.nf = k + 1;
        C.f[C.nf++] = cf[slot0 + i].get();
        net[i] = cf[slot0 + i].get();
      }
      if (C.nf == 0) return MPDG_OK;
      for (int l = 0; l < hlev; ++l) {
        const int *a = hgrid(l), *c = hgrid(l + 1);
        const int o0[3] = {c[0], a[1], a[2]}, o1[3] = {c[0], c[1], a[2]};
        MPDG_TRY(launch_htransfer_fields<double>(true, 0, F[3 * l], a, o0, degree, htab.data(), htab_dev.get(), 0,
                                                 0, false, s));
        MPDG_TRY(launch_htransfer_fields<double>(true, 1, F[3 * l + 1], o0, o1, degree, htab.data(), htab_dev.get(),
                                                 0, 0, false, s));
        MPDG_TRY(launch_htransfer_fields<double>(true, 2, F[3 * l + 2], o1, c, degree, htab.data(), htab_dev.get(),
                                                 0, 0, false, s));
      }
      if (with_stats && f[0] && f[1]) {
        // T_{n-1} and T_n are the first two fields carried; z, when given, the last
        const int sf[3] = {0, 1, f[3] ? C.nf - 1 : -1};
        MPDG_TRY(cnn_scale3(C, dims, isum[0].get(), isum[1].get(), isum[2].get(), s, sf, spart.get(), scount.get(),
                            stats.get()));
        if (done) *done = true;
        return MPDG_OK;
      }
      return cnn_scale3(C, dims, isum[0].get(), isum[1].get(), isum[2].get(), s);
    }
    for (int i = 0; i < count; ++i) net[i] = f[i];  // rho = 1: the fields themselves
    return MPDG_OK;
  }
  int forward(const double *Tnm1, const double *Tn, const double *q, const double *z, hipStream_t s) {
    const double *in[4] = {Tnm1, Tn, q, z}, *net[4];
    bool have_stats = false;
    MPDG_TRY(to_net(in, 0, 4, net, s, true, &have_stats));
    const double *a = net[0], *t = net[1], *qq = net[2], *zz = net[3];
    last_Tnm1 = a;
    last_Tn = t;
    if (!have_stats) MPDG_TRY(cnn_stats(a, t, zz, n, spart.get(), scount.get(), stats.get(), s));
    MPDG_TRY(cnn_features(a, t, qq, zz, n, stats.get(), qscale, act[0].get(), s));
    for (int l = 0; l < kLayers; ++l)
      MPDG_TRY(cnn_conv(act[l].get(), act[l + 1].get(), W[l].get(), b[l].get(), dims, l == kLayers - 1 ? 0 : 1,
                        nullptr, s));
    return MPDG_OK;
  }
  // rms of the target (T_{n+1} - T_b) / rms(z) of a sample after forward()
  int target_rms(const double *base, const double *Tnp1, double *rms, hipStream_t s) {
    const double *in[2] = {Tnp1, base}, *net[2];
    MPDG_TRY(to_net(in, 4, 2, net, s));
    return target_rms_net(net[1], net[0], rms, 512, s);
  }
  const double *last_Tnm1 = nullptr, *last_Tn = nullptr;  // network-grid inputs of the last forward()
  int target_rms_net(const double *bn, const double *tp, double *rms, int nb, hipStream_t s) {
    MPDG_TRY(cnn_loss(nullptr, last_Tnm1, last_Tn, bn, tp, n, stats.get(), 1.0, 1.0f, nullptr, lpart.get(), nb, s));
    std::vector<double> h(nb);
    if (hipMemcpyAsync(h.data(), lpart.get(), nb * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
      return MPDG_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
    double sum = 0.0;
    for (double v : h) sum += v;
    *rms = std::sqrt(sum / (double)n);
    return MPDG_OK;
  }
  // loss and gradients of all layers (dW, db) after forward(); returns the mean squared error
  int backward(const double *Tnm1, const double *Tn, const double *base, const double *Tnp1, double *loss,
               hipStream_t s) {
    const int nb = 512;
    const double *a = rho == 1 ? Tnm1 : cf[0].get(), *t = rho == 1 ? Tn : cf[1].get();
    const double *in[2] = {Tnp1, base}, *net[2];
    MPDG_TRY(to_net(in, 4, 2, net, s));
    const double *tp = net[0], *bn = net[1];
    MPDG_TRY(cnn_loss(act[kLayers].get(), a, t, bn, tp, n, stats.get(), oscale, 1.0f, g0.get(), lpart.get(), nb, s));
    unsigned short *g = g0.get(), *gn = g1.get();
    for (int l = kLayers - 1; l >= 0; --l) {
      MPDG_TRY(cnn_wgrad(g, act[l].get(), dims, part.get(), kWgradBlocks, dW[l].get(), db[l].get(), s));
      if (l > 0) {
        // gradient of the previous layer's pre-activation: (W^T conv g) * relu'(act[l])
        MPDG_TRY(cnn_conv(g, gn, Wt[l].get(), nullptr, dims, 2, act[l].get(), s));
        std::swap(g, gn);
      }
    }
    if (loss) {
      std::vector<double> h(nb);
      if (hipMemcpyAsync(h.data(), lpart.get(), nb * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
        return MPDG_ERR_DEVICE;
      if (hipStreamSynchronize(s) != hipSuccess) return MPDG_ERR_DEVICE;
      double sum = 0.0;
      for (double v : h) sum += v;
      *loss = sum / (double)n;
    }
    return MPDG_OK;
  }
  // out = T_b + P corr after forward() (T_b = base, or 2 Tn - Tnm1 when base is nullptr; out
  // may be base)
  int predict(const double *Tnm1, const double *Tn, const double *base, double *out, hipStream_t s) {
    MPDG_TRY(cnn_correction(act[kLayers].get(), n, stats.get(), oscale, corr.get(), s));
    const double *c = corr.get();
    if (rho > 1) {
      // per 2:1 level (coarsest first) three passes of P (x, y, z)
      const double *src = corr.get();
      for (int l = hlev - 1; l >= 0; --l) {  // from grid l + 1 to grid l
        const int *a = hgrid(l), *b = hgrid(l + 1);
        double *dst = l > 0 ? mid[l - 1].get() : corr_f.get();
        const int o0[3] = {a[0], b[1], b[2]}, o1[3] = {a[0], a[1], b[2]};
        MPDG_TRY(launch_htransfer<double>(false, 0, src, nullptr, tmpb.get(), b, o0, degree, htab.data(),
                                          htab_dev.get(), 0, 0, false, s));
        MPDG_TRY(launch_htransfer<double>(false, 1, tmpb.get(), nullptr, tmpa.get(), o0, o1, degree, htab.data(),
                                          htab_dev.get(), 0, 0, false, s));
        MPDG_TRY(launch_htransfer<double>(false, 2, tmpa.get(), nullptr, dst, o1, a, degree, htab.data(),
                                          htab_dev.get(), 0, 0, false, s));
        src = dst;
      }
      c = corr_f.get();
    }
    return cnn_extrapolate_add(Tnm1, Tn, base, c, nf, out, s);
  }
};

extern "C" {

int mpdg_learned_create(const int *n_cells, int degree, int coarsening, double qscale, unsigned long long seed,
                        mpdg_learned **out) {
  if (!n_cells || !out || degree < 1 || degree > 4) return MPDG_ERR_INVALID;
  if (coarsening != 1 && coarsening != 2 && coarsening != 4 && coarsening != 8) return MPDG_ERR_INVALID;
  for (int d = 0; d < 3; ++d)
    if (n_cells[d] < coarsening || n_cells[d] % coarsening != 0) return MPDG_ERR_INVALID;
  // the 2:1 transfers stage x lines of at most 1024 DoFs
  if (coarsening > 1 && (long)degree * n_cells[0] + 1 > 1024) return MPDG_ERR_INVALID;
  auto *h = new (std::nothrow) mpdg_learned;
  if (!h) return MPDG_ERR_NOMEM;
  const int rc = h->init(n_cells, degree, coarsening, qscale, seed);
  if (rc) {
    delete h;
    return rc;
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    delete h;
    return MPDG_ERR_DEVICE;
  }
  *out = h;
  return MPDG_OK;
}

int mpdg_learned_destroy(mpdg_learned *h) {
  delete h;
  return MPDG_OK;
}

int mpdg_learned_n_params(const mpdg_learned *h) {
  if (!h) return -1;
  int n = 0;
  for (int l = 0; l < kLayers; ++l) n += kCout[l] * kCin[l] * 27 + kCout[l];
  return n;
}

// weights as one FP32 array: per layer W[co][ci][tap] then b[co]
int mpdg_learned_get_weights(const mpdg_learned *h, float *host) {
  if (!h || !host) return MPDG_ERR_INVALID;
  size_t o = 0;
  for (int l = 0; l < kLayers; ++l) {
    const int nw = kCout[l] * kCin[l] * 27;
    if (hipMemcpy(host + o, h->w[l].get(), nw * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return MPDG_ERR_DEVICE;
    o += nw;
    if (hipMemcpy(host + o, h->b[l].get(), kCout[l] * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return MPDG_ERR_DEVICE;
    o += kCout[l];
  }
  return MPDG_OK;
}

int mpdg_learned_set_weights(mpdg_learned *h, const float *host) {
  if (!h || !host) return MPDG_ERR_INVALID;
  size_t o = 0;
  for (int l = 0; l < kLayers; ++l) {
    const int nw = kCout[l] * kCin[l] * 27;
    if (hipMemcpy(h->w[l].get(), host + o, nw * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return MPDG_ERR_DEVICE;
    o += nw;
    if (hipMemcpy(h->b[l].get(), host + o, kCout[l] * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return MPDG_ERR_DEVICE;
    o += kCout[l];
  }
  MPDG_TRY(h->pack(nullptr));
  return hipDeviceSynchronize() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int mpdg_learned_predict(mpdg_learned *h, const double *Tnm1, const double *Tn, const double *q, const double *z,
                         const double *base, double *out, void *stream) {
  if (!h || !Tnm1 || !Tn || !out) return MPDG_ERR_INVALID;
  MPDG_TRY(h->forward(Tnm1, Tn, q, z, S(stream)));
  return h->predict(Tnm1, Tn, base, out, S(stream));
}

// gradients (FP32, same layout as the weights) and the loss of one sample, no update
int mpdg_learned_gradients(mpdg_learned *h, const double *Tnm1, const double *Tn, const double *q, const double *z,
                           const double *base, const double *Tnp1, float *host_grad, double *loss, void *stream) {
  if (!h || !Tnm1 || !Tn || !Tnp1 || !host_grad) return MPDG_ERR_INVALID;
  hipStream_t s = S(stream);
  MPDG_TRY(h->forward(Tnm1, Tn, q, z, s));
  MPDG_TRY(h->backward(Tnm1, Tn, base, Tnp1, loss, s));
  size_t o = 0;
  for (int l = 0; l < kLayers; ++l) {
    std::vector<float> dw(27 * 256), db(16);
    if (hipMemcpyAsync(dw.data(), h->dW[l].get(), dw.size() * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(db.data(), h->db[l].get(), db.size() * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return MPDG_ERR_DEVICE;
    for (int co = 0; co < kCout[l]; ++co)
      for (int ci = 0; ci < kCin[l]; ++ci)
        for (int t = 0; t < 27; ++t) host_grad[o++] = dw[t * 256 + co * 16 + ci];
    for (int co = 0; co < kCout[l]; ++co) host_grad[o++] = db[co];
  }
  return MPDG_OK;
}

// one Adam step on one sample (T_{n-1}, T_n, q, z, T_b) -> T_{n+1}; loss = mean squared error
// of the normalised correction before the update
int mpdg_learned_train_step(mpdg_learned *h, const double *Tnm1, const double *Tn, const double *q, const double *z,
                            const double *base, const double *Tnp1, double lr, double *loss, void *stream) {
  if (!h || !Tnm1 || !Tn || !Tnp1) return MPDG_ERR_INVALID;
  hipStream_t s = S(stream);
  MPDG_TRY(h->forward(Tnm1, Tn, q, z, s));
  if (!h->oscale_set) {
    double rms = 0.0;
    MPDG_TRY(h->target_rms(base, Tnp1, &rms, s));
    h->oscale = (std::isfinite(rms) && rms > 1e-12) ? rms : 1.0;
    h->oscale_set = true;
  }
  MPDG_TRY(h->backward(Tnm1, Tn, base, Tnp1, loss, s));
  h->step += 1;
  for (int l = 0; l < kLayers; ++l)
    MPDG_TRY(cnn_adam(h->w[l].get(), h->mw[l].get(), h->vw[l].get(), h->dW[l].get(), h->db[l].get(), h->b[l].get(),
                      h->mb[l].get(), h->vb[l].get(), kCin[l], kCout[l], (float)lr, 0.9f, 0.999f, 1e-8f, h->step,
                      1.0f, s));
  MPDG_TRY(h->pack(s));
  if (loss && !std::isfinite(*loss)) return MPDG_ERR_NONFINITE;
  return MPDG_OK;
}

// output scale of the correction (see mpdg_learned::oscale)
int mpdg_learned_get_output_scale(const mpdg_learned *h, double *oscale) {
  if (!h || !oscale) return MPDG_ERR_INVALID;
  *oscale = h->oscale;
  return MPDG_OK;
}
int mpdg_learned_set_output_scale(mpdg_learned *h, double oscale) {
  if (!h || !(oscale > 0.0) || !std::isfinite(oscale)) return MPDG_ERR_INVALID;
  h->oscale = oscale;
  h->oscale_set = true;
  return MPDG_OK;
}

// DoF grid of the fields and the network's working grid
int mpdg_learned_grid(const mpdg_learned *h, int *dof_dims, int *net_dims) {
  if (!h) return MPDG_ERR_INVALID;
  for (int d = 0; d < 3; ++d) {
    if (dof_dims) dof_dims[d] = h->fdims[d];
    if (net_dims) net_dims[d] = h->dims[d];
  }
  return MPDG_OK;
}

// raw network input/output of the last forward pass (bf16 bit patterns, [voxel][16])
int mpdg_learned_activation(const mpdg_learned *h, int layer, unsigned short *host) {
  if (!h || !host || layer < 0 || layer > kLayers) return MPDG_ERR_INVALID;
  return hipMemcpy(host, h->act[layer].get(), (size_t)h->n * 16 * sizeof(unsigned short), hipMemcpyDeviceToHost) ==
                 hipSuccess
             ? MPDG_OK
             : MPDG_ERR_DEVICE;
}

}  // extern "C"
