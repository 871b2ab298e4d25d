// Kernels of the learned initial guess (3D convolutional network, BF16 MFMA).
#pragma once

#include "kernels.hpp"

namespace mpdg {

int cnn_conv(const unsigned short *in, unsigned short *out, const unsigned short *W, const float *bias, const int *dims,
             int mode, const unsigned short *mask, hipStream_t s);
// part: 4 nblocks (27 * 256 + 16) floats (one slot per wavefront)
int cnn_wgrad(const unsigned short *g, const unsigned short *in, const int *dims, float *part, int nblocks, float *dW,
              float *db, hipStream_t s);
int cnn_pack(const float *master, int cin, int cout, unsigned short *W, unsigned short *Wt, hipStream_t s);
int cnn_adam(float *w, float *m, float *v, const float *dW, const float *db, float *b, float *mb, float *vb, int cin,
             int cout, float lr, float b1, float b2, float eps, int step, float gscale, hipStream_t s);
// up to four fields on one grid
struct CnnFields {
  double *f[4];
  int nf;
};
// f *= sx[ix] sy[iy] sz[iz] for every field (dims = grid sizes along x, y, z); with stat_fields
// = {T_{n-1}, T_n, z (or -1)} indices into F, also the input statistics of the scaled fields
// (part, counter and stats as cnn_stats takes them)
int cnn_scale3(const CnnFields &F, const int *dims, const double *sx, const double *sy, const double *sz,
               hipStream_t s, const int *stat_fields = nullptr, double *part = nullptr, unsigned *counter = nullptr,
               double *stats = nullptr);
// input statistics (part: 4 x 1024 partial sums; counter: zero before the first call, left at zero)
int cnn_stats(const double *Tnm1, const double *Tn, const double *z, long n, double *part, unsigned *counter,
              double *stats, hipStream_t s);
int cnn_features(const double *Tnm1, const double *Tn, const double *q, const double *z, long n,
                 const double *stats, double qscale, unsigned short *x, hipStream_t s);
// base: the field the network corrects (nullptr: the linear extrapolation 2 Tn - Tnm1)
int cnn_loss(const unsigned short *y, const double *Tnm1, const double *Tn, const double *base, const double *Tnp1,
             long n, const double *stats, double oscale, float gain, unsigned short *g, double *part, int nblocks,
             hipStream_t s);
// corr = rms(z) oscale y on the network grid; out = base + corr on the DoF grid (base nullptr:
// 2 Tn - Tnm1)
int cnn_correction(const unsigned short *y, long n, const double *stats, double oscale, double *corr, hipStream_t s);
int cnn_extrapolate_add(const double *Tnm1, const double *Tn, const double *base, const double *corr, long n,
                        double *out, hipStream_t s);

}  // namespace mpdg
