// BF16 MFMA kernels of the learned initial guess: a 3D convolutional network
// (3x3x3 taps, 16 channels, zero padding) on the nodal grid of the heat
// problem, trained and evaluated on the GPU.
//
// Activations are channel-last bf16 tensors act[voxel][16] (voxel = x + nx*(y
// + ny*z)); a layer with fewer than 16 input or output channels is padded with
// zero channels. One v_mfma_f32_16x16x16_bf16 computes, for one of the 27 taps,
// 16 output channels x 16 voxels (a run of 16 voxels along x) with the 16 input
// channels as the contraction; the 27 taps accumulate in the same registers.
//   A operand (M = out channel, K = in channel): lane l holds W[tap][m = l%16][k = 4(l/16)..+3]
//   B operand (K = in channel, N = voxel):       lane l holds act[voxel n = l%16][k = 4(l/16)..+3]
//   D (M = out channel, N = voxel), f32:          lane l holds D[m = 4(l/16)+i][n = l%16], i = 0..3
// so every lane moves 8 contiguous bytes per operand and writes 4 contiguous
// output channels of one voxel.
#include "cnn.hpp"

namespace mpdg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f32_to_bf16(float f) {
  // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  const unsigned r = 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)((u + r) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

// four consecutive bf16 values (8-byte aligned) as the MFMA operand, in one 8-byte load
__device__ __forceinline__ bf16x4 load4(const unsigned short *p) { return *reinterpret_cast<const bf16x4 *>(p); }

// ---------------------------------------------------------------------------
// forward / data-gradient convolution: out[v][co] = act(bias[co] + sum_tap,ci W[tap][co][ci] in[v+off(tap)][ci])
// mode 0: linear, 1: ReLU, 2: multiply by the ReLU mask of `mask` (backward: g * (a > 0))
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void conv3_kernel(const unsigned short *__restrict__ in,
                                                    unsigned short *__restrict__ out,
                                                    const unsigned short *__restrict__ W, const float *__restrict__ bias,
                                                    int nx, int ny, int nz, int mode,
                                                    const unsigned short *__restrict__ mask) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * 256) >> 6;
  const int n = lane & 15, kg = lane >> 4;
  bf16x4 a[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) a[t] = load4(W + (t * 16 + n) * 16 + 4 * kg);
  float bv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) bv[i] = bias ? bias[4 * kg + i] : 0.0f;
  // tiles of 16 consecutive voxels in the lexicographic order (a tile may wrap into the next x
  // row: every lane finds its own neighbours), so no lanes idle at the ends of short x rows
  const long nvox = (long)nx * ny * nz;
  const long ntiles = (nvox + 15) >> 4;
  const int sy = 16 * nx, sz = 16 * nx * ny;  // element strides of y and z neighbours (grids < 2^26 voxels)
  for (long tile = wave; tile < ntiles; tile += nwaves) {
    const long v = tile * 16 + n;
    const long vc = v < nvox ? v : nvox - 1;
    const int x = (int)(vc % nx), y = (int)((vc / nx) % ny), z = (int)(vc / ((long)nx * ny));
    // which neighbours exist; a missing one is read at the voxel itself and replaced by zero,
    // so the 27 loads carry no branches and only one address add each
    const bool xl = x > 0, xh = x < nx - 1, yl = y > 0, yh = y < ny - 1, zl = z > 0, zh = z < nz - 1;
    const unsigned short *c = in + vc * 16 + 4 * kg;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
      const bool ok = (dx < 0 ? xl : (dx > 0 ? xh : true)) && (dy < 0 ? yl : (dy > 0 ? yh : true)) &&
                      (dz < 0 ? zl : (dz > 0 ? zh : true));
      const int off = 16 * dx + sy * dy + sz * dz;
      bf16x4 b = load4(c + (ok ? off : 0));
      if (!ok) b = bf16x4{0, 0, 0, 0};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], b, acc, 0, 0, 0);
    }
    if (v < nvox) {
      unsigned short o[4];
      uint2 mk = make_uint2(0, 0);
      if (mode == 2) mk = *reinterpret_cast<const uint2 *>(mask + v * 16 + 4 * kg);
      const unsigned short mv[4] = {(unsigned short)(mk.x & 0xffff), (unsigned short)(mk.x >> 16),
                                    (unsigned short)(mk.y & 0xffff), (unsigned short)(mk.y >> 16)};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float r = acc[i] + bv[i];
        if (mode == 1) r = r > 0.f ? r : 0.f;
        if (mode == 2) r = bf16_to_f32(mv[i]) > 0.f ? r : 0.f;
        o[i] = f32_to_bf16(r);
      }
      *reinterpret_cast<uint2 *>(out + v * 16 + 4 * kg) =
          make_uint2((unsigned)o[0] | ((unsigned)o[1] << 16), (unsigned)o[2] | ((unsigned)o[3] << 16));
    }
  }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[tap][co][ci] = sum_v g[v][co] in[v+off(tap)][ci], db[co] = sum_v g[v][co]
// (g = gradient of the layer's pre-activation). MFMA with the voxels as the
// contraction: A (M = co, K = voxel): lane holds g[v = x0+4(l/16)+j][co = l%16];
// B (K = voxel, N = ci): lane holds in[v+off][ci = l%16]. Every wavefront walks its own tiles
// in a fixed order and stores its partial sums to part[wavefront][27*256 + 16] (no two lanes of a
// wavefront share an entry); wgrad_reduce sums the wavefronts' partials in a fixed order, so the
// gradient is bitwise repeatable.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void wgrad_kernel(const unsigned short *__restrict__ g,
                                                    const unsigned short *__restrict__ in, int nx, int ny, int nz,
                                                    float *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * 256) >> 6;
  const int n = lane & 15, kg = lane >> 4;
  f32x4 acc[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const int tx = (nx + 15) >> 4;
  const long ntiles = (long)tx * ny * nz;
  for (long tile = wave; tile < ntiles; tile += nwaves) {
    const int x0 = (int)(tile % tx) * 16, y = (int)((tile / tx) % ny), z = (int)(tile / ((long)tx * ny));
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int x = x0 + 4 * kg + j;
      short val = 0;
      if (x < nx) val = (short)g[((long)x + (long)nx * (y + (long)ny * z)) * 16 + n];
      a[j] = val;
      bsum += bf16_to_f32((unsigned short)val);
    }
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
      const int ys = y + dy, zs = z + dz;
      bf16x4 b = {0, 0, 0, 0};
      if (ys >= 0 && ys < ny && zs >= 0 && zs < nz) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int x = x0 + 4 * kg + j, xs = x + dx;
          if (x < nx && xs >= 0 && xs < nx) b[j] = (short)in[((long)xs + (long)nx * (ys + (long)ny * zs)) * 16 + n];
        }
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
  // D[m = co = 4 kg + i][n = ci]
  float *mine = part + (long)wave * (27 * 256 + 16);
#pragma unroll
  for (int t = 0; t < 27; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) mine[t * 256 + (4 * kg + i) * 16 + n] = acc[t][i];
  // bias: lanes with the same co (= n here, since a holds g[.][co = n]) summed over kg
  bsum += __shfl_xor(bsum, 16);
  bsum += __shfl_xor(bsum, 32);
  if (kg == 0) mine[27 * 256 + n] = bsum;
}

__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float *__restrict__ part, int nparts,
                                                           float *__restrict__ dW, float *__restrict__ db) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 27 * 256 + 16) return;
  float s = 0.f;
  for (int b = 0; b < nparts; ++b) s += part[(long)b * (27 * 256 + 16) + i];
  if (i < 27 * 256) dW[i] = s;
  else db[i - 27 * 256] = s;
}

// bf16 weights for the forward pass (W[tap][co][ci], channel counts padded to 16)
// and for the data gradient (Wt[tap'][ci][co] with the tap mirrored), from FP32 masters
// stored as M[co][ci][tap] over the real channel counts.
__global__ void pack_weights_kernel(const float *__restrict__ master, int cin, int cout,
                                    unsigned short *__restrict__ W, unsigned short *__restrict__ Wt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 27 * 256) return;
  const int t = i / 256, co = (i / 16) % 16, ci = i % 16;
  float v = 0.f;
  if (co < cout && ci < cin) v = master[(co * cin + ci) * 27 + t];
  W[i] = f32_to_bf16(v);
  Wt[(26 - t) * 256 + ci * 16 + co] = f32_to_bf16(v);
}

// Adam step on the FP32 masters (weights M[co][ci][tap] and biases), gradient
// taken from the padded dW[tap][co][ci] / db[co].
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v, const float *__restrict__ dW,
                            const float *__restrict__ db, float *__restrict__ b, float *__restrict__ mb,
                            float *__restrict__ vb, int cin, int cout, float lr, float b1, float b2, float eps,
                            float c1, float c2, float gscale) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nw = cout * cin * 27;
  if (i < nw) {
    const int co = i / (cin * 27), ci = (i / 27) % cin, t = i % 27;
    const float g = gscale * dW[t * 256 + co * 16 + ci];
    m[i] = b1 * m[i] + (1.f - b1) * g;
    v[i] = b2 * v[i] + (1.f - b2) * g * g;
    w[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
  } else if (i < nw + cout) {
    const int co = i - nw;
    const float g = gscale * db[co];
    mb[co] = b1 * mb[co] + (1.f - b1) * g;
    vb[co] = b2 * vb[co] + (1.f - b2) * g * g;
    b[co] -= lr * (mb[co] / c1) / (sqrtf(vb[co] / c2) + eps);
  }
}

// ---- features and targets (FP64 fields -> bf16 channels) ----
// stats = {s, mean T_n, tscale, sz}: s = rms(T_n - T_{n-1}), sz = rms(z) (s when z is absent)
// channels: 0 (Tn - Tnm1)/s, 1 q*qscale/s, 2 (Tn - tmean)/tscale, 3 z/sz
__global__ void features_kernel(const double *__restrict__ Tnm1, const double *__restrict__ Tn,
                                const double *__restrict__ q, const double *__restrict__ z, long n,
                                const double *__restrict__ stats, double qscale, unsigned short *__restrict__ x) {
  const double s = stats[0], tmean = stats[1], ts = stats[2], sz = stats[3];
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const unsigned c0 = f32_to_bf16((float)((Tn[i] - Tnm1[i]) / s));
    const unsigned c1 = q ? f32_to_bf16((float)(q[i] * qscale / s)) : 0u;
    const unsigned c2 = f32_to_bf16((float)((Tn[i] - tmean) / ts));
    const unsigned c3 = z ? f32_to_bf16((float)(z[i] / sz)) : 0u;
    uint4 *dst = reinterpret_cast<uint4 *>(x + i * 16);
    dst[0] = make_uint4(c0 | (c1 << 16), c2 | (c3 << 16), 0, 0);
    dst[1] = make_uint4(0, 0, 0, 0);
  }
}

// loss gradient on channel 0: target y* = (Tnp1 - (2 Tn - Tnm1))/(sz oscale), pred = y[v][0];
// g = 2 (pred - y*) / n * gain; partial sums of the squared error. With y = null: partial
// sums of y*^2 only (the output scale's calibration).
__global__ __launch_bounds__(256) void loss_kernel(const unsigned short *__restrict__ y, const double *__restrict__ Tnm1,
                                                   const double *__restrict__ Tn, const double *__restrict__ base,
                                                   const double *__restrict__ Tnp1, long n,
                                                   const double *__restrict__ stats, double oscale, float gain,
                                                   unsigned short *__restrict__ g, double *__restrict__ part) {
  __shared__ double sh[256];
  const double sz = stats[3] * oscale;
  double acc = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const double b = base ? base[i] : 2.0 * Tn[i] - Tnm1[i];
    const double tgt = (Tnp1[i] - b) / sz;
    if (!y) {
      acc += tgt * tgt;
      continue;
    }
    const double e = (double)bf16_to_f32(y[i * 16]) - tgt;
    acc += e * e;
    uint4 *dst = reinterpret_cast<uint4 *>(g + i * 16);
    dst[0] = make_uint4(f32_to_bf16((float)(2.0 * e * gain / (double)n)), 0, 0, 0);
    dst[1] = make_uint4(0, 0, 0, 0);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// correction of the linear extrapolation on the network grid: corr = sz oscale y[v][0]
__global__ void correction_kernel(const unsigned short *__restrict__ y, long n, const double *__restrict__ stats,
                                  double oscale, double *__restrict__ corr) {
  const double sz = stats[3] * oscale;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    corr[i] = sz * (double)bf16_to_f32(y[i * 16]);
}
// guess = base + corr on the DoF grid, base = 2 Tn - Tnm1 when not given (out may be base)
__global__ void extrapolate_add_kernel(const double *__restrict__ Tnm1, const double *__restrict__ Tn,
                                       const double *base, const double *__restrict__ corr, long n,
                                       double *out) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    out[i] = (base ? base[i] : 2.0 * Tn[i] - Tnm1[i]) + corr[i];
}

// statistics of a sample: sums of (Tn - Tnm1)^2, Tn, Tn^2, z^2
// Input statistics in one pass: every workgroup stores its four partial sums, and the last
// workgroup to finish (counted by an atomic counter, which it resets) adds the partials in a
// fixed order: each thread sums the partials t, t + 256, ... (all loads in flight together),
// then a fixed tree over the threads. Outputs: s = rms(Tn - Tnm1) (+ tiny), mean(Tn),
// rms(Tn - mean) (+ tiny), sz = rms(z) (+ tiny) or s without z.
constexpr int kStatsBlocks = 1024;
// The four sums v (of every workgroup's threads) to the statistics: tree over the workgroup,
// partial of the workgroup to part, and the last workgroup to finish adds the partials in a
// fixed order (each thread the partials t, t + 256, ... with all loads in flight, then the tree)
// and writes s = rms(Tn - Tnm1) (+ tiny), mean(Tn), rms(Tn - mean) (+ tiny), sz = rms(z)
// (+ tiny) or s without z. The counter is back at zero afterwards.
__device__ void stats_finish(double (&v)[4], long n, int has_z, double *__restrict__ part,
                             unsigned *__restrict__ counter, double *__restrict__ stats) {
  __shared__ double sh[4][256];
  __shared__ bool is_last;
  const int t = threadIdx.x;
  auto tree = [&]() {
    for (int j = 0; j < 4; ++j) sh[j][t] = v[j];
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if (t < k)
        for (int j = 0; j < 4; ++j) sh[j][t] += sh[j][t + k];
      __syncthreads();
    }
  };
  tree();
  if (t == 0) {
    for (int j = 0; j < 4; ++j) part[j * kStatsBlocks + blockIdx.x] = sh[j][0];
    __threadfence();
    is_last = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  constexpr int PER = kStatsBlocks / 256;
  double w[4][PER];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int blk = t + 256 * r;
      w[j][r] = blk < (int)gridDim.x ? __builtin_nontemporal_load(part + j * kStatsBlocks + blk) : 0.0;
    }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = 0.0;
#pragma unroll
    for (int r = 0; r < PER; ++r) v[j] += w[j][r];
  }
  __syncthreads();
  tree();
  if (t != 0) return;
  *counter = 0u;
  const double a = sh[0][0], b = sh[1][0], c = sh[2][0], e = sh[3][0];
  const double mean = b / n;
  stats[0] = sqrt(a / n) + 1e-30;
  stats[1] = mean;
  stats[2] = sqrt(fmax(c / n - mean * mean, 0.0)) + 1e-12 + 1e-6 * fabs(mean);
  stats[3] = has_z ? sqrt(e / n) + 1e-30 : stats[0];
}
__global__ __launch_bounds__(256) void stats_kernel(const double *__restrict__ Tnm1, const double *__restrict__ Tn,
                                                    const double *__restrict__ z, long n, double *__restrict__ part,
                                                    unsigned *__restrict__ counter, double *__restrict__ stats) {
  double v[4] = {0, 0, 0, 0};
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const double d = Tn[i] - Tnm1[i];
    v[0] += d * d;
    v[1] += Tn[i];
    v[2] += Tn[i] * Tn[i];
    if (z) v[3] += z[i] * z[i];
  }
  stats_finish(v, n, z != nullptr, part, counter, stats);
}

// f <- (sx[ix] sy[iy] sz[iz]) f on an n0 x n1 x n2 grid, for up to four fields: the row scaling
// that turns three passes of P^T into full weighting. With STATS, also the input statistics of
// the scaled fields F.f[ia] (T_{n-1}), F.f[it] (T_n) and F.f[iz] (z; iz < 0: none), as
// stats_kernel computes them.
template <bool STATS>
__global__ __launch_bounds__(256) void scale3_kernel(CnnFields F, int n0, int n1, int n2, const double *__restrict__ sx,
                                                     const double *__restrict__ sy, const double *__restrict__ sz,
                                                     int ia, int it, int iz, double *__restrict__ part,
                                                     unsigned *__restrict__ counter, double *__restrict__ stats) {
  // grids below 2^26 points (as the convolution): 32-bit index arithmetic; the field loop and
  // the selection of the statistics' fields are unrolled, so that nothing is indexed at run
  // time in private memory
  const int n = n0 * n1 * n2, n01 = n0 * n1;
  double v[4] = {0, 0, 0, 0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int izz = i / n01, rem = i - izz * n01, iy = rem / n0, ix = rem - iy * n0;
    const double w = (sx[ix] * sy[iy]) * sz[izz];
    double val[4] = {0, 0, 0, 0};
#pragma unroll
    for (int f = 0; f < 4; ++f)
      if (f < F.nf) {
        val[f] = F.f[f][i] * w;
        F.f[f][i] = val[f];
      }
    if (STATS) {
      double a = 0.0, t = 0.0, zz = 0.0;
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        a = f == ia ? val[f] : a;
        t = f == it ? val[f] : t;
        zz = f == iz ? val[f] : zz;
      }
      const double d = t - a;
      v[0] += d * d;
      v[1] += t;
      v[2] += t * t;
      v[3] += zz * zz;
    }
  }
  if (STATS) stats_finish(v, n, iz >= 0, part, counter, stats);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// workgroups of the statistics pass: one per 1024 points (each ends with a device-scope release
// and a counter atomic, which dominate on the network's small grids), at most kStatsBlocks
static inline int stats_blocks(long n) { return (int)std::max<long>(1, std::min<long>(kStatsBlocks, (n + 1023) / 1024)); }
static inline int cnn_grid(long ntiles) {
  return (int)std::max<long>(1, std::min<long>((ntiles + 3) / 4, 2048));
}

int cnn_conv(const unsigned short *in, unsigned short *out, const unsigned short *W, const float *bias, const int *dims,
             int mode, const unsigned short *mask, hipStream_t s) {
  const long ntiles = ((long)dims[0] * dims[1] * dims[2] + 15) / 16;
  if ((long)dims[0] * dims[1] * dims[2] >= (1L << 26)) return MPDG_ERR_INVALID;  // 32-bit neighbour offsets
  hipLaunchKernelGGL(conv3_kernel, dim3(cnn_grid(ntiles)), dim3(256), 0, s, in, out, W, bias, dims[0], dims[1],
                     dims[2], mode, mask);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_wgrad(const unsigned short *g, const unsigned short *in, const int *dims, float *part, int nblocks, float *dW,
              float *db, hipStream_t s) {
  hipLaunchKernelGGL(wgrad_kernel, dim3(nblocks), dim3(256), 0, s, g, in, dims[0], dims[1], dims[2], part);
  hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((27 * 256 + 16 + 255) / 256), dim3(256), 0, s, part, 4 * nblocks, dW,
                     db);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_pack(const float *master, int cin, int cout, unsigned short *W, unsigned short *Wt, hipStream_t s) {
  hipLaunchKernelGGL(pack_weights_kernel, dim3(27), dim3(256), 0, s, master, cin, cout, W, Wt);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_adam(float *w, float *m, float *v, const float *dW, const float *db, float *b, float *mb, float *vb, int cin,
             int cout, float lr, float b1, float b2, float eps, int step, float gscale, hipStream_t s) {
  const int n = cout * cin * 27 + cout;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w, m, v, dW, db, b, mb, vb, cin, cout, lr,
                     b1, b2, eps, c1, c2, gscale);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_stats(const double *Tnm1, const double *Tn, const double *z, long n, double *part, unsigned *counter,
              double *stats, hipStream_t s) {
  const int nb = stats_blocks(n);
  hipLaunchKernelGGL(stats_kernel, dim3(nb), dim3(256), 0, s, Tnm1, Tn, z, n, part, counter, stats);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_scale3(const CnnFields &F, const int *dims, const double *sx, const double *sy, const double *sz,
               hipStream_t s, const int *stat_fields, double *part, unsigned *counter, double *stats) {
  if (F.nf < 1 || F.nf > 4) return MPDG_ERR_INVALID;
  const long n = (long)dims[0] * dims[1] * dims[2];
  if (n >= (1L << 26)) return MPDG_ERR_INVALID;  // 32-bit point indices, as the convolution
  if (stat_fields) {
    const int ia = stat_fields[0], it = stat_fields[1], iz = stat_fields[2];
    if (ia < 0 || ia >= F.nf || it < 0 || it >= F.nf || iz >= F.nf || !part || !counter || !stats)
      return MPDG_ERR_INVALID;
    hipLaunchKernelGGL(scale3_kernel<true>, dim3(stats_blocks(n)), dim3(256), 0, s,
                       F, dims[0], dims[1], dims[2], sx, sy, sz, ia, it, iz, part, counter, stats);
  } else {
    hipLaunchKernelGGL(scale3_kernel<false>, dim3((int)std::min<long>(8192, (n + 255) / 256)), dim3(256), 0, s, F,
                       dims[0], dims[1], dims[2], sx, sy, sz, 0, 0, -1, nullptr, nullptr, nullptr);
  }
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_features(const double *Tnm1, const double *Tn, const double *q, const double *z, long n,
                 const double *stats, double qscale, unsigned short *x, hipStream_t s) {
  hipLaunchKernelGGL(features_kernel, dim3((int)std::min<long>(8192, (n + 255) / 256)), dim3(256), 0, s, Tnm1, Tn, q,
                     z, n, stats, qscale, x);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_loss(const unsigned short *y, const double *Tnm1, const double *Tn, const double *base, const double *Tnp1,
             long n, const double *stats, double oscale, float gain, unsigned short *g, double *part, int nblocks,
             hipStream_t s) {
  hipLaunchKernelGGL(loss_kernel, dim3(nblocks), dim3(256), 0, s, y, Tnm1, Tn, base, Tnp1, n, stats, oscale, gain, g,
                     part);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_correction(const unsigned short *y, long n, const double *stats, double oscale, double *corr, hipStream_t s) {
  hipLaunchKernelGGL(correction_kernel, dim3((int)std::min<long>(8192, (n + 255) / 256)), dim3(256), 0, s, y, n, stats,
                     oscale, corr);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int cnn_extrapolate_add(const double *Tnm1, const double *Tn, const double *base, const double *corr, long n,
                        double *out, hipStream_t s) {
  hipLaunchKernelGGL(extrapolate_add_kernel, dim3((int)std::min<long>(8192, (n + 255) / 256)), dim3(256), 0, s, Tnm1,
                     Tn, base, corr, n, out);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

}  // namespace mpdg
