// Matrix-free apply of the implicit heat operator J = c_m M + K_kappa on a
// structured Q_k hex mesh, for CDNA4 (gfx950).
//
// Replaces HeatDiffuseMultiPhaseOperator::tangent_cell_loop /
// tangent_local_cell_operation (source/heat/heat_diffuse_operator.cpp:201-241,
// :957-1086) run through deal.II's MatrixFree::loop + FEEvaluation.
//
// Work decomposition
//   * A workgroup (NW wavefronts, 4 by default) owns a column of BY x BZ cells in
//     y-z and walks it along x in chunks of BX cells. Each cell is handled by
//     N1 = k+1 consecutive lanes of one wavefront; lane l of a cell holds the
//     z-plane l of the cell's DoFs (N1 x N1 values in registers).
//   * Sum factorisation: the x and y sweeps are register-local in this "plane"
//     layout; the z sweeps run after a wavefront-local transpose through LDS
//     into the "y-plane" layout (lane s holds qy = s). Gradients use the
//     collocation derivative on the Gauss points (n_q = k+1).
//   * Scatter: cell results are summed into an LDS copy of the chunk's DoF
//     brick (ds_add). The brick is then written once: DoFs interior to the
//     column are stored, DoFs on the column's y/z faces (shared with the
//     neighbouring columns) are added with global atomics, and the x-face
//     between consecutive chunks stays in LDS and is carried to the next chunk,
//     so no DoF row is split across x. A pre-pass zeroes the shared face rows.
//   * Determinism: a face row gets at most two atomic adds onto zero, and
//     0 + a + b does not depend on the order. Where four columns meet, the
//     upper z-layer of columns adds into a separate edge buffer instead, which
//     a small follow-up kernel adds into dst (and clears), so vmult is
//     bitwise reproducible.
#include "kernels.hpp"

namespace mpdg {

// cells of a chunk along x, y, z and wavefronts per workgroup (NW * (64 / N1) cells per chunk),
// chosen by measurement of the apply on large meshes (Q2 64^3, Q3 128^3, Q4 96^3). Longer chunks
// halve the scatter barriers per column and lengthen the coalesced x rows of the write-out; flat
// columns (BZ = 1) add face rows on every z plane; one-, two- and eight-wavefront workgroups were
// slower everywhere (one wavefront, Q3 4x2x2: 0.70 vs 0.64 ms; its barriers are free, but its
// 2x2-cell columns share 36 % of their rows with neighbours). Q3 (FP64 per-qp): 4x4x4 0.694 ms, 16x2x2 0.652, 8x4x2 0.645, 8x2x4 0.667,
// 16x4x1 0.878, 8x8x1 0.945; FP32 Q3 (4 wavefronts per SIMD with constant k): 16x2x2 is 1-1.5 %
// faster than 8x4x2 for every conductivity mode, 8x2x4 equal, 4x4x4 9 % slower, 8x4x2 with 4x2x2-cell
// wavefront boxes (40.3 KB of LDS: four workgroups per CU instead of three) within 2 %. Q2: 14x3x2 beats
// 7x4x3 by 5 % (FP64) and 25 % (FP32). Q4: FP64 keeps 4x4x3 (the longer shapes are 1-4 % slower),
// FP32 gains 11 % with 12x2x2.
// WX x WY x WZ: how the chunk is split into the NW wavefronts' boxes of cells (each box holds
// the 64 / N1 cells of one wavefront and has its own sub-brick of DoF sums, see ApplyGeom).
template <int N1, typename Number> struct ChunkShape;
template <typename Number> struct ChunkShape<2, Number> {
  static constexpr int BX = 8, BY = 4, BZ = 4, NW = 4, WX = 1, WY = 1, WZ = 4;
};
template <typename Number> struct ChunkShape<3, Number> {
  static constexpr int BX = 14, BY = 3, BZ = 2, NW = 4, WX = 2, WY = 1, WZ = 2;
};
template <> struct ChunkShape<4, double> { static constexpr int BX = 8, BY = 4, BZ = 2, NW = 4, WX = 1, WY = 2, WZ = 2; };
template <> struct ChunkShape<4, float> { static constexpr int BX = 16, BY = 2, BZ = 2, NW = 4, WX = 1, WY = 2, WZ = 2; };
template <> struct ChunkShape<5, double> { static constexpr int BX = 4, BY = 4, BZ = 3, NW = 4, WX = 1, WY = 4, WZ = 1; };
template <> struct ChunkShape<5, float> { static constexpr int BX = 12, BY = 2, BZ = 2, NW = 4, WX = 1, WY = 2, WZ = 2; };
// mapped cells (metric per point): FP64 Q3 runs 8 % faster with 16x2x2 than with the Cartesian
// 8x4x2 (Q3 96^3, 1.02 vs 1.11 ms), where 8x4x2 stays 2 % ahead
template <int N1, typename Number, bool MAPPED> struct ChunkShapeOf { using type = ChunkShape<N1, Number>; };
template <> struct ChunkShapeOf<4, double, true> {
  struct type { static constexpr int BX = 16, BY = 2, BZ = 2, NW = 4, WX = 1, WY = 2, WZ = 2; };
};

// Brick rows per batch of the write-out, whose LDS reads are all issued before any of its sums
// (0: one row at a time, each read waited for in turn). Batches pay where the write-out's LDS
// latency is exposed: the FP64 Q4 apply (one wavefront per SIMD) runs 6 % faster with batches
// of 5 rows; at Q3 they change nothing in FP64 and cost FP32 its occupancy (spills).
template <int N1, typename Number, bool MAPPED> struct WriteOutBatch { static constexpr int value = 0; };
template <> struct WriteOutBatch<5, double, false> { static constexpr int value = 5; };

template <int N1, typename Number = double, bool MAPPED = false> struct ApplyGeom {
  using Shape = typename ChunkShapeOf<N1, Number, MAPPED>::type;
  static constexpr int K = N1 - 1;
  static constexpr int CPW = 64 / N1;                 // cells per wavefront
  static constexpr int NW = Shape::NW;                // wavefronts per workgroup
  static constexpr int NT = 64 * NW;                  // threads per workgroup
  static constexpr int CPB = NW * CPW;                // cells per chunk
  static constexpr int BX = Shape::BX, BY = Shape::BY, BZ = Shape::BZ;
  static_assert(BX * BY * BZ == CPB, "chunk shape must match the cells per workgroup");
  static constexpr int SQY = N1 * N1 + 1;             // padded qy stride of a cell's transpose slot
  static constexpr int CS = N1 * SQY;                 // transpose slot per cell
  static constexpr int AX = K * BX + 1, AY = K * BY + 1, AZ = K * BZ + 1;  // DoF brick of a chunk
  // a wavefront's box of cells and its sub-brick of DoFs; neighbouring sub-bricks overlap in
  // one plane, which the write-out sums in a fixed order
  static constexpr int WX = Shape::WX, WY = Shape::WY, WZ = Shape::WZ;
  static_assert(WX * WY * WZ == NW && BX % WX == 0 && BY % WY == 0 && BZ % WZ == 0, "wavefront grid");
  static constexpr int EX = BX / WX, EY = BY / WY, EZ = BZ / WZ;
  static_assert(EX * EY * EZ == CPW, "a wavefront's box holds its cells");
  static constexpr int SX = K * EX + 1, SY = K * EY + 1, SZ = K * EZ + 1;
  static constexpr int SUB = SX * SY * SZ;
  static constexpr int ACC = NW * SUB;                // the sub-bricks of all wavefronts
  static constexpr int TBUF = NW * CPW * CS;          // transpose buffers of the wavefronts
  // the DoF sums are kept in FP64 at every precision (LDS FP64 adds are fast on CDNA4, FP32
  // LDS atomics are not) and have their own LDS, apart from the transpose buffers, so that
  // only the scatter needs workgroup barriers
  static constexpr int TBUF_DOUBLES = (TBUF * (int)sizeof(Number) + 7) / 8;
  static constexpr int CARRY = AY * AZ;
  static constexpr int ROWS = AY * AZ;                 // x rows of the DoF brick
  // LDS: transpose slots, sub-bricks, two carry planes, a zero entry, then per brick row its
  // global offset (8 B) and its write-out code (4 B)
  static constexpr int LDS_BYTES = 8 * (TBUF_DOUBLES + ACC + 2 * CARRY + 1 + ROWS) + 4 * ROWS;
  // the zero entry, relative to the sub-bricks: what the write-out reads (and clears) for the
  // sub-brick entries a brick position does not have
  static constexpr int ZERO_ENTRY = ACC + 2 * CARRY;
  // x rows of the DoF brick written per pass of the write-out loop (lane = x in the row)
  static constexpr int ROWS_PER_PASS = NT / AX;
  // brick rows per thread, and per batch of the write-out (all LDS reads of a batch are issued
  // before any of its sums)
  static constexpr int ROWS_PER_THREAD = (ROWS + ROWS_PER_PASS - 1) / ROWS_PER_PASS;
  static constexpr int WO_BATCH = ROWS_PER_THREAD < WriteOutBatch<N1, Number, MAPPED>::value
                                     ? ROWS_PER_THREAD : WriteOutBatch<N1, Number, MAPPED>::value;
  static constexpr int WO_ENTRIES = (WX > 1 ? 2 : 1) * (WY > 1 ? 2 : 1) * (WZ > 1 ? 2 : 1);
};

// ---- register-level 1D contractions on an N1 x N1 array a[row][col] ----
// along columns (second index): out[r][q] = sum_i M[q][i] in[r][i]
template <int N1, bool TR, typename Number, typename Mat>
__device__ __forceinline__ void sweep_col(const Mat &M, const Number (&in)[N1][N1],
                                          Number (&out)[N1][N1]) {
#pragma unroll
  for (int r = 0; r < N1; ++r)
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      Number acc = 0;
#pragma unroll
      for (int i = 0; i < N1; ++i) acc = fma(TR ? M(i, q) : M(q, i), in[r][i], acc);
      out[r][q] = acc;
    }
}
// along rows (first index): out[q][c] = sum_i M[q][i] in[i][c]
template <int N1, bool TR, typename Number, typename Mat>
__device__ __forceinline__ void sweep_row(const Mat &M, const Number (&in)[N1][N1],
                                          Number (&out)[N1][N1]) {
#pragma unroll
  for (int q = 0; q < N1; ++q)
#pragma unroll
    for (int c = 0; c < N1; ++c) {
      Number acc = 0;
#pragma unroll
      for (int i = 0; i < N1; ++i) acc = fma(TR ? M(i, q) : M(q, i), in[i][c], acc);
      out[q][c] = acc;
    }
}
template <int N1, bool TR, typename Number, typename Mat>
__device__ __forceinline__ void sweep_row_add(const Mat &M, const Number (&in)[N1][N1],
                                              Number (&out)[N1][N1]) {
#pragma unroll
  for (int q = 0; q < N1; ++q)
#pragma unroll
    for (int c = 0; c < N1; ++c) {
      Number acc = out[q][c];
#pragma unroll
      for (int i = 0; i < N1; ++i) acc = fma(TR ? M(i, q) : M(q, i), in[i][c], acc);
      out[q][c] = acc;
    }
}
template <int N1, bool TR, typename Number, typename Mat>
__device__ __forceinline__ void sweep_col_add(const Mat &M, const Number (&in)[N1][N1],
                                              Number (&out)[N1][N1]) {
#pragma unroll
  for (int r = 0; r < N1; ++r)
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      Number acc = out[r][q];
#pragma unroll
      for (int i = 0; i < N1; ++i) acc = fma(TR ? M(i, q) : M(q, i), in[r][i], acc);
      out[r][q] = acc;
    }
}

// Wavefront-local LDS ordering: all lanes' LDS writes are complete and visible
// to the other lanes of the same wavefront.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// plane layout (lane = z-plane l, a[qy][qx]) -> y-plane layout (lane = qy s, b[qz][qx])
template <int N1, typename Number>
__device__ __forceinline__ void transpose_zy(Number *slot, int l, bool on, const Number (&a)[N1][N1],
                                             Number (&b)[N1][N1]) {
  using G = ApplyGeom<N1>;
  wave_lds_sync();
  if (on) {
#pragma unroll
    for (int qy = 0; qy < N1; ++qy)
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) slot[l + N1 * qx + G::SQY * qy] = a[qy][qx];
  }
  wave_lds_sync();
  if (on) {
#pragma unroll
    for (int qz = 0; qz < N1; ++qz)
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) b[qz][qx] = slot[qz + N1 * qx + G::SQY * l];
  }
}
// y-plane layout (lane = qy s, b[iz][qx]) -> plane layout (lane = z-plane l, a[qy][qx])
template <int N1, typename Number>
__device__ __forceinline__ void transpose_yz(Number *slot, int s, bool on, const Number (&b)[N1][N1],
                                             Number (&a)[N1][N1]) {
  using G = ApplyGeom<N1>;
  wave_lds_sync();
  if (on) {
#pragma unroll
    for (int iz = 0; iz < N1; ++iz)
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) slot[iz + N1 * qx + G::SQY * s] = b[iz][qx];
  }
  wave_lds_sync();
  if (on) {
#pragma unroll
    for (int qy = 0; qy < N1; ++qy)
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) a[qy][qx] = slot[s + N1 * qx + G::SQY * qy];
  }
}

// write-out code of a brick row: low bits the row's y/z part of its sub-brick entry index
// (of the last box that holds it), then flags
constexpr int WO_PART = (1 << 16) - 1;
constexpr int WO_Y2 = 1 << 16;    // on a plane two boxes share in y (the other entry is one box back)
constexpr int WO_Z2 = 1 << 17;    // same in z
constexpr int WO_FACE = 1 << 18;  // on the column's y/z faces: atomic add
constexpr int WO_EDGE = 1 << 19;  // where four columns meet, upper layer: into the edge buffer
constexpr int WO_CON = 1 << 20;   // on a Dirichlet y/z plane
constexpr int WO_SKIP = 1 << 21;  // beyond the mesh (ragged column)

// The wavefront boxes along one direction: W boxes of E DoF intervals each (E = k times the
// cells of a box). Position p of the brick lies in box p / E at offset p % E, and a position
// on a shared plane (offset 0, not the first box) also in the previous box at offset E.
// Writes box * BOX_STRIDE + offset * STRIDE for each (lower box first) and returns their
// number (1 or 2; always 1 when W = 1).
template <int W, int E, int BOX_STRIDE, int STRIDE>
__device__ __forceinline__ int brick_parts(int p, int (&part)[2]) {
  if constexpr (W == 1) {
    part[0] = p * STRIDE;
    return 1;
  } else {
    const int b = min(p / E, W - 1), o = p - b * E;
    if (o == 0 && b > 0) {
      part[0] = (b - 1) * BOX_STRIDE + E * STRIDE;
      part[1] = b * BOX_STRIDE;
      return 2;
    }
    part[0] = b * BOX_STRIDE + o * STRIDE;
    return 1;
  }
}

template <typename Number>
__device__ __forceinline__ Number law_kappa(int law, const Number *p, Number T) {
  if (law == LAW_TANH) {
    return p[0] * (Number(1) + p[1] * tanh((T - p[2]) / p[3]));
  }
  // mushy zone (liquid_solid): p = {k_liquid, k_solid, T_solidus, T_liquidus}
  const Number inv = Number(1) / (p[3] - p[2]);
  Number fs = (p[3] - T) * inv;
  fs = fs < Number(0) ? Number(0) : (fs > Number(1) ? Number(1) : fs);
  return p[0] + (p[1] - p[0]) * (Number(3) - Number(2) * fs) * fs * fs;
}

// Gather the z-plane l of a cell (read_dof_values with homogeneous Dirichlet rows zeroed).
// Lanes without a cell read cell 0's plane (in bounds) and zero it afterwards, so the loads
// carry no per-element branches.
template <int N1, typename Number>
__device__ __forceinline__ void gather_plane(const Number *__restrict__ src, long base, bool valid, int dmask,
                                             const int *nd, int ix0, int iy0, int iz, long sy,
                                             Number (&u)[N1][N1]) {
  const Number *p = src + (valid ? base : 0);
#pragma unroll
  for (int j = 0; j < N1; ++j)
#pragma unroll
    for (int i = 0; i < N1; ++i) u[j][i] = p[i + sy * j];
  if (!valid) {
#pragma unroll
    for (int j = 0; j < N1; ++j)
#pragma unroll
      for (int i = 0; i < N1; ++i) u[j][i] = 0;
  }
  if (dmask) {
    const bool zlo = (dmask & 16) && iz == 0, zhi = (dmask & 32) && iz == nd[2] - 1;
#pragma unroll
    for (int j = 0; j < N1; ++j)
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        const int ix = ix0 + i, iy = iy0 + j;
        if (zlo || zhi || ((dmask & 1) && ix == 0) || ((dmask & 2) && ix == nd[0] - 1) ||
            ((dmask & 4) && iy == 0) || ((dmask & 8) && iy == nd[1] - 1))
          u[j][i] = 0;
      }
  }
}

// Per-quadrature-point conductivity of a cell at qy = l in the y-plane layout: kq[qz][qx] =
// kappa[cell][qz][l][qx], i.e. N1 contiguous values per qz (16-byte loads where they fit).
template <int N1, typename Number>
__device__ __forceinline__ void load_kappa_qp(const Number *__restrict__ kappa, long cell, int l, bool valid,
                                              Number (&kq)[N1][N1]) {
  constexpr int NQ = N1 * N1 * N1;
  const Number *p = kappa + (valid ? cell : 0) * NQ + N1 * l;
  constexpr bool vec = sizeof(Number) == 8 && N1 % 2 == 0;  // 2 doubles per load, 16-byte aligned rows
#pragma unroll
  for (int qz = 0; qz < N1; ++qz) {
    if constexpr (vec) {
      const double2 *r = reinterpret_cast<const double2 *>(p + N1 * N1 * qz);
#pragma unroll
      for (int h = 0; h < N1 / 2; ++h) {
        const double2 v = r[h];
        kq[qz][2 * h] = v.x;
        kq[qz][2 * h + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) kq[qz][qx] = p[N1 * N1 * qz + qx];
    }
  }
  if (!valid) {
#pragma unroll
    for (int qz = 0; qz < N1; ++qz)
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) kq[qz][qx] = 0;
  }
}

// wavefronts per SIMD the register allocation targets (1 = the compiler's choice). Mapped
// cells keep all three flux components and the metric live at the quadrature points and get
// the full register file rather than spill. The FP32 Q3 apply with a constant conductivity
// fits 128 VGPRs (2 spilled) and runs 4 % faster at 4 wavefronts than at 3; with per-point or
// law conductivity the same limit spills 13-18 VGPRs.
template <int N1, typename Number, bool MAPPED, int KMODE> struct Occupancy { static constexpr int W = 1; };
template <int KM> struct Occupancy<4, double, false, KM> { static constexpr int W = 2; };
template <int KM> struct Occupancy<4, float, false, KM> { static constexpr int W = 3; };
template <> struct Occupancy<4, float, false, KAPPA_CONST> { static constexpr int W = 4; };
template <int KM> struct Occupancy<4, float, true, KM> { static constexpr int W = 2; };
template <int KM> struct Occupancy<5, float, false, KM> { static constexpr int W = 2; };

template <int N1, typename Number, int KMODE, bool MAPPED>
__global__ __launch_bounds__((ApplyGeom<N1, Number, MAPPED>::NT)) __attribute__((amdgpu_waves_per_eu(Occupancy<N1, Number, MAPPED, KMODE>::W)))
void heat_apply_kernel(const ApplyArgs<Number, N1> P,
                                                         const Number *__restrict__ src,
                                                         Number *__restrict__ dst) {
  using G = ApplyGeom<N1, Number, MAPPED>;
  constexpr int K = G::K;
  // LDS (sized at launch, G::LDS_BYTES): the transpose slots, the wavefronts' sub-bricks and
  // the last x plane of the chunk, carried to the next one (two buffers by chunk parity, since
  // the write-out reads the previous chunk's carry while it writes its own). Sized at launch
  // rather than declared: with the size known, the compiler derives the occupancy from the
  // LDS and schedules for it (at Q3 FP64 it then spills 29 VGPRs; this way it needs 230).
  extern __shared__ double dyn_lds[];
  double *const lds_raw = dyn_lds;
  double *const acc = dyn_lds + G::TBUF_DOUBLES;
  double *const carry = acc + G::ACC;
  long *const row_off = reinterpret_cast<long *>(carry + 2 * G::CARRY + 1);
  int *const row_code = reinterpret_cast<int *>(row_off + G::ROWS);
  Number *const lds = reinterpret_cast<Number *>(lds_raw);  // transpose slots

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int cw = lane / N1, l = lane - cw * N1;
  const bool lane_on = (64 % N1 == 0) || cw < G::CPW;  // every lane holds a cell plane when N1 divides 64
  // the wavefront's box in the chunk, the cell's place in the box (cells beyond 64 / N1 lanes
  // are off)
  const int bx = wave % G::WX, by = (wave / G::WX) % G::WY, bz = wave / (G::WX * G::WY);
  const int ex = cw % G::EX, ey = (cw / G::EX) % G::EY, ez = cw / (G::EX * G::EY);
  const int lx = bx * G::EX + ex, ly = by * G::EY + ey, lz = bz * G::EZ + (lane_on ? ez : 0);
  double *const sub = acc + wave * G::SUB;  // this wavefront's sub-brick
  Number *slot = lds + (wave * G::CPW + (lane_on ? cw : 0)) * G::CS;

  // column of this workgroup (consecutive XCDs get consecutive column ranges)
  int col = blockIdx.x;
  if (P.xcd_cols > 0) col = (blockIdx.x & 7) * P.xcd_cols + (blockIdx.x >> 3);
  const int col_y = col % P.ncol[0], col_z = col / P.ncol[0];
  const int cy = col_y * G::BY + ly, cz = col_z * G::BZ + lz;
  const int Y0 = K * G::BY * col_y, Z0 = K * G::BZ * col_z;
  const int DY = K * min(G::BY, P.nc[1] - col_y * G::BY) + 1;
  const int DZ = K * min(G::BZ, P.nc[2] - col_z * G::BZ) + 1;
  const long sy = P.nd[0], sz = (long)P.nd[0] * P.nd[1];
  const int nchunks = (P.nc[0] + G::BX - 1) / G::BX;
  // the brick starts at zero; the write-out of every chunk zeroes what it reads
  for (int idx = threadIdx.x; idx < G::ACC; idx += G::NT) acc[idx] = 0.0;
  for (int idx = threadIdx.x; idx < 2 * G::CARRY + 1; idx += G::NT) carry[idx] = 0.0;  // and the zero entry
  // what the write-out needs of each x row (y, z) of the brick, the same for every chunk of
  // the column: its global offset (in dst, or in the edge buffer for the upper layer's rows
  // where four columns meet) and a code: y/z part of its sub-brick entry, whether it lies on a
  // plane two boxes share in y / z, on the column's faces, on a Dirichlet y/z plane, beyond the
  // mesh
  for (int r = threadIdx.x; r < G::ROWS; r += G::NT) {
    const int y = r % G::AY, z = r / G::AY;
    int code = WO_SKIP;
    long off = 0;
    if (y < DY && z < DZ) {
      int ye[2], ze[2];
      const int nye = brick_parts<G::WY, K * G::EY, G::WX * G::SUB, G::SX>(y, ye);
      const int nze = brick_parts<G::WZ, K * G::EZ, G::WX * G::WY * G::SUB, G::SX * G::SY>(z, ze);
      code = ye[nye - 1] + ze[nze - 1];
      if (nye == 2) code |= WO_Y2;
      if (nze == 2) code |= WO_Z2;
      const int iy = Y0 + y, iz = Z0 + z;
      if (y == 0 || y == DY - 1 || z == 0 || z == DZ - 1) code |= WO_FACE;
      if (((P.dmask & 4) && iy == 0) || ((P.dmask & 8) && iy == P.nd[1] - 1) || ((P.dmask & 16) && iz == 0) ||
          ((P.dmask & 32) && iz == P.nd[2] - 1))
        code |= WO_CON;
      off = sy * iy + sz * iz;
      // a row where four columns meet gets two contributions from each z-layer of columns; the
      // upper layer's go to the edge buffer (added by edge_fixup_kernel), so every row receives
      // at most two atomic adds onto zero, whose result does not depend on their order
      if (z == 0 && col_z > 0 && ((y == 0 && col_y > 0) || (y == DY - 1 && col_y + 1 < P.ncol[0]))) {
        code |= WO_EDGE;
        off = ((long)(col_z - 1) * (P.ncol[0] - 1) + (y == 0 ? col_y : col_y + 1) - 1) * P.nd[0];
      }
    }
    row_off[r] = off;
    row_code[r] = code;
  }

  // the src plane of the next chunk is loaded while the current one is scattered and written
  Number nxt[N1][N1];
  {
    const int cx = lx;
    const bool valid = lane_on && cx < P.nc[0] && cy < P.nc[1] && cz < P.nc[2];
    const long base = (long)K * cx + sy * (K * cy) + sz * (K * cz + l);
    gather_plane<N1>(src, base, valid, P.dmask, P.nd, K * cx, K * cy, K * cz + l, sy, nxt);
  }

  for (int chunk = 0; chunk < nchunks; ++chunk) {
    const int cx = chunk * G::BX + lx;
    const bool valid = lane_on && cx < P.nc[0] && cy < P.nc[1] && cz < P.nc[2];
    const long cell = (long)cx + (long)P.nc[0] * (cy + (long)P.nc[1] * cz);
    const long base = (long)K * cx + sy * (K * cy) + sz * (K * cz + l);

    Number u[N1][N1];
#pragma unroll
    for (int j = 0; j < N1; ++j)
#pragma unroll
      for (int i = 0; i < N1; ++i) u[j][i] = nxt[j][i];

    // ---- conductivity at the quadrature points in the y-plane layout ----
    Number kq[N1][N1];
    if (KMODE == KAPPA_QP) load_kappa_qp<N1>(P.kappa_arr, cell, l, valid, kq);
    if (KMODE == KAPPA_LAW) {
      Number T0[N1][N1], w1[N1][N1], t1[N1][N1];
#pragma unroll
      for (int j = 0; j < N1; ++j)
#pragma unroll
        for (int i = 0; i < N1; ++i) T0[j][i] = valid ? P.T[base + i + sy * j] : Number(0);
      sweep_col<N1, false>(P.S, T0, w1);
      sweep_row<N1, false>(P.S, w1, t1);
      transpose_zy<N1>(slot, l, lane_on, t1, w1);
      sweep_row<N1, false>(P.S, w1, T0);
#pragma unroll
      for (int qz = 0; qz < N1; ++qz)
#pragma unroll
        for (int qx = 0; qx < N1; ++qx) kq[qz][qx] = law_kappa<Number>(P.law, P.law_p, T0[qz][qx]);
    }

    // ---- forward, plane layout: t = Sy Sx u, b = Dy t ----
    Number t[N1][N1], b[N1][N1];
    {
      Number w1[N1][N1];
      sweep_col<N1, false>(P.S, u, w1);
      sweep_row<N1, false>(P.S, w1, t);
      sweep_row<N1, false>(P.D, t, b);
    }
    // ---- y-plane layout (lane = qy): v = Sz t, gy = Sz b ----
    Number v[N1][N1], gy[N1][N1];
    transpose_zy<N1>(slot, l, lane_on, t, u);
    sweep_row<N1, false>(P.S, u, v);
    transpose_zy<N1>(slot, l, lane_on, b, u);
    sweep_row<N1, false>(P.S, u, gy);

    // ---- quadrature point operation (heat_diffuse_operator.cpp:1029-1077):
    //      val = c_m u and flux = kappa grad u at every point, then the z and x parts of the
    //      integration right away: e = c W v + Dz^T(kW gz) + Dx^T(kW gx), in place of v ----
    const Number wy = P.w[l];
    Number kc = valid ? P.kappa : Number(0);  // per lane: keeps the constant case on the cell case's schedule
    if (KMODE == KAPPA_CELL) kc = valid ? P.kappa_arr[cell] : Number(0);
    Number e[N1][N1];
    if (MAPPED) {
      // all three reference gradients at a point are needed together
      Number gx[N1][N1], gz[N1][N1];
      sweep_row<N1, false>(P.D, v, gz);
      sweep_col<N1, false>(P.D, v, gx);
#pragma unroll
      for (int qz = 0; qz < N1; ++qz)
#pragma unroll
        for (int qx = 0; qx < N1; ++qx) {
          constexpr int NQ = N1 * N1 * N1;
          const int q = qx + N1 * (l + N1 * qz);
          Number kap = kc;
          if (KMODE == KAPPA_QP || KMODE == KAPPA_LAW) kap = kq[qz][qx];
          // flux in reference coordinates: kappa G grad_xi u, G = JxW J^-1 J^-T
          const Number *M = P.metric + (valid ? cell : 0) * 7 * NQ + q;
          const Number jxw = valid ? M[0] : Number(0);
          const Number g00 = M[NQ], g01 = M[2 * NQ], g11 = M[3 * NQ], g02 = M[4 * NQ], g12 = M[5 * NQ],
                       g22 = M[6 * NQ];
          const Number a = gx[qz][qx], b = gy[qz][qx], c = gz[qz][qx];
          e[qz][qx] = P.cm * jxw * v[qz][qx];
          gx[qz][qx] = kap * (g00 * a + g01 * b + g02 * c);
          gy[qz][qx] = kap * (g01 * a + g11 * b + g12 * c);
          gz[qz][qx] = kap * (g02 * a + g12 * b + g22 * c);
        }
      sweep_row_add<N1, true>(P.D, gz, e);
      sweep_col_add<N1, true>(P.D, gx, e);
    } else {
      // Weights and conductivity folded once into kw (the registers of kq); then the z
      // direction column by column initialises e, and the x direction row by row adds the
      // x flux and the mass term and scales the y flux. v dies row by row in the second
      // pass, and no weight products stay live between the passes (the mass term's weight
      // is formed in a different order on purpose).
      // per-point weight (times the per-point conductivity), in the registers of kq; a
      // cell-wise conductivity goes into the direction factors instead
      Number kx = P.ks[0], ky = P.ks[1], kz = P.ks[2];
      if (KMODE == KAPPA_CONST || KMODE == KAPPA_CELL) {
        kx *= kc;
        ky *= kc;
        kz *= kc;
      }
#pragma unroll
      for (int qz = 0; qz < N1; ++qz)
#pragma unroll
        for (int qx = 0; qx < N1; ++qx) {
          const Number W = (wy * P.w[qz]) * P.w[qx];
          kq[qz][qx] = (KMODE == KAPPA_QP || KMODE == KAPPA_LAW) ? kq[qz][qx] * W : W;
        }
      auto kw = [&](int qz, int qx) -> Number { return kq[qz][qx]; };
#pragma unroll
      for (int qx = 0; qx < N1; ++qx) {
        Number f[N1];
#pragma unroll
        for (int qz = 0; qz < N1; ++qz) {
          Number g = 0;
#pragma unroll
          for (int p = 0; p < N1; ++p) g = fma(P.D(qz, p), v[p][qx], g);
          f[qz] = kw(qz, qx) * kz * g;
        }
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          Number acc = 0;
#pragma unroll
          for (int qz = 0; qz < N1; ++qz) acc = fma(P.D(qz, i), f[qz], acc);
          e[i][qx] = acc;
        }
      }
      // y flux: scale, then take it back to the plane layout right away (ey = Sz^T (kW gy))
#pragma unroll
      for (int qz = 0; qz < N1; ++qz)
#pragma unroll
        for (int qx = 0; qx < N1; ++qx) gy[qz][qx] *= kw(qz, qx) * ky;
      sweep_row<N1, true>(P.S, gy, u);
      transpose_yz<N1>(slot, l, lane_on, u, b);
      const Number cmwy = P.cm * wy;
#pragma unroll
      for (int qz = 0; qz < N1; ++qz) {
        Number f[N1];
#pragma unroll
        for (int qx = 0; qx < N1; ++qx) {
          Number g = 0;
#pragma unroll
          for (int p = 0; p < N1; ++p) g = fma(P.D(qx, p), v[qz][p], g);
          f[qx] = kw(qz, qx) * kx * g;
        }
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          Number acc = fma((cmwy * P.w[i]) * P.w[qz], v[qz][i], e[qz][i]);
#pragma unroll
          for (int qx = 0; qx < N1; ++qx) acc = fma(P.D(qx, i), f[qx], acc);
          e[qz][i] = acc;
        }
      }
    }
    // ---- back to the plane layout: e' = Sz^T e, ey = Sz^T (kW gy) (the latter already in b
    //      for affine cells) ----
    sweep_row<N1, true>(P.S, e, u);
    transpose_yz<N1>(slot, l, lane_on, u, t);
    if (MAPPED) {
      sweep_row<N1, true>(P.S, gy, u);
      transpose_yz<N1>(slot, l, lane_on, u, b);
    }
    // ---- plane layout: out = Sx^T Sy^T (e' + Dy^T ey) ----
    sweep_row_add<N1, true>(P.D, b, t);
    sweep_row<N1, true>(P.S, t, b);
    sweep_col<N1, true>(P.S, b, u);
    if (chunk + 1 < nchunks) {
      const int cxn = cx + G::BX;
      const bool vn = lane_on && cxn < P.nc[0] && cy < P.nc[1] && cz < P.nc[2];
      gather_plane<N1>(src, base + (long)K * G::BX, vn, P.dmask, P.nd, K * cxn, K * cy, K * cz + l, sy, nxt);
    }

    // ---- scatter: sum the chunk's cells in the LDS DoF brick (zero, after the previous
    //      chunk's write-out, which the barrier orders before these adds) ----
    //      sub-brick; one wavefront's LDS adds run in program order and no two lanes of one add
    //      meet at one address, so every sum has a fixed order)
    __syncthreads();
    const int X0 = K * G::BX * chunk;
    const int DX = K * min(G::BX, P.nc[0] - chunk * G::BX) + 1;
    if (valid) {
#pragma unroll
      for (int j = 0; j < N1; ++j)
#pragma unroll
        for (int i = 0; i < N1; ++i)
          atomicAdd(&sub[(K * ex + i) + G::SX * ((K * ey + j) + G::SY * (K * ez + l))], (double)u[j][i]);
    }
    __syncthreads();
    const bool last = chunk == nchunks - 1;
    // write-out by x rows of the brick: lane wx of row group wr (coalesced along x). A brick
    // position lies in one sub-brick, or on a plane two neighbouring boxes share (then in two
    // per direction); its entries are summed in a fixed order (lower boxes first) and cleared.
    // The entry index splits into an x part (per lane) and the row's y/z part (row_code).
    const int wr = tid / G::AX, wx = tid - wr * G::AX;
    if (wr < G::ROWS_PER_PASS && wx < DX) {
      const int ix = X0 + wx;
      int xe[2] = {0, 0};
      const int nxe = brick_parts<G::WX, K * G::EX, G::SUB, 1>(wx, xe);
      const bool con_x = ((P.dmask & 1) && ix == 0) || ((P.dmask & 2) && ix == P.nd[0] - 1);
      const bool carry_out = wx == DX - 1 && !last, carry_in = wx == 0 && chunk > 0;
      double *const carry_rd = carry + ((chunk + 1) & 1) * G::CARRY;
      double *const carry_wr = carry + (chunk & 1) * G::CARRY;
      if constexpr (G::WO_BATCH > 0) {
        // the previous box's entry in y / z: one box back, at the far end of its sub-brick
        constexpr int DYP = -G::WX * G::SUB + K * G::EY * G::SX;
        constexpr int DZP = -G::WX * G::WY * G::SUB + K * G::EZ * G::SX * G::SY;
        constexpr int RB = G::WO_BATCH, NE = G::WO_ENTRIES;
#pragma unroll
        for (int k0 = 0; k0 < G::ROWS_PER_THREAD; k0 += RB) {
          // 1. the rows' codes and offsets, 2. their sub-brick entries (lower boxes first; an
          // absent one reads the zero entry) and carries, 3. clear the entries, 4. sums in that
          // order and the global stores / adds. All LDS reads of a batch are in flight together.
          int code[RB], rr[RB];
          long off[RB];
#pragma unroll
          for (int k = 0; k < RB; ++k) {
            const int r = wr + (k0 + k) * G::ROWS_PER_PASS;
            rr[k] = min(r, G::ROWS - 1);
            const int c = row_code[rr[k]];
            code[k] = (k0 + k < G::ROWS_PER_THREAD && r < G::ROWS) ? c : WO_SKIP;
            off[k] = row_off[rr[k]];
          }
          int ent[RB][NE];
#pragma unroll
          for (int k = 0; k < RB; ++k) {
            const int c = code[k];
            const int yz = c & WO_PART, y2 = (c & WO_Y2) ? 1 : 0, z2 = (c & WO_Z2) ? 1 : 0;
            const bool live = !(c & WO_SKIP);
            int i = 0;
#pragma unroll
            for (int cz = 0; cz < (G::WZ > 1 ? 2 : 1); ++cz)
#pragma unroll
              for (int by = 0; by < (G::WY > 1 ? 2 : 1); ++by)
#pragma unroll
                for (int ax = 0; ax < (G::WX > 1 ? 2 : 1); ++ax, ++i) {
                  const bool present = live && cz <= z2 && by <= y2 && ax < nxe;
                  ent[k][i] = present ? xe[ax] + yz + (cz < z2 ? DZP : 0) + (by < y2 ? DYP : 0) : G::ZERO_ENTRY;
                }
          }
          double v[RB][NE], cr[RB];
#pragma unroll
          for (int k = 0; k < RB; ++k) {
            const double c = carry_rd[rr[k]];
            cr[k] = carry_in ? c : 0.0;
#pragma unroll
            for (int i = 0; i < NE; ++i) v[k][i] = acc[ent[k][i]];
          }
#pragma unroll
          for (int k = 0; k < RB; ++k)
#pragma unroll
            for (int i = 0; i < NE; ++i) acc[ent[k][i]] = 0.0;
#pragma unroll
          for (int k = 0; k < RB; ++k) {
            if (code[k] & WO_SKIP) continue;
            double sum = cr[k];
#pragma unroll
            for (int i = 0; i < NE; ++i) sum += v[k][i];
            if (carry_out) {
              carry_wr[rr[k]] = sum;
              continue;
            }
            const Number val = (Number)sum;
            const bool con = con_x || (code[k] & WO_CON);
            const long g = off[k] + ix;
            if (code[k] & WO_FACE) {
              if (!con) atomicAdd(((code[k] & WO_EDGE) ? P.edge : dst) + g, val);
            } else {
              dst[g] = con ? Number(0) : val;
            }
          }
        }
      } else {
        for (int r = wr; r < G::ROWS; r += G::ROWS_PER_PASS) {
          const int code = row_code[r];
          if (code & WO_SKIP) continue;
          const int yz = code & WO_PART;
          // the previous box's entry in y / z: one box back, at the far end of its sub-brick
          constexpr int DYP = -G::WX * G::SUB + K * G::EY * G::SX;
          constexpr int DZP = -G::WX * G::WY * G::SUB + K * G::EZ * G::SX * G::SY;
          const int y2 = (code & WO_Y2) ? 1 : 0, z2 = (code & WO_Z2) ? 1 : 0;
          double sum = carry_in ? carry_rd[r] : 0.0;
#pragma unroll
          for (int c = 0; c < (G::WZ > 1 ? 2 : 1); ++c)
#pragma unroll
            for (int b = 0; b < (G::WY > 1 ? 2 : 1); ++b)
#pragma unroll
              for (int a = 0; a < (G::WX > 1 ? 2 : 1); ++a)
                if (c <= z2 && b <= y2 && a < nxe) {
                  const int e = xe[a] + yz + (c < z2 ? DZP : 0) + (b < y2 ? DYP : 0);
                  sum += acc[e];
                  acc[e] = 0.0;
                }
          if (carry_out) {
            carry_wr[r] = sum;
            continue;
          }
          const Number val = (Number)sum;
          const bool con = con_x || (code & WO_CON);
          const long g = row_off[r] + ix;
          if (code & WO_FACE) {
            if (!con) atomicAdd(((code & WO_EDGE) ? P.edge : dst) + g, val);
          } else {
            dst[g] = con ? Number(0) : val;
          }
        }
      }
    }
  }
}

// Zero the DoF rows shared between neighbouring columns: the x rows on the
// y-face planes (iy a multiple of py, and the last plane) and on the z-face
// planes. One workgroup per row; rows on both a y and a z face are zeroed twice.
template <typename Number>
__global__ __launch_bounds__(128) void zero_face_rows_kernel(Number *__restrict__ dst, int nd0, int nd1,
                                                             int nd2, int py, int pz, int nfy, int nfz) {
  long b = blockIdx.x;
  int iy, iz;
  if (b < (long)nfy * nd2) {
    iy = min((int)(b % nfy) * py, nd1 - 1);
    iz = (int)(b / nfy);
  } else {
    b -= (long)nfy * nd2;
    iz = min((int)(b % nfz) * pz, nd2 - 1);
    iy = (int)(b / nfz);
  }
  Number *row = dst + (long)nd0 * (iy + (long)nd1 * iz);
  for (int ix = threadIdx.x; ix < nd0; ix += 128) row[ix] = 0;
}

// dst += the upper layer's contributions to the rows where four columns meet (interior y-face
// plane fy and z-face plane fz of the column decomposition), and the edge buffer back to zero
template <typename Number>
__global__ __launch_bounds__(256) void edge_fixup_kernel(Number *__restrict__ dst, Number *__restrict__ edge, int nd0,
                                                        long sy, long sz, int py, int pz, int nly, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const long line = i / nd0;
  const int ix = (int)(i - line * nd0);
  const int fy = (int)(line % nly) + 1, fz = (int)(line / nly) + 1;
  const Number v = edge[i];
  edge[i] = Number(0);
  dst[ix + sy * ((long)fy * py) + sz * ((long)fz * pz)] += v;
}

template <int N1, typename Number, bool MAPPED>
size_t edge_elems_shape(const int *nc) {
  using G = ApplyGeom<N1, Number, MAPPED>;
  const long ny = (nc[1] + G::BY - 1) / G::BY, nz = (nc[2] + G::BZ - 1) / G::BZ;
  return (size_t)((ny - 1) * (nz - 1) * ((long)(N1 - 1) * nc[0] + 1));
}
template <int N1, typename Number>
size_t edge_elems_n1(const int *nc, int mapped) {
  const size_t c = edge_elems_shape<N1, Number, false>(nc), m = edge_elems_shape<N1, Number, true>(nc);
  return mapped < 0 ? std::max(c, m) : (mapped ? m : c);
}
size_t apply_edge_elems(int degree, const int *nc, bool fp32, int mapped) {
  switch (degree) {
    case 1: return fp32 ? edge_elems_n1<2, float>(nc, mapped) : edge_elems_n1<2, double>(nc, mapped);
    case 2: return fp32 ? edge_elems_n1<3, float>(nc, mapped) : edge_elems_n1<3, double>(nc, mapped);
    case 3: return fp32 ? edge_elems_n1<4, float>(nc, mapped) : edge_elems_n1<4, double>(nc, mapped);
    case 4: return fp32 ? edge_elems_n1<5, float>(nc, mapped) : edge_elems_n1<5, double>(nc, mapped);
    default: return 0;
  }
}

// LDS beyond 64 KiB per workgroup has to be allowed per kernel before its first launch
template <int N1, typename Number, int KMODE, bool MAPPED>
int allow_apply_lds() {
  using G = ApplyGeom<N1, Number, MAPPED>;
  if (G::LDS_BYTES <= 65536) return MPDG_OK;
  static const hipError_t rc = hipFuncSetAttribute(
      reinterpret_cast<const void *>(&heat_apply_kernel<N1, Number, KMODE, MAPPED>),
      hipFuncAttributeMaxDynamicSharedMemorySize, G::LDS_BYTES);
  return rc == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}
#define MPDG_APPLY_LDS_LIMIT(MAPPED_)                                                                \
  do {                                                                                             \
    int rc_ = MPDG_OK;                                                                             \
    if constexpr (MAPPED_) {                                                                       \
      rc_ = op.kappa_mode == KAPPA_CELL ? allow_apply_lds<N1, Number, KAPPA_CELL, true>()          \
                                        : allow_apply_lds<N1, Number, KAPPA_CONST, true>();        \
    } else {                                                                                       \
      switch (op.kappa_mode) {                                                                     \
        case KAPPA_QP: rc_ = allow_apply_lds<N1, Number, KAPPA_QP, false>(); break;                \
        case KAPPA_CELL: rc_ = allow_apply_lds<N1, Number, KAPPA_CELL, false>(); break;            \
        case KAPPA_LAW: rc_ = allow_apply_lds<N1, Number, KAPPA_LAW, false>(); break;              \
        default: rc_ = allow_apply_lds<N1, Number, KAPPA_CONST, false>(); break;                   \
      }                                                                                            \
    }                                                                                              \
    if (rc_ != MPDG_OK) return rc_;                                                                \
  } while (0)

template <int N1, typename Number, bool MAPPED>
int launch_apply_shape(const OperatorData<Number> &op, const Number *src, Number *dst, hipStream_t stream,
                       hipEvent_t kernel_start, hipEvent_t kernel_end) {
  using G = ApplyGeom<N1, Number, MAPPED>;
  ApplyArgs<Number, N1> A;
  fill_apply_args(op, A);
  A.ncol[0] = (op.nc[1] + G::BY - 1) / G::BY;
  A.ncol[1] = (op.nc[2] + G::BZ - 1) / G::BZ;
  const int ncols = A.ncol[0] * A.ncol[1];
  A.xcd_cols = (ncols % 8 == 0 && ncols >= 64) ? ncols / 8 : 0;
  const long n_edge = (long)(A.ncol[0] - 1) * (A.ncol[1] - 1) * op.nd[0];
  if (n_edge > 0 && (op.edge == nullptr || (size_t)n_edge > op.edge_elems)) return MPDG_ERR_INVALID;
  {
    // face planes j*p for j = 0 .. ceil((nd-1)/p)-1, then the last plane nd-1
    const int py = (N1 - 1) * G::BY, pz = (N1 - 1) * G::BZ;
    const int nfy = (op.nd[1] - 1 + py - 1) / py + 1, nfz = (op.nd[2] - 1 + pz - 1) / pz + 1;
    const long nrows = (long)nfy * op.nd[2] + (long)nfz * op.nd[1];
    hipLaunchKernelGGL(zero_face_rows_kernel<Number>, dim3((unsigned)nrows), dim3(128), 0, stream, dst,
                       op.nd[0], op.nd[1], op.nd[2], py, pz, nfy, nfz);
  }
  if (kernel_start) (void)hipEventRecord(kernel_start, stream);
  MPDG_APPLY_LDS_LIMIT(MAPPED);
  if constexpr (MAPPED) {
    switch (op.kappa_mode) {
      case KAPPA_CONST: hipLaunchKernelGGL((heat_apply_kernel<N1, Number, KAPPA_CONST, true>), dim3(ncols), dim3(G::NT), G::LDS_BYTES, stream, A, src, dst); break;
      case KAPPA_CELL: hipLaunchKernelGGL((heat_apply_kernel<N1, Number, KAPPA_CELL, true>), dim3(ncols), dim3(G::NT), G::LDS_BYTES, stream, A, src, dst); break;
      default: return MPDG_ERR_INVALID;  // per-qp conductivity on mapped meshes runs through metric_k
    }
  } else {
    switch (op.kappa_mode) {
      case KAPPA_CONST: hipLaunchKernelGGL((heat_apply_kernel<N1, Number, KAPPA_CONST, false>), dim3(ncols), dim3(G::NT), G::LDS_BYTES, stream, A, src, dst); break;
      case KAPPA_QP: hipLaunchKernelGGL((heat_apply_kernel<N1, Number, KAPPA_QP, false>), dim3(ncols), dim3(G::NT), G::LDS_BYTES, stream, A, src, dst); break;
      case KAPPA_CELL: hipLaunchKernelGGL((heat_apply_kernel<N1, Number, KAPPA_CELL, false>), dim3(ncols), dim3(G::NT), G::LDS_BYTES, stream, A, src, dst); break;
      case KAPPA_LAW: hipLaunchKernelGGL((heat_apply_kernel<N1, Number, KAPPA_LAW, false>), dim3(ncols), dim3(G::NT), G::LDS_BYTES, stream, A, src, dst); break;
      default: return MPDG_ERR_INVALID;
    }
  }
  if (kernel_end) (void)hipEventRecord(kernel_end, stream);
  if (n_edge > 0)
    hipLaunchKernelGGL(edge_fixup_kernel<Number>, dim3((unsigned)((n_edge + 255) / 256)), dim3(256), 0, stream, dst,
                       op.edge, op.nd[0], (long)op.nd[0], (long)op.nd[0] * op.nd[1], (N1 - 1) * G::BY,
                       (N1 - 1) * G::BZ, A.ncol[0] - 1, n_edge);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

template <int N1, typename Number>
int launch_apply_n1(const OperatorData<Number> &op_in, const Number *src, Number *dst, hipStream_t stream,
                    hipEvent_t kernel_start, hipEvent_t kernel_end) {
  const OperatorData<Number> op = kernel_view(op_in);
  if (op.metric == nullptr && op_in.metric != nullptr) return MPDG_ERR_INVALID;  // metric_k not computed
  if (op.metric) return launch_apply_shape<N1, Number, true>(op, src, dst, stream, kernel_start, kernel_end);
  return launch_apply_shape<N1, Number, false>(op, src, dst, stream, kernel_start, kernel_end);
}

template <typename Number>
int launch_apply(const OperatorData<Number> &op, const Number *src, Number *dst, hipStream_t stream,
                 hipEvent_t kernel_start, hipEvent_t kernel_end) {
  switch (op.degree) {
    case 1: return launch_apply_n1<2, Number>(op, src, dst, stream, kernel_start, kernel_end);
    case 2: return launch_apply_n1<3, Number>(op, src, dst, stream, kernel_start, kernel_end);
    case 3: return launch_apply_n1<4, Number>(op, src, dst, stream, kernel_start, kernel_end);
    case 4: return launch_apply_n1<5, Number>(op, src, dst, stream, kernel_start, kernel_end);
    default: return MPDG_ERR_INVALID;
  }
}

template int launch_apply<double>(const OperatorData<double> &, const double *, double *, hipStream_t, hipEvent_t,
                                  hipEvent_t);
template int launch_apply<float>(const OperatorData<float> &, const float *, float *, hipStream_t, hipEvent_t,
                                 hipEvent_t);

}  // namespace mpdg
