// Data formats on either side of the operator path:
//  * Paraview output as the reference writes it (Postprocessor::do_paraview_output,
//    source/post_processing/postprocessor.cpp:195-270): DataOut::build_patches(mapping, k)
//    subdivides every cell into k^3 equal sub-cells whose vertices carry the FE field's values,
//    one .vtu per output step and a .pvd record relating files to times
//    (DataOutBase::write_pvd_record). Fields are the heat operation's output vectors
//    (temperature, temperature_old, heat_source, heat_diffuse_operation.cpp:365-405).
//  * The DoF numbering: the library's lexicographic node grid vs. a DoFHandler's numbering,
//    matched through the support points (DoFTools::map_dofs_to_support_points), and the device
//    permutation of vectors between the two.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mpdg_heat.h"
#include "fe1d.hpp"

namespace {

using mpdg::FE1D;

// base64 of a byte range (RFC 4648), the encoding of VTK XML "binary" data arrays
void base64(const unsigned char *p, size_t n, std::string &out) {
  static const char *tab = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  out.reserve(out.size() + 4 * ((n + 2) / 3));
  size_t i = 0;
  for (; i + 2 < n; i += 3) {
    const unsigned v = (p[i] << 16) | (p[i + 1] << 8) | p[i + 2];
    out += tab[(v >> 18) & 63];
    out += tab[(v >> 12) & 63];
    out += tab[(v >> 6) & 63];
    out += tab[v & 63];
  }
  if (i < n) {
    const unsigned v = (p[i] << 16) | (i + 1 < n ? p[i + 1] << 8 : 0);
    out += tab[(v >> 18) & 63];
    out += tab[(v >> 12) & 63];
    out += i + 1 < n ? tab[(v >> 6) & 63] : '=';
    out += '=';
  }
}

// one DataArray payload: UInt64 byte count followed by the data, base64 encoded as one block
template <typename T>
bool write_array(FILE *f, const char *type, const char *name, int ncomp, const std::vector<T> &v) {
  std::vector<unsigned char> buf(8 + v.size() * sizeof(T));
  const uint64_t nbytes = v.size() * sizeof(T);
  std::memcpy(buf.data(), &nbytes, 8);
  if (!v.empty()) std::memcpy(buf.data() + 8, v.data(), nbytes);
  std::string enc;
  base64(buf.data(), buf.size(), enc);
  std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"binary\">\n", type,
               name, ncomp);
  const bool ok = std::fwrite(enc.data(), 1, enc.size(), f) == enc.size();
  std::fprintf(f, "\n        </DataArray>\n");
  return ok;
}

// values at the k+1 equidistant points of [0,1] of the Q_k Lagrange basis on Gauss-Lobatto nodes
void equidistant_table(const FE1D &fe, double *E /* [k+1][k+1] */) {
  const int n = fe.n1;
  for (int p = 0; p < n; ++p) {
    double v[mpdg::kMaxN1];
    FE1D::lagrange(n, fe.nodes, fe.k == 0 ? 0.0 : (double)p / fe.k, v, nullptr);
    for (int j = 0; j < n; ++j) E[p * n + j] = v[j];
  }
}

// u (lexicographic nodes of the whole mesh) -> values at the equidistant patch vertices: the
// tensor product of the per-cell 1D interpolation, applied direction by direction (cells share
// their end points, where both sides give the node value)
void to_patch_vertices(const FE1D &fe, const int *nd, const double *u, std::vector<double> &out) {
  const int k = fe.k, n = fe.n1;
  std::vector<double> E(n * n);
  equidistant_table(fe, E.data());
  std::vector<double> a(u, u + (size_t)nd[0] * nd[1] * nd[2]), b(a.size());
  for (int dir = 0; dir < 3; ++dir) {
    const long stride = dir == 0 ? 1 : (dir == 1 ? nd[0] : (long)nd[0] * nd[1]);
    const long len = nd[dir];
    const long lines = (long)a.size() / len;
    for (long l = 0; l < lines; ++l) {
      long base;
      if (dir == 0) base = l * len;
      else if (dir == 1) base = (l % nd[0]) + (l / nd[0]) * (long)nd[0] * nd[1];
      else base = l;
      for (long g = 0; g < len; ++g) {
        const long c = std::min<long>(g / k, (len - 1) / k - 1), p = g - k * c;
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += E[p * n + j] * a[base + (k * c + j) * stride];
        b[base + g * stride] = s;
      }
    }
    a.swap(b);
  }
  out.swap(a);
}

__global__ __launch_bounds__(256) void permute_kernel(const long *__restrict__ perm, const double *__restrict__ src,
                                                      double *__restrict__ dst, long n, int to_lex) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    if (to_lex) dst[perm[i]] = src[i];
    else dst[i] = src[perm[i]];
  }
}

}  // namespace

extern "C" {

int mpdg_io_permute(const long *lex_index, const double *src, double *dst, long n, int to_lex, void *stream) {
  if (!lex_index || !src || !dst || n < 0 || src == dst) return MPDG_ERR_INVALID;
  if (n == 0) return MPDG_OK;
  const long nb = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(permute_kernel, dim3(nb), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), lex_index, src,
                     dst, n, to_lex ? 1 : 0);
  return hipGetLastError() == hipSuccess ? MPDG_OK : MPDG_ERR_DEVICE;
}

int mpdg_io_write_vtu(const char *filename, int degree, const int *n_cells, const double *extent, int n_fields,
                      const char *const *names, const double *const *fields) {
  if (!filename || !n_cells || !extent || degree < 1 || degree > 4 || n_fields < 0 || (n_fields && (!names || !fields)))
    return MPDG_ERR_INVALID;
  int nd[3];
  double h[3];
  for (int d = 0; d < 3; ++d) {
    if (n_cells[d] < 1 || !(extent[d] > 0.0)) return MPDG_ERR_INVALID;
    nd[d] = degree * n_cells[d] + 1;
    h[d] = extent[d] / (degree * n_cells[d]);  // sub-cell size
  }
  for (int i = 0; i < n_fields; ++i)
    if (!names[i] || !fields[i]) return MPDG_ERR_INVALID;
  const long np = (long)nd[0] * nd[1] * nd[2];
  const long nsub = (long)(nd[0] - 1) * (nd[1] - 1) * (nd[2] - 1);
  FILE *f = std::fopen(filename, "wb");
  if (!f) return MPDG_ERR_INVALID;
  std::fprintf(f,
               "<?xml version=\"1.0\"?>\n"
               "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
               "header_type=\"UInt64\">\n  <UnstructuredGrid>\n"
               "    <Piece NumberOfPoints=\"%ld\" NumberOfCells=\"%ld\">\n      <Points>\n",
               np, nsub);
  bool ok = true;
  {
    std::vector<double> X((size_t)np * 3);
    for (long i = 0; i < np; ++i) {
      const long ix = i % nd[0], iy = (i / nd[0]) % nd[1], iz = i / ((long)nd[0] * nd[1]);
      X[3 * i] = ix * h[0];
      X[3 * i + 1] = iy * h[1];
      X[3 * i + 2] = iz * h[2];
    }
    ok = ok && write_array(f, "Float64", "Points", 3, X);
  }
  std::fprintf(f, "      </Points>\n      <Cells>\n");
  {
    std::vector<int64_t> conn((size_t)nsub * 8), off(nsub);
    std::vector<uint8_t> type(nsub, 12);  // VTK_HEXAHEDRON
    const long sx = nd[0] - 1, sy = nd[1] - 1;
    for (long c = 0; c < nsub; ++c) {
      const long x = c % sx, y = (c / sx) % sy, z = c / (sx * sy);
      const long v0 = x + (long)nd[0] * (y + (long)nd[1] * z), dy = nd[0], dz = (long)nd[0] * nd[1];
      const long v[8] = {v0, v0 + 1, v0 + 1 + dy, v0 + dy, v0 + dz, v0 + 1 + dz, v0 + 1 + dy + dz, v0 + dy + dz};
      for (int j = 0; j < 8; ++j) conn[8 * c + j] = v[j];
      off[c] = 8 * (c + 1);
    }
    ok = ok && write_array(f, "Int64", "connectivity", 1, conn);
    ok = ok && write_array(f, "Int64", "offsets", 1, off);
    ok = ok && write_array(f, "UInt8", "types", 1, type);
  }
  std::fprintf(f, "      </Cells>\n      <PointData>\n");
  const FE1D &fe = mpdg::fe1d(degree);
  for (int i = 0; i < n_fields && ok; ++i) {
    std::vector<double> v;
    to_patch_vertices(fe, nd, fields[i], v);
    ok = write_array(f, "Float64", names[i], 1, v);
  }
  std::fprintf(f, "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
  ok = (std::fclose(f) == 0) && ok;
  return ok ? MPDG_OK : MPDG_ERR_INVALID;
}

int mpdg_io_write_pvd(const char *filename, int n, const double *times, const char *const *vtu_names) {
  if (!filename || n < 0 || (n && (!times || !vtu_names))) return MPDG_ERR_INVALID;
  FILE *f = std::fopen(filename, "w");
  if (!f) return MPDG_ERR_INVALID;
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" ByteOrder=\"LittleEndian\">\n"
                  "  <Collection>\n");
  for (int i = 0; i < n; ++i)
    std::fprintf(f, "    <DataSet timestep=\"%.17g\" group=\"\" part=\"0\" file=\"%s\"/>\n", times[i],
                 vtu_names[i] ? vtu_names[i] : "");
  std::fprintf(f, "  </Collection>\n</VTKFile>\n");
  return std::fclose(f) == 0 ? MPDG_OK : MPDG_ERR_INVALID;
}

int mpdg_io_dof_permutation(int degree, const int *n_cells, const double *extent, const double *points, long n,
                            long *lex_index) {
  if (!n_cells || !extent || !points || !lex_index || degree < 1 || degree > 4) return MPDG_ERR_INVALID;
  int nd[3];
  double h[3];
  for (int d = 0; d < 3; ++d) {
    if (n_cells[d] < 1 || !(extent[d] > 0.0)) return MPDG_ERR_INVALID;
    nd[d] = degree * n_cells[d] + 1;
    h[d] = extent[d] / n_cells[d];
  }
  if (n != (long)nd[0] * nd[1] * nd[2]) return MPDG_ERR_INVALID;
  const FE1D &fe = mpdg::fe1d(degree);
  std::vector<char> seen(n, 0);
  for (long i = 0; i < n; ++i) {
    long idx[3];
    for (int d = 0; d < 3; ++d) {
      const double x = points[3 * i + d] / h[d];  // in cell units
      long c = (long)std::floor(x);
      c = std::max<long>(0, std::min<long>(c, n_cells[d] - 1));
      const double xi = x - c;
      // nearest Gauss-Lobatto node of that cell; the point must lie on it
      int best = 0;
      for (int j = 1; j <= degree; ++j)
        if (std::fabs(xi - fe.nodes[j]) < std::fabs(xi - fe.nodes[best])) best = j;
      if (std::fabs(xi - fe.nodes[best]) > 1e-8) return MPDG_ERR_INVALID;
      idx[d] = degree * c + best;
    }
    const long g = idx[0] + (long)nd[0] * (idx[1] + (long)nd[1] * idx[2]);
    if (seen[g]) return MPDG_ERR_INVALID;  // two DoFs on one node
    seen[g] = 1;
    lex_index[i] = g;
  }
  return MPDG_OK;
}

}  // extern "C"
