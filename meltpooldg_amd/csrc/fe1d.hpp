// 1D finite element data for FE_Q(k) on Gauss-Lobatto support points with
// QGauss(k+1) quadrature on the unit interval (the space MeltPoolDG's heat
// operator uses: source/utilities/fe_util.cpp:18-58, finite_element_data.cpp:62-65).
#pragma once

#include <algorithm>
#include <cmath>

namespace mpdg {

constexpr int kMaxN1 = 8;

struct FE1D {
  int k = 1, n1 = 2;
  double nodes[kMaxN1] = {};  // Gauss-Lobatto support points on [0,1]
  double qp[kMaxN1] = {};     // Gauss points on [0,1]
  double qw[kMaxN1] = {};     // Gauss weights (sum 1)
  double S[kMaxN1][kMaxN1] = {};  // S[q][i] = phi_i(qp_q)
  double G[kMaxN1][kMaxN1] = {};  // G[q][i] = phi_i'(qp_q)
  double D[kMaxN1][kMaxN1] = {};  // D[q][p] = l_p'(qp_q), l = Lagrange basis on the Gauss points

  static void legendre(int n, double x, double &p, double &dp) {
    double a = 1.0, b = x;
    if (n == 0) { p = 1.0; dp = 0.0; return; }
    for (int m = 2; m <= n; ++m) {
      const double c = ((2.0 * m - 1.0) * x * b - (m - 1.0) * a) / m;
      a = b;
      b = c;
    }
    p = b;
    dp = n * (x * b - a) / (x * x - 1.0);
  }

  // value and derivative of the Lagrange basis on `pts` at x
  static void lagrange(int n, const double *pts, double x, double *val, double *der) {
    for (int i = 0; i < n; ++i) {
      double v = 1.0, d = 0.0;
      for (int j = 0; j < n; ++j)
        if (j != i) v *= (x - pts[j]) / (pts[i] - pts[j]);
      for (int m = 0; m < n; ++m) {
        if (m == i) continue;
        double t = 1.0 / (pts[i] - pts[m]);
        for (int j = 0; j < n; ++j)
          if (j != i && j != m) t *= (x - pts[j]) / (pts[i] - pts[j]);
        d += t;
      }
      if (val) val[i] = v;
      if (der) der[i] = d;
    }
  }

  explicit FE1D(int degree) : k(degree), n1(degree + 1) {
    const double pi = 3.14159265358979323846;
    const int n = n1;
    // Gauss points: roots of P_n by Newton from the Chebyshev-like guesses
    for (int j = 0; j < n; ++j) {
      double z = std::cos(pi * (j + 0.75) / (n + 0.5)), p, dp;
      for (int it = 0; it < 100; ++it) {
        legendre(n, z, p, dp);
        const double dz = p / dp;
        z -= dz;
        if (std::fabs(dz) < 1e-17) break;
      }
      legendre(n, z, p, dp);
      qp[n - 1 - j] = 0.5 * (1.0 + z);
      qw[n - 1 - j] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
    // Gauss-Lobatto points: end points and roots of P_{n-1}'
    nodes[0] = 0.0;
    nodes[n - 1] = 1.0;
    const int m = n - 1;
    for (int j = 1; j < n - 1; ++j) {
      double z = -std::cos(pi * j / m), p, dp;
      for (int it = 0; it < 100; ++it) {
        legendre(m, z, p, dp);
        const double d2p = (2.0 * z * dp - m * (m + 1.0) * p) / (1.0 - z * z);
        const double dz = dp / d2p;
        z -= dz;
        if (std::fabs(dz) < 1e-17) break;
      }
      nodes[j] = 0.5 * (1.0 + z);
    }
    for (int q = 0; q < n; ++q) {
      double v[kMaxN1], d[kMaxN1];
      lagrange(n, nodes, qp[q], v, d);
      for (int i = 0; i < n; ++i) { S[q][i] = v[i]; G[q][i] = d[i]; }
      lagrange(n, qp, qp[q], nullptr, d);
      for (int p = 0; p < n; ++p) D[q][p] = d[p];
    }
  }
};

// The tables of degrees 1 .. kMaxN1 - 1, built once per process (the Newton iterations above
// cost tens of microseconds, which a kernel launch should not pay)
inline const FE1D &fe1d(int degree) {
  struct Table {
    FE1D e[kMaxN1 - 1] = {FE1D(1), FE1D(2), FE1D(3), FE1D(4), FE1D(5), FE1D(6), FE1D(7)};
  };
  static const Table t;
  return t.e[std::min(std::max(degree, 1), kMaxN1 - 1) - 1];
}

}  // namespace mpdg
