// Gauss rules and natural shape-function derivatives of the GPU element families,
// tabulated once on the host and uploaded per context. The element kernels read
// dN[q][k][j] instead of re-evaluating shape functions per element, which the
// reference does per Gauss point per element (src/fem/element.cpp:623-624).
//
// Families and rules (reference src/fem/element.cpp:14-126, 160-253):
//   C3D8    trilinear hex, 2x2x2 Gauss, w = 1
//   C3D20   serendipity hex, 3x3x3 Gauss, w = products of 5/9, 8/9
//   C3D20R  serendipity hex, 2x2x2 Gauss (reduced)
//   C3D6    linear wedge, triangle centroid x 2 Gauss points in zeta, w = 1/2
// Tensor rules run xi fastest, then eta, then zeta (CalculiX gauss.f order), which
// fixes the Gauss-point numbering of the J2 history arrays.
#include <cmath>

#include "cxg_internal.hpp"

namespace cxg {
namespace {

struct Pt {
  double xi, et, ze, w;
};

std::vector<Pt> gauss(int type) {
  const double a2 = 1.0 / std::sqrt(3.0);
  const double a3 = std::sqrt(3.0 / 5.0);
  std::vector<Pt> r;
  if (type == CXG_C3D8 || type == CXG_C3D20R) {
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) r.push_back({i ? a2 : -a2, j ? a2 : -a2, k ? a2 : -a2, 1.0});
  } else if (type == CXG_C3D20) {
    const double x[3] = {-a3, 0.0, a3}, w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) r.push_back({x[i], x[j], x[k], w[i] * w[j] * w[k]});
  } else if (type == CXG_C3D6) {
    r.push_back({1.0 / 3.0, 1.0 / 3.0, -a2, 0.5});
    r.push_back({1.0 / 3.0, 1.0 / 3.0, a2, 0.5});
  }
  return r;
}

// Corner sign vectors of the hex in CalculiX node order 1..8.
const int kSign[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                         {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// dN/dxi of the 8-node hex: N_k = (1+s0 xi)(1+s1 et)(1+s2 ze)/8.
void d_hex8(const double p[3], double d[][3]) {
  for (int k = 0; k < 8; ++k) {
    double f[3];
    for (int a = 0; a < 3; ++a) f[a] = 1.0 + kSign[k][a] * p[a];
    d[k][0] = kSign[k][0] * f[1] * f[2] / 8.0;
    d[k][1] = kSign[k][1] * f[0] * f[2] / 8.0;
    d[k][2] = kSign[k][2] * f[0] * f[1] / 8.0;
  }
}

// dN/dxi of the 20-node serendipity hex. Corners: N = prod(1+s p)(s.p - 2)/8.
// Midside node on an edge along axis `ax`: N = (1 - p_ax^2) prod_{m != ax}(1+s_m p_m)/4
// with the in-plane signs of the edge's end corners. Midside order (nodes 9..20):
// edges 1-2, 2-3, 3-4, 4-1, 5-6, 6-7, 7-8, 8-5, 1-5, 2-6, 3-7, 4-8.
void d_hex20(const double p[3], double d[][3]) {
  for (int k = 0; k < 8; ++k) {
    double f[3], s[3];
    for (int a = 0; a < 3; ++a) {
      s[a] = kSign[k][a];
      f[a] = 1.0 + s[a] * p[a];
    }
    const double q = s[0] * p[0] + s[1] * p[1] + s[2] * p[2] - 2.0;
    const double fff = f[0] * f[1] * f[2];
    for (int a = 0; a < 3; ++a) {
      const double others = f[(a + 1) % 3] * f[(a + 2) % 3];
      d[k][a] = (s[a] * others * q + fff * s[a]) / 8.0;
    }
  }
  static const int from[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
  static const int axis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
  for (int e = 0; e < 12; ++e) {
    const int ax = axis[e], c = from[e];
    double f[3], s[3];
    for (int a = 0; a < 3; ++a) {
      s[a] = kSign[c][a];
      f[a] = 1.0 + s[a] * p[a];
    }
    const int m1 = (ax + 1) % 3, m2 = (ax + 2) % 3;
    const double bub = 1.0 - p[ax] * p[ax];
    d[8 + e][ax] = -2.0 * p[ax] * f[m1] * f[m2] / 4.0;
    d[8 + e][m1] = bub * s[m1] * f[m2] / 4.0;
    d[8 + e][m2] = bub * s[m2] * f[m1] / 4.0;
  }
}

// dN/dxi of the 6-node wedge: N = {a, xi, et} x {(1-ze)/2, (1+ze)/2}, a = 1-xi-et.
void d_wedge6(const double p[3], double d[][3]) {
  const double a = 1.0 - p[0] - p[1];
  const double lo = 0.5 * (1.0 - p[2]), hi = 0.5 * (1.0 + p[2]);
  const double t[6][3] = {{-lo, -lo, -0.5 * a}, {lo, 0.0, -0.5 * p[0]}, {0.0, lo, -0.5 * p[1]},
                          {-hi, -hi, 0.5 * a},  {hi, 0.0, 0.5 * p[0]},  {0.0, hi, 0.5 * p[1]}};
  for (int k = 0; k < 6; ++k)
    for (int j = 0; j < 3; ++j) d[k][j] = t[k][j];
}

// Shape-function values N_k at a point, same families and node orders as above.
void n_values(int type, const double p[3], double* N) {
  if (type == CXG_C3D8) {
    for (int k = 0; k < 8; ++k)
      N[k] = (1.0 + kSign[k][0] * p[0]) * (1.0 + kSign[k][1] * p[1]) *
             (1.0 + kSign[k][2] * p[2]) / 8.0;
  } else if (type == CXG_C3D6) {
    const double a = 1.0 - p[0] - p[1];
    const double lo = 0.5 * (1.0 - p[2]), hi = 0.5 * (1.0 + p[2]);
    const double v[6] = {a * lo, p[0] * lo, p[1] * lo, a * hi, p[0] * hi, p[1] * hi};
    for (int k = 0; k < 6; ++k) N[k] = v[k];
  } else {
    for (int k = 0; k < 8; ++k) {
      double f = 1.0, q = -2.0;
      for (int a = 0; a < 3; ++a) {
        f *= 1.0 + kSign[k][a] * p[a];
        q += kSign[k][a] * p[a];
      }
      N[k] = f * q / 8.0;
    }
    static const int from[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
    static const int axis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
    for (int e = 0; e < 12; ++e) {
      const int ax = axis[e], c = from[e];
      const int m1 = (ax + 1) % 3, m2 = (ax + 2) % 3;
      N[8 + e] = (1.0 - p[ax] * p[ax]) * (1.0 + kSign[c][m1] * p[m1]) *
                 (1.0 + kSign[c][m2] * p[m2]) / 4.0;
    }
  }
}

}  // namespace

int nodes_per_element(int type) {
  switch (type) {
    case CXG_C3D8: return 8;
    case CXG_C3D20:
    case CXG_C3D20R: return 20;
    case CXG_C3D6: return 6;
    default: return 0;
  }
}

int gauss_points(int type) {
  switch (type) {
    case CXG_C3D8:
    case CXG_C3D20R: return 8;
    case CXG_C3D20: return 27;
    case CXG_C3D6: return 2;
    default: return 0;
  }
}

void host_tables(int type, std::vector<double>& dN, std::vector<double>& w,
                 std::vector<double>& N) {
  const int nn = nodes_per_element(type);
  const std::vector<Pt> g = gauss(type);
  dN.assign(g.size() * size_t(nn) * 3, 0.0);
  w.assign(g.size(), 0.0);
  N.assign(g.size() * size_t(nn), 0.0);
  double d[20][3];
  for (size_t q = 0; q < g.size(); ++q) {
    const double p[3] = {g[q].xi, g[q].et, g[q].ze};
    if (type == CXG_C3D8) d_hex8(p, d);
    else if (type == CXG_C3D6) d_wedge6(p, d);
    else d_hex20(p, d);
    for (int k = 0; k < nn; ++k)
      for (int j = 0; j < 3; ++j) dN[(q * nn + k) * 3 + j] = d[k][j];
    w[q] = g[q].w;
    n_values(type, p, N.data() + q * size_t(nn));
  }
  if (type == CXG_C3D8)  // the kernels' closed form must give the table's bits
    for (size_t q = 0; q < g.size(); ++q)
      for (int k = 0; k < nn; ++k)
        for (int c = 0; c < 3; ++c)
          if (dN[(q * nn + k) * 3 + c] != hex8_dn(int(q), k, c))
            throw Error(CXG_ERR_STATE, "C3D8 natural derivatives differ from their closed form");
}

}  // namespace cxg
