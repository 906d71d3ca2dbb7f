// TEST INFRASTRUCTURE ONLY — CPU restatement (the oracle) of CalculiX++'s
// element-integration and global-assembly path.
//
// Who may use this: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg, as the checker and the timed CPU baseline. The product (calculixpp_amd and
// libcxg.so) never links, loads or calls it.
//
// Every routine restates the reference algorithm in this file's own code and keeps
// the reference's floating-point operation ORDER, so that compiled with the same
// flags as the reference core (g++ -O2, no FMA contraction on x86-64) it agrees
// with the reference bit for bit; the tests pin that against oracle/_ref (the
// reference core built from its sources) and against the golden fixtures.
//
// Pins per function (see tests/test_oracle.py):
//   gauss/shape        partition of unity, closed-form Gauss tables, _ref bitwise
//   gradients/Jacobian sum g = 0, sum g (x) x = I, volume, _ref bitwise
//   element_stiffness  symmetry, 6 rigid-body modes, patch test, _ref bitwise
//   j2_evaluate        elastic limit, yield-surface return, closed-form dgamma for
//                      linear hardening, FD tangent, _ref bitwise
//   assembly           golden fixtures from _ref (COO key set and values), dense
//                      brute force on tiny meshes
//   internal force     f = K u identity, _ref
//   vjp (adjoint)      central finite differences of w^T K(x,E) u; closed forms
//                      dphi/dE_e = phi_e / E_e and du = K w.
//   mass, geometric    _ref bitwise; total mass rho V per direction, lumped = row sums,
//                      geometric stiffness annihilates rigid translations and is linear
//                      in the stress; hydrostatic stress gives p times the Laplacian
//   recover_fields     _ref bitwise (elastic, material-point and Gauss-point forms);
//                      linear displacement fields give the exact uniform strain and
//                      D strain at every node; reactions balance the loads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "orc_model.h"

namespace {

using V3 = std::array<double, 3>;
constexpr int kMaxN = 20;

struct GP {
  double xi, et, ze, w;
};

// ---- Gauss rules (reference src/fem/element.cpp:14-126) -----------------------
// Hex tensor rules run xi fastest, then eta, then zeta (CalculiX gauss.f order).
const double kA2 = 0.577350269189625764;  // 1/sqrt(3)
const double kA3 = 0.774596669241483377;  // sqrt(3/5)

std::vector<GP> build_rule(int type) {
  std::vector<GP> r;
  switch (type) {
    case ORC_C3D8:
    case ORC_C3D20R:
      for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
          for (int i = 0; i < 2; ++i)
            r.push_back({i ? kA2 : -kA2, j ? kA2 : -kA2, k ? kA2 : -kA2, 1.0});
      break;
    case ORC_C3D20: {
      const double x[3] = {-kA3, 0.0, kA3};
      const double w[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
          for (int i = 0; i < 3; ++i) r.push_back({x[i], x[j], x[k], w[i] * w[j] * w[k]});
      break;
    }
    case ORC_C3D6:
      r.push_back({1.0 / 3.0, 1.0 / 3.0, -kA2, 0.5});
      r.push_back({1.0 / 3.0, 1.0 / 3.0, kA2, 0.5});
      break;
    default:
      throw std::runtime_error("oracle: element type outside the hot path");
  }
  return r;
}

const std::vector<GP>& rule(int type) {
  static const std::vector<GP> r8 = build_rule(ORC_C3D8);
  static const std::vector<GP> r20 = build_rule(ORC_C3D20);
  static const std::vector<GP> r6 = build_rule(ORC_C3D6);
  switch (type) {
    case ORC_C3D8:
    case ORC_C3D20R:
      return r8;
    case ORC_C3D20:
      return r20;
    case ORC_C3D6:
      return r6;
  }
  throw std::runtime_error("oracle: element type outside the hot path");
}

int nodes_of(int type) {
  switch (type) {
    case ORC_C3D8:
      return 8;
    case ORC_C3D20:
    case ORC_C3D20R:
      return 20;
    case ORC_C3D6:
      return 6;
  }
  throw std::runtime_error("oracle: element type outside the hot path");
}

// ---- Shape functions (reference src/fem/element.cpp:160-253) -----------------
// Hex corner sign table in CalculiX node order.
const int kCorner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
// C3D20 midside nodes 9..20: first end corner and the natural axis the edge runs along.
const int kEdgeFrom[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
const int kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};

struct ShapeVals {
  int n = 0;
  double N[kMaxN];
  double d[kMaxN][3];
};

void shape_hex8(double xi, double et, double ze, ShapeVals& s) {
  s.n = 8;
  for (int k = 0; k < 8; ++k) {
    const double sx = kCorner[k][0], sy = kCorner[k][1], sz = kCorner[k][2];
    const double fx = 1.0 + sx * xi, fy = 1.0 + sy * et, fz = 1.0 + sz * ze;
    s.N[k] = fx * fy * fz / 8.0;
    s.d[k][0] = sx * fy * fz / 8.0;
    s.d[k][1] = sy * fx * fz / 8.0;
    s.d[k][2] = sz * fx * fy / 8.0;
  }
}

void shape_hex20(double xi, double et, double ze, ShapeVals& s) {
  s.n = 20;
  const double p[3] = {xi, et, ze};
  for (int k = 0; k < 8; ++k) {
    const double sg[3] = {double(kCorner[k][0]), double(kCorner[k][1]), double(kCorner[k][2])};
    const double f[3] = {1.0 + sg[0] * xi, 1.0 + sg[1] * et, 1.0 + sg[2] * ze};
    const double q = sg[0] * xi + sg[1] * et + sg[2] * ze - 2.0;
    const double fff = f[0] * f[1] * f[2];
    s.N[k] = fff * q / 8.0;
    for (int dd = 0; dd < 3; ++dd) {
      double rest = 1.0;  // product of the two factors other than f[dd], ascending
      for (int m = 0; m < 3; ++m)
        if (m != dd) rest *= f[m];
      s.d[k][dd] = (sg[dd] * rest * q + fff * sg[dd]) / 8.0;
    }
  }
  for (int e = 0; e < 12; ++e) {
    const int node = 8 + e, ax = kEdgeAxis[e], c = kEdgeFrom[e];
    const double sg[3] = {double(kCorner[c][0]), double(kCorner[c][1]), double(kCorner[c][2])};
    const double pa = p[ax];
    double face = 1.0;  // product over the two in-plane axes, ascending
    for (int m = 0; m < 3; ++m)
      if (m != ax) face *= 1.0 + sg[m] * p[m];
    const double bub = 1.0 - pa * pa;
    s.N[node] = bub * face / 4.0;
    for (int dd = 0; dd < 3; ++dd) {
      if (dd == ax) {
        s.d[node][dd] = (-2.0 * pa) * face / 4.0;
      } else {
        double other = 1.0;
        for (int m = 0; m < 3; ++m)
          if (m != ax && m != dd) other *= 1.0 + sg[m] * p[m];
        s.d[node][dd] = bub * sg[dd] * other / 4.0;
      }
    }
  }
}

void shape_wedge6(double xi, double et, double ze, ShapeVals& s) {
  s.n = 6;
  const double a = 1.0 - xi - et;
  const double lo = 0.5 * (1.0 - ze), hi = 0.5 * (1.0 + ze);
  const double N[6] = {a * lo, xi * lo, et * lo, a * hi, xi * hi, et * hi};
  const double d[6][3] = {{-lo, -lo, -0.5 * a}, {lo, 0.0, -0.5 * xi}, {0.0, lo, -0.5 * et},
                          {-hi, -hi, 0.5 * a},  {hi, 0.0, 0.5 * xi},  {0.0, hi, 0.5 * et}};
  for (int k = 0; k < 6; ++k) {
    s.N[k] = N[k];
    for (int j = 0; j < 3; ++j) s.d[k][j] = d[k][j];
  }
}

ShapeVals shape_at(int type, double xi, double et, double ze) {
  ShapeVals s;
  switch (type) {
    case ORC_C3D8:
      shape_hex8(xi, et, ze, s);
      break;
    case ORC_C3D20:
    case ORC_C3D20R:
      shape_hex20(xi, et, ze, s);
      break;
    case ORC_C3D6:
      shape_wedge6(xi, et, ze, s);
      break;
    default:
      throw std::runtime_error("oracle: element type outside the hot path");
  }
  return s;
}

// ---- Jacobian and physical gradients (reference src/fem/element.cpp:80-102,361-380)
const char* kBadJacobian = "non-positive element Jacobian (bad connectivity?)";

double gradients(const ShapeVals& s, const double* x /* n*3 */, double g[kMaxN][3]) {
  double J[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < s.n; ++k) acc += x[3 * k + i] * s.d[k][j];
      J[i][j] = acc;
    }
  // Cofactor inverse; rows of J^-1 are built from cofactors of J.
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  if (det <= 0.0) throw std::runtime_error(kBadJacobian);
  const double r = 1.0 / det;
  double Ai[3][3];  // J^-1
  Ai[0][0] = c00 * r;
  Ai[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
  Ai[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
  Ai[1][0] = c01 * r;
  Ai[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
  Ai[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
  Ai[2][0] = c02 * r;
  Ai[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
  Ai[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
  // g_k[i] = sum_j (J^-T)[i][j] dN_k/dxi_j = sum_j Ai[j][i] dN_k[j]
  for (int k = 0; k < s.n; ++k)
    for (int i = 0; i < 3; ++i) {
      double acc = 0.0;
      for (int j = 0; j < 3; ++j) acc += Ai[j][i] * s.d[k][j];
      g[k][i] = acc;
    }
  return det;
}

// ---- Strain-displacement B columns and elasticity (element.cpp:382-422) -------
// Column (node k, direction dd) of B as a dense Voigt 6-vector.
void bcol(const double gk[3], int dd, double c[6]) {
  const double gx = gk[0], gy = gk[1], gz = gk[2];
  c[0] = dd == 0 ? gx : 0.0;
  c[1] = dd == 1 ? gy : 0.0;
  c[2] = dd == 2 ? gz : 0.0;
  c[3] = dd == 0 ? gy : (dd == 1 ? gx : 0.0);
  c[4] = dd == 0 ? gz : (dd == 2 ? gx : 0.0);
  c[5] = dd == 1 ? gz : (dd == 2 ? gy : 0.0);
}

struct Mat6 {
  double m[6][6];
};

Mat6 iso_D(double E, double nu) {
  const double two_mu = E / (1.0 + nu);
  const double lam = nu * two_mu / (1.0 - 2.0 * nu);
  const double mu = two_mu / 2.0;
  Mat6 D{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D.m[i][j] = (i == j) ? lam + two_mu : lam;
  for (int i = 3; i < 6; ++i) D.m[i][i] = mu;
  return D;
}

// Upper-triangle B^T D B accumulation of one Gauss point (element.cpp:440-462).
void add_btdb(const double g[kMaxN][3], int ndof, const Mat6& D, double scale, double* Ke) {
  for (int a = 0; a < ndof; ++a) {
    double ca[6], Dca[6];
    bcol(g[a / 3], a % 3, ca);
    for (int row = 0; row < 6; ++row) {
      double acc = 0.0;
      for (int r = 0; r < 6; ++r) acc += D.m[row][r] * ca[r];
      Dca[row] = acc;
    }
    for (int b = a; b < ndof; ++b) {
      double cb[6];
      bcol(g[b / 3], b % 3, cb);
      double acc = 0.0;
      for (int row = 0; row < 6; ++row) acc += cb[row] * Dca[row];
      Ke[size_t(a) * ndof + b] += scale * acc;
    }
  }
}

void mirror(int ndof, double* Ke) {
  for (int a = 0; a < ndof; ++a)
    for (int b = a + 1; b < ndof; ++b) Ke[size_t(b) * ndof + a] = Ke[size_t(a) * ndof + b];
}

// B^T sigma accumulation (element.cpp:465-475).
void add_btsig(const double g[kMaxN][3], int ndof, const double sig[6], double scale,
               double* fe) {
  for (int a = 0; a < ndof; ++a) {
    double ca[6];
    bcol(g[a / 3], a % 3, ca);
    double acc = 0.0;
    for (int row = 0; row < 6; ++row) acc += ca[row] * sig[row];
    fe[a] += scale * acc;
  }
}

// strain = B ue with engineering shear (element.cpp:394-408).
void strain_of(const double g[kMaxN][3], int n, const double* ue, double eps[6]) {
  double H[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc += g[k][j] * ue[3 * k + i];
      H[i][j] = acc;
    }
  eps[0] = H[0][0];
  eps[1] = H[1][1];
  eps[2] = H[2][2];
  eps[3] = H[0][1] + H[1][0];
  eps[4] = H[0][2] + H[2][0];
  eps[5] = H[1][2] + H[2][1];
}

void stiffness(int type, const double* x, double E, double nu, double* Ke) {
  const int n = nodes_of(type), ndof = 3 * n;
  std::fill(Ke, Ke + size_t(ndof) * ndof, 0.0);
  const Mat6 D = iso_D(E, nu);
  double g[kMaxN][3];
  for (const GP& p : rule(type)) {
    const ShapeVals s = shape_at(type, p.xi, p.et, p.ze);
    const double det = gradients(s, x, g);
    add_btdb(g, ndof, D, det * p.w, Ke);
  }
  mirror(ndof, Ke);
}

// ---- J2 radial return + consistent tangent (reference src/fem/plasticity.cpp) ---
const double kS23 = 0.816496580927726;  // sqrt(2/3)
const double kS32 = 1.224744871391589;  // sqrt(3/2)

struct J2Mat {
  double E, nu, mu, hkin;
  std::vector<double> ep, sy;
  Mat6 De;
};

J2Mat make_j2(const orc_j2& p) {
  if (p.n_table < 1) throw std::runtime_error("J2PlasticMaterial needs a non-empty *PLASTIC table");
  J2Mat m;
  m.E = p.E;
  m.nu = p.nu;
  m.mu = p.E / (2.0 * (1.0 + p.nu));
  m.hkin = p.kinematic_modulus;
  m.ep.assign(p.eqplastic, p.eqplastic + p.n_table);
  m.sy.assign(p.yield, p.yield + p.n_table);
  if (p.hardening == ORC_HARD_KIN || p.hardening == ORC_HARD_COMBINED) {
    if (m.hkin == 0.0 && m.ep.size() >= 2 && m.ep[1] > m.ep[0])
      m.hkin = (m.sy[1] - m.sy[0]) / (m.ep[1] - m.ep[0]);
    if (p.hardening == ORC_HARD_KIN)
      for (double& y : m.sy) y = m.sy.front();
  }
  m.De = iso_D(p.E, p.nu);
  return m;
}

// Piecewise-linear yield curve with flat ends (plasticity.cpp:67-88).
void yield_curve(const J2Mat& m, double ep, double& sy, double& h) {
  const size_t n = m.ep.size();
  if (n == 1 || ep <= m.ep.front()) {
    sy = m.sy.front();
    h = 0.0;
    return;
  }
  if (ep >= m.ep.back()) {
    sy = m.sy.back();
    h = 0.0;
    return;
  }
  for (size_t i = 0; i + 1 < n; ++i)
    if (ep >= m.ep[i] && ep <= m.ep[i + 1]) {
      h = (m.sy[i + 1] - m.sy[i]) / (m.ep[i + 1] - m.ep[i]);
      sy = m.sy[i] + h * (ep - m.ep[i]);
      return;
    }
  sy = m.sy.back();
  h = 0.0;
}

// st = 26-double history (orc_model.h layout). Writes stress + tangent, trial history.
void j2_eval(const J2Mat& m, const double eps[6], double* st, double sig[6], Mat6& Dt) {
  const double* ep_c = st + 0;
  const double epbar_c = st[6];
  const double* al_c = st + 7;
  double* ep_t = st + 13;
  double* al_t = st + 20;
  double ee[6], s_tr[6];
  for (int i = 0; i < 6; ++i) ee[i] = eps[i] - ep_c[i];
  for (int i = 0; i < 6; ++i) {
    double acc = 0.0;
    for (int j = 0; j < 6; ++j) acc += m.De.m[i][j] * ee[j];
    s_tr[i] = acc;
  }
  const double pm = (s_tr[0] + s_tr[1] + s_tr[2]) / 3.0;
  double dev[6] = {s_tr[0] - pm, s_tr[1] - pm, s_tr[2] - pm, s_tr[3], s_tr[4], s_tr[5]};
  double xi[6];
  for (int i = 0; i < 6; ++i) xi[i] = dev[i] - al_c[i];
  const double nrm = std::sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] +
                               2.0 * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]));
  const double q_tr = kS32 * nrm;
  double sy0 = 0.0, h0 = 0.0;
  yield_curve(m, epbar_c, sy0, h0);
  const double f_tr = q_tr - sy0;
  if (f_tr <= 0.0 || nrm == 0.0) {
    for (int i = 0; i < 6; ++i) sig[i] = s_tr[i];
    Dt = m.De;
    for (int i = 0; i < 6; ++i) {
      ep_t[i] = ep_c[i];
      al_t[i] = al_c[i];
    }
    st[19] = epbar_c;
    return;
  }
  const double epb0 = epbar_c;
  double dg = 0.0;
  for (int it = 0; it < 50; ++it) {
    double sy = 0.0, h = 0.0;
    yield_curve(m, epb0 + dg, sy, h);
    const double res = q_tr - 3.0 * m.mu * dg - m.hkin * dg - sy;
    const double slope = -(3.0 * m.mu + m.hkin + h);
    const double step = -res / slope;
    dg += step;
    if (dg < 0.0) dg = 0.0;
    if (std::abs(step) <= 1e-14 * (1.0 + std::abs(dg))) break;
  }
  double nv[6];
  for (int i = 0; i < 6; ++i) nv[i] = xi[i] / nrm;
  const double cdev = 2.0 * m.mu * dg * kS32;
  const double calpha = (2.0 / 3.0) * m.hkin * dg * kS32;
  const double mean = (s_tr[0] + s_tr[1] + s_tr[2]) / 3.0;
  for (int i = 0; i < 6; ++i) sig[i] = (dev[i] - cdev * nv[i]) + (i < 3 ? mean : 0.0);
  const double cep = dg * kS32;
  for (int i = 0; i < 6; ++i) {
    ep_t[i] = ep_c[i] + cep * nv[i] * (i < 3 ? 1.0 : 2.0);
    al_t[i] = al_c[i] + calpha * nv[i];
  }
  st[19] = epb0 + dg;
  double sy = 0.0, h = 0.0;
  yield_curve(m, epb0 + dg, sy, h);
  const double theta = 1.0 - (2.0 * m.mu * dg * kS32) / nrm;
  const double den = 1.0 + (h + m.hkin) / (3.0 * m.mu);
  const double theta_bar = 1.0 / den - (1.0 - theta);
  Mat6 P{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P.m[i][j] = 2.0 * m.mu * ((i == j ? 1.0 : 0.0) - 1.0 / 3.0);
  for (int i = 3; i < 6; ++i) P.m[i][i] = m.mu;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j)
      Dt.m[i][j] = m.De.m[i][j] - (1.0 - theta) * P.m[i][j] -
                   2.0 * m.mu * theta_bar * nv[i] * nv[j];
}

// Material-point element kernel (element.cpp:708-743), elastic or J2.
void tangent_force(int type, const double* x, const double* ue, double E, double nu,
                   const J2Mat* j2, double* state, double* Ke, double* fe) {
  const int n = nodes_of(type), ndof = 3 * n;
  std::fill(Ke, Ke + size_t(ndof) * ndof, 0.0);
  std::fill(fe, fe + ndof, 0.0);
  const Mat6 De = iso_D(E, nu);
  double g[kMaxN][3];
  const auto& R = rule(type);
  for (size_t q = 0; q < R.size(); ++q) {
    const GP& p = R[q];
    const ShapeVals s = shape_at(type, p.xi, p.et, p.ze);
    const double det = gradients(s, x, g);
    const double scale = det * p.w;
    double eps[6], sig[6];
    strain_of(g, n, ue, eps);
    Mat6 Dt;
    if (j2) {
      j2_eval(*j2, eps, state + q * ORC_STATE, sig, Dt);
    } else {
      Dt = De;
      for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 6; ++j) acc += De.m[i][j] * eps[j];
        sig[i] = acc;
      }
    }
    add_btdb(g, ndof, Dt, scale, Ke);
    add_btsig(g, ndof, sig, scale, fe);
  }
  mirror(ndof, Ke);
}

// Consistent mass (element.cpp:786-812): m_ab = rho detJ w N_a N_b on each direction.
void mass(int type, const double* x, double rho, double* Me) {
  const int n = nodes_of(type), ndof = 3 * n;
  std::fill(Me, Me + size_t(ndof) * ndof, 0.0);
  if (rho == 0.0) return;
  double g[kMaxN][3];
  for (const GP& p : rule(type)) {
    const ShapeVals s = shape_at(type, p.xi, p.et, p.ze);
    const double det = gradients(s, x, g);
    const double scale = rho * det * p.w;
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        const double m = scale * s.N[a] * s.N[b];
        for (int d = 0; d < 3; ++d) Me[size_t(3 * a + d) * ndof + size_t(3 * b + d)] += m;
      }
  }
}

// Row-sum lumped mass (element.cpp:814-831).
void mass_lumped(int type, const double* x, double rho, double* Ml) {
  const int ndof = 3 * nodes_of(type);
  std::vector<double> Mc(size_t(ndof) * ndof);
  mass(type, x, rho, Mc.data());
  std::fill(Ml, Ml + size_t(ndof) * ndof, 0.0);
  for (int i = 0; i < ndof; ++i) {
    double row = 0.0;
    for (int j = 0; j < ndof; ++j) row += Mc[size_t(i) * ndof + j];
    Ml[size_t(i) * ndof + i] = row;
  }
}

// Geometric stiffness (element.cpp:833-854 with accumulate_kgeo, :482-508): upper
// node pairs, sigma-energy of the two gradients on each direction, then mirrored.
// sig: ngp Voigt stresses, or NULL (zero field).
void geometric(int type, const double* x, const double* sig, double* Ke) {
  const int n = nodes_of(type), ndof = 3 * n;
  std::fill(Ke, Ke + size_t(ndof) * ndof, 0.0);
  double g[kMaxN][3];
  const auto& R = rule(type);
  for (size_t q = 0; q < R.size(); ++q) {
    const GP& p = R[q];
    const ShapeVals sv = shape_at(type, p.xi, p.et, p.ze);
    const double det = gradients(sv, x, g);
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (sig)
      for (int c = 0; c < 6; ++c) s[c] = sig[6 * q + c];
    const double scale = det * p.w;
    for (int a = 0; a < n; ++a) {
      const double* ga = g[a];
      for (int b = a; b < n; ++b) {
        const double* gb = g[b];
        const double se = s[0] * ga[0] * gb[0] + s[1] * ga[1] * gb[1] + s[2] * ga[2] * gb[2] +
                          s[3] * (ga[0] * gb[1] + ga[1] * gb[0]) +
                          s[4] * (ga[0] * gb[2] + ga[2] * gb[0]) +
                          s[5] * (ga[1] * gb[2] + ga[2] * gb[1]);
        const double v = scale * se;
        for (int d = 0; d < 3; ++d) {
          const int ia = 3 * a + d, ib = 3 * b + d;
          if (ib >= ia) Ke[size_t(ia) * ndof + ib] += v;
        }
      }
    }
  }
  mirror(ndof, Ke);
}

// ---- Model-level assembly (reference src/fem/assembly.cpp) -------------------
struct DofMap {
  int32_t n_free = 0;
  std::vector<int32_t> eq;          // per global DOF, -1 if constrained
  std::vector<double> prescribed;   // per global DOF
};

// SPC-only DOF numbering (assembly.cpp:19-42, constraint_transform.cpp:78-82).
DofMap dof_map(const orc_model& m) {
  const size_t ndof = size_t(m.n_nodes) * 3;
  DofMap d;
  d.eq.assign(ndof, -1);
  d.prescribed.assign(ndof, 0.0);
  std::vector<char> fixed(ndof, 0);
  for (int i = 0; i < m.n_spc; ++i) {
    const int32_t nd = m.spc_node[i];
    if (nd < 0 || nd >= m.n_nodes) throw std::runtime_error("*BOUNDARY references unknown node");
    if (m.spc_comp[i] < 1 || m.spc_comp[i] > 3) throw std::runtime_error("*BOUNDARY dof out of range");
    const size_t gdof = size_t(nd) * 3 + size_t(m.spc_comp[i] - 1);
    fixed[gdof] = 1;
    d.prescribed[gdof] = m.spc_value[i];
  }
  for (size_t gdof = 0; gdof < ndof; ++gdof)
    if (!fixed[gdof]) d.eq[gdof] = d.n_free++;
  return d;
}

struct Gathered {
  int n;
  std::vector<int32_t> idx;
  std::vector<double> x, u;
};

void gather(const orc_model& m, int64_t e, const double* u, Gathered& G) {
  const int64_t b = m.eptr[e], n = m.eptr[e + 1] - b;
  G.n = int(n);
  G.idx.resize(n);
  G.x.resize(3 * n);
  if (u) G.u.resize(3 * n);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t nd = m.conn[b + i];
    if (nd < 0 || nd >= m.n_nodes) throw std::runtime_error("element references unknown node");
    G.idx[i] = nd;
    for (int c = 0; c < 3; ++c) {
      G.x[3 * i + c] = m.coords[3 * size_t(nd) + c];
      if (u) G.u[3 * i + c] = u[3 * size_t(nd) + c];
    }
  }
}

bool is_active(const orc_model& m, int64_t e) { return m.active == nullptr || m.active[e] != 0; }

using KMap = std::unordered_map<int64_t, double>;

// Congruence scatter for the SPC-only transform (assembly.cpp:76-110): free rows
// only; a prescribed column moves c*k onto the rhs; exact-zero entries skipped.
void scatter(const double* Ke, int ndof, const int32_t* nidx, const DofMap& d, KMap& km,
             std::vector<double>& rhs) {
  const int64_t nf = d.n_free;
  for (int a = 0; a < ndof; ++a) {
    const size_t ga = size_t(nidx[a / 3]) * 3 + a % 3;
    const int32_t ra = d.eq[ga];
    if (ra < 0) continue;  // constrained row: no free terms (constant only)
    for (int b = 0; b < ndof; ++b) {
      const double kab = Ke[size_t(a) * ndof + b];
      if (kab == 0.0) continue;
      const size_t gb = size_t(nidx[b / 3]) * 3 + b % 3;
      const int32_t cb = d.eq[gb];
      if (cb >= 0)
        km[int64_t(ra) * nf + cb] += 1.0 * kab * 1.0;
      else if (d.prescribed[gb] != 0.0)
        rhs[size_t(ra)] -= 1.0 * kab * d.prescribed[gb];
    }
  }
}

void add_cloads(const orc_model& m, const DofMap& d, std::vector<double>& rhs) {
  std::vector<double> f(size_t(m.n_nodes) * 3, 0.0);
  for (int i = 0; i < m.n_cload; ++i) {
    const int32_t nd = m.cload_node[i];
    if (nd >= 0 && nd < m.n_nodes && m.cload_comp[i] >= 1 && m.cload_comp[i] <= 3)
      f[size_t(nd) * 3 + size_t(m.cload_comp[i] - 1)] += 1.0 * m.cload_value[i];
  }
  for (size_t gdof = 0; gdof < f.size(); ++gdof) {
    if (f[gdof] == 0.0) continue;
    if (d.eq[gdof] >= 0) rhs[size_t(d.eq[gdof])] += 1.0 * f[gdof];
  }
}

}  // namespace

struct orc_system {
  int32_t n_free = 0;
  std::vector<int32_t> rows, cols;  // hash order, as flushed (assembly.cpp:211-221)
  std::vector<double> vals;
  std::vector<double> rhs;
  std::vector<int32_t> dof_eq;
  std::vector<double> prescribed;
};

namespace {

void flush(const KMap& km, orc_system& s) {
  s.rows.reserve(km.size());
  s.cols.reserve(km.size());
  s.vals.reserve(km.size());
  for (const auto& kv : km) {
    s.rows.push_back(int32_t(kv.first / s.n_free));
    s.cols.push_back(int32_t(kv.first % s.n_free));
    s.vals.push_back(kv.second);
  }
}

void set_err(char* err, int len, const char* msg) {
  if (err && len > 0) {
    std::strncpy(err, msg, size_t(len) - 1);
    err[len - 1] = 0;
  }
}

int ngp_of(int type) { return int(rule(type).size()); }

// phi = w^T K(x,E) u restricted to one element, with its gradients (the oracle of
// the adjoint kernel). Derivation, per Gauss point with s = det*w:
//   phi_q = s * sigma_u : H_w,  H_w = sum_k w_k (x) g_k, sigma_u = D eps(u)
//   d g_k / d x_m = -g_m (x) g_k,  d s / d x_m = s g_m
//   dphi/dx_m = sum_q [ phi_q I - s (sigma_u H_w + sigma_w H_u)^T ] g_m
//   dphi/dE   = phi / E   (D is linear in E)
//   dphi/du_k = s B_k^T sigma_w  (= K w, K symmetric)
void elem_vjp(int type, const double* x, const double* ue, const double* we, double E,
              double nu, double* dx, double* du, double* dE, double* phi_out) {
  const int n = nodes_of(type);
  const Mat6 D = iso_D(E, nu);
  double g[kMaxN][3];
  double phi = 0.0;
  for (int k = 0; k < 3 * n; ++k) {
    dx[k] = 0.0;
    du[k] = 0.0;
  }
  for (const GP& p : rule(type)) {
    const ShapeVals s = shape_at(type, p.xi, p.et, p.ze);
    const double det = gradients(s, x, g);
    const double sc = det * p.w;
    double Hu[3][3], Hw[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < n; ++k) {
          a += ue[3 * k + i] * g[k][j];
          b += we[3 * k + i] * g[k][j];
        }
        Hu[i][j] = a;
        Hw[i][j] = b;
      }
    const double eu[6] = {Hu[0][0], Hu[1][1], Hu[2][2], Hu[0][1] + Hu[1][0],
                          Hu[0][2] + Hu[2][0], Hu[1][2] + Hu[2][1]};
    const double ew[6] = {Hw[0][0], Hw[1][1], Hw[2][2], Hw[0][1] + Hw[1][0],
                          Hw[0][2] + Hw[2][0], Hw[1][2] + Hw[2][1]};
    double su[6], sw[6];
    for (int i = 0; i < 6; ++i) {
      double a = 0.0, b = 0.0;
      for (int j = 0; j < 6; ++j) {
        a += D.m[i][j] * eu[j];
        b += D.m[i][j] * ew[j];
      }
      su[i] = a;
      sw[i] = b;
    }
    // symmetric tensors from Voigt stress
    const double Su[3][3] = {{su[0], su[3], su[4]}, {su[3], su[1], su[5]}, {su[4], su[5], su[2]}};
    const double Sw[3][3] = {{sw[0], sw[3], sw[4]}, {sw[3], sw[1], sw[5]}, {sw[4], sw[5], sw[2]}};
    double pq = 0.0;
    for (int i = 0; i < 6; ++i) pq += ew[i] * su[i];
    pq *= sc;
    phi += pq;
    double M[3][3];  // s*(Su Hw + Sw Hu)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double a = 0.0;
        for (int l = 0; l < 3; ++l) a += Su[i][l] * Hw[l][j] + Sw[i][l] * Hu[l][j];
        M[i][j] = sc * a;
      }
    for (int m = 0; m < n; ++m)
      for (int pcomp = 0; pcomp < 3; ++pcomp) {
        double a = pq * g[m][pcomp];
        for (int i = 0; i < 3; ++i) a -= M[i][pcomp] * g[m][i];
        dx[3 * m + pcomp] += a;
      }
    add_btsig(g, 3 * n, sw, sc, du);
  }
  *dE = phi / E;
  if (phi_out) *phi_out = phi;
}

}  // namespace

extern "C" {

int orc_nodes_per_element(int type) {
  try {
    return nodes_of(type);
  } catch (...) {
    return -1;
  }
}

int orc_num_gauss(int type) {
  try {
    return ngp_of(type);
  } catch (...) {
    return -1;
  }
}

int orc_gauss_rule(int type, double* out) {
  try {
    const auto& R = rule(type);
    for (size_t q = 0; q < R.size(); ++q) {
      out[4 * q + 0] = R[q].xi;
      out[4 * q + 1] = R[q].et;
      out[4 * q + 2] = R[q].ze;
      out[4 * q + 3] = R[q].w;
    }
    return int(R.size());
  } catch (...) {
    return -1;
  }
}

int orc_shape(int type, double xi, double et, double ze, double* N, double* dN) {
  try {
    const ShapeVals s = shape_at(type, xi, et, ze);
    for (int k = 0; k < s.n; ++k) {
      N[k] = s.N[k];
      for (int j = 0; j < 3; ++j) dN[3 * k + j] = s.d[k][j];
    }
    return s.n;
  } catch (...) {
    return -1;
  }
}

// Physical gradients at Gauss point q; returns 0 and writes grad (n*3) and det.
int orc_gradients(int type, int q, const double* coords, double* grad, double* det,
                  char* err, int errlen) {
  try {
    const GP& p = rule(type).at(size_t(q));
    const ShapeVals s = shape_at(type, p.xi, p.et, p.ze);
    double g[kMaxN][3];
    *det = gradients(s, coords, g);
    for (int k = 0; k < s.n; ++k)
      for (int j = 0; j < 3; ++j) grad[3 * k + j] = g[k][j];
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int orc_element_stiffness(int type, const double* coords, double E, double nu, double* Ke,
                          char* err, int errlen) {
  try {
    stiffness(type, coords, E, nu, Ke);
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int orc_j2_evaluate(const orc_j2* mat, const double* strain, double* state, double* stress,
                    double* D, char* err, int errlen) {
  try {
    const J2Mat m = make_j2(*mat);
    Mat6 Dt;
    j2_eval(m, strain, state, stress, Dt);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) D[6 * i + j] = Dt.m[i][j];
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// mat == NULL: linear elastic (E, nu); otherwise J2 with state (ngp*26).
int orc_element_tangent_force(int type, const double* coords, const double* ue, double E,
                              double nu, const orc_j2* mat, double* state, double* Ke,
                              double* fe, char* err, int errlen) {
  try {
    if (mat) {
      const J2Mat m = make_j2(*mat);
      tangent_force(type, coords, ue, E, nu, &m, state, Ke, fe);
    } else {
      tangent_force(type, coords, ue, E, nu, nullptr, state, Ke, fe);
    }
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// assemble_linear_static (assembly.cpp:225-257) for an SPC + CLOAD model.
int orc_assemble_linear_static(const orc_model* m, orc_system** out, char* err, int errlen) {
  try {
    auto* s = new orc_system();
    DofMap d = dof_map(*m);
    s->n_free = d.n_free;
    s->rhs.assign(size_t(d.n_free), 0.0);
    KMap km;
    Gathered G;
    std::vector<double> Ke;
    for (int64_t e = 0; e < m->n_elem; ++e) {
      if (!is_active(*m, e)) continue;
      gather(*m, e, nullptr, G);
      const int ndof = 3 * G.n;
      Ke.resize(size_t(ndof) * ndof);
      stiffness(m->etype[e], G.x.data(), m->E[e], m->nu[e], Ke.data());
      scatter(Ke.data(), ndof, G.idx.data(), d, km, s->rhs);
    }
    add_cloads(*m, d, s->rhs);
    flush(km, *s);
    s->dof_eq = std::move(d.eq);
    s->prescribed = std::move(d.prescribed);
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// assemble_material_tangent (assembly.cpp:385-423): K_t on the free DOFs and f_int
// on the full DOF set. `state` holds sum_e ngp(e)*26 doubles, element-major.
int orc_assemble_material_tangent(const orc_model* m, const double* u, double* state,
                                  double* f_int, orc_system** out, char* err, int errlen) {
  try {
    auto* s = new orc_system();
    DofMap d = dof_map(*m);
    s->n_free = d.n_free;
    s->rhs.assign(size_t(d.n_free), 0.0);
    std::fill(f_int, f_int + size_t(m->n_nodes) * 3, 0.0);
    std::vector<J2Mat> mats;
    for (int i = 0; i < m->n_j2; ++i) mats.push_back(make_j2(m->j2[i]));
    KMap km;
    Gathered G;
    std::vector<double> Ke, fe;
    size_t soff = 0;
    for (int64_t e = 0; e < m->n_elem; ++e) {
      const int type = m->etype[e];
      const size_t ng = size_t(ngp_of(type));
      if (!is_active(*m, e)) {
        soff += ng * ORC_STATE;
        continue;
      }
      gather(*m, e, u, G);
      const int ndof = 3 * G.n;
      Ke.resize(size_t(ndof) * ndof);
      fe.resize(size_t(ndof));
      const int32_t jid = m->j2_id ? m->j2_id[e] : -1;
      tangent_force(type, G.x.data(), G.u.data(), m->E[e], m->nu[e],
                    jid >= 0 ? &mats[size_t(jid)] : nullptr, state + soff, Ke.data(),
                    fe.data());
      soff += ng * ORC_STATE;
      scatter(Ke.data(), ndof, G.idx.data(), d, km, s->rhs);
      for (int a = 0; a < ndof; ++a) f_int[size_t(G.idx[a / 3]) * 3 + a % 3] += fe[size_t(a)];
    }
    add_cloads(*m, d, s->rhs);
    flush(km, *s);
    s->dof_eq = std::move(d.eq);
    s->prescribed = std::move(d.prescribed);
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// MaterialPoints::commit (assembly.hpp:84-87, material_model.hpp:58-63).
void orc_commit(double* state, int64_t n_points) {
  for (int64_t p = 0; p < n_points; ++p) {
    double* st = state + p * ORC_STATE;
    for (int i = 0; i < 13; ++i) st[i] = st[13 + i];
  }
}

// internal_force (stress.cpp:125-145): f_int = sum_e Ke ue on the full DOF set.
int orc_internal_force(const orc_model* m, const double* u, double* f_int, char* err,
                       int errlen) {
  try {
    std::fill(f_int, f_int + size_t(m->n_nodes) * 3, 0.0);
    Gathered G;
    std::vector<double> Ke;
    for (int64_t e = 0; e < m->n_elem; ++e) {
      if (!is_active(*m, e)) continue;
      gather(*m, e, u, G);
      const int ndof = 3 * G.n;
      Ke.resize(size_t(ndof) * ndof);
      stiffness(m->etype[e], G.x.data(), m->E[e], m->nu[e], Ke.data());
      for (int a = 0; a < ndof; ++a) {
        double v = 0.0;
        for (int b = 0; b < ndof; ++b) v += Ke[size_t(a) * ndof + b] * G.u[size_t(b)];
        f_int[size_t(G.idx[a / 3]) * 3 + a % 3] += v;
      }
    }
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// Vector-Jacobian product of r = K(x,E) u with cotangent w (all DOFs, no BCs):
// dx (n_nodes*3), dE (n_elem), du (n_nodes*3) and, optionally, r itself.
int orc_vjp(const orc_model* m, const double* u, const double* w, double* dx, double* dE,
            double* du, double* r, char* err, int errlen) {
  try {
    const size_t nd = size_t(m->n_nodes) * 3;
    std::fill(dx, dx + nd, 0.0);
    std::fill(du, du + nd, 0.0);
    if (r) {
      if (orc_internal_force(m, u, r, err, errlen)) return 1;
    }
    Gathered G;
    std::vector<double> we, ldx, ldu;
    for (int64_t e = 0; e < m->n_elem; ++e) {
      dE[e] = 0.0;
      if (!is_active(*m, e)) continue;
      gather(*m, e, u, G);
      const int n = G.n;
      we.resize(3 * size_t(n));
      ldx.resize(3 * size_t(n));
      ldu.resize(3 * size_t(n));
      for (int k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c) we[3 * k + c] = w[3 * size_t(G.idx[k]) + c];
      elem_vjp(m->etype[e], G.x.data(), G.u.data(), we.data(), m->E[e], m->nu[e], ldx.data(),
               ldu.data(), &dE[e], nullptr);
      for (int k = 0; k < 3 * n; ++k) {
        dx[3 * size_t(G.idx[k / 3]) + k % 3] += ldx[size_t(k)];
        du[3 * size_t(G.idx[k / 3]) + k % 3] += ldu[size_t(k)];
      }
    }
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// assemble_mass (assembly.cpp:259-297) and assemble_geometric_stiffness (:299-330):
// kind 0 consistent mass, 1 lumped mass, 2 geometric stiffness from gp_stress (per point,
// element-major, 6 each). Scattered through the SPC congruence like the stiffness; the
// load side effect is discarded, so rhs stays zero.
int orc_assemble_scalar(const orc_model* m, int kind, const double* gp_stress, orc_system** out,
                        char* err, int errlen) {
  try {
    auto* s = new orc_system();
    DofMap d = dof_map(*m);
    s->n_free = d.n_free;
    s->rhs.assign(size_t(d.n_free), 0.0);
    std::vector<double> dummy(size_t(d.n_free), 0.0);
    KMap km;
    Gathered G;
    std::vector<double> Ke;
    size_t poff = 0;
    for (int64_t e = 0; e < m->n_elem; ++e) {
      const int type = m->etype[e];
      const size_t ng = size_t(ngp_of(type));
      if (!is_active(*m, e)) {
        poff += ng;
        continue;
      }
      gather(*m, e, nullptr, G);
      const int ndof = 3 * G.n;
      Ke.resize(size_t(ndof) * ndof);
      const double rho = m->rho ? m->rho[e] : 0.0;
      if (kind == 0) mass(type, G.x.data(), rho, Ke.data());
      else if (kind == 1) mass_lumped(type, G.x.data(), rho, Ke.data());
      else geometric(type, G.x.data(), gp_stress ? gp_stress + 6 * poff : nullptr, Ke.data());
      poff += ng;
      scatter(Ke.data(), ndof, G.idx.data(), d, km, dummy);
    }
    flush(km, *s);
    s->dof_eq = std::move(d.eq);
    s->prescribed = std::move(d.prescribed);
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// Field recovery. state == NULL: recover_fields (stress.cpp:189-273), every element
// elastic with f_int = Ke ue (accumulate_internal_force, stress.cpp:52-69) and
// stress = D strain (strain_stress_at, stress.cpp:20-40, no thermal strain). state given
// (26 doubles per point, element-major): the material-point form (stress.cpp:275-353),
// f_int from element_tangent_force and the point stress from the material evaluated on
// a copy of the history. Hex/wedge extrapolation is the Gauss-point mean at every node
// (extrapolation.hpp:37-45); node values are summed over elements in element order and
// divided by the count; RF = f_int - f_ext (CLOADs at full magnitude).
// gauss_stress: the per-point stress (recover_gauss_stress, stress.cpp:147-187, for
// state == NULL), element-major; inactive elements' points stay zero. Outputs may be NULL.
int orc_recover_fields(const orc_model* m, const double* u, const double* state,
                       double* strain, double* stress, double* reaction, double* f_int,
                       double* gauss_stress, char* err, int errlen) {
  try {
    const size_t nn = size_t(m->n_nodes);
    std::vector<double> se(nn * 6, 0.0), ss(nn * 6, 0.0), fi(nn * 3, 0.0);
    std::vector<int> count(nn, 0);
    std::vector<J2Mat> mats;
    for (int i = 0; i < m->n_j2; ++i) mats.push_back(make_j2(m->j2[i]));
    Gathered G;
    std::vector<double> Ke, fe, st;
    double g[kMaxN][3];
    size_t poff = 0;  // first point of element e
    for (int64_t e = 0; e < m->n_elem; ++e) {
      const int type = m->etype[e];
      const auto& R = rule(type);
      const size_t ng = R.size();
      if (!is_active(*m, e)) {
        if (gauss_stress) std::fill(gauss_stress + 6 * poff, gauss_stress + 6 * (poff + ng), 0.0);
        poff += ng;
        continue;
      }
      gather(*m, e, u, G);
      const int n = G.n, ndof = 3 * n;
      const Mat6 De = iso_D(m->E[e], m->nu[e]);
      const int32_t jid = (state && m->j2_id) ? m->j2_id[e] : -1;
      Ke.resize(size_t(ndof) * ndof);
      fe.resize(size_t(ndof));
      if (!state) {
        stiffness(type, G.x.data(), m->E[e], m->nu[e], Ke.data());
        for (int a = 0; a < ndof; ++a) {
          double v = 0.0;
          for (int b = 0; b < ndof; ++b) v += Ke[size_t(a) * ndof + b] * G.u[size_t(b)];
          fi[size_t(G.idx[a / 3]) * 3 + a % 3] += v;
        }
      } else {
        st.assign(state + poff * ORC_STATE, state + (poff + ng) * ORC_STATE);
        tangent_force(type, G.x.data(), G.u.data(), m->E[e], m->nu[e],
                      jid >= 0 ? &mats[size_t(jid)] : nullptr, st.data(), Ke.data(), fe.data());
        for (int a = 0; a < ndof; ++a) fi[size_t(G.idx[a / 3]) * 3 + a % 3] += fe[size_t(a)];
      }
      double me[6] = {0, 0, 0, 0, 0, 0}, ms[6] = {0, 0, 0, 0, 0, 0};
      for (size_t q = 0; q < ng; ++q) {
        const GP& p = R[q];
        const ShapeVals sv = shape_at(type, p.xi, p.et, p.ze);
        gradients(sv, G.x.data(), g);
        double eps[6], sig[6];
        strain_of(g, n, G.u.data(), eps);
        if (jid >= 0) {
          double scratch[ORC_STATE];
          std::copy(state + (poff + q) * ORC_STATE, state + (poff + q + 1) * ORC_STATE, scratch);
          Mat6 Dt;
          j2_eval(mats[size_t(jid)], eps, scratch, sig, Dt);
        } else {
          for (int i = 0; i < 6; ++i) {
            double v = 0.0;
            for (int j = 0; j < 6; ++j) v += De.m[i][j] * (eps[j] - 0.0);
            sig[i] = v;
          }
        }
        for (int c = 0; c < 6; ++c) {
          me[c] += eps[c];
          ms[c] += sig[c];
        }
        if (gauss_stress)
          for (int c = 0; c < 6; ++c) gauss_stress[6 * (poff + q) + c] = sig[c];
      }
      const double inv = 1.0 / double(ng);
      for (int c = 0; c < 6; ++c) {
        me[c] *= inv;
        ms[c] *= inv;
      }
      for (int i = 0; i < n; ++i) {
        const size_t gi = size_t(G.idx[size_t(i)]);
        for (int c = 0; c < 6; ++c) {
          se[6 * gi + c] += me[c];
          ss[6 * gi + c] += ms[c];
        }
        ++count[gi];
      }
      poff += ng;
    }
    for (size_t i = 0; i < nn; ++i)
      if (count[i] > 0)
        for (int c = 0; c < 6; ++c) {
          se[6 * i + c] /= count[i];
          ss[6 * i + c] /= count[i];
        }
    std::vector<double> fx(nn * 3, 0.0);
    for (int i = 0; i < m->n_cload; ++i) {
      const int32_t nd = m->cload_node[i];
      if (nd >= 0 && nd < m->n_nodes && m->cload_comp[i] >= 1 && m->cload_comp[i] <= 3)
        fx[size_t(nd) * 3 + size_t(m->cload_comp[i] - 1)] += 1.0 * m->cload_value[i];
    }
    if (strain) std::copy(se.begin(), se.end(), strain);
    if (stress) std::copy(ss.begin(), ss.end(), stress);
    if (f_int) std::copy(fi.begin(), fi.end(), f_int);
    if (reaction)
      for (size_t k = 0; k < nn * 3; ++k) reaction[k] = fi[k] - fx[k];
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int32_t orc_system_n_free(const orc_system* s) { return s->n_free; }
int64_t orc_system_nnz(const orc_system* s) { return int64_t(s->vals.size()); }

void orc_system_coo(const orc_system* s, int32_t* rows, int32_t* cols, double* vals) {
  std::copy(s->rows.begin(), s->rows.end(), rows);
  std::copy(s->cols.begin(), s->cols.end(), cols);
  std::copy(s->vals.begin(), s->vals.end(), vals);
}

// Canonical CSR of the summed COO: rows ascending, columns ascending in a row.
void orc_system_csr(const orc_system* s, int64_t* row_ptr, int32_t* col_idx, double* vals) {
  const size_t nnz = s->vals.size();
  std::vector<size_t> ord(nnz);
  for (size_t i = 0; i < nnz; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
    return s->rows[a] != s->rows[b] ? s->rows[a] < s->rows[b] : s->cols[a] < s->cols[b];
  });
  std::fill(row_ptr, row_ptr + s->n_free + 1, 0);
  for (size_t i = 0; i < nnz; ++i) row_ptr[s->rows[i] + 1] += 1;
  for (int32_t r = 0; r < s->n_free; ++r) row_ptr[r + 1] += row_ptr[r];
  for (size_t i = 0; i < nnz; ++i) {
    col_idx[i] = s->cols[ord[i]];
    vals[i] = s->vals[ord[i]];
  }
}

void orc_system_rhs(const orc_system* s, double* rhs) {
  std::copy(s->rhs.begin(), s->rhs.end(), rhs);
}

void orc_system_dof_eq(const orc_system* s, int32_t* dof_eq) {
  std::copy(s->dof_eq.begin(), s->dof_eq.end(), dof_eq);
}

void orc_system_free(orc_system* s) { delete s; }

}  // extern "C"
