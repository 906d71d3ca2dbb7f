// Element kernels of libcxg.so for gfx950 (CDNA4, 64-lane wavefronts).
//
// One element per 64-lane wavefront (one wave per workgroup). The element's node
// coordinates, DOF map, per-Gauss-point J^-1 and the physical gradients g[q][k] of
// every Gauss point are staged in LDS; the element stiffness never exists in
// memory: each lane owns node-pair blocks K_ab (3x3, nine FP64 accumulators in
// registers), integrates them over the Gauss points and adds them straight into
// the CSR value array through the element -> CSR slot map.
//
// Determinism: kernels are launched colour by colour; no two elements of one
// colour share a node, so every CSR value, f_int, dx and du entry has exactly one
// writer per launch and is updated by a plain read-modify-write. The summation
// order is fixed (colour order, then Gauss-point order), independent of scheduling;
// no floating-point atomics are used anywhere.
//
// Reference algorithm (restated, not translated):
//   Jacobian / gradients   src/fem/element.cpp:80-102, 361-380
//   B, D, B^T D B          src/fem/element.cpp:382-462 (isotropic D exploited:
//                          K_ab = lam g_a g_b^T + mu g_b g_a^T + mu (g_a.g_b) I)
//   internal force         src/fem/stress.cpp:125-145 (as sum_q s B^T D B u)
//   J2 return map/tangent  src/fem/plasticity.cpp:90-236
//   scatter                src/fem/assembly.cpp:76-110 (SPC-only transform)
#include <algorithm>
#include <climits>
#include <limits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "cxg_internal.hpp"

// Element-force arena of the allocation-free internal force (cxg_internal_force_noalloc):
// a fixed device array, 3 doubles per connectivity entry, kScratchDoubles at most. Each
// device that loads the library holds its own copy until the process ends. In host code
// `scratch` names a host-side stand-in, never a device address: kernels receive the
// address scratch_device_ptr() obtained from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[cxg::kScratchDoubles];

namespace cxg {

double* scratch_device_ptr(int device) {
  // One lookup per device, with that device current (cxg_create calls this after
  // hipSetDevice); every later context on the device reuses the address.
  static std::mutex mu;
  static std::vector<double*> table;
  std::lock_guard<std::mutex> lock(mu);
  if (device < 0) throw Error(CXG_ERR_ARG, "bad device");
  if (table.size() <= size_t(device)) table.resize(size_t(device) + 1, nullptr);
  if (!table[size_t(device)]) {
    void* ptr = nullptr;
    CXG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    table[size_t(device)] = static_cast<double*>(ptr);
  }
  return table[size_t(device)];
}
namespace {

constexpr int kWave = 64;

__global__ void k_fill(double* __restrict__ p, int64_t n, double v) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Block b of a launch of n blocks -> work index: with slab > 0 each of the 8 XCDs (which
// receive consecutive workgroups in turn) walks one contiguous range of slab indices;
// -1 past the end. slab 0: identity.
__device__ __forceinline__ int64_t xcd_index(int64_t b, int64_t n, int64_t slab) {
  const int64_t i = slab > 0 ? (b % 8) * slab + b / 8 : b;
  return i < n ? i : -1;
}

// ---- LDS staging shared by all element kernels ------------------------------
template <int NN, int NGP>
struct Geo {
  double x[NN * 3];
  double Ai[NGP * 9];      // J^-1 (row-major) per Gauss point
  double sc[NGP];          // det J * w
  double g[NGP * NN * 3];  // physical gradients, [q][k][i]
};

template <int NN>
struct DofLds {
  int64_t row[NN * 3];  // CSR row start of each element DOF (-1: constrained row)
  int32_t col[NN * 3];  // offset of the DOF inside its node's column block (-1: constrained)
};

template <int NN, int NGP>
__device__ __forceinline__ void stage_coords(Geo<NN, NGP>& G, int32_t* node,
                                             const int32_t* __restrict__ ce,
                                             const double* __restrict__ coords, int lane) {
  if (lane < NN) {
    const int32_t nd = ce[lane];
    node[lane] = nd;
    const double* xp = coords + 3 * int64_t(nd);
    G.x[3 * lane + 0] = xp[0];
    G.x[3 * lane + 1] = xp[1];
    G.x[3 * lane + 2] = xp[2];
  }
}

template <int NN>
__device__ __forceinline__ void stage_vec(double* dst, const int32_t* __restrict__ ce,
                                          const double* __restrict__ v, int lane) {
  if (lane < NN) {
    const double* p = v + 3 * int64_t(ce[lane]);
    dst[3 * lane + 0] = p[0];
    dst[3 * lane + 1] = p[1];
    dst[3 * lane + 2] = p[2];
  }
}

template <int NN>
__device__ __forceinline__ void stage_dofs(DofLds<NN>& D, const int32_t* __restrict__ ce,
                                           const int32_t* __restrict__ dof_eq,
                                           const int64_t* __restrict__ row_ptr, int lane) {
  if (lane < 3 * NN) {
    const int k = lane / 3, c = lane % 3;
    const int64_t base = 3 * int64_t(ce[k]);
    const int32_t eq = dof_eq[base + c];
    int before = 0;
    for (int cc = 0; cc < c; ++cc) before += dof_eq[base + cc] >= 0 ? 1 : 0;
    D.row[lane] = eq >= 0 ? row_ptr[eq] : -1;
    D.col[lane] = eq >= 0 ? before : -1;
  }
}

// J, det J and J^-1 per Gauss point; one lane per Gauss point.
template <int NN, int NGP>
__device__ __forceinline__ void jacobians(Geo<NN, NGP>& G, const double* __restrict__ dN,
                                          const double* __restrict__ wq, int lane,
                                          int* __restrict__ flag) {
  for (int q = lane; q < NGP; q += kWave) {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double* d = dN + q * NN * 3;
#pragma unroll
    for (int k = 0; k < NN; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) J[i][j] += G.x[3 * k + i] * d[3 * k + j];
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    if (!(det > 0.0)) atomicOr(flag, kFlagJacobian);
    const double r = 1.0 / det;
    double* A = G.Ai + 9 * q;
    A[0] = c00 * r;
    A[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
    A[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    A[3] = c01 * r;
    A[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
    A[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    A[6] = c02 * r;
    A[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
    A[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
    G.sc[q] = det * wq[q];
  }
}

// g_k = J^-T dN_k for every (Gauss point, node).
template <int NN, int NGP>
__device__ __forceinline__ void gradients(Geo<NN, NGP>& G, const double* __restrict__ dN,
                                          int lane) {
  for (int idx = lane; idx < NGP * NN; idx += kWave) {
    const int q = idx / NN;
    const double* A = G.Ai + 9 * q;
    const double d0 = dN[3 * idx], d1 = dN[3 * idx + 1], d2 = dN[3 * idx + 2];
    G.g[3 * idx + 0] = A[0] * d0 + A[3] * d1 + A[6] * d2;
    G.g[3 * idx + 1] = A[1] * d0 + A[4] * d1 + A[7] * d2;
    G.g[3 * idx + 2] = A[2] * d0 + A[5] * d1 + A[8] * d2;
  }
}

// Displacement gradient H[i][j] = sum_k v_k[i] g_k[j] at Gauss point q.
template <int NN, int NGP>
__device__ __forceinline__ void grad_field(const Geo<NN, NGP>& G, const double* v, int q,
                                           double H[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) H[i][j] = 0.0;
  const double* g = G.g + q * NN * 3;
#pragma unroll
  for (int k = 0; k < NN; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) H[i][j] += v[3 * k + i] * g[3 * k + j];
}

__device__ __forceinline__ void voigt_strain(const double H[3][3], double e[6]) {
  e[0] = H[0][0];
  e[1] = H[1][1];
  e[2] = H[2][2];
  e[3] = H[0][1] + H[1][0];
  e[4] = H[0][2] + H[2][0];
  e[5] = H[1][2] + H[2][1];
}

// sigma = D_iso eps (engineering-shear strain).
__device__ __forceinline__ void iso_stress(double lam, double mu, const double e[6], double s[6]) {
  const double tr = lam * (e[0] + e[1] + e[2]);
  s[0] = tr + 2.0 * mu * e[0];
  s[1] = tr + 2.0 * mu * e[1];
  s[2] = tr + 2.0 * mu * e[2];
  s[3] = mu * e[3];
  s[4] = mu * e[4];
  s[5] = mu * e[5];
}

__device__ __forceinline__ void lame(double E, double nu, double& lam, double& mu) {
  const double two_mu = E / (1.0 + nu);
  lam = nu * two_mu / (1.0 - 2.0 * nu);
  mu = two_mu / 2.0;
}

// f_k = sum_q B_k^T S_q where S_q already carries det J * w. Written (RMW) into the
// full-DOF vector at the element's nodes; one lane per node.
template <int NN, int NGP>
__device__ __forceinline__ void node_forces(const Geo<NN, NGP>& G, const double* S,
                                            const int32_t* node, double* __restrict__ f,
                                            int lane) {
  for (int k = lane; k < NN; k += kWave) {
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int q = 0; q < NGP; ++q) {
      const double* g = G.g + (q * NN + k) * 3;
      const double* s = S + 6 * q;
      f0 += g[0] * s[0] + g[1] * s[3] + g[2] * s[4];
      f1 += g[1] * s[1] + g[0] * s[3] + g[2] * s[5];
      f2 += g[2] * s[2] + g[0] * s[4] + g[1] * s[5];
    }
    double* fp = f + 3 * int64_t(node[k]);
    fp[0] += f0;
    fp[1] += f1;
    fp[2] += f2;
  }
}

// Node pair p -> (a, b). Upper numbering walks a-major over b >= a.
template <int NN, bool UPPER>
__device__ __forceinline__ void decode_pair(int p, int& a, int& b) {
  if (UPPER) {
    int aa = 0, rem = p;
    while (rem >= NN - aa) {
      rem -= NN - aa;
      ++aa;
    }
    a = aa;
    b = aa + rem;
  } else {
    a = p / NN;
    b = p % NN;
  }
}

// Lane -> (node a, first partner b0) of its group of up to G consecutive upper pairs
// (a, b0 .. b0+G-1), b >= a, groups ordered a-major; (-1, -1) past the last group.
template <int NN, int G>
__device__ __forceinline__ void decode_group(int lane, int& a, int& b0) {
  int rem = lane;
  a = -1;
  b0 = -1;
  for (int aa = 0; aa < NN; ++aa) {
    const int items = (NN - aa + G - 1) / G;
    if (rem < items) {
      a = aa;
      b0 = aa + rem * G;
      return;
    }
    rem -= items;
  }
}

// Add the 3x3 block K (rows: DOFs of node a, cols: DOFs of node b) into the CSR
// values; with `mirror` also its transpose into the rows of node b.
template <int NN>
__device__ __forceinline__ void scatter_pair(const DofLds<NN>& D, const int32_t* __restrict__ se,
                                             int a, int b, const double K[9], bool mirror,
                                             double* __restrict__ vals) {
  const int32_t cab = se[a * NN + b] & kSlotMask;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int64_t r = D.row[3 * a + i];
    if (r < 0) continue;
    double* rp = vals + r + cab;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int32_t c = D.col[3 * b + j];
      if (c >= 0) rp[c] += K[3 * i + j];
    }
  }
  if (mirror) {
    const int32_t cba = se[b * NN + a] & kSlotMask;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int64_t r = D.row[3 * b + j];
      if (r < 0) continue;
      double* rp = vals + r + cba;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int32_t c = D.col[3 * a + i];
        if (c >= 0) rp[c] += K[3 * i + j];
      }
    }
  }
}

// ---- K1+K4(+K5): isotropic elastic stiffness straight into CSR ---------------
template <int NN, int NGP, bool UPPER>
struct AsmLds {
  Geo<NN, NGP> G;
  DofLds<NN> D;
  double u[NN * 3];
  double S[NGP * 6];
  int32_t node[NN];
};

template <int NN, int NGP, bool UPPER>
__global__ __launch_bounds__(kWave) void k_assemble_iso(
    const int32_t* __restrict__ conn, const int32_t* __restrict__ elist,
    const int32_t* __restrict__ pptr, int64_t elem_offset,
    const double* __restrict__ dN, const double* __restrict__ wq,
    const double* __restrict__ coords, const double* __restrict__ E,
    const double* __restrict__ nu, const int32_t* __restrict__ dof_eq,
    const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ slot,
    double* __restrict__ vals, const double* __restrict__ u, double* __restrict__ f_int,
    int* __restrict__ flag) {
  __shared__ AsmLds<NN, NGP, UPPER> L;
  const int lane = threadIdx.x;
  const int32_t it_end = pptr[blockIdx.x + 1];
  for (int32_t it = pptr[blockIdx.x]; it < it_end; ++it) {
  const int64_t le = elist[it];
  const int32_t* ce = conn + le * NN;
  stage_coords<NN, NGP>(L.G, L.node, ce, coords, lane);
  stage_dofs<NN>(L.D, ce, dof_eq, row_ptr, lane);
  if (f_int) stage_vec<NN>(L.u, ce, u, lane);
  __syncthreads();
  jacobians<NN, NGP>(L.G, dN, wq, lane, flag);
  __syncthreads();
  gradients<NN, NGP>(L.G, dN, lane);
  __syncthreads();
  double lam, mu;
  lame(E[elem_offset + le], nu[elem_offset + le], lam, mu);
  if (f_int) {
    for (int q = lane; q < NGP; q += kWave) {
      double H[3][3], e[6], s[6];
      grad_field<NN, NGP>(L.G, L.u, q, H);
      voigt_strain(H, e);
      iso_stress(lam, mu, e, s);
      for (int i = 0; i < 6; ++i) L.S[6 * q + i] = L.G.sc[q] * s[i];
    }
    __syncthreads();
    node_forces<NN, NGP>(L.G, L.S, L.node, f_int, lane);
  }
  const int32_t* se = slot + le * NN * NN;
  constexpr int NP = UPPER ? NN * (NN + 1) / 2 : NN * NN;
  for (int p = lane; p < NP; p += kWave) {
    int a, b;
    decode_pair<NN, UPPER>(p, a, b);
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < NGP; ++q) {
      const double* ga = L.G.g + (q * NN + a) * 3;
      const double* gb = L.G.g + (q * NN + b) * 3;
      const double ls = lam * L.G.sc[q], ms = mu * L.G.sc[q];
      const double la0 = ls * ga[0], la1 = ls * ga[1], la2 = ls * ga[2];
      const double ma0 = ms * ga[0], ma1 = ms * ga[1], ma2 = ms * ga[2];
      const double gb0 = gb[0], gb1 = gb[1], gb2 = gb[2];
      const double dd = ma0 * gb0 + ma1 * gb1 + ma2 * gb2;
      K[0] += la0 * gb0 + ma0 * gb0 + dd;
      K[1] += la0 * gb1 + ma1 * gb0;
      K[2] += la0 * gb2 + ma2 * gb0;
      K[3] += la1 * gb0 + ma0 * gb1;
      K[4] += la1 * gb1 + ma1 * gb1 + dd;
      K[5] += la1 * gb2 + ma2 * gb1;
      K[6] += la2 * gb0 + ma0 * gb2;
      K[7] += la2 * gb1 + ma1 * gb2;
      K[8] += la2 * gb2 + ma2 * gb2 + dd;
    }
    scatter_pair<NN>(L.D, se, a, b, K, UPPER && a != b, vals);
  }
  __syncthreads();  // LDS is reused by the next element of the patch
  }
}

// ---- K1+K4(+K5), variant 2 ------------------------------------------------------
// For isotropic D the node-pair block is
//   K_ab = lam M_ab + mu M_ab^T + mu tr(M_ab) I,   M_ab = sum_q h_a^q (h_b^q)^T,
// with h_k^q = sqrt(det J_q w_q) g_k^q. Each lane accumulates the 3x3 M of up to PPL
// upper node pairs (a-major, so consecutive pairs of a lane share node a): 9 FMA
// per pair and Gauss point. Gauss points are processed QC at a time so only the h
// of QC points sits in LDS. The scatter is row-coalesced: for each element node a
// the blocks K_ab of all b are staged in LDS and lane (b, j) adds column (b, j) of
// the three rows of a, so one wave instruction updates one CSR row segment.
template <int NN, int NGP, int QC>
struct Asm2Lds {
  double x[NN * 3];
  double u[NN * 3];
  double Ai[NGP * 9];
  double rs[NGP];        // sqrt(det J * w)
  double h[QC * NN * 3]; // h of the current Gauss-point chunk, [qq][k][i]
  double S[QC * 6];      // D eps(h, u) of the chunk (residual)
  double kb[NN * 9];     // staged blocks K_ab of the current row node a, [b][i][j]
  int64_t row[NN * 3];
  int32_t col[NN * 3];
  int32_t slot[NN * NN];
  int32_t node[NN];
};

template <int NN, int NGP, int QC, int WAVES>
__global__ __launch_bounds__(kWave, WAVES) void k_assemble_iso2(
    const int32_t* __restrict__ conn, const int32_t* __restrict__ elist,
    const int32_t* __restrict__ pptr, int64_t elem_offset, const double* __restrict__ dN,
    const double* __restrict__ wq, const double* __restrict__ coords,
    const double* __restrict__ E, const double* __restrict__ nu,
    const int32_t* __restrict__ dof_eq, const int64_t* __restrict__ row_ptr,
    const int32_t* __restrict__ slot, double* __restrict__ vals, const double* __restrict__ u,
    double* __restrict__ f_int, int* __restrict__ flag) {
  constexpr int NP = NN * (NN + 1) / 2;
  constexpr int PPL = (NP + kWave - 1) / kWave;
  __shared__ Asm2Lds<NN, NGP, QC> L;
  const int lane = threadIdx.x;
  int pa[PPL], pb[PPL];
#pragma unroll
  for (int t = 0; t < PPL; ++t) {
    const int p = lane * PPL + t;
    if (p < NP) {
      decode_pair<NN, true>(p, pa[t], pb[t]);
    } else {
      pa[t] = -1;
      pb[t] = -1;
    }
  }
  const int32_t it_end = pptr[blockIdx.x + 1];
  for (int32_t it = pptr[blockIdx.x]; it < it_end; ++it) {
    const int64_t le = elist[it];
    const int32_t* ce = conn + le * NN;
    if (lane < NN) {
      const int32_t nd = ce[lane];
      L.node[lane] = nd;
      const double* xp = coords + 3 * int64_t(nd);
      L.x[3 * lane + 0] = xp[0];
      L.x[3 * lane + 1] = xp[1];
      L.x[3 * lane + 2] = xp[2];
      if (f_int) {
        const double* up = u + 3 * int64_t(nd);
        L.u[3 * lane + 0] = up[0];
        L.u[3 * lane + 1] = up[1];
        L.u[3 * lane + 2] = up[2];
      }
    }
    if (lane < 3 * NN) {
      const int k = lane / 3, c = lane % 3;
      const int64_t base = 3 * int64_t(ce[k]);
      const int32_t eq = dof_eq[base + c];
      int before = 0;
      for (int cc = 0; cc < c; ++cc) before += dof_eq[base + cc] >= 0 ? 1 : 0;
      L.row[lane] = eq >= 0 ? row_ptr[eq] : -1;
      L.col[lane] = eq >= 0 ? before : -1;
    }
    const int32_t* se = slot + le * NN * NN;
    for (int i = lane; i < NN * NN; i += kWave) L.slot[i] = se[i] & kSlotMask;
    __syncthreads();
    for (int q = lane; q < NGP; q += kWave) {
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      const double* d = dN + q * NN * 3;
      for (int k = 0; k < NN; ++k) {
        const double x0 = L.x[3 * k], x1 = L.x[3 * k + 1], x2 = L.x[3 * k + 2];
        const double d0 = d[3 * k], d1 = d[3 * k + 1], d2 = d[3 * k + 2];
        J[0][0] += x0 * d0; J[0][1] += x0 * d1; J[0][2] += x0 * d2;
        J[1][0] += x1 * d0; J[1][1] += x1 * d1; J[1][2] += x1 * d2;
        J[2][0] += x2 * d0; J[2][1] += x2 * d1; J[2][2] += x2 * d2;
      }
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      if (!(det > 0.0)) atomicOr(flag, kFlagJacobian);
      const double r = 1.0 / det;
      double* A = L.Ai + 9 * q;
      A[0] = c00 * r;
      A[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
      A[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
      A[3] = c01 * r;
      A[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
      A[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
      A[6] = c02 * r;
      A[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
      A[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
      L.rs[q] = sqrt(det * wq[q]);
    }
    double lam, mu;
    lame(E[elem_offset + le], nu[elem_offset + le], lam, mu);
    double M[PPL][9];
#pragma unroll
    for (int t = 0; t < PPL; ++t)
#pragma unroll
      for (int i = 0; i < 9; ++i) M[t][i] = 0.0;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    __syncthreads();
    for (int q0 = 0; q0 < NGP; q0 += QC) {
      for (int idx = lane; idx < QC * NN; idx += kWave) {
        const int q = q0 + idx / NN;
        double h0 = 0.0, h1 = 0.0, h2 = 0.0;
        if (q < NGP) {
          const double* A = L.Ai + 9 * q;
          const double* d = dN + (q * NN + idx % NN) * 3;
          const double s = L.rs[q];
          h0 = s * (A[0] * d[0] + A[3] * d[1] + A[6] * d[2]);
          h1 = s * (A[1] * d[0] + A[4] * d[1] + A[7] * d[2]);
          h2 = s * (A[2] * d[0] + A[5] * d[1] + A[8] * d[2]);
        }
        L.h[3 * idx + 0] = h0;
        L.h[3 * idx + 1] = h1;
        L.h[3 * idx + 2] = h2;
      }
      __syncthreads();
      if (f_int) {
        if (lane < QC) {
          double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
          const double* hq = L.h + lane * NN * 3;
          for (int k = 0; k < NN; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j < 3; ++j) H[i][j] += L.u[3 * k + i] * hq[3 * k + j];
          double e[6], s[6];
          voigt_strain(H, e);
          iso_stress(lam, mu, e, s);
#pragma unroll
          for (int i = 0; i < 6; ++i) L.S[6 * lane + i] = s[i];
        }
        __syncthreads();
        if (lane < NN) {
#pragma unroll
          for (int qq = 0; qq < QC; ++qq) {
            const double* g = L.h + (qq * NN + lane) * 3;
            const double* s = L.S + 6 * qq;
            f0 += g[0] * s[0] + g[1] * s[3] + g[2] * s[4];
            f1 += g[1] * s[1] + g[0] * s[3] + g[2] * s[5];
            f2 += g[2] * s[2] + g[0] * s[4] + g[1] * s[5];
          }
        }
      }
#pragma unroll
      for (int t = 0; t < PPL; ++t) {
        if (pa[t] < 0) continue;
#pragma unroll 1
        for (int qq = 0; qq < QC; ++qq) {
          const double* ha = L.h + (qq * NN + pa[t]) * 3;
          const double* hb = L.h + (qq * NN + pb[t]) * 3;
          const double a0 = ha[0], a1 = ha[1], a2 = ha[2];
          const double b0 = hb[0], b1 = hb[1], b2 = hb[2];
          M[t][0] += a0 * b0; M[t][1] += a0 * b1; M[t][2] += a0 * b2;
          M[t][3] += a1 * b0; M[t][4] += a1 * b1; M[t][5] += a1 * b2;
          M[t][6] += a2 * b0; M[t][7] += a2 * b1; M[t][8] += a2 * b2;
        }
      }
      __syncthreads();
    }
    if (f_int && lane < NN) {
      double* fp = f_int + 3 * int64_t(L.node[lane]);
      fp[0] += f0;
      fp[1] += f1;
      fp[2] += f2;
    }
    // K blocks from M, staged per row node a and added row-coalesced.
    for (int a = 0; a < NN; ++a) {
#pragma unroll
      for (int t = 0; t < PPL; ++t) {
        const bool lead = pa[t] == a, trail = pb[t] == a && pa[t] != a;
        if (!lead && !trail) continue;
        const double* m = M[t];
        const double tr = mu * (m[0] + m[4] + m[8]);
        double* dst = L.kb + 9 * (lead ? pb[t] : pa[t]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const double kij = lam * m[3 * i + j] + mu * m[3 * j + i] + (i == j ? tr : 0.0);
            if (lead) dst[3 * i + j] = kij;
            else dst[3 * j + i] = kij;
          }
      }
      __syncthreads();
      if (lane < 3 * NN) {
        const int b = lane / 3, j = lane % 3;
        const int32_t c = L.col[lane];
        if (c >= 0) {
          const int64_t off = int64_t(L.slot[a * NN + b]) + c;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const int64_t r = L.row[3 * a + i];
            if (r >= 0) vals[r + off] += L.kb[9 * b + 3 * i + j];
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---- K1, two-phase variant, phase A: element matrices to an arena ------------------
// Same M-identity arithmetic as k_assemble_iso2. Each lane forms K_ab once for its
// (up to PPL, a-major consecutive) upper node pairs and stores the block into the
// rows of a and its transpose into the rows of b of the element's full row-major Ke,
// so that phase B reads whole contiguous rows. The whole 3n x 3n region is written by
// one wave within a short time, which the L2 merges into full lines. No colouring,
// no read-modify-write.
constexpr int kStageNodes = 4;  // element-matrix node rows staged in LDS per store turn
typedef double f64x2 __attribute__((ext_vector_type(2)));  // one 16-byte store

// det J * w of one Gauss point from the node coordinates x[3n] and the natural derivatives
// d[n][3] (at d[(3k + c) S]: S = 1 for a point's row of the dN table, S = n_gp for the
// point's column of its transpose), and with INV also J^-1 (row-major, cofactors) into
// A; raises kFlagJacobian when det J <= 0 (element.cpp:86-87).
template <int NN, bool INV, int S = 1>
__device__ __forceinline__ double jacobian_detw(const double* x, const double* __restrict__ d,
                                                const double w, double* A,
                                                int* __restrict__ flag) {
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int k = 0; k < NN; ++k) {
    const double x0 = x[3 * k], x1 = x[3 * k + 1], x2 = x[3 * k + 2];
    const double d0 = d[(3 * k) * S], d1 = d[(3 * k + 1) * S], d2 = d[(3 * k + 2) * S];
    J[0][0] += x0 * d0; J[0][1] += x0 * d1; J[0][2] += x0 * d2;
    J[1][0] += x1 * d0; J[1][1] += x1 * d1; J[1][2] += x1 * d2;
    J[2][0] += x2 * d0; J[2][1] += x2 * d1; J[2][2] += x2 * d2;
  }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  if (!(det > 0.0)) atomicOr(flag, kFlagJacobian);
  if constexpr (INV) {
    const double r = 1.0 / det;
    A[0] = c00 * r;
    A[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
    A[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    A[3] = c01 * r;
    A[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
    A[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    A[6] = c02 * r;
    A[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
    A[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
  }
  return det * w;
}

// J^-1 into A and sqrt(det J * w) of one Gauss point (see jacobian_detw).
template <int NN, int S = 1>
__device__ __forceinline__ void jacobian_inverse(const double* x, const double* __restrict__ d,
                                                 const double w, double* A, double& rs,
                                                 int* __restrict__ flag) {
  rs = sqrt(jacobian_detw<NN, true, S>(x, d, w, A, flag));
}

// The natural derivatives of point q for a lane that takes one point: families of at
// most 8 points read the transposed table that the device table of every family holds
// after the point-major one ((3k + c) n_gp + q: the lanes' reads consecutive); the
// 27-point families read their point-major row (3 n consecutive doubles per lane, in
// 16-byte loads), which measured faster for them (recovery 0.76 against 1.03 ms at 64^3
// C3D20). Element k component c is at dn_point(dN, q)[(3k + c) dn_stride].
template <int NGP>
constexpr int dn_stride() { return NGP <= 8 ? NGP : 1; }
template <int NN, int NGP>
__device__ __forceinline__ const double* dn_point(const double* dN, int q) {
  return NGP <= 8 ? dN + NGP * NN * 3 + q : dN + q * NN * 3;
}

// h = sqrt(det J w) J^-T dN of one node at one Gauss point (A = J^-1 row-major, d = the
// reference gradient), with the multiply-adds written out, so that every kernel forming
// h gets the same bits.
__device__ __forceinline__ void scaled_gradient(const double* A, const double s, const double* d,
                                                double h[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
    h[i] = s * __builtin_fma(A[6 + i], d[2], __builtin_fma(A[3 + i], d[1], A[i] * d[0]));
}

template <int NN, int NGP, int QC>
struct ElemLds {
  double x[NN * 3];
  double Ai[NGP * 9];
  double rs[NGP];
  alignas(16) double h[QC * NN * 3];
  alignas(16) double stage[kStageNodes * 3 * 3 * NN];
};

// Packed index of the upper node-pair block (a, b), b >= a, of an n-node element: row a
// of the block triangle holds (a, a) .. (a, n-1) after the rows above it.
__host__ __device__ __forceinline__ int up_block(int a, int b, int n) {
  return a * n - a * (a - 1) / 2 + (b - a);
}

// Staged store of a lane's upper node-pair blocks K[t] = K_(pa[t], pb[t]) (pb >= pa; pa <
// 0: none) into the element's packed upper-block arena kel, the block rows of
// kStageNodes nodes at a time. A turn's blocks are one contiguous run of the packed
// arena; each lane writes its blocks into stage at their packed place (9 consecutive
// doubles each), and the wave then copies the run out with 16-byte streaming stores
// (up_block(g0, g0) is even for g0 % 4 == 0, so the run starts 16-byte aligned). stage
// holds kStageNodes * NN blocks, 16-byte aligned. Every lane of the wave calls it; it
// ends with a barrier.
// NTS: streaming (non-temporal) stores, for an arena read back only after the whole
// element pass; false: cached stores, for a ring read back while still cache-resident.
template <int NN, int PPL, bool NTS = true>
__device__ __forceinline__ void store_upper_blocks(double* __restrict__ stage, const int* pa,
                                                   const int* pb, const double (*K)[9],
                                                   double* __restrict__ kel, const int lane) {
  constexpr int NP = NN * (NN + 1) / 2;
  static_assert(kStageNodes % 4 == 0, "turns must start at even packed blocks");
  for (int g0 = 0; g0 < NN; g0 += kStageNodes) {
    const int g1 = NN - g0 < kStageNodes ? NN : g0 + kStageNodes;
    const int k0 = up_block(g0, g0, NN);
    const int k1 = g1 < NN ? up_block(g1, g1, NN) : NP;
#pragma unroll
    for (int t = 0; t < PPL; ++t) {
      if (pa[t] < g0 || pa[t] >= g1) continue;
      double* sb = stage + 9 * (up_block(pa[t], pb[t], NN) - k0);
#pragma unroll
      for (int r = 0; r < 9; ++r) sb[r] = K[t][r];
    }
    __syncthreads();
    const int cnt = 9 * (k1 - k0);
    f64x2* dst = reinterpret_cast<f64x2*>(kel + 9 * k0);
    const f64x2* src = reinterpret_cast<const f64x2*>(stage);
    if constexpr (NTS) {
      for (int idx = lane; idx < cnt / 2; idx += kWave) __builtin_nontemporal_store(src[idx], dst + idx);
      if ((cnt & 1) && lane == 0) __builtin_nontemporal_store(stage[cnt - 1], kel + 9 * k0 + cnt - 1);
    } else {
      for (int idx = lane; idx < cnt / 2; idx += kWave) dst[idx] = src[idx];
      if ((cnt & 1) && lane == 0) kel[9 * k0 + cnt - 1] = stage[cnt - 1];
    }
    __syncthreads();
  }
}

// One element: ce its connectivity row, kel its arena slot. UP = false: the full 3n x 3n
// matrix, row-major. UP = true: only the upper node-pair blocks K_ab (b >= a), packed in
// up_block order, each 3x3 row-major; the element matrix is symmetric (K_ba = K_ab^T),
// so the row gather reads the lower blocks transposed. Every lane of the wave calls it;
// it ends with a barrier, so L can be reused at once.
//
// TL = 1 (PPL = 4 only): a lane owns the 2x2 node pairs {2s, 2s+1} x {2t, 2t+1} of one
// upper pair of node couples s <= t (of the lower pair (2s+1, 2s) of a diagonal couple
// nothing is stored), and loads the two couples' gradients with 16-byte LDS reads: 6
// reads of 96 bytes per Gauss point against 15 of 120 bytes for TL = 0.
template <int NN, int NGP, int QC, bool UP = false, int TL = 0, bool NTS = true>
__device__ __forceinline__ void elem_iso_body(ElemLds<NN, NGP, QC>& L, const int lane,
                                              const int32_t* __restrict__ ce, const double Ee,
                                              const double nue, const double* __restrict__ dN,
                                              const double* __restrict__ wq,
                                              const double* __restrict__ coords,
                                              double* __restrict__ kel, int* __restrict__ flag) {
  constexpr int NP = NN * (NN + 1) / 2;
  constexpr int PPL = (NP + kWave - 1) / kWave;
  static_assert(TL == 0 || (PPL == 4 && NN % 2 == 0), "couple tiling needs four pairs per lane");
  // TL = 0: a lane owns node a and up to PPL consecutive partners b (pairs
  // (a, b0..b0+PPL-1)), so that h_a is loaded once per Gauss point for all of them.
  int pa[PPL], pb[PPL];
  int ga, gb0;
  int cs = 0, ct = 0;  // TL = 1: the lane's node couples (clamped for idle lanes)
  if constexpr (TL == 0) {
    decode_group<NN, PPL>(lane, ga, gb0);
#pragma unroll
    for (int t = 0; t < PPL; ++t) {
      const bool ok = ga >= 0 && gb0 + t < NN;
      pa[t] = ok ? ga : -1;
      pb[t] = ok ? gb0 + t : -1;
    }
  } else {
    int s_ = -1, t_ = -1;
    decode_group<NN / 2, 1>(lane, s_, t_);
    ga = gb0 = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int a = 2 * s_ + t / 2, b = 2 * t_ + t % 2;
      const bool ok = s_ >= 0 && a <= b;
      pa[t] = ok ? a : -1;
      pb[t] = ok ? b : -1;
    }
    cs = s_ >= 0 ? s_ : 0;
    ct = s_ >= 0 ? t_ : 0;
  }
  const int ga_c = ga >= 0 ? ga : 0;  // clamped: unowned slots compute unused values
  if (lane < NN) {
    const double* xp = coords + 3 * int64_t(ce[lane]);
    L.x[3 * lane + 0] = xp[0];
    L.x[3 * lane + 1] = xp[1];
    L.x[3 * lane + 2] = xp[2];
  }
  __syncthreads();
  for (int q = lane; q < NGP; q += kWave)
    jacobian_inverse<NN, dn_stride<NGP>()>(L.x, dn_point<NN, NGP>(dN, q), wq[q], L.Ai + 9 * q, L.rs[q], flag);
  double M[PPL][9];
#pragma unroll
  for (int t = 0; t < PPL; ++t)
#pragma unroll
    for (int i = 0; i < 9; ++i) M[t][i] = 0.0;
  __syncthreads();
  // A chunk's (Gauss point, node) gradients take one pass of the wave; lane cq * NN + cn
  // forms those of point q0 + cq, node cn, from the reference gradients it loaded during
  // the previous chunk.
  static_assert(QC * NN <= kWave, "one pass per chunk");
  const int cq = lane / NN, cn = lane - NN * (lane / NN);
  const auto load_dn = [&](int q0, double (&d)[3]) {
    const int q = q0 + cq;
    d[0] = d[1] = d[2] = 0.0;
    if (lane < QC * NN && q < NGP) {
      const double* p = dN + (q * NN + cn) * 3;
      d[0] = p[0];
      d[1] = p[1];
      d[2] = p[2];
    }
  };
  double d[3];
  load_dn(0, d);
  for (int q0 = 0; q0 < NGP; q0 += QC) {
    if (lane < QC * NN) {
      const int q = q0 + cq;
      double hv[3] = {0.0, 0.0, 0.0};
      if (q < NGP) scaled_gradient(L.Ai + 9 * q, L.rs[q], d, hv);
      L.h[3 * lane + 0] = hv[0];
      L.h[3 * lane + 1] = hv[1];
      L.h[3 * lane + 2] = hv[2];
    }
    __syncthreads();
    load_dn(q0 + QC, d);  // in flight during this chunk's products
#pragma unroll 1
    for (int qq = 0; qq < QC && TL == 1; ++qq) {
      // couples: 6 contiguous doubles, 16-byte aligned (h is, and 2 nodes are 48 bytes)
      const f64x2* hs = reinterpret_cast<const f64x2*>(L.h + (qq * NN + 2 * cs) * 3);
      const f64x2* ht = reinterpret_cast<const f64x2*>(L.h + (qq * NN + 2 * ct) * 3);
      const f64x2 s0 = hs[0], s1 = hs[1], s2 = hs[2];
      const f64x2 t0 = ht[0], t1 = ht[1], t2 = ht[2];
      const double A[2][3] = {{s0.x, s0.y, s1.x}, {s1.y, s2.x, s2.y}};
      const double B[2][3] = {{t0.x, t0.y, t1.x}, {t1.y, t2.x, t2.y}};
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) M[t][3 * i + j] = __builtin_fma(A[t / 2][i], B[t % 2][j], M[t][3 * i + j]);
    }
#pragma unroll 1
    for (int qq = 0; qq < QC && TL == 0; ++qq) {
      const double* ha = L.h + (qq * NN + ga_c) * 3;
      const double a0 = ha[0], a1 = ha[1], a2 = ha[2];
#pragma unroll
      for (int t = 0; t < PPL; ++t) {
        const int b = gb0 + t < NN ? gb0 + t : NN - 1;
        const double* hb = L.h + (qq * NN + (b >= 0 ? b : 0)) * 3;
        const double av[3] = {a0, a1, a2}, bv[3] = {hb[0], hb[1], hb[2]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) M[t][3 * i + j] = __builtin_fma(av[i], bv[j], M[t][3 * i + j]);
      }
    }
    __syncthreads();
  }
  double lam, mu;
  lame(Ee, nue, lam, mu);
  constexpr int ND = 3 * NN;
  // M -> K in place: K_ab = lam M + mu M^T + mu tr(M) I
#pragma unroll
  for (int t = 0; t < PPL; ++t) {
    double* m = M[t];
    const double tr = mu * (m[0] + m[4] + m[8]);
    double K[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        K[3 * i + j] = __builtin_fma(lam, m[3 * i + j], __builtin_fma(mu, m[3 * j + i], i == j ? tr : 0.0));
#pragma unroll
    for (int i = 0; i < 9; ++i) m[i] = K[i];
  }
  if constexpr (UP) {
    store_upper_blocks<NN, PPL, NTS>(L.stage, pa, pb, M, kel, lane);
    return;
  }
  // Staged store: rows of kStageNodes nodes at a time. Block (a, b) goes to the rows of
  // a, its transpose to the rows of b; the staged rows are then one contiguous run of
  // the element's row-major matrix.
  for (int g0 = 0; g0 < NN; g0 += kStageNodes) {
#pragma unroll
    for (int t = 0; t < PPL; ++t) {
      if (pa[t] < 0) continue;
      const double* K = M[t];
      if (pa[t] >= g0 && pa[t] < g0 + kStageNodes) {
        double* ra = L.stage + (3 * (pa[t] - g0)) * ND + 3 * pb[t];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) ra[i * ND + j] = K[3 * i + j];
      }
      if (pb[t] != pa[t] && pb[t] >= g0 && pb[t] < g0 + kStageNodes) {
        double* rb = L.stage + (3 * (pb[t] - g0)) * ND + 3 * pa[t];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int i = 0; i < 3; ++i) rb[j * ND + i] = K[3 * i + j];
      }
    }
    __syncthreads();
    const int nrow = 3 * (NN - g0 < kStageNodes ? NN - g0 : kStageNodes);
    // 16-byte streaming stores: the arena is read once, by the gather, much later
    // (nrow * ND is even and every staged run starts 16-byte aligned)
    f64x2* dst = reinterpret_cast<f64x2*>(kel + int64_t(3 * g0) * ND);
    const f64x2* src = reinterpret_cast<const f64x2*>(L.stage);
    for (int idx = lane; idx < nrow * ND / 2; idx += kWave)
      __builtin_nontemporal_store(src[idx], dst + idx);
    __syncthreads();
  }
}

// With ue given, the element's nodal displacements of n_cases fields (u + c * n_dof) are
// copied to ue + c * ue_stride + 3 * (connectivity offset), where the row gather reads
// them by its incidence records without a connectivity load.
template <int NN, int NGP, int QC, int WAVES, bool UP = false, int TL = 0, bool NTS = true>
__global__ __launch_bounds__(kWave, WAVES) void k_elem_iso(
    const int32_t* __restrict__ conn, int64_t elem_offset, const double* __restrict__ dN,
    const double* __restrict__ wq, const double* __restrict__ coords,
    const double* __restrict__ E, const double* __restrict__ nu, double* __restrict__ kp,
    int* __restrict__ flag, int64_t n_elem, int64_t slab, const double* __restrict__ u,
    int64_t n_dof, int32_t n_cases, double* __restrict__ ue, int64_t ue_stride) {
  __shared__ ElemLds<NN, NGP, QC> L;
  const int64_t le = xcd_index(blockIdx.x, n_elem, slab);
  if (le < 0) return;
  if (ue && threadIdx.x < NN) {
    const int64_t k = le * NN + threadIdx.x;
    const int64_t nd = conn[k];
    for (int c = 0; c < n_cases; ++c)
#pragma unroll
      for (int j = 0; j < 3; ++j) ue[c * ue_stride + 3 * k + j] = u[c * n_dof + 3 * nd + j];
  }
  elem_iso_body<NN, NGP, QC, UP, TL, NTS>(L, int(threadIdx.x), conn + le * NN, E[elem_offset + le],
                             nu[elem_offset + le], dN, wq, coords, kp + le * (9 * NN * NN), flag);
}

// Elements of at most 8 nodes and Gauss points (C3D8, C3D6), upper blocks: EPW elements
// per wave. The Jacobians of all their points are formed at once (one lane per element
// and point; one element per wave leaves 56 of 64 lanes idle there for C3D8), then each
// element's products, K_ab and staged store run as in elem_iso_body (TL = 0, one chunk
// of points), with the same device code in the same order: the values are the same bits.
constexpr int iso_multi_epw(int ngp) { return ngp >= 8 ? kWave / ngp : 16; }

template <int NN, int NGP, int QC, int EPW>
struct IsoMultiLds {
  double x[EPW][NN * 3];
  double Ai[EPW][NGP * 9];
  double rs[EPW][NGP];
  alignas(16) double h[QC * NN * 3];
  alignas(16) double stage[kStageNodes * 3 * 3 * NN];
};

template <int NN, int NGP, int QC>
__global__ __launch_bounds__(kWave) void k_elem_iso_multi(
    const int32_t* __restrict__ conn, int64_t elem_offset, const double* __restrict__ dN,
    const double* __restrict__ wq, const double* __restrict__ coords,
    const double* __restrict__ E, const double* __restrict__ nu, double* __restrict__ kp,
    int* __restrict__ flag, int64_t n_elem, int64_t slab, const double* __restrict__ u,
    int64_t n_dof, int32_t n_cases, double* __restrict__ ue, int64_t ue_stride) {
  constexpr int EPW = iso_multi_epw(NGP);
  constexpr int NP = NN * (NN + 1) / 2;
  static_assert(NN <= 8 && NGP <= QC && NP <= kWave && QC * NN <= kWave && EPW * NGP <= kWave,
                "one pair and one (point, node) per lane, one chunk of points");
  __shared__ IsoMultiLds<NN, NGP, QC, EPW> L;
  const int lane = threadIdx.x;
  const int64_t grp = xcd_index(blockIdx.x, (n_elem + EPW - 1) / EPW, slab);
  if (grp < 0) return;
  const int64_t le0 = grp * EPW;
  const int ne = int(n_elem - le0 < EPW ? n_elem - le0 : EPW);
  for (int idx = lane; idx < ne * NN; idx += kWave) {
    const int e = idx / NN, k = idx - e * NN;
    const int64_t ck = (le0 + e) * NN + k;
    const int64_t nd = conn[ck];
#pragma unroll
    for (int c = 0; c < 3; ++c) L.x[e][3 * k + c] = coords[3 * nd + c];
    if (ue)
      for (int cs = 0; cs < n_cases; ++cs)
#pragma unroll
        for (int j = 0; j < 3; ++j) ue[cs * ue_stride + 3 * ck + j] = u[cs * n_dof + 3 * nd + j];
  }
  __syncthreads();
  if (lane < ne * NGP) {
    const int e = lane / NGP, q = lane - NGP * (lane / NGP);
    jacobian_inverse<NN, dn_stride<NGP>()>(L.x[e], dn_point<NN, NGP>(dN, q), wq[q],
                                           L.Ai[e] + 9 * q, L.rs[e][q], flag);
  }
  __syncthreads();
  // the lane's pair (a, b) and (point, node) of the gradient pass, as elem_iso_body
  int ga, gb0;
  decode_group<NN, 1>(lane, ga, gb0);
  int pa[1], pb[1];
  pa[0] = ga >= 0 && gb0 < NN ? ga : -1;
  pb[0] = ga >= 0 && gb0 < NN ? gb0 : -1;
  const int ga_c = ga >= 0 ? ga : 0;
  const int b_c = gb0 < NN ? (gb0 >= 0 ? gb0 : 0) : NN - 1;
  const int cq = lane / NN, cn = lane - NN * (lane / NN);
  double d[3] = {0.0, 0.0, 0.0};
  if (lane < QC * NN && cq < NGP) {
    const double* p = dN + (cq * NN + cn) * 3;
    d[0] = p[0];
    d[1] = p[1];
    d[2] = p[2];
  }
  for (int e = 0; e < ne; ++e) {
    if (lane < QC * NN) {
      double hv[3] = {0.0, 0.0, 0.0};
      if (cq < NGP) scaled_gradient(L.Ai[e] + 9 * cq, L.rs[e][cq], d, hv);
      L.h[3 * lane + 0] = hv[0];
      L.h[3 * lane + 1] = hv[1];
      L.h[3 * lane + 2] = hv[2];
    }
    __syncthreads();
    double M[1][9];
#pragma unroll
    for (int i = 0; i < 9; ++i) M[0][i] = 0.0;
#pragma unroll 1
    for (int qq = 0; qq < QC; ++qq) {
      const double* ha = L.h + (qq * NN + ga_c) * 3;
      const double* hb = L.h + (qq * NN + b_c) * 3;
      const double av[3] = {ha[0], ha[1], ha[2]}, bv[3] = {hb[0], hb[1], hb[2]};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[0][3 * i + j] = __builtin_fma(av[i], bv[j], M[0][3 * i + j]);
    }
    double lam, mu;
    lame(E[elem_offset + le0 + e], nu[elem_offset + le0 + e], lam, mu);
    {
      double* m = M[0];
      const double tr = mu * (m[0] + m[4] + m[8]);
      double K[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          K[3 * i + j] = __builtin_fma(lam, m[3 * i + j], __builtin_fma(mu, m[3 * j + i], i == j ? tr : 0.0));
#pragma unroll
      for (int i = 0; i < 9; ++i) m[i] = K[i];
    }
    // ends with a barrier: the next element's gradients may overwrite h
    store_upper_blocks<NN, 1>(L.stage, pa, pb, M, kp + (le0 + e) * (9 * NN * NN), lane);
  }
}

// ---- K1, gradient form (C3D20R) -----------------------------------------------------
// Instead of its upper node-pair blocks the element's slot holds what they are made of:
// slot[0..1] = lambda, mu and the scaled gradients h_{k,c}(q) = sqrt(det J w)
// (J^-T dN_k)_c at 2 + 3n q + 3k + c (point-major: at each point the 3n values of the
// element are consecutive, as the row gather's lanes read them). The gather forms
// K_ab = lambda M_ab + mu M_ab^T + mu tr(M_ab) I, M_ab = sum_q h_a h_b^T, with the
// operations of k_elem_iso, so the values are the same bits; the slot is 2 + 3 n NGP
// doubles (3.9 KB for C3D20R against 15.1 KB of upper blocks).
// 64 / NGP elements per wave: lane (e, q) forms J^-1 and sqrt(det J w) of its point into
// LDS, then the wave forms the gradients of one element at a time, lane t the point
// t / n and node t % n, so that the stores of a wave are consecutive.
template <int NN, int NGP>
__global__ __launch_bounds__(kWave) void k_elem_hform(
    const int32_t* __restrict__ conn, int64_t elem_offset, const double* __restrict__ dN,
    const double* __restrict__ wq, const double* __restrict__ coords,
    const double* __restrict__ E, const double* __restrict__ nu, double* __restrict__ kp,
    int* __restrict__ flag, int64_t n_elem, int64_t slab, const double* __restrict__ u,
    int64_t n_dof, int32_t n_cases, double* __restrict__ ue, int64_t ue_stride) {
  static_assert(kWave % NGP == 0, "one lane per point");
  constexpr int EPW = kWave / NGP;
  __shared__ double X[EPW][NN * 3];
  __shared__ double Ai[EPW][NGP][9];
  __shared__ double R[EPW][NGP];
  const int lane = threadIdx.x;
  const int64_t grp = xcd_index(blockIdx.x, (n_elem + EPW - 1) / EPW, slab);
  if (grp < 0) return;
  const int64_t le0 = grp * EPW;
  const int ne = int(n_elem - le0 < EPW ? n_elem - le0 : EPW);
  for (int idx = lane; idx < ne * NN; idx += kWave) {
    const int e = idx / NN, k = idx - e * NN;
    const int64_t ck = (le0 + e) * NN + k;
    const int64_t nd = conn[ck];
    for (int c = 0; c < 3; ++c) X[e][3 * k + c] = coords[3 * nd + c];
    if (ue)
      for (int cs = 0; cs < n_cases; ++cs)
        for (int j = 0; j < 3; ++j) ue[cs * ue_stride + 3 * ck + j] = u[cs * n_dof + 3 * nd + j];
  }
  __syncthreads();
  {
    const int e = lane / NGP, q = lane - NGP * (lane / NGP);
    if (e < ne) {
      jacobian_inverse<NN, dn_stride<NGP>()>(X[e], dn_point<NN, NGP>(dN, q), wq[q], Ai[e][q], R[e][q], flag);
      if (q == 0) {
        double lam, mu;
        lame(E[elem_offset + le0 + e], nu[elem_offset + le0 + e], lam, mu);
        double* slot = kp + (le0 + e) * (9 * NN * NN);
        slot[0] = lam;
        slot[1] = mu;
      }
    }
  }
  __syncthreads();
  for (int e = 0; e < ne; ++e) {
    double* slot = kp + (le0 + e) * (9 * NN * NN) + 2;
    for (int t = lane; t < NGP * NN; t += kWave) {
      const int q = t / NN;
      double h[3];
      scaled_gradient(Ai[e][q], R[e][q], dN + 3 * t, h);
#pragma unroll
      for (int c = 0; c < 3; ++c) slot[3 * t + c] = h[c];
    }
  }
}

// ---- K1 on the matrix cores (asm_variant 6): M = H^T H by FP64 MFMA -------------------
// Same M-identity as k_elem_iso, as one small GEMM per element: with
// H[q][3a+i] = sqrt(det J w)_q g_{q,a,i} (NGP x 3n), M = H^T H holds every
// M_ab = sum_q h_a h_b^T at once, and K_ab = lam M_ab + mu M_ab^T + mu tr(M_ab) I.
// One element per wave. Each lane computes exactly the H entries it feeds to
// v_mfma_f64_16x16x4f64 (row k = lane/16 of a 4-deep K step, column lane%16 of a
// 16-wide tile), so H never touches LDS; the upper tiles of M (TI <= TJ) accumulate
// in registers over ceil(NGP/4) K steps. M goes to LDS, and each lane then writes
// consecutive entries of the row-major 3n x 3n Ke,
//   K[3a+i][3b+j] = lam M(3a+i, 3b+j) + mu M(3a+j, 3b+i) + [i == j] mu tr M_ab,
// which, M being symmetric, holds for every (a, b): Ke comes out bitwise symmetric,
// stored with fully coalesced writes.
typedef double mf64x4 __attribute__((ext_vector_type(4)));

template <int NN, int NGP>
struct MfmaLds {
  static constexpr int ND = 3 * NN;
  static constexpr int LD = ND + 1;  // row stride of M (odd: spreads LDS banks)
  double x[NN * 3];
  double Ai[NGP * 9];
  double rs[NGP];
  double M[ND * LD];
};

template <int NN, int NGP>
__global__ __launch_bounds__(kWave) void k_elem_iso_mfma(
    const int32_t* __restrict__ conn, int64_t elem_offset, const double* __restrict__ dN,
    const double* __restrict__ wq, const double* __restrict__ coords,
    const double* __restrict__ E, const double* __restrict__ nu, double* __restrict__ kp,
    int* __restrict__ flag) {
  constexpr int ND = 3 * NN;
  constexpr int NT = (ND + 15) / 16;   // 16-wide column tiles of H
  constexpr int KS = (NGP + 3) / 4;    // 4-deep steps over the Gauss points
  constexpr int LD = MfmaLds<NN, NGP>::LD;
  __shared__ MfmaLds<NN, NGP> L;
  const int lane = threadIdx.x;
  const int64_t le = blockIdx.x;
  const int32_t* ce = conn + le * NN;
  if (lane < NN) {
    const double* xp = coords + 3 * int64_t(ce[lane]);
    L.x[3 * lane + 0] = xp[0];
    L.x[3 * lane + 1] = xp[1];
    L.x[3 * lane + 2] = xp[2];
  }
  __syncthreads();
  for (int q = lane; q < NGP; q += kWave)
    jacobian_inverse<NN, dn_stride<NGP>()>(L.x, dn_point<NN, NGP>(dN, q), wq[q], L.Ai + 9 * q, L.rs[q], flag);
  __syncthreads();
  // this lane's H entries: h[s][t] = H[4s + lane/16][16t + lane%16]
  const int qs = lane / 16, cs = lane % 16;
  double h[KS][NT];
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int q = 4 * s + qs, col = 16 * t + cs;
      double v = 0.0;
      if (q < NGP && col < ND) {
        const int a = col / 3, i = col - 3 * (col / 3);
        const double* A = L.Ai + 9 * q;
        const double* d = dN + (q * NN + a) * 3;
        v = L.rs[q] * (A[i] * d[0] + A[3 + i] * d[1] + A[6 + i] * d[2]);
      }
      h[s][t] = v;
    }
  // upper tiles of M = H^T H
  mf64x4 acc[NT * (NT + 1) / 2];
#pragma unroll
  for (int k = 0; k < NT * (NT + 1) / 2; ++k) acc[k] = mf64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    int k = 0;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
      for (int tj = ti; tj < NT; ++tj, ++k)
        acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(h[s][ti], h[s][tj], acc[k], 0, 0, 0);
  }
  // D layout (gfx950, checked by an in-bounds probe): element m of a lane's result is
  // row lane/16 + 4m, column lane%16 of its tile
  {
    int k = 0;
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
      for (int tj = ti; tj < NT; ++tj, ++k) {
        const int c = 16 * tj + cs;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int r = 16 * ti + qs + 4 * m;
          if (r < ND && c < ND) {
            L.M[r * LD + c] = acc[k][m];
            L.M[c * LD + r] = acc[k][m];  // M is symmetric: store both triangles
          }
        }
      }
  }
  __syncthreads();
  double lam, mu;
  lame(E[elem_offset + le], nu[elem_offset + le], lam, mu);
  double* kel = kp + le * (ND * ND);
  // lane = column 3b + j of Ke; rows walk node by node
  for (int c = lane; c < ND; c += kWave) {
    const int b3 = c - c % 3, j = c % 3;
    for (int a3 = 0; a3 < ND; a3 += 3) {
      const double* Ma = L.M + a3 * LD;
      const double tr = mu * (Ma[b3] + Ma[LD + b3 + 1] + Ma[2 * LD + b3 + 2]);
      const double* Mj = Ma + j * LD + b3;  // M(a3 + j, b3 + i) = Mj[i]
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double v = lam * Ma[i * LD + c] + mu * Mj[i];
        if (i == j) v += tr;
        kel[(a3 + i) * ND + c] = v;
      }
    }
  }
}

// ---- K4 (+K5), two-phase variant, phase B: CSR rows gathered by their owner node ----
// One 64-lane wave per node a. Its (up to three) CSR rows share one column layout of
// length L; they are accumulated in LDS from rows 3la..3la+2 of the element matrices
// of a's incident elements
// in ascending element order (the order in which the reference sums its COO
// duplicates, assembly.cpp:94) and written once, coalesced. Rows longer than the LDS
// capacity are accumulated in place in the CSR array (one owner, still race-free).
// With u given, the same pass forms f_int[a] = sum_e (Ke ue)_a over the full DOF
// set, as fem::internal_force does (stress.cpp:125-145), with a fixed-order wave
// reduction. With fe given instead (element force vectors of the J2 path), f_int[a]
// is the sum of the incident elements' entries for a, in ascending element order.
constexpr int kRowCap = 256;
constexpr int kBatch = 4;  // incident elements whose loads are issued together

// v of lane l (l wave-uniform) to every lane.
__device__ __forceinline__ double read_lane(double v, int l) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(b), l);
  const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(b >> 32), l);
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// buf[0] += v in LDS without a read-back (DS_ADD_F64). The rows belong to one wave, whose
// LDS operations execute in issue order, so every entry is still summed in element order.
__device__ __forceinline__ void lds_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// The first kPreRec incidence records of a node, one dword per lane: lane 8k + w holds
// dword w of record k (a single coalesced load of 256 bytes into one VGPR, read across
// the wave by the one-element-per-step path).
constexpr int kPreRec = 8;
static_assert(sizeof(IncRecord) == 32, "8 dwords per incidence record");

__device__ __forceinline__ uint32_t load_records(const IncRecord* __restrict__ rec,
                                                 const GatherDesc& d, int lane) {
  const int n = d.inc_n < kPreRec ? d.inc_n : kPreRec;
  return lane < 8 * n ? reinterpret_cast<const uint32_t*>(rec + d.inc0)[lane] : 0u;
}

// Record `it` (< i1) of the node whose records start at i0, the same in every lane: from
// the lanes of rv for the first kPreRec (lane reads), else from memory.
__device__ __forceinline__ IncRecord record_uniform(const IncRecord* __restrict__ rec, int64_t it,
                                                    int64_t i0, uint32_t rv) {
  const int64_t k = it - i0;
  if (k >= kPreRec) return rec[it];
  uint32_t w[8];
#pragma unroll
  for (int f = 0; f < 8; ++f) w[f] = __builtin_amdgcn_readlane(rv, int(8 * k) + f);
  return __builtin_bit_cast(IncRecord, w);
}

// One node's three rows. HALVES = 2: two elements per step, one per half-wave, for
// nodes whose incident elements all have <= 10 nodes (3n <= 30 columns fit in 32 lanes).
// SCALAR: the arena holds k_ab per node pair (n^2 per element, at a ninth of the full
// matrix's offsets) for a k_ab I3 block matrix, expanded here.
// UP: the arena holds the upper node-pair blocks (elem_iso_body with UP = true) of the
// elements of at least up_min nodes at their slots, the blocks of partners b < a read
// transposed, and the full matrix of the others.
// HF > 0 (with UP, HALVES = 1): the 20-node elements hold their gradient form
// (k_elem_hform, HF Gauss points); the lane forms its column of M_ab from it, takes the
// rest of the 3x3 block from the two other lanes of its node, then K_ab as k_elem_iso
// does. Other elements are read as with UP alone.
// (Same-direction rows of the scalar gathers go through row_gather_sd.)
template <int NC, bool LIVE, int HALVES, bool SCALAR, bool UP = false, int HF = 0>
__device__ __forceinline__ void row_gather_node(
    const GatherDesc& d, int cap, double* __restrict__ buf, uint8_t* __restrict__ lbuf,
    const IncRecord* __restrict__ rec, const int32_t* __restrict__ conn,
    const int32_t* __restrict__ slot_all, const double* __restrict__ kp,
    const double* __restrict__ u, int64_t n_dof, const double* __restrict__ fe,
    double* __restrict__ vals, double* __restrict__ f_int, uint8_t* __restrict__ live,
    const double* __restrict__ ue, int64_t ue_stride, int up_min, uint32_t rv, int lane) {
  constexpr int halves = HALVES;
  const int h = HALVES == 2 ? lane / 32 : 0;
  const int sub = HALVES == 2 ? lane % 32 : lane;
  const int bl = sub / 3, jl = sub - 3 * (sub / 3);
  const int32_t a = d.node;
  // the node's free rows are consecutive equations sharing one column layout
  int64_t rs[3];
  const int64_t Lrow = d.len;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    rs[i] = (d.flags >> i) & 1 ? d.row0 + int64_t(__popc(d.flags & ((1 << i) - 1))) * Lrow : -1;
  const bool in_lds = Lrow <= cap;
  if (in_lds) {
    // the whole row area (values, then the flag bytes) with 16-byte writes: a few LDS
    // instructions instead of one per row, chunk and array
    // LIVE: one flag byte per entry
    const int nbytes = 3 * cap * 8 + (LIVE ? 3 * cap : 0);
    uint4* z = reinterpret_cast<uint4*>(buf);
    for (int c = lane; c < nbytes / 16; c += kWave) z[c] = make_uint4(0u, 0u, 0u, 0u);
    for (int c = nbytes / 16 * 16 + lane; c < nbytes; c += kWave) reinterpret_cast<uint8_t*>(buf)[c] = 0;
  } else {
    for (int i = 0; i < 3; ++i) {
      if (rs[i] < 0) continue;
      for (int64_t c = lane; c < Lrow; c += kWave) {
        vals[rs[i] + c] = 0.0;
        if constexpr (LIVE) live[rs[i] + c] = 0;
      }
    }
    __syncthreads();  // the zeros in vals are seen by the lanes that add
  }
  constexpr int NU = NC > 0 ? NC : 1;
  double f[NU][3];
#pragma unroll
  for (int c = 0; c < NU; ++c) f[c][0] = f[c][1] = f[c][2] = 0.0;
  // element force vectors (fe mode): lanes sub < 3 of each half sum their elements'
  // component sub in ascending element order; the halves are added at the end
  double fsum = 0.0;
  const int64_t i0 = d.inc0, i1 = d.inc0 + d.inc_n;
  // gradient form only where one element is taken per step (every incidence is then
  // wave-uniform)
  constexpr int HFE = HALVES == 1 ? HF : 0;
  // one incidence per step in the gradient form: its partner gradients take 16 VGPRs per
  // incidence, and a batch of 2 (4) cut the waves per SIMD so far that the gather took
  // 8% (44%) longer at 64^3 C3D20R
  constexpr int KB = HFE > 0 ? 1 : kBatch;
  constexpr int HB = HFE > 0 ? HFE : 1;
  for (int64_t ib = i0; ib < i1; ib += KB * halves) {
    // issue every load of the batch first ...
    double v[KB][3], uv[KB][NU], fv[KB];
    // gradient form: partner gradients h_{b,jl}(q), the lane's share of the row node's
    // gradients / lambda / mu, and whether incidence k is in that form (wave-uniform)
    double hb[KB][HB], cv[KB];
    bool hk[KB];
    int32_t colp[KB];
    int fm[KB];
    bool ok[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      ok[k] = false;
      hk[k] = false;
      v[k][0] = v[k][1] = v[k][2] = 0.0;
      cv[k] = 0.0;
#pragma unroll
      for (int q = 0; q < HB; ++q) hb[k][q] = 0.0;
#pragma unroll
      for (int c = 0; c < NU; ++c) uv[k][c] = 0.0;
      colp[k] = 0;
      fm[k] = 0;
      fv[k] = 0.0;
      const int64_t it = ib + int64_t(k) * halves + h;
      if (it < i1) {
        const IncRecord r = HALVES == 1 ? record_uniform(rec, it, i0, rv) : rec[it];
        if constexpr (NC == 0)
          if (sub < 3) fv[k] = fe[3 * (r.conn + r.la) + sub];
        if (bl < r.n) {
          ok[k] = true;
          const int32_t sl = slot_all[r.slot + bl];
          colp[k] = sl & kSlotMask;
          fm[k] = sl >> kSlotFreeShift;
          if (HFE > 0 && r.n == 20) {
            // slot: lambda, mu, then h_{k,c}(q) at 2 + 3n q + 3k + c. Lane 3q + i < 3 HF
            // loads the row node's h_{la,i}(q), lanes 3 HF and 3 HF + 1 lambda and mu
            // (the compute step below reads them across the wave); every lane its
            // partner's h_{bl,jl}(q), the wave's lanes consecutive at each q.
            hk[k] = true;
            const int la = r.la, n = r.n;
            const int64_t base = r.rows - int64_t(9) * la * n;
            const int qa = lane / 3, ia = lane - 3 * (lane / 3);
            const int64_t oa = lane < 3 * HFE ? 2 + int64_t(3) * n * qa + 3 * la + ia
                                              : (lane == 3 * HFE + 1 ? 1 : 0);
            cv[k] = kp[base + oa];
            const double* hp = kp + base + 2 + 3 * bl + jl;
#pragma unroll
            for (int q = 0; q < HB; ++q) hb[k][q] = hp[int64_t(3) * n * q];
          } else if constexpr (SCALAR) {
            const double s = __builtin_nontemporal_load(kp + r.rows / 9 + bl);
            v[k][0] = jl == 0 ? s : 0.0;
            v[k][1] = jl == 1 ? s : 0.0;
            v[k][2] = jl == 2 ? s : 0.0;
          } else if constexpr (UP) {
            // Branch-free, so that the batch's loads issue together: an element of at
            // least up_min nodes has its upper blocks in the arena (K_ab[i][j] at 3i + j
            // of block (a, b) for b >= a, at 3j + i of block (b, a) for b < a), the
            // others their full rows. Plain (cached) loads: the partner node's wave reads
            // the same block.
            const int la = r.la, n = r.n;
            const bool upr = n >= up_min, dir = bl >= la;
            const int64_t base = r.rows - int64_t(9) * la * n;
            const int64_t off_up = dir ? 9 * up_block(la, bl, n) + jl : 9 * up_block(bl, la, n) + 3 * jl;
            const double* ke = kp + (upr ? base + off_up : r.rows + 3 * bl + jl);
            const int64_t st = upr ? (dir ? 3 : 1) : 3 * n;
            v[k][0] = ke[0];
            v[k][1] = ke[st];
            v[k][2] = ke[2 * st];
          } else {
            const double* ke = kp + r.rows + 3 * bl + jl;
            const int ndof = 3 * r.n;
            v[k][0] = __builtin_nontemporal_load(ke);
            v[k][1] = __builtin_nontemporal_load(ke + ndof);
            v[k][2] = __builtin_nontemporal_load(ke + 2 * ndof);
          }
          if constexpr (NC > 0)
            if (u) {
              if constexpr (UP) {
                // the element's displacements, copied by the element kernel in
                // connectivity order: no connectivity load ahead of this one
#pragma unroll
                for (int c = 0; c < NC; ++c) uv[k][c] = ue[c * ue_stride + 3 * (r.conn + bl) + jl];
              } else {
                const int32_t nd = conn[r.conn + bl];
#pragma unroll
                for (int c = 0; c < NC; ++c) uv[k][c] = u[c * n_dof + 3 * int64_t(nd) + jl];
              }
            }
        }
      }
    }
    if constexpr (HFE > 0) {
      // Gradient form, the whole wave (wave-uniform condition): the lane's column
      // M_ab[.][jl] = sum_q h_a(q) h_{b,jl}(q) in ascending q with the row node's
      // gradients read across the wave, then from the node's two other lanes the row
      // M_ab[jl][.] and the diagonal (5 shuffles), then K_ab[i][jl] as k_elem_iso forms it.
      const int g0 = lane - jl;
      const int j1 = jl == 2 ? 0 : jl + 1, j2 = jl == 0 ? 2 : jl - 1;
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        if (!__builtin_amdgcn_readfirstlane(int(hk[k]))) continue;
        double mc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < HFE; ++q)
#pragma unroll
          for (int i = 0; i < 3; ++i) mc[i] = __builtin_fma(read_lane(cv[k], 3 * q + i), hb[k][q], mc[i]);
        const double lam = read_lane(cv[k], 3 * HFE), mu = read_lane(cv[k], 3 * HFE + 1);
        const double dg = jl == 0 ? mc[0] : (jl == 1 ? mc[1] : mc[2]);  // M_ab[jl][jl]
        const double p1 = jl == 0 ? mc[2] : (jl == 1 ? mc[0] : mc[1]);  // M_ab[jl - 1][jl]
        const double p2 = jl == 0 ? mc[1] : (jl == 1 ? mc[2] : mc[0]);  // M_ab[jl + 1][jl]
        const double t0 = __shfl(dg, g0), t1 = __shfl(dg, g0 + 1), t2 = __shfl(dg, g0 + 2);
        const double r1 = __shfl(p1, g0 + j1);  // M_ab[jl][jl + 1]
        const double r2 = __shfl(p2, g0 + j2);  // M_ab[jl][jl - 1]
        const double tr = mu * (t0 + t1 + t2);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double mt = i == jl ? dg : (i == j1 ? r1 : r2);  // M_ab[jl][i]
          v[k][i] = __builtin_fma(lam, mc[i], __builtin_fma(mu, mt, i == jl ? tr : 0.0));
        }
      }
    }
    // ... then accumulate in ascending element order: element (k, half 0) before
    // (k, half 1) before (k + 1, half 0). LDS updates of one wave are ordered by program
    // order; the global read-modify-writes of long rows are separated by barriers.
#pragma unroll
    for (int k = 0; k < KB; ++k)
    for (int hh = 0; hh < halves; ++hh) {
      if (!in_lds && (k > 0 || hh > 0)) __syncthreads();
      if (!ok[k] || h != hh) continue;
      if constexpr (NC == 0) fsum += fv[k];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        f[c][0] += v[k][0] * uv[k][c];
        f[c][1] += v[k][1] * uv[k][c];
        f[c][2] += v[k][2] * uv[k][c];
      }
      if ((fm[k] >> jl) & 1) {
        const int64_t pos = int64_t(colp[k]) + __popc(fm[k] & ((1 << jl) - 1));
        if constexpr (SCALAR) {
          // k_ab I3: of the three rows only row jl receives a value (the others' zeros
          // would leave them unchanged)
          const double sv = jl == 0 ? v[k][0] : (jl == 1 ? v[k][1] : v[k][2]);
          const int64_t rsj = jl == 0 ? rs[0] : (jl == 1 ? rs[1] : rs[2]);
          if (in_lds) {
            if (rsj >= 0) lds_add(buf + jl * cap + pos, sv);
            if constexpr (LIVE)
              if (sv != 0.0) lbuf[jl * cap + pos] = 1;
          } else {
            if (rsj >= 0) vals[rsj + pos] += sv;
            if constexpr (LIVE)
              if (rsj >= 0 && sv != 0.0) live[rsj + pos] = 1;
          }
        } else if (in_lds) {
          if (rs[0] >= 0) lds_add(buf + pos, v[k][0]);
          if (rs[1] >= 0) lds_add(buf + cap + pos, v[k][1]);
          if (rs[2] >= 0) lds_add(buf + 2 * cap + pos, v[k][2]);
          if constexpr (LIVE) {
            if (v[k][0] != 0.0) lbuf[pos] = 1;
            if (v[k][1] != 0.0) lbuf[cap + pos] = 1;
            if (v[k][2] != 0.0) lbuf[2 * cap + pos] = 1;
          }
        } else {
          if (rs[0] >= 0) vals[rs[0] + pos] += v[k][0];
          if (rs[1] >= 0) vals[rs[1] + pos] += v[k][1];
          if (rs[2] >= 0) vals[rs[2] + pos] += v[k][2];
          if constexpr (LIVE) {
            if (rs[0] >= 0 && v[k][0] != 0.0) live[rs[0] + pos] = 1;
            if (rs[1] >= 0 && v[k][1] != 0.0) live[rs[1] + pos] = 1;
            if (rs[2] >= 0 && v[k][2] != 0.0) live[rs[2] + pos] = 1;
          }
        }
      }
    }
    if (!in_lds) __syncthreads();  // global RMWs of the next batch must see these
  }
  __syncthreads();
  if (in_lds)
    for (int i = 0; i < 3; ++i) {
      if (rs[i] < 0) continue;
      for (int64_t c = lane; c < Lrow; c += kWave)
        __builtin_nontemporal_store(buf[i * cap + c], vals + rs[i] + c);
      if constexpr (LIVE)
        for (int64_t c = lane; c < Lrow; c += kWave) live[rs[i] + c] = lbuf[i * cap + c];
    }
  if constexpr (NC == 0) {
    const double other = __shfl(fsum, lane + 32);  // half 1's sum (0.0 for one half)
    if (lane < 3) f_int[3 * int64_t(a) + lane] = halves == 2 ? fsum + other : fsum;
  } else if (u) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double s0 = wave_sum(f[c][0]);
      const double s1 = wave_sum(f[c][1]);
      const double s2 = wave_sum(f[c][2]);
      if (lane == 0) {
        double* fp = f_int + c * n_dof + 3 * int64_t(a);
        fp[0] = s0;
        fp[1] = s1;
        fp[2] = s2;
      }
    }
  }
}

// One node's three rows from incident elements of at most 8 nodes each, all with
// upper-block arena slots, and f_int from the element forces fe (the J2 path): lane 8s + b
// takes partner b of incidence s of a group of 8, loading the whole 3x3 block (four
// 16-byte loads and one 8-byte load, whatever the block's alignment) with one slot word,
// so that every load of the group issues in one step. The LDS adds go incidence by
// incidence (ascending element order for every entry, as row_gather_node sums them) and
// f_int sums the element forces in row_gather_node's two-half order: the values are the
// same bits. The first group's records come from rv (kPreRec = 8 of them).
template <bool LIVE>
__device__ __forceinline__ void row_gather_pb(
    const GatherDesc& d, int cap, double* __restrict__ buf, uint8_t* __restrict__ lbuf,
    const IncRecord* __restrict__ rec, const int32_t* __restrict__ slot_all,
    const double* __restrict__ kp, const double* __restrict__ fe, double* __restrict__ vals,
    double* __restrict__ f_int, uint8_t* __restrict__ live, uint32_t rv, int lane) {
  const int s = lane >> 3, b = lane & 7;
  const int n_inc = d.inc_n;
  int64_t rs[3];
  const int64_t Lrow = d.len;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    rs[i] = (d.flags >> i) & 1 ? d.row0 + int64_t(__popc(d.flags & ((1 << i) - 1))) * Lrow : -1;
  {
    // the row area (values, then the flag bytes) with 16-byte writes, as row_gather_node
    const int nbytes = 3 * cap * (8 + (LIVE ? 1 : 0));
    uint4* z = reinterpret_cast<uint4*>(buf);
    for (int c = lane; c < nbytes / 16; c += kWave) z[c] = make_uint4(0u, 0u, 0u, 0u);
    for (int c = nbytes / 16 * 16 + lane; c < nbytes; c += kWave) reinterpret_cast<uint8_t*>(buf)[c] = 0;
  }
  double h0 = 0.0, h1 = 0.0;  // f_int: the even incidences in order, the odd ones in order
  for (int g0 = 0; g0 < n_inc; g0 += 8) {
    const int sg = g0 + s;
    // record sg (zeros past the node's incidences); the permutes run in every lane
    IncRecord r;
    if (g0 == 0) {
      uint32_t w[8];
#pragma unroll
      for (int f = 0; f < 8; ++f) w[f] = uint32_t(__shfl(int(rv), 8 * s + f));
      r = __builtin_bit_cast(IncRecord, w);
    } else {
      r = sg < n_inc ? rec[d.inc0 + sg] : IncRecord{};
    }
    const bool ok = sg < n_inc && b < r.n;
    double blk[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) blk[k] = 0.0;
    int32_t sl = 0;
    double fv = 0.0;
    const bool dir = b >= r.la;
    if (ok) {
      sl = slot_all[r.slot + b];
      const int64_t base = r.rows - int64_t(9) * r.la * r.n;
      const int64_t o = base + 9 * int64_t(dir ? up_block(r.la, b, r.n) : up_block(b, r.la, r.n));
      const int lo = int(o & 1);
      const f64x2* q = reinterpret_cast<const f64x2*>(kp + o + lo);
      const f64x2 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
      const double one = kp[lo ? o : o + 8];
      const double t[8] = {q0.x, q0.y, q1.x, q1.y, q2.x, q2.y, q3.x, q3.y};
      blk[0] = lo ? one : t[0];
#pragma unroll
      for (int k = 1; k < 8; ++k) blk[k] = lo ? t[k - 1] : t[k];
      blk[8] = lo ? t[7] : one;
    }
    if (sg < n_inc && b < 3) fv = fe[3 * (r.conn + r.la) + b];
    const int colp = sl & kSlotMask, fm = sl >> kSlotFreeShift;
    // K_ab[i][j]: block (a, b) row-major for b >= a, block (b, a) transposed for b < a
    const int ng = n_inc - g0 < 8 ? n_inc - g0 : 8;
    for (int e = 0; e < ng; ++e) {
      if (s == e && ok) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          if (!((fm >> j) & 1)) continue;
          const int64_t pos = int64_t(colp) + __popc(fm & ((1 << j) - 1));
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const double v = dir ? blk[3 * i + j] : blk[3 * j + i];
            if (rs[i] >= 0) lds_add(buf + i * cap + pos, v);
            if constexpr (LIVE)
              if (v != 0.0) lbuf[i * cap + pos] = 1;
          }
        }
      }
    }
    for (int e = 0; e < ng; ++e) {
      const double x = __shfl(fv, 8 * e + (lane < 3 ? lane : 0));
      if (e & 1) h1 += x;  // g0 is a multiple of 8: e's parity is the incidence's
      else h0 += x;
    }
  }
  __syncthreads();
  for (int i = 0; i < 3; ++i) {
    if (rs[i] < 0) continue;
    for (int64_t c = lane; c < Lrow; c += kWave)
      __builtin_nontemporal_store(buf[i * cap + c], vals + rs[i] + c);
    if constexpr (LIVE)
      for (int64_t c = lane; c < Lrow; c += kWave) live[rs[i] + c] = lbuf[i * cap + c];
  }
  if (lane < 3) f_int[3 * int64_t(d.node) + lane] = h0 + h1;
}

// Same-direction rows of a k_ab I3 matrix (mass, geometric stiffness; every row in LDS).
// Row (a, i) takes the columns (b, i), and the three rows of a node carry the same sums
// k_ab, so each partner's sum is accumulated once, at the position of its first free DOF
// in the structural row, by one lane per (incidence, partner): L lanes per incidence, G
// incidences per step (10 and 6 where every incident element has at most 10 nodes, else
// 20 and 3), the adds of a step issued incidence by incidence, so every sum keeps
// ascending element order as in row_gather_node. Each position's flag byte holds, per
// direction i, touched (bit 3 + i: the position is a column of direction i that received
// a value) and live (bit i: a value other than 0.0), and in bits 6-7 the position's rank
// among its node's free DOFs (the offset back to the value). Row i then goes out chunk by
// chunk: its touched positions, in order, ballot-ranked onto sd_row_ptr's row. One value
// row per node where three rows each took one of every partner's three equal sums (1.80
// -> 1.13 ms for the 64^3 C3D20 mass) gives the same bits.
template <bool LIVE>
__device__ __forceinline__ void row_gather_sd(
    const GatherDesc& d, int cap, double* __restrict__ buf, const IncRecord* __restrict__ rec,
    const int32_t* __restrict__ slot_all, const double* __restrict__ kp,
    double* __restrict__ vals, uint8_t* __restrict__ live,
    const int64_t* __restrict__ sd_row_ptr, const int32_t* __restrict__ dof_eq, uint32_t rv,
    int lane) {
  static_assert(LIVE, "same-direction rows carry live flags");
  int64_t o[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if ((d.flags >> i) & 1) o[i] = sd_row_ptr[dof_eq[3 * int64_t(d.node) + i]];
  uint8_t* lb = reinterpret_cast<uint8_t*>(buf + cap);
  {
    // the value row and the flag bytes with 16-byte writes
    const int nbytes = cap * 8 + (cap + 3) / 4 * 4;
    uint4* z = reinterpret_cast<uint4*>(buf);
    for (int c = lane; c < nbytes / 16; c += kWave) z[c] = make_uint4(0u, 0u, 0u, 0u);
    for (int c = nbytes / 16 * 16 + lane; c < nbytes; c += kWave) reinterpret_cast<uint8_t*>(buf)[c] = 0;
  }
  const bool small = (d.flags & 8) != 0;
  const int L = small ? 10 : 20, G = small ? 6 : 3;
  const int s = lane / L, b = lane - L * (lane / L);
  const int n_inc = d.inc_n;
  for (int g0 = 0; g0 < n_inc; g0 += G) {
    const int sg = g0 + s;
    const bool inc_ok = s < G && sg < n_inc;
    // record sg: from rv while the group lies in the first kPreRec (the permutes run in
    // every lane), else from memory
    IncRecord r{};
    if (g0 + G <= kPreRec) {
      const int src = inc_ok ? sg : 0;
      uint32_t w[8];
#pragma unroll
      for (int f = 0; f < 8; ++f) w[f] = uint32_t(__shfl(int(rv), 8 * src + f));
      r = __builtin_bit_cast(IncRecord, w);
    } else if (inc_ok) {
      r = rec[d.inc0 + sg];
    }
    const bool ok = inc_ok && b < r.n;
    int32_t sl = 0;
    double k = 0.0;
    if (ok) {
      sl = slot_all[r.slot + b];
      k = __builtin_nontemporal_load(kp + r.rows / 9 + b);
    }
    const int colp = sl & kSlotMask, fm = sl >> kSlotFreeShift;
    const int ng = n_inc - g0 < G ? n_inc - g0 : G;
    for (int e = 0; e < ng; ++e) {
      if (s != e || !ok || fm == 0) continue;
      lds_add(buf + colp, k);
      const uint32_t lv = k != 0.0 ? 1u : 0u;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if (!((fm >> j) & 1)) continue;
        const int rank = __popc(fm & ((1 << j) - 1));
        const int pos = colp + rank;
        const uint32_t mk = (((8u | lv) << j) | (uint32_t(rank) << 6)) << (8 * (pos & 3));
        __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(lb) + (pos >> 2), mk, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
  __syncthreads();
  // chunk by chunk, one flag byte per position for the three rows
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  const int64_t Lrow = d.len;
  for (int64_t c0 = 0; c0 < Lrow; c0 += kWave) {
    const int64_t c = c0 + lane;
    const uint32_t fl = c < Lrow ? lb[c] : 0u;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (!((d.flags >> i) & 1)) continue;
      const bool on = (fl >> (3 + i)) & 1u;
      const uint64_t m = __ballot(on);
      if (on) {
        const int64_t at = o[i] + __popcll(m & below);
        __builtin_nontemporal_store(buf[c - (fl >> 6)], vals + at);
        live[at] = uint8_t((fl >> i) & 1u);
      }
      o[i] += __popcll(m);
    }
  }
}

// NC displacement fields (load cases) at u + c*n_dof give NC internal forces at
// f_int + c*n_dof from the same pass over the element rows (u may be null: K only).
// NC = 0: f_int is the sum of the element force vectors fe instead. LIVE: also flag the
// entries that received a contribution other than 0.0. SCALAR: the k_ab I3 arena of
// the mass and geometric stiffness. One wave per node.
// UP: upper-block arena, nodes taken in the order given by `order` (a locality order,
// so that the two nodes reading one block do so close together in time).
template <int NC, bool LIVE, bool SCALAR = false, bool UP = false, int HF = 0, bool SD = false>
__global__ __launch_bounds__(kWave) void k_row_gather(
    const GatherDesc* __restrict__ desc, const IncRecord* __restrict__ rec,
    const int32_t* __restrict__ conn, const int32_t* __restrict__ slot_all,
    const double* __restrict__ kp, const double* __restrict__ u, int64_t n_dof,
    const double* __restrict__ fe, double* __restrict__ vals, double* __restrict__ f_int,
    uint8_t* __restrict__ live, int32_t n_nodes, int32_t slab, int32_t cap,
    const double* __restrict__ ue, int64_t ue_stride, int32_t up_min,
    const int64_t* __restrict__ sd_row_ptr, const int32_t* __restrict__ dof_eq) {
  // three rows of cap columns (cap: the pattern's longest row, at most kRowCap; longer
  // rows accumulate in vals) and, for LIVE, their flags (SD: one value row and a flag byte
  // per position instead, see row_gather_sd)
  extern __shared__ __attribute__((aligned(16))) double row_lds[];
  double* buf = row_lds;
  uint8_t* lbuf = reinterpret_cast<uint8_t*>(row_lds + 3 * cap);
  const int lane = threadIdx.x;
  // one contiguous node range per XCD (slab > 0): its L2 sees the neighbouring nodes'
  // data again
  const int64_t ai = xcd_index(blockIdx.x, n_nodes, slab);
  if (ai < 0) return;
  // the position's descriptor (wave-uniform: scalar loads; flags bit 3: every incident
  // element has at most 10 nodes) and its first incidence records, one dword per lane
  // (one coalesced load instead of a scalar load per record and batch)
  const GatherDesc d = desc[ai];
  if constexpr (SD) {
    // same-direction rows (scalar gathers with flags, every row in LDS)
    row_gather_sd<LIVE>(d, cap, buf, rec, slot_all, kp, vals, live, sd_row_ptr, dof_eq,
                        load_records(rec, d, lane), lane);
  } else {
    // Nodes whose elements all have at most 8 nodes (flags bit 3: at most 10, which among
    // the supported families means C3D8 and C3D6), every element with upper blocks, rows
    // in LDS, element forces (the J2 path): one partner block per lane (row_gather_pb).
    const bool pb = NC == 0 && UP && !SCALAR && HF == 0 && (d.flags & 8) && up_min <= 6 &&
                    d.len <= cap && fe != nullptr;
    const uint32_t rv = (d.flags & 8) && !pb ? 0u : load_records(rec, d, lane);
    if (pb)
      row_gather_pb<LIVE>(d, cap, buf, lbuf, rec, slot_all, kp, fe, vals, f_int, live, rv, lane);
    else if (d.flags & 8)
      row_gather_node<NC, LIVE, 2, SCALAR, UP, HF>(d, cap, buf, lbuf, rec, conn, slot_all, kp,
                                                   u, n_dof, fe, vals, f_int, live, ue, ue_stride,
                                                   up_min, rv, lane);
    else
      row_gather_node<NC, LIVE, 1, SCALAR, UP, HF>(d, cap, buf, lbuf, rec, conn, slot_all, kp,
                                                   u, n_dof, fe, vals, f_int, live, ue, ue_stride,
                                                   up_min, rv, lane);
  }
}

// One thread per gather position: the node's descriptor (see GatherDesc). order: node
// per position (null: index order).
__global__ void k_gather_desc(int32_t n_nodes, const int32_t* __restrict__ order,
                              const int8_t* __restrict__ nfm, const int32_t* __restrict__ dof_eq,
                              const int64_t* __restrict__ row_ptr,
                              const int64_t* __restrict__ inc_ptr, GatherDesc* __restrict__ desc) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n_nodes; p += stride) {
    const int32_t a = order ? order[p] : int32_t(p);
    GatherDesc d;
    d.node = a;
    d.flags = 0;
    d.row0 = 0;
    d.len = 0;
    for (int i = 2; i >= 0; --i) {
      const int32_t eq = dof_eq[3 * int64_t(a) + i];
      if (eq < 0) continue;
      d.flags |= 1 << i;
      d.row0 = row_ptr[eq];
      d.len = int32_t(row_ptr[eq + 1] - row_ptr[eq]);
    }
    if (nfm[a] & 8) d.flags |= 8;
    d.inc0 = inc_ptr[a];
    d.inc_n = int32_t(inc_ptr[a + 1] - inc_ptr[a]);
    desc[p] = d;
  }
}

// ---- K2: J2 radial return + consistent tangent + internal force --------------
// Material constants prepared on the host (J2PlasticMaterial ctor,
// src/fem/plasticity.cpp:42-65): kinematic modulus derived from the table slope
// for KINEMATIC/COMBINED, yield table flattened for pure KINEMATIC.
struct J2Dev {
  double E, nu, mu, hkin;
  int32_t n;
  double ep[CXG_MAX_TABLE], sy[CXG_MAX_TABLE];
};

constexpr double kSqrt32 = 1.224744871391589;  // sqrt(3/2)

// Piecewise-linear yield curve with flat ends (plasticity.cpp:67-88).
__device__ __forceinline__ void yield_at(const J2Dev& m, double ep, double& sy, double& h) {
  const int n = m.n;
  if (n == 1 || ep <= m.ep[0]) {
    sy = m.sy[0];
    h = 0.0;
    return;
  }
  if (ep >= m.ep[n - 1]) {
    sy = m.sy[n - 1];
    h = 0.0;
    return;
  }
  for (int i = 0; i + 1 < n; ++i)
    if (ep >= m.ep[i] && ep <= m.ep[i + 1]) {
      h = (m.sy[i + 1] - m.sy[i]) / (m.ep[i + 1] - m.ep[i]);
      sy = m.sy[i] + h * (ep - m.ep[i]);
      return;
    }
  sy = m.sy[n - 1];
  h = 0.0;
}

// Algorithmic tangent of a J2 point in compact form: D_t = iso(lam, mu) - c n n^T,
// iso(l, m) the isotropic Voigt matrix (l + 2m on the normal diagonal, l off it, m on
// the engineering-shear diagonal). The reference's De - (1 - theta) Idev - 2 mu
// theta_bar n n^T (plasticity.cpp:209-234) is exactly this with
// lam = lam_e + (2/3) mu_e (1 - theta), mu = mu_e theta, c = 2 mu_e theta_bar.
struct J2Tangent {
  double lam, mu, c;
  double n[6];
};

// Radial return of one Gauss point: reads the committed history hc[13], writes the
// trial history ht[13], the stress and the compact algorithmic tangent.
__device__ void j2_return(const J2Dev& m, const double eps[6], const double hc[13],
                          double ht[13], double sig[6], J2Tangent& T) {
  double lam, mu;
  lame(m.E, m.nu, lam, mu);
  double ee[6];
  for (int i = 0; i < 6; ++i) ee[i] = eps[i] - hc[i];
  double st[6];
  iso_stress(lam, mu, ee, st);
  const double mean = (st[0] + st[1] + st[2]) / 3.0;
  const double dev[6] = {st[0] - mean, st[1] - mean, st[2] - mean, st[3], st[4], st[5]};
  double xi[6];
  for (int i = 0; i < 6; ++i) xi[i] = dev[i] - hc[7 + i];
  const double nrm = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2] +
                          2.0 * (xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5]));
  double sy0, h0;
  yield_at(m, hc[6], sy0, h0);
  const double qtr = kSqrt32 * nrm;
  T.lam = lam;
  T.mu = mu;
  T.c = 0.0;
  for (int i = 0; i < 6; ++i) T.n[i] = 0.0;
  if (qtr - sy0 <= 0.0 || nrm == 0.0) {
    for (int i = 0; i < 6; ++i) sig[i] = st[i];
    for (int i = 0; i < 13; ++i) ht[i] = hc[i];
    return;
  }
  const double ep0 = hc[6];
  double dg = 0.0;
  for (int it = 0; it < 50; ++it) {
    double sy, h;
    yield_at(m, ep0 + dg, sy, h);
    const double res = qtr - 3.0 * mu * dg - m.hkin * dg - sy;
    const double step = res / (3.0 * mu + m.hkin + h);
    dg += step;
    if (dg < 0.0) dg = 0.0;
    if (fabs(step) <= 1e-14 * (1.0 + fabs(dg))) break;
  }
  const double inv = 1.0 / nrm;
  double n[6];
  for (int i = 0; i < 6; ++i) n[i] = xi[i] * inv;
  const double cdev = 2.0 * mu * dg * kSqrt32;
  const double calpha = (2.0 / 3.0) * m.hkin * dg * kSqrt32;
  const double cep = dg * kSqrt32;
  for (int i = 0; i < 6; ++i) {
    sig[i] = dev[i] - cdev * n[i] + (i < 3 ? mean : 0.0);
    ht[i] = hc[i] + cep * n[i] * (i < 3 ? 1.0 : 2.0);
    ht[7 + i] = hc[7 + i] + calpha * n[i];
  }
  ht[6] = ep0 + dg;
  double sy, h;
  yield_at(m, ep0 + dg, sy, h);
  const double theta = 1.0 - cdev / nrm;
  const double theta_bar = 1.0 / (1.0 + (h + m.hkin) / (3.0 * mu)) - (1.0 - theta);
  const double soft = 1.0 - theta;
  T.lam = lam + (2.0 / 3.0) * mu * soft;
  T.mu = mu * theta;
  T.c = 2.0 * mu * theta_bar;
  for (int i = 0; i < 6; ++i) T.n[i] = n[i];
}

// As j2_return, with the tangent expanded to a 6x6 row-major D.
__device__ void j2_point(const J2Dev& m, const double eps[6], const double hc[13],
                         double ht[13], double sig[6], double D[36]) {
  J2Tangent T;
  j2_return(m, eps, hc, ht, sig, T);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double d = 0.0;
      if (i < 3 && j < 3) d = T.lam + (i == j ? 2.0 * T.mu : 0.0);
      else if (i == j) d = T.mu;
      D[6 * i + j] = d - T.c * T.n[i] * T.n[j];
    }
}

// (D_q B_b)[r][j] for the sparse B column pattern of node b (gradient gb).
__device__ __forceinline__ void d_times_b(const double* D, const double* gb, double DB[18]) {
  // Nonzero rows of B column j: j=0: (0,gx) (3,gy) (4,gz); j=1: (1,gy) (3,gx) (5,gz);
  // j=2: (2,gz) (4,gx) (5,gy).
  for (int r = 0; r < 6; ++r) {
    const double* Dr = D + 6 * r;
    DB[3 * r + 0] = Dr[0] * gb[0] + Dr[3] * gb[1] + Dr[4] * gb[2];
    DB[3 * r + 1] = Dr[1] * gb[1] + Dr[3] * gb[0] + Dr[5] * gb[2];
    DB[3 * r + 2] = Dr[2] * gb[2] + Dr[4] * gb[0] + Dr[5] * gb[1];
  }
}

// K_ab[i][j] += s * sum_r B_a[r][i] (D B_b)[r][j]
__device__ __forceinline__ void bt_db(const double* ga, const double* DB, double s,
                                      double K[9]) {
  for (int j = 0; j < 3; ++j) {
    K[0 + j] += s * (ga[0] * DB[0 + j] + ga[1] * DB[9 + j] + ga[2] * DB[12 + j]);
    K[3 + j] += s * (ga[1] * DB[3 + j] + ga[0] * DB[9 + j] + ga[2] * DB[15 + j]);
    K[6 + j] += s * (ga[2] * DB[6 + j] + ga[0] * DB[12 + j] + ga[1] * DB[15 + j]);
  }
}

template <int NN, int NGP, bool UPPER>
struct J2Lds {
  Geo<NN, NGP> G;
  DofLds<NN> D;
  double u[NN * 3];
  double S[NGP * 6];
  double Dt[NGP * 36];
  int32_t node[NN];
};

template <int NN, int NGP, bool UPPER>
__global__ __launch_bounds__(kWave) void k_assemble_j2(
    const int32_t* __restrict__ conn, const int32_t* __restrict__ elist,
    const int32_t* __restrict__ pptr, int64_t elem_offset,
    int64_t gp_offset, int64_t n_points, const double* __restrict__ dN,
    const double* __restrict__ wq, const double* __restrict__ coords,
    const double* __restrict__ E, const double* __restrict__ nu,
    const int32_t* __restrict__ mat_id, const J2Dev* __restrict__ mats,
    const double* __restrict__ u, const double* __restrict__ committed,
    double* __restrict__ trial, const int32_t* __restrict__ dof_eq,
    const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ slot,
    double* __restrict__ vals, double* __restrict__ f_int, int* __restrict__ flag) {
  __shared__ J2Lds<NN, NGP, UPPER> L;
  const int lane = threadIdx.x;
  const int32_t it_end = pptr[blockIdx.x + 1];
  for (int32_t it = pptr[blockIdx.x]; it < it_end; ++it) {
  const int64_t le = elist[it];
  const int64_t ge = elem_offset + le;
  const int32_t* ce = conn + le * NN;
  stage_coords<NN, NGP>(L.G, L.node, ce, coords, lane);
  stage_dofs<NN>(L.D, ce, dof_eq, row_ptr, lane);
  stage_vec<NN>(L.u, ce, u, lane);
  __syncthreads();
  jacobians<NN, NGP>(L.G, dN, wq, lane, flag);
  __syncthreads();
  gradients<NN, NGP>(L.G, dN, lane);
  __syncthreads();
  const int32_t mid = mat_id ? mat_id[ge] : -1;
  for (int q = lane; q < NGP; q += kWave) {
    double H[3][3], e[6], s[6];
    grad_field<NN, NGP>(L.G, L.u, q, H);
    voigt_strain(H, e);
    const int64_t pt = gp_offset + le * NGP + q;
    double* Dq = L.Dt + 36 * q;
    if (mid >= 0) {
      double hc[13], ht[13];
      for (int c = 0; c < 13; ++c) hc[c] = committed[c * n_points + pt];
      j2_point(mats[mid], e, hc, ht, s, Dq);
      for (int c = 0; c < 13; ++c) trial[c * n_points + pt] = ht[c];
    } else {
      double lam, mu;
      lame(E[ge], nu[ge], lam, mu);
      iso_stress(lam, mu, e, s);
      for (int i = 0; i < 36; ++i) Dq[i] = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Dq[6 * i + j] = (i == j) ? lam + 2.0 * mu : lam;
      for (int i = 3; i < 6; ++i) Dq[7 * i] = mu;
      for (int c = 0; c < 13; ++c) trial[c * n_points + pt] = committed[c * n_points + pt];
    }
    for (int i = 0; i < 6; ++i) L.S[6 * q + i] = L.G.sc[q] * s[i];
  }
  __syncthreads();
  node_forces<NN, NGP>(L.G, L.S, L.node, f_int, lane);
  const int32_t* se = slot + le * NN * NN;
  constexpr int NP = UPPER ? NN * (NN + 1) / 2 : NN * NN;
  for (int p = lane; p < NP; p += kWave) {
    int a, b;
    decode_pair<NN, UPPER>(p, a, b);
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < NGP; ++q) {
      double DB[18];
      d_times_b(L.Dt + 36 * q, L.G.g + (q * NN + b) * 3, DB);
      bt_db(L.G.g + (q * NN + a) * 3, DB, L.G.sc[q], K);
    }
    scatter_pair<NN>(L.D, se, a, b, K, UPPER && a != b, vals);
  }
  __syncthreads();  // LDS is reused by the next element of the patch
  }
}

// ---- K2, two-phase variant: return map, K_t and f_e to the element arenas ------------
// Phase 1, one lane per Gauss point: Jacobian, strain (gradients formed on the fly),
// radial return (j2_return), trial history to the SoA state, and the compact tangent
// (lam, mu, c, n) and rs*sigma to LDS. Both the per-family and the persistent kernel run
// EPW elements per wave here, so that most lanes carry a Gauss point (6 C3D8 elements of
// 8 points, 2 C3D20 of 27, 4 C3D20R of 8, 24 C3D6 of 2; for C3D8, 6 measured faster than
// 8 and 4 (see j2_epw); C3D20R and C3D6 stay below the C3D20 record, which sets the LDS
// of the persistent kernel's union). The coordinates and displacements, read only here,
// share the LDS with phase 2's workspace.
// Phase 2, one element at a time on the whole wave: QC Gauss points at a time,
// h = rs g and t = B(h)^T n per node go to LDS, and each lane accumulates its upper
// node-pair blocks
//   K_ab += lam h_a h_b^T + mu h_b h_a^T + mu (h_a . h_b) I - c t_a t_b^T
// (= s B_a^T D_t B_b with s = det J w, rs = sqrt(s)), while lanes < n accumulate the
// element force f_a += B(h_a)^T (rs sigma). K_t is written as full rows through the
// staged, coalesced stores of k_elem_iso; f_e (3n) to its own arena. The row gather
// then sums K_t rows and f_e entries per node in element order.
// What phase 2 needs of an element (kept from phase 1) ...
template <int NN, int NGP>
struct J2Elem {
  double Ai[NGP * 9];
  double rs[NGP];
  double T[NGP * 9];  // lam, mu, c, n[6] per Gauss point
  double S[NGP * 6];  // rs * sigma
};
// ... and what only phase 1 reads: it shares the LDS with phase 2's workspace.
template <int NN>
struct J2Geo {
  double x[NN * 3];
  double u[NN * 3];
};

template <int NN, int NGP, int QC, int EPW>
struct ElemJ2Lds {
  J2Elem<NN, NGP> e[EPW];
  union {
    J2Geo<NN> g[EPW];  // phase 1
    struct {
      double h[QC * NN * 3];
      double t[QC * NN * 3];
    } c;               // phase 2
    alignas(16) double stage[kStageNodes * 3 * 3 * NN];
  } w;
};

// Resident waves per SIMD the J2 kernel is compiled for: the 8-node kernel fits 3 (168
// VGPRs, 2 spilled), the others keep the compiler's choice.
constexpr int j2_min_waves(int nn) { return nn == 8 ? 3 : 1; }

// Elements per wave of the J2 kernels (see above). C3D8: 6, whose 12.7 KB of LDS lets the
// 3 waves per SIMD of its registers reside (8 per wave, 15.9 KB, allow 10 waves per CU:
// 2.43 against 2.22 ms at 128x128x64).
constexpr int j2_epw(int nn, int ngp) {
  return ngp >= 27 ? 2 : (nn == 20 ? 4 : (nn == 8 ? 6 : (64 / ngp > 24 ? 24 : 64 / ngp)));
}

// Node k of an element into its LDS record.
template <int NN>
__device__ __forceinline__ void j2_stage(J2Geo<NN>& P, const int32_t* __restrict__ ce,
                                         const double* __restrict__ coords,
                                         const double* __restrict__ u, int k) {
  const int64_t n3 = 3 * int64_t(ce[k]);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    P.x[3 * k + c] = coords[n3 + c];
    P.u[3 * k + c] = u[n3 + c];
  }
}

// Phase 1 for Gauss point q (x, u staged in G). mat == nullptr: isotropic elastic with
// Ee, nue. pt: the point's index in the SoA state.
template <int NN, int NGP>
__device__ __forceinline__ void j2_gauss_point(
    J2Elem<NN, NGP>& P, const J2Geo<NN>& G, const int q, const double Ee, const double nue,
    const J2Dev* mat,
    const int64_t pt, const int64_t n_points, const double* __restrict__ dN,
    const double* __restrict__ wq, const double* __restrict__ committed,
    double* __restrict__ trial, int* __restrict__ flag) {
  const double* d = dn_point<NN, NGP>(dN, q);
  double* A = P.Ai + 9 * q;
  double rs;
  jacobian_inverse<NN, dn_stride<NGP>()>(G.x, d, wq[q], A, rs, flag);
  P.rs[q] = rs;
  double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int k = 0; k < NN; ++k) {
    constexpr int S = dn_stride<NGP>();
    const double d0 = d[(3 * k) * S], d1 = d[(3 * k + 1) * S], d2 = d[(3 * k + 2) * S];
    const double g0 = A[0] * d0 + A[3] * d1 + A[6] * d2;
    const double g1 = A[1] * d0 + A[4] * d1 + A[7] * d2;
    const double g2 = A[2] * d0 + A[5] * d1 + A[8] * d2;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double v = G.u[3 * k + i];
      H[i][0] += v * g0;
      H[i][1] += v * g1;
      H[i][2] += v * g2;
    }
  }
  double e[6], sg[6];
  voigt_strain(H, e);
  J2Tangent T;
  // the history planes stream: read once and written once per iteration (no reuse
  // that a cache could serve)
  if (mat) {
    double hc[13], ht[13];
#pragma unroll
    for (int c = 0; c < 13; ++c) hc[c] = __builtin_nontemporal_load(committed + c * n_points + pt);
    j2_return(*mat, e, hc, ht, sg, T);
#pragma unroll
    for (int c = 0; c < 13; ++c) __builtin_nontemporal_store(ht[c], trial + c * n_points + pt);
  } else {
    lame(Ee, nue, T.lam, T.mu);
    iso_stress(T.lam, T.mu, e, sg);
    T.c = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) T.n[i] = 0.0;
#pragma unroll
    for (int c = 0; c < 13; ++c)
      __builtin_nontemporal_store(__builtin_nontemporal_load(committed + c * n_points + pt),
                                  trial + c * n_points + pt);
  }
  double* Tq = P.T + 9 * q;
  Tq[0] = T.lam;
  Tq[1] = T.mu;
  Tq[2] = T.c;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    Tq[3 + i] = T.n[i];
    P.S[6 * q + i] = rs * sg[i];
  }
}

// Phase 2's per-point pieces, with the multiply-adds written out so that every kernel
// forming them gets the same bits. h = rs J^-T dN_k and t = B(h)^T n of node k at point q.
template <int NN, int NGP>
__device__ __forceinline__ void j2_gradients(const J2Elem<NN, NGP>& P, const int q,
                                             const double* __restrict__ d, double h[3],
                                             double t[3]) {
  scaled_gradient(P.Ai + 9 * q, P.rs[q], d, h);
  const double* n = P.T + 9 * q + 3;
  t[0] = __builtin_fma(h[2], n[4], __builtin_fma(h[1], n[3], h[0] * n[0]));
  t[1] = __builtin_fma(h[2], n[5], __builtin_fma(h[0], n[3], h[1] * n[1]));
  t[2] = __builtin_fma(h[1], n[5], __builtin_fma(h[0], n[4], h[2] * n[2]));
}

// K_ab += lam h_a h_b^T + mu h_b h_a^T + mu (h_a . h_b) I - c t_a t_b^T at one point.
__device__ __forceinline__ void j2_pair_add(double K[9], const double av[3], const double bv[3],
                                            const double ta[3], const double tb[3],
                                            const double lam, const double mu, const double c) {
  const double la[3] = {lam * av[0], lam * av[1], lam * av[2]};
  const double mb[3] = {mu * bv[0], mu * bv[1], mu * bv[2]};
  const double ct[3] = {c * ta[0], c * ta[1], c * ta[2]};
  const double dg = mu * __builtin_fma(av[2], bv[2], __builtin_fma(av[1], bv[1], av[0] * bv[0]));
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      K[3 * i + j] = __builtin_fma(-ct[i], tb[j], __builtin_fma(mb[i], av[j], __builtin_fma(la[i], bv[j], K[3 * i + j])));
  K[0] += dg;
  K[4] += dg;
  K[8] += dg;
}

// Phase 2 for one element: every lane of the wave calls it; ends with a barrier. UP: K_t
// (symmetric) as its packed upper node-pair blocks (store_upper_blocks) instead of full
// rows.
template <int NN, int NGP, int QC, class W, bool UP = false>
__device__ __forceinline__ void j2_element_matrix(const J2Elem<NN, NGP>& P, W& w, const int lane,
                                                  const double* __restrict__ dN,
                                                  double* __restrict__ kel,
                                                  double* __restrict__ fel) {
  constexpr int NP = NN * (NN + 1) / 2;
  constexpr int PPL = (NP + kWave - 1) / kWave;
  constexpr int ND = 3 * NN;
  // TC (upper-block arena, four pairs per lane, even node count): a lane owns a 2x2 tile
  // of node couples (2s, 2s+1) x (2t, 2t+1), s <= t (of a diagonal tile the lower pair
  // (2s+1, 2s) is computed and not stored), and reads each couple's h and t with three
  // 16-byte LDS reads: 24 doubles per Gauss point against 30 for (a, b0 .. b0+3).
  // Otherwise (and in the full-row persistent kernel, where the tiles' registers spill) a
  // lane owns node a and up to PPL consecutive partners b (see decode_group). Either way
  // every pair is summed by j2_pair_add in the same order with the same operands.
  constexpr bool TC = UP && PPL == 4 && NN % 2 == 0;
  int pa[PPL], pb[PPL];
  int ga = 0, gb0 = 0;
  int cs = 0, ct = 0;  // TC: the lane's couples (clamped for idle lanes)
  if constexpr (TC) {
    int s_ = -1, t_ = -1;
    decode_group<NN / 2, 1>(lane, s_, t_);
#pragma unroll
    for (int t = 0; t < PPL; ++t) {
      const int a = 2 * s_ + t / 2, b = 2 * t_ + t % 2;
      const bool ok = s_ >= 0 && a <= b;
      pa[t] = ok ? a : -1;
      pb[t] = ok ? b : -1;
    }
    cs = s_ >= 0 ? s_ : 0;
    ct = s_ >= 0 ? t_ : 0;
  } else {
    decode_group<NN, PPL>(lane, ga, gb0);
#pragma unroll
    for (int t = 0; t < PPL; ++t) {
      const bool ok = ga >= 0 && gb0 + t < NN;
      pa[t] = ok ? ga : -1;
      pb[t] = ok ? gb0 + t : -1;
    }
  }
  const int ga_c = ga >= 0 ? ga : 0;
  double K[PPL][9];
#pragma unroll
  for (int t = 0; t < PPL; ++t)
#pragma unroll
    for (int i = 0; i < 9; ++i) K[t][i] = 0.0;
  // element force f_k += B(h_k)^T (rs sigma) per point (sq = rs sigma in Voigt order):
  // lane 3k + c sums component c of node k, g[c] sq[c] first, then the two shear terms
  // with explicit multiply-adds, its operands picked by lane
  double f = 0.0;
  const int fk = lane / 3, fc = lane - 3 * (lane / 3);
  const int g1i = fc == 0 ? 1 : 0, s1i = fc == 2 ? 4 : 3;
  const int g2i = fc == 2 ? 1 : 2, s2i = fc == 0 ? 4 : 5;
  for (int q0 = 0; q0 < NGP; q0 += QC) {
    for (int idx = lane; idx < QC * NN; idx += kWave) {
      const int q = q0 + idx / NN;
      double h[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
      if (q < NGP) j2_gradients(P, q, dN + (q * NN + idx % NN) * 3, h, t);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        w.c.h[3 * idx + i] = h[i];
        w.c.t[3 * idx + i] = t[i];
      }
    }
    __syncthreads();
    const int qn = NGP - q0 < QC ? NGP - q0 : QC;
    if (lane < 3 * NN)
      for (int qq = 0; qq < qn; ++qq) {
        const double* g = w.c.h + (qq * NN + fk) * 3;
        const double* sq = P.S + 6 * (q0 + qq);
        f += __builtin_fma(g[g2i], sq[s2i], __builtin_fma(g[g1i], sq[s1i], g[fc] * sq[fc]));
      }
#pragma unroll 1
    for (int qq = 0; qq < qn && TC; ++qq) {
      const double* Tq = P.T + 9 * (q0 + qq);
      // couples: 6 contiguous doubles, 16-byte aligned (h and t are, and 2 nodes are 48
      // bytes)
      const f64x2* hs = reinterpret_cast<const f64x2*>(w.c.h + (qq * NN + 2 * cs) * 3);
      const f64x2* hc = reinterpret_cast<const f64x2*>(w.c.h + (qq * NN + 2 * ct) * 3);
      const f64x2* ts = reinterpret_cast<const f64x2*>(w.c.t + (qq * NN + 2 * cs) * 3);
      const f64x2* tc = reinterpret_cast<const f64x2*>(w.c.t + (qq * NN + 2 * ct) * 3);
      const f64x2 s0 = hs[0], s1 = hs[1], s2 = hs[2], c0 = hc[0], c1 = hc[1], c2 = hc[2];
      const f64x2 u0 = ts[0], u1 = ts[1], u2 = ts[2], v0 = tc[0], v1 = tc[1], v2 = tc[2];
      const double A[2][3] = {{s0.x, s0.y, s1.x}, {s1.y, s2.x, s2.y}};
      const double B[2][3] = {{c0.x, c0.y, c1.x}, {c1.y, c2.x, c2.y}};
      const double TA[2][3] = {{u0.x, u0.y, u1.x}, {u1.y, u2.x, u2.y}};
      const double TB[2][3] = {{v0.x, v0.y, v1.x}, {v1.y, v2.x, v2.y}};
#pragma unroll
      for (int t = 0; t < PPL; ++t)
        j2_pair_add(K[t], A[t / 2], B[t % 2], TA[t / 2], TB[t % 2], Tq[0], Tq[1], Tq[2]);
    }
#pragma unroll 1
    for (int qq = 0; qq < qn && !TC; ++qq) {
      const double* Tq = P.T + 9 * (q0 + qq);
      const double* ha = w.c.h + (qq * NN + ga_c) * 3;
      const double* ta = w.c.t + (qq * NN + ga_c) * 3;
      const double av[3] = {ha[0], ha[1], ha[2]}, tav[3] = {ta[0], ta[1], ta[2]};
#pragma unroll
      for (int t = 0; t < PPL; ++t) {
        const int b = gb0 + t < NN ? (gb0 + t >= 0 ? gb0 + t : 0) : NN - 1;
        const double* hb = w.c.h + (qq * NN + b) * 3;
        const double* tb = w.c.t + (qq * NN + b) * 3;
        const double bv[3] = {hb[0], hb[1], hb[2]}, tbv[3] = {tb[0], tb[1], tb[2]};
        j2_pair_add(K[t], av, bv, tav, tbv, Tq[0], Tq[1], Tq[2]);
      }
    }
    __syncthreads();
  }
  if (lane < 3 * NN) fel[lane] = f;
  if constexpr (UP) {
    store_upper_blocks<NN, PPL>(w.stage, pa, pb, K, kel, lane);
    return;
  }
  for (int g0 = 0; g0 < NN; g0 += kStageNodes) {
#pragma unroll
    for (int t = 0; t < PPL; ++t) {
      if (pa[t] < 0) continue;
      if (pa[t] >= g0 && pa[t] < g0 + kStageNodes) {
        double* ra = w.stage + (3 * (pa[t] - g0)) * ND + 3 * pb[t];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) ra[i * ND + j] = K[t][3 * i + j];
      }
      if (pb[t] != pa[t] && pb[t] >= g0 && pb[t] < g0 + kStageNodes) {
        double* rb = w.stage + (3 * (pb[t] - g0)) * ND + 3 * pa[t];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int i = 0; i < 3; ++i) rb[j * ND + i] = K[t][3 * i + j];
      }
    }
    __syncthreads();
    const int nrow = 3 * (NN - g0 < kStageNodes ? NN - g0 : kStageNodes);
    f64x2* dst = reinterpret_cast<f64x2*>(kel + int64_t(3 * g0) * ND);
    const f64x2* src = reinterpret_cast<const f64x2*>(w.stage);
    for (int idx = lane; idx < nrow * ND / 2; idx += kWave)
      __builtin_nontemporal_store(src[idx], dst + idx);
    __syncthreads();
  }
}

// ne <= EPW consecutive elements of one block (first block-local le0) on one wave:
// phase 1 for all their Gauss points at once, then phase 2 element by element. kp and
// fe point at the block's arenas. Ends with a barrier.
template <int NN, int NGP, int QC, int EPW, bool UP = false>
__device__ __forceinline__ void j2_group(
    ElemJ2Lds<NN, NGP, QC, EPW>& L, const int lane, const int ne, const int64_t le0,
    const int32_t* __restrict__ conn, const int64_t elem_offset, const int64_t gp_offset,
    const int64_t n_points, const double* __restrict__ dN, const double* __restrict__ wq,
    const double* __restrict__ coords, const double* __restrict__ E,
    const double* __restrict__ nu, const int32_t* __restrict__ mat_id,
    const J2Dev* __restrict__ mats, const double* __restrict__ u,
    const double* __restrict__ committed, double* __restrict__ trial, double* __restrict__ kp,
    double* __restrict__ fe, int* __restrict__ flag) {
  for (int idx = lane; idx < ne * NN; idx += kWave) {
    const int e = idx / NN;
    j2_stage<NN>(L.w.g[e], conn + (le0 + e) * NN, coords, u, idx - e * NN);
  }
  __syncthreads();
  for (int idx = lane; idx < ne * NGP; idx += kWave) {
    const int e = idx / NGP, q = idx - e * NGP;
    const int64_t ge = elem_offset + le0 + e;
    const int32_t mid = mat_id ? mat_id[ge] : -1;
    j2_gauss_point<NN, NGP>(L.e[e], L.w.g[e], q, E[ge], nu[ge], mid >= 0 ? mats + mid : nullptr,
                            gp_offset + (le0 + e) * NGP + q, n_points, dN, wq, committed, trial,
                            flag);
  }
  __syncthreads();
  for (int e = 0; e < ne; ++e)
    j2_element_matrix<NN, NGP, QC, decltype(L.w), UP>(L.e[e], L.w, lane, dN,
                                                      kp + (le0 + e) * (9 * NN * NN),
                                                      fe + (le0 + e) * (3 * NN));
}

// EPW elements per wave (see j2_group). UP: upper-block arena (asm_variant 7).
template <int NN, int NGP, int QC, bool UP = false>
__global__ __launch_bounds__(kWave, j2_min_waves(NN)) void k_elem_j2(
    const int32_t* __restrict__ conn, int64_t elem_offset, int64_t n_elem, int64_t gp_offset,
    int64_t n_points, const double* __restrict__ dN, const double* __restrict__ wq,
    const double* __restrict__ coords, const double* __restrict__ E,
    const double* __restrict__ nu, const int32_t* __restrict__ mat_id,
    const J2Dev* __restrict__ mats, const double* __restrict__ u,
    const double* __restrict__ committed, double* __restrict__ trial, double* __restrict__ kp,
    double* __restrict__ fe, int* __restrict__ flag, int64_t slab) {
  constexpr int EPW = j2_epw(NN, NGP);
  __shared__ ElemJ2Lds<NN, NGP, QC, EPW> L;
  const int lane = threadIdx.x;
  const int64_t grp = xcd_index(blockIdx.x, (n_elem + EPW - 1) / EPW, slab);
  if (grp < 0) return;
  const int64_t le0 = grp * EPW;
  const int ne = int(n_elem - le0 < EPW ? n_elem - le0 : EPW);
  j2_group<NN, NGP, QC, EPW, UP>(L, lane, ne, le0, conn, elem_offset, gp_offset, n_points, dN,
                                 wq, coords, E, nu, mat_id, mats, u, committed, trial, kp, fe,
                                 flag);
}

// ---- Persistent element kernels (asm_variant 5) ------------------------------------
// One launch covers every block of the mesh, whatever its family: a grid of resident
// workgroups (one wave each) takes work items (runs of consecutive elements of one
// block) from a shared counter until the counter passes the last item, so meshes of
// very uneven size and mixed families balance across the compute units without one
// launch per family. Items are ordered costliest family first. The element arithmetic
// is the same device code as the per-family kernels (elem_iso_body / j2_group);
// the LDS is the union of the families' layouts.
union IsoLdsAll {
  ElemLds<8, 8, 8> c3d8;
  ElemLds<20, 27, 3> c3d20;
  ElemLds<20, 8, 3> c3d20r;
  ElemLds<6, 2, 2> c3d6;
};
union J2LdsAll {
  ElemJ2Lds<8, 8, 8, j2_epw(8, 8)> c3d8;
  ElemJ2Lds<20, 27, 3, j2_epw(20, 27)> c3d20;
  ElemJ2Lds<20, 8, 3, j2_epw(20, 8)> c3d20r;
  ElemJ2Lds<6, 2, 2, j2_epw(6, 2)> c3d6;
};

// Next work item for the whole wave (lane 0 takes it from the counter).
__device__ __forceinline__ int32_t next_item(int32_t* queue, int32_t* slot, int lane) {
  if (lane == 0) *slot = atomicAdd(queue, 1);
  __syncthreads();
  const int32_t w = *slot;
  __syncthreads();  // every lane has read the slot before lane 0 writes it again
  return w;
}

__global__ void k_reset_queue(int32_t* queue) {
  if (threadIdx.x == 0) *queue = 0;
}

__global__ __launch_bounds__(kWave, 4) void k_elem_iso_persistent(
    const WorkChunk* __restrict__ work, int32_t n_work, int32_t* __restrict__ queue,
    const BlockDesc* __restrict__ blocks, const int32_t* __restrict__ conn,
    const double* __restrict__ coords, const double* __restrict__ E,
    const double* __restrict__ nu, double* __restrict__ kp, int* __restrict__ flag) {
  __shared__ IsoLdsAll L;
  __shared__ int32_t slot;
  const int lane = threadIdx.x;
  for (;;) {
    const int32_t w = next_item(queue, &slot, lane);
    if (w >= n_work) break;
    const WorkChunk c = work[w];
    const BlockDesc B = blocks[c.block];
    // The lane index is passed through an empty register-only asm per element, so the
    // compiler cannot hoist the per-family lane -> node-pair decoding out of the loops
    // (hoisted for all four families at once it would exceed the register budget).
#define CXG_ISO_RUN(NN, NGP, QC, LDS)                                                      \
  for (int32_t i = 0; i < c.count; ++i) {                                                \
    const int64_t le = int64_t(c.first) + i;                                             \
    const int64_t ge = B.elem_offset + le;                                               \
    int ln = lane;                                                                       \
    asm volatile("" : "+v"(ln));                                                         \
    elem_iso_body<NN, NGP, QC>(LDS, ln, conn + B.conn_offset + le * NN, E[ge], nu[ge],   \
                               B.dN, B.w, coords, kp + B.kp_base + le * (9 * NN * NN),   \
                               flag);                                                    \
  }
    switch (B.type) {
      case CXG_C3D8: CXG_ISO_RUN(8, 8, 8, L.c3d8) break;
      case CXG_C3D20: CXG_ISO_RUN(20, 27, 3, L.c3d20) break;
      case CXG_C3D20R: CXG_ISO_RUN(20, 8, 3, L.c3d20r) break;
      case CXG_C3D6: CXG_ISO_RUN(6, 2, 2, L.c3d6) break;
      default: break;
    }
#undef CXG_ISO_RUN
  }
}

__global__ __launch_bounds__(kWave, 2) void k_elem_j2_persistent(
    const WorkChunk* __restrict__ work, int32_t n_work, int32_t* __restrict__ queue,
    const BlockDesc* __restrict__ blocks, const int32_t* __restrict__ conn,
    const double* __restrict__ coords, const double* __restrict__ E,
    const double* __restrict__ nu, const int32_t* __restrict__ mat_id,
    const J2Dev* __restrict__ mats, int64_t n_points, const double* __restrict__ u,
    const double* __restrict__ committed, double* __restrict__ trial, double* __restrict__ kp,
    double* __restrict__ fe, int* __restrict__ flag) {
  __shared__ J2LdsAll L;
  __shared__ int32_t slot;
  const int lane = threadIdx.x;
  for (;;) {
    const int32_t w = next_item(queue, &slot, lane);
    if (w >= n_work) break;
    const WorkChunk c = work[w];
    const BlockDesc B = blocks[c.block];
#define CXG_J2_RUN(NN, NGP, QC, LDS)                                                        \
  for (int32_t i = 0; i < c.count; i += j2_epw(NN, NGP)) {                                    \
    const int ne = c.count - i < j2_epw(NN, NGP) ? c.count - i : j2_epw(NN, NGP);                 \
    int ln = lane;                                                                        \
    asm volatile("" : "+v"(ln));                                                          \
    j2_group<NN, NGP, QC, j2_epw(NN, NGP)>(LDS, ln, ne, int64_t(c.first) + i,                 \
                                       conn + B.conn_offset, B.elem_offset, B.gp_offset,  \
                                       n_points, B.dN, B.w, coords, E, nu, mat_id, mats,  \
                                       u, committed, trial, kp + B.kp_base,               \
                                       fe + 3 * B.conn_offset, flag);                     \
  }
    switch (B.type) {
      case CXG_C3D8: CXG_J2_RUN(8, 8, 8, L.c3d8) break;
      case CXG_C3D20: CXG_J2_RUN(20, 27, 3, L.c3d20) break;
      case CXG_C3D20R: CXG_J2_RUN(20, 8, 3, L.c3d20r) break;
      case CXG_C3D6: CXG_J2_RUN(6, 2, 2, L.c3d6) break;
      default: break;
    }
#undef CXG_J2_RUN
  }
}

// MaterialModel::evaluate at independent material points (one thread per point, grid
// stride): J2PlasticMaterial::evaluate (plasticity.cpp:90-236) for points with a J2
// material, ElasticIsoMaterial::evaluate (element.cpp:856-871) for the others, whose
// trial history is left untouched as the reference leaves it.
__global__ void k_material_points(int64_t n, const int32_t* __restrict__ mat_id,
                                  const J2Dev* __restrict__ mats, const double* __restrict__ E,
                                  const double* __restrict__ nu, const double* __restrict__ strain,
                                  const double* __restrict__ committed, double* __restrict__ trial,
                                  double* __restrict__ stress, double* __restrict__ tangent) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += stride) {
    double e[6], sg[6], D[36];
#pragma unroll
    for (int i = 0; i < 6; ++i) e[i] = strain[6 * p + i];
    const int32_t id = mat_id ? mat_id[p] : -1;
    if (id >= 0) {
      double hc[13], ht[13];
#pragma unroll
      for (int c = 0; c < 13; ++c) hc[c] = committed[c * n + p];
      j2_point(mats[id], e, hc, ht, sg, D);
#pragma unroll
      for (int c = 0; c < 13; ++c) trial[c * n + p] = ht[c];
    } else {
      double lam, mu;
      lame(E[p], nu[p], lam, mu);
      iso_stress(lam, mu, e, sg);
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
          D[6 * i + j] = i < 3 && j < 3 ? lam + (i == j ? 2.0 * mu : 0.0) : (i == j ? mu : 0.0);
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) stress[6 * p + i] = sg[i];
    for (int k = 0; k < 36; ++k) tangent[36 * p + k] = D[k];
  }
}

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

// rhs[eq(g)] += f_ext[g] - f_c[g] on free DOFs.
__global__ void k_reduce_free(int64_t n_dof, const int32_t* __restrict__ dof_eq,
                              const double* __restrict__ f_ext, const double* __restrict__ f_c,
                              double* __restrict__ rhs) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < n_dof; g += stride) {
    const int32_t eq = dof_eq[g];
    if (eq < 0) continue;
    double v = 0.0;
    if (f_ext) v += f_ext[g];
    if (f_c) v -= f_c[g];
    rhs[eq] += v;
  }
}


// ---- K7: field recovery (stress.cpp:147-353) ----------------------------------------
// Phase 1, one wave per element: Gauss-point strain (B u, engineering shear) and stress
// (D eps, or the radial return from the committed history when a J2 table is given),
// then per element the Gauss-point means (the hex/wedge extrapolation of
// extrapolation.hpp:26-55 is the mean at every node), the element force
// f_e = sum_q det*w B^T sigma into the element-force arena, and optionally every point's
// stress. Phase 2, one thread per node: mean over incident elements in ascending element
// order, f_int from the arena, RF = f_int - f_ext.
template <int NN, int NGP>
struct RecoverLds {
  union {
    struct {
      double x[NN * 3];
      double u[NN * 3];
    };
    double eps[NGP * 6];  // written once the points' x and u reads are done
  };
  double P[NGP * 9];    // det J w sigma J^-1 per Gauss point (row-major; see ForceLds)
  double sig[NGP * 6];
  int32_t node[NN];
};

// Physical gradient g_k = J^-T dN_k at one Gauss point (A = J^-1 row-major; the natural
// derivatives at d[c S], as in jacobian_detw).
template <int S = 1>
__device__ __forceinline__ void grad_at(const double* A, const double* d, double g[3]) {
  const double d0 = d[0], d1 = d[S], d2 = d[2 * S];
  g[0] = A[0] * d0 + A[3] * d1 + A[6] * d2;
  g[1] = A[1] * d0 + A[4] * d1 + A[7] * d2;
  g[2] = A[2] * d0 + A[5] * d1 + A[8] * d2;
}

// Displacement gradient H = sum_k u_k (J^-T d_k)^T at one Gauss point (A = J^-1
// row-major, natural derivatives at d[(3k + c) S]) as G J^-1 with G = sum_k u_k d_k^T: 9
// multiply-adds per node and 27 once, where forming every g_k first takes 18 per node.
template <int NN, int S>
__device__ __forceinline__ void displacement_gradient(const double* u, const double* d,
                                                      const double* A, double H[3][3]) {
  double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int k = 0; k < NN; ++k) {
    const double d0 = d[(3 * k) * S], d1 = d[(3 * k + 1) * S], d2 = d[(3 * k + 2) * S];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double v = u[3 * k + i];
      G[i][0] += v * d0;
      G[i][1] += v * d1;
      G[i][2] += v * d2;
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) H[i][j] = G[i][0] * A[j] + G[i][1] * A[3 + j] + G[i][2] * A[6 + j];
}

// The C3D8 natural derivatives in registers (hex8_dn; no table reads). A point lane's, at
// d[3k + c]: 12 distinct products of the factors 1 + g and 1 - g, their signs those of the
// node corners.
__device__ __forceinline__ void hex8_dn_point(int q, double d[24]) {
  constexpr double gp = 1.0 + kGauss2, gm = 1.0 - kGauss2;
  double f[3][2];  // f[a][s]: the factor along axis a of a node on side s (1: +) of the point
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const bool p = (q >> a) & 1;
    f[a][1] = p ? gp : gm;
    f[a][0] = p ? gm : gp;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int a = c == 0 ? 1 : 0, b = c == 2 ? 1 : 2;
      const double v = (0.125 * f[a][hex_corner_pos(k, a)]) * f[b][hex_corner_pos(k, b)];
      d[3 * k + c] = hex_corner_pos(k, c) ? v : -v;
    }
}

// Calls fn(q, d0, d1, d2) for the NGP points in order with node k's natural derivatives:
// for C3D8 from the closed form (the 12 values of node k, by the point's sides along the
// other two axes; the loop unrolled), else from the point-major table dN.
template <int NN, int NGP, class Fn>
__device__ __forceinline__ void for_node_points(const double* dN, int k, Fn&& fn) {
  if constexpr (NN == 8 && NGP == 8) {
    constexpr double gp = 1.0 + kGauss2, gm = 1.0 - kGauss2;
    double f[3][2];  // f[a][s]: the factor along axis a at the points on side s (1: +)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const bool p = hex_corner_pos(k, a);
      f[a][1] = p ? gp : gm;
      f[a][0] = p ? gm : gp;
    }
    double e[3][2][2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double s = hex_corner_pos(k, c) ? 0.125 : -0.125;
      const int a = c == 0 ? 1 : 0, b = c == 2 ? 1 : 2;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) e[c][i][j] = (s * f[a][i]) * f[b][j];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      fn(q, e[0][(q >> 1) & 1][(q >> 2) & 1], e[1][q & 1][(q >> 2) & 1], e[2][q & 1][(q >> 1) & 1]);
  } else {
    for (int q = 0; q < NGP; ++q) {
      const double* d = dN + (q * NN + k) * 3;
      fn(q, d[0], d[1], d[2]);
    }
  }
}

// Elements per wave of the recovery kernel: one lane per (element, Gauss point) and per
// (element, node) (2 C3D20, 3 C3D20R, 8 C3D8, 10 C3D6).
constexpr int rec_epw(int nn, int ngp) {
  return (64 / ngp < 64 / nn ? 64 / ngp : 64 / nn) > 0 ? (64 / ngp < 64 / nn ? 64 / ngp : 64 / nn) : 1;
}

// J2: the radial return for elements with a J2 material (else every element elastic).
// Gradients are formed on the fly from J^-1 and the natural derivatives, which keeps the
// LDS at the elements' coordinates, displacements, J^-1 and the point fields. EPW
// consecutive elements per wave.
template <int NN, int NGP, bool J2>
__global__ __launch_bounds__(kWave, 4) void k_gauss_fields(
    const int32_t* __restrict__ conn, int64_t elem_offset, int64_t gp_offset,
    int64_t n_points, const double* __restrict__ dN, const double* __restrict__ wq,
    const double* __restrict__ coords, const double* __restrict__ E,
    const double* __restrict__ nu, const double* __restrict__ u,
    const int32_t* __restrict__ mat_id, const J2Dev* __restrict__ mats,
    const double* __restrict__ committed, double* __restrict__ emean,
    const int32_t* __restrict__ ipos, double* __restrict__ fe, double* __restrict__ gauss_stress,
    int* __restrict__ flag, int64_t n_elem, int64_t slab) {
  constexpr int EPW = rec_epw(NN, NGP);
  __shared__ RecoverLds<NN, NGP> L[EPW];
  const int lane = threadIdx.x;
  const int64_t grp = xcd_index(blockIdx.x, (n_elem + EPW - 1) / EPW, slab);
  if (grp < 0) return;
  const int64_t le0 = grp * EPW;
  const int ne = int(n_elem - le0 < EPW ? n_elem - le0 : EPW);
  for (int idx = lane; idx < ne * NN; idx += kWave) {
    const int e = idx / NN, k = idx - e * NN;
    const int32_t nd = conn[(le0 + e) * NN + k];
    L[e].node[k] = nd;
    for (int c = 0; c < 3; ++c) {
      L[e].x[3 * k + c] = coords[3 * int64_t(nd) + c];
      L[e].u[3 * k + c] = u[3 * int64_t(nd) + c];
    }
  }
  __syncthreads();
  static_assert(EPW * NGP <= kWave && EPW * NN <= kWave, "one lane per point and per node");
  const bool pt_lane = lane < ne * NGP;
  const int e = pt_lane ? lane / NGP : 0, q = pt_lane ? lane - e * NGP : 0;
  RecoverLds<NN, NGP>& R = L[e];
  double A[9], e6[6], sg[6], sc = 0.0;
  if (pt_lane) {
    const int64_t le = le0 + e, eg = elem_offset + le;
    const double* d = dn_point<NN, NGP>(dN, q);
    sc = jacobian_detw<NN, true, dn_stride<NGP>()>(R.x, d, wq[q], A, flag);
    double H[3][3];
    displacement_gradient<NN, dn_stride<NGP>()>(R.u, d, A, H);
    voigt_strain(H, e6);
    const int64_t pt = gp_offset + le * NGP + q;
    const int32_t mid = J2 ? mat_id[eg] : -1;
    if (J2 && mid >= 0) {
      // the committed history is read, never advanced (the reference evaluates a copy)
      double hc[13], ht[13];
      for (int c = 0; c < 13; ++c) hc[c] = committed[c * n_points + pt];
      J2Tangent T;
      j2_return(mats[mid], e6, hc, ht, sg, T);
    } else {
      double lam, mu;
      lame(E[eg], nu[eg], lam, mu);
      iso_stress(lam, mu, e6, sg);
    }
    if (gauss_stress)
      for (int i = 0; i < 6; ++i) gauss_stress[6 * pt + i] = sg[i];
  }
  __syncthreads();  // x and u are read: eps takes their place
  if (pt_lane) {
    // the node phase's force is sum_q P d_k, P = det J w sigma J^-1 (as k_elem_forces)
    const double s0 = sc * sg[0], s1 = sc * sg[1], s2 = sc * sg[2];
    const double s3 = sc * sg[3], s4 = sc * sg[4], s5 = sc * sg[5];
    const double S[3][3] = {{s0, s3, s4}, {s3, s1, s5}, {s4, s5, s2}};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        R.P[9 * q + 3 * c + j] = S[c][0] * A[3 * j] + S[c][1] * A[3 * j + 1] + S[c][2] * A[3 * j + 2];
    for (int i = 0; i < 6; ++i) {
      R.eps[6 * q + i] = e6[i];
      R.sig[6 * q + i] = sg[i];
    }
  }
  __syncthreads();
  for (int idx = lane; emean && idx < ne * 12; idx += kWave) {
    // Gauss-point means: sum in point order, times 1/NGP
    const int e = idx / 12, c = idx - e * 12;
    const double* v = c < 6 ? L[e].eps + c : L[e].sig + (c - 6);
    double m = 0.0;
    for (int q = 0; q < NGP; ++q) m += v[6 * q];
    emean[12 * (elem_offset + le0 + e) + c] = m * (1.0 / NGP);
  }
  if (lane < ne * NN) {
    const int e = lane / NN, k = lane - e * NN;
    const RecoverLds<NN, NGP>& R = L[e];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int q = 0; q < NGP; ++q) {
      const double* d = dN + (q * NN + k) * 3;
      const double d0 = d[0], d1 = d[1], d2 = d[2];
      const double* P = R.P + 9 * q;
      f0 += P[0] * d0 + P[1] * d1 + P[2] * d2;
      f1 += P[3] * d0 + P[4] * d1 + P[5] * d2;
      f2 += P[6] * d0 + P[7] * d1 + P[8] * d2;
    }
    // node-major, at the incidence of (element, node), as k_elem_forces
    double* fp = fe + 3 * int64_t(ipos[(le0 + e) * NN + k]);
    fp[0] = f0;
    fp[1] = f1;
    fp[2] = f2;
  }
}

// Element forces alone (the matrix-free internal force: no means, no point stresses, no
// J2). With s = det J w sigma (symmetric 3x3) and g_k = J^-T d_k, the force of node k is
// sum_q s g_k = sum_q P d_k with P = s J^-1 (P_cj = sum_i s_ci (J^-1)_ji), so the point
// phase leaves 9 values per point (P) where J^-1 and s took 15, and the node phase takes 9
// multiply-adds per point where forming g and then s g took 18; P overlays the point
// phase's coordinates and displacements (4.6 KB per wave of 8 C3D8 elements). The force
// of (element, node) goes node-major, to 3 * (its incidence), so that k_node_sum_seq
// reads each node's vectors as one contiguous run.
template <int NN, int NGP>
struct ForceLds {
  union {
    struct {
      double x[NN * 3];
      double u[NN * 3];
    };
    double P[NGP * 9];  // det J w sigma J^-1 per Gauss point (row-major)
  };
};

// The natural-derivative tables of a family of at most 8 points (the point-major table
// and its transpose, 2 NGP NN 3 doubles: 3 KB for C3D8) are staged in the LDS, so that the
// point and node lanes read them there instead of by vector loads from the cache.
template <int NN, int NGP>
constexpr bool forces_dn_lds() { return NGP <= 8 && 2 * NGP * NN * 3 <= 512; }

// Orders this wave's LDS accesses (the wave's lanes exchange data through the LDS; the
// other waves of the workgroup work on other elements and are not waited for).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// WPB waves per workgroup, each on its own EPW elements: the staged natural-derivative
// table is loaded once per workgroup and shared; past that load the waves synchronise
// only with themselves.
template <int NN, int NGP, int WAVES, int WPB>
__global__ __launch_bounds__(kWave * WPB, WAVES) void k_elem_forces(
    const int32_t* __restrict__ conn, int64_t elem_offset, const double* __restrict__ dN_g,
    const double* __restrict__ wq, const double* __restrict__ coords,
    const double* __restrict__ E, const double* __restrict__ nu, const double* __restrict__ u,
    const int32_t* __restrict__ ipos, double* __restrict__ fe, int* __restrict__ flag,
    int64_t n_elem, int64_t slab) {
  constexpr int EPW = rec_epw(NN, NGP);
  constexpr bool HEX8 = NN == 8 && NGP == 8;  // natural derivatives in closed form
  constexpr bool DNL = forces_dn_lds<NN, NGP>() && !HEX8;
  constexpr int NDN = DNL ? 2 * NGP * NN * 3 : 1;
  __shared__ ForceLds<NN, NGP> LW[WPB][EPW];
  __shared__ double dn_lds[NDN];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  const int64_t n_grp = (n_elem + EPW - 1) / EPW;
  const int64_t blk = xcd_index(blockIdx.x, (n_grp + WPB - 1) / WPB, slab);
  if (blk < 0) return;  // the whole workgroup
  const int64_t grp = blk * WPB + wv;
  const int64_t le0 = grp * EPW;
  // a wave past the last group keeps no element but takes part in the workgroup barrier
  const int ne = grp < n_grp ? int(n_elem - le0 < EPW ? n_elem - le0 : EPW) : 0;
  ForceLds<NN, NGP>* L = LW[wv];
  if constexpr (DNL) {
    for (int i = threadIdx.x; i < NDN; i += kWave * WPB) dn_lds[i] = dN_g[i];
    __syncthreads();
  }
  const double* dN = DNL ? dn_lds : dN_g;
  for (int idx = lane; idx < ne * NN; idx += kWave) {
    const int e = idx / NN, k = idx - e * NN;
    const int64_t nd = 3 * int64_t(conn[(le0 + e) * NN + k]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      L[e].x[3 * k + c] = coords[nd + c];
      L[e].u[3 * k + c] = u[nd + c];
    }
  }
  wave_lds_sync();
  static_assert(EPW * NGP <= kWave && EPW * NN <= kWave, "one lane per point and per node");
  const bool pt_lane = lane < ne * NGP;
  const int e = pt_lane ? lane / NGP : 0, q = pt_lane ? lane - e * NGP : 0;
  ForceLds<NN, NGP>& R = L[e];
  double A[9], sg[6], sc = 0.0;
  if (pt_lane) {
    const int64_t eg = elem_offset + le0 + e;
    double dh[24];
    if constexpr (HEX8) hex8_dn_point(q, dh);
    const double* d = HEX8 ? dh : dn_point<NN, NGP>(dN, q);
    constexpr int DS = HEX8 ? 1 : dn_stride<NGP>();
    sc = jacobian_detw<NN, true, DS>(R.x, d, wq[q], A, flag);
    double H[3][3];
    displacement_gradient<NN, DS>(R.u, d, A, H);
    double e6[6], lam, mu;
    voigt_strain(H, e6);
    lame(E[eg], nu[eg], lam, mu);
    iso_stress(lam, mu, e6, sg);
  }
  wave_lds_sync();  // x and u are read: P takes their place
  if (pt_lane) {
    // s in Voigt order (xx yy zz xy xz yz) as a symmetric 3x3, A = J^-1 row-major
    const double s0 = sc * sg[0], s1 = sc * sg[1], s2 = sc * sg[2];
    const double s3 = sc * sg[3], s4 = sc * sg[4], s5 = sc * sg[5];
    const double S[3][3] = {{s0, s3, s4}, {s3, s1, s5}, {s4, s5, s2}};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        R.P[9 * q + 3 * c + j] = S[c][0] * A[3 * j] + S[c][1] * A[3 * j + 1] + S[c][2] * A[3 * j + 2];
  }
  wave_lds_sync();
  if (lane < ne * NN) {
    const int e = lane / NN, k = lane - e * NN;
    const ForceLds<NN, NGP>& R = L[e];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for_node_points<NN, NGP>(dN, k, [&](int q, double d0, double d1, double d2) {
      const double* P = R.P + 9 * q;
      f0 += P[0] * d0 + P[1] * d1 + P[2] * d2;
      f1 += P[3] * d0 + P[4] * d1 + P[5] * d2;
      f2 += P[6] * d0 + P[7] * d1 + P[8] * d2;
    });
    // node-major: at the incidence of (element, node) (ipos: the block's conn offsets)
    double* fp = fe + 3 * int64_t(ipos[(le0 + e) * NN + k]);
    fp[0] = f0;
    fp[1] = f1;
    fp[2] = f2;
  }
}

// Phase 2: 16 lanes per node (4 nodes per wave). fe holds the element forces node-major,
// at 3 * (incidence) + component, so a node's are one contiguous run. Lanes 0-11 sum the
// incident elements' Gauss-point means of one component (strain, then stress), lanes
// 12-14 one component of the force, in ascending element order; the means are divided by
// the number of incident elements.
constexpr int kNodeBatch = 8;

__global__ void k_node_fields(int32_t n_nodes, const int64_t* __restrict__ inc_ptr,
                              const int32_t* __restrict__ inc_elem,
                              const double* __restrict__ emean, const double* __restrict__ fe,
                              const double* __restrict__ f_ext, double* __restrict__ strain,
                              double* __restrict__ stress, double* __restrict__ reaction,
                              double* __restrict__ f_int) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int32_t a = int32_t(t / 16);
  const int c = int(t % 16);
  if (a >= n_nodes || c >= 15) return;
  const int64_t i0 = inc_ptr[a], i1 = inc_ptr[a + 1];
  double v = 0.0;
  // kNodeBatch incidences at a time: their index and value loads issue together, the sum
  // keeps ascending element order
  for (int64_t it = i0; it < i1; it += kNodeBatch) {
    int64_t at[kNodeBatch];
    double x[kNodeBatch];
#pragma unroll
    for (int k = 0; k < kNodeBatch; ++k)
      at[k] = it + k < i1 ? (c < 12 ? 12 * int64_t(inc_elem[it + k]) + c : 3 * (it + k) + (c - 12)) : -1;
#pragma unroll
    for (int k = 0; k < kNodeBatch; ++k) x[k] = at[k] >= 0 ? (c < 12 ? emean : fe)[at[k]] : 0.0;
#pragma unroll
    for (int k = 0; k < kNodeBatch; ++k)
      if (at[k] >= 0) v += x[k];
  }
  if (c < 12) {
    const int64_t cnt = i1 - i0;
    if (cnt > 0) v /= double(cnt);
    double* dst = c < 6 ? strain : stress;
    if (dst) dst[6 * int64_t(a) + (c < 6 ? c : c - 6)] = v;
  } else {
    const int64_t k = 3 * int64_t(a) + (c - 12);
    if (f_int) f_int[k] = v;
    if (reaction) reaction[k] = v - (f_ext ? f_ext[k] : 0.0);
  }
}

// Sums of per-element node vectors over each node's incident elements, in ascending
// element order, for node-major vectors (k_elem_forces, k_vjp): node a's are the run
// 3 * [inc_ptr[a], inc_ptr[a + 1]), so out0[3a + c] = sum_it a0[3 it + c] (and out1 from
// a1). G lanes per node: lanes 0-2 the components of out0, 3-5 those of out1 (G = 8;
// G = 4: out0 alone).
template <int G>
__global__ void k_node_sum_seq(int32_t n_nodes, const int64_t* __restrict__ inc_ptr,
                               const double* __restrict__ a0, const double* __restrict__ a1,
                               double* __restrict__ out0, double* __restrict__ out1) {
  static_assert(G == 4 || G == 8, "4 lanes per node for one array, 8 for two");
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int32_t a = int32_t(t / G);
  const int c = int(t % G);
  if (a >= n_nodes || c >= 6 || (c >= 3 && !a1)) return;
  const double* src = (c < 3 ? a0 : a1) + (c < 3 ? c : c - 3);
  const int64_t i0 = inc_ptr[a], i1 = inc_ptr[a + 1];
  double v = 0.0;
  for (int64_t it = i0; it < i1; it += kNodeBatch) {
    double x[kNodeBatch];
#pragma unroll
    for (int k = 0; k < kNodeBatch; ++k) x[k] = it + k < i1 ? src[3 * (it + k)] : 0.0;
#pragma unroll
    for (int k = 0; k < kNodeBatch; ++k)
      if (it + k < i1) v += x[k];
  }
  (c < 3 ? out0 : out1)[3 * int64_t(a) + (c < 3 ? c : c - 3)] = v;
}


// ---- K6: adjoint of r = K(x, E) u -------------------------------------------
// Per Gauss point, with s = det J * w, sigma = D eps:
//   phi_q = s sigma_u : H_w             (w . r restricted to the element)
//   dphi/dx_m = sum_q P_q g_m,  P_q = phi_q I - s (sigma_u H_w + sigma_w H_u)^T
//   dphi/dE   = phi / E,   dphi/du_m = sum_q s B_m^T sigma_w
// B is recomputed at the Gauss points (never stored), exactly as the forward pass. EPW
// elements per wave as in the recovery kernel: one lane per (element, Gauss point), then
// one per (element, node) writing the element's dx and du vectors to arenas (3 per
// connectivity entry, node-major as k_elem_forces) that k_node_sum_seq adds up per node in
// ascending element order. With g_m = J^-T d_m both node sums are sum_q X_q d_m, X = P_q
// J^-1 for dx and s sigma_w J^-1 for du (as k_elem_forces), so the point phase leaves two
// 3x3 matrices per point where J^-1, P and s sigma_w took 24 values.
template <int NN, int NGP>
struct VjpLds {
  union {
    struct {  // the point phase reads
      double x[NN * 3];
      double u[NN * 3];
      double w[NN * 3];
    };
    struct {  // ... and then leaves for the node phase
      double Q[NGP * 9];  // P J^-1 per Gauss point (row-major)
      double F[NGP * 9];  // s sigma_w J^-1
      double phi[NGP];
    };
  };
};

// WPB waves per workgroup, each on its own EPW elements, sharing the staged table (as
// k_elem_forces; 2 waves keep the residency of single-wave workgroups, whose LDS each
// held a copy of the table).
template <int NN, int NGP, int WPB>
__global__ __launch_bounds__(kWave * WPB) void k_vjp(
    const int32_t* __restrict__ conn, int64_t elem_offset, const double* __restrict__ dN_g,
    const double* __restrict__ wq, const double* __restrict__ coords,
    const double* __restrict__ E, const double* __restrict__ nu, const double* __restrict__ u,
    const double* __restrict__ w, const int32_t* __restrict__ ipos, double* __restrict__ dxe,
    double* __restrict__ dE, double* __restrict__ due, int* __restrict__ flag, int64_t n_elem,
    int64_t slab) {
  constexpr int EPW = rec_epw(NN, NGP);
  static_assert(EPW * NGP <= kWave && EPW * NN <= kWave, "one lane per point and per node");
  // small natural-derivative tables in the LDS, as k_elem_forces (C3D8: closed form)
  constexpr bool HEX8 = NN == 8 && NGP == 8;
  constexpr bool DNL = forces_dn_lds<NN, NGP>() && !HEX8;
  constexpr int NDN = DNL ? 2 * NGP * NN * 3 : 1;
  __shared__ VjpLds<NN, NGP> LW[WPB][EPW];
  __shared__ double dn_lds[NDN];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  const int64_t n_grp = (n_elem + EPW - 1) / EPW;
  const int64_t blk = xcd_index(blockIdx.x, (n_grp + WPB - 1) / WPB, slab);
  if (blk < 0) return;  // the whole workgroup
  const int64_t grp = blk * WPB + wv;
  const int64_t le0 = grp * EPW;
  // a wave past the last group keeps no element but takes part in the workgroup barrier
  const int ne = grp < n_grp ? int(n_elem - le0 < EPW ? n_elem - le0 : EPW) : 0;
  VjpLds<NN, NGP>* L = LW[wv];
  if constexpr (DNL) {
    for (int i = threadIdx.x; i < NDN; i += kWave * WPB) dn_lds[i] = dN_g[i];
    __syncthreads();
  }
  const double* dN = DNL ? dn_lds : dN_g;
  for (int idx = lane; idx < ne * NN; idx += kWave) {
    const int e = idx / NN, k = idx - e * NN;
    const int64_t nd = conn[(le0 + e) * NN + k];
    for (int c = 0; c < 3; ++c) {
      L[e].x[3 * k + c] = coords[3 * nd + c];
      L[e].u[3 * k + c] = u[3 * nd + c];
      L[e].w[3 * k + c] = w[3 * nd + c];
    }
  }
  wave_lds_sync();
  // the point phase's results, held across the barrier after which they overlay x, u, w
  const bool pt_lane = lane < ne * NGP;
  const int pe = pt_lane ? lane / NGP : 0, pq_i = pt_lane ? lane - pe * NGP : 0;
  double A[9], Pm[9], Sw6[6], pq = 0.0;
  if (pt_lane) {
    const int e = pe, q = pq_i;
    VjpLds<NN, NGP>& R = L[e];
    const int64_t ge = elem_offset + le0 + e;
    double dh[24];
    if constexpr (HEX8) hex8_dn_point(q, dh);
    const double* d = HEX8 ? dh : dn_point<NN, NGP>(dN, q);
    constexpr int DS = HEX8 ? 1 : dn_stride<NGP>();
    const double s = jacobian_detw<NN, true, DS>(R.x, d, wq[q], A, flag);
    double Hu[3][3], Hw[3][3];
    displacement_gradient<NN, DS>(R.u, d, A, Hu);
    displacement_gradient<NN, DS>(R.w, d, A, Hw);
    double eu[6], ew[6], su[6], sw[6], lam, mu;
    lame(E[ge], nu[ge], lam, mu);
    voigt_strain(Hu, eu);
    voigt_strain(Hw, ew);
    iso_stress(lam, mu, eu, su);
    iso_stress(lam, mu, ew, sw);
    for (int i = 0; i < 6; ++i) pq += ew[i] * su[i];
    pq *= s;
    const double Su[3][3] = {{su[0], su[3], su[4]}, {su[3], su[1], su[5]}, {su[4], su[5], su[2]}};
    const double Sw[3][3] = {{sw[0], sw[3], sw[4]}, {sw[3], sw[1], sw[5]}, {sw[4], sw[5], sw[2]}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double m = 0.0;
        for (int l = 0; l < 3; ++l) m += Su[i][l] * Hw[l][j] + Sw[i][l] * Hu[l][j];
        // P[j][i] = pq delta_ij - s M[i][j]
        Pm[3 * j + i] = (i == j ? pq : 0.0) - s * m;
      }
    for (int i = 0; i < 6; ++i) Sw6[i] = s * sw[i];
  }
  wave_lds_sync();  // x, u and w are read: the point results take their place
  if (pt_lane) {
    VjpLds<NN, NGP>& R = L[pe];
    const double Sw[3][3] = {{Sw6[0], Sw6[3], Sw6[4]}, {Sw6[3], Sw6[1], Sw6[5]}, {Sw6[4], Sw6[5], Sw6[2]}};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        R.Q[9 * pq_i + 3 * c + j] =
            Pm[3 * c] * A[3 * j] + Pm[3 * c + 1] * A[3 * j + 1] + Pm[3 * c + 2] * A[3 * j + 2];
        R.F[9 * pq_i + 3 * c + j] =
            Sw[c][0] * A[3 * j] + Sw[c][1] * A[3 * j + 1] + Sw[c][2] * A[3 * j + 2];
      }
    R.phi[pq_i] = pq;
  }
  wave_lds_sync();
  if (lane < ne) {
    double phi = 0.0;
    for (int q = 0; q < NGP; ++q) phi += L[lane].phi[q];
    const int64_t ge = elem_offset + le0 + lane;
    dE[ge] = phi / E[ge];
  }
  if (lane < ne * NN) {
    const int e = lane / NN, k = lane - e * NN;
    const VjpLds<NN, NGP>& R = L[e];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for_node_points<NN, NGP>(dN, k, [&](int q, double d0, double d1, double d2) {
      const double* Q = R.Q + 9 * q;
      a0 += Q[0] * d0 + Q[1] * d1 + Q[2] * d2;
      a1 += Q[3] * d0 + Q[4] * d1 + Q[5] * d2;
      a2 += Q[6] * d0 + Q[7] * d1 + Q[8] * d2;
      const double* F = R.F + 9 * q;
      f0 += F[0] * d0 + F[1] * d1 + F[2] * d2;
      f1 += F[3] * d0 + F[4] * d1 + F[5] * d2;
      f2 += F[6] * d0 + F[7] * d1 + F[8] * d2;
    });
    const int64_t o = 3 * int64_t(ipos[(le0 + e) * NN + k]);  // node-major, as k_elem_forces
    dxe[o] = a0;
    dxe[o + 1] = a1;
    dxe[o + 2] = a2;
    due[o] = f0;
    due[o + 1] = f1;
    due[o + 2] = f2;
  }
}

// ---- K8: mass and geometric stiffness into the element arena (element.cpp:786-854) ---
// Both matrices are k_ab I3 blocks: the consistent mass k_ab = sum_q rho detJ w N_a N_b
// (element_mass), the lumped mass its row sums on the diagonal (element_mass_lumped),
// the geometric stiffness k_ab = sum_q detJ w g_a^T sigma_q g_b (element_geometric_
// stiffness with accumulate_kgeo). One element per wave, lanes over the upper node
// pairs. The arena holds only k_ab (n^2 per element); the row gather expands the I3
// blocks, so the entries off the block diagonals never receive a contribution and the
// liveness flags give the reference's key set.
enum : int { kMassConsistent = 0, kMassLumped = 1, kGeometric = 2 };

// LDS of k_elem_scalar, per wave: node coordinates, det J * w per point, for the
// geometric stiffness J^-1 per point and the gradients of QC points at a time; the
// element's n x n scalars. The mass's shape-value table (read 5 times per point and lane:
// from LDS, not through the vector memory path) is one per workgroup.
template <int NN, int NGP, int QC, int KIND>
struct ScalarLds {
  double x[NN * 3];
  double sc[NGP];
  double Ai[KIND == kGeometric ? NGP * 9 : 1];
  double g[KIND == kGeometric ? QC * NN * 3 : 1];
  alignas(16) double k[NN * NN];  // 16-byte aligned: paired reads and writes of the rows
  double row[NN];
};

// Waves per workgroup of k_elem_scalar: the mass kinds share the shape-value table
// between 4 waves (one element each); the geometric stiffness, which has none, runs one.
constexpr int scalar_waves(int kind) { return kind == kGeometric ? 1 : 4; }

// A wave's LDS exchange point: the workgroup barrier where the workgroup is one wave,
// else wave_lds_sync (the other waves work on other elements).
template <int WPB>
__device__ __forceinline__ void wave_exchange() {
  if constexpr (WPB == 1)
    __syncthreads();
  else
    wave_lds_sync();
}

// One element per wave (XCD-aware like k_elem_iso), WPB waves per workgroup. A lane owns
// node a and up to four consecutive upper partners b (decode_group): k_ab = rho sum_q
// det J w N_a N_b (mass) or sum_q det J w g_a^T sigma_q g_b (geometric, with t_a =
// sigma_q g_a formed once per point for the lane's partners); the upper values are
// mirrored in LDS, so k is bitwise symmetric, and written row-major with coalesced
// streaming stores. Past the shared table's load each wave synchronises only with itself.
template <int NN, int NGP, int QC, int KIND>
__global__ __launch_bounds__(kWave * scalar_waves(KIND)) void k_elem_scalar(
    const int32_t* __restrict__ conn, int64_t elem_offset, int64_t gp_offset,
    const double* __restrict__ dN, const double* __restrict__ wq, const double* __restrict__ Nq,
    const double* __restrict__ coords, const double* __restrict__ rho,
    const double* __restrict__ sig, double* __restrict__ kp, int* __restrict__ flag,
    int64_t n_elem, int64_t slab) {
  constexpr int WPB = scalar_waves(KIND);
  __shared__ ScalarLds<NN, NGP, QC, KIND> LW[WPB];
  __shared__ double ntab[KIND == kGeometric ? 1 : NGP * NN];  // shape values N[q][k] (mass)
  // a single wave keeps constant LDS addresses (a base indexed by the wave took the
  // geometric kernel from 80 to 100 VGPRs, 6 to 4 waves per SIMD)
  const int lane = WPB == 1 ? int(threadIdx.x) : int(threadIdx.x % kWave);
  const int wv = WPB == 1 ? 0 : int(threadIdx.x / kWave);
  const int64_t blk = xcd_index(blockIdx.x, (n_elem + WPB - 1) / WPB, slab);
  if (blk < 0) return;  // the whole workgroup
  const int64_t le = blk * WPB + wv;
  ScalarLds<NN, NGP, QC, KIND>& L = LW[wv];
  // (a single-wave workgroup's element exists: xcd_index checked it)
  const bool has_elem = WPB == 1 || le < n_elem;
  if (has_elem && lane < NN) {
    const double* xp = coords + 3 * int64_t(conn[le * NN + lane]);
    L.x[3 * lane + 0] = xp[0];
    L.x[3 * lane + 1] = xp[1];
    L.x[3 * lane + 2] = xp[2];
  }
  if constexpr (KIND != kGeometric) {
    for (int idx = threadIdx.x; idx < NGP * NN; idx += kWave * WPB) ntab[idx] = Nq[idx];
    __syncthreads();
  } else {
    wave_exchange<WPB>();
  }
  if (!has_elem) return;  // a wave past the last element (after the workgroup barrier)
  for (int q = lane; q < NGP; q += kWave)
    L.sc[q] = jacobian_detw<NN, KIND == kGeometric, dn_stride<NGP>()>(L.x, dn_point<NN, NGP>(dN, q), wq[q],
                                                        L.Ai + 9 * q, flag);
  constexpr int NP = NN * (NN + 1) / 2;
  constexpr int PPL = (NP + kWave - 1) / kWave;
  int ga, gb0;
  decode_group<NN, PPL>(lane, ga, gb0);
  const int a = ga >= 0 ? ga : 0;  // clamped: unowned slots compute unused values
  int b[PPL];
  double k[PPL];
#pragma unroll
  for (int t = 0; t < PPL; ++t) {
    b[t] = gb0 + t < NN ? (gb0 + t >= 0 ? gb0 + t : 0) : NN - 1;
    k[t] = 0.0;
  }
  wave_exchange<WPB>();
  if constexpr (KIND != kGeometric) {
    const double r = rho[elem_offset + le];
#pragma unroll 3
    for (int q = 0; q < NGP; ++q) {
      const double* n = ntab + q * NN;
      const double c = r * L.sc[q] * n[a];
#pragma unroll
      for (int t = 0; t < PPL; ++t) k[t] += c * n[b[t]];
    }
  } else {
#pragma unroll 1
    for (int q0 = 0; q0 < NGP; q0 += QC) {
      for (int idx = lane; idx < QC * NN; idx += kWave) {
        const int q = q0 + idx / NN;
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        if (q < NGP) {
          const double* A = L.Ai + 9 * q;
          const double* d = dN + (q * NN + idx % NN) * 3;
          g0 = A[0] * d[0] + A[3] * d[1] + A[6] * d[2];
          g1 = A[1] * d[0] + A[4] * d[1] + A[7] * d[2];
          g2 = A[2] * d[0] + A[5] * d[1] + A[8] * d[2];
        }
        L.g[3 * idx + 0] = g0;
        L.g[3 * idx + 1] = g1;
        L.g[3 * idx + 2] = g2;
      }
      wave_exchange<WPB>();
      const int qn = NGP - q0 < QC ? NGP - q0 : QC;
#pragma unroll 1
      for (int qq = 0; qq < qn; ++qq) {
        const int q = q0 + qq;
        const double* s = sig + 6 * (gp_offset + le * NGP + q);  // Voigt xx yy zz xy xz yz
        const double s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3], s4 = s[4], s5 = s[5];
        const double* gA = L.g + (qq * NN + a) * 3;
        const double sc = L.sc[q];
        const double t0 = sc * (s0 * gA[0] + s3 * gA[1] + s4 * gA[2]);
        const double t1 = sc * (s3 * gA[0] + s1 * gA[1] + s5 * gA[2]);
        const double t2 = sc * (s4 * gA[0] + s5 * gA[1] + s2 * gA[2]);
#pragma unroll
        for (int t = 0; t < PPL; ++t) {
          const double* gB = L.g + (qq * NN + b[t]) * 3;
          k[t] += t0 * gB[0] + t1 * gB[1] + t2 * gB[2];
        }
      }
      wave_exchange<WPB>();
    }
  }
#pragma unroll
  for (int t = 0; t < PPL; ++t)
    if (ga >= 0 && gb0 + t < NN) {
      L.k[a * NN + gb0 + t] = k[t];
      L.k[(gb0 + t) * NN + a] = k[t];
    }
  wave_exchange<WPB>();
  if constexpr (KIND == kMassLumped) {
    for (int r = lane; r < NN; r += kWave) {
      double sum = 0.0;
      for (int c = 0; c < NN; ++c) sum += L.k[r * NN + c];
      L.row[r] = sum;
    }
    wave_exchange<WPB>();
  }
  // the element's n x n scalars, row-major, at a ninth of its full-matrix arena offset
  double* ks = kp + le * NN * NN;
  for (int idx = lane; idx < NN * NN; idx += kWave) {
    double v;
    if constexpr (KIND == kMassLumped) {
      const int r = idx / NN;
      v = idx == r * NN + r ? L.row[r] : 0.0;
    } else {
      v = L.k[idx];
    }
    __builtin_nontemporal_store(v, ks + idx);
  }
}

// ---- host-side dispatch -------------------------------------------------------
// Work items per XCD slab (0: no remap) and the grid that covers them.
int64_t xcd_slab(const cxg_context* ctx, int64_t n) { return ctx->xcd_remap ? (n + 7) / 8 : 0; }
unsigned xcd_grid(const cxg_context* ctx, int64_t n) {
  const int64_t s = xcd_slab(ctx, n);
  return unsigned(s > 0 ? 8 * s : (n > 0 ? n : 1));
}

// Phase B over every node of the pattern's mesh (see k_row_gather).
// up: the arena holds upper blocks (asm_variant 7), nodes in the pattern's locality order.
// The row gather over n_pos gather positions (descriptors desc[0..n_pos)) on `stream`,
// reading element matrices at the records' offsets into kp.
void launch_gather_on(cxg_context* ctx, const cxg_pattern* p, hipStream_t stream,
                      const GatherDesc* desc, int32_t n_pos, const IncRecord* rec,
                      const double* kp, const double* u, const double* fe, double* vals,
                      double* f_int, uint8_t* live, int n_cases, bool scalar, bool up,
                      int hf = 0, bool sd = false) {
  if (n_pos <= 0) return;
  const cxg_mesh* m = p->mesh;
  const int64_t n_dof = 3 * int64_t(m->n_nodes);
  const int32_t slab = int32_t(xcd_slab(ctx, n_pos));
  const unsigned grid = xcd_grid(ctx, n_pos);
  const int32_t cap = int32_t(std::min<int64_t>((std::max<int64_t>(p->max_row, 1) + 7) / 8 * 8,
                                                std::min(kRowCap, ctx->gather_lds_cols)));
  // same-direction rows: one value row and a flag byte per position (row_gather_sd)
  const size_t lds = sd ? size_t(cap) * sizeof(double) + size_t(cap + 3) / 4 * 4
                        : size_t(3 * cap) * sizeof(double) + (live ? size_t(3 * cap) : 0);
#define GATHER(NC, LV, ...)                                                                     \
  hipLaunchKernelGGL((k_row_gather<NC, LV, ##__VA_ARGS__>), dim3(grid), dim3(kWave), lds, stream, \
                     desc, rec, m->conn.p, p->slot_all.p, kp, u, n_dof, fe, vals, f_int, live,  \
                     n_pos, slab, cap, up ? p->ue_arena.p : nullptr,                            \
                     int64_t(3) * int64_t(m->conn.n), int32_t(ctx->up_min_nodes),               \
                     sd ? p->sd_row_ptr.p : nullptr, p->dof_eq.p)
  if (sd) {
    // same-direction rows: every row must fit the LDS (checked by the caller)
    if (!scalar || !live || up || p->max_row > cap)
      throw Error(CXG_ERR_UNSUPPORTED, "same-direction rows: scalar gathers of rows held in LDS");
    GATHER(1, true, true, false, 0, true);
  } else if (hf > 0) {
    // gradient-form arena (k_elem_hform): linear stiffness, with or without f_int
    if (!up || scalar || fe || hf != 8)
      throw Error(CXG_ERR_UNSUPPORTED, "gradient-form arena: linear stiffness of 8-point elements only");
    if (n_cases == 4) {
      if (live) GATHER(4, true, false, true, 8); else GATHER(4, false, false, true, 8);
    } else if (n_cases == 2) {
      if (live) GATHER(2, true, false, true, 8); else GATHER(2, false, false, true, 8);
    } else {
      if (live) GATHER(1, true, false, true, 8); else GATHER(1, false, false, true, 8);
    }
  } else if (up) {
    if (scalar) throw Error(CXG_ERR_UNSUPPORTED, "upper-block arena: stiffness gathers only");
    if (fe) {
      if (live) GATHER(0, true, false, true); else GATHER(0, false, false, true);
    } else if (n_cases == 4) {
      if (live) GATHER(4, true, false, true); else GATHER(4, false, false, true);
    } else if (n_cases == 2) {
      if (live) GATHER(2, true, false, true); else GATHER(2, false, false, true);
    } else {
      if (live) GATHER(1, true, false, true); else GATHER(1, false, false, true);
    }
  } else if (scalar) {
    if (live) GATHER(1, true, true); else GATHER(1, false, true);
  } else if (fe) {
    if (live) GATHER(0, true); else GATHER(0, false);
  } else if (n_cases == 4) {
    if (live) GATHER(4, true); else GATHER(4, false);
  } else if (n_cases == 2) {
    if (live) GATHER(2, true); else GATHER(2, false);
  } else {
    if (live) GATHER(1, true); else GATHER(1, false);
  }
#undef GATHER
}

void launch_gather(cxg_context* ctx, const cxg_pattern* p, const double* u, const double* fe,
                   double* vals, double* f_int, uint8_t* live, int n_cases, bool scalar = false,
                   bool up = false, int hf = 0, bool sd = false) {
  prepare_desc(ctx, p);
  launch_gather_on(ctx, p, ctx->stream, p->gdesc.p, p->mesh->n_nodes, p->inc_rec.p,
                   p->ke_arena.p, u, fe, vals, f_int, live, n_cases, scalar, up, hf, sd);
}

// Resident workgroups of a persistent launch: waves per CU times CUs, at most one per
// work item.
unsigned persistent_grid(const cxg_context* ctx, const cxg_pattern* p) {
  const int64_t g = std::min<int64_t>(int64_t(ctx->n_cu) * ctx->persistent_waves, p->n_work);
  return unsigned(g > 0 ? g : 1);
}

inline int grid_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return int(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

const char* type_name(int t) {
  switch (t) {
    case CXG_C3D8: return "C3D8";
    case CXG_C3D20: return "C3D20";
    case CXG_C3D20R: return "C3D20R";
    case CXG_C3D6: return "C3D6";
    default: return "other";
  }
}

// One launch per (colour, block) in colour order, one workgroup per patch; each
// launch is timed as "<what>_<family>" when the context's timing is on.
struct LaunchSet {
  size_t bi;
  const cxg_block_dev& B;
  const int32_t* elist;
  const int32_t* pptr;
  unsigned patches;
};

template <class F>
void for_each_colour_block(cxg_context* ctx, const char* what, const cxg_mesh* m, F&& f) {
  for (int c = 0; c < m->n_colors; ++c)
    for (size_t bi = 0; bi < m->blocks.size(); ++bi) {
      const cxg_block_dev& B = m->blocks[bi];
      if (c < int(B.color_count.size()) && B.color_count[size_t(c)] > 0) {
        LaunchTimer timer(ctx, std::string(what) + "_" + type_name(B.type), B.color_count[size_t(c)]);
        f(LaunchSet{bi, B, B.color[size_t(c)].p, B.patch_ptr[size_t(c)].p,
                    unsigned(B.patch_count[size_t(c)])});
        CXG_HIP(hipGetLastError());
      }
    }
}

#define CXG_DISPATCH(TYPE, MACRO)                                                   \
  switch (TYPE) {                                                                   \
    case CXG_C3D8: MACRO(8, 8, false, 8); break;                                    \
    case CXG_C3D20: MACRO(20, 27, true, 3); break;                                  \
    case CXG_C3D20R: MACRO(20, 8, true, 3); break;                                  \
    case CXG_C3D6: MACRO(6, 2, false, 2); break;                                    \
    default: throw Error(CXG_ERR_UNSUPPORTED, "element type outside the GPU path"); \
  }

}  // namespace

void fill_f64(hipStream_t s, double* p, int64_t n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
  CXG_HIP(hipGetLastError());
}

// asm_variant 8 (see StagePlan): per stage s, on the context's stream A the element
// matrices of the stage's elements into ring slot s % R (cached stores), then on the
// plan's stream B the rows of the nodes that stage completes. A stage's elements wait
// for the gathers that still read the slot they overwrite (those of stages up to
// s - R + W - 1); each gather waits for its stage's elements; A finally waits for B.
void launch_staged_linear(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                          const double* E, const double* nu, double* vals, const double* u,
                          double* f_int, uint8_t* live, int n_cases) {
  const cxg_mesh* m = p->mesh;
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  StagePlan& P = *pm->stages;
  double* ue = nullptr;
  if (u && f_int) {
    prepare_displacement_arena(p, n_cases);
    ue = pm->ue_arena.p;
  }
  const int64_t n_dof = 3 * int64_t(m->n_nodes);
  const int64_t ue_stride = int64_t(3) * int64_t(m->conn.n);
  const bool two = ctx->stage_overlap != 0;
  hipStream_t A = ctx->stream, B = two ? P.stream : ctx->stream;
  LaunchTimer total(ctx, "staged_assembly", m->n_elem);
  for (int s = 0; s < P.S; ++s) {
    if (two && s >= P.R) CXG_HIP(hipStreamWaitEvent(A, P.ev_gather[size_t(s - P.R + P.W - 1)], 0));
    const int64_t e0 = P.elem_first[size_t(s)], e1 = P.elem_first[size_t(s) + 1];
    double* slot = P.ring.p + int64_t(s % P.R) * P.slot_cap;
    for (size_t bi = 0; bi < m->blocks.size(); ++bi) {
      const cxg_block_dev& Bk = m->blocks[bi];
      const int64_t lo = std::max(e0, Bk.elem_offset), hi = std::min(e1, Bk.elem_offset + Bk.n_elem);
      if (lo >= hi) continue;
      const int64_t first = lo - Bk.elem_offset, cnt = hi - lo;
      const ElemTable& T = ctx->tables[Bk.type];
      const int32_t* conn = m->conn.p + Bk.conn_offset + first * Bk.nn;
      double* kp = slot + (P.kp_host[size_t(lo)] - P.kp_host[size_t(e0)]);
      double* ueb = ue ? ue + 3 * (Bk.conn_offset + first * Bk.nn) : nullptr;
      LaunchTimer timer(ctx, std::string("elem_iso_staged_") + type_name(Bk.type), cnt, A);
#define ISO(NN, NGP, QC, TLV)                                                                   \
  hipLaunchKernelGGL((k_elem_iso<NN, NGP, QC, 4, true, TLV, false>), dim3(xcd_grid(ctx, cnt)), \
                     dim3(kWave), 0, A, conn, lo, T.dN, T.w, coords, E, nu, kp, ctx->flag_dev, \
                     cnt, xcd_slab(ctx, cnt), u, n_dof, n_cases, ueb, ue_stride)
#define LAUNCH(NN, NGP, UPPER, QC)                                                              \
  if constexpr (NN == 20) {                                                                     \
    if (ctx->elem_tiling == 1) ISO(NN, NGP, QC, 1); else ISO(NN, NGP, QC, 0);                   \
  } else {                                                                                      \
    ISO(NN, NGP, QC, 0);                                                                        \
  }
      CXG_DISPATCH(Bk.type, LAUNCH)
#undef LAUNCH
#undef ISO
      CXG_HIP(hipGetLastError());
    }
    if (two) {
      CXG_HIP(hipEventRecord(P.ev_elem[size_t(s)], A));
      CXG_HIP(hipStreamWaitEvent(B, P.ev_elem[size_t(s)], 0));
    }
    const int64_t n0 = P.node_off[size_t(s)], n1 = P.node_off[size_t(s) + 1];
    if (n1 > n0) {
      LaunchTimer timer(ctx, "row_gather_staged", n1 - n0, B);
      launch_gather_on(ctx, p, B, P.desc.p + n0, int32_t(n1 - n0), P.rec.p, P.ring.p,
                       f_int ? u : nullptr, nullptr, vals, f_int, live, n_cases, false, true);
      CXG_HIP(hipGetLastError());
    }
    if (two) CXG_HIP(hipEventRecord(P.ev_gather[size_t(s)], B));
  }
  if (two) CXG_HIP(hipStreamWaitEvent(A, P.ev_gather[size_t(P.S - 1)], 0));
}

// Whether this assembly takes the staged path: asm_variant 8, every family in the
// upper-block layout, not under capture (graphs replay the one-pass kernels), and a
// stage plan that pays for the mesh.
bool use_stages(cxg_context* ctx, const cxg_pattern* p, const double* coords) {
  if (ctx->asm_variant != 8) return false;
  for (const cxg_block_dev& B : p->mesh->blocks)
    if (B.n_elem > 0 && B.nn < ctx->up_min_nodes) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CXG_HIP(hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone) return false;
  prepare_order(ctx, p, coords);
  return prepare_stages(ctx, p);
}

void launch_assemble_linear(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                            const double* E, const double* nu, double* vals, const double* u,
                            double* f_int, uint8_t* live, int n_cases) {
  const cxg_mesh* m = p->mesh;
  if (n_cases != 1 && (ctx->asm_variant < 4 || (n_cases != 2 && n_cases != 4)))
    throw Error(CXG_ERR_UNSUPPORTED, "load cases: 1, or 2 or 4 with the two-phase assembly");
  if (m->n_nodes > 0 && use_stages(ctx, p, coords)) {
    launch_staged_linear(ctx, p, coords, E, nu, vals, u, f_int, live, n_cases);
    return;
  }
  if (ctx->asm_variant >= 4) {
    // Two-phase: element pair blocks to the arena, then one owner wave per node row.
    cxg_pattern* pm = const_cast<cxg_pattern*>(p);
    prepare_arenas(p, false);
    if (ctx->asm_variant == 5) {
      prepare_work(ctx, p);
      hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, ctx->stream, pm->queue.p);
      LaunchTimer timer(ctx, "elem_iso_persistent", m->n_elem);
      hipLaunchKernelGGL(k_elem_iso_persistent, dim3(persistent_grid(ctx, p)), dim3(kWave), 0,
                         ctx->stream, pm->work.p, pm->n_work, pm->queue.p, pm->bdesc.p, m->conn.p,
                         coords, E, nu, pm->ke_arena.p, ctx->flag_dev);
      CXG_HIP(hipGetLastError());
    }
    for (size_t bi = 0; ctx->asm_variant == 6 && bi < m->blocks.size(); ++bi) {
      const cxg_block_dev& B = m->blocks[bi];
      if (B.n_elem == 0) continue;
      const ElemTable& T = ctx->tables[B.type];
      const int32_t* conn = m->conn.p + B.conn_offset;
      double* kp = pm->ke_arena.p + p->kp_base[bi];
      LaunchTimer timer(ctx, std::string("elem_iso_mfma_") + type_name(B.type), B.n_elem);
#define LAUNCH(NN, NGP, UP, QC)                                                                 \
  hipLaunchKernelGGL((k_elem_iso_mfma<NN, NGP>), dim3(unsigned(B.n_elem)), dim3(kWave), 0,      \
                     ctx->stream, conn, B.elem_offset, T.dN, T.w, coords, E, nu, kp,            \
                     ctx->flag_dev)
      CXG_DISPATCH(B.type, LAUNCH)
#undef LAUNCH
      CXG_HIP(hipGetLastError());
    }
    const int hs = hform_stride(ctx, m);
    const bool up = ctx->asm_variant >= 7 && (hs > 0 || uses_upper_blocks(ctx, m));
    if (up) prepare_order(ctx, p, coords);
    // variant 7 with f_int: the element kernels copy u per element for the gather
    double* ue = nullptr;
    if (up && u && f_int) {
      prepare_displacement_arena(p, n_cases);
      ue = pm->ue_arena.p;
    }
    const int64_t n_dof = 3 * int64_t(m->n_nodes);
    const int64_t ue_stride = int64_t(3) * int64_t(m->conn.n);
    for (size_t bi = 0; (ctx->asm_variant == 4 || ctx->asm_variant >= 7) && bi < m->blocks.size(); ++bi) {
      const cxg_block_dev& B = m->blocks[bi];
      if (B.n_elem == 0) continue;
      const ElemTable& T = ctx->tables[B.type];
      const int32_t* conn = m->conn.p + B.conn_offset;
      double* kp = pm->ke_arena.p + p->kp_base[bi];
      if (hs > 0 && B.type == CXG_C3D20R) {
        LaunchTimer timer(ctx, "elem_hform_C3D20R", B.n_elem);
        constexpr int EPW = kWave / 8;
        const int64_t groups = (B.n_elem + EPW - 1) / EPW;
        hipLaunchKernelGGL((k_elem_hform<20, 8>), dim3(xcd_grid(ctx, groups)), dim3(kWave), 0,
                           ctx->stream, conn, B.elem_offset, T.dN, T.w, coords, E, nu, kp,
                           ctx->flag_dev, B.n_elem, xcd_slab(ctx, groups), u, n_dof, n_cases,
                           ue ? ue + 3 * B.conn_offset : nullptr, ue_stride);
        CXG_HIP(hipGetLastError());
        continue;
      }
      const bool bup = up && B.nn >= ctx->up_min_nodes;  // this family's arena layout
      LaunchTimer timer(ctx, std::string(bup ? "elem_iso_up_" : "elem_iso_") + type_name(B.type), B.n_elem);
#define ISO(NN, NGP, QC, UPV, TLV)                                                              \
  hipLaunchKernelGGL((k_elem_iso<NN, NGP, QC, 4, UPV, TLV>), dim3(xcd_grid(ctx, B.n_elem)),    \
                     dim3(kWave), 0, ctx->stream, conn, B.elem_offset, T.dN, T.w, coords, E,    \
                     nu, kp, ctx->flag_dev, B.n_elem, xcd_slab(ctx, B.n_elem), u, n_dof,        \
                     n_cases, ue ? ue + 3 * B.conn_offset : nullptr, ue_stride)
#define MULTI(NN, NGP, QC)                                                                      \
  hipLaunchKernelGGL((k_elem_iso_multi<NN, NGP, QC>),                                            \
                     dim3(xcd_grid(ctx, (B.n_elem + iso_multi_epw(NGP) - 1) / iso_multi_epw(NGP))), \
                     dim3(kWave), 0, ctx->stream, conn, B.elem_offset, T.dN, T.w, coords, E,    \
                     nu, kp, ctx->flag_dev, B.n_elem,                                           \
                     xcd_slab(ctx, (B.n_elem + iso_multi_epw(NGP) - 1) / iso_multi_epw(NGP)),  \
                     u, n_dof, n_cases, ue ? ue + 3 * B.conn_offset : nullptr, ue_stride)
#define LAUNCH(NN, NGP, UPPER, QC)                                                              \
  if constexpr (NN == 20) {                                                                     \
    if (ctx->elem_tiling == 1) {                                                                \
      if (bup) ISO(NN, NGP, QC, true, 1); else ISO(NN, NGP, QC, false, 1);                     \
    } else {                                                                                    \
      if (bup) ISO(NN, NGP, QC, true, 0); else ISO(NN, NGP, QC, false, 0);                     \
    }                                                                                           \
  } else if constexpr (NN <= 8 && NGP <= QC) {                                                  \
    if (bup) MULTI(NN, NGP, QC); else ISO(NN, NGP, QC, false, 0);                              \
  } else {                                                                                      \
    if (bup) ISO(NN, NGP, QC, true, 0); else ISO(NN, NGP, QC, false, 0);                       \
  }
      CXG_DISPATCH(B.type, LAUNCH)
#undef LAUNCH
#undef MULTI
#undef ISO
      CXG_HIP(hipGetLastError());
    }
    if (m->n_nodes > 0) {
      LaunchTimer timer(ctx, hs > 0 ? "row_gather_hform" : (up ? "row_gather_up" : "row_gather"), m->n_elem);
      launch_gather(ctx, p, f_int ? u : nullptr, nullptr, vals, f_int, live, n_cases, false, up, hs);
      CXG_HIP(hipGetLastError());
    }
    return;
  }
  fill_f64(ctx->stream, vals, p->nnz, 0.0);
  if (f_int) fill_f64(ctx->stream, f_int, int64_t(m->n_nodes) * 3, 0.0);
  const bool v2 = ctx->asm_variant >= 2;
  for_each_colour_block(ctx, v2 ? "assemble_iso2" : "assemble_iso", m, [&](const LaunchSet& S) {
    const ElemTable& T = ctx->tables[S.B.type];
    const int32_t* conn = m->conn.p + S.B.conn_offset;
#define LAUNCH(NN, NGP, UP, QC)                                                                 \
  if (ctx->asm_variant == 2)                                                                    \
    hipLaunchKernelGGL((k_assemble_iso2<NN, NGP, QC, 4>), dim3(S.patches), dim3(kWave), 0,      \
                       ctx->stream, conn, S.elist, S.pptr, S.B.elem_offset, T.dN, T.w, coords,  \
                       E, nu, p->dof_eq.p, p->row_ptr.p, (p->slot_all.p + p->slot_base[S.bi]), vals, u, f_int,       \
                       ctx->flag_dev);                                                          \
  else if (ctx->asm_variant == 3)                                                               \
    hipLaunchKernelGGL((k_assemble_iso2<NN, NGP, QC, 3>), dim3(S.patches), dim3(kWave), 0,      \
                       ctx->stream, conn, S.elist, S.pptr, S.B.elem_offset, T.dN, T.w, coords,  \
                       E, nu, p->dof_eq.p, p->row_ptr.p, (p->slot_all.p + p->slot_base[S.bi]), vals, u, f_int,       \
                       ctx->flag_dev);                                                          \
  else                                                                                          \
    hipLaunchKernelGGL((k_assemble_iso<NN, NGP, UP>), dim3(S.patches), dim3(kWave), 0,          \
                       ctx->stream, conn, S.elist, S.pptr, S.B.elem_offset, T.dN, T.w, coords,  \
                       E, nu, p->dof_eq.p, p->row_ptr.p, (p->slot_all.p + p->slot_base[S.bi]), vals, u, f_int,       \
                       ctx->flag_dev)
    CXG_DISPATCH(S.B.type, LAUNCH)
#undef LAUNCH
  });
}

// Waves per workgroup of k_elem_forces and k_vjp, and the workgroups for n elements of
// epw per wave.
constexpr int kForceWaves = 4;
constexpr int kVjpWaves = 2;
int64_t wave_blocks(int64_t n_elem, int epw, int wpb) {
  return ((n_elem + epw - 1) / epw + wpb - 1) / wpb;
}

// Matrix-free f_int = sum_e Ke ue (stress.cpp:125-145): the element forces
// sum_q det J w B^T sigma of k_gauss_fields (means off) into the element-force arena,
// then summed per node in ascending element order.
void launch_internal_force(cxg_context* ctx, const cxg_mesh* m, const double* coords,
                           const double* E, const double* nu, const double* u, double* f_int,
                           double* arena) {
  if (!arena) {
    m->rec_fe.ensure(std::max<size_t>(3 * m->conn.n, 1));
    arena = m->rec_fe.p;
  }
  for (const cxg_block_dev& B : m->blocks) {
    if (B.n_elem == 0) continue;
    const ElemTable& T = ctx->tables[B.type];
    const int32_t* conn = m->conn.p + B.conn_offset;
    LaunchTimer timer(ctx, std::string("internal_force_") + type_name(B.type), B.n_elem);
#define LAUNCH(NN, NGP, UP, QC)                                                                 \
  hipLaunchKernelGGL((k_elem_forces<NN, NGP, 4, kForceWaves>),                                  \
                     dim3(xcd_grid(ctx, wave_blocks(B.n_elem, rec_epw(NN, NGP), kForceWaves))), \
                     dim3(kWave * kForceWaves), 0, ctx->stream, conn, B.elem_offset, T.dN, T.w, \
                     coords, E, nu, u, m->inc_pos.p + B.conn_offset, arena, ctx->flag_dev,      \
                     B.n_elem, xcd_slab(ctx, wave_blocks(B.n_elem, rec_epw(NN, NGP), kForceWaves)))
    CXG_DISPATCH(B.type, LAUNCH)
#undef LAUNCH
    CXG_HIP(hipGetLastError());
  }
  if (m->n_nodes > 0) {
    LaunchTimer timer(ctx, "internal_force_nodes", m->n_nodes);
    hipLaunchKernelGGL(k_node_sum_seq<4>, dim3(unsigned((int64_t(m->n_nodes) * 4 + 255) / 256)),
                       dim3(256), 0, ctx->stream, m->n_nodes, m->inc_ptr.p, arena, nullptr, f_int,
                       nullptr);
    CXG_HIP(hipGetLastError());
  }
}

void launch_internal_force_noalloc(cxg_context* ctx, const cxg_mesh* m, const double* coords,
                                   const double* E, const double* nu, const double* u,
                                   double* f_int) {
  const int64_t need = 3 * int64_t(m->conn.n);
  if (need > kScratchDoubles)
    throw Error(CXG_ERR_ARG, "cxg_internal_force_noalloc: the mesh needs " + std::to_string(need) +
                                 " doubles of scratch (3 per connectivity entry), the maximum is " +
                                 std::to_string(kScratchDoubles));
  if (!ctx->scratch) throw Error(CXG_ERR_STATE, "no scratch address on this context");
  launch_internal_force(ctx, m, coords, E, nu, u, f_int, ctx->scratch);
}

void launch_reduce_free(cxg_context* ctx, int32_t n_nodes, const int32_t* dof_eq,
                        const double* f_ext, const double* f_c, double* rhs) {
  const int64_t n = int64_t(n_nodes) * 3;
  hipLaunchKernelGGL(k_reduce_free, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, dof_eq,
                     f_ext, f_c, rhs);
  CXG_HIP(hipGetLastError());
}

static J2Dev prepare_j2(const cxg_j2_material& p) {
  if (p.n_table < 1 || p.n_table > CXG_MAX_TABLE)
    throw Error(CXG_ERR_ARG, "J2PlasticMaterial needs a non-empty *PLASTIC table");
  J2Dev d{};
  d.E = p.E;
  d.nu = p.nu;
  d.mu = p.E / (2.0 * (1.0 + p.nu));
  d.hkin = p.kinematic_modulus;
  d.n = p.n_table;
  for (int i = 0; i < p.n_table; ++i) {
    d.ep[i] = p.eqplastic[i];
    d.sy[i] = p.yield[i];
  }
  if (p.hardening == CXG_HARD_KINEMATIC || p.hardening == CXG_HARD_COMBINED) {
    if (d.hkin == 0.0 && d.n >= 2 && d.ep[1] > d.ep[0])
      d.hkin = (d.sy[1] - d.sy[0]) / (d.ep[1] - d.ep[0]);
    if (p.hardening == CXG_HARD_KINEMATIC)
      for (int i = 1; i < d.n; ++i) d.sy[i] = d.sy[0];
  }
  return d;
}

std::vector<unsigned char> j2_table_bytes(const cxg_j2_material* mats, int32_t n_mats) {
  std::vector<J2Dev> h(size_t(n_mats > 0 ? n_mats : 0));
  for (int32_t i = 0; i < n_mats; ++i) h[size_t(i)] = prepare_j2(mats[i]);
  const auto* b = reinterpret_cast<const unsigned char*>(h.data());
  return std::vector<unsigned char>(b, b + h.size() * sizeof(J2Dev));
}

const void* upload_j2_table(cxg_context* ctx, const cxg_j2_material* mats, int32_t n_mats) {
  // The device copy of the material table is re-uploaded only when it changes, so
  // the steady state of a Newton loop enqueues kernels only.
  const std::vector<unsigned char> h = j2_table_bytes(mats, n_mats);
  const bool same = h.size() == ctx->j2_host.size() && ctx->j2_dev.p &&
                    std::memcmp(h.data(), ctx->j2_host.data(), h.size()) == 0;
  if (!same && !h.empty()) {
    CXG_HIP(hipStreamSynchronize(ctx->stream));  // earlier kernels may still read the old table
    ctx->j2_dev.alloc(h.size());
    CXG_HIP(hipMemcpy(ctx->j2_dev.p, h.data(), h.size(), hipMemcpyHostToDevice));
    ctx->j2_host = h;
  }
  return h.empty() ? nullptr : ctx->j2_dev.p;
}

bool uses_upper_blocks(const cxg_context* ctx, const cxg_mesh* m) {
  for (const cxg_block_dev& B : m->blocks)
    if (B.n_elem > 0 && B.nn >= ctx->up_min_nodes) return true;
  return false;
}

int hform_stride(const cxg_context* ctx, const cxg_mesh* m) {
  if (ctx->asm_variant != 7 || !ctx->gradient_form) return 0;
  bool c3d20r = false;
  for (const cxg_block_dev& B : m->blocks) {
    if (B.n_elem == 0) continue;
    if (B.type == CXG_C3D20) return 0;  // 20 nodes too: the gather tells forms apart by n
    c3d20r = c3d20r || B.type == CXG_C3D20R;
  }
  return c3d20r ? gauss_points(CXG_C3D20R) : 0;
}

bool linear_upper_path(const cxg_context* ctx, const cxg_mesh* m) {
  return ctx->asm_variant >= 7 && (hform_stride(ctx, m) > 0 || uses_upper_blocks(ctx, m));
}

void prepare_displacement_arena(const cxg_pattern* p, int n_cases) {
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  const size_t n = size_t(n_cases) * 3 * p->mesh->conn.n;
  if (pm->ue_arena.n < n) {
    guard_realloc(p, pm->ue_arena.p != nullptr, "displacement arena (more load cases)");
    pm->ue_arena.alloc(n);
  }
}

void prepare_arenas(const cxg_pattern* p, bool element_forces) {
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  if (pm->ke_arena.n < size_t(p->kp_total)) {
    guard_realloc(p, pm->ke_arena.p != nullptr, "element-matrix arena");
    pm->ke_arena.alloc(size_t(p->kp_total));
  }
  const size_t fe_n = size_t(3 * p->mesh->conn.n);
  if (element_forces && pm->fe_arena.n < fe_n) {
    guard_realloc(p, pm->fe_arena.p != nullptr, "element-force arena");
    pm->fe_arena.alloc(fe_n);
  }
}

// Columns of tile x tile elements over the two shorter extents of the nodes' bounding
// box, each swept along the longest extent in layers a quarter element thick: the nodes
// of one element, which read each other's upper blocks, come close together in the
// order, so the second read of a block finds it in L2. The element size is estimated as
// the edge of the bounding box volume shared out over the elements. A mesh of several
// bodies (connected components of the node graph, such as the independent meshes of a
// batch, which may share one region of space) is ordered body by body, in the order of
// each body's smallest node, each body by its own box and element size. Any order gives
// the same values (each node's rows are summed by one wave in element order).
void prepare_order(cxg_context* ctx, const cxg_pattern* p, const double* coords) {
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  const int tile = ctx->gather_tile;
  if (pm->gorder_tile == tile) return;
  guard_realloc(p, pm->gorder.p != nullptr, "gather order (option gather_tile changed)");
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CXG_HIP(hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone)
    throw Error(CXG_ERR_STATE, "the gather order is built before capture (assemble once first)");
  const cxg_mesh* m = p->mesh;
  const int32_t n = m->n_nodes;
  if (tile <= 0 || n == 0) {
    pm->gorder.reset();
    pm->gorder_tile = tile;
    return;
  }
  std::vector<double> x(size_t(3) * size_t(n));
  std::vector<int32_t> conn(size_t(m->conn.n));
  CXG_HIP(hipStreamSynchronize(ctx->stream));
  CXG_HIP(hipMemcpy(x.data(), coords, x.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (!conn.empty())
    CXG_HIP(hipMemcpy(conn.data(), m->conn.p, conn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));

  // Components: union-find with the smaller root kept, so each root is its component's
  // smallest node; numbered in ascending root order. Nodes that no element lists form
  // one last component together (not one each).
  std::vector<int32_t> comp(static_cast<size_t>(n));
  std::vector<uint8_t> used(static_cast<size_t>(n), 0);
  for (const int32_t v : conn) used[size_t(v)] = 1;
  for (int32_t i = 0; i < n; ++i) comp[size_t(i)] = i;
  const auto find = [&](int32_t a) {
    while (comp[size_t(a)] != a) {
      comp[size_t(a)] = comp[size_t(comp[size_t(a)])];
      a = comp[size_t(a)];
    }
    return a;
  };
  for (const cxg_block_dev& B : m->blocks)
    for (int64_t e = 0; e < B.n_elem; ++e) {
      const int32_t* c = conn.data() + B.conn_offset + e * B.nn;
      int32_t r0 = find(c[0]);
      for (int k = 1; k < B.nn; ++k) {
        const int32_t r = find(c[k]);
        if (r < r0) {
          comp[size_t(r0)] = r;
          r0 = r;
        } else if (r > r0) {
          comp[size_t(r)] = r0;
        }
      }
    }
  for (int32_t i = 0; i < n; ++i) comp[size_t(i)] = find(i);  // every node to its root
  int32_t n_comp = 0;
  for (int32_t i = 0; i < n; ++i) {  // a root precedes its nodes, so comp[root] is its number
    if (!used[size_t(i)]) continue;  // (an unused node is its own root: no element joins it)
    comp[size_t(i)] = comp[size_t(i)] == i ? n_comp++ : comp[size_t(comp[size_t(i)])];
  }
  bool any_unused = false;
  for (int32_t i = 0; i < n; ++i)
    if (!used[size_t(i)]) {
      comp[size_t(i)] = n_comp;
      any_unused = true;
    }
  if (any_unused) ++n_comp;

  // Each body's frame: sweep axis last, tile width and layer thickness from its own
  // bounding box and element count.
  struct Frame {
    double lo[3], hi[3];
    int64_t ne = 0;
    int ax[3] = {0, 1, 2};
    double w = 1.0, q = 1.0;
    int64_t ntv = 1, ns = 1, nv = 1, nu = 1;
  };
  std::vector<Frame> F(static_cast<size_t>(n_comp));
  for (Frame& f : F)
    for (int c = 0; c < 3; ++c) {
      f.lo[c] = std::numeric_limits<double>::infinity();
      f.hi[c] = -std::numeric_limits<double>::infinity();
    }
  for (int32_t i = 0; i < n; ++i) {
    Frame& f = F[size_t(comp[size_t(i)])];
    for (int c = 0; c < 3; ++c) {
      f.lo[c] = std::min(f.lo[c], x[3 * size_t(i) + c]);
      f.hi[c] = std::max(f.hi[c], x[3 * size_t(i) + c]);
    }
  }
  for (const cxg_block_dev& B : m->blocks)
    for (int64_t e = 0; e < B.n_elem; ++e) ++F[size_t(comp[size_t(conn[size_t(B.conn_offset + e * B.nn)])])].ne;
  const auto cell = [](double d, double step) { return int64_t(std::floor(d / step)); };
  unsigned __int128 span_max = 0;
  for (Frame& f : F) {
    std::sort(f.ax, f.ax + 3, [&](int a, int b) { return f.hi[a] - f.lo[a] < f.hi[b] - f.lo[b]; });
    const double ext = f.hi[f.ax[2]] - f.lo[f.ax[2]];
    double vol = 1.0;
    int dims = 0;
    for (int c = 0; c < 3; ++c)
      if (f.hi[c] - f.lo[c] > 1e-12 * ext) {
        vol *= f.hi[c] - f.lo[c];
        ++dims;
      }
    const double ne = double(std::max<int64_t>(f.ne, 1));
    double h = dims > 0 ? std::pow(vol / ne, 1.0 / dims) : 1.0;
    if (!(h > 0.0) || !std::isfinite(h)) h = 1.0;
    f.w = tile * h;
    f.q = 0.25 * h;
    // key (tile column, layer, v, u) of every node; the cells are >= 0 and at most those
    // of the bounding box's far corner
    f.ntv = cell(f.hi[f.ax[1]] - f.lo[f.ax[1]], f.w) + 1;
    const int64_t nt = (cell(f.hi[f.ax[0]] - f.lo[f.ax[0]], f.w) + 1) * f.ntv;
    f.ns = cell(f.hi[f.ax[2]] - f.lo[f.ax[2]], f.q) + 1;
    f.nv = cell(f.hi[f.ax[1]] - f.lo[f.ax[1]], f.q) + 1;
    f.nu = cell(f.hi[f.ax[0]] - f.lo[f.ax[0]], f.q) + 1;
    const unsigned __int128 span =
        static_cast<unsigned __int128>(nt) * uint64_t(f.ns) * uint64_t(f.nv) * uint64_t(f.nu);
    span_max = std::max(span_max, span);
  }
  const auto cells = [&](int32_t i, int64_t c[4]) {
    const Frame& f = F[size_t(comp[size_t(i)])];
    const double* xi = &x[3 * size_t(i)];
    const double du = xi[f.ax[0]] - f.lo[f.ax[0]], dv = xi[f.ax[1]] - f.lo[f.ax[1]],
                 ds = xi[f.ax[2]] - f.lo[f.ax[2]];
    c[0] = cell(du, f.w) * f.ntv + cell(dv, f.w);
    c[1] = cell(ds, f.q);
    c[2] = cell(dv, f.q);
    c[3] = cell(du, f.q);
  };
  std::vector<int32_t> ord(static_cast<size_t>(n));
  if (span_max < (static_cast<unsigned __int128>(1) << 63)) {
    // one 64-bit key per node within its body, ordered by a stable LSD radix sort over
    // 16-bit digits (ties keep ascending node order), then a stable counting sort by body
    std::vector<uint64_t> k(static_cast<size_t>(n)), k2(static_cast<size_t>(n));
    std::vector<int32_t> o2(static_cast<size_t>(n));
    uint64_t kmax = 0;
    for (int32_t i = 0; i < n; ++i) {
      const Frame& f = F[size_t(comp[size_t(i)])];
      int64_t c[4];
      cells(i, c);
      k[size_t(i)] = ((uint64_t(c[0]) * uint64_t(f.ns) + uint64_t(c[1])) * uint64_t(f.nv) + uint64_t(c[2])) *
                         uint64_t(f.nu) + uint64_t(c[3]);
      kmax = std::max(kmax, k[size_t(i)]);
      ord[size_t(i)] = i;
    }
    std::vector<size_t> hist(size_t(1) << 16);
    for (int shift = 0; shift < 64 && (kmax >> shift) != 0; shift += 16) {
      std::fill(hist.begin(), hist.end(), 0);
      for (uint64_t v : k) ++hist[(v >> shift) & 0xffff];
      size_t run = 0;
      for (size_t& hcount : hist) {
        const size_t c = hcount;
        hcount = run;
        run += c;
      }
      for (size_t j = 0; j < k.size(); ++j) {
        const size_t at = hist[(k[j] >> shift) & 0xffff]++;
        k2[at] = k[j];
        o2[at] = ord[j];
      }
      k.swap(k2);
      ord.swap(o2);
    }
    if (n_comp > 1) {
      std::vector<size_t> start(size_t(n_comp) + 1, 0);
      for (int32_t i = 0; i < n; ++i) ++start[size_t(comp[size_t(i)]) + 1];
      for (int32_t c = 0; c < n_comp; ++c) start[size_t(c) + 1] += start[size_t(c)];
      for (int32_t j = 0; j < n; ++j) {
        const int32_t i = ord[size_t(j)];
        o2[start[size_t(comp[size_t(i)])]++] = i;
      }
      ord.swap(o2);
    }
  } else {
    struct Key {
      int64_t c[5];
      int32_t i;
    };
    std::vector<Key> key(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) {
      key[size_t(i)].c[0] = comp[size_t(i)];
      cells(i, key[size_t(i)].c + 1);
      key[size_t(i)].i = i;
    }
    std::sort(key.begin(), key.end(), [](const Key& a, const Key& b) {
      for (int j = 0; j < 5; ++j)
        if (a.c[j] != b.c[j]) return a.c[j] < b.c[j];
      return a.i < b.i;
    });
    for (int32_t i = 0; i < n; ++i) ord[size_t(i)] = key[size_t(i)].i;
  }
  // the gather indexes node data by this order: it must be a permutation of the nodes
  std::vector<uint8_t> seen(static_cast<size_t>(n), 0);
  for (const int32_t v : ord) {
    if (v < 0 || v >= n || seen[size_t(v)]) throw Error(CXG_ERR_STATE, "gather order is not a permutation of the nodes");
    seen[size_t(v)] = 1;
  }
  pm->gorder.alloc(size_t(n));
  CXG_HIP(hipMemcpy(pm->gorder.p, ord.data(), ord.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  pm->gorder_tile = tile;
}

void prepare_desc(cxg_context* ctx, const cxg_pattern* p) {
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  const bool ordered = p->gorder.p && p->gorder_tile > 0;
  const int key = ordered ? p->gorder_tile : -1;
  if (pm->gdesc_key == key) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CXG_HIP(hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone)
    throw Error(CXG_ERR_STATE, "the gather descriptors are built before capture (assemble once first)");
  guard_realloc(p, pm->gdesc.p != nullptr, "gather descriptors (gather order changed)");
  const int32_t n = p->mesh->n_nodes;
  pm->gdesc.alloc(size_t(std::max(n, 1)));
  if (n > 0) {
    hipLaunchKernelGGL(k_gather_desc, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n,
                       ordered ? p->gorder.p : nullptr, p->nfm.p, p->dof_eq.p, p->row_ptr.p,
                       p->mesh->inc_ptr.p, pm->gdesc.p);
    CXG_HIP(hipGetLastError());
  }
  pm->gdesc_key = key;
}

bool prepare_stages(cxg_context* ctx, const cxg_pattern* p) {
  // A ring beyond this is no longer cache-resident: the one-pass kernels run instead.
  constexpr int64_t kRingBytesMax = int64_t(1) << 30;
  constexpr int kMaxSpan = 8;
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  const int tile = p->gorder.p ? p->gorder_tile : -1;
  if (pm->stages_tile == tile && pm->stages_mb == ctx->stage_mb)
    return !pm->stages_off && pm->stages != nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CXG_HIP(hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone)
    throw Error(CXG_ERR_STATE, "the stage plan is built before capture (assemble once first)");
  guard_realloc(p, pm->stages != nullptr, "stage plan (gather order or stage_mb changed)");
  pm->stages.reset();
  pm->stages_off = true;
  pm->stages_tile = tile;
  pm->stages_mb = ctx->stage_mb;
  const cxg_mesh* m = p->mesh;
  const int64_t ne = m->n_elem;
  const int32_t nn = m->n_nodes;
  const int64_t n_inc = int64_t(m->inc_elem.n);
  if (ne < 2 || nn < 1) return false;
  auto P = std::make_unique<StagePlan>();
  // host copies of the element slots, the incidence and the gather order
  CXG_HIP(hipStreamSynchronize(ctx->stream));
  P->kp_host.resize(size_t(ne) + 1);
  std::vector<int64_t> iptr(size_t(nn) + 1);
  std::vector<int32_t> ielem(static_cast<size_t>(n_inc));
  std::vector<IncRecord> rec(static_cast<size_t>(n_inc));
  std::vector<int32_t> order(static_cast<size_t>(nn));
  CXG_HIP(hipMemcpy(P->kp_host.data(), p->kpptr.p, P->kp_host.size() * sizeof(int64_t),
                    hipMemcpyDeviceToHost));
  CXG_HIP(hipMemcpy(iptr.data(), m->inc_ptr.p, iptr.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (n_inc > 0) {
    CXG_HIP(hipMemcpy(ielem.data(), m->inc_elem.p, ielem.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    CXG_HIP(hipMemcpy(rec.data(), p->inc_rec.p, rec.size() * sizeof(IncRecord), hipMemcpyDeviceToHost));
  }
  if (tile > 0)
    CXG_HIP(hipMemcpy(order.data(), p->gorder.p, order.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  else
    for (int32_t i = 0; i < nn; ++i) order[size_t(i)] = i;
  const std::vector<int64_t>& kp = P->kp_host;
  // stages of consecutive elements, each at most stage_mb of element matrices
  const int64_t target = std::max<int64_t>(1, int64_t(ctx->stage_mb) * (int64_t(1) << 20) / 8);
  std::vector<int32_t> estage(static_cast<size_t>(ne));
  P->elem_first.push_back(0);
  for (int64_t e = 0; e < ne; ++e) {
    if (e > P->elem_first.back() && kp[size_t(e) + 1] - kp[size_t(P->elem_first.back())] > target)
      P->elem_first.push_back(e);
    estage[size_t(e)] = int32_t(P->elem_first.size()) - 1;
  }
  P->elem_first.push_back(ne);
  P->S = int(P->elem_first.size()) - 1;
  for (int s = 0; s < P->S; ++s)
    P->slot_cap = std::max(P->slot_cap, kp[size_t(P->elem_first[size_t(s) + 1])] -
                                            kp[size_t(P->elem_first[size_t(s)])]);
  // completion stage of every node (its last incident element's; nodes of no element go
  // with the first stage) and the widest span of stages one node reads
  std::vector<int32_t> cstage(static_cast<size_t>(nn), 0);
  P->W = 1;
  for (int32_t a = 0; a < nn; ++a) {
    if (iptr[size_t(a)] == iptr[size_t(a) + 1]) continue;
    const int32_t f = estage[size_t(ielem[size_t(iptr[size_t(a)])])];
    const int32_t c = estage[size_t(ielem[size_t(iptr[size_t(a) + 1] - 1)])];
    cstage[size_t(a)] = c;
    P->W = std::max(P->W, c - f + 1);
  }
  P->R = std::min(P->S, P->W + 1);
  if (P->S < 2 || P->W > kMaxSpan || int64_t(P->R) * P->slot_cap * 8 > kRingBytesMax) return false;
  // gather positions grouped by completion stage, the locality order kept within a stage
  P->node_off.assign(size_t(P->S) + 1, 0);
  for (int32_t a = 0; a < nn; ++a) ++P->node_off[size_t(cstage[size_t(a)]) + 1];
  for (int s = 0; s < P->S; ++s) P->node_off[size_t(s) + 1] += P->node_off[size_t(s)];
  std::vector<int32_t> sorder(static_cast<size_t>(nn));
  {
    std::vector<int64_t> fill(P->node_off.begin(), P->node_off.end() - 1);
    for (int32_t i = 0; i < nn; ++i) {
      const int32_t a = order[size_t(i)];
      sorder[size_t(fill[size_t(cstage[size_t(a)])]++)] = a;
    }
  }
  // incidence records addressing the ring: element e of stage s sits at slot s % R,
  // kp[e] - kp[first element of s] into it
  for (int64_t i = 0; i < n_inc; ++i) {
    const int32_t e = ielem[size_t(i)];
    const int32_t st = estage[size_t(e)];
    rec[size_t(i)].rows += int64_t(st % P->R) * P->slot_cap - kp[size_t(P->elem_first[size_t(st)])];
  }
  P->rec.alloc(size_t(std::max<int64_t>(n_inc, 1)));
  if (n_inc > 0)
    CXG_HIP(hipMemcpy(P->rec.p, rec.data(), rec.size() * sizeof(IncRecord), hipMemcpyHostToDevice));
  DevBuf<int32_t> sord(static_cast<size_t>(nn));
  CXG_HIP(hipMemcpy(sord.p, sorder.data(), sorder.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  P->desc.alloc(static_cast<size_t>(nn));
  hipLaunchKernelGGL(k_gather_desc, dim3(grid_for(nn)), dim3(256), 0, ctx->stream, nn, sord.p,
                     p->nfm.p, p->dof_eq.p, p->row_ptr.p, m->inc_ptr.p, P->desc.p);
  CXG_HIP(hipGetLastError());
  CXG_HIP(hipStreamSynchronize(ctx->stream));  // sord is released on return
  P->ring.alloc(size_t(int64_t(P->R) * P->slot_cap));
  CXG_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
  P->ev_elem.assign(size_t(P->S), nullptr);
  P->ev_gather.assign(size_t(P->S), nullptr);
  for (int s = 0; s < P->S; ++s) {
    CXG_HIP(hipEventCreateWithFlags(&P->ev_elem[size_t(s)], hipEventDisableTiming));
    CXG_HIP(hipEventCreateWithFlags(&P->ev_gather[size_t(s)], hipEventDisableTiming));
  }
  pm->stages = std::move(P);
  pm->stages_off = false;
  return true;
}

void prepare_work(cxg_context* ctx, const cxg_pattern* p) {
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  if (pm->work_chunk == ctx->chunk_elems && pm->queue.p) return;
  guard_realloc(p, pm->work.p != nullptr, "work queue (option chunk_elems changed)");
  const cxg_mesh* m = p->mesh;
  std::vector<BlockDesc> desc(m->blocks.size());
  std::vector<size_t> order(m->blocks.size());
  for (size_t bi = 0; bi < m->blocks.size(); ++bi) {
    const cxg_block_dev& B = m->blocks[bi];
    const ElemTable& T = ctx->tables[B.type];
    desc[bi] = BlockDesc{B.type, B.nn, B.ngp, 0, B.elem_offset, B.gp_offset, B.conn_offset,
                         p->kp_base[bi], T.dN, T.w};
    order[bi] = bi;
  }
  // costliest family first (pair blocks times Gauss points per element)
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    const auto cost = [&](size_t i) { return int64_t(desc[i].ngp) * desc[i].nn * desc[i].nn; };
    return cost(a) > cost(b);
  });
  // an item is chunk_elems groups of the J2 kernel's elements per wave
  std::vector<WorkChunk> items;
  for (size_t bi : order) {
    const int64_t n = m->blocks[bi].n_elem;
    const int64_t chunk = int64_t(ctx->chunk_elems) * j2_epw(m->blocks[bi].nn, m->blocks[bi].ngp);
    for (int64_t f = 0; f < n; f += chunk)
      items.push_back(WorkChunk{int32_t(bi), int32_t(f), int32_t(std::min<int64_t>(chunk, n - f)), 0});
  }
  const int32_t chunk = ctx->chunk_elems;
  if (items.size() > size_t(INT32_MAX)) throw Error(CXG_ERR_UNSUPPORTED, "too many work items");
  pm->work.alloc(std::max<size_t>(items.size(), 1));
  pm->bdesc.alloc(std::max<size_t>(desc.size(), 1));
  if (!pm->queue.p) pm->queue.alloc(1);
  CXG_HIP(hipStreamSynchronize(ctx->stream));
  if (!items.empty())
    CXG_HIP(hipMemcpy(pm->work.p, items.data(), items.size() * sizeof(WorkChunk), hipMemcpyHostToDevice));
  if (!desc.empty())
    CXG_HIP(hipMemcpy(pm->bdesc.p, desc.data(), desc.size() * sizeof(BlockDesc), hipMemcpyHostToDevice));
  pm->n_work = int32_t(items.size());
  pm->work_chunk = chunk;
}

void launch_assemble_j2(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                        const double* E, const double* nu, const int32_t* mat_id,
                        const void* j2_table, const double* u, const double* committed,
                        double* trial, double* vals, double* f_int, uint8_t* live) {
  const J2Dev* dmats = static_cast<const J2Dev*>(j2_table);
  const cxg_mesh* m = p->mesh;
  if (ctx->asm_variant >= 4) {
    // Two-phase: K_t and f_e of every element to the arenas, then the row gather.
    cxg_pattern* pm = const_cast<cxg_pattern*>(p);
    prepare_arenas(p, true);
    if (ctx->asm_variant == 5) {
      prepare_work(ctx, p);
      hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, ctx->stream, pm->queue.p);
      LaunchTimer timer(ctx, "elem_j2_persistent", m->n_elem);
      hipLaunchKernelGGL(k_elem_j2_persistent, dim3(persistent_grid(ctx, p)), dim3(kWave), 0,
                         ctx->stream, pm->work.p, pm->n_work, pm->queue.p, pm->bdesc.p, m->conn.p,
                         coords, E, nu, mat_id, dmats, m->n_points, u, committed, trial,
                         pm->ke_arena.p, pm->fe_arena.p, ctx->flag_dev);
      CXG_HIP(hipGetLastError());
    }
    const bool up = ctx->asm_variant >= 7 && uses_upper_blocks(ctx, m);
    if (up) prepare_order(ctx, p, coords);
    for (size_t bi = 0; ctx->asm_variant != 5 && bi < m->blocks.size(); ++bi) {
      const cxg_block_dev& B = m->blocks[bi];
      if (B.n_elem == 0) continue;
      const ElemTable& T = ctx->tables[B.type];
      const int32_t* conn = m->conn.p + B.conn_offset;
      double* kp = pm->ke_arena.p + p->kp_base[bi];
      double* fe = pm->fe_arena.p + 3 * B.conn_offset;
      const bool bup = up && B.nn >= ctx->up_min_nodes;  // this family's arena layout
      LaunchTimer timer(ctx, std::string(bup ? "elem_j2_up_" : "elem_j2_") + type_name(B.type), B.n_elem);
#define J2K(NN, NGP, QC, UPV)                                                                   \
  hipLaunchKernelGGL((k_elem_j2<NN, NGP, (NN == 20 && NGP == 27 ? 9 : QC), UPV>),                                             \
                     dim3(xcd_grid(ctx, (B.n_elem + j2_epw(NN, NGP) - 1) / j2_epw(NN, NGP))), dim3(kWave), \
                     0, ctx->stream, conn, B.elem_offset, B.n_elem, B.gp_offset, m->n_points,   \
                     T.dN, T.w, coords, E, nu, mat_id, dmats, u, committed, trial, kp, fe,      \
                     ctx->flag_dev, xcd_slab(ctx, (B.n_elem + j2_epw(NN, NGP) - 1) / j2_epw(NN, NGP)))
#define LAUNCH(NN, NGP, UPPER, QC) \
  if (bup) J2K(NN, NGP, QC, true); else J2K(NN, NGP, QC, false);
      CXG_DISPATCH(B.type, LAUNCH)
#undef LAUNCH
#undef J2K
      CXG_HIP(hipGetLastError());
    }
    if (m->n_nodes > 0) {
      LaunchTimer timer(ctx, up ? "row_gather_j2_up" : "row_gather_j2", m->n_elem);
      launch_gather(ctx, p, nullptr, pm->fe_arena.p, vals, f_int, live, 1, false, up);
      CXG_HIP(hipGetLastError());
    }
    return;
  }
  fill_f64(ctx->stream, vals, p->nnz, 0.0);
  fill_f64(ctx->stream, f_int, int64_t(m->n_nodes) * 3, 0.0);
  for_each_colour_block(ctx, "assemble_j2", m, [&](const LaunchSet& S) {
    const ElemTable& T = ctx->tables[S.B.type];
    const int32_t* conn = m->conn.p + S.B.conn_offset;
#define LAUNCH(NN, NGP, UP, QC)                                                                 \
  hipLaunchKernelGGL((k_assemble_j2<NN, NGP, UP>), dim3(S.patches), dim3(kWave), 0,             \
                     ctx->stream, conn, S.elist, S.pptr, S.B.elem_offset, S.B.gp_offset,        \
                     m->n_points, T.dN, T.w, coords, E, nu, mat_id, dmats, u, committed, trial, \
                     p->dof_eq.p, p->row_ptr.p, (p->slot_all.p + p->slot_base[S.bi]), vals, f_int, ctx->flag_dev)
    CXG_DISPATCH(S.B.type, LAUNCH)
#undef LAUNCH
  });
}

void launch_material_points(cxg_context* ctx, int64_t n_points, const int32_t* mat_id,
                            const void* j2_table, const double* E, const double* nu,
                            const double* strain, const double* committed, double* trial,
                            double* stress, double* tangent) {
  if (n_points <= 0) return;
  LaunchTimer timer(ctx, "material_points", n_points);
  hipLaunchKernelGGL(k_material_points, dim3(grid_for(n_points)), dim3(256), 0, ctx->stream,
                     n_points, mat_id, static_cast<const J2Dev*>(j2_table), E, nu, strain,
                     committed, trial, stress, tangent);
  CXG_HIP(hipGetLastError());
}

void launch_commit(cxg_context* ctx, int64_t n_points, double* committed, const double* trial) {
  const int64_t n = n_points * CXG_STATE_COMPONENTS;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, ctx->stream, committed, trial, n);
  CXG_HIP(hipGetLastError());
}

bool same_direction_fits(const cxg_context* ctx, const cxg_pattern* p) {
  return p->max_row <= std::min(kRowCap, ctx->gather_lds_cols);
}

void launch_assemble_scalar(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                            int kind, const double* rho, const double* sig, double* vals,
                            uint8_t* live, bool same_direction) {
  if (same_direction) {
    if (!live || !same_direction_fits(ctx, p))
      throw Error(CXG_ERR_UNSUPPORTED,
                  "same-direction rows need live flags and every row within the gather's LDS "
                  "columns (option gather_lds_cols)");
    prepare_same_direction(ctx, p);
  }
  const cxg_mesh* m = p->mesh;
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  prepare_arenas(p, false);
  static const char* const names[3] = {"elem_mass_", "elem_mass_lumped_", "elem_geometric_"};
  for (size_t bi = 0; bi < m->blocks.size(); ++bi) {
    const cxg_block_dev& B = m->blocks[bi];
    if (B.n_elem == 0) continue;
    const ElemTable& T = ctx->tables[B.type];
    const int32_t* conn = m->conn.p + B.conn_offset;
    double* kp = pm->ke_arena.p + p->kp_base[bi] / 9;  // scalar arena: n^2 per element
    LaunchTimer timer(ctx, std::string(names[kind]) + type_name(B.type), B.n_elem);
#define LAUNCH_K(NN, NGP, QC, KIND)                                                             \
  hipLaunchKernelGGL((k_elem_scalar<NN, NGP, QC, KIND>),                                        \
                     dim3(xcd_grid(ctx, wave_blocks(B.n_elem, 1, scalar_waves(KIND)))),         \
                     dim3(kWave * scalar_waves(KIND)), 0, ctx->stream, conn, B.elem_offset,     \
                     B.gp_offset, T.dN, T.w, T.N, coords, rho, sig, kp, ctx->flag_dev,          \
                     B.n_elem, xcd_slab(ctx, wave_blocks(B.n_elem, 1, scalar_waves(KIND))))
#define LAUNCH(NN, NGP, UP, QC)                                          \
  if (kind == kMassConsistent) LAUNCH_K(NN, NGP, QC, kMassConsistent);   \
  else if (kind == kMassLumped) LAUNCH_K(NN, NGP, QC, kMassLumped);      \
  else LAUNCH_K(NN, NGP, QC, kGeometric);
    CXG_DISPATCH(B.type, LAUNCH)
#undef LAUNCH
#undef LAUNCH_K
    CXG_HIP(hipGetLastError());
  }
  if (m->n_nodes > 0) {
    LaunchTimer timer(ctx, std::string(kind == kGeometric ? "row_gather_geometric" : "row_gather_mass") +
                               (same_direction ? "_sd" : ""), m->n_elem);
    launch_gather(ctx, p, nullptr, nullptr, vals, nullptr, live, 1, true, false, 0, same_direction);
    CXG_HIP(hipGetLastError());
  }
}

void launch_recover(cxg_context* ctx, const cxg_mesh* m, const double* coords, const double* E,
                    const double* nu, const double* u, const int32_t* mat_id,
                    const void* j2_table, const double* committed, const double* f_ext,
                    double* strain, double* stress, double* reaction, double* f_int,
                    double* gauss_stress) {
  m->rec_mean.ensure(size_t(std::max<int64_t>(12 * m->n_elem, 1)));
  m->rec_fe.ensure(std::max<size_t>(3 * m->conn.n, 1));
  const J2Dev* dmats = static_cast<const J2Dev*>(j2_table);
  for (const cxg_block_dev& B : m->blocks) {
    if (B.n_elem == 0) continue;
    const ElemTable& T = ctx->tables[B.type];
    const int32_t* conn = m->conn.p + B.conn_offset;
    LaunchTimer timer(ctx, "gauss_fields", B.n_elem);
#define LAUNCH_R(NN, NGP, J2)                                                                   \
  hipLaunchKernelGGL((k_gauss_fields<NN, NGP, J2>),                                             \
                     dim3(xcd_grid(ctx, (B.n_elem + rec_epw(NN, NGP) - 1) / rec_epw(NN, NGP))), \
                     dim3(kWave), 0, ctx->stream, conn, B.elem_offset, B.gp_offset,             \
                     m->n_points, T.dN, T.w, coords, E, nu, u, mat_id, dmats, committed,        \
                     m->rec_mean.p, m->inc_pos.p + B.conn_offset, m->rec_fe.p, gauss_stress,    \
                     ctx->flag_dev, B.n_elem,                                                   \
                     xcd_slab(ctx, (B.n_elem + rec_epw(NN, NGP) - 1) / rec_epw(NN, NGP)))
#define LAUNCH(NN, NGP, UP, QC)                                  \
  if (dmats && mat_id) LAUNCH_R(NN, NGP, true);                  \
  else LAUNCH_R(NN, NGP, false);
    CXG_DISPATCH(B.type, LAUNCH)
#undef LAUNCH
#undef LAUNCH_R
    CXG_HIP(hipGetLastError());
  }
  if (m->n_nodes > 0 && (strain || stress || reaction || f_int)) {
    LaunchTimer timer(ctx, "node_fields", m->n_nodes);
    hipLaunchKernelGGL(k_node_fields, dim3(unsigned((int64_t(m->n_nodes) * 16 + 255) / 256)),
                       dim3(256), 0, ctx->stream, m->n_nodes, m->inc_ptr.p, m->inc_elem.p,
                       m->rec_mean.p, m->rec_fe.p, f_ext, strain, stress, reaction, f_int);
    CXG_HIP(hipGetLastError());
  }
}

void launch_vjp(cxg_context* ctx, const cxg_mesh* m, const double* coords, const double* E,
                const double* nu, const double* u, const double* w, double* dx, double* dE,
                double* du) {
  m->rec_fe.ensure(std::max<size_t>(3 * m->conn.n, 1));
  m->vjp_dx.ensure(std::max<size_t>(3 * m->conn.n, 1));
  for (const cxg_block_dev& B : m->blocks) {
    if (B.n_elem == 0) continue;
    const ElemTable& T = ctx->tables[B.type];
    const int32_t* conn = m->conn.p + B.conn_offset;
    LaunchTimer timer(ctx, std::string("vjp_") + type_name(B.type), B.n_elem);
#define LAUNCH(NN, NGP, UP, QC)                                                                 \
  hipLaunchKernelGGL((k_vjp<NN, NGP, kVjpWaves>),                                               \
                     dim3(xcd_grid(ctx, wave_blocks(B.n_elem, rec_epw(NN, NGP), kVjpWaves))),    \
                     dim3(kWave * kVjpWaves), 0, ctx->stream, conn, B.elem_offset, T.dN, T.w,   \
                     coords, E, nu, u, w, m->inc_pos.p + B.conn_offset, m->vjp_dx.p, dE,        \
                     m->rec_fe.p, ctx->flag_dev, B.n_elem,                                      \
                     xcd_slab(ctx, wave_blocks(B.n_elem, rec_epw(NN, NGP), kVjpWaves)))
    CXG_DISPATCH(B.type, LAUNCH)
#undef LAUNCH
    CXG_HIP(hipGetLastError());
  }
  if (m->n_nodes > 0) {
    LaunchTimer timer(ctx, "vjp_nodes", m->n_nodes);
    hipLaunchKernelGGL(k_node_sum_seq<8>, dim3(unsigned((int64_t(m->n_nodes) * 8 + 255) / 256)),
                       dim3(256), 0, ctx->stream, m->n_nodes, m->inc_ptr.p, m->vjp_dx.p,
                       m->rec_fe.p, dx, du);
    CXG_HIP(hipGetLastError());
  }
}

}  // namespace cxg
