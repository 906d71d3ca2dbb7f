// Sparse SPD solve K u = f on the device: Jacobi-preconditioned conjugate gradients
// over the assembled CSR (the reduced K_ff of the assembly kernels).
//
// Replaces ComputeBackend::solve_sparse (include/calculixpp/compute/backend.hpp:51-56)
// for the symmetric positive definite systems of linear and Newton steps (the CPU
// backend hands them to SciPP, src/compute/cpu_backend.cpp:33-46). Everything stays
// on the device: the CSR product (one 64-lane wave per group of consecutive rows with the
// same columns, such as the three rows of a node, so that the column indices are read and
// x gathered once per group; a fixed-order wave sum per row), the vector updates, and the
// dot products, reduced in a fixed order (per-thread strided
// sums, a fixed LDS tree per block, then one block over the block sums), so the result
// is bitwise reproducible. The CG scalars (alpha, beta, r.z) live in device memory;
// the host reads the residual norm only every kCheck iterations.
#include <cmath>
#include <vector>

#include "cxg_internal.hpp"

namespace cxg {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kParts = 512;  // blocks of every dot product (fixed: reproducible sums)
constexpr int kCheck = 16;   // iterations between host residual checks

struct PcgScalars {
  double rz, pq, rr, bb, alpha, beta;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Fixed-order block reduction of one value per thread (kBlock threads).
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// y = A x, one wave per row.
__global__ __launch_bounds__(kBlock) void k_spmv(int32_t n, const int64_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ci,
                                                 const double* __restrict__ va,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ y) {
  const int64_t row = int64_t(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (row >= n) return;
  double s = 0.0;
  for (int64_t k = rp[row] + lane; k < rp[row + 1]; k += kWave) s += va[k] * x[ci[k]];
  s = wave_sum(s);
  if (lane == 0) y[row] = s;
}

constexpr int kGroupRows = 3;  // rows per group of the grouped product (a node's DOFs)

// same[r] = 1 when row r has the length and the columns of row r - 1 (one wave per row).
__global__ __launch_bounds__(kBlock) void k_same_columns(int32_t n, const int64_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         uint8_t* __restrict__ same) {
  const int64_t row = int64_t(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (row >= n) return;
  bool eq = row > 0 && rp[row + 1] - rp[row] == rp[row] - rp[row - 1];
  if (eq) {
    const int64_t len = rp[row + 1] - rp[row], a = rp[row - 1], b = rp[row];
    bool mine = true;
    for (int64_t k = lane; k < len; k += kWave) mine = mine && ci[a + k] == ci[b + k];
    eq = __all(mine);
  }
  if (lane == 0) same[row] = eq ? 1 : 0;
}

// y = A x over groups of up to kGroupRows consecutive rows with the same columns, one wave
// per group (group g: rows first[g] .. first[g + 1] - 1): column k of the first row is
// read once and x[ci[k]] gathered once for all rows of the group. Lane l takes offsets
// l, l + 64, ... of every row and the wave sums in the order of k_spmv, so y is bitwise
// k_spmv's.
__global__ __launch_bounds__(kBlock) void k_spmv_groups(int32_t n_groups,
                                                        const int32_t* __restrict__ first,
                                                        const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci,
                                                        const double* __restrict__ va,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y) {
  const int64_t g = int64_t(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (g >= n_groups) return;
  const int32_t r0 = first[g], nr = first[g + 1] - r0;
  const int64_t b0 = rp[r0], len = rp[r0 + 1] - b0;
  double s[kGroupRows] = {0.0, 0.0, 0.0};
  int64_t base[kGroupRows];
#pragma unroll
  for (int r = 0; r < kGroupRows; ++r) base[r] = r < nr ? rp[r0 + r] : b0;
  for (int64_t k = lane; k < len; k += kWave) {
    const double xv = x[ci[b0 + k]];
#pragma unroll
    for (int r = 0; r < kGroupRows; ++r)
      if (r < nr) s[r] += va[base[r] + k] * xv;
  }
#pragma unroll
  for (int r = 0; r < kGroupRows; ++r) {
    const double t = wave_sum(s[r]);
    if (lane == 0 && r < nr) y[r0 + r] = t;
  }
}

// dinv[i] = 1 / K_ii (1 where the diagonal is missing or not positive).
__global__ void k_diag_inv(int32_t n, const int64_t* __restrict__ rp,
                           const int32_t* __restrict__ ci, const double* __restrict__ va,
                           double* __restrict__ dinv) {
  const int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t lo = rp[i], hi = rp[i + 1];
  double d = 0.0;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (ci[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  if (lo < rp[i + 1] && ci[lo] == i) d = va[lo];
  dinv[i] = d > 0.0 ? 1.0 / d : 1.0;
}

// part[block] = sum over the block's strided share of a.b (and c.d when c is given).
__global__ __launch_bounds__(kBlock) void k_dot2(int32_t n, const double* __restrict__ a,
                                                 const double* __restrict__ b,
                                                 const double* __restrict__ c,
                                                 const double* __restrict__ d,
                                                 double* __restrict__ part1,
                                                 double* __restrict__ part2) {
  __shared__ double sh[kBlock];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(kParts) * kBlock) {
    s1 += a[i] * b[i];
    if (c) s2 += c[i] * d[i];
  }
  s1 = block_sum(s1, sh);
  if (c) s2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    part1[blockIdx.x] = s1;
    if (c) part2[blockIdx.x] = s2;
  }
}

// Sums kParts partials in a fixed order; what = 0: r.z, r.r, b.b of the start;
// 1: alpha = rz / p.q; 2: beta = rz_new / rz, rz = rz_new, rr.
__global__ __launch_bounds__(kBlock) void k_scalars(const double* __restrict__ part1,
                                                    const double* __restrict__ part2,
                                                    const double* __restrict__ part3,
                                                    PcgScalars* s, int what) {
  __shared__ double sh[kBlock];
  double v1 = 0.0, v2 = 0.0, v3 = 0.0;
  for (int i = threadIdx.x; i < kParts; i += kBlock) {
    v1 += part1[i];
    if (part2) v2 += part2[i];
    if (part3) v3 += part3[i];
  }
  v1 = block_sum(v1, sh);
  if (part2) v2 = block_sum(v2, sh);
  if (part3) v3 = block_sum(v3, sh);
  if (threadIdx.x != 0) return;
  if (what == 0) {
    s->rz = v1;
    s->rr = v2;
    s->bb = v3;
  } else if (what == 1) {
    s->pq = v1;
    s->alpha = v1 != 0.0 ? s->rz / v1 : 0.0;
  } else {
    s->beta = s->rz != 0.0 ? v1 / s->rz : 0.0;
    s->rz = v1;
    s->rr = v2;
  }
}

// r = b - A x (Ax in q), z = dinv r, p = z.
__global__ void k_start(int32_t n, const double* __restrict__ b, const double* __restrict__ q,
                        const double* __restrict__ dinv, double* __restrict__ r,
                        double* __restrict__ z, double* __restrict__ p) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ri = b[i] - q[i];
  r[i] = ri;
  z[i] = dinv[i] * ri;
  p[i] = z[i];
}

// x += alpha p, r -= alpha q, z = dinv r.
__global__ void k_update(int32_t n, const PcgScalars* __restrict__ s, const double* __restrict__ p,
                         const double* __restrict__ q, const double* __restrict__ dinv,
                         double* __restrict__ x, double* __restrict__ r, double* __restrict__ z) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = s->alpha;
  x[i] += a * p[i];
  const double ri = r[i] - a * q[i];
  r[i] = ri;
  z[i] = dinv[i] * ri;
}

// p = z + beta p.
__global__ void k_direction(int32_t n, const PcgScalars* __restrict__ s,
                            const double* __restrict__ z, double* __restrict__ p) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = z[i] + s->beta * p[i];
}

}  // namespace

void pcg(cxg_context* ctx, int32_t n, const int64_t* rp, const int32_t* ci, const double* va,
         const double* b, double* x, double tol, int32_t max_iter, int32_t& iters,
         double& rel_res) {
  hipStream_t st = ctx->stream;
  iters = 0;
  rel_res = 0.0;
  if (n <= 0) return;
  const size_t nn = static_cast<size_t>(n);
  DevBuf<double> r(nn), z(nn), p(nn), q(nn), dinv(nn);
  DevBuf<double> part(3 * static_cast<size_t>(kParts));
  DevBuf<PcgScalars> s(1);
  double* p1 = part.p;
  double* p2 = part.p + kParts;
  double* p3 = part.p + 2 * kParts;
  const unsigned g_rows = unsigned((int64_t(n) + kBlock / kWave - 1) / (kBlock / kWave));
  const unsigned g_vec = unsigned((int64_t(n) + kBlock - 1) / kBlock);
  // groups of consecutive rows with the same columns (up to kGroupRows each)
  DevBuf<int32_t> first;
  int32_t n_groups = 0;
  {
    LaunchTimer t(ctx, "pcg_groups", 0);
    DevBuf<uint8_t> same(nn);
    hipLaunchKernelGGL(k_same_columns, dim3(g_rows), dim3(kBlock), 0, st, n, rp, ci, same.p);
    CXG_HIP(hipGetLastError());
    std::vector<uint8_t> hs(nn);
    CXG_HIP(hipMemcpyAsync(hs.data(), same.p, nn, hipMemcpyDeviceToHost, st));
    CXG_HIP(hipStreamSynchronize(st));
    std::vector<int32_t> hf;
    hf.reserve(nn / 2 + 2);
    for (int32_t r = 0; r < n; ++r)
      if (r == 0 || !hs[size_t(r)] || r - hf.back() >= kGroupRows) hf.push_back(r);
    n_groups = int32_t(hf.size());
    hf.push_back(n);
    first.alloc(hf.size());
    CXG_HIP(hipMemcpyAsync(first.p, hf.data(), hf.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  }
  const unsigned g_groups = unsigned((int64_t(n_groups) + kBlock / kWave - 1) / (kBlock / kWave));
  {
    LaunchTimer t(ctx, "pcg_setup", 0);
    hipLaunchKernelGGL(k_diag_inv, dim3(g_vec), dim3(kBlock), 0, st, n, rp, ci, va, dinv.p);
    hipLaunchKernelGGL(k_spmv, dim3(g_rows), dim3(kBlock), 0, st, n, rp, ci, va, x, q.p);
    hipLaunchKernelGGL(k_start, dim3(g_vec), dim3(kBlock), 0, st, n, b, q.p, dinv.p, r.p, z.p, p.p);
    hipLaunchKernelGGL(k_dot2, dim3(kParts), dim3(kBlock), 0, st, n, r.p, z.p, r.p, r.p, p1, p2);
    hipLaunchKernelGGL(k_dot2, dim3(kParts), dim3(kBlock), 0, st, n, b, b, nullptr, nullptr, p3, nullptr);
    hipLaunchKernelGGL(k_scalars, dim3(1), dim3(kBlock), 0, st, p1, p2, p3, s.p, 0);
    CXG_HIP(hipGetLastError());
  }
  PcgScalars h{};
  CXG_HIP(hipMemcpyAsync(&h, s.p, sizeof(h), hipMemcpyDeviceToHost, st));
  CXG_HIP(hipStreamSynchronize(st));
  const double bnorm = std::sqrt(h.bb);
  if (bnorm == 0.0) {  // K u = 0: u = 0
    std::vector<double> zero(size_t(n), 0.0);
    CXG_HIP(hipMemcpyAsync(x, zero.data(), size_t(n) * sizeof(double), hipMemcpyHostToDevice, st));
    CXG_HIP(hipStreamSynchronize(st));
    return;
  }
  rel_res = std::sqrt(h.rr) / bnorm;
  while (rel_res > tol && iters < max_iter) {
    {  // the timer's end event precedes the synchronization below
      LaunchTimer t(ctx, "pcg_iterations", kCheck);
      for (int k = 0; k < kCheck && iters < max_iter; ++k, ++iters) {
        hipLaunchKernelGGL(k_spmv_groups, dim3(g_groups), dim3(kBlock), 0, st, n_groups, first.p,
                           rp, ci, va, p.p, q.p);
        hipLaunchKernelGGL(k_dot2, dim3(kParts), dim3(kBlock), 0, st, n, p.p, q.p, nullptr, nullptr, p1, nullptr);
        hipLaunchKernelGGL(k_scalars, dim3(1), dim3(kBlock), 0, st, p1, nullptr, nullptr, s.p, 1);
        hipLaunchKernelGGL(k_update, dim3(g_vec), dim3(kBlock), 0, st, n, s.p, p.p, q.p, dinv.p, x, r.p, z.p);
        hipLaunchKernelGGL(k_dot2, dim3(kParts), dim3(kBlock), 0, st, n, r.p, z.p, r.p, r.p, p1, p2);
        hipLaunchKernelGGL(k_scalars, dim3(1), dim3(kBlock), 0, st, p1, p2, nullptr, s.p, 2);
        hipLaunchKernelGGL(k_direction, dim3(g_vec), dim3(kBlock), 0, st, n, s.p, z.p, p.p);
      }
      CXG_HIP(hipGetLastError());
    }
    CXG_HIP(hipMemcpyAsync(&h, s.p, sizeof(h), hipMemcpyDeviceToHost, st));
    CXG_HIP(hipStreamSynchronize(st));
    rel_res = std::sqrt(h.rr) / bnorm;
    if (!std::isfinite(rel_res)) throw Error(CXG_ERR_STATE, "conjugate gradients broke down (matrix not SPD?)");
  }
  collect_timing(ctx);
}

}  // namespace cxg

extern "C" cxg_status cxg_pcg(cxg_context* ctx, int32_t n, const int64_t* row_ptr,
                              const int32_t* col_idx, const double* vals, const double* b,
                              double* x, double tol, int32_t max_iter, int32_t* iters,
                              double* rel_residual) {
  if (n < 0 || (n > 0 && (!row_ptr || !col_idx || !vals || !b || !x)) || !(tol >= 0.0) ||
      max_iter < 0)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad pcg arguments");
  return cxg::guarded(ctx, [&] {
    int32_t it = 0;
    double rr = 0.0;
    cxg::pcg(ctx, n, row_ptr, col_idx, vals, b, x, tol, max_iter, it, rr);
    if (iters) *iters = it;
    if (rel_residual) *rel_residual = rr;
    if (rr > tol)
      throw cxg::Error(CXG_ERR_STATE, "conjugate gradients did not reach the tolerance");
  });
}
