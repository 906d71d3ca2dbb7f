// K3: CSR pattern of the free/free stiffness and the element -> CSR slot map.
//
// Replaces what the reference spreads over build_dof_map (src/fem/assembly.cpp:
// 19-42), the int64 COO keys of scatter_block (src/fem/assembly.cpp:80-94), the
// hash-ordered flush (src/fem/assembly.cpp:211-221) and SciPP's COO->CSR. The
// pattern is the structural one: free DOF (a,i) couples to free DOF (b,j) when
// nodes a and b share an element. Rows are free equations in ascending order;
// columns ascend inside a row (free equations are numbered in ascending global DOF
// order, constraint_transform.cpp:79-82, so sorting by node then component sorts by
// equation). Built once per mesh + constraint set and reused by every assembly.
//
// Steps:
//   1. node adjacency: one 64-lane wave per node gathers the nodes of its incident
//      elements into LDS, bitonic-sorts them and keeps the unique ones (count pass,
//      host prefix sum, fill pass);
//   2. per adjacency entry the offset of the neighbour's column block in the row
//      (prefix of free components), and per node the row length;
//   3. row_ptr by a host prefix sum over free equations;
//   4. col_idx, one wave per node;
//   5. slot[e][a][b] = column offset of node b's block inside the rows of node a,
//      by binary search of the sorted adjacency (one thread per element node pair),
//      with node b's free-DOF mask in bits 28..30.
#include <algorithm>
#include <climits>
#include <vector>

#include "cxg_internal.hpp"

namespace cxg {
namespace {

constexpr int kWave = 64;
constexpr int kCap = 2048;  // candidate nodes per node (sum of incident element sizes)

template <bool FILL>
__global__ __launch_bounds__(kWave) void k_adjacency(
    const int64_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_elem,
    const int64_t* __restrict__ eptr, const int32_t* __restrict__ conn,
    int32_t* __restrict__ adj_count, const int64_t* __restrict__ adj_ptr,
    int32_t* __restrict__ adj, int* __restrict__ flag) {
  __shared__ int32_t buf[kCap];
  __shared__ int32_t total;
  const int32_t a = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t i0 = inc_ptr[a], i1 = inc_ptr[a + 1];
  if (lane == 0) {
    int64_t t = 0;
    for (int64_t i = i0; i < i1; ++i) {
      const int32_t e = inc_elem[i];
      t += eptr[e + 1] - eptr[e];
    }
    total = t > kCap ? -1 : int32_t(t);
  }
  __syncthreads();
  const int cnt = total;
  if (cnt < 0) {
    if (lane == 0) atomicOr(flag, kFlagValence);
    return;
  }
  int off = 0;
  for (int64_t i = i0; i < i1; ++i) {
    const int32_t e = inc_elem[i];
    const int64_t b = eptr[e];
    const int n = int(eptr[e + 1] - b);
    if (lane < n) buf[off + lane] = conn[b + lane];
    off += n;
  }
  int N = kWave;
  while (N < cnt) N <<= 1;
  for (int i = cnt + lane; i < N; i += kWave) buf[i] = INT_MAX;
  __syncthreads();
  for (int k = 2; k <= N; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < N; i += kWave) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int32_t x = buf[i], y = buf[ixj];
          const bool asc = (i & k) == 0;
          if ((x > y) == asc) {
            buf[i] = y;
            buf[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  int out = 0;
  const int64_t base_out = FILL ? adj_ptr[a] : 0;
  for (int base = 0; base < cnt; base += kWave) {
    const int i = base + lane;
    const bool keep = i < cnt && (i == 0 || buf[i] != buf[i - 1]);
    const unsigned long long m = __ballot(keep);
    if (FILL && keep) {
      const unsigned long long below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
      adj[base_out + out + __popcll(below)] = buf[i];
    }
    out += __popcll(m);
  }
  if (!FILL && lane == 0) adj_count[a] = out;
}

// Column-block offsets inside each node's rows, and the row length of each node.
__global__ void k_column_offsets(int32_t n_nodes, const int64_t* __restrict__ adj_ptr,
                                 const int32_t* __restrict__ adj,
                                 const int32_t* __restrict__ dof_eq, int32_t* __restrict__ cpre,
                                 int64_t* __restrict__ rowlen) {
  const int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_nodes) return;
  int64_t run = 0;
  for (int64_t p = adj_ptr[a]; p < adj_ptr[a + 1]; ++p) {
    cpre[p] = int32_t(run);
    const int64_t m = adj[p];
    run += (dof_eq[3 * m] >= 0) + (dof_eq[3 * m + 1] >= 0) + (dof_eq[3 * m + 2] >= 0);
  }
  rowlen[a] = run;
}

__global__ __launch_bounds__(kWave) void k_col_idx(const int64_t* __restrict__ adj_ptr,
                                                   const int32_t* __restrict__ adj,
                                                   const int32_t* __restrict__ cpre,
                                                   const int32_t* __restrict__ dof_eq,
                                                   const int64_t* __restrict__ row_ptr,
                                                   int32_t* __restrict__ col_idx) {
  const int32_t a = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t p0 = adj_ptr[a], deg = adj_ptr[a + 1] - p0;
  for (int c = 0; c < 3; ++c) {
    const int32_t eq = dof_eq[3 * int64_t(a) + c];
    if (eq < 0) continue;
    const int64_t rs = row_ptr[eq];
    for (int64_t p = lane; p < deg; p += kWave) {
      const int64_t m = adj[p0 + p];
      int64_t pos = rs + cpre[p0 + p];
      for (int j = 0; j < 3; ++j) {
        const int32_t e2 = dof_eq[3 * m + j];
        if (e2 >= 0) col_idx[pos++] = e2;
      }
    }
  }
}

__global__ void k_slot(int64_t n_elem, int nn, const int32_t* __restrict__ conn,
                       const int64_t* __restrict__ adj_ptr, const int32_t* __restrict__ adj,
                       const int32_t* __restrict__ cpre, const int32_t* __restrict__ dof_eq,
                       int32_t* __restrict__ slot, int* __restrict__ flag) {
  const int64_t total = n_elem * nn * nn;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total; idx += stride) {
    const int64_t le = idx / (nn * nn);
    const int r = int(idx % (nn * nn));
    const int32_t na = conn[le * nn + r / nn], nb = conn[le * nn + r % nn];
    int64_t lo = adj_ptr[na], hi = adj_ptr[na + 1] - 1;
    const int64_t p0 = lo;
    int64_t found = -1;
    while (lo <= hi) {
      const int64_t mid = (lo + hi) >> 1;
      const int32_t v = adj[mid];
      if (v == nb) {
        found = mid;
        break;
      }
      if (v < nb) lo = mid + 1;
      else hi = mid - 1;
    }
    const int64_t b3 = 3 * int64_t(nb);
    const int32_t free_b = (dof_eq[b3] >= 0 ? 1 : 0) | (dof_eq[b3 + 1] >= 0 ? 2 : 0) |
                           (dof_eq[b3 + 2] >= 0 ? 4 : 0);
    if (found < 0 || cpre[found] > kSlotMask) {
      atomicOr(flag, kFlagPattern);
      slot[idx] = 0;
    } else {
      slot[idx] = cpre[found] | (free_b << kSlotFreeShift);
      (void)p0;
    }
  }
}

// Incidence records of the two-phase gather and the per-node free masks.
__global__ void k_inc_records(int32_t n_nodes, const int64_t* __restrict__ inc_ptr,
                              const int32_t* __restrict__ inc_elem,
                              const int64_t* __restrict__ eptr, const int32_t* __restrict__ conn,
                              const int64_t* __restrict__ sptr, const int64_t* __restrict__ kpptr,
                              const int32_t* __restrict__ dof_eq, IncRecord* __restrict__ rec,
                              int8_t* __restrict__ nfm) {
  const int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_nodes) return;
  const int64_t a3 = 3 * int64_t(a);
  int small = 8;  // bit 3: every incident element has at most 10 nodes
  for (int64_t it = inc_ptr[a]; it < inc_ptr[a + 1]; ++it) {
    const int32_t e = inc_elem[it];
    const int64_t b = eptr[e];
    const int n = int(eptr[e + 1] - b);
    if (n > 10) small = 0;
    int la = 0;
    while (la < n && conn[b + la] != a) ++la;
    rec[it] = IncRecord{kpptr[e] + int64_t(3 * la) * (3 * n), sptr[e] + int64_t(la) * n, b, n, la};
  }
  nfm[a] = int8_t((dof_eq[a3] >= 0 ? 1 : 0) | (dof_eq[a3 + 1] >= 0 ? 2 : 0) |
                  (dof_eq[a3 + 2] >= 0 ? 4 : 0) | small);
}

// Entries kept per row (live != 0); one thread per row.
// Key-set compaction, one 64-lane wave per row (four rows per 256-thread block): the
// lanes take 64 consecutive entries at a time and count or place the live ones with a
// ballot, so every access is coalesced.
constexpr int kRowsPerBlock = 4;

__device__ __forceinline__ uint64_t lanes_below(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__global__ __launch_bounds__(256) void k_live_count(int32_t n_free,
                                                    const int64_t* __restrict__ row_ptr,
                                                    const uint8_t* __restrict__ live,
                                                    int64_t* __restrict__ count) {
  const int64_t r = int64_t(blockIdx.x) * kRowsPerBlock + threadIdx.x / kWave;
  if (r >= n_free) return;
  const int lane = threadIdx.x % kWave;
  const int64_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  int64_t c = 0;
  for (int64_t k = k0; k < k1; k += kWave) {
    const bool on = k + lane < k1 && live[k + lane];
    c += __popcll(__ballot(on));
  }
  if (lane == 0) count[r] = c;
}

// Exclusive scan of count[0..n) into out[0..n] (out[n] = the total), in three passes:
// per-block scans of kScanItems entries with the block totals, a scan of the block totals
// by one block, and the block offsets added back.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanItems = kScanThreads * kScanPer;

__device__ int64_t block_exclusive_scan(int64_t v, int64_t* lds, int64_t& total) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int64_t add = t >= off ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  total = lds[kScanThreads - 1];
  const int64_t excl = lds[t] - v;
  __syncthreads();
  return excl;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(int64_t n,
                                                              const int64_t* __restrict__ count,
                                                              int64_t* __restrict__ out,
                                                              int64_t* __restrict__ block_sum) {
  __shared__ int64_t lds[kScanThreads];
  const int64_t base = int64_t(blockIdx.x) * kScanItems + int64_t(threadIdx.x) * kScanPer;
  int64_t v[kScanPer];
  int64_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    v[i] = base + i < n ? count[base + i] : 0;
    sum += v[i];
  }
  int64_t total = 0;
  int64_t run = block_exclusive_scan(sum, lds, total);
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) block_sum[blockIdx.x] = total;
}

// One block: exclusive scan of the n_blocks block totals in place; out[n] = the total.
__global__ __launch_bounds__(kScanThreads) void k_scan_totals(int64_t n_blocks,
                                                              int64_t* __restrict__ block_sum,
                                                              int64_t* __restrict__ out, int64_t n) {
  __shared__ int64_t lds[kScanThreads];
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < n_blocks; b0 += kScanThreads) {
    const int64_t b = b0 + threadIdx.x;
    const int64_t v = b < n_blocks ? block_sum[b] : 0;
    int64_t total = 0;
    const int64_t excl = block_exclusive_scan(v, lds, total);
    if (b < n_blocks) block_sum[b] = carry + excl;
    carry += total;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int64_t n,
                                                           const int64_t* __restrict__ block_sum,
                                                           int64_t* __restrict__ out) {
  const int64_t base = int64_t(blockIdx.x) * kScanItems;
  const int64_t add = block_sum[blockIdx.x];
  for (int i = threadIdx.x; i < kScanItems; i += kScanThreads)
    if (base + i < n) out[base + i] += add;
}

// Copies the kept entries of each row, in column order.
__global__ __launch_bounds__(256) void k_compact(int32_t n_free, const int64_t* __restrict__ row_ptr,
                                                 const int32_t* __restrict__ col_idx,
                                                 const double* __restrict__ vals,
                                                 const uint8_t* __restrict__ live,
                                                 const int64_t* __restrict__ row_out,
                                                 int32_t* __restrict__ col_out,
                                                 double* __restrict__ vals_out) {
  const int64_t r = int64_t(blockIdx.x) * kRowsPerBlock + threadIdx.x / kWave;
  if (r >= n_free) return;
  const int lane = threadIdx.x % kWave;
  const int64_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  int64_t o = row_out[r];
  for (int64_t k = k0; k < k1; k += kWave) {
    const bool on = k + lane < k1 && live[k + lane];
    const uint64_t m = __ballot(on);
    if (on) {
      const int64_t at = o + __popcll(m & lanes_below(lane));
      col_out[at] = col_idx[k + lane];
      vals_out[at] = vals[k + lane];
    }
    o += __popcll(m);
  }
}

// The same-direction pattern: row (a, i) keeps the structural columns (b, i). comp[e] is
// the direction of free equation e; one wave per row, columns placed with a ballot.
__global__ void k_eq_comp(int64_t n_dof, const int32_t* __restrict__ dof_eq,
                          int8_t* __restrict__ comp) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g < n_dof && dof_eq[g] >= 0) comp[dof_eq[g]] = int8_t(g % 3);
}

__global__ __launch_bounds__(256) void k_sd_count(int32_t n_free,
                                                  const int64_t* __restrict__ row_ptr,
                                                  const int32_t* __restrict__ col_idx,
                                                  const int8_t* __restrict__ comp,
                                                  int64_t* __restrict__ count) {
  const int64_t r = int64_t(blockIdx.x) * kRowsPerBlock + threadIdx.x / kWave;
  if (r >= n_free) return;
  const int lane = threadIdx.x % kWave;
  const int8_t cr = comp[r];
  const int64_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  int64_t c = 0;
  for (int64_t k = k0; k < k1; k += kWave) {
    const bool on = k + lane < k1 && comp[col_idx[k + lane]] == cr;
    c += __popcll(__ballot(on));
  }
  if (lane == 0) count[r] = c;
}

__global__ __launch_bounds__(256) void k_sd_fill(int32_t n_free,
                                                 const int64_t* __restrict__ row_ptr,
                                                 const int32_t* __restrict__ col_idx,
                                                 const int8_t* __restrict__ comp,
                                                 const int64_t* __restrict__ row_out,
                                                 int32_t* __restrict__ col_out) {
  const int64_t r = int64_t(blockIdx.x) * kRowsPerBlock + threadIdx.x / kWave;
  if (r >= n_free) return;
  const int lane = threadIdx.x % kWave;
  const int8_t cr = comp[r];
  const int64_t k0 = row_ptr[r], k1 = row_ptr[r + 1];
  int64_t o = row_out[r];
  for (int64_t k = k0; k < k1; k += kWave) {
    const int32_t col = k + lane < k1 ? col_idx[k + lane] : 0;
    const bool on = k + lane < k1 && comp[col] == cr;
    const uint64_t m = __ballot(on);
    if (on) col_out[o + __popcll(m & lanes_below(lane))] = col;
    o += __popcll(m);
  }
}

// Exclusive scan of count[0..n) into out[0..n], out[n] = the total (stream order).
void exclusive_scan(hipStream_t s, int64_t n, const int64_t* count, int64_t* out) {
  const int64_t sblocks = (n + kScanItems - 1) / kScanItems;
  DevBuf<int64_t> bsum(static_cast<size_t>(sblocks));
  hipLaunchKernelGGL(k_scan_blocks, dim3(unsigned(sblocks)), dim3(kScanThreads), 0, s, n,
                     count, out, bsum.p);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanThreads), 0, s, sblocks, bsum.p, out, n);
  hipLaunchKernelGGL(k_scan_add, dim3(unsigned(sblocks)), dim3(kScanThreads), 0, s, n, bsum.p,
                     out);
  CXG_HIP(hipGetLastError());
  CXG_HIP(hipStreamSynchronize(s));  // bsum is released on return
}

}  // namespace

void prepare_same_direction(cxg_context* ctx, const cxg_pattern* p) {
  if (p->sd_nnz >= 0) return;
  cxg_pattern* pm = const_cast<cxg_pattern*>(p);
  hipStream_t s = ctx->stream;
  const int32_t n = p->n_free;
  const int64_t n_dof = 3 * int64_t(p->mesh->n_nodes);
  pm->sd_row_ptr.alloc(size_t(n) + 1);
  int64_t total = 0;
  if (n > 0) {
    DevBuf<int8_t> comp(static_cast<size_t>(n));
    DevBuf<int64_t> count(static_cast<size_t>(n));
    hipLaunchKernelGGL(k_eq_comp, dim3(unsigned((n_dof + 255) / 256)), dim3(256), 0, s, n_dof,
                       p->dof_eq.p, comp.p);
    const unsigned rblocks = unsigned((int64_t(n) + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL(k_sd_count, dim3(rblocks), dim3(256), 0, s, n, p->row_ptr.p,
                       p->col_idx.p, comp.p, count.p);
    CXG_HIP(hipGetLastError());
    exclusive_scan(s, n, count.p, pm->sd_row_ptr.p);
    CXG_HIP(hipMemcpy(&total, pm->sd_row_ptr.p + n, sizeof(int64_t), hipMemcpyDeviceToHost));
    pm->sd_col_idx.alloc(size_t(std::max<int64_t>(total, 1)));
    hipLaunchKernelGGL(k_sd_fill, dim3(rblocks), dim3(256), 0, s, n, p->row_ptr.p,
                       p->col_idx.p, comp.p, pm->sd_row_ptr.p, pm->sd_col_idx.p);
    CXG_HIP(hipGetLastError());
    CXG_HIP(hipStreamSynchronize(s));  // comp and count are released on return
  } else {
    CXG_HIP(hipMemcpy(pm->sd_row_ptr.p, &total, sizeof(int64_t), hipMemcpyHostToDevice));
    pm->sd_col_idx.alloc(1);
  }
  pm->sd_nnz = total;
}

int64_t compact_csr(cxg_context* ctx, int32_t n, const int64_t* row_ptr, const int32_t* col_idx,
                    const uint8_t* live, const double* vals, int64_t* row_ptr_out,
                    int32_t* col_idx_out, double* vals_out) {
  hipStream_t s = ctx->stream;
  int64_t kept = 0;
  if (n > 0) {
    DevBuf<int64_t> count(static_cast<size_t>(n));
    const unsigned rblocks = unsigned((int64_t(n) + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL(k_live_count, dim3(rblocks), dim3(256), 0, s, n, row_ptr, live, count.p);
    CXG_HIP(hipGetLastError());
    exclusive_scan(s, n, count.p, row_ptr_out);
    hipLaunchKernelGGL(k_compact, dim3(rblocks), dim3(256), 0, s, n, row_ptr, col_idx, vals,
                       live, row_ptr_out, col_idx_out, vals_out);
    CXG_HIP(hipGetLastError());
    CXG_HIP(hipMemcpyAsync(&kept, row_ptr_out + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  } else {
    CXG_HIP(hipMemcpyAsync(row_ptr_out, &kept, sizeof(int64_t), hipMemcpyHostToDevice, s));
  }
  CXG_HIP(hipStreamSynchronize(s));
  return kept;
}

// The reference's key set: scatter_block (src/fem/assembly.cpp:90) skips element
// contributions equal to 0.0, so an entry all of whose contributions are 0.0 never
// enters its COO map; the live flags of the gather mark the others.
int64_t compact_pattern(cxg_context* ctx, const cxg_pattern* p, const uint8_t* live,
                        const double* vals, int64_t* row_ptr_out, int32_t* col_idx_out,
                        double* vals_out) {
  return compact_csr(ctx, p->n_free, p->row_ptr.p, p->col_idx.p, live, vals, row_ptr_out,
                     col_idx_out, vals_out);
}

void build_pattern(cxg_context* ctx, const cxg_mesh* m, const int32_t* dof_eq_dev,
                   int32_t n_free, cxg_pattern* P) {
  hipStream_t s = ctx->stream;
  const int32_t nn_nodes = m->n_nodes;
  const int64_t n_dof = int64_t(nn_nodes) * 3;
  P->mesh = m;
  P->n_free = n_free;
  P->dof_eq.alloc(static_cast<size_t>(n_dof));
  CXG_HIP(hipMemcpyAsync(P->dof_eq.p, dof_eq_dev, size_t(n_dof) * sizeof(int32_t),
                         hipMemcpyDeviceToDevice, s));
  std::vector<int32_t> deq(static_cast<size_t>(n_dof));
  CXG_HIP(hipMemcpyAsync(deq.data(), dof_eq_dev, size_t(n_dof) * sizeof(int32_t),
                         hipMemcpyDeviceToHost, s));
  CXG_HIP(hipStreamSynchronize(s));
  // Free equations must be 0..n_free-1 in ascending global DOF order, the numbering
  // of build_constraint_transform (constraint_transform.cpp:79-82).
  int32_t expect = 0;
  for (int64_t g = 0; g < n_dof; ++g) {
    if (deq[size_t(g)] < 0) continue;
    if (deq[size_t(g)] != expect)
      throw Error(CXG_ERR_ARG, "dof_eq must number free DOFs 0..n_free-1 in ascending DOF order");
    ++expect;
  }
  if (expect != n_free) throw Error(CXG_ERR_ARG, "n_free does not match dof_eq");

  // 1. node adjacency
  DevBuf<int32_t> adj_count(static_cast<size_t>(nn_nodes));
  if (nn_nodes > 0)
    hipLaunchKernelGGL((k_adjacency<false>), dim3(unsigned(nn_nodes)), dim3(kWave), 0, s,
                       m->inc_ptr.p, m->inc_elem.p, m->eptr.p, m->conn.p, adj_count.p,
                       nullptr, nullptr, ctx->flag_dev);
  CXG_HIP(hipGetLastError());
  std::vector<int32_t> hcount(static_cast<size_t>(nn_nodes));
  CXG_HIP(hipMemcpyAsync(hcount.data(), adj_count.p, hcount.size() * sizeof(int32_t),
                         hipMemcpyDeviceToHost, s));
  check_flags(ctx);
  std::vector<int64_t> hadj_ptr(size_t(nn_nodes) + 1, 0);
  for (int32_t a = 0; a < nn_nodes; ++a) hadj_ptr[size_t(a) + 1] = hadj_ptr[size_t(a)] + hcount[size_t(a)];
  const int64_t n_adj = hadj_ptr.back();
  DevBuf<int64_t> adj_ptr(hadj_ptr.size());
  CXG_HIP(hipMemcpyAsync(adj_ptr.p, hadj_ptr.data(), hadj_ptr.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, s));
  DevBuf<int32_t> adj(size_t(std::max<int64_t>(n_adj, 1)));
  if (nn_nodes > 0)
    hipLaunchKernelGGL((k_adjacency<true>), dim3(unsigned(nn_nodes)), dim3(kWave), 0, s,
                       m->inc_ptr.p, m->inc_elem.p, m->eptr.p, m->conn.p, nullptr, adj_ptr.p,
                       adj.p, ctx->flag_dev);
  CXG_HIP(hipGetLastError());

  // 2. column-block offsets and row lengths
  DevBuf<int32_t> cpre(size_t(std::max<int64_t>(n_adj, 1)));
  DevBuf<int64_t> rowlen(static_cast<size_t>(nn_nodes));
  if (nn_nodes > 0)
    hipLaunchKernelGGL(k_column_offsets, dim3(unsigned((nn_nodes + 255) / 256)), dim3(256), 0, s,
                       nn_nodes, adj_ptr.p, adj.p, P->dof_eq.p, cpre.p, rowlen.p);
  CXG_HIP(hipGetLastError());
  std::vector<int64_t> hrowlen(static_cast<size_t>(nn_nodes));
  CXG_HIP(hipMemcpyAsync(hrowlen.data(), rowlen.p, hrowlen.size() * sizeof(int64_t),
                         hipMemcpyDeviceToHost, s));
  CXG_HIP(hipStreamSynchronize(s));

  // 3. row_ptr over free equations
  std::vector<int64_t> hrow(size_t(n_free) + 1, 0);
  for (int64_t g = 0; g < n_dof; ++g) {
    const int32_t eq = deq[size_t(g)];
    if (eq >= 0) hrow[size_t(eq) + 1] = hrow[size_t(eq)] + hrowlen[size_t(g / 3)];
  }
  P->nnz = hrow.back();
  P->max_row = 0;
  for (int64_t nd = 0; nd < nn_nodes; ++nd) P->max_row = std::max(P->max_row, hrowlen[size_t(nd)]);
  P->row_ptr.alloc(hrow.size());
  CXG_HIP(hipMemcpyAsync(P->row_ptr.p, hrow.data(), hrow.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, s));

  // 4. col_idx
  P->col_idx.alloc(size_t(std::max<int64_t>(P->nnz, 1)));
  if (nn_nodes > 0)
    hipLaunchKernelGGL(k_col_idx, dim3(unsigned(nn_nodes)), dim3(kWave), 0, s, adj_ptr.p, adj.p,
                       cpre.p, P->dof_eq.p, P->row_ptr.p, P->col_idx.p);
  CXG_HIP(hipGetLastError());

  // 5. element -> slot map, one flat array with per-element offsets
  std::vector<int64_t> hsptr(size_t(m->n_elem) + 1, 0);
  P->slot_base.clear();
  for (const cxg_block_dev& B : m->blocks) {
    P->slot_base.push_back(hsptr[size_t(B.elem_offset)]);
    for (int64_t e = 0; e < B.n_elem; ++e)
      hsptr[size_t(B.elem_offset + e) + 1] = hsptr[size_t(B.elem_offset + e)] + int64_t(B.nn) * B.nn;
  }
  P->slot_total = hsptr.back();
  std::vector<int64_t> hkp(size_t(m->n_elem) + 1, 0);
  P->kp_base.clear();
  for (const cxg_block_dev& B : m->blocks) {
    P->kp_base.push_back(hkp[size_t(B.elem_offset)]);
    for (int64_t e = 0; e < B.n_elem; ++e)
      hkp[size_t(B.elem_offset + e) + 1] = hkp[size_t(B.elem_offset + e)] + 9 * int64_t(B.nn) * B.nn;
  }
  P->kp_total = hkp.back();
  P->kpptr.alloc(hkp.size());
  CXG_HIP(hipMemcpyAsync(P->kpptr.p, hkp.data(), hkp.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, s));
  P->slot_all.alloc(size_t(std::max<int64_t>(hsptr.back(), 1)));
  P->sptr.alloc(hsptr.size());
  CXG_HIP(hipMemcpyAsync(P->sptr.p, hsptr.data(), hsptr.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, s));
  for (size_t bi = 0; bi < m->blocks.size(); ++bi) {
    const cxg_block_dev& B = m->blocks[bi];
    const int64_t total = B.n_elem * B.nn * B.nn;
    if (total > 0) {
      const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
      hipLaunchKernelGGL(k_slot, dim3(unsigned(blocks)), dim3(256), 0, s, B.n_elem, B.nn,
                         m->conn.p + B.conn_offset, adj_ptr.p, adj.p, cpre.p, P->dof_eq.p,
                         P->slot_all.p + P->slot_base[bi], ctx->flag_dev);
      CXG_HIP(hipGetLastError());
    }
  }
  const size_t n_inc = m->inc_elem.n;
  P->inc_rec.alloc(std::max<size_t>(n_inc, 1));
  P->nfm.alloc(size_t(std::max(nn_nodes, 1)));
  if (nn_nodes > 0) {
    hipLaunchKernelGGL(k_inc_records, dim3(unsigned((nn_nodes + 255) / 256)), dim3(256), 0, s,
                       nn_nodes, m->inc_ptr.p, m->inc_elem.p, m->eptr.p, m->conn.p, P->sptr.p,
                       P->kpptr.p, P->dof_eq.p, P->inc_rec.p, P->nfm.p);
    CXG_HIP(hipGetLastError());
  }
  check_flags(ctx);  // synchronizes: temporaries (and hsptr) may be freed afterwards
}

}  // namespace cxg
