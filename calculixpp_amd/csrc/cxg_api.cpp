// C ABI of libcxg.so (include/cxg.h): argument checks, object lifetime, the mesh
// set-up (node incidence + element colouring) and the translation of internal
// errors into status codes + last-error text.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "cxg_internal.hpp"

using cxg::DevBuf;
using cxg::Error;

namespace cxg {

static hipEvent_t take_event(cxg_context* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  CXG_HIP(hipEventCreate(&e));
  return e;
}

LaunchTimer::LaunchTimer(cxg_context* c, std::string name, int64_t elements)
    : LaunchTimer(c, std::move(name), elements, c->stream) {}

LaunchTimer::LaunchTimer(cxg_context* c, std::string name, int64_t elements, hipStream_t s)
    : ctx(c), on(c->timing), stream(s) {
  if (!on) return;
  p.name = std::move(name);
  p.elements = elements;
  p.a = take_event(ctx);
  p.b = take_event(ctx);
  (void)hipEventRecord(p.a, stream);
}

LaunchTimer::~LaunchTimer() {
  if (!on) return;
  (void)hipEventRecord(p.b, stream);
  ctx->pending.push_back(p);
}

void collect_timing(cxg_context* ctx) {
  for (const auto& p : ctx->pending) {
    float ms = 0.f;
    CXG_HIP(hipEventSynchronize(p.b));
    CXG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto it = std::find_if(ctx->stats.begin(), ctx->stats.end(),
                           [&](const cxg_context::Stat& s) { return s.name == p.name; });
    if (it == ctx->stats.end()) {
      ctx->stats.push_back({p.name, 0, 0, 0.0});
      it = ctx->stats.end() - 1;
    }
    it->launches += 1;
    it->elements += p.elements;
    it->ms += double(ms);
    ctx->event_pool.push_back(p.a);
    ctx->event_pool.push_back(p.b);
  }
  ctx->pending.clear();
}

void check_flags(cxg_context* ctx) {
  CXG_HIP(hipMemcpyAsync(ctx->flag_host, ctx->flag_dev, sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  CXG_HIP(hipStreamSynchronize(ctx->stream));
  if (!ctx->pending.empty()) collect_timing(ctx);
  const int f = *ctx->flag_host;
  if (f == 0) return;
  *ctx->flag_host = 0;
  CXG_HIP(hipMemcpyAsync(ctx->flag_dev, ctx->flag_host, sizeof(int), hipMemcpyHostToDevice,
                         ctx->stream));
  CXG_HIP(hipStreamSynchronize(ctx->stream));
  if (f & kFlagJacobian) throw Error(CXG_ERR_JACOBIAN, kBadJacobian);
  if (f & kFlagValence)
    throw Error(CXG_ERR_UNSUPPORTED, "node valence exceeds the pattern builder capacity");
  throw Error(CXG_ERR_STATE, "element node pair missing from the CSR pattern");
}

namespace {

// Greedy first-fit colouring of patches in ascending order: patch p takes the
// smallest colour not used by an earlier patch sharing a node with it. With one
// element per patch on a structured hex grid this yields the 8-colouring
// (i%2, j%2, k%2); runs of consecutive elements along x colour the same way.
std::vector<int32_t> colour_patches(int64_t n_patches, const std::vector<int64_t>& patch_first,
                                    const std::vector<int32_t>& elem_patch,
                                    const std::vector<int64_t>& eptr,
                                    const std::vector<int32_t>& conn,
                                    const std::vector<int64_t>& inc_ptr,
                                    const std::vector<int32_t>& inc_elem, int32_t& n_colors) {
  constexpr int kMaxColours = 64;
  std::vector<int32_t> col(size_t(n_patches), -1);
  n_colors = 0;
  for (int64_t p = 0; p < n_patches; ++p) {
    uint64_t used = 0;
    for (int64_t e = patch_first[size_t(p)]; e < patch_first[size_t(p) + 1]; ++e)
      for (int64_t k = eptr[size_t(e)]; k < eptr[size_t(e) + 1]; ++k) {
        const int32_t nd = conn[size_t(k)];
        for (int64_t i = inc_ptr[size_t(nd)]; i < inc_ptr[size_t(nd) + 1]; ++i) {
          const int32_t q = elem_patch[size_t(inc_elem[size_t(i)])];
          const int32_t c = col[size_t(q)];
          if (c >= 0 && q != p) used |= (1ull << c);
        }
      }
    int c = 0;
    while (c < kMaxColours && (used >> c) & 1ull) ++c;
    if (c == kMaxColours) throw Error(CXG_ERR_UNSUPPORTED, "patch colouring needs more than 64 colours");
    col[size_t(p)] = c;
    n_colors = std::max(n_colors, c + 1);
  }
  return col;
}

void create_mesh(cxg_context* ctx, int32_t n_nodes, int32_t n_blocks, const cxg_block* blocks,
                 cxg_mesh* M) {
  if (n_nodes < 0 || n_blocks < 0 || (n_blocks > 0 && !blocks))
    throw Error(CXG_ERR_ARG, "bad mesh arguments");
  hipStream_t s = ctx->stream;
  M->ctx = ctx;
  M->n_nodes = n_nodes;
  int64_t n_elem = 0, n_points = 0, n_conn = 0;
  for (int32_t b = 0; b < n_blocks; ++b) {
    const cxg_block& B = blocks[b];
    const int nn = nodes_per_element(B.type);
    if (nn == 0) throw Error(CXG_ERR_UNSUPPORTED, "element type outside the GPU path");
    if (B.n_elem < 0 || (B.n_elem > 0 && !B.conn)) throw Error(CXG_ERR_ARG, "bad block");
    if (B.elem_offset != n_elem || B.gp_offset != n_points)
      throw Error(CXG_ERR_ARG, "blocks must cover consecutive elements and Gauss points");
    cxg_block_dev D;
    D.type = B.type;
    D.nn = nn;
    D.ngp = gauss_points(B.type);
    D.n_elem = B.n_elem;
    D.elem_offset = B.elem_offset;
    D.gp_offset = B.gp_offset;
    D.conn_offset = n_conn;
    M->blocks.push_back(std::move(D));
    n_elem += B.n_elem;
    n_points += B.n_elem * gauss_points(B.type);
    n_conn += B.n_elem * nn;
  }
  if (n_elem > INT32_MAX) throw Error(CXG_ERR_UNSUPPORTED, "more than 2^31-1 elements");
  M->n_elem = n_elem;
  M->n_points = n_points;
  std::vector<int32_t> conn(static_cast<size_t>(n_conn));
  std::vector<int64_t> eptr(size_t(n_elem) + 1, 0);
  for (int32_t b = 0; b < n_blocks; ++b) {
    const cxg_block_dev& D = M->blocks[size_t(b)];
    if (D.n_elem)
      CXG_HIP(hipMemcpyAsync(conn.data() + D.conn_offset, blocks[b].conn,
                             size_t(D.n_elem * D.nn) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    for (int64_t e = 0; e < D.n_elem; ++e)
      eptr[size_t(D.elem_offset + e) + 1] = D.conn_offset + (e + 1) * D.nn;
  }
  CXG_HIP(hipStreamSynchronize(s));
  for (int32_t v : conn)
    if (v < 0 || v >= n_nodes) throw Error(CXG_ERR_ARG, "element references unknown node");
  // A node listed twice by one element (a collapsed hex or wedge) would need one incidence
  // per occurrence in the row gather and the node sums; such meshes are outside the GPU
  // path and are refused here, before anything is assembled.
  for (int64_t e = 0; e < n_elem; ++e)
    for (int64_t k = eptr[size_t(e)]; k < eptr[size_t(e) + 1]; ++k)
      for (int64_t j = eptr[size_t(e)]; j < k; ++j)
        if (conn[size_t(j)] == conn[size_t(k)])
          throw Error(CXG_ERR_UNSUPPORTED,
                      "element " + std::to_string(e) + " lists node " +
                          std::to_string(conn[size_t(k)]) +
                          " twice (collapsed element): outside the GPU path");
  // node -> element incidence, ascending element order
  std::vector<int64_t> inc_ptr(size_t(n_nodes) + 1, 0);
  for (int32_t v : conn) ++inc_ptr[size_t(v) + 1];
  for (int32_t a = 0; a < n_nodes; ++a) inc_ptr[size_t(a) + 1] += inc_ptr[size_t(a)];
  std::vector<int32_t> inc_elem(static_cast<size_t>(n_conn));
  std::vector<int32_t> inc_pos(static_cast<size_t>(n_conn));
  if (n_conn > int64_t(INT32_MAX))
    throw Error(CXG_ERR_UNSUPPORTED, "more than 2^31 - 1 connectivity entries in one mesh");
  {
    std::vector<int64_t> cur(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int64_t e = 0; e < n_elem; ++e)
      for (int64_t k = eptr[size_t(e)]; k < eptr[size_t(e) + 1]; ++k) {
        const int32_t v = conn[size_t(k)];
        inc_pos[size_t(k)] = int32_t(cur[size_t(v)]);
        inc_elem[size_t(cur[size_t(v)]++)] = int32_t(e);
      }
  }
  // patches: up to patch_size consecutive elements of one block
  const int64_t P = std::max(1, ctx->patch_size);
  std::vector<int64_t> patch_first;
  std::vector<int32_t> elem_patch(static_cast<size_t>(n_elem));
  for (const cxg_block_dev& D : M->blocks)
    for (int64_t e = 0; e < D.n_elem; e += P) {
      const int64_t end = std::min(D.n_elem, e + P);
      for (int64_t k = e; k < end; ++k) elem_patch[size_t(D.elem_offset + k)] = int32_t(patch_first.size());
      patch_first.push_back(D.elem_offset + e);
    }
  const int64_t n_patches = int64_t(patch_first.size());
  patch_first.push_back(n_elem);
  int32_t ncol = 0;
  const std::vector<int32_t> colour =
      colour_patches(n_patches, patch_first, elem_patch, eptr, conn, inc_ptr, inc_elem, ncol);
  M->n_colors = ncol;
  for (cxg_block_dev& D : M->blocks) {
    std::vector<std::vector<int32_t>> lists(static_cast<size_t>(ncol)), ptrs(static_cast<size_t>(ncol));
    for (auto& v : ptrs) v.push_back(0);
    for (int64_t e = 0; e < D.n_elem; e += P) {
      const int32_t p = elem_patch[size_t(D.elem_offset + e)];
      const size_t c = size_t(colour[size_t(p)]);
      for (int64_t k = e; k < std::min(D.n_elem, e + P); ++k) lists[c].push_back(int32_t(k));
      ptrs[c].push_back(int32_t(lists[c].size()));
    }
    for (int c = 0; c < ncol; ++c) {
      DevBuf<int32_t> buf(lists[size_t(c)].size()), pb(ptrs[size_t(c)].size());
      if (!lists[size_t(c)].empty()) {
        CXG_HIP(hipMemcpyAsync(buf.p, lists[size_t(c)].data(), lists[size_t(c)].size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, s));
        CXG_HIP(hipMemcpyAsync(pb.p, ptrs[size_t(c)].data(), ptrs[size_t(c)].size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, s));
      }
      D.color_count.push_back(int64_t(lists[size_t(c)].size()));
      D.patch_count.push_back(int64_t(ptrs[size_t(c)].size()) - 1);
      D.color.push_back(std::move(buf));
      D.patch_ptr.push_back(std::move(pb));
    }
    CXG_HIP(hipStreamSynchronize(s));  // host lists go out of scope
  }
  M->conn.alloc(size_t(std::max<int64_t>(n_conn, 1)));
  M->eptr.alloc(eptr.size());
  M->inc_ptr.alloc(inc_ptr.size());
  M->inc_elem.alloc(std::max<size_t>(inc_elem.size(), 1));
  M->inc_pos.alloc(std::max<size_t>(inc_pos.size(), 1));
  if (n_conn)
    CXG_HIP(hipMemcpyAsync(M->conn.p, conn.data(), conn.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  CXG_HIP(hipMemcpyAsync(M->eptr.p, eptr.data(), eptr.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
  CXG_HIP(hipMemcpyAsync(M->inc_ptr.p, inc_ptr.data(), inc_ptr.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, s));
  if (!inc_elem.empty()) {
    CXG_HIP(hipMemcpyAsync(M->inc_elem.p, inc_elem.data(), inc_elem.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
    CXG_HIP(hipMemcpyAsync(M->inc_pos.p, inc_pos.data(), inc_pos.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
  }
  CXG_HIP(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace cxg

extern "C" {

const char* cxg_version(void) { return "cxg 0.1 gfx950"; }

int32_t cxg_nodes_per_element(int32_t type) { return cxg::nodes_per_element(type); }
int32_t cxg_gauss_points(int32_t type) { return cxg::gauss_points(type); }

cxg_status cxg_create(int device, void* stream, cxg_context** out) {
  if (!out) return CXG_ERR_ARG;
  *out = nullptr;
  auto* ctx = new (std::nothrow) cxg_context();
  if (!ctx) return CXG_ERR_HIP;
  ctx->device = device;
  ctx->stream = static_cast<hipStream_t>(stream);
  const cxg_status st = cxg::guarded(ctx, [&] {
    CXG_HIP(hipMalloc(reinterpret_cast<void**>(&ctx->flag_dev), sizeof(int)));
    CXG_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->flag_host), sizeof(int), hipHostMallocDefault));
    *ctx->flag_host = 0;
    CXG_HIP(hipMemcpy(ctx->flag_dev, ctx->flag_host, sizeof(int), hipMemcpyHostToDevice));
    ctx->scratch = cxg::scratch_device_ptr(device);
    int cus = 0;
    CXG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (cus > 0) ctx->n_cu = cus;
    for (int t : {CXG_C3D8, CXG_C3D20, CXG_C3D20R, CXG_C3D6}) {
      std::vector<double> dN, w, N;
      cxg::host_tables(t, dN, w, N);
      // the point-major table, then its transpose ((3k + c) n_gp + q), read by kernels
      // that give each lane one point
      const int nn = cxg::nodes_per_element(t), ng = cxg::gauss_points(t);
      const size_t n1 = dN.size();
      dN.resize(2 * n1);
      for (int q = 0; q < ng; ++q)
        for (int r = 0; r < 3 * nn; ++r) dN[n1 + size_t(r) * ng + q] = dN[size_t(q) * 3 * nn + r];
      ctx->tab_dN[t].alloc(dN.size());
      ctx->tab_w[t].alloc(w.size());
      ctx->tab_N[t].alloc(N.size());
      CXG_HIP(hipMemcpy(ctx->tab_dN[t].p, dN.data(), dN.size() * sizeof(double), hipMemcpyHostToDevice));
      CXG_HIP(hipMemcpy(ctx->tab_w[t].p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
      CXG_HIP(hipMemcpy(ctx->tab_N[t].p, N.data(), N.size() * sizeof(double), hipMemcpyHostToDevice));
      ctx->tables[t] = {cxg::nodes_per_element(t), cxg::gauss_points(t), ctx->tab_dN[t].p,
                        ctx->tab_w[t].p, ctx->tab_N[t].p};
    }
  });
  if (st != CXG_OK) {
    cxg_destroy(ctx);
    return st;
  }
  *out = ctx;
  return CXG_OK;
}

void cxg_destroy(cxg_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (!ctx->pending.empty()) (void)hipStreamSynchronize(ctx->stream);
  for (const auto& p : ctx->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  if (!ctx->allocs.empty()) (void)hipStreamSynchronize(ctx->stream);
  for (void* p : ctx->allocs) (void)hipFree(p);
  if (ctx->flag_dev) (void)hipFree(ctx->flag_dev);
  if (ctx->flag_host) (void)hipHostFree(ctx->flag_host);
  delete ctx;
}

cxg_status cxg_set_stream(cxg_context* ctx, void* stream) {
  if (!ctx) return CXG_ERR_ARG;
  ctx->stream = static_cast<hipStream_t>(stream);
  return CXG_OK;
}

const char* cxg_last_error(const cxg_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

cxg_status cxg_malloc(cxg_context* ctx, size_t bytes, void** out) {
  if (!out) return cxg::fail(ctx, CXG_ERR_ARG, "null output");
  *out = nullptr;
  return cxg::guarded(ctx, [&] {
    void* p = nullptr;
    CXG_HIP(hipMalloc(&p, bytes ? bytes : 1));
    ctx->allocs.insert(p);
    *out = p;
  });
}

cxg_status cxg_free(cxg_context* ctx, void* p) {
  if (!p) return CXG_OK;
  return cxg::guarded(ctx, [&] {
    if (ctx->allocs.erase(p) == 0) throw cxg::Error(CXG_ERR_ARG, "pointer not allocated by cxg_malloc");
    CXG_HIP(hipFree(p));
  });
}

cxg_status cxg_copy_to_device(cxg_context* ctx, void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return cxg::fail(ctx, CXG_ERR_ARG, "null copy argument");
  return cxg::guarded(ctx, [&] {
    if (!bytes) return;
    CXG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    CXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

cxg_status cxg_copy_to_host(cxg_context* ctx, void* dst, const void* src, size_t bytes) {
  if ((!dst || !src) && bytes) return cxg::fail(ctx, CXG_ERR_ARG, "null copy argument");
  return cxg::guarded(ctx, [&] {
    if (!bytes) return;
    CXG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

cxg_status cxg_set_option(cxg_context* ctx, const char* key, int64_t value) {
  if (!ctx || !key) return CXG_ERR_ARG;
  const std::string k(key);
  if (k == "patch_size" && value >= 1 && value <= 4096) {
    ctx->patch_size = int(value);
    return CXG_OK;
  }
  if (k == "asm_variant" && value >= 1 && value <= 8) {
    ctx->asm_variant = int(value);
    return CXG_OK;
  }
  if (k == "chunk_elems" && value >= 1 && value <= 4096) {
    ctx->chunk_elems = int(value);
    return CXG_OK;
  }
  if (k == "xcd_remap" && (value == 0 || value == 1)) {
    ctx->xcd_remap = int(value);
    return CXG_OK;
  }
  if (k == "persistent_waves" && value >= 1 && value <= 32) {
    ctx->persistent_waves = int(value);
    return CXG_OK;
  }
  if (k == "up_min_nodes" && value >= 0 && value <= 64) {
    ctx->up_min_nodes = int(value);
    return CXG_OK;
  }
  if (k == "elem_tiling" && (value == 0 || value == 1)) {
    ctx->elem_tiling = int(value);
    return CXG_OK;
  }
  if (k == "gather_tile" && value >= 0 && value <= 4096) {
    ctx->gather_tile = int(value);
    return CXG_OK;
  }
  if (k == "gradient_form" && (value == 0 || value == 1)) {
    ctx->gradient_form = int(value);
    return CXG_OK;
  }
  if (k == "stage_overlap" && (value == 0 || value == 1)) {
    ctx->stage_overlap = int(value);
    return CXG_OK;
  }
  if (k == "stage_mb" && value >= 1 && value <= 4096) {
    ctx->stage_mb = int(value);
    return CXG_OK;
  }
  if (k == "gather_lds_cols" && value >= 8 && value <= 256) {
    ctx->gather_lds_cols = int(value);
    return CXG_OK;
  }
  return cxg::fail(ctx, CXG_ERR_ARG, "unknown option or value");
}

cxg_status cxg_profile_enable(cxg_context* ctx, int enable) {
  return cxg::guarded(ctx, [&] {
    if (!enable && !ctx->pending.empty()) {
      CXG_HIP(hipStreamSynchronize(ctx->stream));
      cxg::collect_timing(ctx);
    }
    ctx->timing = enable != 0;
  });
}

cxg_status cxg_profile_reset(cxg_context* ctx) {
  return cxg::guarded(ctx, [&] {
    if (!ctx->pending.empty()) {
      CXG_HIP(hipStreamSynchronize(ctx->stream));
      cxg::collect_timing(ctx);
    }
    ctx->stats.clear();
  });
}

int32_t cxg_profile_count(const cxg_context* ctx) { return ctx ? int32_t(ctx->stats.size()) : 0; }

cxg_status cxg_profile_get(const cxg_context* ctx, int32_t i, const char** name,
                           int64_t* launches, double* total_ms, int64_t* elements) {
  if (!ctx || i < 0 || i >= int32_t(ctx->stats.size())) return CXG_ERR_ARG;
  const auto& s = ctx->stats[size_t(i)];
  if (name) *name = s.name.c_str();
  if (launches) *launches = s.launches;
  if (total_ms) *total_ms = s.ms;
  if (elements) *elements = s.elements;
  return CXG_OK;
}

cxg_status cxg_mesh_create(cxg_context* ctx, int32_t n_nodes, int32_t n_blocks,
                           const cxg_block* blocks, cxg_mesh** out) {
  if (!out) return CXG_ERR_ARG;
  *out = nullptr;
  auto* M = new (std::nothrow) cxg_mesh();
  if (!M) return CXG_ERR_HIP;
  const cxg_status st = cxg::guarded(ctx, [&] { cxg::create_mesh(ctx, n_nodes, n_blocks, blocks, M); });
  if (st != CXG_OK) {
    delete M;
    return st;
  }
  *out = M;
  return CXG_OK;
}

void cxg_mesh_destroy(cxg_mesh* mesh) {
  if (!mesh) return;
  (void)hipSetDevice(mesh->ctx->device);
  delete mesh;
}

cxg_status cxg_mesh_info(const cxg_mesh* m, int64_t* n_elem, int64_t* n_points, int32_t* n_colors) {
  if (!m) return CXG_ERR_ARG;
  if (n_elem) *n_elem = m->n_elem;
  if (n_points) *n_points = m->n_points;
  if (n_colors) *n_colors = m->n_colors;
  return CXG_OK;
}

cxg_status cxg_pattern_build(cxg_context* ctx, const cxg_mesh* mesh, const int32_t* dof_eq,
                             int32_t n_free, cxg_pattern** out) {
  if (!out) return CXG_ERR_ARG;
  *out = nullptr;
  if (!mesh || (!dof_eq && mesh->n_nodes > 0) || n_free < 0) return cxg::fail(ctx, CXG_ERR_ARG, "bad pattern arguments");
  auto* P = new (std::nothrow) cxg_pattern();
  if (!P) return CXG_ERR_HIP;
  const cxg_status st = cxg::guarded(ctx, [&] { cxg::build_pattern(ctx, mesh, dof_eq, n_free, P); });
  if (st != CXG_OK) {
    delete P;
    return st;
  }
  *out = P;
  return CXG_OK;
}

void cxg_pattern_destroy(cxg_pattern* p) {
  if (!p) return;
  (void)hipSetDevice(p->mesh->ctx->device);
  delete p;
}

cxg_status cxg_pattern_info(const cxg_pattern* p, int32_t* n_free, int64_t* nnz) {
  if (!p) return CXG_ERR_ARG;
  if (n_free) *n_free = p->n_free;
  if (nnz) *nnz = p->nnz;
  return CXG_OK;
}

const int64_t* cxg_pattern_row_ptr(const cxg_pattern* p) { return p ? p->row_ptr.p : nullptr; }
const int32_t* cxg_pattern_col_idx(const cxg_pattern* p) { return p ? p->col_idx.p : nullptr; }

cxg_status cxg_pattern_copy(cxg_context* ctx, const cxg_pattern* p, int64_t* row_ptr,
                            int32_t* col_idx) {
  if (!p) return cxg::fail(ctx, CXG_ERR_ARG, "null pattern");
  return cxg::guarded(ctx, [&] {
    if (row_ptr)
      CXG_HIP(hipMemcpyAsync(row_ptr, p->row_ptr.p, size_t(p->n_free + 1) * sizeof(int64_t),
                             hipMemcpyDeviceToDevice, ctx->stream));
    if (col_idx && p->nnz)
      CXG_HIP(hipMemcpyAsync(col_idx, p->col_idx.p, size_t(p->nnz) * sizeof(int32_t),
                             hipMemcpyDeviceToDevice, ctx->stream));
  });
}

cxg_status cxg_pattern_compact(cxg_context* ctx, const cxg_pattern* p, const uint8_t* live,
                               const double* vals, int64_t* row_ptr_out, int32_t* col_idx_out,
                               double* vals_out, int64_t* nnz_out) {
  if (!p || !live || !vals || !row_ptr_out || !col_idx_out || !vals_out || !nnz_out)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad pattern_compact arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  return cxg::guarded(ctx, [&] {
    *nnz_out = cxg::compact_pattern(ctx, p, live, vals, row_ptr_out, col_idx_out, vals_out);
  });
}

cxg_status cxg_assemble_linear(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                               const double* E, const double* nu, double* vals, const double* u,
                               double* f_int, uint8_t* live) {
  if (!p || !coords || !E || !nu || !vals) return cxg::fail(ctx, CXG_ERR_ARG, "bad assemble arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  if (live && ctx->asm_variant < 4)
    return cxg::fail(ctx, CXG_ERR_UNSUPPORTED, "entry liveness needs the two-phase assembly (asm_variant 4 or 5)");
  const bool with_f = u && f_int;
  return cxg::guarded(ctx, [&] {
    cxg::launch_assemble_linear(ctx, p, coords, E, nu, vals, with_f ? u : nullptr,
                                with_f ? f_int : nullptr, live);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_assemble_linear_cases(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                     const double* E, const double* nu, double* vals,
                                     int32_t n_cases, const double* u, double* f_int,
                                     uint8_t* live) {
  if (!p || !coords || !E || !nu || !vals || !u || !f_int)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad assemble arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  if (live && ctx->asm_variant < 4)
    return cxg::fail(ctx, CXG_ERR_UNSUPPORTED, "entry liveness needs the two-phase assembly (asm_variant 4 or 5)");
  return cxg::guarded(ctx, [&] {
    cxg::launch_assemble_linear(ctx, p, coords, E, nu, vals, u, f_int, live, int(n_cases));
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_internal_force(cxg_context* ctx, const cxg_mesh* mesh, const double* coords,
                              const double* E, const double* nu, const double* u, double* f_int) {
  if (!mesh || !coords || !E || !nu || !u || !f_int) return cxg::fail(ctx, CXG_ERR_ARG, "bad internal_force arguments");
  return cxg::guarded(ctx, [&] {
    cxg::launch_internal_force(ctx, mesh, coords, E, nu, u, f_int);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_internal_force_noalloc(cxg_context* ctx, const cxg_mesh* mesh, const double* coords,
                                      const double* E, const double* nu, const double* u,
                                      double* f_int) {
  if (!mesh || !coords || !E || !nu || !u || !f_int)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad internal_force_noalloc arguments");
  if (mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "mesh belongs to another context");
  return cxg::guarded(ctx, [&] {
    cxg::launch_internal_force_noalloc(ctx, mesh, coords, E, nu, u, f_int);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_reduce_free(cxg_context* ctx, int32_t n_nodes, const int32_t* dof_eq,
                           const double* f_ext, const double* f_c, double* rhs) {
  if (!dof_eq || !rhs || n_nodes < 0) return cxg::fail(ctx, CXG_ERR_ARG, "bad reduce_free arguments");
  return cxg::guarded(ctx, [&] { cxg::launch_reduce_free(ctx, n_nodes, dof_eq, f_ext, f_c, rhs); });
}

cxg_status cxg_assemble_j2(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                           const double* E, const double* nu, const int32_t* mat_id,
                           const cxg_j2_material* mats, int32_t n_mats, const double* u,
                           const double* committed, double* trial, double* vals,
                           double* f_int, uint8_t* live) {
  if (!p || !coords || !E || !nu || !u || !committed || !trial || !vals || !f_int ||
      (n_mats > 0 && (!mats || !mat_id)) || n_mats < 0)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad assemble_j2 arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  if (live && ctx->asm_variant < 4)
    return cxg::fail(ctx, CXG_ERR_UNSUPPORTED, "entry liveness needs the two-phase assembly (asm_variant 4 or 5)");
  return cxg::guarded(ctx, [&] {
    const void* table = cxg::upload_j2_table(ctx, mats, n_mats);
    cxg::launch_assemble_j2(ctx, p, coords, E, nu, n_mats > 0 ? mat_id : nullptr, table, u,
                            committed, trial, vals, f_int, live);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_assemble_mass(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                             const double* rho, int32_t lumped, double* vals, uint8_t* live) {
  if (!p || !coords || !rho || !vals) return cxg::fail(ctx, CXG_ERR_ARG, "bad assemble_mass arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  return cxg::guarded(ctx, [&] {
    cxg::launch_assemble_scalar(ctx, p, coords, lumped ? 1 : 0, rho, nullptr, vals, live);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_pattern_same_direction(cxg_context* ctx, const cxg_pattern* p, int64_t* row_ptr,
                                      int32_t* col_idx, int64_t* nnz_out) {
  if (!p || !nnz_out) return cxg::fail(ctx, CXG_ERR_ARG, "bad pattern_same_direction arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  if (!cxg::same_direction_fits(ctx, p))
    return cxg::fail(ctx, CXG_ERR_UNSUPPORTED,
                     "same-direction rows: a CSR row is longer than the gather's LDS columns");
  return cxg::guarded(ctx, [&] {
    cxg::prepare_same_direction(ctx, p);
    *nnz_out = p->sd_nnz;
    if (row_ptr)
      CXG_HIP(hipMemcpyAsync(row_ptr, p->sd_row_ptr.p, size_t(p->n_free + 1) * sizeof(int64_t),
                             hipMemcpyDeviceToDevice, ctx->stream));
    if (col_idx && p->sd_nnz > 0)
      CXG_HIP(hipMemcpyAsync(col_idx, p->sd_col_idx.p, size_t(p->sd_nnz) * sizeof(int32_t),
                             hipMemcpyDeviceToDevice, ctx->stream));
  });
}

cxg_status cxg_assemble_mass_sd(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                const double* rho, int32_t lumped, double* vals, uint8_t* live) {
  if (!p || !coords || !rho || !vals || !live)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad assemble_mass_sd arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  return cxg::guarded(ctx, [&] {
    cxg::launch_assemble_scalar(ctx, p, coords, lumped ? 1 : 0, rho, nullptr, vals, live, true);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_assemble_geometric_sd(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                     const double* gp_stress, double* vals, uint8_t* live) {
  if (!p || !coords || !gp_stress || !vals || !live)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad assemble_geometric_sd arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  return cxg::guarded(ctx, [&] {
    cxg::launch_assemble_scalar(ctx, p, coords, 2, nullptr, gp_stress, vals, live, true);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_compact_csr(cxg_context* ctx, int32_t n_rows, const int64_t* row_ptr,
                           const int32_t* col_idx, const uint8_t* live, const double* vals,
                           int64_t* row_ptr_out, int32_t* col_idx_out, double* vals_out,
                           int64_t* nnz_out) {
  if (n_rows < 0 || !row_ptr || !col_idx || !live || !vals || !row_ptr_out || !col_idx_out ||
      !vals_out || !nnz_out)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad compact_csr arguments");
  return cxg::guarded(ctx, [&] {
    *nnz_out = cxg::compact_csr(ctx, n_rows, row_ptr, col_idx, live, vals, row_ptr_out,
                                col_idx_out, vals_out);
  });
}

cxg_status cxg_assemble_geometric(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                  const double* gp_stress, double* vals, uint8_t* live) {
  if (!p || !coords || !gp_stress || !vals)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad assemble_geometric arguments");
  if (p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  return cxg::guarded(ctx, [&] {
    cxg::launch_assemble_scalar(ctx, p, coords, 2, nullptr, gp_stress, vals, live);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_recover_fields(cxg_context* ctx, const cxg_mesh* mesh, const double* coords,
                              const double* E, const double* nu, const int32_t* mat_id,
                              const cxg_j2_material* mats, int32_t n_mats, const double* u,
                              const double* committed, const double* f_ext, double* strain,
                              double* stress, double* reaction, double* f_int,
                              double* gauss_stress) {
  if (!mesh || !coords || !E || !nu || !u || n_mats < 0 ||
      (n_mats > 0 && (!mats || !mat_id || !committed)))
    return cxg::fail(ctx, CXG_ERR_ARG, "bad recover_fields arguments");
  if (mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "mesh belongs to another context");
  return cxg::guarded(ctx, [&] {
    const void* table = n_mats > 0 ? cxg::upload_j2_table(ctx, mats, n_mats) : nullptr;
    cxg::launch_recover(ctx, mesh, coords, E, nu, u, n_mats > 0 ? mat_id : nullptr, table,
                        committed, f_ext, strain, stress, reaction, f_int, gauss_stress);
    cxg::check_flags(ctx);
  });
}

cxg_status cxg_material_evaluate(cxg_context* ctx, int64_t n_points, const int32_t* mat_id,
                                 const cxg_j2_material* mats, int32_t n_mats, const double* E,
                                 const double* nu, const double* strain, const double* committed,
                                 double* trial, double* stress, double* tangent) {
  if (n_points < 0 || n_mats < 0 || (n_points > 0 && (!strain || !stress || !tangent)) ||
      (mat_id && (n_mats == 0 || !mats || !committed || !trial)) || (!mat_id && (!E || !nu)))
    return cxg::fail(ctx, CXG_ERR_ARG, "bad material_evaluate arguments");
  return cxg::guarded(ctx, [&] {
    const void* table = mat_id ? cxg::upload_j2_table(ctx, mats, n_mats) : nullptr;
    cxg::launch_material_points(ctx, n_points, mat_id, table, E, nu, strain, committed, trial,
                                stress, tangent);
    CXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

cxg_status cxg_commit(cxg_context* ctx, int64_t n_points, double* committed, const double* trial) {
  if (!committed || !trial || n_points < 0) return cxg::fail(ctx, CXG_ERR_ARG, "bad commit arguments");
  return cxg::guarded(ctx, [&] { cxg::launch_commit(ctx, n_points, committed, trial); });
}

cxg_status cxg_vjp(cxg_context* ctx, const cxg_mesh* mesh, const double* coords, const double* E,
                   const double* nu, const double* u, const double* w, double* dx, double* dE,
                   double* du) {
  if (!mesh || !coords || !E || !nu || !u || !w || !dx || !dE || !du)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad vjp arguments");
  return cxg::guarded(ctx, [&] {
    cxg::launch_vjp(ctx, mesh, coords, E, nu, u, w, dx, dE, du);
    cxg::check_flags(ctx);
  });
}

}  // extern "C"
