// HIP graphs of one assembly / one Newton iteration of the material-point path.
//
// A Newton loop of the reference (solve_nonlinear over assemble_material_tangent,
// src/fem/assembly.cpp:385-423) re-assembles K_t, the internal force and the
// residual at every iteration with the same mesh, pattern and buffers; only the
// contents of u change. Here that iteration's kernels (element kernels per family,
// the row gather, and optionally the residual reduction) are captured once into a
// HIP graph and replayed with one launch per iteration. The graph reads its inputs
// at replay time from the addresses given at capture, so the caller updates u (and
// the committed history, after a commit) in place.
//
// Capture rules kept here: everything that allocates, uploads or synchronizes (the
// two-phase arenas, the gather order and descriptors, the J2 material table) happens
// before capture begins; the graph
// holds kernel nodes only (buffers are cleared by fill kernels, never by memset
// nodes); capture runs on a private non-blocking stream of default priority; the
// device status word is read outside the graph, by cxg_check.
#include <memory>
#include <string>

#include "cxg_internal.hpp"

struct cxg_graph {
  cxg_context* ctx = nullptr;
  const cxg_pattern* pattern = nullptr;  // whose buffers the graph's kernels address
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  cxg::DevBuf<unsigned char> j2_table;  // the graph's own copy of the material table
  size_t nodes = 0;
  int64_t elements = 0;
  ~cxg_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (pattern) --pattern->live_graphs;
  }
};

namespace cxg {
namespace {

// Runs `enqueue` with the context's stream switched to a capturing stream and turns
// what it enqueued into an executable graph owned by g.
template <class F>
void capture(cxg_context* ctx, cxg_graph* g, F&& enqueue) {
  CXG_HIP(hipStreamSynchronize(ctx->stream));  // preparations are complete
  hipStream_t s = nullptr;
  CXG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipStream_t saved = ctx->stream;
  const bool timing = ctx->timing;
  ctx->stream = s;
  ctx->timing = false;
  std::string err;
  cxg_status code = CXG_OK;
  hipGraph_t graph = nullptr;
  hipError_t st = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (st != hipSuccess) {
    code = CXG_ERR_HIP;
    err = std::string("hipStreamBeginCapture failed: ") + hipGetErrorString(st);
  } else {
    try {
      enqueue();
    } catch (const Error& e) {
      code = e.status;
      err = e.what();
    }
    st = hipStreamEndCapture(s, &graph);
    if (st != hipSuccess && code == CXG_OK) {
      code = CXG_ERR_HIP;
      err = std::string("hipStreamEndCapture failed: ") + hipGetErrorString(st);
    }
  }
  ctx->stream = saved;
  ctx->timing = timing;
  (void)hipStreamDestroy(s);
  if (code != CXG_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    throw Error(code, err);
  }
  g->graph = graph;
  CXG_HIP(hipGraphGetNodes(graph, nullptr, &g->nodes));
  CXG_HIP(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
}

// From here on the pattern keeps the buffers the graph's kernels address.
void hold(cxg_graph* g, const cxg_pattern* p) {
  g->pattern = p;
  ++p->live_graphs;
}

}  // namespace
}  // namespace cxg

extern "C" {

cxg_status cxg_graph_capture_linear(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                    const double* E, const double* nu, double* vals,
                                    const double* u, double* f_int, cxg_graph** out) {
  if (!out) return cxg::fail(ctx, CXG_ERR_ARG, "null graph output");
  *out = nullptr;
  if (!p || !coords || !E || !nu || !vals) return cxg::fail(ctx, CXG_ERR_ARG, "bad capture arguments");
  if (ctx && p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  const bool with_f = u && f_int;
  return cxg::guarded(ctx, [&] {
    auto g = std::make_unique<cxg_graph>();
    g->ctx = ctx;
    g->elements = p->mesh->n_elem;
    if (ctx->asm_variant >= 4) cxg::prepare_arenas(p, false);
    if (ctx->asm_variant == 5) cxg::prepare_work(ctx, p);
    if (cxg::linear_upper_path(ctx, p->mesh)) {
      cxg::prepare_order(ctx, p, coords);
      if (with_f) cxg::prepare_displacement_arena(p, 1);
    }
    if (ctx->asm_variant >= 4) cxg::prepare_desc(ctx, p);
    cxg::capture(ctx, g.get(), [&] {
      cxg::launch_assemble_linear(ctx, p, coords, E, nu, vals, with_f ? u : nullptr,
                                  with_f ? f_int : nullptr, nullptr);
    });
    cxg::hold(g.get(), p);
    *out = g.release();
  });
}

cxg_status cxg_graph_capture_linear_cases(cxg_context* ctx, const cxg_pattern* p,
                                          const double* coords, const double* E,
                                          const double* nu, double* vals, int32_t n_cases,
                                          const double* u, double* f_int, cxg_graph** out) {
  if (!out) return cxg::fail(ctx, CXG_ERR_ARG, "null graph output");
  *out = nullptr;
  if (!p || !coords || !E || !nu || !vals || !u || !f_int)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad capture arguments");
  if (ctx && p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  return cxg::guarded(ctx, [&] {
    auto g = std::make_unique<cxg_graph>();
    g->ctx = ctx;
    g->elements = p->mesh->n_elem;
    if (ctx->asm_variant >= 4) cxg::prepare_arenas(p, false);
    if (ctx->asm_variant == 5) cxg::prepare_work(ctx, p);
    if (cxg::linear_upper_path(ctx, p->mesh)) {
      cxg::prepare_order(ctx, p, coords);
      cxg::prepare_displacement_arena(p, int(n_cases));
    }
    if (ctx->asm_variant >= 4) cxg::prepare_desc(ctx, p);
    cxg::capture(ctx, g.get(), [&] {
      cxg::launch_assemble_linear(ctx, p, coords, E, nu, vals, u, f_int, nullptr, int(n_cases));
    });
    cxg::hold(g.get(), p);
    *out = g.release();
  });
}

cxg_status cxg_graph_capture_j2(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                const double* E, const double* nu, const int32_t* mat_id,
                                const cxg_j2_material* mats, int32_t n_mats, const double* u,
                                const double* committed, double* trial, double* vals,
                                double* f_int, const double* f_ext, double* rhs,
                                cxg_graph** out) {
  if (!out) return cxg::fail(ctx, CXG_ERR_ARG, "null graph output");
  *out = nullptr;
  if (!p || !coords || !E || !nu || !u || !committed || !trial || !vals || !f_int ||
      (n_mats > 0 && (!mats || !mat_id)) || n_mats < 0)
    return cxg::fail(ctx, CXG_ERR_ARG, "bad capture arguments");
  if (ctx && p->mesh->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "pattern belongs to another context");
  return cxg::guarded(ctx, [&] {
    auto g = std::make_unique<cxg_graph>();
    g->ctx = ctx;
    g->elements = p->mesh->n_elem;
    const std::vector<unsigned char> table = cxg::j2_table_bytes(mats, n_mats);
    if (!table.empty()) {
      g->j2_table.alloc(table.size());
      CXG_HIP(hipMemcpy(g->j2_table.p, table.data(), table.size(), hipMemcpyHostToDevice));
    }
    if (ctx->asm_variant >= 4) cxg::prepare_arenas(p, true);
    if (ctx->asm_variant == 5) cxg::prepare_work(ctx, p);
    if (ctx->asm_variant >= 7 && cxg::uses_upper_blocks(ctx, p->mesh))
      cxg::prepare_order(ctx, p, coords);
    if (ctx->asm_variant >= 4) cxg::prepare_desc(ctx, p);
    const int32_t* mid = n_mats > 0 ? mat_id : nullptr;
    cxg::capture(ctx, g.get(), [&] {
      cxg::launch_assemble_j2(ctx, p, coords, E, nu, mid, g->j2_table.p, u, committed, trial,
                              vals, f_int, nullptr);
      if (rhs) {
        cxg::fill_f64(ctx->stream, rhs, p->n_free, 0.0);
        cxg::launch_reduce_free(ctx, p->mesh->n_nodes, p->dof_eq.p, f_ext, f_int, rhs);
      }
    });
    cxg::hold(g.get(), p);
    *out = g.release();
  });
}

cxg_status cxg_graph_launch(cxg_context* ctx, cxg_graph* g) {
  if (!g) return cxg::fail(ctx, CXG_ERR_ARG, "null graph");
  if (g->ctx != ctx) return cxg::fail(ctx, CXG_ERR_STATE, "graph belongs to another context");
  return cxg::guarded(ctx, [&] {
    cxg::LaunchTimer timer(ctx, "graph", g->elements);
    CXG_HIP(hipGraphLaunch(g->exec, ctx->stream));
  });
}

cxg_status cxg_graph_info(const cxg_graph* g, int64_t* nodes) {
  if (!g) return CXG_ERR_ARG;
  if (nodes) *nodes = int64_t(g->nodes);
  return CXG_OK;
}

cxg_status cxg_check(cxg_context* ctx) {
  return cxg::guarded(ctx, [&] { cxg::check_flags(ctx); });
}

void cxg_graph_destroy(cxg_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  delete g;
}

}  // extern "C"
