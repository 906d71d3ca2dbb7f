// Internal declarations of libcxg.so (not installed; the public ABI is include/cxg.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "cxg.h"

namespace cxg {

// Doubles in the fixed device scratch array of the allocation-free entry point
// (64 MiB; CXG_NOALLOC_SCRATCH_DOUBLES in cxg.h).
constexpr int64_t kScratchDoubles = CXG_NOALLOC_SCRATCH_DOUBLES;

// Error carried from the internals to the C ABI, where it becomes a status code and
// the context's last-error text. Never crosses the ABI as an exception.
struct Error : std::runtime_error {
  cxg_status status;
  Error(cxg_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define CXG_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::cxg::Error(CXG_ERR_HIP, std::string(#call " failed: ") +                 \
                                          hipGetErrorString(e_));                      \
  } while (0)

// The reference's message for a non-positive Jacobian (src/fem/element.cpp:87).
inline const char* kBadJacobian = "non-positive element Jacobian (bad connectivity?)";

// Bits of the device status word every kernel may raise.
enum : int { kFlagJacobian = 1, kFlagValence = 2, kFlagPattern = 4 };
// slot_all entries: bits 0..27 the column offset, bits 28..30 the free-DOF mask of the
// column node (bit c: DOF c is free)
constexpr int32_t kSlotMask = (1 << 28) - 1;
constexpr int kSlotFreeShift = 28;

// Shape tables of one element family, resident in device memory:
// dN[q][k][j] = dN_k/dxi_j at Gauss point q, the Gauss weights w[q] and the
// shape-function values N[q][k].
struct ElemTable {
  int nn = 0, ngp = 0;
  const double* dN = nullptr;  // device, ngp*nn*3 point-major, then its transpose (nn*3*ngp)
  const double* w = nullptr;   // device, ngp
  const double* N = nullptr;   // device, ngp*nn
};

int nodes_per_element(int type);  // 0 for types outside the GPU path
int gauss_points(int type);       // 0 for types outside the GPU path
// Host tables (natural coordinates of the Gauss points and derivatives).
void host_tables(int type, std::vector<double>& dN, std::vector<double>& w,
                 std::vector<double>& N);

// The C3D8 natural derivatives at its 2 x 2 x 2 Gauss points in closed form. Point q has
// coordinate +-g along axis a (+ when bit a of q is set; g = 1.0 / std::sqrt(3.0) as the
// Gauss rule computes it, one ulp above the nearest double to 1/sqrt(3)), node k the corner
// sign s_ka (CalculiX order), so every factor 1 + s_ka p_a of dN_k/dxi_c = s_kc (1 + s_ka
// p_a)(1 + s_kb p_b) / 8 is 1 + g or 1 - g. hex8_dn rounds as the table does (one rounded
// product; the sign and the 1/8 exact), and host_tables checks the table against it, so
// the kernels that form these values instead of reading the table get the table's bits.
constexpr double kGauss2 = 0x1.279a74590331dp-1;
__host__ __device__ constexpr bool hex_corner_pos(int k, int a) {
  return a == 0 ? ((k ^ (k >> 1)) & 1) != 0 : a == 1 ? ((k >> 1) & 1) != 0 : ((k >> 2) & 1) != 0;
}
__host__ __device__ inline double hex8_dn(int q, int k, int c) {
  double f[3];
  for (int a = 0; a < 3; ++a)
    f[a] = hex_corner_pos(k, a) == (((q >> a) & 1) != 0) ? 1.0 + kGauss2 : 1.0 - kGauss2;
  const int a = c == 0 ? 1 : 0, b = c == 2 ? 1 : 2;
  return ((hex_corner_pos(k, c) ? 0.125 : -0.125) * f[a]) * f[b];
}

// Device buffer owning helper (hipMalloc/hipFree), move-only.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void alloc(size_t count) {
    reset();
    n = count;
    if (count) CXG_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Grow to at least `count` elements (contents are not kept).
  void ensure(size_t count) {
    if (n < count) alloc(count);
  }
};

}  // namespace cxg

struct cxg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int* flag_dev = nullptr;   // device status word (cxg::kFlag*)
  int* flag_host = nullptr;  // pinned mirror
  cxg::DevBuf<double> tab_dN[8], tab_w[8], tab_N[8];
  cxg::ElemTable tables[8];
  std::vector<unsigned char> j2_host;  // last uploaded J2 material table (bytes)
  cxg::DevBuf<unsigned char> j2_dev;
  // Kernel timing (cxg_profile_*): events bracket each launch while enabled.
  struct Pending {
    std::string name;
    hipEvent_t a, b;
    int64_t elements;
  };
  struct Stat {
    std::string name;
    int64_t launches = 0, elements = 0;
    double ms = 0.0;
  };
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<Pending> pending;
  std::vector<Stat> stats;
  // Tunables (cxg_set_option): elements per patch for meshes created afterwards,
  // and the stiffness-kernel variant.
  int patch_size = 8;
  int asm_variant = 7;
  // Persistent element kernels (asm_variant 5): elements per work item, and resident
  // waves per compute unit of the persistent grid.
  int chunk_elems = 4;
  int persistent_waves = 8;
  int n_cu = 256;  // compute units of the device (hipDeviceAttributeMultiprocessorCount)
  int xcd_remap = 1;  // two-phase kernels: one contiguous work range per XCD
  int gather_lds_cols = 256;  // longest row the row gather accumulates in LDS
  // asm_variant 7: columns of gather_tile x gather_tile elements in the gather's node
  // order (0: node index order)
  int gather_tile = 8;
  // asm_variant 7: element families of at least this many nodes store the upper node-pair
  // blocks, the others the full matrix (the gather reads each element in its layout)
  int up_min_nodes = 0;
  // C3D20/C3D20R element kernels (asm_variant 4, 7): lane -> node-pair tiling (0: a node
  // and four consecutive partners, 1: 2x2 pairs of two node couples)
  int elem_tiling = 1;
  // asm_variant 8: target size of one ring slot of the staged assembly, in MiB
  int stage_mb = 64;
  // asm_variant 8: 1 = gathers on a second stream, overlapping later stages' element
  // kernels; 0 = every launch in order on the context's stream
  int stage_overlap = 1;
  // asm_variant 7: 1 = the C3D20R elements of a mesh without C3D20 keep their scaled
  // gradients instead of their upper blocks, and the gather forms K_ab
  int gradient_form = 1;
  std::unordered_set<void*> allocs;  // cxg_malloc blocks, released by cxg_destroy
  double* scratch = nullptr;  // device address of the fixed scratch array on `device`
};

namespace cxg {
// Records a start event on construction and an end event on destruction when the
// context's timing is on; results are collected by collect_timing().
struct LaunchTimer {
  cxg_context* ctx;
  cxg_context::Pending p{};
  bool on;
  hipStream_t stream;
  LaunchTimer(cxg_context* c, std::string name, int64_t elements);
  // the same on another stream of the context's device
  LaunchTimer(cxg_context* c, std::string name, int64_t elements, hipStream_t s);
  ~LaunchTimer();
};
void collect_timing(cxg_context* ctx);  // after a stream synchronization
}  // namespace cxg

// One element family run inside a mesh. Elements are grouped into patches of up to
// patch_size consecutive elements; patches are coloured so that two patches of one
// colour share no node. A workgroup walks one patch element by element, so repeated
// updates of the same CSR rows by neighbouring elements happen back to back (cache
// reuse), while patches of one colour run concurrently without conflicts.
struct cxg_block_dev {
  int32_t type = 0, nn = 0, ngp = 0;
  int64_t n_elem = 0, elem_offset = 0, gp_offset = 0;
  int64_t conn_offset = 0;                     // into cxg_mesh::conn (entries)
  std::vector<cxg::DevBuf<int32_t>> color;     // per colour: block-local element ids, by patch
  std::vector<cxg::DevBuf<int32_t>> patch_ptr; // per colour: patch p = color[ptr[p]..ptr[p+1])
  std::vector<int64_t> color_count;            // elements per colour
  std::vector<int64_t> patch_count;            // patches per colour
};

struct cxg_mesh {
  cxg_context* ctx = nullptr;
  int32_t n_nodes = 0;
  int64_t n_elem = 0, n_points = 0;
  int32_t n_colors = 0;
  std::vector<cxg_block_dev> blocks;
  cxg::DevBuf<int32_t> conn;      // all blocks, concatenated
  cxg::DevBuf<int64_t> eptr;      // model element -> offset into conn (n_elem+1)
  cxg::DevBuf<int64_t> inc_ptr;   // node -> incident elements (n_nodes+1)
  cxg::DevBuf<int32_t> inc_elem;  // incident model element numbers, ascending
  cxg::DevBuf<int32_t> inc_pos;   // per conn offset: its place in the node's incidence list
                                  // (an element lists a node once; create_mesh checks)
  int32_t max_candidates = 0;     // max over nodes of sum of incident element sizes
  // Field recovery arenas, allocated on first use: per element the Gauss-point means of
  // strain and stress (12), per element node its force vector (3, node-major at
  // 3 * incidence, as the matrix-free internal force and the adjoint put theirs).
  mutable cxg::DevBuf<double> rec_mean;
  mutable cxg::DevBuf<double> rec_fe;
  // The adjoint's per-element dx vectors (3 per connectivity entry, node-major; du goes to
  // rec_fe).
  mutable cxg::DevBuf<double> vjp_dx;
};

namespace cxg {
// A work item of the persistent element kernels: `count` consecutive elements of
// block `block`, starting at block-local element `first`.
struct WorkChunk {
  int32_t block;
  int32_t first;
  int32_t count;
  int32_t pad;
};
// What the persistent kernels need of one block.
struct BlockDesc {
  int32_t type, nn, ngp, pad;
  int64_t elem_offset, gp_offset, conn_offset, kp_base;
  const double* dN;
  const double* w;
};
// One (node, incident element) pair of the two-phase gather: where the node's three
// rows of the element matrix start in the arena, where its slot row starts, the
// element's connectivity offset and node count.
struct IncRecord {
  int64_t rows;   // arena offset of row 3*la of the element matrix (full layout; the
                  // element's slot starts 9*la*n doubles earlier)
  int64_t slot;   // offset of slot row la of the element
  int64_t conn;   // offset of the element in the mesh connectivity
  int32_t n;      // nodes of the element
  int32_t la;     // local index of the node inside the element
};
// What the row gather needs of one node, per gather position (built once per pattern
// and gather order by k_gather_desc), so that a wave starts from one scalar load
// instead of a chain order -> dof_eq -> row_ptr / inc_ptr.
struct GatherDesc {
  int64_t row0;   // CSR offset of the node's first free row (its rows are consecutive
                  // equations of len columns each); unused when no DOF is free
  int64_t inc0;   // first incidence record of the node
  int32_t len;    // columns per row
  int32_t node;   // node index
  int32_t inc_n;  // incident elements
  int32_t flags;  // bits 0..2: DOF c free; bit 3: every incident element has <= 10 nodes
};
// Staged two-phase assembly (asm_variant 8). The elements are cut into S stages of
// consecutive element numbers; a node is complete once the stage of its last incident
// element has run, and its rows are gathered right after that stage, on a second stream,
// while the next stages' element matrices are formed. The element matrices live in a
// ring of R slots (R = W + 1, W the most stages one node's incident elements span), so
// the arena that a gather reads was written moments before and is still cache-resident.
struct StagePlan {
  int S = 0, R = 0, W = 0;
  int64_t slot_cap = 0;             // doubles per ring slot
  std::vector<int64_t> elem_first;  // S+1 element boundaries (model element numbers)
  std::vector<int64_t> kp_host;     // the pattern's kpptr (n_elem+1), on the host
  std::vector<int64_t> node_off;    // S+1 offsets into desc: the nodes completed by stage s
  DevBuf<IncRecord> rec;            // incidence records addressing the ring
  DevBuf<GatherDesc> desc;          // gather positions grouped by completion stage
  DevBuf<double> ring;              // R * slot_cap
  hipStream_t stream = nullptr;     // the gathers' stream
  std::vector<hipEvent_t> ev_elem, ev_gather;
  StagePlan() = default;
  StagePlan(const StagePlan&) = delete;
  StagePlan& operator=(const StagePlan&) = delete;
  ~StagePlan() {
    for (hipEvent_t e : ev_elem) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_gather) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};
}  // namespace cxg

struct cxg_pattern {
  const cxg_mesh* mesh = nullptr;
  int32_t n_free = 0;
  int64_t nnz = 0;
  int64_t max_row = 0;  // longest row (columns of one node's rows)
  cxg::DevBuf<int32_t> dof_eq;                // 3*n_nodes
  cxg::DevBuf<int64_t> row_ptr;               // n_free+1
  cxg::DevBuf<int32_t> col_idx;               // nnz
  // slot_all[sptr[e] + a*n + b]: offset of node b's column block inside the CSR rows of
  // node a (bits 0..27) and node b's free-DOF mask (bits 28..30), for element e with n
  // nodes (a, b local). slot_base[k] = sptr of block k's
  // first element.
  cxg::DevBuf<int32_t> slot_all;
  cxg::DevBuf<int64_t> sptr;                  // n_elem+1
  // The same-direction pattern (mass, geometric stiffness: k_ab I3 blocks): row (a, i)
  // keeps the structural columns (b, i), in the same order; built on first use
  // (sd_nnz -1 before).
  cxg::DevBuf<int64_t> sd_row_ptr;            // n_free+1
  cxg::DevBuf<int32_t> sd_col_idx;            // sd_nnz
  int64_t sd_nnz = -1;
  std::vector<int64_t> slot_base;
  int64_t slot_total = 0;                     // sum over elements of n^2
  // Two-phase assembly arena, allocated on first use: a slot of 9n^2 doubles per element
  // at kpptr[e], holding its full Ke (3n x 3n, row-major), its packed upper node-pair
  // blocks (asm_variant 7: the first 9n(n+1)/2 doubles), or its n^2 scalars k_ab (mass,
  // geometric stiffness: at kpptr[e] / 9).
  cxg::DevBuf<double> ke_arena;
  cxg::DevBuf<double> fe_arena;               // element force vectors (3n at 3*eptr[e])
  // asm_variant 7 with f_int: each element's nodal displacements of every load case,
  // 3 per connectivity entry, case c at c * 3 * conn.n
  cxg::DevBuf<double> ue_arena;
  cxg::DevBuf<int64_t> kpptr;                 // n_elem+1, in doubles
  cxg::DevBuf<cxg::IncRecord> inc_rec;        // per mesh incidence entry (inc_elem order)
  cxg::DevBuf<int8_t> nfm;                    // per node: bit c set if DOF c is free
  std::vector<int64_t> kp_base;               // kpptr of each block's first element
  int64_t kp_total = 0;
  // Work queue of the persistent element kernels, built for work_chunk elements per
  // item (0: not built), costliest family first; queue is the shared item counter.
  cxg::DevBuf<cxg::WorkChunk> work;
  cxg::DevBuf<cxg::BlockDesc> bdesc;
  cxg::DevBuf<int32_t> queue;
  int32_t n_work = 0;
  int32_t work_chunk = 0;
  // Node order of the upper-block gather (asm_variant 7), built for gorder_tile (-1: not
  // built; empty for tile 0, node index order).
  cxg::DevBuf<int32_t> gorder;
  int gorder_tile = -1;
  // Gather descriptors in the gather order: gorder when built with a tile > 0, else
  // node index order (gdesc_key: that tile, -1 for index order, -2 not built).
  cxg::DevBuf<cxg::GatherDesc> gdesc;
  int gdesc_key = -2;
  // asm_variant 8: the stage plan (null: not built), built for the gather order of
  // tile stages_tile and a ring slot of stages_mb MiB; stages_off: staging does not pay
  // for this mesh (the ring would exceed its budget), variant 7 runs instead
  std::unique_ptr<cxg::StagePlan> stages;
  int stages_tile = -2, stages_mb = 0;
  bool stages_off = false;
  // Captured graphs (cxg_graph_capture_*) holding the addresses of the buffers above;
  // while any is alive those buffers are never freed or moved (cxg::guard_realloc).
  mutable int live_graphs = 0;
};

namespace cxg {
// Kernels launchers (cxg_elem.hip / cxg_pattern.hip).
void fill_f64(hipStream_t s, double* p, int64_t n, double v);
void launch_assemble_linear(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                            const double* E, const double* nu, double* vals, const double* u,
                            double* f_int, uint8_t* live, int n_cases = 1);
// arena: the element-force arena (3 doubles per connectivity entry); null: the mesh's own,
// allocated on first use.
void launch_internal_force(cxg_context* ctx, const cxg_mesh* m, const double* coords,
                           const double* E, const double* nu, const double* u, double* f_int,
                           double* arena = nullptr);
// The same into the library's fixed scratch array (no allocation; CXG_ERR_ARG beyond
// kScratchDoubles).
void launch_internal_force_noalloc(cxg_context* ctx, const cxg_mesh* m, const double* coords,
                                   const double* E, const double* nu, const double* u,
                                   double* f_int);
// Device address of the fixed scratch array on `device` (looked up once per device, with
// the device current).
double* scratch_device_ptr(int device);
void launch_reduce_free(cxg_context* ctx, int32_t n_nodes, const int32_t* dof_eq,
                        const double* f_ext, const double* f_c, double* rhs);
// J2 material tables: host bytes of the device form, and the context's cached upload.
std::vector<unsigned char> j2_table_bytes(const cxg_j2_material* mats, int32_t n_mats);
const void* upload_j2_table(cxg_context* ctx, const cxg_j2_material* mats, int32_t n_mats);
// Allocates the two-phase arenas of a pattern ahead of use (no-op when present), and
// for asm_variant 5 the work queue of the persistent kernels.
void prepare_arenas(const cxg_pattern* p, bool element_forces);
void prepare_work(cxg_context* ctx, const cxg_pattern* p);
// asm_variant 7 stores upper blocks for the families of at least up_min_nodes nodes; a
// mesh without such a family runs the full-matrix kernels and gather of asm_variant 4.
bool uses_upper_blocks(const cxg_context* ctx, const cxg_mesh* m);
// The gradient form of asm_variant 7 (option gradient_form): the Gauss points of the
// C3D20R elements (8) when the mesh has C3D20R and no C3D20 elements, else 0. The
// C3D20R elements then keep their scaled gradients, the others their usual layout.
int hform_stride(const cxg_context* ctx, const cxg_mesh* m);
// Whether the linear assembly of asm_variant 7 runs the ordered gather (upper blocks or
// gradient form): its node order and displacement arena are then prepared.
bool linear_upper_path(const cxg_context* ctx, const cxg_mesh* m);
// The per-element displacement arena of asm_variant 7 for n_cases fields (no-op when
// large enough); before capture, as it allocates.
void prepare_displacement_arena(const cxg_pattern* p, int n_cases);
// The upper-block gather's node order from the node coordinates (asm_variant 7; a
// no-op when built for the context's gather_tile). Synchronizes the stream.
void prepare_order(cxg_context* ctx, const cxg_pattern* p, const double* coords);
// The gather descriptors for the pattern's current gather order (a no-op when built for
// it); enqueues one kernel. Before capture, as it allocates.
void prepare_desc(cxg_context* ctx, const cxg_pattern* p);
// asm_variant 8: the stage plan for the current gather order and stage_mb (a no-op when
// built for them); false when staging does not pay for this mesh. Synchronizes.
bool prepare_stages(cxg_context* ctx, const cxg_pattern* p);
void launch_assemble_j2(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                        const double* E, const double* nu, const int32_t* mat_id,
                        const void* j2_table, const double* u, const double* committed,
                        double* trial, double* vals, double* f_int, uint8_t* live);
void launch_commit(cxg_context* ctx, int64_t n_points, double* committed, const double* trial);
// MaterialModel::evaluate at n independent points (cxg_material_evaluate).
void launch_material_points(cxg_context* ctx, int64_t n_points, const int32_t* mat_id,
                            const void* j2_table, const double* E, const double* nu,
                            const double* strain, const double* committed, double* trial,
                            double* stress, double* tangent);
// Mass (kind 0 consistent, 1 lumped; rho per element) or geometric stiffness (kind 2;
// sig per Gauss point) into the CSR values (K8 + the row gather). same_direction: vals
// and live (both required) on the pattern's same-direction pattern (prepared), which
// needs every row in LDS (max_row <= the gather's LDS columns).
void launch_assemble_scalar(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                            int kind, const double* rho, const double* sig, double* vals,
                            uint8_t* live, bool same_direction = false);
// Whether the same-direction gather fits this pattern (every row in LDS).
bool same_direction_fits(const cxg_context* ctx, const cxg_pattern* p);
// Field recovery (K7): Gauss-point strain/stress, node means, f_int and reactions.
void launch_recover(cxg_context* ctx, const cxg_mesh* m, const double* coords, const double* E,
                    const double* nu, const double* u, const int32_t* mat_id, const void* j2_table,
                    const double* committed, const double* f_ext, double* strain,
                    double* stress, double* reaction, double* f_int, double* gauss_stress);
void launch_vjp(cxg_context* ctx, const cxg_mesh* m, const double* coords, const double* E,
                const double* nu, const double* u, const double* w, double* dx, double* dE,
                double* du);
void build_pattern(cxg_context* ctx, const cxg_mesh* m, const int32_t* dof_eq, int32_t n_free,
                   cxg_pattern* out);
// Jacobi-preconditioned CG on a CSR (cxg_solve.hip).
void pcg(cxg_context* ctx, int32_t n, const int64_t* rp, const int32_t* ci, const double* va,
         const double* b, double* x, double tol, int32_t max_iter, int32_t& iters,
         double& rel_res);
// Keeps the entries with live != 0; returns the kept count (synchronizes the stream).
int64_t compact_pattern(cxg_context* ctx, const cxg_pattern* p, const uint8_t* live,
                        const double* vals, int64_t* row_ptr_out, int32_t* col_idx_out,
                        double* vals_out);
// The same, for any CSR of n_rows rows.
int64_t compact_csr(cxg_context* ctx, int32_t n_rows, const int64_t* row_ptr,
                    const int32_t* col_idx, const uint8_t* live, const double* vals,
                    int64_t* row_ptr_out, int32_t* col_idx_out, double* vals_out);
// Builds the pattern's same-direction pattern (no-op when built); synchronizes.
void prepare_same_direction(cxg_context* ctx, const cxg_pattern* p);
// Read back and clear the device status word; throws on a raised flag.
void check_flags(cxg_context* ctx);

// A captured graph replays with the addresses its pattern's arenas, gather order and work
// queue had at capture. `frees`: the caller is about to free or move one of them (`what`);
// refused while a graph of the pattern is alive.
inline void guard_realloc(const cxg_pattern* p, bool frees, const char* what) {
  if (frees && p->live_graphs > 0)
    throw Error(CXG_ERR_STATE, std::string("a captured graph uses this pattern's ") + what +
                                   "; destroy the pattern's graphs before a call that needs "
                                   "it rebuilt");
}

// Status + last-error text for the C ABI.
inline cxg_status fail(cxg_context* ctx, cxg_status st, const char* msg) {
  if (ctx) ctx->err = msg;
  return st;
}

// Runs f on the context's device; exceptions become status codes at the ABI.
template <class F>
cxg_status guarded(cxg_context* ctx, F&& f) {
  if (!ctx) return CXG_ERR_ARG;
  try {
    CXG_HIP(hipSetDevice(ctx->device));
    f();
    return CXG_OK;
  } catch (const Error& e) {
    return fail(ctx, e.status, e.what());
  } catch (const std::bad_alloc&) {
    return fail(ctx, CXG_ERR_HIP, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(ctx, CXG_ERR_HIP, e.what());
  }
}
}  // namespace cxg
