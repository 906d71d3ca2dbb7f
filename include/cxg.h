/*
 * cxg.h — C ABI of libcxg.so, the MI355X (gfx950) element-integration and
 * global-assembly library that stands in for CalculiX++'s fem:: hot path.
 *
 * Conventions (all from the reference):
 *  - element type codes are cxpp::ElementType's enumerator values
 *    (include/calculixpp/core/element.hpp:13-22); the GPU path implements C3D8,
 *    C3D20, C3D20R and C3D6, anything else returns CXG_ERR_UNSUPPORTED;
 *  - node order inside an element is CalculiX order (src/fem/element.cpp:160-253);
 *  - global DOF g = 3*node_index + comp (core/types.hpp:13); dof_eq[g] is the free
 *    equation of g or -1 when g is constrained (fem/assembly.hpp:25-27);
 *  - Voigt order xx,yy,zz,xy,xz,yz with engineering shear (core/results.hpp:9);
 *  - all floating point is FP64 (core/types.hpp:10).
 *
 * Memory: every array argument is a DEVICE pointer on the context's device unless
 * it is marked (host). Arrays are caller-owned; objects created here (mesh,
 * pattern) own their device memory and free it on destroy. Work is enqueued on the
 * context's stream; functions that must report a data-dependent error (a
 * non-positive Jacobian) synchronize that stream before returning.
 *
 * Errors: every function returns a cxg_status; the message of the last failure
 * on a context is cxg_last_error(ctx). Messages for conditions the reference
 * reports reuse the reference text (for example src/fem/element.cpp:87).
 * No C++ exception crosses this interface.
 *
 * Threading: one context per host thread; objects may be shared read-only.
 */
#ifndef CXG_H
#define CXG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  CXG_OK = 0,
  CXG_ERR_ARG = 1,          /* bad argument (null pointer, size, type)           */
  CXG_ERR_HIP = 2,          /* HIP runtime failure                               */
  CXG_ERR_JACOBIAN = 3,     /* det J <= 0 at a Gauss point (element.cpp:86-88)   */
  CXG_ERR_UNSUPPORTED = 4,  /* element type or mesh shape outside the GPU path   */
  CXG_ERR_STATE = 5         /* object used with the wrong mesh/pattern           */
} cxg_status;

enum {
  CXG_C3D4 = 0, CXG_C3D10 = 1, CXG_C3D8 = 2, CXG_C3D8R = 3,
  CXG_C3D20 = 4, CXG_C3D20R = 5, CXG_C3D6 = 6, CXG_C3D15 = 7
};

/* Hardening kinds, cxpp::Plastic::Hardening (core/material.hpp:27-28). */
enum { CXG_HARD_ISOTROPIC = 0, CXG_HARD_KINEMATIC = 1, CXG_HARD_COMBINED = 2 };

#define CXG_MAX_TABLE 16
/* Committed (or trial) J2 history per Gauss point, SoA: 13 planes of n_points
 * doubles: plastic strain [0..5] (engineering shear), equivalent plastic strain
 * [6], back stress [7..12] (fem/material_model.hpp:38-48). */
#define CXG_STATE_COMPONENTS 13

typedef struct cxg_context cxg_context;
typedef struct cxg_mesh cxg_mesh;
typedef struct cxg_pattern cxg_pattern;
typedef struct cxg_graph cxg_graph;

/* A run of elements of one type. Blocks of a mesh cover consecutive element
 * numbers: block k holds elements elem_offset .. elem_offset+n_elem-1 of the
 * model, and its Gauss points are gp_offset .. gp_offset + n_elem*ngp - 1. */
typedef struct {
  int32_t type;            /* CXG_C3D8 ... */
  int64_t n_elem;
  int64_t elem_offset;     /* first model-wide element number of the block */
  int64_t gp_offset;       /* first model-wide Gauss-point number of the block */
  const int32_t* conn;     /* device, n_elem * nodes_per_element node indices */
} cxg_block;

/* One J2 material (core/material.hpp:24-44 + ElasticIso), passed by host pointer. */
typedef struct {
  double E, nu;
  int32_t hardening;       /* CXG_HARD_* */
  int32_t n_table;         /* 1..CXG_MAX_TABLE points of the *PLASTIC table */
  double eqplastic[CXG_MAX_TABLE];
  double yield[CXG_MAX_TABLE];
  double kinematic_modulus;
} cxg_j2_material;

/* ---- context ------------------------------------------------------------- */
/* Create a context on HIP device `device`, enqueueing on `stream` (a
 * hipStream_t; NULL = the null stream). No counterpart in the reference, which is
 * single-threaded host code (src/compute/cpu_backend.cpp:108-111). */
cxg_status cxg_create(int device, void* stream, cxg_context** out);
void cxg_destroy(cxg_context* ctx);
cxg_status cxg_set_stream(cxg_context* ctx, void* stream);
const char* cxg_last_error(const cxg_context* ctx);
/* Tunables of the context (no counterpart in the reference):
 *   "patch_size"  elements per patch of meshes created afterwards (>= 1, default 8)
 *   "asm_variant" linear stiffness kernel: 1 = per-pair scatter, 2 = M-identity with
 *                 row-coalesced scatter at 4 waves/SIMD, 3 = the same at 3, 4 =
 *                 two-phase: element matrices to an arena, then each CSR row gathered
 *                 and written once by its node, 5 = two-phase with one persistent
 *                 element launch over all families, fed by a work queue, 6 = two-phase
 *                 with the element matrices on the FP64 matrix cores (M = H^T H by
 *                 MFMA), 7 = two-phase with only the upper node-pair blocks in the
 *                 arena, the lower ones read transposed by the gather, which takes the
 *                 nodes in a locality order built from the coordinates at the first
 *                 assembly of a pattern (default; same values as 4 bit for bit)
 *   "gather_tile" asm_variant 7: the locality order groups nodes into columns of
 *                 gather_tile x gather_tile elements across the two shorter extents,
 *                 swept along the longest (0: node index order; 0..4096, default 8)
 *   "up_min_nodes" asm_variant 7: element families of at least this many nodes store
 *                 the upper node-pair blocks, the others their full matrix; the gather
 *                 reads each element in its own layout (0..64, default 0: every family)
 *   "gradient_form" asm_variant 7, meshes with C3D20R and without C3D20 elements: 1 =
 *                 each C3D20R element keeps lambda, mu and its scaled shape-function
 *                 gradients at its 8 points (482 doubles) instead of its upper blocks
 *                 (1890), and the row gather forms K_ab from them; the other families
 *                 keep their layout; same values bit for bit (0/1, default 1)
 *   "elem_tiling" 20-node element kernels of asm_variant 4 and 7: 0 = a lane takes a
 *                 node and four consecutive partners, 1 = the 2x2 pairs of two node
 *                 couples with 16-byte LDS reads (default 1; same values bit for bit)
 *   "chunk_elems"      work item of the persistent launch, in groups of the J2 kernel's
 *                      elements per wave (8 C3D8, 2 C3D20, 4 C3D20R, 24 C3D6) (1..4096, 4)
 *   "persistent_waves" resident waves per compute unit of that launch (1..32, 8)
 *   "xcd_remap"   1: the two-phase kernels give each of the 8 XCDs one contiguous range
 *                 of nodes / elements (its L2 then sees neighbouring data again); 0: the
 *                 plain block order (default 1)
 *   "gather_lds_cols" longest CSR row the two-phase row gather accumulates in LDS; longer
 *                 rows accumulate in the value array itself (8..256, default 256)
 * Returns CXG_ERR_ARG for an unknown key or value. */
cxg_status cxg_set_option(cxg_context* ctx, const char* key, int64_t value);
/* Device memory owned by the context, for callers without a device allocator of
 * their own (host-language bindings). Blocks still held at cxg_destroy are released
 * there. cxg_copy_to_device / cxg_copy_to_host enqueue on the context's stream and
 * return when the copy is complete. No counterpart in the reference (host memory
 * only); the reference's buffers map onto these one to one. */
cxg_status cxg_malloc(cxg_context* ctx, size_t bytes, void** out);
cxg_status cxg_free(cxg_context* ctx, void* p);
cxg_status cxg_copy_to_device(cxg_context* ctx, void* dst, const void* src, size_t bytes);
cxg_status cxg_copy_to_host(cxg_context* ctx, void* dst, const void* src, size_t bytes);
/* Library version string, e.g. "cxg 0.1 gfx950". */
const char* cxg_version(void);
int32_t cxg_nodes_per_element(int32_t type);   /* core/element.hpp:24-42 */
int32_t cxg_gauss_points(int32_t type);        /* fem/element.cpp:106-126 */

/* ---- mesh: element blocks, node incidence and a conflict-free element colouring.
 * Replaces the per-element node-id gather of fem/assembly.cpp:46-59: connectivity
 * is copied to library memory once. The colouring (no two elements of one colour
 * share a node) is what makes every scatter below deterministic without atomics. */
cxg_status cxg_mesh_create(cxg_context* ctx, int32_t n_nodes, int32_t n_blocks,
                           const cxg_block* blocks /* host array */, cxg_mesh** out);
void cxg_mesh_destroy(cxg_mesh* mesh);
/* n_elem, n_gauss_points, n_colors (any may be NULL). */
cxg_status cxg_mesh_info(const cxg_mesh* mesh, int64_t* n_elem, int64_t* n_points,
                         int32_t* n_colors);

/* ---- K3: CSR pattern + element -> CSR slot map ------------------------------
 * Replaces build_dof_map (fem/assembly.cpp:19-42) + the COO keys of scatter_block
 * (fem/assembly.cpp:76-100) + SciPP CsrMatrix::from_coo (src/compute/
 * cpu_backend.cpp:46). Rows are free equations in ascending order, columns
 * ascending within a row; every structural entry (free DOF pair of two nodes that
 * share an element) is present, also when its value is zero.
 * dof_eq: device, 3*n_nodes (free equation or -1). */
cxg_status cxg_pattern_build(cxg_context* ctx, const cxg_mesh* mesh, const int32_t* dof_eq,
                             int32_t n_free, cxg_pattern** out);
void cxg_pattern_destroy(cxg_pattern* pattern);
cxg_status cxg_pattern_info(const cxg_pattern* p, int32_t* n_free, int64_t* nnz);
/* Device pointers owned by the pattern: row_ptr[n_free+1] (int64), col_idx[nnz]. */
const int64_t* cxg_pattern_row_ptr(const cxg_pattern* p);
const int32_t* cxg_pattern_col_idx(const cxg_pattern* p);
/* Copy row_ptr / col_idx into caller buffers (device pointers; either may be NULL),
 * enqueued on the context's stream. */
cxg_status cxg_pattern_copy(cxg_context* ctx, const cxg_pattern* p, int64_t* row_ptr,
                            int32_t* col_idx);

/* The reference's key set from a structural system: keeps the entries with live[k] != 0
 * (see cxg_assemble_linear), in CSR order. Replaces the COO -> CSR step of the
 * reference (scipp::sparse::CsrMatrix::from_coo over the map scatter_block fills,
 * fem/assembly.cpp:76-100). Device outputs row_ptr_out[n_free+1], col_idx_out and
 * vals_out with room for the structural nnz; *nnz_out (host) receives the kept count.
 * Synchronizes the context's stream. */
cxg_status cxg_pattern_compact(cxg_context* ctx, const cxg_pattern* p, const uint8_t* live,
                               const double* vals, int64_t* row_ptr_out,
                               int32_t* col_idx_out, double* vals_out, int64_t* nnz_out);

/* ---- K1 + K4 (+K5): linear-elastic stiffness into CSR values -----------------
 * Replaces the element loop of assemble_linear_static (fem/assembly.cpp:237-244):
 * element_stiffness (fem/element.cpp:614) per element and scatter_tangent /
 * scatter_block into the free/free matrix. vals[nnz] is overwritten with K_ff.
 * Optional: when u and f_int are both non-NULL, f_int[3*n_nodes] is overwritten
 * with the internal force sum_e Ke ue over the full DOF set (fem/stress.cpp:
 * 125-145) in the same pass. coords: n_nodes*3; E, nu: per model element.
 * Optional live[nnz]: set to 1 where at least one element contributed a value other
 * than 0.0, else 0. The reference leaves entries whose every contribution is 0.0 out
 * of its COO map (scatter_block, fem/assembly.cpp:90); cxg_pattern_compact applies
 * that rule. Needs a two-phase assembly (option "asm_variant" 4 to 7; 7 is the default). */
cxg_status cxg_assemble_linear(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                               const double* E, const double* nu, double* vals,
                               const double* u, double* f_int, uint8_t* live);

/* cxg_assemble_linear for n_cases displacement fields at once (load cases of one
 * mesh or batch): u and f_int hold n_cases consecutive vectors of 3*n_nodes; K is
 * formed once and every f_int[c] = K_full u[c] comes from the same pass over the
 * element rows. n_cases 1, 2 or 4 (2 and 4 need a two-phase asm_variant, 4 to 7). */
cxg_status cxg_assemble_linear_cases(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                     const double* E, const double* nu, double* vals,
                                     int32_t n_cases, const double* u, double* f_int,
                                     uint8_t* live);

/* ---- K5: matrix-free internal force f_int = sum_e Ke(x,E) ue ----------------
 * fem::internal_force (fem/stress.cpp:125-145) without forming Ke: per Gauss point
 * B^T D B ue. f_int[3*n_nodes] is overwritten. This is also the forward
 * r = K(x, E) u of the adjoint pair below. */
cxg_status cxg_internal_force(cxg_context* ctx, const cxg_mesh* mesh, const double* coords,
                              const double* E, const double* nu, const double* u,
                              double* f_int);

/* cxg_internal_force without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The element forces go to a
 * fixed device array of CXG_NOALLOC_SCRATCH_DOUBLES doubles (64 MiB) that the library
 * holds on every device it runs on until the process ends, 3 doubles per connectivity
 * entry: meshes of at most 2,796,202 entries (349,525 C3D8 or 139,810 C3D20
 * elements); a larger mesh returns CXG_ERR_ARG before anything is enqueued. The caller
 * passes the mesh (created beforehand), inputs and outputs as for cxg_internal_force,
 * with the same results bit for bit. Every call on one device shares that device's
 * array, which keeps the last call's contents: calls that overlap on one device (two
 * streams or two threads) race. Like cxg_internal_force it synchronizes the context's
 * stream to report a non-positive Jacobian. */
#define CXG_NOALLOC_SCRATCH_DOUBLES (8 * 1024 * 1024)
cxg_status cxg_internal_force_noalloc(cxg_context* ctx, const cxg_mesh* mesh, const double* coords,
                                      const double* E, const double* nu, const double* u,
                                      double* f_int);

/* rhs[eq(g)] += f_ext[g] - f_c[g] for every free g, where f_c = K_full u_c is the
 * internal force of the prescribed displacement field u_c (zero on free DOFs).
 * Together with cxg_internal_force this forms the reduced rhs of
 * assemble_linear_static: loads through add_external_load (fem/assembly.cpp:
 * 200-208) and the eliminated prescribed columns (fem/assembly.cpp:95-96).
 * f_ext or f_c may be NULL (treated as zero). rhs[n_free] is accumulated into. */
cxg_status cxg_reduce_free(cxg_context* ctx, int32_t n_nodes, const int32_t* dof_eq,
                           const double* f_ext, const double* f_c, double* rhs);

/* ---- K2: fused J2 return map + consistent tangent + internal force -----------
 * Replaces assemble_material_tangent (fem/assembly.cpp:385-423) with
 * element_tangent_force (fem/element.cpp:708-743) and J2PlasticMaterial::evaluate
 * (src/fem/plasticity.cpp:90-236) at every Gauss point.
 * mat_id[n_elem]: index into mats (host array of n_mats), -1 = linear elastic
 * with E/nu. committed, trial: SoA state, 13 planes of n_points (see above);
 * committed is read, trial is overwritten. vals[nnz] <- K_t, f_int[3*n_nodes]
 * <- internal force (full DOF set). live: optional, as for cxg_assemble_linear. */
cxg_status cxg_assemble_j2(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                           const double* E, const double* nu, const int32_t* mat_id,
                           const cxg_j2_material* mats, int32_t n_mats, const double* u,
                           const double* committed, double* trial, double* vals,
                           double* f_int, uint8_t* live);

/* ---- material points: MaterialModel::evaluate at n independent points --------
 * Replaces the virtual MaterialModel::evaluate (fem/material_model.hpp:77) of
 * J2PlasticMaterial (src/fem/plasticity.cpp:90-236) and ElasticIsoMaterial
 * (src/fem/element.cpp:856-871), point by point, on the device.
 * mat_id[n_points]: index into mats (host array of n_mats) or -1 for an isotropic
 * elastic point with E[p], nu[p]; NULL: every point elastic (E and nu required; they
 * may be NULL when every point has a J2 material). strain[6 n]: total strain (Voigt,
 * engineering shear). committed, trial: SoA history, 13 planes of n_points; a J2 point
 * reads committed and writes trial, an elastic point leaves trial untouched (as the
 * reference). stress[6 n] and tangent[36 n] (row-major 6x6 D_t) are overwritten.
 * Synchronizes the context's stream. */
cxg_status cxg_material_evaluate(cxg_context* ctx, int64_t n_points, const int32_t* mat_id,
                                 const cxg_j2_material* mats, int32_t n_mats, const double* E,
                                 const double* nu, const double* strain, const double* committed,
                                 double* trial, double* stress, double* tangent);

/* MaterialPoints::commit (fem/assembly.hpp:84-87): committed <- trial. */
cxg_status cxg_commit(cxg_context* ctx, int64_t n_points, double* committed,
                      const double* trial);

/* ---- K6: adjoint (vector-Jacobian product) of r = K(x, E) u ------------------
 * For a cotangent w: dx[n_nodes*3] = d(w.r)/dx, dE[n_elem] = d(w.r)/dE_e,
 * du[n_nodes*3] = K^T w = K w. Outputs are overwritten. The reference's closest
 * counterpart is the semi-analytic central difference of the reduced operator in
 * solve_sensitivity (src/numerics/sensitivity.cpp:52-108), which covers
 * coordinate variables only. */
cxg_status cxg_vjp(cxg_context* ctx, const cxg_mesh* mesh, const double* coords,
                   const double* E, const double* nu, const double* u, const double* w,
                   double* dx, double* dE, double* du);

/* ---- K8: mass and geometric stiffness ----------------------------------------
 * Replace assemble_mass (fem/assembly.cpp:259-297) with element_mass /
 * element_mass_lumped (fem/element.cpp:786-831), and assemble_geometric_stiffness
 * (fem/assembly.cpp:299-330) with element_geometric_stiffness (fem/element.cpp:
 * 833-854), on the pattern p. rho[n_elem]: density per element (0: no mass);
 * lumped != 0: row-sum lumping. gp_stress[n_points*6]: Cauchy stress per Gauss point
 * (Voigt xx,yy,zz,xy,xz,yz; point index as the state arrays, e.g. the gauss_stress of
 * cxg_recover_fields). vals[nnz] is overwritten; live (optional) flags the entries that
 * received a contribution other than 0.0, for cxg_pattern_compact (the reference's
 * key set). *MASS point masses are not part of this path. */
cxg_status cxg_assemble_mass(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                             const double* rho, int32_t lumped, double* vals, uint8_t* live);
cxg_status cxg_assemble_geometric(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                  const double* gp_stress, double* vals, uint8_t* live);

/* Both matrices are k_ab I3 blocks: an entry couples two DOFs of one direction or is
 * never touched, so the reference's key set lies inside the same-direction pattern, the
 * structural pattern's entries (a, i)-(b, i) (a third of it). cxg_pattern_same_direction
 * builds that pattern once per pattern (synchronizes) and returns its nnz in *nnz_out
 * (host); row_ptr[n_free+1] / col_idx[nnz] (device, optional) receive a copy.
 * cxg_assemble_mass_sd / cxg_assemble_geometric_sd write vals[nnz] and live[nnz] (both
 * required) on it directly, a third of the structural writes and no compaction when
 * every entry is live; cxg_compact_csr keeps the live entries of any CSR (as
 * cxg_pattern_compact; synchronizes). They need every CSR row within the gather's LDS
 * columns (option "gather_lds_cols"; rows of 20-node hexahedra fit the default) and
 * return CXG_ERR_UNSUPPORTED otherwise (cxg_pattern_same_direction first of all): use
 * cxg_assemble_mass + cxg_pattern_compact. */
cxg_status cxg_pattern_same_direction(cxg_context* ctx, const cxg_pattern* p, int64_t* row_ptr,
                                      int32_t* col_idx, int64_t* nnz_out);
cxg_status cxg_assemble_mass_sd(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                const double* rho, int32_t lumped, double* vals, uint8_t* live);
cxg_status cxg_assemble_geometric_sd(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                     const double* gp_stress, double* vals, uint8_t* live);
cxg_status cxg_compact_csr(cxg_context* ctx, int32_t n_rows, const int64_t* row_ptr,
                           const int32_t* col_idx, const uint8_t* live, const double* vals,
                           int64_t* row_ptr_out, int32_t* col_idx_out, double* vals_out,
                           int64_t* nnz_out);

/* ---- K7: field recovery -----------------------------------------------------
 * Replaces recover_fields (fem/stress.cpp:189-273), its material-point form
 * (fem/stress.cpp:275-353) and recover_gauss_stress (fem/stress.cpp:147-187) for the
 * C3D8/C3D20/C3D20R/C3D6 families, whose integration-point -> node extrapolation is
 * the Gauss-point mean at every node (fem/extrapolation.hpp:26-55).
 * u[3*n_nodes]: the displacement of every node. Per Gauss point: strain = B u
 * (engineering shear); stress = D strain, or, when mats is given, the radial return
 * of J2 elements (mat_id >= 0) from the committed history (13 planes of n_points;
 * read, never advanced), as recover_fields(model, fields, mp).
 * Outputs (device, each may be NULL, overwritten):
 *   strain, stress [n_nodes*6]  mean over the node's elements of each element's
 *                               Gauss-point mean (Voigt xx,yy,zz,xy,xz,yz)
 *   f_int [3*n_nodes]           sum_e of det*w B^T stress over the Gauss points
 *   reaction [3*n_nodes]        f_int - f_ext (f_ext: full-DOF load, NULL = zero)
 *   gauss_stress [n_points*6]   per point, at the point index of the state arrays
 * Inactive elements (*MODEL CHANGE, REMOVE) must be left out of the mesh. */
cxg_status cxg_recover_fields(cxg_context* ctx, const cxg_mesh* mesh, const double* coords,
                              const double* E, const double* nu, const int32_t* mat_id,
                              const cxg_j2_material* mats, int32_t n_mats, const double* u,
                              const double* committed, const double* f_ext, double* strain,
                              double* stress, double* reaction, double* f_int,
                              double* gauss_stress);

/* ---- f1: sparse SPD solve on the device --------------------------------------
 * Replaces ComputeBackend::solve_sparse (compute/backend.hpp:51-56) for the SPD
 * systems of linear and Newton steps (the CPU backend hands them to SciPP,
 * src/compute/cpu_backend.cpp:33-46): Jacobi-preconditioned conjugate gradients on
 * the CSR (row_ptr[n+1], col_idx, vals: device, e.g. a pattern and its values). x is
 * the start on entry and the solution on return; iterations stop once
 * ||b - A x|| <= tol ||b|| (checked every 16 iterations) or after max_iter
 * (CXG_ERR_STATE, x holds the last iterate). Deterministic: every reduction has a
 * fixed order. iters / rel_residual (host, may be NULL) receive the count and the
 * final relative residual. */
cxg_status cxg_pcg(cxg_context* ctx, int32_t n, const int64_t* row_ptr, const int32_t* col_idx,
                   const double* vals, const double* b, double* x, double tol, int32_t max_iter,
                   int32_t* iters, double* rel_residual);

/* ---- HIP graphs: one assembly / one Newton iteration, replayed ---------------
 * The reference re-assembles K_t, f_int and the residual at every Newton iteration
 * of solve_nonlinear through assemble_material_tangent (fem/assembly.cpp:385-423)
 * with the same mesh and buffers. cxg_graph_capture_j2 captures that iteration (the
 * arguments of cxg_assemble_j2; with rhs non-NULL also rhs[n_free] = f_ext - f_int
 * on the free DOFs, f_ext may be NULL) into a HIP graph; cxg_graph_launch replays it
 * on the context's stream, reading the inputs from the captured addresses at replay
 * time (update u in place between iterations). cxg_graph_capture_linear does the
 * same for cxg_assemble_linear (no live flags in graphs). The graph keeps its own
 * copy of the material table; it must be destroyed before its pattern and mesh.
 * While a graph of a pattern is alive, the pattern's element arenas, gather order and
 * work queue stay where the graph found them: a later call that would have to rebuild
 * one of them (more load cases than any capture or earlier call used, or a changed
 * "gather_tile" or "chunk_elems" option) returns CXG_ERR_STATE instead; destroy the
 * pattern's graphs first.
 * Kernel status (non-positive Jacobian) is reported by cxg_check after replays. */
cxg_status cxg_graph_capture_linear(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                    const double* E, const double* nu, double* vals,
                                    const double* u, double* f_int, cxg_graph** out);
/* As cxg_graph_capture_linear for cxg_assemble_linear_cases. */
cxg_status cxg_graph_capture_linear_cases(cxg_context* ctx, const cxg_pattern* p,
                                          const double* coords, const double* E,
                                          const double* nu, double* vals, int32_t n_cases,
                                          const double* u, double* f_int, cxg_graph** out);
cxg_status cxg_graph_capture_j2(cxg_context* ctx, const cxg_pattern* p, const double* coords,
                                const double* E, const double* nu, const int32_t* mat_id,
                                const cxg_j2_material* mats, int32_t n_mats, const double* u,
                                const double* committed, double* trial, double* vals,
                                double* f_int, const double* f_ext, double* rhs,
                                cxg_graph** out);
cxg_status cxg_graph_launch(cxg_context* ctx, cxg_graph* g);
cxg_status cxg_graph_info(const cxg_graph* g, int64_t* nodes);
void cxg_graph_destroy(cxg_graph* g);
/* Synchronizes the context's stream and reports (then clears) kernel status flags
 * raised since the last check, as the eager entry points do before returning. */
cxg_status cxg_check(cxg_context* ctx);

/* ---- kernel timing ----------------------------------------------------------
 * When enabled, every kernel launch of the library is bracketed by HIP events on
 * the context's stream; per kernel name the library accumulates launches, total
 * milliseconds and the number of elements processed. No counterpart in the
 * reference (it has no kernel timers). Timing adds two event
 * records per launch and a synchronization per call; leave it off when timing
 * whole calls. */
cxg_status cxg_profile_enable(cxg_context* ctx, int enable);
cxg_status cxg_profile_reset(cxg_context* ctx);
int32_t cxg_profile_count(const cxg_context* ctx);
/* Entry i: name (owned by the context), launches, total ms, elements processed. */
cxg_status cxg_profile_get(const cxg_context* ctx, int32_t i, const char** name,
                           int64_t* launches, double* total_ms, int64_t* elements);

#ifdef __cplusplus
}
#endif
#endif
