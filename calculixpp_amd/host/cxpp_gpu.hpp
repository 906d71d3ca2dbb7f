// Drop-in GPU versions of the reference's assembly entry points, for C++20 callers.
//
// Same signatures, argument meaning, ownership (value semantics) and error behaviour
// (std::runtime_error with the reference's messages) as the cxpp::fem functions they
// replace; they run on libcxg.so through its C ABI (include/cxg.h):
//
//   cxpp::gpu::assemble_linear_static      fem::assemble_linear_static
//                                          (include/calculixpp/fem/assembly.hpp:42,
//                                           src/fem/assembly.cpp:225-257)
//   cxpp::gpu::assemble_material_tangent   fem::assemble_material_tangent
//                                          (assembly.hpp:105-108, assembly.cpp:385-423)
//   cxpp::gpu::internal_force              fem::internal_force
//                                          (fem/stress.hpp:16, stress.cpp:125-145)
//   cxpp::gpu::element_stiffness           fem::element_stiffness
//                                          (fem/element.hpp:62, element.cpp:614-634)
//   cxpp::gpu::assemble_mass               fem::assemble_mass
//                                          (assembly.hpp:55, assembly.cpp:259-297)
//   cxpp::gpu::assemble_geometric_stiffness fem::assemble_geometric_stiffness
//                                          (assembly.hpp:68, assembly.cpp:299-330)
//   cxpp::gpu::recover_fields (both forms) fem::recover_fields
//                                          (fem/stress.hpp:24,43, stress.cpp:189-353)
//   cxpp::gpu::recover_gauss_stress        fem::recover_gauss_stress
//                                          (fem/stress.hpp:35, stress.cpp:147-187)
//   cxpp::gpu::element_tangent_force       fem::element_tangent_force
//                                          (fem/element.hpp:160, element.cpp:708-743)
//   cxpp::gpu::J2Material::evaluate        fem::MaterialModel::evaluate of the J2 law
//                                          (fem/material_model.hpp:77,
//                                           plasticity.cpp:90-236), on the device
//   cxpp::gpu::HipBackend                  compute::ComputeBackend
//                                          (compute/backend.hpp:40-57): solve_sparse
//                                          and kind(); hip_backend() is its instance
//   cxpp::gpu::solve_sparse                HipBackend::solve_sparse as a free function
//
// This file and cxpp_gpu.cpp compile inside the reference's tree (they include its
// headers) and link against libcxg.so. A model the GPU path does not cover (element
// types other than C3D8/C3D20/C3D20R/C3D6, MPC/rigid/tie equations, springs, thermal
// loads, hyperelastic or user materials) raises std::runtime_error naming the
// feature; nothing falls back to the CPU.
//
// Results: K as COO triplets in CSR order (rows ascending, columns ascending inside a
// row) with the reference's key set (entries whose every element contribution is 0.0
// are left out, assembly.cpp:90); values within rel-L2 1e-12 of the reference's.
// One libcxg context per host thread, on the device chosen by set_device (default
// 0), created on first use and kept until the process ends.
//
// Residency: a Newton loop calls assemble_material_tangent with one Model many times
// (src/numerics/nonlinear_static.cpp:94 and the line-search re-assemblies at :123-138).
// Each thread keeps the device copy of the last model it saw: node coordinates,
// connectivity, elasticity, the device mesh, the all-free CSR pattern (and its host
// copy), the value buffers, the J2 material table and the committed history. The cache
// is keyed by the Model's address and a 64-bit fingerprint of everything uploaded, so
// a changed or different model is uploaded afresh; the committed history is compared
// with the device copy on every call and uploaded only when the host changed it (a
// commit() of the trial state the last call returned is replayed on the device).
#pragma once

#include <cstdint>
#include <span>
#include <vector>

#include "calculixpp/compute/backend.hpp"
#include "calculixpp/core/model.hpp"
#include "calculixpp/core/results.hpp"
#include "calculixpp/fem/assembly.hpp"
#include "calculixpp/fem/element.hpp"
#include "calculixpp/fem/material_model.hpp"

namespace cxpp::gpu {

// HIP device of the calling thread's context (takes effect on the next call).
void set_device(int device);

// Array-native model of the GPU path: what a cxpp::Model carries for it, as arrays, so
// that a mesh with per-element elasticity at scale (262k elements, each its own E) needs
// no *MATERIAL and *ELSET per element. It replaces the section -> material resolution of
// Model::element_elastic / element_density / element_active_mask
// (src/core/model.cpp:33-72, 200-207, 267-280) and the *BOUNDARY / *CLOAD lists of the
// deck (io/inp_parser.cpp:1631-1645); node and element numbers are indices (node ids of
// a Model map to their insertion order, mesh.cpp:29-32).
struct ArrayModel {
  struct Block {
    ElementType type{};
    std::vector<Index> conn;  // node indices, nodes_per_element(type) per element, CalculiX order
  };
  struct NodalValue {
    Index node{0};  // node index
    int comp{1};    // 1..3
    Real value{0.0};
  };
  std::vector<Real> coords;     // x, y, z per node
  std::vector<Block> blocks;    // runs of one family; element numbers follow block order
  std::vector<Real> E, nu;      // per element
  std::vector<Real> rho;        // per element (*DENSITY); empty: 0
  std::vector<bool> active;     // per element (*MODEL CHANGE, REMOVE); empty: all active
  std::vector<NodalValue> spc;  // single point constraints (prescribed values)
  std::vector<NodalValue> load; // nodal loads
  size_t num_nodes() const { return coords.size() / 3; }
  size_t num_elements() const;
};

// The arrays of a reference Model: coordinates in node order, element runs in model
// order, each element's elasticity and density through its section, the active mask,
// the *BOUNDARY values and the nonzero entries of fem::external_load_vector. Throws
// std::runtime_error with the reference's messages (unknown node, elset, material).
ArrayModel to_arrays(const Model& model);

fem::LinearSystem assemble_linear_static(const Model& model);

fem::LinearSystem assemble_material_tangent(const Model& model, const std::vector<Vec3>& u,
                                            fem::MaterialPoints& mp, std::vector<Real>& f_int);

std::vector<Real> internal_force(const Model& model, const std::vector<Vec3>& u);

// The same three entry points on an ArrayModel: single point constraints only, numbered
// as build_dof_map does (ascending global DOF, constrained DOFs skipped); the transform of
// the returned system is the SPC-only one (a free DOF expands to itself, a constrained DOF
// to its prescribed value).
fem::LinearSystem assemble_linear_static(const ArrayModel& model);
std::vector<Real> internal_force(const ArrayModel& model, const std::vector<Vec3>& u);
fem::LinearSystem assemble_mass(const ArrayModel& model, bool lumped = false);

std::vector<Real> element_stiffness(ElementType type, std::span<const Vec3> coords,
                                    const ElasticIso& elastic);

// Consistent (or row-sum lumped) mass from the elements' *DENSITY; *MASS point masses
// raise std::runtime_error.
fem::LinearSystem assemble_mass(const Model& model, bool lumped = false);

// Geometric stiffness of the per-element Gauss-point stresses (missing: zero).
fem::LinearSystem assemble_geometric_stiffness(const Model& model,
                                               const std::vector<std::vector<Voigt6>>& gp_stress);

// Averaged nodal stress and strain and the reactions RF = f_int - f_ext of
// fields.displacement; with mp, J2 elements take the radial-return stress from the
// committed history (read only).
void recover_fields(const Model& model, StaticFields& fields);
void recover_fields(const Model& model, StaticFields& fields, fem::MaterialPoints& mp);

// Per element, the stress at each Gauss point of u (removed elements: empty).
std::vector<std::vector<Voigt6>> recover_gauss_stress(const Model& model,
                                                      const std::vector<Vec3>& u);

// J2 plasticity (*PLASTIC with its *ELASTIC) whose constants are readable, so that the
// device path can take it: the reference's J2PlasticMaterial keeps them private. Built
// from the same ElasticIso + Plastic as J2PlasticMaterial (and validated the same way);
// evaluate() runs the radial return of one point on the device (cxg_material_evaluate).
class J2Material final : public fem::MaterialModel {
 public:
  J2Material(const ElasticIso& elastic, const Plastic& plastic);
  fem::MaterialResponse evaluate(const Voigt6& strain, fem::MaterialState& state) const override;
  const ElasticIso& elastic() const { return elastic_; }
  const Plastic& plastic() const { return plastic_; }

 private:
  ElasticIso elastic_;
  Plastic plastic_;
};

// Ke = sum_q B^T D_t B det J w and fe = sum_q B^T sigma det J w of one element at the
// nodal displacements ue, with the material's history advanced in the trial slots of
// `state` (one entry per Gauss point). The material must be fem::ElasticIsoMaterial
// (its D gives E and nu) or cxpp::gpu::J2Material; any other MaterialModel raises
// std::runtime_error naming the types the device path takes.
fem::ElementResponse element_tangent_force(ElementType type, std::span<const Vec3> coords,
                                           std::span<const Vec3> ue,
                                           const fem::MaterialModel& material,
                                           std::span<fem::MaterialState> state);

// The ComputeBackend of the device. kind() is kHipBackendKind, the value a maintainer
// gives the enumerator BackendKind::HIP when adding it to compute/backend.hpp (the
// reference lists CPU, CUDA, OpenCL and Metal only). solve_sparse runs
// Jacobi-preconditioned conjugate gradients on the device (cxg_pcg) for either
// SolverKind, each with its own contract:
//   SolverKind::CG      ||b - A x|| <= 1e-12 ||b||; on failure std::runtime_error
//                       "CG did not converge (relative residual R after N iterations)"
//                       as the reference's CG path (src/compute/cpu_backend.cpp:55-62);
//   SolverKind::Direct  there is no sparse factorization on the device: the request is
//                       served by the same iteration to ||b - A x|| <= 1e-13 ||b|| and
//                       raises "Direct solve by device PCG did not reach 1e-13 ..."
//                       when that is not reached, never returning a looser result.
inline constexpr compute::BackendKind kHipBackendKind = static_cast<compute::BackendKind>(4);

class HipBackend final : public compute::ComputeBackend {
 public:
  compute::BackendKind kind() const override { return kHipBackendKind; }
  std::vector<Real> solve_sparse(const std::vector<Index>& rows, const std::vector<Index>& cols,
                                 const std::vector<Real>& vals, Index n,
                                 const std::vector<Real>& rhs,
                                 compute::SolverKind solver) const override;
};

// The process-wide HipBackend (stateless; the device context is per calling thread).
const compute::ComputeBackend& hip_backend();

// HipBackend::solve_sparse as a free function.
std::vector<Real> solve_sparse(const std::vector<Index>& rows, const std::vector<Index>& cols,
                               const std::vector<Real>& vals, Index n,
                               const std::vector<Real>& rhs, compute::SolverKind solver);

// Counters of the calling thread's residency cache (see the top of this header).
struct ResidentStats {
  int64_t calls = 0;             // model-level calls that went through the cache
  int64_t mesh_uploads = 0;      // device meshes built (coordinates, connectivity, E, nu)
  int64_t pattern_builds = 0;    // all-free CSR patterns built
  int64_t dof_maps = 0;          // DOF maps built from the model's constraints
  int64_t history_uploads = 0;   // committed histories copied to the device
  int64_t history_commits = 0;   // commits replayed on the device (trial -> committed)
};
ResidentStats resident_stats();
// Drops the calling thread's cached model (device memory is released) and zeroes the
// counters.
void resident_reset();

}  // namespace cxpp::gpu
