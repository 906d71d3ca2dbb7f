// TEST INFRASTRUCTURE ONLY — C-ABI glue from the oracle's array model
// (orc_model.h) to the reference's own calculixpp_core, compiled together with the
// reference sources by oracle/Makefile into oracle/_ref/libcxpp_ref.so.
//
// It is used to pin the CPU restatement (cxg_oracle.cpp), to write the golden
// fixtures under tests/golden/, and as the "reference" CPU baseline in bench.py.
// Only this glue is ours; every number it returns is computed by the reference
// functions it calls:
//   fem::element_stiffness           (src/fem/element.cpp:614)
//   fem::element_tangent_force       (src/fem/element.cpp:708)
//   fem::J2PlasticMaterial::evaluate (src/fem/plasticity.cpp:90)
//   fem::assemble_linear_static      (src/fem/assembly.cpp:225)
//   fem::make_material_points        (src/fem/assembly.cpp:357)
//   fem::assemble_material_tangent   (src/fem/assembly.cpp:385)
//   fem::internal_force              (src/fem/stress.cpp:125)
//   fem::recover_fields              (src/fem/stress.cpp:189, 275)
//   fem::assemble_mass               (src/fem/assembly.cpp:259)
//   fem::assemble_geometric_stiffness (src/fem/assembly.cpp:299)
//   fem::recover_gauss_stress        (src/fem/stress.cpp:147)
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "calculixpp/core/model.hpp"
#include "calculixpp/fem/assembly.hpp"
#include "calculixpp/fem/element.hpp"
#include "calculixpp/fem/material_model.hpp"
#include "calculixpp/fem/loads.hpp"
#include "calculixpp/fem/stress.hpp"
#include "orc_model.h"

using namespace cxpp;

// Built a second time with CXPP_DROPIN (oracle/Makefile, target dropin): the same
// entry points then run the GPU drop-in cxpp::gpu (calculixpp_amd/host/cxpp_gpu.cpp)
// on the very same cxpp::Model, so a test can hold the two side by side.
#ifdef CXPP_DROPIN
#include "cxpp_gpu.hpp"
namespace impl = cxpp::gpu;
#else
namespace impl = cxpp::fem;
#endif

namespace {

void set_err(char* err, int len, const char* msg) {
  if (err && len > 0) {
    std::strncpy(err, msg, size_t(len) - 1);
    err[len - 1] = 0;
  }
}

Plastic to_plastic(const orc_j2& p) {
  Plastic pl;
  pl.hardening = p.hardening == ORC_HARD_KIN        ? Plastic::Hardening::Kinematic
                 : p.hardening == ORC_HARD_COMBINED ? Plastic::Hardening::Combined
                                                    : Plastic::Hardening::Isotropic;
  pl.eqplastic.assign(p.eqplastic, p.eqplastic + p.n_table);
  pl.yield.assign(p.yield, p.yield + p.n_table);
  pl.kinematic_modulus = p.kinematic_modulus;
  return pl;
}

void load_state(const double* st, fem::MaterialState& s) {
  for (int i = 0; i < 6; ++i) {
    s.committed_plastic_strain[size_t(i)] = st[i];
    s.committed_back_stress[size_t(i)] = st[7 + i];
    s.plastic_strain[size_t(i)] = st[13 + i];
    s.back_stress[size_t(i)] = st[20 + i];
  }
  s.committed_eqplastic = st[6];
  s.eqplastic = st[19];
}

void store_state(const fem::MaterialState& s, double* st) {
  for (int i = 0; i < 6; ++i) {
    st[i] = s.committed_plastic_strain[size_t(i)];
    st[7 + i] = s.committed_back_stress[size_t(i)];
    st[13 + i] = s.plastic_strain[size_t(i)];
    st[20 + i] = s.back_stress[size_t(i)];
  }
  st[6] = s.committed_eqplastic;
  st[19] = s.eqplastic;
}

std::vector<Vec3> as_vec3(const double* p, size_t n) {
  std::vector<Vec3> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
  return v;
}

// Build a reference Model: node ids = index+1, element ids = index+1, one
// material + elset per distinct (E, nu, J2 material, density).
std::unique_ptr<Model> build_model(const orc_model& m) {
  auto M = std::make_unique<Model>();
  for (int32_t i = 0; i < m.n_nodes; ++i)
    M->mesh.add_node(i + 1, {m.coords[3 * i], m.coords[3 * i + 1], m.coords[3 * i + 2]});
  std::map<std::tuple<double, double, int32_t, double>, std::vector<Index>> groups;
  for (int32_t e = 0; e < m.n_elem; ++e) {
    std::vector<Index> ids;
    for (int64_t k = m.eptr[e]; k < m.eptr[e + 1]; ++k) ids.push_back(m.conn[k] + 1);
    M->mesh.add_element(e + 1, static_cast<ElementType>(m.etype[e]), std::move(ids));
    const int32_t jid = m.j2_id ? m.j2_id[e] : -1;
    groups[{m.E[e], m.nu[e], jid, m.rho ? m.rho[e] : 0.0}].push_back(e + 1);
    if (m.active && m.active[e] == 0) M->deactivated_elements.push_back(e + 1);
  }
  int g = 0;
  for (auto& [key, ids] : groups) {
    const std::string name = "M" + std::to_string(g++);
    Material mat;
    mat.name = name;
    mat.elastic = ElasticIso{std::get<0>(key), std::get<1>(key)};
    if (std::get<3>(key) != 0.0) mat.density = std::get<3>(key);
    if (std::get<2>(key) >= 0) mat.plastic = to_plastic(m.j2[std::get<2>(key)]);
    M->materials[name] = mat;
    M->mesh.add_elset(name, ids);
    M->sections.push_back(SolidSection{name, name});
  }
  for (int32_t i = 0; i < m.n_spc; ++i) {
    Spc s;
    s.node_id = m.spc_node[i] + 1;
    s.comp = m.spc_comp[i];
    s.value = m.spc_value[i];
    M->spcs.push_back(s);
  }
  for (int32_t i = 0; i < m.n_cload; ++i) {
    Cload c;
    c.node_id = m.cload_node[i] + 1;
    c.comp = m.cload_comp[i];
    c.value = m.cload_value[i];
    M->cloads.push_back(c);
  }
  return M;
}

}  // namespace

struct ref_model {
  std::unique_ptr<Model> model;
};

struct ref_system {
  fem::LinearSystem sys;
};

extern "C" {

int ref_element_stiffness(int type, const double* coords, double E, double nu, double* Ke,
                          char* err, int errlen) {
  try {
    const int n = nodes_per_element(static_cast<ElementType>(type));
    const std::vector<Vec3> x = as_vec3(coords, size_t(n));
    const std::vector<Real> K = impl::element_stiffness(static_cast<ElementType>(type), x,
                                                       ElasticIso{E, nu});
    std::memcpy(Ke, K.data(), K.size() * sizeof(double));
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int ref_j2_evaluate(const orc_j2* p, const double* strain, double* state, double* stress,
                    double* D, char* err, int errlen) {
  try {
#ifdef CXPP_DROPIN
    const cxpp::gpu::J2Material mat(ElasticIso{p->E, p->nu}, to_plastic(*p));
#else
    fem::J2PlasticMaterial mat(ElasticIso{p->E, p->nu}, to_plastic(*p));
#endif
    fem::MaterialState s;
    load_state(state, s);
    Voigt6 eps{};
    for (int i = 0; i < 6; ++i) eps[size_t(i)] = strain[i];
    const fem::MaterialResponse r = mat.evaluate(eps, s);
    store_state(s, state);
    for (int i = 0; i < 6; ++i) {
      stress[i] = r.stress[size_t(i)];
      for (int j = 0; j < 6; ++j) D[6 * i + j] = r.tangent[size_t(i)][size_t(j)];
    }
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int ref_element_tangent_force(int type, const double* coords, const double* ue, double E,
                              double nu, const orc_j2* p, double* state, double* Ke,
                              double* fe, char* err, int errlen) {
  try {
    const ElementType t = static_cast<ElementType>(type);
    const int n = nodes_per_element(t);
    const size_t ng = fem::gauss_rule(t).size();
    const std::vector<Vec3> x = as_vec3(coords, size_t(n)), u = as_vec3(ue, size_t(n));
    std::unique_ptr<fem::MaterialModel> mat;
#ifdef CXPP_DROPIN
    using J2Law = cxpp::gpu::J2Material;
#else
    using J2Law = fem::J2PlasticMaterial;
#endif
    if (p)
      mat = std::make_unique<J2Law>(ElasticIso{E, nu}, to_plastic(*p));
    else
      mat = std::make_unique<fem::ElasticIsoMaterial>(ElasticIso{E, nu});
    std::vector<fem::MaterialState> st(ng);
    if (state)
      for (size_t q = 0; q < ng; ++q) load_state(state + q * ORC_STATE, st[q]);
    const fem::ElementResponse r = impl::element_tangent_force(t, x, u, *mat, st);
    if (state)
      for (size_t q = 0; q < ng; ++q) store_state(st[q], state + q * ORC_STATE);
    std::memcpy(Ke, r.Ke.data(), r.Ke.size() * sizeof(double));
    std::memcpy(fe, r.fe.data(), r.fe.size() * sizeof(double));
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int ref_model_create(const orc_model* m, ref_model** out, char* err, int errlen) {
  try {
    auto* h = new ref_model();
    h->model = build_model(*m);
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

void ref_model_free(ref_model* h) { delete h; }

int ref_assemble_linear_static(const ref_model* h, ref_system** out, char* err, int errlen) {
  try {
    auto* s = new ref_system();
    s->sys = impl::assemble_linear_static(*h->model);
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// state: sum_e ngp(e)*26 doubles, element-major (orc_model.h layout); committed
// slots are loaded, trial slots are written back. f_int: 3*n_nodes.
int ref_assemble_material_tangent(const ref_model* h, const double* u, double* state,
                                  double* f_int, ref_system** out, char* err, int errlen) {
  try {
    const Model& M = *h->model;
    fem::MaterialPoints mp = fem::make_material_points(M);
    size_t off = 0;
    for (auto& es : mp.state)
      for (auto& s : es) {
        load_state(state + off, s);
        off += ORC_STATE;
      }
    const std::vector<Vec3> uu = as_vec3(u, M.mesh.num_nodes());
    std::vector<Real> fi;
    auto* s = new ref_system();
    s->sys = impl::assemble_material_tangent(M, uu, mp, fi);
    off = 0;
    for (auto& es : mp.state)
      for (auto& st : es) {
        store_state(st, state + off);
        off += ORC_STATE;
      }
    std::memcpy(f_int, fi.data(), fi.size() * sizeof(double));
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int ref_internal_force(const ref_model* h, const double* u, double* f_int, char* err,
                       int errlen) {
  try {
    const std::vector<Vec3> uu = as_vec3(u, h->model->mesh.num_nodes());
    const std::vector<Real> fi = impl::internal_force(*h->model, uu);
    std::memcpy(f_int, fi.data(), fi.size() * sizeof(double));
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// recover_fields (state NULL) or its material-point form (state: 26 doubles per point,
// the committed slots loaded into make_material_points' history). f_int is derived as
// reaction + f_ext (the reference keeps its f_int internal). gauss_stress (state NULL
// only): recover_gauss_stress flattened in element order, zeros for removed elements.
int ref_recover_fields(const ref_model* h, const double* u, const double* state,
                       double* strain, double* stress, double* reaction, double* f_int,
                       double* gauss_stress, char* err, int errlen) {
  try {
    const Model& M = *h->model;
    const size_t nn = M.mesh.num_nodes();
    StaticFields fields;
    fields.displacement = as_vec3(u, nn);
    if (state) {
      fem::MaterialPoints mp = fem::make_material_points(M);
      size_t off = 0;
      for (auto& es : mp.state)
        for (auto& s : es) {
          load_state(state + off, s);
          off += ORC_STATE;
        }
      impl::recover_fields(M, fields, mp);
    } else {
      impl::recover_fields(M, fields);
    }
    const std::vector<Real> fx = fem::external_load_vector(M);
    for (size_t i = 0; i < nn; ++i) {
      for (int c = 0; c < 6; ++c) {
        if (strain) strain[6 * i + c] = fields.strain[i][size_t(c)];
        if (stress) stress[6 * i + c] = fields.stress[i][size_t(c)];
      }
      for (int c = 0; c < 3; ++c) {
        if (reaction) reaction[3 * i + c] = fields.reaction[i][size_t(c)];
        if (f_int) f_int[3 * i + c] = fields.reaction[i][size_t(c)] + fx[3 * i + size_t(c)];
      }
    }
    if (gauss_stress && !state) {
      const auto gp = impl::recover_gauss_stress(M, fields.displacement);
      size_t off = 0;
      for (size_t e = 0; e < gp.size(); ++e) {
        const size_t ng = fem::gauss_rule(M.mesh.elements()[e].type).size();
        for (size_t q = 0; q < ng; ++q)
          for (int c = 0; c < 6; ++c)
            gauss_stress[6 * (off + q) + c] = gp[e].empty() ? 0.0 : gp[e][q][size_t(c)];
        off += ng;
      }
    }
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// assemble_mass (kind 0 consistent, 1 lumped) or assemble_geometric_stiffness (kind 2,
// gp_stress flattened per point in element order; removed elements' points skipped).
int ref_assemble_scalar(const ref_model* h, int kind, const double* gp_stress, ref_system** out,
                        char* err, int errlen) {
  try {
    const Model& M = *h->model;
    auto* s = new ref_system();
    if (kind < 2) {
      s->sys = impl::assemble_mass(M, kind == 1);
    } else {
      std::vector<std::vector<Voigt6>> gp(M.mesh.num_elements());
      const std::vector<bool> active = M.element_active_mask();
      size_t off = 0;
      for (size_t e = 0; e < gp.size(); ++e) {
        const size_t ng = fem::gauss_rule(M.mesh.elements()[e].type).size();
        if (active[e] && gp_stress) {
          gp[e].resize(ng);
          for (size_t q = 0; q < ng; ++q)
            for (size_t c = 0; c < 6; ++c) gp[e][q][c] = gp_stress[6 * (off + q) + c];
        }
        off += ng;
      }
      s->sys = impl::assemble_geometric_stiffness(M, gp);
    }
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int32_t ref_system_n_free(const ref_system* s) { return s->sys.n_free; }
int64_t ref_system_nnz(const ref_system* s) { return int64_t(s->sys.vals.size()); }

void ref_system_coo(const ref_system* s, int32_t* rows, int32_t* cols, double* vals) {
  std::memcpy(rows, s->sys.rows.data(), s->sys.rows.size() * sizeof(int32_t));
  std::memcpy(cols, s->sys.cols.data(), s->sys.cols.size() * sizeof(int32_t));
  std::memcpy(vals, s->sys.vals.data(), s->sys.vals.size() * sizeof(double));
}

void ref_system_rhs(const ref_system* s, double* rhs) {
  std::memcpy(rhs, s->sys.rhs.data(), s->sys.rhs.size() * sizeof(double));
}

void ref_system_dof_eq(const ref_system* s, int32_t* dof_eq) {
  std::memcpy(dof_eq, s->sys.dof_eq.data(), s->sys.dof_eq.size() * sizeof(int32_t));
}

void ref_system_free(ref_system* s) { delete s; }

#ifdef CXPP_DROPIN
// The orc_model arrays as a cxpp::gpu::ArrayModel: no Model, no materials or elsets.
static cxpp::gpu::ArrayModel array_model(const orc_model& m) {
  cxpp::gpu::ArrayModel a;
  a.coords.assign(m.coords, m.coords + 3 * int64_t(m.n_nodes));
  for (int32_t e = 0; e < m.n_elem; ++e) {
    const ElementType t = static_cast<ElementType>(m.etype[e]);
    if (a.blocks.empty() || a.blocks.back().type != t) a.blocks.push_back({t, {}});
    a.blocks.back().conn.insert(a.blocks.back().conn.end(), m.conn + m.eptr[e], m.conn + m.eptr[e + 1]);
  }
  a.E.assign(m.E, m.E + m.n_elem);
  a.nu.assign(m.nu, m.nu + m.n_elem);
  if (m.rho) a.rho.assign(m.rho, m.rho + m.n_elem);
  if (m.active) a.active.assign(m.active, m.active + m.n_elem);
  for (int32_t i = 0; i < m.n_spc; ++i) a.spc.push_back({m.spc_node[i], m.spc_comp[i], m.spc_value[i]});
  for (int32_t i = 0; i < m.n_cload; ++i)
    a.load.push_back({m.cload_node[i], m.cload_comp[i], m.cload_value[i]});
  return a;
}

// cxpp::gpu::assemble_linear_static on the ArrayModel of the arrays (arrays = 1) or on
// to_arrays of the reference Model built from them (arrays = 0).
int ref_gpu_array_linear_static(const orc_model* m, int arrays, ref_system** out, char* err,
                                int errlen) {
  try {
    auto* s = new ref_system();
    if (arrays) {
      s->sys = cxpp::gpu::assemble_linear_static(array_model(*m));
    } else {
      const std::unique_ptr<Model> M = build_model(*m);
      s->sys = cxpp::gpu::assemble_linear_static(cxpp::gpu::to_arrays(*M));
    }
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int ref_gpu_array_internal_force(const orc_model* m, const double* u, double* f_int, char* err,
                                 int errlen) {
  try {
    const std::vector<Real> f =
        cxpp::gpu::internal_force(array_model(*m), as_vec3(u, size_t(m->n_nodes)));
    std::memcpy(f_int, f.data(), f.size() * sizeof(double));
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

int ref_gpu_array_mass(const orc_model* m, int lumped, ref_system** out, char* err, int errlen) {
  try {
    auto* s = new ref_system();
    s->sys = cxpp::gpu::assemble_mass(array_model(*m), lumped != 0);
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// COO system -> cxpp::gpu::hip_backend().solve_sparse (drop-in build only); direct != 0
// asks for SolverKind::Direct, else SolverKind::CG. *kind receives backend.kind().
int ref_solve_sparse(int32_t n, int64_t nnz, const int32_t* rows, const int32_t* cols,
                     const double* vals, const double* rhs, int direct, double* x, int* kind,
                     char* err, int errlen) {
  try {
    const std::vector<Index> r(rows, rows + nnz), c(cols, cols + nnz);
    const std::vector<Real> v(vals, vals + nnz), b(rhs, rhs + n);
    const compute::ComputeBackend& be = cxpp::gpu::hip_backend();
    if (kind) *kind = static_cast<int>(be.kind());
    const std::vector<Real> u = be.solve_sparse(
        r, c, v, n, b, direct ? compute::SolverKind::Direct : compute::SolverKind::CG);
    std::memcpy(x, u.data(), u.size() * sizeof(double));
    return 0;
  } catch (const std::exception& e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// The calling thread's residency counters of cxpp::gpu (calls, mesh uploads, pattern
// builds, DOF maps, history uploads, history commits); reset != 0 drops the cache first.
void ref_gpu_resident_stats(int reset, int64_t* out) {
  if (reset) cxpp::gpu::resident_reset();
  const cxpp::gpu::ResidentStats s = cxpp::gpu::resident_stats();
  out[0] = s.calls;
  out[1] = s.mesh_uploads;
  out[2] = s.pattern_builds;
  out[3] = s.dof_maps;
  out[4] = s.history_uploads;
  out[5] = s.history_commits;
}
#endif

}  // extern "C"
