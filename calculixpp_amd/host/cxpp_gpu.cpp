// cxpp::gpu — the reference's assembly entry points on libcxg.so (see cxpp_gpu.hpp).
//
// Each call turns the Model into the C ABI's arrays (node coordinates in index order,
// connectivity as node indices in runs of one element family, per-element E, nu and
// J2 material), assembles on the device over the pattern in which every DOF is free,
// and reduces on the host into the reference's LinearSystem: a free/free entry lands
// at (dof_eq[row], dof_eq[col]) when some element contributed a value other than 0.0
// to it (the key rule of scatter_block, src/fem/assembly.cpp:76-100); an entry in a
// prescribed column moves to the load, rhs[dof_eq[row]] -= value * prescribed[col]
// (assembly.cpp:95-96); rows of constrained DOFs are dropped (assembly.cpp:86). The
// DOF map and the constraint transform are the reference's own (build_dof_map,
// assembly.cpp:19-42, over fem::build_constraint_transform) and must reduce to single
// point constraints; external loads come from fem::external_load_vector, as in
// assemble_linear_static (assembly.cpp:250).
#include "cxpp_gpu.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>

#include "calculixpp/fem/constraint_transform.hpp"
#include "calculixpp/fem/loads.hpp"
#include "cxg.h"

namespace cxpp::gpu {
namespace {

constexpr int kDof = 3;

thread_local int t_device = 0;
thread_local cxg_context* t_ctx = nullptr;  // kept until the process ends
thread_local int t_ctx_device = -1;

cxg_context* context() {
  if (!t_ctx || t_ctx_device != t_device) {
    if (t_ctx) cxg_destroy(t_ctx);
    t_ctx = nullptr;
    cxg_context* h = nullptr;
    if (cxg_create(t_device, nullptr, &h) != CXG_OK)
      throw std::runtime_error("cxpp::gpu: no libcxg context on device " + std::to_string(t_device));
    t_ctx = h;
    t_ctx_device = t_device;
  }
  return t_ctx;
}

void check(cxg_context* c, cxg_status st) {
  if (st != CXG_OK) throw std::runtime_error(cxg_last_error(c));
}

// Device buffer owned by the context (cxg_malloc / cxg_free).
template <class T>
class Buf {
 public:
  Buf(cxg_context* c, size_t n) : c_(c), n_(n) {
    void* p = nullptr;
    check(c, cxg_malloc(c, n * sizeof(T), &p));
    p_ = static_cast<T*>(p);
  }
  Buf(cxg_context* c, const std::vector<T>& host) : Buf(c, host.size()) { put(host.data()); }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() {
    if (p_) cxg_free(c_, p_);
  }
  T* get() const { return p_; }
  void put(const T* src) { check(c_, cxg_copy_to_device(c_, p_, src, n_ * sizeof(T))); }
  std::vector<T> fetch(size_t n) const {
    std::vector<T> out(n);
    check(c_, cxg_copy_to_host(c_, out.data(), p_, n * sizeof(T)));
    return out;
  }

 private:
  cxg_context* c_;
  size_t n_;
  T* p_ = nullptr;
};

bool on_gpu_path(ElementType t) {
  return t == ElementType::C3D8 || t == ElementType::C3D20 || t == ElementType::C3D20R ||
         t == ElementType::C3D6;
}

int gauss_points(ElementType t) {
  const int n = cxg_gauss_points(static_cast<int32_t>(t));
  if (n <= 0) throw std::runtime_error("cxpp::gpu: element type outside the GPU path");
  return n;
}

// The mesh part of a Model as arrays (coordinates, element runs, elasticity, active mask);
// a node id the mesh does not know becomes -1, an error only for an active element.
ArrayModel mesh_arrays(const Model& model) {
  const Mesh& m = model.mesh;
  ArrayModel a;
  a.coords.resize(m.num_nodes() * kDof);
  for (size_t i = 0; i < m.num_nodes(); ++i)
    for (int k = 0; k < kDof; ++k) a.coords[kDof * i + size_t(k)] = m.nodes()[i].x[size_t(k)];
  const std::vector<ElasticIso> el = model.element_elastic();
  a.active = model.element_active_mask();
  a.E.resize(m.num_elements());
  a.nu.resize(m.num_elements());
  for (size_t e = 0; e < m.num_elements(); ++e) {
    const Element& elem_e = m.elements()[e];
    if (a.blocks.empty() || a.blocks.back().type != elem_e.type)
      a.blocks.push_back(ArrayModel::Block{elem_e.type, {}});
    for (Index id : elem_e.nodes) a.blocks.back().conn.push_back(m.node_index(id));
    a.E[e] = el[e].E;
    a.nu[e] = el[e].nu;
  }
  return a;
}

// An array model's active elements on the device, as runs of one family in model order.
struct DeviceMesh {
  cxg_context* c;
  std::vector<size_t> elem;       // model element index of each device element
  std::vector<int64_t> gp_first;  // first device Gauss point of each device element
  int32_t n_nodes = 0;
  int64_t n_points = 0;
  std::unique_ptr<Buf<double>> coords, E, nu;
  std::unique_ptr<Buf<int32_t>> conn;
  cxg_mesh* mesh = nullptr;

  DeviceMesh(cxg_context* ctx, const ArrayModel& am) : c(ctx) {
    const size_t ne_all = am.num_elements();
    if (am.coords.size() % kDof != 0) throw std::runtime_error("cxpp::gpu: coords must hold 3 values per node");
    if (am.E.size() != ne_all || am.nu.size() != ne_all)
      throw std::runtime_error("cxpp::gpu: E and nu must hold one value per element");
    if (!am.active.empty() && am.active.size() != ne_all)
      throw std::runtime_error("cxpp::gpu: active must hold one flag per element");
    if (am.num_nodes() > size_t(INT32_MAX)) throw std::runtime_error("cxpp::gpu: too many nodes");
    n_nodes = static_cast<int32_t>(am.num_nodes());
    std::vector<int32_t> cn;
    std::vector<double> ev, nv;
    std::vector<cxg_block> blocks;
    std::vector<size_t> block_conn;
    size_t e = 0;
    for (const ArrayModel::Block& B : am.blocks) {
      const size_t nn = size_t(nodes_per_element(B.type));
      if (nn == 0 || B.conn.size() % nn != 0)
        throw std::runtime_error("cxpp::gpu: connectivity does not match the element type");
      for (size_t k = 0; k < B.conn.size(); k += nn, ++e) {
        if (!am.active.empty() && !am.active[e]) continue;  // *MODEL CHANGE, REMOVE (assembly.cpp:238)
        if (!on_gpu_path(B.type))
          throw std::runtime_error("cxpp::gpu: element type outside the GPU path (C3D8, C3D20, "
                                   "C3D20R, C3D6)");
        const int32_t t = static_cast<int32_t>(B.type);
        if (blocks.empty() || blocks.back().type != t) {
          blocks.push_back(cxg_block{t, 0, int64_t(elem.size()), n_points, nullptr});
          block_conn.push_back(cn.size());
        }
        ++blocks.back().n_elem;
        for (size_t i = 0; i < nn; ++i) {
          const Index ni = B.conn[k + i];
          if (ni < 0 || ni >= n_nodes) throw std::runtime_error("element references unknown node");
          cn.push_back(ni);
        }
        elem.push_back(e);
        gp_first.push_back(n_points);
        n_points += gauss_points(B.type);
        ev.push_back(am.E[e]);
        nv.push_back(am.nu[e]);
      }
    }
    coords = std::make_unique<Buf<double>>(c, am.coords);
    conn = std::make_unique<Buf<int32_t>>(c, cn);
    E = std::make_unique<Buf<double>>(c, ev);
    nu = std::make_unique<Buf<double>>(c, nv);
    for (size_t b = 0; b < blocks.size(); ++b) blocks[b].conn = conn->get() + block_conn[b];
    check(c, cxg_mesh_create(c, n_nodes, int32_t(blocks.size()), blocks.data(), &mesh));
  }
  DeviceMesh(cxg_context* ctx, const Model& model) : DeviceMesh(ctx, mesh_arrays(model)) {}
  DeviceMesh(const DeviceMesh&) = delete;
  DeviceMesh& operator=(const DeviceMesh&) = delete;
  ~DeviceMesh() {
    if (mesh) cxg_mesh_destroy(mesh);
  }
};

// The pattern in which every DOF of the mesh is a free equation.
struct AllFreePattern {
  cxg_context* c;
  cxg_pattern* p = nullptr;
  int64_t nnz = 0;
  int32_t n = 0;
  AllFreePattern(cxg_context* ctx, const cxg_mesh* mesh, int32_t n_nodes) : c(ctx) {
    const int64_t nd = int64_t(n_nodes) * kDof;
    if (nd > INT32_MAX) throw std::runtime_error("cxpp::gpu: model too large for int32 equations");
    n = int32_t(nd);
    std::vector<int32_t> eq(static_cast<size_t>(n));
    for (int32_t g = 0; g < n; ++g) eq[size_t(g)] = g;
    Buf<int32_t> d(c, eq);
    check(c, cxg_pattern_build(c, mesh, d.get(), n, &p));
    int32_t nf = 0;
    check(c, cxg_pattern_info(p, &nf, &nnz));
  }
  AllFreePattern(const AllFreePattern&) = delete;
  AllFreePattern& operator=(const AllFreePattern&) = delete;
  ~AllFreePattern() {
    if (p) cxg_pattern_destroy(p);
  }
};

// build_dof_map (assembly.cpp:19-42) through the public constraint transform; the GPU
// path takes single point constraints only.
void dof_map(const Model& model, fem::LinearSystem& sys) {
  const Mesh& mesh = model.mesh;
  const size_t n_dof = mesh.num_nodes() * kDof;
  sys.prescribed.assign(n_dof, 0.0);
  sys.prescribed_amp.assign(n_dof, std::string{});
  std::vector<bool> constrained(n_dof, false);
  for (const Spc& spc : model.spcs) {
    const Index ni = mesh.node_index(spc.node_id);
    if (ni < 0) throw std::runtime_error("*BOUNDARY references unknown node");
    if (spc.comp < 1 || spc.comp > kDof) throw std::runtime_error("*BOUNDARY dof out of range");
    const size_t g = size_t(ni) * kDof + size_t(spc.comp - 1);
    constrained[g] = true;
    sys.prescribed[g] = spc.value;
    sys.prescribed_amp[g] = spc.amplitude;
  }
  sys.transform = fem::build_constraint_transform(model, constrained, sys.prescribed);
  sys.dof_eq = sys.transform.dof_eq;
  sys.n_free = sys.transform.n_free;
  sys.rhs.assign(size_t(sys.n_free), 0.0);
  for (size_t g = 0; g < n_dof; ++g) {
    const fem::DofExpansion& x = sys.transform.expansion[g];
    const bool single = sys.dof_eq[g] >= 0
                            ? x.terms.size() == 1 && x.terms[0].eq == sys.dof_eq[g] &&
                                  x.terms[0].coeff == 1.0 && x.constant == 0.0
                            : x.terms.empty() && !sys.transform.is_slave[g];
    if (!single)
      throw std::runtime_error("cxpp::gpu: *EQUATION / MPC / rigid / tie constraints are "
                               "outside the GPU path");
  }
}

// build_dof_map for single point constraints only: free equations in ascending global DOF
// order, constrained DOFs -1 with their prescribed value; the transform expands a free
// DOF to itself and a constrained one to its value (constraint_transform.cpp:25-131 for
// a model without equations).
void dof_map(const ArrayModel& am, fem::LinearSystem& sys) {
  const size_t n_dof = am.num_nodes() * kDof;
  sys.prescribed.assign(n_dof, 0.0);
  sys.prescribed_amp.assign(n_dof, std::string{});
  std::vector<bool> constrained(n_dof, false);
  for (const ArrayModel::NodalValue& spc : am.spc) {
    if (spc.node < 0 || size_t(spc.node) >= am.num_nodes())
      throw std::runtime_error("*BOUNDARY references unknown node");
    if (spc.comp < 1 || spc.comp > kDof) throw std::runtime_error("*BOUNDARY dof out of range");
    const size_t g = size_t(spc.node) * kDof + size_t(spc.comp - 1);
    constrained[g] = true;
    sys.prescribed[g] = spc.value;
  }
  fem::ConstraintTransform& T = sys.transform;
  T.dof_eq.assign(n_dof, -1);
  T.expansion.assign(n_dof, fem::DofExpansion{});
  T.is_slave.assign(n_dof, false);
  Index n_free = 0;
  for (size_t g = 0; g < n_dof; ++g) {
    if (constrained[g]) {
      T.expansion[g].constant = sys.prescribed[g];
    } else {
      T.dof_eq[g] = n_free;
      T.expansion[g].terms.push_back(fem::DofTerm{n_free, 1.0});
      ++n_free;
    }
  }
  T.n_free = n_free;
  sys.dof_eq = T.dof_eq;
  sys.n_free = n_free;
  sys.rhs.assign(size_t(n_free), 0.0);
}

// As fem::external_load_vector (src/fem/loads.cpp:578-584): a load on an unknown node or
// an out-of-range component carries nothing.
void add_external_load(const ArrayModel& am, fem::LinearSystem& sys) {
  for (const ArrayModel::NodalValue& f : am.load) {
    if (f.node < 0 || size_t(f.node) >= am.num_nodes() || f.comp < 1 || f.comp > kDof) continue;
    const Index r = sys.dof_eq[size_t(f.node) * kDof + size_t(f.comp - 1)];
    if (r >= 0) sys.rhs[size_t(r)] += f.value;
  }
}

void require_mechanical(const Model& model) {
  if (!model.springs.empty()) throw std::runtime_error("cxpp::gpu: *SPRING is outside the GPU path");
  for (Real v : fem::thermal_strain_load_vector(model))
    if (v != 0.0) throw std::runtime_error("cxpp::gpu: thermal loads are outside the GPU path");
}

// All-free CSR + live flags -> the reference's reduced COO and load correction.
void reduce(const AllFreePattern& pat, const std::vector<double>& vals,
            const std::vector<uint8_t>& live, fem::LinearSystem& sys) {
  Buf<int64_t> rp_dev(pat.c, size_t(pat.n) + 1);
  Buf<int32_t> ci_dev(pat.c, size_t(std::max<int64_t>(pat.nnz, 1)));
  check(pat.c, cxg_pattern_copy(pat.c, pat.p, rp_dev.get(), ci_dev.get()));
  const std::vector<int64_t> rp = rp_dev.fetch(size_t(pat.n) + 1);
  const std::vector<int32_t> ci = ci_dev.fetch(size_t(pat.nnz));
  sys.rows.clear();
  sys.cols.clear();
  sys.vals.clear();
  for (int32_t g = 0; g < pat.n; ++g) {
    const Index r = sys.dof_eq[size_t(g)];
    if (r < 0) continue;  // constrained row
    for (int64_t k = rp[size_t(g)]; k < rp[size_t(g) + 1]; ++k) {
      if (!live[size_t(k)]) continue;  // every contribution was 0.0
      const int32_t h = ci[size_t(k)];
      const Index col = sys.dof_eq[size_t(h)];
      if (col >= 0) {
        sys.rows.push_back(r);
        sys.cols.push_back(col);
        sys.vals.push_back(vals[size_t(k)]);
      } else if (sys.prescribed[size_t(h)] != 0.0) {
        sys.rhs[size_t(r)] -= vals[size_t(k)] * sys.prescribed[size_t(h)];
      }
    }
  }
}

void add_external_load(const Model& model, fem::LinearSystem& sys) {
  const std::vector<Real> f = fem::external_load_vector(model);
  for (size_t g = 0; g < f.size(); ++g)
    if (f[g] != 0.0 && sys.dof_eq[g] >= 0) sys.rhs[size_t(sys.dof_eq[g])] += f[g];
}

cxg_j2_material to_cxg(const ElasticIso& el, const Plastic& pl) {
  cxg_j2_material m{};
  m.E = el.E;
  m.nu = el.nu;
  m.hardening = pl.hardening == Plastic::Hardening::Kinematic  ? CXG_HARD_KINEMATIC
                : pl.hardening == Plastic::Hardening::Combined ? CXG_HARD_COMBINED
                                                               : CXG_HARD_ISOTROPIC;
  if (pl.yield.empty() || pl.yield.size() > CXG_MAX_TABLE || pl.yield.size() != pl.eqplastic.size())
    throw std::runtime_error("cxpp::gpu: *PLASTIC table outside the GPU path (1.." +
                             std::to_string(CXG_MAX_TABLE) + " points)");
  m.n_table = int32_t(pl.yield.size());
  for (size_t i = 0; i < pl.yield.size(); ++i) {
    m.eqplastic[i] = pl.eqplastic[i];
    m.yield[i] = pl.yield[i];
  }
  m.kinematic_modulus = pl.kinematic_modulus;
  return m;
}

// The material-point path covers J2 (*PLASTIC) and linear elastic elements only.
void require_material_points(const Model& model, const fem::MaterialPoints& mp) {
  for (const auto& h : model.element_hyperelastic())
    if (h && !h->empty()) throw std::runtime_error("cxpp::gpu: *HYPERELASTIC is outside the GPU path");
  for (const auto& um : model.element_user_material())
    if (um && !um->empty()) throw std::runtime_error("cxpp::gpu: *USER MATERIAL is outside the GPU path");
  if (mp.state.size() != model.mesh.num_elements())
    throw std::runtime_error("cxpp::gpu: material points do not match the mesh");
}

// J2 inputs of the device elements: the material table (one entry per distinct
// elastic + plastic pair), each element's entry (-1: elastic) and the committed history
// as SoA planes: plastic strain (6), eqplastic, back stress (6).
struct J2Inputs {
  std::vector<cxg_j2_material> mats;
  std::vector<int32_t> mat_id;
  std::vector<double> committed;
};

J2Inputs j2_inputs(const Model& model, const DeviceMesh& dm, const fem::MaterialPoints& mp) {
  J2Inputs in;
  const size_t ne = dm.elem.size();
  const std::vector<ElasticIso> el = model.element_elastic();
  const std::vector<std::optional<Plastic>> pl = model.element_plastic();
  std::map<std::vector<double>, int32_t> seen;
  in.mat_id.assign(ne, -1);
  for (size_t k = 0; k < ne; ++k) {
    const size_t e = dm.elem[k];
    if (!pl[e] || pl[e]->empty()) continue;
    const cxg_j2_material m = to_cxg(el[e], *pl[e]);
    std::vector<double> key{m.E, m.nu, double(m.hardening), double(m.n_table), m.kinematic_modulus};
    key.insert(key.end(), m.eqplastic, m.eqplastic + m.n_table);
    key.insert(key.end(), m.yield, m.yield + m.n_table);
    auto it = seen.find(key);
    if (it == seen.end()) {
      it = seen.emplace(key, int32_t(in.mats.size())).first;
      in.mats.push_back(m);
    }
    in.mat_id[k] = it->second;
  }
  const size_t P = size_t(dm.n_points);
  in.committed.assign(size_t(CXG_STATE_COMPONENTS) * P, 0.0);
  for (size_t k = 0; k < ne; ++k) {
    const std::vector<fem::MaterialState>& st = mp.state[dm.elem[k]];
    const size_t ng = size_t(gauss_points(model.mesh.elements()[dm.elem[k]].type));
    if (st.size() != ng) throw std::runtime_error("cxpp::gpu: material points do not match the mesh");
    for (size_t q = 0; q < ng; ++q) {
      const size_t pt = size_t(dm.gp_first[k]) + q;
      for (size_t i = 0; i < 6; ++i) {
        in.committed[i * P + pt] = st[q].committed_plastic_strain[i];
        in.committed[(7 + i) * P + pt] = st[q].committed_back_stress[i];
      }
      in.committed[6 * P + pt] = st[q].committed_eqplastic;
    }
  }
  return in;
}

std::vector<double> flat(const std::vector<Vec3>& v) {
  std::vector<double> h(v.size() * kDof);
  for (size_t i = 0; i < v.size(); ++i)
    for (int k = 0; k < kDof; ++k) h[kDof * i + size_t(k)] = v[i][size_t(k)];
  return h;
}

// recover_fields in both forms (mp == nullptr: every element elastic).
void recover(const Model& model, StaticFields& fields, const fem::MaterialPoints* mp) {
  require_mechanical(model);
  if (mp) require_material_points(model, *mp);
  const Mesh& mesh = model.mesh;
  const size_t nn = mesh.num_nodes();
  if (fields.displacement.size() != nn)
    throw std::runtime_error("cxpp::gpu: the displacement does not match the mesh");
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  J2Inputs j2;
  if (mp) j2 = j2_inputs(model, dm, *mp);
  const bool plastic = !j2.mats.empty();
  Buf<double> ud(c, flat(fields.displacement)), fx(c, fem::external_load_vector(model));
  Buf<double> sd(c, 6 * nn), ed(c, 6 * nn), rd(c, kDof * nn);
  Buf<double> cd(c, std::max<size_t>(j2.committed.size(), 1));
  Buf<int32_t> mid(c, std::max<size_t>(j2.mat_id.size(), 1));
  if (plastic) {
    cd.put(j2.committed.data());
    mid.put(j2.mat_id.data());
  }
  check(c, cxg_recover_fields(c, dm.mesh, dm.coords->get(), dm.E->get(), dm.nu->get(),
                              plastic ? mid.get() : nullptr, plastic ? j2.mats.data() : nullptr,
                              int32_t(j2.mats.size()), ud.get(), plastic ? cd.get() : nullptr,
                              fx.get(), ed.get(), sd.get(), rd.get(), nullptr, nullptr));
  const std::vector<double> s = sd.fetch(6 * nn), e = ed.fetch(6 * nn), r = rd.fetch(kDof * nn);
  fields.stress.assign(nn, Voigt6{});
  fields.strain.assign(nn, Voigt6{});
  fields.reaction.assign(nn, Vec3{0, 0, 0});
  for (size_t i = 0; i < nn; ++i) {
    for (size_t k = 0; k < 6; ++k) {
      fields.stress[i][k] = s[6 * i + k];
      fields.strain[i][k] = e[6 * i + k];
    }
    for (size_t k = 0; k < size_t(kDof); ++k) fields.reaction[i][k] = r[kDof * i + k];
  }
}

}  // namespace

void set_device(int device) { t_device = device; }

fem::LinearSystem assemble_linear_static(const Model& model) {
  require_mechanical(model);
  fem::LinearSystem sys;
  dof_map(model, sys);
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  const AllFreePattern pat(c, dm.mesh, dm.n_nodes);
  const size_t nnz = size_t(std::max<int64_t>(pat.nnz, 1));
  Buf<double> vals(c, nnz);
  Buf<uint8_t> live(c, nnz);
  check(c, cxg_assemble_linear(c, pat.p, dm.coords->get(), dm.E->get(), dm.nu->get(), vals.get(),
                               nullptr, nullptr, live.get()));
  reduce(pat, vals.fetch(size_t(pat.nnz)), live.fetch(size_t(pat.nnz)), sys);
  add_external_load(model, sys);
  return sys;
}

fem::LinearSystem assemble_material_tangent(const Model& model, const std::vector<Vec3>& u,
                                            fem::MaterialPoints& mp, std::vector<Real>& f_int) {
  require_mechanical(model);
  require_material_points(model, mp);
  const Mesh& mesh = model.mesh;
  if (u.size() != mesh.num_nodes()) throw std::runtime_error("cxpp::gpu: u does not match the mesh");
  fem::LinearSystem sys;
  dof_map(model, sys);
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  const size_t ne = dm.elem.size();
  const J2Inputs j2 = j2_inputs(model, dm, mp);
  const std::vector<cxg_j2_material>& mats = j2.mats;
  const std::vector<int32_t>& mat_id = j2.mat_id;
  const std::vector<double>& committed = j2.committed;
  const size_t P = size_t(dm.n_points);
  const std::vector<double> uh = flat(u);
  const AllFreePattern pat(c, dm.mesh, dm.n_nodes);
  const size_t nnz = size_t(std::max<int64_t>(pat.nnz, 1));
  Buf<double> vals(c, nnz), fd(c, uh.size()), ud(c, uh), cd(c, committed), td(c, committed.size());
  Buf<uint8_t> live(c, nnz);
  Buf<int32_t> mid(c, std::max<size_t>(ne, 1));
  if (ne) mid.put(mat_id.data());
  check(c, cxg_assemble_j2(c, pat.p, dm.coords->get(), dm.E->get(), dm.nu->get(),
                           mats.empty() ? nullptr : mid.get(), mats.empty() ? nullptr : mats.data(),
                           int32_t(mats.size()), ud.get(), cd.get(), td.get(), vals.get(),
                           fd.get(), live.get()));
  reduce(pat, vals.fetch(size_t(pat.nnz)), live.fetch(size_t(pat.nnz)), sys);
  add_external_load(model, sys);
  f_int = fd.fetch(uh.size());
  const std::vector<double> trial = td.fetch(committed.size());
  for (size_t k = 0; k < ne; ++k) {
    std::vector<fem::MaterialState>& st = mp.state[dm.elem[k]];
    for (size_t q = 0; q < st.size(); ++q) {
      const size_t pt = size_t(dm.gp_first[k]) + q;
      for (size_t i = 0; i < 6; ++i) {
        st[q].plastic_strain[i] = trial[i * P + pt];
        st[q].back_stress[i] = trial[(7 + i) * P + pt];
      }
      st[q].eqplastic = trial[6 * P + pt];
    }
  }
  return sys;
}

std::vector<Real> internal_force(const Model& model, const std::vector<Vec3>& u) {
  require_mechanical(model);
  const Mesh& mesh = model.mesh;
  if (u.size() != mesh.num_nodes()) throw std::runtime_error("cxpp::gpu: u does not match the mesh");
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  std::vector<double> uh(mesh.num_nodes() * kDof);
  for (size_t i = 0; i < mesh.num_nodes(); ++i)
    for (int k = 0; k < kDof; ++k) uh[kDof * i + size_t(k)] = u[i][size_t(k)];
  Buf<double> ud(c, uh), fd(c, uh.size());
  check(c, cxg_internal_force(c, dm.mesh, dm.coords->get(), dm.E->get(), dm.nu->get(), ud.get(),
                              fd.get()));
  return fd.fetch(uh.size());
}

size_t ArrayModel::num_elements() const {
  size_t n = 0;
  for (const Block& b : blocks) {
    const int nn = nodes_per_element(b.type);
    n += nn > 0 ? b.conn.size() / size_t(nn) : 0;
  }
  return n;
}

ArrayModel to_arrays(const Model& model) {
  ArrayModel a = mesh_arrays(model);
  const Mesh& mesh = model.mesh;
  for (const ArrayModel::Block& b : a.blocks)
    for (Index ni : b.conn)
      if (ni < 0) throw std::runtime_error("element references unknown node");
  a.rho = model.element_density();
  for (const Spc& spc : model.spcs) {
    const Index ni = mesh.node_index(spc.node_id);
    if (ni < 0) throw std::runtime_error("*BOUNDARY references unknown node");
    a.spc.push_back(ArrayModel::NodalValue{ni, spc.comp, spc.value});
  }
  const std::vector<Real> f = fem::external_load_vector(model);
  for (size_t g = 0; g < f.size(); ++g)
    if (f[g] != 0.0)
      a.load.push_back(ArrayModel::NodalValue{Index(g / kDof), int(g % kDof) + 1, f[g]});
  return a;
}

fem::LinearSystem assemble_linear_static(const ArrayModel& model) {
  fem::LinearSystem sys;
  dof_map(model, sys);
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  const AllFreePattern pat(c, dm.mesh, dm.n_nodes);
  const size_t nnz = size_t(std::max<int64_t>(pat.nnz, 1));
  Buf<double> vals(c, nnz);
  Buf<uint8_t> live(c, nnz);
  check(c, cxg_assemble_linear(c, pat.p, dm.coords->get(), dm.E->get(), dm.nu->get(), vals.get(),
                               nullptr, nullptr, live.get()));
  reduce(pat, vals.fetch(size_t(pat.nnz)), live.fetch(size_t(pat.nnz)), sys);
  add_external_load(model, sys);
  return sys;
}

std::vector<Real> internal_force(const ArrayModel& model, const std::vector<Vec3>& u) {
  if (u.size() != model.num_nodes()) throw std::runtime_error("cxpp::gpu: u does not match the mesh");
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  const std::vector<double> uh = flat(u);
  Buf<double> ud(c, uh), fd(c, uh.size());
  check(c, cxg_internal_force(c, dm.mesh, dm.coords->get(), dm.E->get(), dm.nu->get(), ud.get(),
                              fd.get()));
  return fd.fetch(uh.size());
}

std::vector<Real> element_stiffness(ElementType type, std::span<const Vec3> coords,
                                    const ElasticIso& elastic) {
  if (!on_gpu_path(type))
    throw std::runtime_error("cxpp::gpu: element type outside the GPU path (C3D8, C3D20, C3D20R, C3D6)");
  const int n = cxg_nodes_per_element(static_cast<int32_t>(type));
  if (coords.size() != size_t(n)) throw std::runtime_error("cxpp::gpu: coords do not match the element type");
  cxg_context* c = context();
  std::vector<double> x(size_t(n) * kDof);
  std::vector<int32_t> cn(static_cast<size_t>(n));
  for (int a = 0; a < n; ++a) {
    cn[size_t(a)] = a;
    for (int k = 0; k < kDof; ++k) x[size_t(kDof * a + k)] = coords[size_t(a)][size_t(k)];
  }
  Buf<double> xd(c, x), Ed(c, std::vector<double>{elastic.E}), nd(c, std::vector<double>{elastic.nu});
  Buf<int32_t> cd(c, cn);
  cxg_block b{static_cast<int32_t>(type), 1, 0, 0, cd.get()};
  cxg_mesh* mesh = nullptr;
  check(c, cxg_mesh_create(c, n, 1, &b, &mesh));
  std::unique_ptr<cxg_mesh, void (*)(cxg_mesh*)> hold(mesh, cxg_mesh_destroy);
  const AllFreePattern pat(c, mesh, n);
  Buf<double> vals(c, size_t(pat.nnz));
  check(c, cxg_assemble_linear(c, pat.p, xd.get(), Ed.get(), nd.get(), vals.get(), nullptr,
                               nullptr, nullptr));
  Buf<int64_t> rp_dev(c, size_t(pat.n) + 1);
  Buf<int32_t> ci_dev(c, size_t(pat.nnz));
  check(c, cxg_pattern_copy(c, pat.p, rp_dev.get(), ci_dev.get()));
  const std::vector<int64_t> rp = rp_dev.fetch(size_t(pat.n) + 1);
  const std::vector<int32_t> ci = ci_dev.fetch(size_t(pat.nnz));
  const std::vector<double> v = vals.fetch(size_t(pat.nnz));
  const size_t nd3 = size_t(kDof * n);
  std::vector<Real> K(nd3 * nd3, 0.0);
  for (size_t r = 0; r < nd3; ++r)
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) K[r * nd3 + size_t(ci[size_t(k)])] = v[size_t(k)];
  return K;
}

std::vector<Real> solve_sparse(const std::vector<Index>& rows, const std::vector<Index>& cols,
                               const std::vector<Real>& vals, Index n,
                               const std::vector<Real>& rhs, compute::SolverKind) {
  if (rows.size() != cols.size() || rows.size() != vals.size() || rhs.size() != size_t(n) || n < 0)
    throw std::runtime_error("cxpp::gpu::solve_sparse: inconsistent system sizes");
  // COO -> CSR with duplicates summed in triplet order
  std::vector<int64_t> rp(size_t(n) + 1, 0);
  for (Index r : rows) {
    if (r < 0 || r >= n) throw std::runtime_error("cxpp::gpu::solve_sparse: row out of range");
    ++rp[size_t(r) + 1];
  }
  for (size_t i = 0; i < size_t(n); ++i) rp[i + 1] += rp[i];
  std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
  std::vector<int32_t> ci(rows.size());
  std::vector<double> va(rows.size());
  for (size_t k = 0; k < rows.size(); ++k) {
    if (cols[k] < 0 || cols[k] >= n) throw std::runtime_error("cxpp::gpu::solve_sparse: column out of range");
    const int64_t at = fill[size_t(rows[k])]++;
    ci[size_t(at)] = cols[k];
    va[size_t(at)] = vals[k];
  }
  std::vector<int64_t> out_rp(size_t(n) + 1, 0);
  std::vector<int32_t> out_ci;
  std::vector<double> out_va;
  out_ci.reserve(ci.size());
  out_va.reserve(va.size());
  std::vector<std::pair<int32_t, double>> row;
  for (size_t r = 0; r < size_t(n); ++r) {
    row.clear();
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) row.emplace_back(ci[size_t(k)], va[size_t(k)]);
    std::stable_sort(row.begin(), row.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t k = 0; k < row.size(); ++k) {
      if (k > 0 && row[k].first == out_ci.back()) {
        out_va.back() += row[k].second;
      } else {
        out_ci.push_back(row[k].first);
        out_va.push_back(row[k].second);
      }
    }
    out_rp[r + 1] = int64_t(out_ci.size());
  }
  cxg_context* c = context();
  Buf<int64_t> rpd(c, out_rp);
  Buf<int32_t> cid(c, out_ci);
  Buf<double> vad(c, out_va), bd(c, rhs), xd(c, std::vector<double>(size_t(n), 0.0));
  int32_t iters = 0;
  double rel = 0.0;
  check(c, cxg_pcg(c, n, rpd.get(), cid.get(), vad.get(), bd.get(), xd.get(), 1e-13,
                   std::max<int32_t>(100, 20 * n), &iters, &rel));
  return xd.fetch(size_t(n));
}

void recover_fields(const Model& model, StaticFields& fields) { recover(model, fields, nullptr); }

void recover_fields(const Model& model, StaticFields& fields, fem::MaterialPoints& mp) {
  recover(model, fields, &mp);
}

std::vector<std::vector<Voigt6>> recover_gauss_stress(const Model& model,
                                                      const std::vector<Vec3>& u) {
  require_mechanical(model);
  const Mesh& mesh = model.mesh;
  if (u.size() != mesh.num_nodes()) throw std::runtime_error("cxpp::gpu: u does not match the mesh");
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  const size_t P = size_t(dm.n_points);
  Buf<double> ud(c, flat(u)), gd(c, std::max<size_t>(6 * P, 1));
  check(c, cxg_recover_fields(c, dm.mesh, dm.coords->get(), dm.E->get(), dm.nu->get(), nullptr,
                              nullptr, 0, ud.get(), nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr, gd.get()));
  const std::vector<double> g = gd.fetch(6 * P);
  std::vector<std::vector<Voigt6>> out(mesh.num_elements());  // removed elements: empty
  for (size_t k = 0; k < dm.elem.size(); ++k) {
    const size_t ng = size_t(gauss_points(mesh.elements()[dm.elem[k]].type));
    std::vector<Voigt6>& v = out[dm.elem[k]];
    v.resize(ng);
    for (size_t q = 0; q < ng; ++q)
      for (size_t i = 0; i < 6; ++i) v[q][i] = g[6 * (size_t(dm.gp_first[k]) + q) + i];
  }
  return out;
}

namespace {

// assemble_mass / assemble_geometric_stiffness: the k_ab I3 matrices on the all-free
// pattern, reduced like the stiffness with the eliminated-column load discarded.
// sys holds the DOF map; rho per model element (mass) or gp_stress per model element
// (geometric stiffness).
fem::LinearSystem assemble_scalar(fem::LinearSystem sys, cxg_context* c, const DeviceMesh& dm,
                                  int kind, const std::vector<Real>& rho,
                                  const std::vector<std::vector<Voigt6>>* gp_stress) {
  const AllFreePattern pat(c, dm.mesh, dm.n_nodes);
  const size_t nnz = size_t(std::max<int64_t>(pat.nnz, 1));
  Buf<double> vals(c, nnz);
  Buf<uint8_t> live(c, nnz);
  if (kind < 2) {
    std::vector<double> r(std::max<size_t>(dm.elem.size(), 1), 0.0);
    for (size_t k = 0; k < dm.elem.size(); ++k) r[k] = dm.elem[k] < rho.size() ? rho[dm.elem[k]] : 0.0;
    Buf<double> rd(c, r);
    check(c, cxg_assemble_mass(c, pat.p, dm.coords->get(), rd.get(), kind == 1, vals.get(),
                               live.get()));
  } else {
    const size_t P = size_t(dm.n_points);
    std::vector<double> g(std::max<size_t>(6 * P, 1), 0.0);  // missing stress: zero
    for (size_t k = 0; k < dm.elem.size(); ++k) {
      const size_t e = dm.elem[k];
      if (e >= gp_stress->size()) continue;
      const std::vector<Voigt6>& es = (*gp_stress)[e];
      const size_t ng = size_t(k + 1 < dm.elem.size() ? dm.gp_first[k + 1] : dm.n_points) -
                        size_t(dm.gp_first[k]);
      for (size_t q = 0; q < ng && q < es.size(); ++q)
        for (size_t i = 0; i < 6; ++i) g[6 * (size_t(dm.gp_first[k]) + q) + i] = es[q][i];
    }
    Buf<double> gd(c, g);
    check(c, cxg_assemble_geometric(c, pat.p, dm.coords->get(), gd.get(), vals.get(), live.get()));
  }
  reduce(pat, vals.fetch(size_t(pat.nnz)), live.fetch(size_t(pat.nnz)), sys);
  std::fill(sys.rhs.begin(), sys.rhs.end(), 0.0);  // mass and K_geo never drive a load
  return sys;
}

}  // namespace

fem::LinearSystem assemble_mass(const Model& model, bool lumped) {
  require_mechanical(model);
  if (!model.point_masses.empty())
    throw std::runtime_error("cxpp::gpu: *MASS point masses are outside the GPU path");
  fem::LinearSystem sys;
  dof_map(model, sys);
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  return assemble_scalar(std::move(sys), c, dm, lumped ? 1 : 0, model.element_density(), nullptr);
}

fem::LinearSystem assemble_mass(const ArrayModel& model, bool lumped) {
  fem::LinearSystem sys;
  dof_map(model, sys);
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  return assemble_scalar(std::move(sys), c, dm, lumped ? 1 : 0, model.rho, nullptr);
}

fem::LinearSystem assemble_geometric_stiffness(const Model& model,
                                               const std::vector<std::vector<Voigt6>>& gp_stress) {
  require_mechanical(model);
  fem::LinearSystem sys;
  dof_map(model, sys);
  cxg_context* c = context();
  const DeviceMesh dm(c, model);
  return assemble_scalar(std::move(sys), c, dm, 2, {}, &gp_stress);
}

}  // namespace cxpp::gpu
