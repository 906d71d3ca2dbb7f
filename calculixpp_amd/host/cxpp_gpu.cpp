// cxpp::gpu — the reference's assembly entry points on libcxg.so (see cxpp_gpu.hpp).
//
// A call turns the Model into the C ABI's arrays (node coordinates in index order,
// connectivity as node indices in runs of one element family, per-element E, nu and
// J2 material), assembles on the device over the pattern in which every DOF is free,
// and reduces on the host into the reference's LinearSystem: a free/free entry lands
// at (dof_eq[row], dof_eq[col]) when some element contributed a value other than 0.0
// to it (the key rule of scatter_block, src/fem/assembly.cpp:76-100); an entry in a
// prescribed column moves to the load, rhs[dof_eq[row]] -= value * prescribed[col]
// (assembly.cpp:95-96); rows of constrained DOFs are dropped (assembly.cpp:86).
// External loads come from fem::external_load_vector, as in assemble_linear_static
// (assembly.cpp:250).
//
// The numbering of LinearSystem (LinearSystem::dof_eq, assembly.hpp:25-27) for models
// whose only constraints are *BOUNDARY cards: equations in ascending global DOF order
// over the DOFs no card names, each prescribed DOF expanding to its value (the
// transform the reference's fem::build_constraint_transform returns when a model has no
// equations, constraint_transform.cpp:25-131). Models with *EQUATION / MPC / rigid /
// coupling / tie constraints go through build_constraint_transform itself and are
// refused unless it reduces to single point constraints.
//
// Residency (cxpp_gpu.hpp): the last model of the calling thread stays on the device,
// keyed by its address and a fingerprint of the uploaded arrays.
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
constexpr size_t kHist = CXG_STATE_COMPONENTS;

thread_local int t_device = 0;
thread_local cxg_context* t_ctx = nullptr;  // kept until the process ends
thread_local int t_ctx_device = -1;

void drop_resident();

cxg_context* context() {
  if (!t_ctx || t_ctx_device != t_device) {
    drop_resident();  // device memory of the previous context
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
    check(c, cxg_malloc(c, std::max<size_t>(n, 1) * sizeof(T), &p));
    p_ = static_cast<T*>(p);
  }
  Buf(cxg_context* c, const std::vector<T>& host) : Buf(c, host.size()) { put(host.data()); }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() {
    if (p_) cxg_free(c_, p_);
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void put(const T* src) { check(c_, cxg_copy_to_device(c_, p_, src, n_ * sizeof(T))); }
  void put(const std::vector<T>& v) {
    if (v.size() != n_) throw std::runtime_error("cxpp::gpu: internal buffer size mismatch");
    put(v.data());
  }
  std::vector<T> fetch(size_t n) const {
    std::vector<T> out(n);
    check(c_, cxg_copy_to_host(c_, out.data(), p_, n * sizeof(T)));
    return out;
  }
  void fetch_into(std::vector<T>& out) const {
    out.resize(n_);
    check(c_, cxg_copy_to_host(c_, out.data(), p_, n_ * sizeof(T)));
  }

 private:
  cxg_context* c_;
  size_t n_;
  T* p_ = nullptr;
};

template <class T>
using BufPtr = std::unique_ptr<Buf<T>>;

// 64-bit fingerprint of host arrays (multiply-rotate per 8-byte word), used to tell
// whether a model seen before still holds the same numbers.
class Fingerprint {
 public:
  void word(uint64_t w) {
    h_ ^= w * 0x9e3779b97f4a7c15ull;
    h_ = ((h_ << 29) | (h_ >> 35)) * 0xbf58476d1ce4e5b9ull;
  }
  void bytes(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
      uint64_t w;
      std::memcpy(&w, c + i, 8);
      word(w);
    }
    uint64_t tail = uint64_t(n) << 56;
    if (i < n) std::memcpy(&tail, c + i, n - i);
    word(tail);
  }
  template <class T>
  void vec(const std::vector<T>& v) {
    word(v.size());
    bytes(v.data(), v.size() * sizeof(T));
  }
  void flags(const std::vector<bool>& v) {
    word(v.size());
    uint64_t acc = 0;
    for (size_t i = 0; i < v.size(); ++i) {
      acc = (acc << 1) | uint64_t(v[i]);
      if (i % 64 == 63) word(acc), acc = 0;
    }
    word(acc);
  }
  void text(const std::string& s) {
    word(s.size());
    bytes(s.data(), s.size());
  }
  uint64_t value() const { return h_; }

 private:
  uint64_t h_ = 0x243f6a8885a308d3ull;
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

// Everything of an array model that goes to the device with its mesh.
uint64_t mesh_fingerprint(const ArrayModel& a) {
  Fingerprint f;
  f.vec(a.coords);
  f.word(a.blocks.size());
  for (const ArrayModel::Block& b : a.blocks) {
    f.word(uint64_t(b.type));
    f.vec(b.conn);
  }
  f.vec(a.E);
  f.vec(a.nu);
  f.flags(a.active);
  return f.value();
}

// An array model's active elements on the device, as runs of one family in model order.
struct DeviceMesh {
  cxg_context* c;
  std::vector<size_t> elem;       // model element index of each device element
  std::vector<int64_t> gp_first;  // first device Gauss point of each device element
  int32_t n_nodes = 0;
  int64_t n_points = 0;
  BufPtr<double> coords, E, nu;
  BufPtr<int32_t> conn;
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
  DeviceMesh(const DeviceMesh&) = delete;
  DeviceMesh& operator=(const DeviceMesh&) = delete;
  ~DeviceMesh() {
    if (mesh) cxg_mesh_destroy(mesh);
  }
};

// The pattern in which every DOF of the mesh is a free equation, with a host copy of its
// row pointers and column indices (what the reduction to the reference's COO walks).
struct AllFreePattern {
  cxg_context* c;
  cxg_pattern* p = nullptr;
  int64_t nnz = 0;
  int32_t n = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
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
    Buf<int64_t> rp_dev(c, size_t(n) + 1);
    Buf<int32_t> ci_dev(c, size_t(nnz));
    check(c, cxg_pattern_copy(c, p, rp_dev.get(), ci_dev.get()));
    rp = rp_dev.fetch(size_t(n) + 1);
    ci = ci_dev.fetch(size_t(nnz));
  }
  AllFreePattern(const AllFreePattern&) = delete;
  AllFreePattern& operator=(const AllFreePattern&) = delete;
  ~AllFreePattern() {
    if (p) cxg_pattern_destroy(p);
  }
};

// J2 inputs of the device elements: the material table (one entry per distinct
// elastic + plastic pair) and each element's entry (-1: elastic).
struct J2Table {
  std::vector<cxg_j2_material> mats;
  std::vector<int32_t> mat_id;
};

// The calling thread's resident model. Members are released in reverse order of
// declaration, so buffers and the pattern go before the mesh they belong to.
struct Resident {
  const void* owner = nullptr;
  uint64_t mesh_fp = 0;
  std::unique_ptr<DeviceMesh> dm;
  std::unique_ptr<AllFreePattern> pat;
  BufPtr<double> vals;
  BufPtr<uint8_t> live;
  BufPtr<double> u, f;  // 3 doubles per node
  // DOF map of the model's constraints
  bool dof_valid = false;
  uint64_t dof_fp = 0;
  fem::LinearSystem dof;  // prescribed, prescribed_amp, transform, dof_eq, n_free
  // J2 material table of the device elements
  bool j2_valid = false;
  uint64_t j2_fp = 0;
  J2Table j2;
  BufPtr<int32_t> mid;
  // committed / trial history (13 planes of n_points) and host copies of what the
  // device holds
  BufPtr<double> committed, trial;
  std::vector<double> committed_dev, trial_dev;
  bool has_committed = false, has_trial = false;
};

thread_local Resident* t_res = nullptr;  // released by resident_reset or a model change
thread_local ResidentStats t_stats;

void drop_resident() {
  delete t_res;
  t_res = nullptr;
}

// The resident copy of the array model `am` seen through `owner` (the caller's Model or
// ArrayModel): reused when both the owner and the fingerprint match, else uploaded.
Resident& resident(const ArrayModel& am, const void* owner) {
  cxg_context* c = context();
  ++t_stats.calls;
  const uint64_t fp = mesh_fingerprint(am);
  if (!t_res || !t_res->dm || t_res->owner != owner || t_res->mesh_fp != fp) {
    drop_resident();
    auto r = std::make_unique<Resident>();
    r->dm = std::make_unique<DeviceMesh>(c, am);
    r->owner = owner;
    r->mesh_fp = fp;
    const size_t nd = size_t(r->dm->n_nodes) * kDof;
    r->u = std::make_unique<Buf<double>>(c, nd);
    r->f = std::make_unique<Buf<double>>(c, nd);
    t_res = r.release();
    ++t_stats.mesh_uploads;
  }
  return *t_res;
}

AllFreePattern& pattern(Resident& R) {
  if (!R.pat) {
    cxg_context* c = context();
    R.pat = std::make_unique<AllFreePattern>(c, R.dm->mesh, R.dm->n_nodes);
    R.vals = std::make_unique<Buf<double>>(c, size_t(R.pat->nnz));
    R.live = std::make_unique<Buf<uint8_t>>(c, size_t(R.pat->nnz));
    ++t_stats.pattern_builds;
  }
  return *R.pat;
}

// The SPC-only constraint transform: free equations in ascending global DOF order over
// the DOFs no card names; a free DOF expands to itself, a prescribed one to its value
// with itself as its one SPC master (what build_constraint_transform gives for a model
// without equations).
fem::ConstraintTransform spc_transform(const std::vector<char>& fixed,
                                       const std::vector<Real>& value) {
  const size_t n_dof = fixed.size();
  fem::ConstraintTransform T;
  T.dof_eq.assign(n_dof, -1);
  T.expansion.assign(n_dof, fem::DofExpansion{});
  T.is_slave.assign(n_dof, false);
  T.spc_terms.assign(n_dof, {});
  Index next = 0;
  for (size_t g = 0; g < n_dof; ++g) {
    if (fixed[g]) {
      T.expansion[g].constant = value[g];
      T.spc_terms[g].push_back(fem::DofTerm{Index(g), 1.0});
    } else {
      T.expansion[g].terms.push_back(fem::DofTerm{next, 1.0});
      T.dof_eq[g] = next++;
    }
  }
  T.n_free = next;
  return T;
}

void adopt_transform(fem::LinearSystem& sys) {
  sys.dof_eq = sys.transform.dof_eq;
  sys.n_free = sys.transform.n_free;
  sys.rhs.assign(size_t(sys.n_free), 0.0);
}

bool spc_only(const Model& model) {
  return model.equations.empty() && model.mpcs.empty() && model.rigid_bodies.empty() &&
         model.couplings.empty() && model.ties.empty();
}

// LinearSystem numbering of a reference Model (see the top of this file). The cards are
// applied in deck order, so a DOF named twice takes its last value and amplitude.
void number_dofs(const Model& model, fem::LinearSystem& sys) {
  const Mesh& mesh = model.mesh;
  const size_t n_dof = mesh.num_nodes() * kDof;
  std::vector<char> fixed(n_dof, 0);
  std::vector<Real> value(n_dof, 0.0);
  std::vector<std::string> amp(n_dof);
  for (const Spc& card : model.spcs) {
    const Index node = mesh.node_index(card.node_id);
    if (node < 0) throw std::runtime_error("*BOUNDARY references unknown node");
    if (card.comp < 1 || card.comp > kDof) throw std::runtime_error("*BOUNDARY dof out of range");
    const size_t dof = size_t(node) * kDof + size_t(card.comp - 1);
    fixed[dof] = 1;
    value[dof] = card.value;
    amp[dof] = card.amplitude;
  }
  sys.prescribed = value;
  sys.prescribed_amp = std::move(amp);
  if (spc_only(model)) {
    sys.transform = spc_transform(fixed, value);
  } else {
    // equations present: the reference's own transform, which must come out SPC-only
    const std::vector<bool> as_bool(fixed.begin(), fixed.end());
    sys.transform = fem::build_constraint_transform(model, as_bool, value);
    for (size_t g = 0; g < n_dof; ++g) {
      const fem::DofExpansion& x = sys.transform.expansion[g];
      const Index eq = sys.transform.dof_eq[g];
      const bool plain = eq >= 0 ? x.terms.size() == 1 && x.terms[0].eq == eq &&
                                       x.terms[0].coeff == 1.0 && x.constant == 0.0
                                 : x.terms.empty() && !sys.transform.is_slave[g];
      if (!plain)
        throw std::runtime_error("cxpp::gpu: *EQUATION / MPC / rigid / tie constraints are "
                                 "outside the GPU path");
    }
  }
  adopt_transform(sys);
}

// The DOF map of `model` through the resident cache (rebuilt when the constraint cards
// change).
void dof_map(const Model& model, Resident& R, fem::LinearSystem& sys) {
  Fingerprint f;
  f.word(model.mesh.num_nodes());
  f.word(model.spcs.size());
  for (const Spc& s : model.spcs) {
    f.word(uint64_t(s.node_id));
    f.word(uint64_t(s.comp));
    f.bytes(&s.value, sizeof(Real));
    f.text(s.amplitude);
  }
  f.word(model.equations.size() + model.mpcs.size() + model.rigid_bodies.size() +
         model.couplings.size() + model.ties.size());
  if (!spc_only(model) || !R.dof_valid || R.dof_fp != f.value()) {
    number_dofs(model, R.dof);
    R.dof_valid = spc_only(model);
    R.dof_fp = f.value();
    ++t_stats.dof_maps;
  }
  sys.prescribed = R.dof.prescribed;
  sys.prescribed_amp = R.dof.prescribed_amp;
  sys.transform = R.dof.transform;
  adopt_transform(sys);
}

// The same numbering for an ArrayModel's single point constraints.
void dof_map(const ArrayModel& am, fem::LinearSystem& sys) {
  const size_t n_dof = am.num_nodes() * kDof;
  std::vector<char> fixed(n_dof, 0);
  std::vector<Real> value(n_dof, 0.0);
  for (const ArrayModel::NodalValue& card : am.spc) {
    if (card.node < 0 || size_t(card.node) >= am.num_nodes())
      throw std::runtime_error("*BOUNDARY references unknown node");
    if (card.comp < 1 || card.comp > kDof) throw std::runtime_error("*BOUNDARY dof out of range");
    const size_t dof = size_t(card.node) * kDof + size_t(card.comp - 1);
    fixed[dof] = 1;
    value[dof] = card.value;
  }
  sys.prescribed = value;
  sys.prescribed_amp.assign(n_dof, std::string{});
  sys.transform = spc_transform(fixed, value);
  adopt_transform(sys);
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

void add_external_load(const Model& model, fem::LinearSystem& sys) {
  const std::vector<Real> f = fem::external_load_vector(model);
  for (size_t g = 0; g < f.size(); ++g)
    if (f[g] != 0.0 && sys.dof_eq[g] >= 0) sys.rhs[size_t(sys.dof_eq[g])] += f[g];
}

void require_mechanical(const Model& model) {
  if (!model.springs.empty()) throw std::runtime_error("cxpp::gpu: *SPRING is outside the GPU path");
  for (Real v : fem::thermal_strain_load_vector(model))
    if (v != 0.0) throw std::runtime_error("cxpp::gpu: thermal loads are outside the GPU path");
}

// All-free CSR + live flags -> the reference's reduced COO and load correction.
void reduce(const AllFreePattern& pat, const std::vector<double>& vals,
            const std::vector<uint8_t>& live, fem::LinearSystem& sys) {
  sys.rows.clear();
  sys.cols.clear();
  sys.vals.clear();
  for (int32_t g = 0; g < pat.n; ++g) {
    const Index r = sys.dof_eq[size_t(g)];
    if (r < 0) continue;  // constrained row
    for (int64_t k = pat.rp[size_t(g)]; k < pat.rp[size_t(g) + 1]; ++k) {
      if (!live[size_t(k)]) continue;  // every contribution was 0.0
      const int32_t h = pat.ci[size_t(k)];
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

J2Table j2_table(const Model& model, const DeviceMesh& dm) {
  J2Table t;
  const size_t ne = dm.elem.size();
  const std::vector<ElasticIso> el = model.element_elastic();
  const std::vector<std::optional<Plastic>> pl = model.element_plastic();
  std::map<std::vector<double>, int32_t> seen;
  t.mat_id.assign(ne, -1);
  for (size_t k = 0; k < ne; ++k) {
    const size_t e = dm.elem[k];
    if (!pl[e] || pl[e]->empty()) continue;
    const cxg_j2_material m = to_cxg(el[e], *pl[e]);
    std::vector<double> key{m.E, m.nu, double(m.hardening), double(m.n_table), m.kinematic_modulus};
    key.insert(key.end(), m.eqplastic, m.eqplastic + m.n_table);
    key.insert(key.end(), m.yield, m.yield + m.n_table);
    auto it = seen.find(key);
    if (it == seen.end()) {
      it = seen.emplace(key, int32_t(t.mats.size())).first;
      t.mats.push_back(m);
    }
    t.mat_id[k] = it->second;
  }
  return t;
}

// The model's J2 table on the device (re-uploaded when the materials change).
const J2Table& resident_j2(const Model& model, Resident& R) {
  J2Table t = j2_table(model, *R.dm);
  Fingerprint f;
  f.vec(t.mat_id);
  f.vec(t.mats);
  if (!R.j2_valid || R.j2_fp != f.value() || R.j2.mat_id != t.mat_id) {
    R.j2 = std::move(t);
    R.mid = std::make_unique<Buf<int32_t>>(context(), R.j2.mat_id);
    R.j2_fp = f.value();
    R.j2_valid = true;
  }
  return R.j2;
}

// The committed history of mp on the device: compared with what the device holds and
// copied only when it differs; when it equals the trial state the last call returned (the
// caller committed), the commit is replayed on the device instead of copying. Only the
// points of J2 elements are taken: the elastic law neither reads nor advances a history,
// so the device holds zeros there whatever the host's slots hold.
void resident_committed(const Model& model, Resident& R, const fem::MaterialPoints& mp,
                        const J2Table& j2) {
  cxg_context* c = context();
  const DeviceMesh& dm = *R.dm;
  const size_t P = size_t(dm.n_points);
  std::vector<double> cur(kHist * P, 0.0);
  for (size_t k = 0; k < dm.elem.size(); ++k) {
    if (j2.mat_id[k] < 0) continue;
    const std::vector<fem::MaterialState>& st = mp.state[dm.elem[k]];
    const size_t ng = size_t(gauss_points(model.mesh.elements()[dm.elem[k]].type));
    if (st.size() != ng) throw std::runtime_error("cxpp::gpu: material points do not match the mesh");
    for (size_t q = 0; q < ng; ++q) {
      const size_t pt = size_t(dm.gp_first[k]) + q;
      for (size_t i = 0; i < 6; ++i) {
        cur[i * P + pt] = st[q].committed_plastic_strain[i];
        cur[(7 + i) * P + pt] = st[q].committed_back_stress[i];
      }
      cur[6 * P + pt] = st[q].committed_eqplastic;
    }
  }
  const size_t bytes = cur.size() * sizeof(double);
  if (!R.committed) {
    R.committed = std::make_unique<Buf<double>>(c, cur.size());
    R.trial = std::make_unique<Buf<double>>(c, cur.size());
  }
  if (R.has_committed && std::memcmp(cur.data(), R.committed_dev.data(), bytes) == 0) return;
  if (R.has_trial && std::memcmp(cur.data(), R.trial_dev.data(), bytes) == 0) {
    check(c, cxg_commit(c, int64_t(P), R.committed->get(), R.trial->get()));
    R.committed_dev = R.trial_dev;
    ++t_stats.history_commits;
  } else {
    R.committed->put(cur);
    R.committed_dev = std::move(cur);
    ++t_stats.history_uploads;
  }
  R.has_committed = true;
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
  Resident& R = resident(mesh_arrays(model), &model);
  const DeviceMesh& dm = *R.dm;
  const J2Table* j2 = nullptr;
  if (mp) {
    j2 = &resident_j2(model, R);
    if (!j2->mats.empty()) resident_committed(model, R, *mp, *j2);
  }
  const bool plastic = j2 && !j2->mats.empty();
  R.u->put(flat(fields.displacement));
  Buf<double> fx(c, fem::external_load_vector(model));
  Buf<double> sd(c, 6 * nn), ed(c, 6 * nn), rd(c, kDof * nn);
  check(c, cxg_recover_fields(c, dm.mesh, dm.coords->get(), dm.E->get(), dm.nu->get(),
                              plastic ? R.mid->get() : nullptr, plastic ? j2->mats.data() : nullptr,
                              plastic ? int32_t(j2->mats.size()) : 0, R.u->get(),
                              plastic ? R.committed->get() : nullptr, fx.get(), ed.get(), sd.get(),
                              rd.get(), nullptr, nullptr));
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

// One element of a family with nodes 0..n-1 on the device: mesh and all-free pattern,
// kept per family for the element-level entry points.
struct OneElement {
  std::unique_ptr<Buf<int32_t>> conn;
  cxg_mesh* mesh = nullptr;
  std::unique_ptr<AllFreePattern> pat;
  ~OneElement() {
    pat.reset();
    if (mesh) cxg_mesh_destroy(mesh);
  }
};

thread_local OneElement* t_one[8] = {};
thread_local cxg_context* t_one_ctx = nullptr;

const OneElement& one_element(ElementType type) {
  cxg_context* c = context();
  if (t_one_ctx != c) {  // a new context: the old objects belonged to the old one
    for (OneElement*& o : t_one) o = nullptr;  // (their context is gone; not freed)
    t_one_ctx = c;
  }
  const int t = static_cast<int>(type);
  if (t < 0 || t >= 8) throw std::runtime_error("cxpp::gpu: unknown element type");
  if (!t_one[t]) {
    const int n = cxg_nodes_per_element(t);
    auto o = std::make_unique<OneElement>();
    std::vector<int32_t> cn(static_cast<size_t>(n));
    for (int a = 0; a < n; ++a) cn[size_t(a)] = a;
    o->conn = std::make_unique<Buf<int32_t>>(c, cn);
    cxg_block b{t, 1, 0, 0, o->conn->get()};
    check(c, cxg_mesh_create(c, n, 1, &b, &o->mesh));
    o->pat = std::make_unique<AllFreePattern>(c, o->mesh, n);
    t_one[t] = o.release();
  }
  return *t_one[t];
}

// Dense (3n)^2 row-major matrix of one element from its all-free CSR values.
std::vector<Real> dense_element(const AllFreePattern& pat, const std::vector<double>& v) {
  const size_t nd = size_t(pat.n);
  std::vector<Real> K(nd * nd, 0.0);
  for (size_t r = 0; r < nd; ++r)
    for (int64_t k = pat.rp[r]; k < pat.rp[r + 1]; ++k) K[r * nd + size_t(pat.ci[size_t(k)])] = v[size_t(k)];
  return K;
}

void require_element(ElementType type, size_t n_coords) {
  if (!on_gpu_path(type))
    throw std::runtime_error("cxpp::gpu: element type outside the GPU path (C3D8, C3D20, C3D20R, C3D6)");
  if (n_coords != size_t(cxg_nodes_per_element(static_cast<int32_t>(type))))
    throw std::runtime_error("cxpp::gpu: coords do not match the element type");
}

// E and nu of an isotropic elastic D (elastic_iso_D, element.cpp:410-422: lambda off the
// normal diagonal, mu on the shear diagonal).
ElasticIso elastic_of(const fem::D6& D) {
  const Real lam = D[0][1], mu = D[3][3];
  return ElasticIso{mu * (3.0 * lam + 2.0 * mu) / (lam + mu), lam / (2.0 * (lam + mu))};
}

}  // namespace

void set_device(int device) { t_device = device; }

ResidentStats resident_stats() { return t_stats; }

void resident_reset() {
  drop_resident();
  t_stats = ResidentStats{};
}

fem::LinearSystem assemble_linear_static(const Model& model) {
  require_mechanical(model);
  cxg_context* c = context();
  Resident& R = resident(mesh_arrays(model), &model);
  fem::LinearSystem sys;
  dof_map(model, R, sys);
  const AllFreePattern& pat = pattern(R);
  const DeviceMesh& dm = *R.dm;
  check(c, cxg_assemble_linear(c, pat.p, dm.coords->get(), dm.E->get(), dm.nu->get(),
                               R.vals->get(), nullptr, nullptr, R.live->get()));
  reduce(pat, R.vals->fetch(size_t(pat.nnz)), R.live->fetch(size_t(pat.nnz)), sys);
  add_external_load(model, sys);
  return sys;
}

fem::LinearSystem assemble_material_tangent(const Model& model, const std::vector<Vec3>& u,
                                            fem::MaterialPoints& mp, std::vector<Real>& f_int) {
  require_mechanical(model);
  require_material_points(model, mp);
  if (u.size() != model.mesh.num_nodes()) throw std::runtime_error("cxpp::gpu: u does not match the mesh");
  cxg_context* c = context();
  Resident& R = resident(mesh_arrays(model), &model);
  fem::LinearSystem sys;
  dof_map(model, R, sys);
  const AllFreePattern& pat = pattern(R);
  const DeviceMesh& dm = *R.dm;
  const J2Table& j2 = resident_j2(model, R);
  resident_committed(model, R, mp, j2);
  R.u->put(flat(u));
  const bool plastic = !j2.mats.empty();
  check(c, cxg_assemble_j2(c, pat.p, dm.coords->get(), dm.E->get(), dm.nu->get(),
                           plastic ? R.mid->get() : nullptr, plastic ? j2.mats.data() : nullptr,
                           int32_t(j2.mats.size()), R.u->get(), R.committed->get(), R.trial->get(),
                           R.vals->get(), R.f->get(), R.live->get()));
  reduce(pat, R.vals->fetch(size_t(pat.nnz)), R.live->fetch(size_t(pat.nnz)), sys);
  add_external_load(model, sys);
  f_int = R.f->fetch(R.f->size());
  R.trial->fetch_into(R.trial_dev);
  R.has_trial = true;
  // trial slots of the J2 points; elastic points keep theirs, as ElasticIsoMaterial
  // leaves the state alone (element.cpp:859-871)
  const size_t P = size_t(dm.n_points);
  const std::vector<double>& trial = R.trial_dev;
  for (size_t k = 0; k < dm.elem.size(); ++k) {
    if (j2.mat_id[k] < 0) continue;
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
  if (u.size() != model.mesh.num_nodes()) throw std::runtime_error("cxpp::gpu: u does not match the mesh");
  cxg_context* c = context();
  Resident& R = resident(mesh_arrays(model), &model);
  const DeviceMesh& dm = *R.dm;
  R.u->put(flat(u));
  check(c, cxg_internal_force(c, dm.mesh, dm.coords->get(), dm.E->get(), dm.nu->get(),
                              R.u->get(), R.f->get()));
  return R.f->fetch(R.f->size());
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
  Resident& R = resident(model, &model);
  const AllFreePattern& pat = pattern(R);
  const DeviceMesh& dm = *R.dm;
  check(c, cxg_assemble_linear(c, pat.p, dm.coords->get(), dm.E->get(), dm.nu->get(),
                               R.vals->get(), nullptr, nullptr, R.live->get()));
  reduce(pat, R.vals->fetch(size_t(pat.nnz)), R.live->fetch(size_t(pat.nnz)), sys);
  add_external_load(model, sys);
  return sys;
}

std::vector<Real> internal_force(const ArrayModel& model, const std::vector<Vec3>& u) {
  if (u.size() != model.num_nodes()) throw std::runtime_error("cxpp::gpu: u does not match the mesh");
  cxg_context* c = context();
  Resident& R = resident(model, &model);
  const DeviceMesh& dm = *R.dm;
  R.u->put(flat(u));
  check(c, cxg_internal_force(c, dm.mesh, dm.coords->get(), dm.E->get(), dm.nu->get(),
                              R.u->get(), R.f->get()));
  return R.f->fetch(R.f->size());
}

std::vector<Real> element_stiffness(ElementType type, std::span<const Vec3> coords,
                                    const ElasticIso& elastic) {
  require_element(type, coords.size());
  cxg_context* c = context();
  const OneElement& one = one_element(type);
  Buf<double> xd(c, flat(std::vector<Vec3>(coords.begin(), coords.end())));
  Buf<double> Ed(c, std::vector<double>{elastic.E}), nd(c, std::vector<double>{elastic.nu});
  Buf<double> vals(c, size_t(one.pat->nnz));
  check(c, cxg_assemble_linear(c, one.pat->p, xd.get(), Ed.get(), nd.get(), vals.get(), nullptr,
                               nullptr, nullptr));
  return dense_element(*one.pat, vals.fetch(size_t(one.pat->nnz)));
}

J2Material::J2Material(const ElasticIso& elastic, const Plastic& plastic)
    : elastic_(elastic), plastic_(plastic) {
  if (plastic.empty())
    throw std::runtime_error("J2PlasticMaterial needs a non-empty *PLASTIC table");
  (void)to_cxg(elastic_, plastic_);  // the table fits the device form
}

fem::MaterialResponse J2Material::evaluate(const Voigt6& strain, fem::MaterialState& state) const {
  cxg_context* c = context();
  const cxg_j2_material m = to_cxg(elastic_, plastic_);
  std::vector<double> hc(kHist);
  for (size_t i = 0; i < 6; ++i) {
    hc[i] = state.committed_plastic_strain[i];
    hc[7 + i] = state.committed_back_stress[i];
  }
  hc[6] = state.committed_eqplastic;
  Buf<int32_t> mid(c, std::vector<int32_t>{0});
  Buf<double> ed(c, std::vector<double>(strain.begin(), strain.end()));
  Buf<double> cd(c, hc), td(c, kHist), sd(c, 6), dd(c, 36);
  check(c, cxg_material_evaluate(c, 1, mid.get(), &m, 1, nullptr, nullptr, ed.get(), cd.get(),
                                 td.get(), sd.get(), dd.get()));
  const std::vector<double> ht = td.fetch(kHist), s = sd.fetch(6), D = dd.fetch(36);
  for (size_t i = 0; i < 6; ++i) {
    state.plastic_strain[i] = ht[i];
    state.back_stress[i] = ht[7 + i];
  }
  state.eqplastic = ht[6];
  fem::MaterialResponse r;
  for (size_t i = 0; i < 6; ++i) {
    r.stress[i] = s[i];
    for (size_t j = 0; j < 6; ++j) r.tangent[i][j] = D[6 * i + j];
  }
  return r;
}

fem::ElementResponse element_tangent_force(ElementType type, std::span<const Vec3> coords,
                                           std::span<const Vec3> ue,
                                           const fem::MaterialModel& material,
                                           std::span<fem::MaterialState> state) {
  require_element(type, coords.size());
  if (ue.size() != coords.size()) throw std::runtime_error("cxpp::gpu: ue does not match the element");
  const size_t ng = size_t(gauss_points(type));
  if (state.size() != ng) throw std::runtime_error("cxpp::gpu: state needs one entry per Gauss point");
  const auto* iso = dynamic_cast<const fem::ElasticIsoMaterial*>(&material);
  const auto* j2 = dynamic_cast<const J2Material*>(&material);
  if (!iso && !j2)
    throw std::runtime_error("cxpp::gpu::element_tangent_force: material outside the GPU path "
                             "(fem::ElasticIsoMaterial or cxpp::gpu::J2Material)");
  cxg_context* c = context();
  const OneElement& one = one_element(type);
  const ElasticIso el = iso ? elastic_of(iso->tangent()) : j2->elastic();
  const cxg_j2_material m = j2 ? to_cxg(j2->elastic(), j2->plastic()) : cxg_j2_material{};
  std::vector<double> hc(kHist * ng, 0.0);
  for (size_t q = 0; q < ng; ++q) {
    for (size_t i = 0; i < 6; ++i) {
      hc[i * ng + q] = state[q].committed_plastic_strain[i];
      hc[(7 + i) * ng + q] = state[q].committed_back_stress[i];
    }
    hc[6 * ng + q] = state[q].committed_eqplastic;
  }
  const size_t nd = kDof * coords.size();
  Buf<double> xd(c, flat(std::vector<Vec3>(coords.begin(), coords.end())));
  Buf<double> ud(c, flat(std::vector<Vec3>(ue.begin(), ue.end())));
  Buf<double> Ed(c, std::vector<double>{el.E}), nud(c, std::vector<double>{el.nu});
  Buf<int32_t> mid(c, std::vector<int32_t>{0});
  Buf<double> cd(c, hc), td(c, hc.size()), fd(c, nd), vals(c, size_t(one.pat->nnz));
  check(c, cxg_assemble_j2(c, one.pat->p, xd.get(), Ed.get(), nud.get(), j2 ? mid.get() : nullptr,
                           j2 ? &m : nullptr, j2 ? 1 : 0, ud.get(), cd.get(), td.get(), vals.get(),
                           fd.get(), nullptr));
  fem::ElementResponse out;
  out.Ke = dense_element(*one.pat, vals.fetch(size_t(one.pat->nnz)));
  out.fe = fd.fetch(nd);
  if (j2) {  // the elastic law leaves the state alone (element.cpp:859-871)
    const std::vector<double> ht = td.fetch(hc.size());
    for (size_t q = 0; q < ng; ++q) {
      for (size_t i = 0; i < 6; ++i) {
        state[q].plastic_strain[i] = ht[i * ng + q];
        state[q].back_stress[i] = ht[(7 + i) * ng + q];
      }
      state[q].eqplastic = ht[6 * ng + q];
    }
  }
  return out;
}

std::vector<Real> HipBackend::solve_sparse(const std::vector<Index>& rows,
                                           const std::vector<Index>& cols,
                                           const std::vector<Real>& vals, Index n,
                                           const std::vector<Real>& rhs,
                                           compute::SolverKind solver) const {
  if (rows.size() != cols.size() || rows.size() != vals.size() || n < 0 || rhs.size() != size_t(n))
    throw std::runtime_error("cxpp::gpu::solve_sparse: inconsistent system sizes");
  double tol = 0.0;
  switch (solver) {
    case compute::SolverKind::CG: tol = 1e-12; break;
    case compute::SolverKind::Direct: tol = 1e-13; break;
    default: throw std::runtime_error("cxpp::gpu::solve_sparse: unknown SolverKind");
  }
  // COO -> CSR, each row's triplets in column order with duplicates summed in triplet
  // order
  std::vector<int64_t> rp(size_t(n) + 1, 0);
  for (Index r : rows) {
    if (r < 0 || r >= n) throw std::runtime_error("cxpp::gpu::solve_sparse: row out of range");
    ++rp[size_t(r) + 1];
  }
  for (size_t i = 0; i < size_t(n); ++i) rp[i + 1] += rp[i];
  std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
  std::vector<std::pair<int32_t, double>> ent(rows.size());
  for (size_t k = 0; k < rows.size(); ++k) {
    if (cols[k] < 0 || cols[k] >= n) throw std::runtime_error("cxpp::gpu::solve_sparse: column out of range");
    ent[size_t(fill[size_t(rows[k])]++)] = {cols[k], vals[k]};
  }
  std::vector<int64_t> out_rp(size_t(n) + 1, 0);
  std::vector<int32_t> out_ci;
  std::vector<double> out_va;
  out_ci.reserve(ent.size());
  out_va.reserve(ent.size());
  for (size_t r = 0; r < size_t(n); ++r) {
    auto b = ent.begin() + rp[r], e = ent.begin() + rp[r + 1];
    std::stable_sort(b, e, [](const auto& x, const auto& y) { return x.first < y.first; });
    for (auto it = b; it != e; ++it) {
      if (it != b && it->first == out_ci.back()) {
        out_va.back() += it->second;
      } else {
        out_ci.push_back(it->first);
        out_va.push_back(it->second);
      }
    }
    out_rp[r + 1] = int64_t(out_ci.size());
  }
  cxg_context* c = context();
  Buf<int64_t> rpd(c, out_rp);
  Buf<int32_t> cid(c, out_ci);
  Buf<double> vad(c, out_va), bd(c, rhs), xd(c, std::vector<double>(size_t(n), 0.0));
  const int32_t max_iter = int32_t(std::min<int64_t>(INT32_MAX, std::max<int64_t>(100, 20 * int64_t(n))));
  int32_t iters = 0;
  double rel = 0.0;
  const cxg_status st = cxg_pcg(c, n, rpd.get(), cid.get(), vad.get(), bd.get(), xd.get(), tol,
                                max_iter, &iters, &rel);
  if (st == CXG_ERR_STATE) {  // the tolerance was not reached
    if (solver == compute::SolverKind::CG)
      throw std::runtime_error("CG did not converge (relative residual " + std::to_string(rel) +
                               " after " + std::to_string(iters) + " iterations)");
    throw std::runtime_error("Direct solve by device PCG did not reach 1e-13 (relative residual " +
                             std::to_string(rel) + " after " + std::to_string(iters) +
                             " iterations)");
  }
  check(c, st);
  return xd.fetch(size_t(n));
}

const compute::ComputeBackend& hip_backend() {
  static const HipBackend backend;
  return backend;
}

std::vector<Real> solve_sparse(const std::vector<Index>& rows, const std::vector<Index>& cols,
                               const std::vector<Real>& vals, Index n,
                               const std::vector<Real>& rhs, compute::SolverKind solver) {
  return hip_backend().solve_sparse(rows, cols, vals, n, rhs, solver);
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
  Resident& R = resident(mesh_arrays(model), &model);
  const DeviceMesh& dm = *R.dm;
  const size_t P = size_t(dm.n_points);
  R.u->put(flat(u));
  Buf<double> gd(c, 6 * P);
  check(c, cxg_recover_fields(c, dm.mesh, dm.coords->get(), dm.E->get(), dm.nu->get(), nullptr,
                              nullptr, 0, R.u->get(), nullptr, nullptr, nullptr, nullptr, nullptr,
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
fem::LinearSystem assemble_scalar(fem::LinearSystem sys, Resident& R, int kind,
                                  const std::vector<Real>& rho,
                                  const std::vector<std::vector<Voigt6>>* gp_stress) {
  cxg_context* c = context();
  const AllFreePattern& pat = pattern(R);
  const DeviceMesh& dm = *R.dm;
  if (kind < 2) {
    std::vector<double> r(dm.elem.size(), 0.0);
    for (size_t k = 0; k < dm.elem.size(); ++k) r[k] = dm.elem[k] < rho.size() ? rho[dm.elem[k]] : 0.0;
    Buf<double> rd(c, r);
    check(c, cxg_assemble_mass(c, pat.p, dm.coords->get(), rd.get(), kind == 1, R.vals->get(),
                               R.live->get()));
  } else {
    const size_t P = size_t(dm.n_points);
    std::vector<double> g(6 * P, 0.0);  // missing stress: zero
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
    check(c, cxg_assemble_geometric(c, pat.p, dm.coords->get(), gd.get(), R.vals->get(),
                                    R.live->get()));
  }
  reduce(pat, R.vals->fetch(size_t(pat.nnz)), R.live->fetch(size_t(pat.nnz)), sys);
  std::fill(sys.rhs.begin(), sys.rhs.end(), 0.0);  // mass and K_geo never drive a load
  return sys;
}

}  // namespace

fem::LinearSystem assemble_mass(const Model& model, bool lumped) {
  require_mechanical(model);
  if (!model.point_masses.empty())
    throw std::runtime_error("cxpp::gpu: *MASS point masses are outside the GPU path");
  Resident& R = resident(mesh_arrays(model), &model);
  fem::LinearSystem sys;
  dof_map(model, R, sys);
  return assemble_scalar(std::move(sys), R, lumped ? 1 : 0, model.element_density(), nullptr);
}

fem::LinearSystem assemble_mass(const ArrayModel& model, bool lumped) {
  fem::LinearSystem sys;
  dof_map(model, sys);
  Resident& R = resident(model, &model);
  return assemble_scalar(std::move(sys), R, lumped ? 1 : 0, model.rho, nullptr);
}

fem::LinearSystem assemble_geometric_stiffness(const Model& model,
                                               const std::vector<std::vector<Voigt6>>& gp_stress) {
  require_mechanical(model);
  Resident& R = resident(mesh_arrays(model), &model);
  fem::LinearSystem sys;
  dof_map(model, R, sys);
  return assemble_scalar(std::move(sys), R, 2, {}, &gp_stress);
}

}  // namespace cxpp::gpu
