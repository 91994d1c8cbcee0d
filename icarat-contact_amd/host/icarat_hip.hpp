// C++ drop-in layer over libicx: the reference's analysis / optimisation call
// surface for structured quad4 / hexa8 meshes, with every hot loop on the GPU.
//
// Class and method names, argument meaning and the order of calls follow the
// reference (paths relative to the reference repository root):
//   FEM, Force, Value, Optimize ........ src/problem/base.hpp:16-129, src/optimization/opt_base.hpp:14-76
//   StructuredModel .................... MeshIcarat (src/io/mesh_icarat.hpp:69-136) + BCIcarat
//                                        (src/io/boundary_condition_icarat.hpp:68-396) + LinearElastic
//                                        ([material], src/problem/elasticity.cpp:19-32)
//   LinearAnalysis::solve(name, log) ... src/analysis/linear.hpp:54-104 + Solvers::solve
//                                        (src/analysis/solver/solvers.cpp:28-62)
//   Filter ............................. src/optimization/filter.hpp:17-185
//   FilterHelmholtz .................... src/optimization/filter_helm.hpp:23-173
//   OC ................................. src/optimization/opt-solver/OCSolver.hpp:13-80
//   MMASolver .......................... src/optimization/opt-solver/MMASolver.hpp:62-116
// Differences: vectors are std::vector<double> (Eigen is not required); the model
// object replaces the reference's vector<LinearElastic>/vector<Node> pair; solver
// names gain "HIP_PCG", "HIP_PCG_FUSED" and "HIP_PCG_MIXED" ("Eigen_CG" runs the
// FP64 plain form at the reference's tolerance, "Eigen_LDLT" the plain form at 1e-14).
// As in the reference, an unknown solver name or a degenerate filter prints to
// std::cerr and exits with status 1.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/icx.h"
#include "toml_lite.hpp"
#include "toml_job.hpp"

namespace icarat {

// ---------------------------------------------------------------------------------
inline void check(icx_ctx *ctx, icx_status rc, const char *what) {
  if (rc != ICX_OK && rc != ICX_ERR_NOCONV)
    throw std::runtime_error(std::string(what) + ": " + icx_last_error(ctx));
}

// one context (device 0, one stream) per process
inline icx_ctx *default_context() {
  static icx_ctx *ctx = [] {
    icx_ctx *c = nullptr;
    if (icx_ctx_create(0, &c) != ICX_OK) throw std::runtime_error("no HIP device available");
    return c;
  }();
  return ctx;
}

struct FEM {
  int neq = 0, numeq = 0, ndim = 0, nelx = 0, voigt = 0, dofnp = 0, numnp = 0;
};

// A std::vector<double> with the few Eigen::VectorXd calls the reference's drivers make
// on Force / Value (dot, setZero, norm), so that e.g. dis.val.dot(force.fext) reads as
// there.
struct DVec : std::vector<double> {
  using std::vector<double>::vector;
  double dot(const std::vector<double> &o) const {
    double s = 0.0;
    for (size_t i = 0; i < size(); ++i) s += (*this)[i] * o[i];
    return s;
  }
  void setZero() { std::fill(begin(), end(), 0.0); }
  double norm() const { return std::sqrt(dot(*this)); }
};

struct Force {
  explicit Force(int numeq) : fext_org(numeq, 0.0), fext(numeq, 0.0), fint(numeq, 0.0), fres(numeq, 0.0) {}
  DVec fext_org, fext, fint, fres;
};

struct Value {
  explicit Value(int numeq) : val(numeq, 0.0) {}
  DVec val;
};

// Node and element records with the reference's field names (src/problem/base.hpp:37-106).
// For a GPU job the element record is a handle: its design, volume, connectivity and von
// Mises value mirror the job on the device (see LinearElastic below).
struct Node {
  int ID = 0;
  int dof[6] = {0, 0, 0, 0, 0, 0};
  double x[3] = {0, 0, 0};
  double val[3] = {0, 0, 0};
  double reaction[3] = {0, 0, 0};
};
struct Element {
  int numdof = 0, ne = 0, ipmax = 0;
  std::string eType;
  int ID = 0;
  std::vector<int> nodeID;
  double volume = 0.0;
};
class StructuredModel;

class Optimize {
 public:
  Optimize(int design, int constraint, double xmin_, double xmax_)
      : design_s(design), const_h(constraint), dfds(design), dgds(constraint * design),
        xmin(design, xmin_), xmax(design, xmax_) {}
  int isConvergence(double epsilon, double minstep) {
    if (std::abs(object_old - object_f) / object_f < epsilon && optstep > minstep) {
      std::cout << "\n--------------------Optimized--------------------" << std::endl;
      return 0;
    }
    object_old = object_f;
    return 1;
  }
  std::vector<double> design_s, const_h, dfds, dgds, xmin, xmax;
  int optstep = 0;
  double object_f = 0.0;

 private:
  double object_old = 0.0;
};

// ---------------------------------------------------------------------------------
// structured box model on the device
class StructuredModel {
 public:
  // The box job of a reference job file: mesh, material, Dirichlet ranges, DOF map and
  // tractions in the reference's order, by the same front end as icx_job_create_toml.
  StructuredModel(const toml_lite::Table &config, FEM &fem) : fem_(fem) {
    TomlJob tj;
    std::string why;
    if (box_job_from_toml(ctx(), config, tj, why) != ICX_OK) fatal(why);
    job_ = tj.job;
    heat_ = tj.heat;
    prob_ = tj.prob;
    design_s0_ = tj.design0;
    numeq_ = tj.numeq;
    from_spec_ = true;
    finish(tj.ndim);
    current() = this;
  }

  // MeshMDPA::setParameter / generate (src/io/mesh_mdpa.hpp:84-216): the mesh of an .mdpa
  // file.  [mesh] type is a string or a list of element types; the GPU path takes meshes
  // of one type, quad4 or hexa8, and rejects a file that holds elements of another
  // listed type.  [material] as for the box (elastic only).
  StructuredModel(const toml_lite::Table &config, FEM &fem, const std::string &mdpa) : fem_(fem) {
    static const char *kTags[][2] = {{"tria3", "2D3"},   {"tria6", "2D6"},   {"quad4", "2D4"}, {"quad8", "2D8"},
                                     {"tetra4", "3D4"},  {"tetra10", "3D10"}, {"prism6", "3D6"},
                                     {"prism15", "3D15"}, {"hexa8", "3D8"},   {"hexa20", "3D20"}};
    std::vector<std::string> types;
    const auto &tv = config.at("mesh", "type");
    if (tv.kind == toml_lite::Value::ARR)
      for (const auto &t : tv.as_array()) types.push_back(t.as_string());
    else
      types.push_back(tv.as_string());
    int eltype = 0;
    for (const auto &t : types)
      if (t == "quad4" || t == "hexa8") eltype = t == "quad4" ? ICX_QUAD4 : ICX_HEXA8;
    if (!eltype) fatal("no element type of the GPU path (quad4, hexa8) in [mesh] type");
    {
      std::ifstream in(mdpa);
      if (!in) fatal("cannot open " + mdpa);
      std::string line;
      while (std::getline(in, line)) {
        if (line.rfind("Begin Elements", 0) != 0) continue;
        for (const auto &t : types) {
          if (t == "quad4" || t == "hexa8") continue;
          for (const auto &kt : kTags)
            if (t == kt[0] && line.find(kt[1]) != std::string::npos)
              fatal("element type " + t + " is not supported by the GPU path (quad4, hexa8)");
        }
      }
    }
    const int ndim = config.at("mesh", "dimension").as_int();
    if ((ndim == 2) != (eltype == ICX_QUAD4)) fatal("[mesh] dimension does not match the element type");
    std::string state = config.at("material", "state").as_string();
    if (ndim == 3) state = "3D";
    icx_mesh m{};
    m.eltype = eltype;
    m.state = state == "plane_stress" ? ICX_PLANE_STRESS : (state == "plane_strain" ? ICX_PLANE_STRAIN : ICX_SOLID3D);
    const auto &young = config.at("material", "young").as_array();
    m.young1 = young[0].as_double();
    m.young2 = young[1].as_double();
    m.poisson = config.at("material", "poisson").as_double();
    m.penalty = config.at("material", "penalty").as_double();
    design_s0_ = config.at("material", "designs_0").as_double();
    if (icx_mdpa_read(mdpa.c_str(), eltype, &m.numnp, &m.nelx, nullptr, nullptr) != ICX_OK)
      fatal("cannot read the " + std::string(eltype == ICX_QUAD4 ? "quad4" : "hexa8") + " mesh of " + mdpa);
    std::vector<double> x(3 * (size_t)m.numnp);
    std::vector<int> conn((size_t)m.nelx * eltype);
    if (icx_mdpa_read(mdpa.c_str(), eltype, &m.numnp, &m.nelx, x.data(), conn.data()) != ICX_OK)
      fatal("cannot read " + mdpa);
    m.x = x.data();
    m.conn = conn.data();
    prob_.eltype = eltype;
    prob_.state = m.state;
    prob_.young1 = m.young1;
    prob_.young2 = m.young2;
    prob_.poisson = m.poisson;
    prob_.penalty = m.penalty;
    check(ctx(), icx_job_create_mesh(ctx(), &m, &job_), "icx_job_create_mesh");
    finish(ndim);
    current() = this;
  }
  ~StructuredModel() {
    if (current() == this) current() = nullptr;
    if (job_) icx_job_destroy(job_);
  }
  StructuredModel(const StructuredModel &) = delete;
  StructuredModel &operator=(const StructuredModel &) = delete;

  // BCIcarat::input_dirich (a box job built from its job file has its ranges already)
  void input_dirich(const toml_lite::Table &config) {
    if (from_spec_) return;
    for (const auto &entry : config.at("boundary_condition", "dirichlet").as_array()) {
      const auto &e = entry.as_array();
      int flag[6];
      double val[6], range[6];
      for (int k = 0; k < 6; ++k) {
        flag[k] = e[0].as_array()[k].as_int();
        val[k] = e[1].as_array()[k].as_double();
      }
      for (int d = 0; d < 3; ++d)
        for (int k = 0; k < 2; ++k) range[2 * d + k] = e[2].as_array()[d].as_array()[k].as_double();
      check(ctx(), icx_add_dirichlet(job_, flag, val, range), "icx_add_dirichlet");
    }
  }
  // BCIcarat::makeDOF
  int makeDOF() {
    if (from_spec_) return numeq_;
    int numeq = 0;
    check(ctx(), icx_make_dof(job_, &numeq), "icx_make_dof");
    return numeq;
  }
  // BCIcarat::input_neumann (accumulates into force.fext_org)
  void input_neumann(const toml_lite::Table &config, Force &force) {
    if (!from_spec_ && config.has("boundary_condition", "neumann"))
      for (const auto &entry : config.at("boundary_condition", "neumann").as_array()) {
        const auto &e = entry.as_array();
        double val[3], range[6];
        for (int k = 0; k < 3; ++k) val[k] = e[0].as_array()[k].as_double();
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < 2; ++k) range[2 * d + k] = e[1].as_array()[d].as_array()[k].as_double();
        check(ctx(), icx_add_neumann(job_, val, range), "icx_add_neumann");
      }
    force.fext_org.assign(fem_.numeq, 0.0);
    check(ctx(), icx_get_fext(job_, force.fext_org.data()), "icx_get_fext");
  }
  bool heat() const { return heat_; }
  // element[i].setDesignS(d[i]) for all elements
  void setDesignS(const std::vector<double> &d) {
    check(ctx(), icx_set_design(job_, d.data()), "icx_set_design");
    design_ = d;
    design_dirty_ = false;
  }
  // one element's design, from an element handle (LinearElastic::setDesignS): kept on
  // the host and sent with the next solve
  void setElementDesign(int e, double d) {
    if (design_.size() != (size_t)fem_.nelx) design_.assign(fem_.nelx, design_s0_);
    design_[e] = d;
    design_dirty_ = true;
  }
  void flushDesign() {
    if (design_dirty_) setDesignS(design_);
  }
  // the element-mean von Mises stresses of the last solve (icx_mises, once per solve)
  const std::vector<double> &mises() {
    if (mises_solve_ != solves_) {
      mises_.assign(fem_.nelx, 0.0);
      check(ctx(), icx_mises(job_, mises_.data()), "icx_mises");
      mises_solve_ = solves_;
    }
    return mises_;
  }
  void solved() { ++solves_; }
  // the model the latest StructuredModel built (the reference's MMASolver(n, m) names no model)
  static StructuredModel *&current() {
    static StructuredModel *m = nullptr;
    return m;
  }
  const std::vector<double> &design_s() const { return design_; }
  double design_s0() const { return design_s0_; }
  const std::vector<double> &volume() const { return volume_; }
  const icx_problem &problem() const { return prob_; }
  icx_job *job() const { return job_; }
  icx_ctx *ctx() const { return default_context(); }
  static void fatal(const std::string &msg) {
    std::cerr << msg << std::endl;
    std::exit(1);
  }

 private:
  void finish(int ndim) {
    int numnp, nelx, neq;
    icx_job_sizes(job_, &numnp, &nelx, &neq);
    fem_.ndim = ndim;
    fem_.dofnp = heat_ ? 1 : ndim;
    fem_.voigt = heat_ ? ndim : (ndim == 2 ? 3 : 6);
    fem_.numnp = numnp;
    fem_.nelx = nelx;
    fem_.neq = neq;
    volume_.assign(nelx, 0.0);
    icx_element_volume(job_, volume_.data());
  }

  FEM &fem_;
  icx_job *job_ = nullptr;
  icx_problem prob_{};
  double design_s0_ = 0.0;
  bool heat_ = false;
  bool from_spec_ = false;  // box job built by box_job_from_toml (ranges, DOF map, loads done)
  int numeq_ = 0;
  bool design_dirty_ = false;
  long long solves_ = 0, mises_solve_ = -1;
  std::vector<double> mises_;
  std::vector<double> volume_, design_;
};

// ---------------------------------------------------------------------------------
// The reference's set-up objects over a GPU job, so that a driver written against
// MeshIcarat / LinearElastic / Node / BCIcarat (src/io/mesh_icarat.hpp:69-136,
// src/problem/elasticity.hpp:15-100, src/problem/base.hpp:37-106,
// src/io/boundary_condition_icarat.hpp:68-396) keeps its set-up code:
//   MeshIcarat<LinearElastic, Node> mesh(fem, config);  mesh.setParameter("structure");
//   vector<LinearElastic> element (one per element, setParameter(config));
//   vector<Node> node(fem.numnp);  mesh.generate(element, node);
//   BCIcarat<LinearElastic, Node> BC(fem, element, node);  BC.input_dirich(config);
//   fem.numeq = BC.makeDOF();  BC.input_neumann(config, force);
// generate() builds the job from the job file (box mesh, material, Dirichlet ranges, DOF
// map and tractions, by box_job_from_toml); the element records then hold handles: an
// element's setDesignS goes to the job with the next solve, makeMisesStress reads the
// device's von Mises value.  The analysis and filter classes below take (fem, element,
// node, ...) as the reference's do.
class LinearElastic : public Element {
 public:
  LinearElastic(int voigt, int ne_, int ipmax_, int numdof_) : voigt_(voigt) {
    ne = ne_;
    ipmax = ipmax_;
    numdof = numdof_;
  }
  // Elastic::setParameter (src/problem/elasticity.cpp:19-32): [material] young, poisson,
  // penalty, state and the initial design designs_0
  void setParameter(const toml_lite::Table &config) {
    const auto &young = config.at("material", "young").as_array();
    young1_ = young.at(0).as_double();
    young2_ = young.at(1).as_double();
    poisson_ = config.at("material", "poisson").as_double();
    pp_ = config.at("material", "penalty").as_double();
    mattype_ = config.has("material", "state") ? config.at("material", "state").as_string() : "3D";
    design_s_ = config.at("material", "designs_0").as_double();
  }
  void setDesignS(const double &d) {
    design_s_ = d;
    if (model_) model_->setElementDesign(ID, d);
  }
  // Elastic::makeMisesStress: the element-mean von Mises stress of the last solve
  void makeMisesStress(FEM &, std::vector<Node> &) {
    if (model_) mises_ = model_->mises()[ID];
  }
  const double &design_s() const { return design_s_; }
  const double &young1() const { return young1_; }
  const double &young2() const { return young2_; }
  const double &poisson() const { return poisson_; }
  const double &pp() const { return pp_; }
  const double &mises() const { return mises_; }
  const std::string &mattype() const { return mattype_; }
  // SIMP modulus of the current design (src/problem/base.hpp SIMP)
  double young() const {
    if (young2_ - young1_ >= 0.0) return (1.0 - std::pow(design_s_, pp_)) * young1_ + std::pow(design_s_, pp_) * young2_;
    return std::pow(1.0 - design_s_, pp_) * young1_ + (1.0 - std::pow(1.0 - design_s_, pp_)) * young2_;
  }
  StructuredModel &model() const {
    if (!model_) StructuredModel::fatal("element not bound to a mesh: call MeshIcarat::generate first");
    return *model_;
  }
  void bind(StructuredModel *m) { model_ = m; }

 private:
  int voigt_;
  StructuredModel *model_ = nullptr;
  std::string mattype_;
  double design_s_ = 0.0, young1_ = 0.0, young2_ = 0.0, poisson_ = 0.0, pp_ = 0.0, mises_ = 0.0;
};

template <class E, class N>
class MeshIcarat {
 public:
  MeshIcarat(FEM &fem, const toml_lite::Table &config) : fem_(fem), config_(config) {}
  // MeshIcarat::setParameter: the [mesh] sizes (mesh_icarat.hpp:69-128)
  void setParameter(const std::string &) {
    const std::string type = config_.at("mesh", "type").as_string();
    if (type != "quad4" && type != "hexa8")
      StructuredModel::fatal("element type " + type + " is not supported by the GPU path (quad4, hexa8)");
    ne_ = type == "quad4" ? 4 : 8;
    const auto &div = config_.at("mesh", "division").as_array();
    fem_.ndim = config_.at("mesh", "dimension").as_int();
    const bool heat = config_.has("material", "type") && config_.at("material", "type").as_string() == "Heat";
    fem_.dofnp = heat ? 1 : fem_.ndim;
    fem_.voigt = heat ? fem_.ndim : (fem_.ndim == 2 ? 3 : 6);
    const int d0 = div.at(0).as_int(), d1 = div.at(1).as_int(), d2 = fem_.ndim == 3 ? div.at(2).as_int() : 1;
    fem_.nelx = d0 * d1 * d2;
    fem_.numnp = (d0 + 1) * (d1 + 1) * (fem_.ndim == 3 ? d2 + 1 : 1);
    fem_.neq = fem_.numnp * fem_.dofnp;
  }
  int numeleTypes() const { return 1; }
  int ne(int) const { return ne_; }
  int ipmax(int) const { return ne_; }
  int nelx(int) const { return fem_.nelx; }
  // MeshIcarat::generate: the job, and the element / node records bound to it
  void generate(std::vector<E> &element, std::vector<N> &node) {
    model_.reset(new StructuredModel(config_, fem_));
    if ((int)element.size() != fem_.nelx || (int)node.size() != fem_.numnp)
      StructuredModel::fatal("MeshIcarat::generate: element / node vectors do not match the mesh");
    std::vector<double> x(3 * (size_t)fem_.numnp);
    std::vector<int> conn((size_t)fem_.nelx * ne_);
    check(model_->ctx(), icx_get_mesh(model_->job(), x.data(), conn.data()), "icx_get_mesh");
    for (int n = 0; n < fem_.numnp; ++n) {
      node[n].ID = n;
      for (int d = 0; d < 3; ++d) node[n].x[d] = x[3 * (size_t)n + d];
    }
    const std::vector<double> &vol = model_->volume();
    for (int e = 0; e < fem_.nelx; ++e) {
      E &el = element[e];
      el.ID = e;
      el.eType = ne_ == 4 ? "quad4" : "hexa8";
      el.nodeID.assign(conn.begin() + (size_t)e * ne_, conn.begin() + (size_t)(e + 1) * ne_);
      el.volume = vol[e];
      el.bind(model_.get());
      el.setDesignS(el.design_s());
    }
  }
  StructuredModel &model() { return *model_; }

 private:
  FEM &fem_;
  const toml_lite::Table &config_;
  int ne_ = 0;
  std::unique_ptr<StructuredModel> model_;
};

template <class E, class N>
class BCIcarat {
 public:
  BCIcarat(FEM &fem, std::vector<E> &element, std::vector<N> &node) : fem_(fem), element_(element), node_(node) {}
  void input_dirich(const toml_lite::Table &config) { element_.front().model().input_dirich(config); }
  // BCIcarat::makeDOF: numeq, and node.dof as countDOF numbers it
  int makeDOF() {
    StructuredModel &m = element_.front().model();
    const int numeq = m.makeDOF();
    std::vector<int> dof((size_t)fem_.numnp * fem_.dofnp);
    check(m.ctx(), icx_get_dof_map(m.job(), dof.data()), "icx_get_dof_map");
    for (int n = 0; n < fem_.numnp; ++n)
      for (int c = 0; c < fem_.dofnp; ++c) node_[n].dof[c] = dof[(size_t)n * fem_.dofnp + c];
    return numeq;
  }
  void input_neumann(const toml_lite::Table &config, Force &force) {
    element_.front().model().input_neumann(config, force);
  }

 private:
  FEM &fem_;
  std::vector<E> &element_;
  std::vector<N> &node_;
};

// ---------------------------------------------------------------------------------
template <class E = void, class N = void>
class LinearAnalysis {
 public:
  LinearAnalysis(FEM &fem, StructuredModel &model, Force &force, Value &value)
      : fem_(fem), model_(model), force_(force), value_(value) {}
  // LinearAnalysis<E, N>(fem, element, node, force, value) (src/analysis/linear.hpp:18-19):
  // the solve also writes node.val (divideValIntoNode)
  LinearAnalysis(FEM &fem, std::vector<E> &element, std::vector<N> &node, Force &force, Value &value)
      : fem_(fem), model_(element.front().model()), force_(force), value_(value) {
    node_vals_ = [this, &node]() {
      std::vector<double> v(3 * (size_t)fem_.numnp);
      check(model_.ctx(), icx_get_node_values(model_.job(), v.data()), "icx_get_node_values");
      for (int n = 0; n < fem_.numnp; ++n)
        for (int c = 0; c < 3; ++c) node[n].val[c] = v[3 * (size_t)n + c];
    };
  }

  // LinearAnalysis::solve with the Solvers facade dispatch on the name
  void solve(const std::string &solvName, bool log) {
    int form;
    double tol = tol_;
    if (solvName == "HIP_PCG" || solvName == "Eigen_CG") form = ICX_SOLVE_PLAIN;
    else if (solvName == "Eigen_LDLT") { form = ICX_SOLVE_PLAIN; tol = 1e-14; }
    else if (solvName == "HIP_PCG_FUSED") form = ICX_SOLVE_FUSED;
    else if (solvName == "HIP_PCG_MIXED") form = ICX_SOLVE_MIXED;
    else {
      std::cerr << "Error ! Solver type name is wrong." << std::endl;
      std::exit(1);
    }
    if (!force_.fext_org.empty()) check(model_.ctx(), icx_set_fext(model_.job(), force_.fext_org.data()), "icx_set_fext");
    force_.fext = force_.fext_org;
    value_.val.assign(fem_.numeq, 0.0);
    model_.flushDesign();
    icx_solve_info info{};
    icx_status rc = icx_solve(model_.job(), form, tol, maxit_, value_.val.data(), &info);
    check(model_.ctx(), rc, "icx_solve");
    model_.solved();
    if (node_vals_) node_vals_();
    info_ = info;
    if (log) {
      std::cout << "#iterations:     " << info.iterations << std::endl;
      std::cout << "estimated error: " << info.relres << std::endl;
      if (rc == ICX_ERR_NOCONV) std::cerr << "solving failed" << std::endl;
    }
  }
  void setTolerance(double tol) { tol_ = tol; }
  void setMaxIterations(int it) { maxit_ = it; }
  const icx_solve_info &info() const { return info_; }

 private:
  FEM &fem_;
  StructuredModel &model_;
  Force &force_;
  Value &value_;
  double tol_ = 1.0e-10;   // src/analysis/solver/solvers.hpp:90
  int maxit_ = 10000;      // src/analysis/solver/solvers.hpp:91
  icx_solve_info info_{};
  std::function<void()> node_vals_;
};

// the drivers' sensitivity(): dfds = -(dE/E) u^T Ke u, dgds = v / vscale
inline void sensitivity(StructuredModel &model, Optimize &optim, double vscale = 1.0) {
  check(model.ctx(), icx_sensitivity(model.job(), vscale, optim.dfds.data(), optim.dgds.data()), "icx_sensitivity");
}

// the drivers' output() (opt-macro-oc/opt_macro_oc.cpp:214-238,
// opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:207-235): res<step>.vtk with the
// displacement, density and von Mises stress, written binary directly (the reference
// writes ASCII and converts it with meshio-binary)
inline void output(StructuredModel &model, const Optimize &optim) {
  const std::string path = "res" + std::to_string(optim.optstep) + ".vtk";
  check(model.ctx(), icx_write_vtk(model.job(), path.c_str()), "icx_write_vtk");
}

// the reference driver's sensitivity(fem, element, node, optim, V0) and output(fem,
// element, node, optim) (opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:173-235) on the
// element handles' job: dfds = -(dE/E) u^T Ke u, dgds = v / V0; res<step>.vtk
inline void sensitivity(FEM &, std::vector<LinearElastic> &element, std::vector<Node> &, Optimize &optim,
                        double V0) {
  sensitivity(element.front().model(), optim, V0);
}
inline void output(FEM &, std::vector<LinearElastic> &element, std::vector<Node> &, Optimize &optim) {
  output(element.front().model(), optim);
}
// src/io/export_graph.hpp: addPlot(x, y, file, isFirst) and exportGraph(y, file), in the
// reference's text format
template <typename T1, typename T2>
void addPlot(T1 xx, T2 yy, const std::string &filename, int isFirst) {
  std::ofstream f(filename, isFirst == 1 ? std::ios::out : std::ios::app);
  f << xx << " " << yy << std::endl;
}
template <typename T> void exportGraph(T &yy, const std::string &filename) {
  std::ofstream f(filename, std::ios::out);
  for (int i = 0; i < (int)yy.size(); i++) f << i << " " << yy[i] << std::endl;
}

// ---------------------------------------------------------------------------------
inline std::vector<double> threshold_host(StructuredModel &m, const std::vector<double> &v, double beta, double T,
                                          bool derivative) {
  std::vector<double> out(v.size());
  check(m.ctx(), icx_project(m.job(), beta, T, v.data(), derivative ? nullptr : out.data(),
                             derivative ? out.data() : nullptr), "icx_project");
  return out;
}

template <class E = void, class N = void>
class Filter {
 public:
  Filter(FEM &, StructuredModel &model, double radius) : model_(model) {
    check(model.ctx(), icx_cone_filter_setup(model.job(), radius), "icx_cone_filter_setup");
  }
  // Filter<E, N>(fem, element, node, r) (src/optimization/filter.hpp:20-21)
  Filter(FEM &fem, std::vector<E> &element, std::vector<N> &, double radius)
      : Filter(fem, element.front().model(), radius) {}
  // Filter::sensitivityFilter (filter.hpp:72-100)
  std::vector<double> sensitivityFilter(const std::vector<double> &design_s, const std::vector<double> &dfds) {
    std::vector<double> out(dfds.size());
    check(model_.ctx(), icx_sensitivity_filter(model_.job(), design_s.data(), dfds.data(), out.data()),
          "icx_sensitivity_filter");
    return out;
  }
  std::vector<double> densityFilter(const std::vector<double> &design_s, bool isSens) {
    std::vector<double> out(design_s.size());
    check(model_.ctx(), icx_cone_filter(model_.job(), isSens ? 1 : 0, design_s.data(), out.data()), "icx_cone_filter");
    return out;
  }
  std::vector<double> threshold(const std::vector<double> &v, double beta, double T) {
    return threshold_host(model_, v, beta, T, false);
  }
  std::vector<double> thresholdDerivative(const std::vector<double> &v, double beta, double T) {
    return threshold_host(model_, v, beta, T, true);
  }

 private:
  StructuredModel &model_;
};

template <class E = void, class N = void>
class FilterHelmholtz {
 public:
  FilterHelmholtz(FEM &, StructuredModel &model, double radius) : model_(model) {
    check(model.ctx(), icx_helmholtz_setup(model.job(), radius), "icx_helmholtz_setup");
  }
  // FilterHelmholtz<E, N>(fem, element, node, R) (src/optimization/filter_helm.hpp:23-30)
  FilterHelmholtz(FEM &fem, std::vector<E> &element, std::vector<N> &, double radius)
      : FilterHelmholtz(fem, element.front().model(), radius) {}
  std::vector<double> densityFilter(const std::vector<double> &design_s, const std::string &solvName, int log,
                                    bool sensFlag) {
    if (solvName != "Eigen_CG" && solvName != "HIP_PCG") {
      std::cerr << "Error ! Solver type name is wrong." << std::endl;
      std::exit(1);
    }
    std::vector<double> out(design_s.size());
    int it = 0;
    check(model_.ctx(), icx_helmholtz_filter(model_.job(), sensFlag ? 1 : 0, 1.0e-10, design_s.data(), out.data(), &it),
          "icx_helmholtz_filter");
    if (log) std::cout << "#iterations:     " << it << std::endl;
    return out;
  }
  std::vector<double> threshold(const std::vector<double> &v, double beta, double T) {
    return threshold_host(model_, v, beta, T, false);
  }
  std::vector<double> thresholdDerivative(const std::vector<double> &v, double beta, double T) {
    return threshold_host(model_, v, beta, T, true);
  }

 private:
  StructuredModel &model_;
};

// ---------------------------------------------------------------------------------
class OC {
 public:
  explicit OC(int n_) : n(n_) {}
  // the reference's host bisection with a caller-supplied volume function
  template <class F>
  void UpdateVariables(std::vector<double> &xk, std::vector<double> &dfdx, double g, std::vector<double> &dgdx,
                       std::vector<double> &xmin, std::vector<double> &xmax, F gkp1) {
    double l0 = lambdamin, l1 = lambdamax, lambda = 0.0;
    std::vector<double> xkp1(n);
    int counter = 0;
    while ((l1 - l0) / (l1 + l0) > lambdaeps && maxiter > counter) {
      lambda = 0.5 * (l1 + l0);
      for (int i = 0; i < n; i++) {
        xkp1[i] = std::pow(std::max(0.0, -dfdx[i] / (dgdx[i] * lambda)), iota) * xk[i];
        if (xkp1[i] < std::max(xmin[i], (1.0 - movelimit) * xk[i])) xkp1[i] = std::max(xmin[i], (1.0 - movelimit) * xk[i]);
        else if (xkp1[i] > std::min(xmax[i], (1.0 + movelimit) * xk[i])) xkp1[i] = std::min(xmax[i], (1.0 + movelimit) * xk[i]);
      }
      if (gkp1(xkp1) - g > 0.0) l0 = lambda;
      else l1 = lambda;
      counter++;
    }
    if (counter >= maxiter) std::cerr << "not converge in OCSolver.hpp" << std::endl;
    xk = xkp1;
  }
  // the same bisection with the volume (filter + projection) evaluated on the device
  int UpdateVariablesDevice(StructuredModel &model, std::vector<double> &xk, std::vector<double> &dfdx, double g,
                            std::vector<double> &dgdx, double xmin, double xmax, int filter, double beta,
                            double T) {
    int bis = 0;
    check(model.ctx(), icx_oc_update(model.job(), xk.data(), dfdx.data(), g, dgdx.data(), xmin, xmax, movelimit,
                                     filter, beta, T, &bis), "icx_oc_update");
    return bis;
  }
  void setMoveLimit(double m) { movelimit = m; }

 private:
  int n;
  double iota = 0.5, lambdamin = 0.0, lambdamax = 1.0e4, lambdaeps = 1.0e-5, movelimit = 0.15;
  int maxiter = 100;
};

class MMASolver {
 public:
  // m = 1, a = 0, c = 1e5 (the drivers' use, MMASolver.hpp:67); the optimiser history
  // lives with the model's job on the device
  // MMASolver(n, m, a, c, d) as the reference constructs it (MMASolver.hpp:67): the model
  // is the one the latest mesh generated
  explicit MMASolver(int n, int m = 1, double a = 0.0, double c = 100000.0, double d = 0.0)
      : MMASolver(current_model(), n, m, a, c, d) {}
  MMASolver(StructuredModel &model, int n, int m = 1, double a = 0.0, double c = 100000.0, double d = 0.0)
      : model_(model), n_(n) {
    if (m != 1 || a != 0.0 || c != 100000.0 || d != 0.0)
      StructuredModel::fatal("MMASolver: the GPU path implements m = 1, a = 0, c = 1e5");
    icx_mma_reset(model.job());
  }
  void Update(double *xval, const double *dfdx, const double *gx, const double *dgdx, const double *xmin,
              const double *xmax) {
    for (int i = 1; i < n_; ++i)
      if (xmin[i] != xmin[0] || xmax[i] != xmax[0]) StructuredModel::fatal("MMASolver: xmin/xmax must be uniform");
    int steps = 0;
    check(model_.ctx(), icx_mma_update(model_.job(), xval, dfdx, gx[0], dgdx, xmin[0], xmax[0], &steps),
          "icx_mma_update");
    std::cout << "MMA loop: " << steps << std::endl;
  }
  void Reset() { icx_mma_reset(model_.job()); }
  static StructuredModel &current_model() {
    if (!StructuredModel::current()) StructuredModel::fatal("MMASolver: no model (generate a mesh first)");
    return *StructuredModel::current();
  }
  void SetAsymptotes(double init, double decrease, double increase) {
    check(model_.ctx(), icx_mma_set_asymptotes(model_.job(), init, decrease, increase), "icx_mma_set_asymptotes");
  }

 private:
  StructuredModel &model_;
  int n_;
};

// The stress-constrained driver's functions (constraint/opt-stress-constraint/
// sensitivity.cpp:8-155) over icx_stress_constraint; FEMStressConst's parameters.
namespace stress {
struct StressParams {
  double stressLim = 1.0, pNorm = 8.0, qp = 0.7;
};
inline double calConstraintValue(StructuredModel &model, const StressParams &sp) {
  double g = 0.0;
  check(model.ctx(), icx_stress_constraint(model.job(), sp.stressLim, sp.pNorm, sp.qp, 1e-10, 10000, &g, nullptr,
                                           nullptr), "icx_stress_constraint");
  return g;
}
// optim.dgds = dg/drho_hat (before the threshold derivative); the adjoint solve is the
// Eigen-CG restatement at 1e-10 (Solvers(...).solve("Eigen_CG"), sensitivity.cpp:124-126)
inline void sensitivity(StructuredModel &model, const StressParams &sp, Optimize &optim) {
  double g = 0.0;
  check(model.ctx(), icx_stress_constraint(model.job(), sp.stressLim, sp.pNorm, sp.qp, 1e-10, 10000, &g,
                                           optim.dgds.data(), nullptr), "icx_stress_constraint");
}
}  // namespace stress

// GCMMASolver (src/optimization/opt-solver/GCMMASolver.hpp:61-145) with m = 1, a = 0,
// c = 1000 (the driver's use, opt_macro_gcmma.cpp:69); the solver state (asymptotes,
// x history, raa0 / raa, the approximations' values) lives with the model's job.
class GCMMASolver {
 public:
  GCMMASolver(StructuredModel &model, int n, int m = 1, double a = 0.0, double c = 1000.0, double d = 0.0)
      : model_(model), n_(n) {
    if (m != 1 || a != 0.0 || c != 1000.0 || d != 0.0)
      StructuredModel::fatal("GCMMASolver: the GPU path implements m = 1, a = 0, c = 1000");
  }
  void OuterUpdate(double *xmma, const double *xval, double f0x, const double *df0dx, const double *fx,
                   const double *dfdx, const double *xmin, const double *xmax) {
    uniform(xmin, xmax);
    int steps = 0;
    check(model_.ctx(), icx_gcmma_outer_update(model_.job(), xmma, xval, f0x, df0dx, fx[0], dfdx, xmin[0], xmax[0],
                                               &steps), "icx_gcmma_outer_update");
  }
  void InnerUpdate(double *xmma, double f0xnew, const double *fxnew, const double *xval, double f0x,
                   const double *df0dx, const double *fx, const double *dfdx, const double *xmin,
                   const double *xmax) {
    uniform(xmin, xmax);
    int steps = 0;
    check(model_.ctx(), icx_gcmma_inner_update(model_.job(), xmma, f0xnew, fxnew[0], xval, f0x, df0dx, fx[0], dfdx,
                                               xmin[0], xmax[0], &steps), "icx_gcmma_inner_update");
  }
  bool ConCheck(double f0xnew, const double *fxnew) const {
    int ok = 0;
    check(model_.ctx(), icx_gcmma_concheck(model_.job(), f0xnew, fxnew[0], &ok), "icx_gcmma_concheck");
    return ok != 0;
  }
  void Reset() { icx_gcmma_reset(model_.job()); }

 private:
  void uniform(const double *xmin, const double *xmax) const {
    for (int i = 1; i < n_; ++i)
      if (xmin[i] != xmin[0] || xmax[i] != xmax[0]) StructuredModel::fatal("GCMMASolver: xmin/xmax must be uniform");
  }
  StructuredModel &model_;
  int n_;
};

}  // namespace icarat
