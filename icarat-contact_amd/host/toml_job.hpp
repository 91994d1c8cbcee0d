// A box job from the keys of a reference job file, through the C ABI.  The one front end
// of the TOML job spec: libicx's icx_job_create_toml (the serving loop's wire format) and
// the C++ drivers' StructuredModel both build their jobs here.
//   [mesh]               dimension, type ("quad4" / "hexa8"), division, length
//                        (MeshIcarat::setParameter, src/io/mesh_icarat.hpp:69-128)
//   [boundary_condition] dirichlet = [[flag6], [value6], [[x0, x1], [y0, y1], [z0, z1]]], ...
//                        neumann   = [[traction3], [[x0, x1], [y0, y1], [z0, z1]]], ...
//                        (BCIcarat::input_dirich / input_neumann,
//                        src/io/boundary_condition_icarat.hpp:68-396; tractions integrated
//                        as BCBase::generateLoadMain, boundary_condition_base.cpp:143-180)
//   [material]           type ("Elastic" / "Heat"), state, young, poisson, penalty,
//                        conductivity, designs_0 (Elastic::setParameter,
//                        src/problem/elasticity.cpp:19-32; Heat::setParameter, heat.cpp:19-60)
// Order of the reference's set-up: mesh and material, Dirichlet ranges, countDOF, loads.
// Errors come back as a status and a message (no exit): a malformed spec is refused.
#pragma once

#include <exception>
#include <string>
#include <vector>

#include "../../include/icx.h"
#include "toml_lite.hpp"

namespace icarat {

struct TomlJob {
  icx_job *job = nullptr;
  int ndim = 0;
  int numeq = 0;
  bool heat = false;
  double design0 = 0.0;     // [material] designs_0
  icx_problem prob{};
};

inline icx_status box_job_from_toml(icx_ctx *ctx, const toml_lite::Table &config, TomlJob &out, std::string &why) {
  out = TomlJob{};
  icx_job *job = nullptr;
  auto bail = [&](icx_status rc, const std::string &msg) {
    why = msg;
    if (job) icx_job_destroy(job);
    return rc;
  };
  try {
    const std::string type = config.at("mesh", "type").as_string();
    if (type != "quad4" && type != "hexa8")
      return bail(ICX_ERR_ARG, "element type " + type + " is not supported by the GPU path (quad4, hexa8)");
    icx_problem p{};
    p.eltype = type == "quad4" ? ICX_QUAD4 : ICX_HEXA8;
    const auto &div = config.at("mesh", "division").as_array();
    const auto &len = config.at("mesh", "length").as_array();
    if (div.size() < 3 || len.size() < 3) return bail(ICX_ERR_ARG, "[mesh] division and length need 3 entries");
    for (int d = 0; d < 3; ++d) {
      p.div[d] = div[d].as_int();
      p.length[d] = len[d].as_double();
    }
    const int ndim = config.at("mesh", "dimension").as_int();
    if (ndim != (p.eltype == ICX_QUAD4 ? 2 : 3)) return bail(ICX_ERR_ARG, "[mesh] dimension does not match the type");
    const bool heat = config.has("material", "type") && config.at("material", "type").as_string() == "Heat";
    double cond[2][6] = {};
    if (heat) {
      p.state = ICX_HEAT;
      const auto &rows = config.at("material", "conductivity").as_array();
      if (rows.size() < 2) return bail(ICX_ERR_ARG, "[material] conductivity needs two rows");
      for (int m = 0; m < 2; ++m) {
        const auto &row = rows[m].as_array();
        if (row.size() < 6) return bail(ICX_ERR_ARG, "[material] conductivity rows need 6 entries");
        for (int k = 0; k < 6; ++k) cond[m][k] = row[k].as_double();
      }
    } else {
      std::string state = config.at("material", "state").as_string();
      if (ndim == 3) state = "3D";
      if (state != "plane_stress" && state != "plane_strain" && state != "3D")
        return bail(ICX_ERR_ARG, "[material] state " + state + " is not plane_stress, plane_strain or 3D");
      p.state = state == "plane_stress" ? ICX_PLANE_STRESS : (state == "plane_strain" ? ICX_PLANE_STRAIN : ICX_SOLID3D);
      const auto &young = config.at("material", "young").as_array();
      if (young.size() < 2) return bail(ICX_ERR_ARG, "[material] young needs two entries");
      p.young1 = young[0].as_double();
      p.young2 = young[1].as_double();
      p.poisson = config.at("material", "poisson").as_double();
    }
    p.penalty = config.at("material", "penalty").as_double();
    out.design0 = config.at("material", "designs_0").as_double();
    icx_status rc = icx_job_create(ctx, &p, &job);
    if (rc) return bail(rc, std::string("icx_job_create: ") + icx_last_error(ctx));
    if (heat && (rc = icx_set_conductivity(job, cond[0], cond[1])))
      return bail(rc, std::string("icx_set_conductivity: ") + icx_last_error(ctx));
    for (const auto &entry : config.at("boundary_condition", "dirichlet").as_array()) {
      const auto &e = entry.as_array();
      if (e.size() < 3 || e[0].as_array().size() < 6 || e[1].as_array().size() < 6 || e[2].as_array().size() < 3)
        return bail(ICX_ERR_ARG, "dirichlet entries are [[flag x 6], [value x 6], [[x0, x1], [y0, y1], [z0, z1]]]");
      int flag[6];
      double val[6], range[6];
      for (int k = 0; k < 6; ++k) {
        flag[k] = e[0].as_array()[k].as_int();
        val[k] = e[1].as_array()[k].as_double();
      }
      for (int d = 0; d < 3; ++d) {
        const auto &r = e[2].as_array()[d].as_array();
        if (r.size() < 2) return bail(ICX_ERR_ARG, "dirichlet ranges are [lo, hi] pairs");
        for (int k = 0; k < 2; ++k) range[2 * d + k] = r[k].as_double();
      }
      if ((rc = icx_add_dirichlet(job, flag, val, range)))
        return bail(rc, std::string("icx_add_dirichlet: ") + icx_last_error(ctx));
    }
    int numeq = 0;
    if ((rc = icx_make_dof(job, &numeq))) return bail(rc, std::string("icx_make_dof: ") + icx_last_error(ctx));
    if (config.has("boundary_condition", "neumann"))
      for (const auto &entry : config.at("boundary_condition", "neumann").as_array()) {
        const auto &e = entry.as_array();
        if (e.size() < 2 || e[0].as_array().size() < 3 || e[1].as_array().size() < 3)
          return bail(ICX_ERR_ARG, "neumann entries are [[t_x, t_y, t_z], [[x0, x1], [y0, y1], [z0, z1]]]");
        double val[3], range[6];
        for (int k = 0; k < 3; ++k) val[k] = e[0].as_array()[k].as_double();
        for (int d = 0; d < 3; ++d) {
          const auto &r = e[1].as_array()[d].as_array();
          if (r.size() < 2) return bail(ICX_ERR_ARG, "neumann ranges are [lo, hi] pairs");
          for (int k = 0; k < 2; ++k) range[2 * d + k] = r[k].as_double();
        }
        if ((rc = icx_add_neumann(job, val, range)))
          return bail(rc, std::string("icx_add_neumann: ") + icx_last_error(ctx));
      }
    out.job = job;
    out.ndim = ndim;
    out.numeq = numeq;
    out.heat = heat;
    out.prob = p;
    return ICX_OK;
  } catch (const std::exception &e) {
    return bail(ICX_ERR_ARG, e.what());
  }
}

}  // namespace icarat
