// Driver A on the GPU path: cone density filter + projection + linear solve +
// sensitivity + OC, step for step as opt-macro-oc/opt_macro_oc.cpp:13-180, on the
// unstructured mesh of an .mdpa file (as the reference: MeshMDPA) or on the structured
// [mesh] box of the TOML file when no .mdpa is given.
//
// usage: opt_macro_oc <job.toml> [mesh.mdpa] [solver=Eigen_CG] [max_steps=1000] [radius=3.0]
// The OC volume function (filter + threshold of the candidate) runs on the device.
// Writes function.csv and res<step>.vtk (binary) in the working directory.
#include <chrono>
#include <fstream>
#include <memory>

#include "icarat_hip.hpp"

using namespace icarat;

int main(int argc, char *argv[]) {
  if (argc < 2) {
    std::cout << "usage: " << argv[0] << " <.toml file> [.mdpa file] [solver] [max_steps] [radius]\n";
    return 1;
  }
  std::string mdpa;
  int a = 2;
  if (argc > 2) {
    const std::string s2 = argv[2];
    if (s2.size() > 5 && s2.compare(s2.size() - 5, 5, ".mdpa") == 0) {
      mdpa = s2;
      a = 3;
    }
  }
  const std::string solver = argc > a ? argv[a] : "Eigen_CG";
  const int max_steps = argc > a + 1 ? std::atoi(argv[a + 1]) : 1000;
  const double filter_radious = argc > a + 2 ? std::atof(argv[a + 2]) : 3.0;  // opt_macro_oc.cpp:74
  auto t_start = std::chrono::system_clock::now();
  int convergence = 1;
  double V0 = 0.0, T = 0.5;
  const auto config = toml_lite::parse(argv[1]);

  FEM fem;
  std::unique_ptr<StructuredModel> mp(mdpa.empty() ? new StructuredModel(config, fem)
                                                   : new StructuredModel(config, fem, mdpa));
  StructuredModel &model = *mp;
  model.input_dirich(config);
  fem.numeq = model.makeDOF();
  Force force(fem.numeq);
  Value dis(fem.numeq);
  model.input_neumann(config, force);

  Optimize optim(fem.nelx, 1, 1.0e-4, 1.0);
  OC oc(fem.nelx);
  Filter filter(fem, model, filter_radious);
  optim.optstep = 0;
  for (int i = 0; i < fem.nelx; i++) optim.design_s[i] = model.design_s0();
  LinearAnalysis compt(fem, model, force, dis);
  std::ofstream csv("function.csv");
  csv.precision(17);

  while (convergence == 1) {
    optim.optstep++;
    std::cout << "step no. " << optim.optstep << std::endl;
    auto design_filtered = filter.densityFilter(optim.design_s, false);
    double beta = std::min(std::pow(2, (int)(optim.optstep / 40)), 16.0);
    auto design_threshold = filter.threshold(design_filtered, beta, T);
    auto threDeriv = filter.thresholdDerivative(design_filtered, beta, T);
    model.setDesignS(design_threshold);
    compt.solve(solver, 1);
    double obj = 0.0;
    for (int i = 0; i < fem.numeq; i++) obj += dis.val[i] * force.fext[i];
    optim.object_f = obj;
    std::cout << "function value = " << optim.object_f << std::endl;
    if (optim.optstep == 1) {
      V0 = 0.0;
      for (int i = 0; i < fem.nelx; i++) V0 += design_threshold[i] * model.volume()[i];
      optim.const_h[0] = V0;
    } else {
      optim.const_h[0] = 0.0;
      for (int i = 0; i < fem.nelx; i++) optim.const_h[0] += design_threshold[i] * model.volume()[i];
    }
    csv << optim.optstep << "," << optim.object_f << "," << optim.const_h[0] << "\n";
    output(model, optim);
    convergence = optim.isConvergence(1.0e-7, 50);
    sensitivity(model, optim);
    for (int i = 0; i < fem.nelx; i++) {
      optim.dfds[i] *= threDeriv[i];
      optim.dgds[i] *= threDeriv[i];
    }
    optim.dfds = filter.densityFilter(optim.dfds, true);
    optim.dgds = filter.densityFilter(optim.dgds, true);
    oc.UpdateVariablesDevice(model, optim.design_s, optim.dfds, V0, optim.dgds, optim.xmin[0], optim.xmax[0],
                             ICX_FILTER_CONE, beta, T);
    if (optim.optstep == max_steps) {
      std::cerr << "not convergence in this optimization" << std::endl;
      break;
    }
  }
  auto t_sec = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::system_clock::now() - t_start).count();
  std::cout << "Computation time: " << t_sec << " sec \n";
  return 0;
}
