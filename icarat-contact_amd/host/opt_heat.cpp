// Heat-conduction TO on the GPU path: Helmholtz filter + projection + linear heat solve
// + sensitivity + MMA, step for step as opt-heat/opt_heat.cpp:11-155.
//
// usage: opt_heat <job.toml> [max_steps=1000]
// Writes function.csv ("step,objective,constraint") and res<step>.vtk (binary:
// temperature, threshold_s, design_s) in the working directory.
#include <chrono>
#include <fstream>

#include "icarat_hip.hpp"

using namespace icarat;

int main(int argc, char *argv[]) {
  if (argc < 2) {
    std::cout << "usage: " << argv[0] << " <.toml file> [max_steps]\n";
    return 1;
  }
  const int max_steps = argc > 2 ? std::atoi(argv[2]) : 1000;  // opt_heat.cpp:143
  auto t_start = std::chrono::system_clock::now();
  int convergence = 1;
  double V0 = 0.0;
  const auto config = toml_lite::parse(argv[1]);
  const double filterR = config.at("this_problem", "filterR").as_double();

  FEM fem;
  StructuredModel model(config, fem);
  if (!model.heat()) StructuredModel::fatal("opt_heat needs [material] type = \"Heat\"");
  model.input_dirich(config);
  fem.numeq = model.makeDOF();
  Force force(fem.numeq);
  Value dis(fem.numeq);
  model.input_neumann(config, force);

  Optimize optim(fem.nelx, 1, 1.0e-4, 1.0);
  MMASolver mma(model, fem.nelx, 1);
  FilterHelmholtz filter(fem, model, filterR);
  optim.optstep = 0;
  for (int i = 0; i < fem.nelx; i++) optim.design_s[i] = model.design_s0();
  LinearAnalysis compt(fem, model, force, dis);
  std::ofstream csv("function.csv");
  csv.precision(17);
  std::vector<double> KT(fem.numeq);

  while (convergence == 1) {
    optim.optstep++;
    std::cout << "============== optimization step ==============" << std::endl;
    std::cout << "step no. " << optim.optstep << std::endl;
    auto design_filtered = filter.densityFilter(optim.design_s, "Eigen_CG", 0, false);
    double beta = std::min(std::pow(2, (int)(optim.optstep / 40)), 16.0);
    auto design_threshold = filter.threshold(design_filtered, beta, 0.5);
    model.setDesignS(design_threshold);
    compt.solve("Eigen_CG", 1);
    if (optim.optstep == 1) {
      V0 = 0.0;
      for (int i = 0; i < fem.nelx; i++) V0 += design_threshold[i] * model.volume()[i];
      if (V0 == 0.0) std::cerr << "V0 = 0. in optimization" << std::endl;
    }
    // objective T^T K T over the free DOFs (opt_heat.cpp:108)
    check(model.ctx(), icx_matvec(model.job(), 0, dis.val.data(), KT.data()), "icx_matvec");
    optim.object_f = 0.0;
    for (int i = 0; i < fem.numeq; i++) optim.object_f += dis.val[i] * KT[i];
    std::cout << "function value = " << optim.object_f << std::endl;
    optim.const_h[0] = 0.0;
    for (int i = 0; i < fem.nelx; i++) optim.const_h[0] += design_threshold[i] * model.volume()[i];
    optim.const_h[0] -= V0;
    std::cout << "Constraint value:" << optim.const_h[0] << std::endl;
    csv << optim.optstep << "," << optim.object_f << "," << optim.const_h[0] << "\n";
    output(model, optim);
    convergence = optim.isConvergence(1.0e-7, 50);
    // sensitivity (opt_heat.cpp:158-194): -thr' T^T dK/ds T and thr' v
    auto threDifferent = filter.thresholdDerivative(design_filtered, beta, 0.5);
    sensitivity(model, optim);
    for (int i = 0; i < fem.nelx; i++) {
      optim.dfds[i] *= threDifferent[i];
      optim.dgds[i] *= threDifferent[i];
    }
    optim.dfds = filter.densityFilter(optim.dfds, "Eigen_CG", 0, true);
    optim.dgds = filter.densityFilter(optim.dgds, "Eigen_CG", 0, true);
    mma.Update(optim.design_s.data(), optim.dfds.data(), optim.const_h.data(), optim.dgds.data(),
               optim.xmin.data(), optim.xmax.data());
    if (optim.optstep == max_steps) {
      std::cerr << "not convergence in this optimization" << std::endl;
      break;
    }
  }
  auto t_sec = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::system_clock::now() - t_start).count();
  std::cout << "Computation time: " << t_sec << " sec " << std::endl;
  return 0;
}
