// Driver B on the GPU path: Helmholtz filter + projection + linear solve +
// sensitivity + MMA, step for step as opt-mma-HelmholtzFilter/opt_mma_Helmholtz.cpp:12-171.
//
// usage: opt_mma_helmholtz <job.toml> [solver=Eigen_LDLT] [max_steps=1000]
// Writes function.csv ("step,objective,constraint") and res<step>.vtk (binary) in the
// working directory.
#include <chrono>
#include <fstream>

#include "icarat_hip.hpp"

using namespace icarat;

int main(int argc, char *argv[]) {
  if (argc < 2) {
    std::cout << "usage: " << argv[0] << " <.toml file> [solver] [max_steps]\n";
    return 1;
  }
  const std::string solver = argc > 2 ? argv[2] : "Eigen_LDLT";
  const int max_steps = argc > 3 ? std::atoi(argv[3]) : 1000;
  auto t_start = std::chrono::system_clock::now();
  int convergence = 1;
  double V0 = 0.0;
  const auto config = toml_lite::parse(argv[1]);

  FEM fem;
  StructuredModel model(config, fem);
  model.input_dirich(config);
  fem.numeq = model.makeDOF();
  Force force(fem.numeq);
  Value dis(fem.numeq);
  model.input_neumann(config, force);

  Optimize optim(fem.nelx, 1, 1.0e-4, 1.0);
  MMASolver mma(model, fem.nelx, 1);
  // filter radius 2 * |x0 - x1| (opt_mma_Helmholtz.cpp:70)
  const double filterR = 2.0 * model.problem().length[0] / model.problem().div[0];
  FilterHelmholtz filter(fem, model, filterR);
  optim.optstep = 0;
  for (int i = 0; i < fem.nelx; i++) optim.design_s[i] = model.design_s0();
  LinearAnalysis compt(fem, model, force, dis);
  std::ofstream csv("function.csv");
  csv.precision(17);

  while (convergence == 1) {
    optim.optstep++;
    std::cout << "step no. " << optim.optstep << std::endl;
    auto fil_s = filter.densityFilter(optim.design_s, "Eigen_CG", 0, false);
    double beta = std::min(std::pow(2, (int)(optim.optstep / 40)), 16.0);
    auto thre_s = filter.threshold(fil_s, beta, 0.5);
    model.setDesignS(thre_s);
    compt.solve(solver, 1);
    if (optim.optstep == 1) {
      V0 = 0.0;
      for (int i = 0; i < fem.nelx; i++) V0 += thre_s[i] * model.volume()[i];
    } else {
      optim.const_h[0] = 0.0;
      for (int i = 0; i < fem.nelx; i++) optim.const_h[0] += thre_s[i] * model.volume()[i] / V0;
      optim.const_h[0] -= 1;
      std::cout << "Constraint value:" << optim.const_h[0] << std::endl;
    }
    double obj = 0.0;
    for (int i = 0; i < fem.numeq; i++) obj += dis.val[i] * force.fext[i];
    optim.object_f = obj;
    std::cout << "function value = " << optim.object_f << std::endl;
    csv << optim.optstep << "," << optim.object_f << "," << optim.const_h[0] << "\n";
    output(model, optim);
    convergence = optim.isConvergence(1.0e-7, 50);
    sensitivity(model, optim, V0);
    auto thrediff = filter.thresholdDerivative(fil_s, beta, 0.5);
    for (int i = 0; i < fem.nelx; i++) {
      optim.dfds[i] *= thrediff[i];
      optim.dgds[i] *= thrediff[i];
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
