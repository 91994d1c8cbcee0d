// The stress-constrained driver on the GPU path: minimum volume under a p-norm of the
// qp-relaxed von Mises stresses, cone filter (radius 2 h) + projection + MMA, step for
// step as constraint/opt-stress-constraint/opt_stress_constraint.cpp:11-166 (2D jobs).
//
// usage: opt_stress_constraint <job.toml> [solver=Eigen_CG] [max_steps=1000]
// [this_problem] stress_limit, p_norm, qp as in the reference's job file.
// Writes function.csv (step, volume, stress constraint) and res<step>.vtk (binary).
#include <chrono>
#include <fstream>

#include "icarat_hip.hpp"

using namespace icarat;

int main(int argc, char *argv[]) {
  if (argc < 2) {
    std::cout << "usage: " << argv[0] << " <.toml file> [solver] [max_steps]\n";
    return 1;
  }
  const std::string solver = argc > 2 ? argv[2] : "Eigen_CG";
  const int max_steps = argc > 3 ? std::atoi(argv[3]) : 1000;
  auto t_start = std::chrono::system_clock::now();
  int convergence = 1;
  const int numConst = 1;
  double V0 = 0.0;
  const auto config = toml_lite::parse(argv[1]);

  FEM fem;
  stress::StressParams sp;
  sp.stressLim = config.at("this_problem", "stress_limit").as_double();
  sp.pNorm = config.at("this_problem", "p_norm").as_double();
  sp.qp = config.at("this_problem", "qp").as_double();
  StructuredModel model(config, fem);
  if (fem.ndim != 2) StructuredModel::fatal("stressconstrainted 3D is not applied.");
  model.input_dirich(config);
  fem.numeq = model.makeDOF();
  Force force(fem.numeq);
  Value dis(fem.numeq);
  model.input_neumann(config, force);

  Optimize optim(fem.nelx, numConst, 1.0e-4, 1.0);
  MMASolver mma(model, fem.nelx, numConst);
  mma.SetAsymptotes(0.1, 0.7, 1.2);
  const double filterR = 2.0 * model.problem().length[0] / model.problem().div[0];
  Filter filter(fem, model, filterR);
  optim.optstep = 0;
  for (int i = 0; i < fem.nelx; i++) optim.design_s[i] = model.design_s0();
  LinearAnalysis compt(fem, model, force, dis);
  std::ofstream csv("function.csv");
  csv.precision(17);

  while (convergence == 1) {
    optim.optstep++;
    std::cout << "step no. " << optim.optstep << std::endl;
    auto design_filtered = filter.densityFilter(optim.design_s, false);
    const double beta = std::min(std::pow(2.0, (int)(optim.optstep / 40)), 8.0), T = 0.5;
    auto design_threshold = filter.threshold(design_filtered, beta, T);
    auto threDifferent = filter.thresholdDerivative(design_filtered, beta, T);
    model.setDesignS(design_threshold);
    compt.solve(solver, 0);
    optim.object_f = 0.0;
    for (int i = 0; i < fem.nelx; i++) optim.object_f += design_threshold[i] * model.volume()[i];
    if (optim.optstep == 1) V0 = optim.object_f;
    std::cout << "function value = " << optim.object_f << std::endl;
    optim.const_h[0] = stress::calConstraintValue(model, sp);
    std::cout << "Constraint value:" << optim.const_h[0] << std::endl;
    csv << optim.optstep << "," << optim.object_f << "," << optim.const_h[0] << "\n";
    output(model, optim);
    convergence = optim.isConvergence(1.0e-6, 50);
    stress::sensitivity(model, sp, optim);  // dg/drho_hat, one adjoint solve
    for (int i = 0; i < fem.nelx; i++) {
      optim.dfds[i] = model.volume()[i] * threDifferent[i];
      optim.dgds[i] *= threDifferent[i];
    }
    optim.dfds = filter.densityFilter(optim.dfds, true);
    optim.dgds = filter.densityFilter(optim.dgds, true);
    mma.Update(optim.design_s.data(), optim.dfds.data(), optim.const_h.data(), optim.dgds.data(),
               optim.xmin.data(), optim.xmax.data());
    if (optim.optstep == max_steps) {
      std::cerr << "not convergence in this optimization" << std::endl;
      break;
    }
  }
  (void)V0;
  auto t_sec = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::system_clock::now() - t_start).count();
  std::cout << "Computation time: " << t_sec << " sec \n";
  return 0;
}
