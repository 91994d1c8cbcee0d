// The GCMMA driver on the GPU path: cone density filter (radius 2 h) + projection +
// linear solve + sensitivity + GCMMA with its inner loop, step for step as
// opt_macro_gcmma.cpp:11-231 on the structured [mesh] box of the TOML file.
//
// usage: opt_macro_gcmma <job.toml> [solver=Eigen_CG] [max_steps=1000]
// Writes function.csv (step, objective, const_h, inner loops) and res<step>.vtk (binary)
// in the working directory.
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
  StructuredModel model(config, fem);
  model.input_dirich(config);
  fem.numeq = model.makeDOF();
  Force force(fem.numeq);
  Value dis(fem.numeq);
  model.input_neumann(config, force);

  Optimize optim(fem.nelx, numConst, 1.0e-4, 1.0);
  GCMMASolver gcmma(model, fem.nelx, numConst);
  const double filterR = 2.0 * model.problem().length[0] / model.problem().div[0];  // 2-cell length
  Filter filter(fem, model, filterR);
  for (int i = 0; i < fem.nelx; i++) optim.design_s[i] = model.design_s0();
  LinearAnalysis compt(fem, model, force, dis);
  optim.optstep = 0;
  std::ofstream csv("function.csv");
  csv.precision(17);

  auto volume = [&](const std::vector<double> &rho) {
    double v = 0.0;
    for (int i = 0; i < fem.nelx; i++) v += rho[i] * model.volume()[i];
    return v;
  };
  auto compliance = [&]() {
    double c = 0.0;
    for (int i = 0; i < fem.numeq; i++) c += dis.val[i] * force.fext[i];
    return c;
  };

  while (convergence == 1) {
    optim.optstep++;
    std::cout << "step no. " << optim.optstep << std::endl;
    const int beta_tmp = std::max((optim.optstep - 100) / 40.0, 1.0);
    const double beta = std::min((double)beta_tmp, 8.0), T = 0.5;
    auto design_filtered = filter.densityFilter(optim.design_s, false);
    auto design_threshold = filter.threshold(design_filtered, beta, T);
    auto threDifferent = filter.thresholdDerivative(design_filtered, beta, T);
    model.setDesignS(design_threshold);
    compt.solve(solver, 1);
    if (optim.optstep == 1) {
      V0 = volume(design_threshold);
      if (V0 == 0.0) std::cerr << "V0 = 0. in optimization" << std::endl;
    }
    optim.object_f = compliance();
    optim.const_h[0] = volume(design_threshold) - V0;
    std::cout << "function value : " << optim.object_f << "\nconstraint value : " << optim.const_h[0] << std::endl;
    output(model, optim);
    convergence = optim.isConvergence(1.0e-7, 50);
    // dfds = -threDiff E'/E u_e.Ke.u_e, dgds = threDiff v_e (opt_macro_gcmma.cpp:233-259)
    sensitivity(model, optim);
    for (int i = 0; i < fem.nelx; i++) {
      optim.dfds[i] *= threDifferent[i];
      optim.dgds[i] *= threDifferent[i];
    }
    optim.dfds = filter.densityFilter(optim.dfds, true);
    optim.dgds = filter.densityFilter(optim.dgds, true);

    std::vector<double> design_s_new(fem.nelx), const_h_new(numConst);
    double object_f_new;
    auto trial = [&]() {
      design_filtered = filter.densityFilter(design_s_new, false);
      design_threshold = filter.threshold(design_filtered, beta, T);
      model.setDesignS(design_threshold);
      compt.solve(solver, 0);
      object_f_new = compliance();
      const_h_new[0] = volume(design_threshold) - V0;
    };
    gcmma.OuterUpdate(design_s_new.data(), optim.design_s.data(), optim.object_f, optim.dfds.data(),
                      optim.const_h.data(), optim.dgds.data(), optim.xmin.data(), optim.xmax.data());
    trial();
    bool conserv = gcmma.ConCheck(object_f_new, const_h_new.data());
    int inneriter = 0;
    for (; !conserv && inneriter < 15; ++inneriter) {
      gcmma.InnerUpdate(design_s_new.data(), object_f_new, const_h_new.data(), optim.design_s.data(),
                        optim.object_f, optim.dfds.data(), optim.const_h.data(), optim.dgds.data(),
                        optim.xmin.data(), optim.xmax.data());
      trial();
      conserv = gcmma.ConCheck(object_f_new, const_h_new.data());
    }
    std::cout << "GCMMA inner loop: " << inneriter << std::endl;
    csv << optim.optstep << "," << optim.object_f << "," << optim.const_h[0] << "," << inneriter << "\n";
    optim.design_s = design_s_new;
    if (optim.optstep == max_steps) {
      std::cerr << "not convergence in this optimization" << std::endl;
      break;
    }
  }
  auto t_sec = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::system_clock::now() - t_start).count();
  std::cout << "Computation time: " << t_sec << " sec \n";
  return 0;
}
