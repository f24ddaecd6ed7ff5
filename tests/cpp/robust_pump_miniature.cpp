// The optimisation loop of applications/RobustPump2d/robustPump2d.cc (one realisation, a small
// channel) written against the reference's class names through the lbm_topopt host mirror:
// solver, boundary descriptors, design-box mapping, double projection filter, forward + adjoint
// pass with dynamic checkpointing, filtered sensitivities, fluid-volume constraint and MMA.
//
// Prints one line per optimisation iteration and, at the end, the check lines the GPU test
// reads: the per-step objective (collideAndStream + computeMacroFunctional) against the fused
// forward pass, and whether the pressure drop fell.
#include <cmath>
#include <cstdio>
#include <vector>

#include "lbm_topopt.hpp"

using namespace lbm_topopt;

#define CHK(call)                                                             \
  do {                                                                        \
    PetscErrorCode e_ = (call);                                               \
    if (e_) {                                                                 \
      std::fprintf(stderr, "error %d at %s:%d: %s\n", e_, __FILE__, __LINE__, \
                   lbmx_last_error());                                        \
      return 1;                                                               \
    }                                                                         \
  } while (0)

int main() {
  const PetscInt nx = 96, ny = 40, numSteps = 200, channelLength = 16;
  using Operator = BrinkmanCollision<IncompressibleMRT2d<D2Q9>>;
  NewLBSolverInfo2d info;
  info.nx = nx;
  info.ny = ny;
  IsothermalLBSolver2d<Operator> solver;
  CHK(solver.make(info, Operator(IncompressibleMRT2d<D2Q9>(1.25), BPInterpolation(0.1), 0.1)));

  // walls, parabolic Zou-He inlet and pressure outlet, as the channel of the robust pump
  const double mid = (ny - 1) / 2.0;
  CHK(solver.addBoundaryCondition(Box2d(0, nx - 1, 0, 0), OnGridBouncebackDescriptor()));
  CHK(solver.addBoundaryCondition(Box2d(0, nx - 1, ny - 1, ny - 1), OnGridBouncebackDescriptor()));
  CHK(solver.addBoundaryCondition(Box2d(0, 0, 1, ny - 2), ZouHeParabolicVelocityDescriptor(0.05, mid, mid)));
  CHK(solver.addBoundaryCondition(Box2d(nx - 1, nx - 1, 1, ny - 2), ZouHePressureDescriptor(1.0)));

  AdjointLBSolver adjSolver;
  CHK(solver.makeSourceAdjoint(&adjSolver));
  const Box2d designBox(channelLength, nx - 1 - channelLength, 1, ny - 2);
  FullToDesignMapping2d map(designBox, solver);
  DoubleProjectionFilter2d filter(3, 2, designBox, solver);
  CHK(filter.status);
  filter.setFirstBeta(2.);
  filter.setSecondBeta(1.);
  filter.setFirstEta(0.5);
  filter.setSecondEta(0.5);

  std::vector<double> fullDomain, physicalDomain, designDomain, designSens, designVcSens;
  CHK(solver.createDesignVec(&fullDomain));
  CHK(map.createDesignDomainVec(&designDomain));
  for (auto& v : designDomain) v = 0.5;
  const PetscInt numDesign = (PetscInt)designDomain.size();

  const auto objective = PressureDrop2d(Box2d(0, 0, 1, ny - 2), Box2d(nx - 1, nx - 1, 1, ny - 2));
  const PetscScalar timeAveraging = 1.0 / numSteps;
  UnsteadyLBTopOpt topOpt(solver, adjSolver, objective, numSteps, timeAveraging, &filter);
  CHK(topOpt.allocateMemoryForForwardSolution(16, true));  // LBStaticCheckpointing, as robustPump2d.cc:166

  MMA mma(numDesign, 1, designDomain);
  CHK(mma.status);
  std::vector<double> sens, vcSens((size_t)nx * ny, 0.0);
  double J0 = 0.0, Jlast = 0.0, volLast = 0.0;
  bool monotone = true;
  for (PetscInt iter = 0; iter < 10; ++iter) {
    CHK(map.mapDesignToFull(designDomain, fullDomain));
    CHK(filter.filterDesign(fullDomain, physicalDomain));
    CHK(solver.setFieldFromDesignVec(physicalDomain));
    CHK(topOpt.forwardAndAdjoint());
    const double J = topOpt.getObjective();
    CHK(topOpt.getFilteredSensitivities(&sens));
    // fluid fraction of the design box <= 0.5 (VolumeConstraint(0.5, ConstrainFluid, ...))
    double vol = 0.0;
    std::vector<double> physDesign;
    CHK(map.mapFullToDesign(physicalDomain, physDesign));
    for (double v : physDesign) vol += v;
    vol = vol / numDesign - 0.5;
    std::fill(vcSens.begin(), vcSens.end(), 0.0);
    std::vector<double> ones(physDesign.size(), 1.0 / numDesign);
    CHK(map.mapDesignToFull(ones, vcSens));
    CHK(filter.filterSensitivities(vcSens, nullptr, 0));
    CHK(map.mapFullToDesign(sens, designSens));
    CHK(map.mapFullToDesign(vcSens, designVcSens));
    if (iter == 0) J0 = J;
    for (auto& v : designSens) v /= std::fabs(J0);
    if (iter > 0 && J > Jlast) monotone = false;
    Jlast = J;
    volLast = vol;
    PetscScalar con[1] = {vol};
    CHK(mma.Update(designDomain, designSens, con, &designVcSens, 0.2));
    PetscScalar ch = 0.0;
    CHK(mma.DesignChange(designDomain, ch));
    std::printf("iteration %d J %.12e volume %.6f change %.6f\n", iter, J, vol + 0.5, ch);
  }

  // final design written and read back as the application does for restarts, fields as .vts
  std::vector<double> reread(designDomain.size(), 0.0);
  CHK(writeVectorToBinary("", "robust_pump_miniature_design.bin", designDomain));
  CHK(loadVectorFromBinary("", "robust_pump_miniature_design.bin", reread));
  const bool restart_ok = reread == designDomain;

  // per-step objective (collideAndStream + computeMacroFunctional) against the fused pass
  CHK(topOpt.computeObjective());
  const double Jfused = topOpt.getObjective();
  CHK(solver.initializeAtEquilibrium());
  double Jsteps = 0.0;
  for (PetscInt t = 0; t < numSteps; ++t) {
    CHK(solver.collideAndStream());
    double v = 0.0;
    CHK(solver.computeMacroFunctional(objective, timeAveraging, &v));
    Jsteps += v;
  }
  const double rel = std::fabs(Jsteps - Jfused) / std::fabs(Jfused);
  CHK(solver.outputFields("", "robust_pump_miniature_fields.vts"));  // the state after the last step
  std::printf("per-step objective %.15e fused %.15e rel %.3e\n", Jsteps, Jfused, rel);
  std::printf("pressure drop %.6e -> %.6e monotone %d volume %.6f\n", J0, Jlast, monotone ? 1 : 0, volLast + 0.5);
  std::printf("design restart file round trip %d\n", restart_ok ? 1 : 0);
  const bool ok = rel < 1e-12 && Jlast < 0.95 * J0 && monotone && volLast <= 0.02 && restart_ok;
  std::printf(ok ? "RESULT OK\n" : "RESULT FAIL\n");
  return ok ? 0 : 2;
}
