// The reference application's unfused forward and reverse loops
// (applications/RobustPump2d/robustPump2d.cc:233-277) written against the lbm_topopt host
// mirror with caller-owned checkpointing: makeCheckpoint after every stream,
// getForwardSolution + adjointCollide + computeCollideSource + adjointStream per reverse step.
//
// Two CheckpointingBase implementations drive it: the mirror's LBNaiveCheckpointing (every
// state on the host) and one written here that keeps every fourth state and re-advances the
// solver between them. Both must give the objective, dJ/dgamma and the adjoint populations of
// the fused pass (UnsteadyLBTopOpt::forwardAndAdjoint) bit for bit: the same sweeps run on the
// same states. So must the fused pass with the binomial checkpoint placement. Prints the
// comparison lines and RESULT OK / RESULT FAIL.
#include <algorithm>
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

// Keeps the states of steps that are multiples of `stride`; any other step is recomputed from
// the nearest kept state below it by advancing the solver.
class StridedRecomputeCheckpointing : public CheckpointingBase {
 public:
  StridedRecomputeCheckpointing(LBSolver& s, PetscInt stride) : solver_(s), stride_(stride) {}
  PetscErrorCode makeCheckpoint(PetscInt step, const std::vector<double>& f, void*) override {
    if (step % stride_ != 0) return 0;
    if ((size_t)step >= kept_.size()) kept_.resize((size_t)step + 1);
    kept_[(size_t)step] = f;
    return 0;
  }
  PetscErrorCode getForwardSolution(PetscInt step, std::vector<double>& f, void* ctx) override {
    const PetscInt base = step - step % stride_;
    if ((size_t)base >= kept_.size() || kept_[(size_t)base].empty()) return LBMX_ERR_STATE;
    PetscErrorCode e = setSolverState(base, kept_[(size_t)base], ctx);
    if (e) return e;
    for (PetscInt k = base; k < step; ++k) {
      e = advance(ctx);
      if (e) return e;
      ++recomputed;
    }
    return copyFromSolverState(f, ctx);
  }
  PetscErrorCode reset(void*) override {
    kept_.clear();
    return 0;
  }
  long long recomputed = 0;

 protected:
  PetscErrorCode advance(void*) override { return solver_.collideAndStream(); }
  PetscErrorCode setSolverState(PetscInt step, const std::vector<double>& f, void*) override {
    return solver_.setDistributions(f, step);
  }
  PetscErrorCode copyFromSolverState(std::vector<double>& f, void*) override { return solver_.getDistributions(&f); }

 private:
  LBSolver& solver_;
  PetscInt stride_;
  std::vector<std::vector<double>> kept_;
};

static double max_abs_diff(const std::vector<double>& a, const std::vector<double>& b) {
  double m = a.size() == b.size() ? 0.0 : INFINITY;
  for (size_t i = 0; i < std::min(a.size(), b.size()); ++i) m = std::max(m, std::fabs(a[i] - b[i]));
  return m;
}

int main() {
  const PetscInt nx = 64, ny = 28, numSteps = 60;
  using Operator = BrinkmanCollision<IncompressibleMRT2d<D2Q9>>;
  NewLBSolverInfo2d info;
  info.nx = nx;
  info.ny = ny;
  IsothermalLBSolver2d<Operator> solver;
  CHK(solver.make(info, Operator(IncompressibleMRT2d<D2Q9>(1.25), BPInterpolation(0.1), 0.1)));
  const double mid = (ny - 1) / 2.0;
  CHK(solver.addBoundaryCondition(Box2d(0, nx - 1, 0, 0), OnGridBouncebackDescriptor()));
  CHK(solver.addBoundaryCondition(Box2d(0, nx - 1, ny - 1, ny - 1), OnGridBouncebackDescriptor()));
  CHK(solver.addBoundaryCondition(Box2d(0, 0, 1, ny - 2), ZouHeParabolicVelocityDescriptor(0.05, mid, mid)));
  CHK(solver.addBoundaryCondition(Box2d(nx - 1, nx - 1, 1, ny - 2), ZouHePressureDescriptor(1.0)));

  // a smooth obstacle in the channel
  std::vector<double> design;
  CHK(solver.createDesignVec(&design));
  for (PetscInt j = 0; j < ny; ++j)
    for (PetscInt i = 0; i < nx; ++i) {
      const double dx = (i - 0.45 * nx) / 6.0, dy = (j - mid) / 5.0;
      design[(size_t)j * nx + i] = 1.0 - 0.9 * std::exp(-(dx * dx + dy * dy));
    }
  CHK(solver.setFieldFromDesignVec(design));
  // solid blocks at the inlet corners, as applications/PressureDiode2d/pressureDiode2d.cc:159-163 places them
  const PetscScalar solid[1] = {0.};
  CHK(solver.setUniformFieldValues(Box2d(0, 3, 1, 4), solid));
  CHK(solver.setUniformFieldValues(Box2d(0, 3, ny - 5, ny - 2), solid));
  std::vector<double> placed;
  CHK(solver.createDesignVec(&placed));
  bool blocks_ok = placed[(size_t)2 * nx + 1] == 0.0 && placed[(size_t)(ny - 3) * nx + 3] == 0.0 &&
                   placed[(size_t)6 * nx + 1] == design[(size_t)6 * nx + 1];

  AdjointLBSolver adjSolver;
  CHK(solver.makeSourceAdjoint(&adjSolver));
  const auto objective = PressureDrop2d(Box2d(0, 0, 1, ny - 2), Box2d(nx - 1, nx - 1, 1, ny - 2));
  const PetscScalar timeAveraging = 1.0 / numSteps;

  // the fused pass: forward with the engine's checkpoints, reverse with recomputation
  UnsteadyLBTopOpt topOpt(solver, adjSolver, objective, numSteps, timeAveraging);
  CHK(topOpt.allocateMemoryForForwardSolution(8));
  CHK(topOpt.forwardAndAdjoint());
  const double Jfused = topOpt.getObjective();
  std::vector<double> sensFused, adjFused;
  CHK(adjSolver.getSensitivities(&sensFused));
  CHK(adjSolver.getAdjoints(&adjFused, 9));
  double sensScale = 0.0;
  for (double v : sensFused) sensScale = std::max(sensScale, std::fabs(v));

  // the same fused pass with the binomial placement of the 8 checkpoints: same states, same sweeps
  CHK(topOpt.allocateMemoryForForwardSolution(8, LBMX_CHECKPOINT_BINOMIAL));
  CHK(topOpt.forwardAndAdjoint());
  std::vector<double> sensBinomial;
  CHK(adjSolver.getSensitivities(&sensBinomial));
  const bool binomial_ok = topOpt.getObjective() == Jfused && max_abs_diff(sensBinomial, sensFused) == 0.0;
  std::printf("binomial placement bitwise equal %d\n", binomial_ok ? 1 : 0);
  CHK(topOpt.allocateMemoryForForwardSolution(8));

  LBNaiveCheckpointing naive(solver);
  StridedRecomputeCheckpointing strided(solver, 4);
  CheckpointingBase* schemes[2] = {&naive, &strided};
  const char* names[2] = {"host storage of every state", "every 4th state, recompute"};
  bool ok = sensScale > 0.0 && blocks_ok && binomial_ok;
  std::printf("setUniformFieldValues blocks placed %d\n", blocks_ok ? 1 : 0);
  std::vector<double> f;
  for (int k = 0; k < 2; ++k) {
    CheckpointingBase& check = *schemes[k];
    CHK(check.reset(nullptr));
    // forward loop (robustPump2d.cc:234-251)
    CHK(solver.initializeAtEquilibrium());
    CHK(solver.getDistributions(&f));
    CHK(check.makeCheckpoint(0, f, nullptr));
    PetscInt step = 0;
    double J = 0.0;
    while (step < numSteps) {
      CHK(solver.collideAndStream());
      ++step;
      CHK(solver.getDistributions(&f));
      CHK(check.makeCheckpoint(step, f, nullptr));
      double v = 0.0;
      CHK(solver.computeMacroFunctional(objective, timeAveraging, &v));
      J += v;
    }
    // reverse loop (robustPump2d.cc:263-277)
    CHK(adjSolver.resetAdjoints());
    CHK(adjSolver.setCurrentTimestep(step));
    while (step > 0) {
      CHK(check.getForwardSolution(step, f, nullptr));
      CHK(adjSolver.adjointCollide(f));
      CHK(adjSolver.computeCollideSource(objective, timeAveraging));
      CHK(adjSolver.adjointStream());
      --step;
    }
    std::vector<double> sens, adj;
    CHK(adjSolver.getSensitivities(&sens));
    CHK(adjSolver.getAdjoints(&adj, 9));
    const double dJ = std::fabs(J - Jfused) / std::fabs(Jfused);
    const double ds = max_abs_diff(sens, sensFused) / sensScale;
    const double da = max_abs_diff(adj, adjFused);
    std::printf("%s: J rel %.3e, dJ/dgamma max rel %.3e, adjoint max abs %.3e, adjoint step %d\n", names[k], dJ, ds,
                da, (int)adjSolver.currentTimestep());
    ok = ok && dJ < 1e-12 && ds == 0.0 && da == 0.0 && adjSolver.currentTimestep() == 0;
  }
  std::printf("recomputed forward steps (strided scheme): %lld\n", strided.recomputed);
  ok = ok && strided.recomputed > 0;
  std::printf(ok ? "RESULT OK\n" : "RESULT FAIL\n");
  return ok ? 0 : 2;
}
