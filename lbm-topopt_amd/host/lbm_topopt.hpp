// C++ host mirror of the reference's interfaces over the lbmx C ABI (include/lbmx.h).
//
// Application code written against the reference's classes (IsothermalLBSolver2d, LBSolver3d,
// AdjointLBSolver, UnsteadyLBTopOpt2d, the filters, FullToDesignMapping2d, MMA) keeps its
// structure: the same class and method names, argument meaning and PetscErrorCode-style
// status returns (0 = success; the codes are the PETSc error classes of lbmx_status). What
// changes:
//   * PETSc Vec arguments are std::vector<double> in the single-rank DMDA order;
//   * the MPI communicator argument is gone (one GPU per process);
//   * the per-step reverse loop (adjointCollide / computeCollideSource / adjointStream with a
//     checkpointer) is the single call UnsteadyLBTopOpt::forwardAndAdjoint(), as the
//     reference's own UnsteadyLBTopOpt2d::forwardAndAdjoint drives it
//     (src/topOpt/unsteadyLBTopOpt2d.cc:77-106).
// Header-only; link with liblbmx.so.
#pragma once

#include <cmath>
#include <cstring>
#include <memory>
#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "lbmx.h"

namespace lbm_topopt {

using PetscErrorCode = int;
using PetscInt = int;
using PetscScalar = double;

#define LBM_CHKERRQ(call)                  \
  do {                                     \
    ::lbm_topopt::PetscErrorCode e_ = (call); \
    if (e_) return e_;                     \
  } while (0)

// ---- geometry (src/core/globalDefinitions.hh:25-71, src/core/geometry2d.hh) ----
class Range {
 public:
  Range(PetscInt b = 0, PetscInt e = -1) : b_(b), e_(e) {}
  PetscInt getBeginId() const { return b_; }
  PetscInt getEndId() const { return e_; }
  PetscInt size() const { return e_ >= b_ ? e_ - b_ + 1 : 0; }

 private:
  PetscInt b_, e_;
};

class Box2d {
 public:
  Box2d() = default;
  Box2d(PetscInt x0, PetscInt x1, PetscInt y0, PetscInt y1) : xRange(x0, x1), yRange(y0, y1) {}
  PetscInt getNx() const { return xRange.size(); }
  PetscInt getNy() const { return yRange.size(); }
  lbmx_box raw() const {
    return lbmx_box{xRange.getBeginId(), xRange.getEndId(), yRange.getBeginId(), yRange.getEndId(), 0, 0};
  }
  Range xRange, yRange;
};

class Box3d {
 public:
  Box3d() = default;
  Box3d(PetscInt x0, PetscInt x1, PetscInt y0, PetscInt y1, PetscInt z0, PetscInt z1)
      : xRange(x0, x1), yRange(y0, y1), zRange(z0, z1) {}
  PetscInt getNx() const { return xRange.size(); }
  PetscInt getNy() const { return yRange.size(); }
  PetscInt getNz() const { return zRange.size(); }
  lbmx_box raw() const {
    return lbmx_box{xRange.getBeginId(), xRange.getEndId(), yRange.getBeginId(),
                    yRange.getEndId(),   zRange.getBeginId(), zRange.getEndId()};
  }
  Range xRange, yRange, zRange;
};

// ---- lattices (src/latticeBoltzmann/lattices2d.hh:16-84, lattices3d.hh:34-53) ----
struct D2Q9 {
  static constexpr int numDOF = 9;
  static constexpr int dims = 2;
};
struct D3Q19 {
  static constexpr int numDOF = 19;
  static constexpr int dims = 3;
};
struct D2Q5 {
  static constexpr int numDOF = 5;
  static constexpr int dims = 2;
};

// ---- interpolations (src/core/interpolationFunctions.hh:7-54) ----
struct BPInterpolation {
  explicit BPInterpolation(PetscScalar q_ = 0.1) : q(q_) {}
  PetscScalar q;
};
struct RampInterpolation {
  RampInterpolation(PetscScalar p_, PetscScalar ck_) : p(p_), ck(ck_) {}
  PetscScalar p, ck;
};

// ---- collision operators: each fills the lattice/operator/coupling part of an lbmx_desc ----
template <class L>
struct IncompressibleBGK2d {  // src/dynamics/incompressibleBGK2d.hh:94-153
  using Lattice = L;
  explicit IncompressibleBGK2d(PetscScalar omega_) : omega(omega_) {}
  void fill(lbmx_desc& d) const { d.lattice = LBMX_D2Q9, d.collision = LBMX_BGK, d.omega = omega; }
  PetscScalar omega;
};
template <class L>
struct IncompressibleMRT2d {  // src/dynamics/incompressibleMRT2d.hh:18-98
  using Lattice = L;
  explicit IncompressibleMRT2d(PetscScalar omega_) : omega(omega_) {}
  void fill(lbmx_desc& d) const { d.lattice = LBMX_D2Q9, d.collision = LBMX_MRT, d.omega = omega; }
  PetscScalar omega;
};
template <class L>
struct IncompressibleCascaded2d {  // src/dynamics/cascaded2d.hh:15-97
  using Lattice = L;
  explicit IncompressibleCascaded2d(PetscScalar omega_) : omega(omega_) {}
  void fill(lbmx_desc& d) const { d.lattice = LBMX_D2Q9, d.collision = LBMX_CASCADED, d.omega = omega; }
  PetscScalar omega;
};
template <class L>
struct IncompressibleMRT3d {  // src/dynamics/MRT3d.hh:28-200
  using Lattice = L;
  explicit IncompressibleMRT3d(PetscScalar omega_) : omega(omega_) {}
  void fill(lbmx_desc& d) const { d.lattice = LBMX_D3Q19, d.collision = LBMX_MRT, d.omega = omega; }
  PetscScalar omega;
};
template <class L>
struct StandardMRT3d {  // src/dynamics/MRT3d.hh:206-383 (forward only)
  using Lattice = L;
  explicit StandardMRT3d(PetscScalar omega_) : omega(omega_) {}
  void fill(lbmx_desc& d) const { d.lattice = LBMX_D3Q19, d.collision = LBMX_MRT_STANDARD, d.omega = omega; }
  PetscScalar omega;
};
// PartialBouncebackCollision<Op, BPInterpolation> (src/dynamics/partialBouncebackCollision.hh:53-65)
template <class Op, class Interp = BPInterpolation>
struct PartialBouncebackCollision {
  using Lattice = typename Op::Lattice;
  PartialBouncebackCollision(Op op_, Interp in_) : op(op_), interp(in_) {}
  void fill(lbmx_desc& d) const {
    op.fill(d);
    d.design = LBMX_DESIGN_PBB;
    d.interp_q = interp.q;
    d.alpha_max = 1.0;
  }
  Op op;
  Interp interp;
};
// Brinkman drag -alpha_max BP_q(gamma) u with Guo forcing (defined by the engine)
template <class Op, class Interp = BPInterpolation>
struct BrinkmanCollision {
  using Lattice = typename Op::Lattice;
  BrinkmanCollision(Op op_, Interp in_, PetscScalar alphaMax_) : op(op_), interp(in_), alphaMax(alphaMax_) {}
  void fill(lbmx_desc& d) const {
    op.fill(d);
    d.design = LBMX_DESIGN_BRINKMAN;
    d.interp_q = interp.q;
    d.alpha_max = alphaMax;
  }
  Op op;
  Interp interp;
  PetscScalar alphaMax;
};
// ThermalObstacleCollision<ThermalMRT2d<D2Q5_TMRT>, Ramp> with the flow operator
// (src/dynamics/thermalObstacleCollision.hh:47-53, thermalDualCollision.hh:57-64)
template <class FlowOp>
struct ThermalDualCollision {
  using Lattice = D2Q9;
  ThermalDualCollision(FlowOp flow_, PetscScalar diffusivity_, RampInterpolation ramp_)
      : flow(flow_), diffusivity(diffusivity_), ramp(ramp_) {}
  void fill(lbmx_desc& d) const {
    flow.fill(d);
    d.lattice = LBMX_D2Q9_D2Q5;
    d.thermal_diffusivity = diffusivity;
    d.ramp_p = ramp.p;
    d.ramp_ck = ramp.ck;
  }
  FlowOp flow;
  PetscScalar diffusivity;
  RampInterpolation ramp;
};

// ---- solver info (src/latticeBoltzmann/NewLBSolver2d.hh:10-25) ----
struct NewLBSolverInfo2d {
  PetscInt nx = 0, ny = 0;
};
struct LBSolverInfo3d {
  PetscInt nx = 0, ny = 0, nz = 0;
};

// ---- boundary descriptors (src/boundaryConditions/boundaryDescriptor2d.hh:53-73) ----
struct NewBoundaryDescriptor2d {
  lbmx_bc_desc desc{};
};
using BoundaryDescriptor3d = NewBoundaryDescriptor2d;

inline NewBoundaryDescriptor2d OnGridBouncebackDescriptor() {  // bounceback2d.cc, bounceback3d.cc
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_BOUNCEBACK;
  return b;
}
// Zou-He velocity with a parabolic InflowVelocity functor (applications/diffuser/diffuser.cc:21-31);
// a sinusoidal time factor offset + amp sin(2 pi t / period) when period > 0
// (applications/RobustPump2d/robustPump2d.cc:14-19)
inline NewBoundaryDescriptor2d ZouHeParabolicVelocityDescriptor(PetscScalar uMax, PetscScalar centre,
                                                                PetscScalar radius, PetscScalar period = 0,
                                                                PetscScalar amp = 0, PetscScalar offset = 1) {
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_ZOUHE_VELOCITY;
  b.desc.profile = LBMX_PROFILE_PARABOLIC;
  b.desc.value = uMax;
  b.desc.dir[0] = 1.0;
  b.desc.center[0] = centre;
  b.desc.radius[0] = radius;
  b.desc.radius[1] = 1.0;
  if (period > 0) {
    b.desc.time_mode = LBMX_TIME_SINE;
    b.desc.t_period = period;
    b.desc.t_amp = amp;
    b.desc.t_offset = offset;
  }
  return b;
}
inline NewBoundaryDescriptor2d ZouHePressureDescriptor(PetscScalar rho) {  // ZouHe2d.hh:245-281
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_ZOUHE_PRESSURE;
  b.desc.value = rho;
  return b;
}
inline NewBoundaryDescriptor2d HechtVelocityDescriptor(PetscScalar u) {  // Hecht3d.hh:69-89
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_HECHT_VELOCITY;
  b.desc.value = u;
  b.desc.dir[0] = 1.0;
  return b;
}
inline NewBoundaryDescriptor2d HechtPressureDescriptor(PetscScalar rho) {  // Hecht3d.hh:182-204
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_HECHT_PRESSURE;
  b.desc.value = rho;
  return b;
}
inline NewBoundaryDescriptor2d ThermalZouHeDescriptor(PetscScalar T) {  // thermalZouHe2d.hh:62-93
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_THERMAL_DIRICHLET;
  b.desc.value = T;
  return b;
}
inline NewBoundaryDescriptor2d ThermalBouncebackDescriptor() {  // thermalBounceback2d.hh:273-298
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_THERMAL_BOUNCEBACK;
  return b;
}
// ZouHeNeumann2d (ZouHe2d.hh:528-566): east/west outflow, u_x from the two nodes inwards
inline NewBoundaryDescriptor2d ZouHeNeumannDescriptor() {
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_ZOUHE_NEUMANN;
  return b;
}
// ThermalZouHeNeumann2d (src/descriptors/thermalZouHeDescriptor2d.hh:55-116): T from the two nodes inwards
inline NewBoundaryDescriptor2d ThermalZouHeNeumannDescriptor() {
  NewBoundaryDescriptor2d b;
  b.desc.type = LBMX_BC_THERMAL_NEUMANN;
  return b;
}

// ---- macroscopic functionals (src/functionals/LBMacroFunctional.hh:6-14) ----
// One of the engine's functionals (lbmx_objective_kind with its boxes), or a linear box
// functional given by its terms (lbmx_term): the form an application-defined functional takes
// here in place of an LBMacroFunctional subclass.
class LBMacroFunctional {
 public:
  LBMacroFunctional(int kind, lbmx_box a, lbmx_box b = lbmx_box{0, -1, 0, -1, 0, -1}) : kind_(kind), a_(a), b_(b) {}
  explicit LBMacroFunctional(std::vector<lbmx_term> terms) : kind_(-1), a_{}, b_{}, terms_(std::move(terms)) {}
  lbmx_objective objective(PetscScalar timeConst) const {
    lbmx_objective o{};
    o.kind = kind_ < 0 ? LBMX_OBJ_NONE : kind_;
    o.time_const = timeConst;
    o.box_a = a_;
    o.box_b = b_;
    return o;
  }
  // sets this functional on a solver (lbmx_set_objective or lbmx_set_objective_terms)
  lbmx_status install(lbmx_solver* h, PetscScalar timeConst) const {
    if (kind_ < 0) return lbmx_set_objective_terms(h, terms_.data(), (int)terms_.size(), timeConst);
    const lbmx_objective o = objective(timeConst);
    return lbmx_set_objective(h, &o);
  }
  // J_t of a state from its macros (rho, ux, uy, uz, T per node, natural order), the value the
  // engine's forward sweep accumulates for the same state
  PetscScalar evaluate(const std::vector<double>& mac, PetscInt nx, PetscInt ny) const {
    auto sum = [&](const lbmx_box& b, auto&& g) {
      double s = 0.0;
      for (int z = b.z0; z <= b.z1; ++z)
        for (int y = b.y0; y <= b.y1; ++y)
          for (int x = b.x0; x <= b.x1; ++x) s += g(&mac[5 * (((size_t)z * ny + y) * nx + x)]);
      return s;
    };
    auto count = [](const lbmx_box& b) { return (double)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1) * (b.z1 - b.z0 + 1); };
    auto rho = [](const double* m) { return m[0]; };
    auto ux = [](const double* m) { return m[1]; };
    auto power = [](const double* m) {
      return (m[0] / 3.0 + 0.5 * (m[1] * m[1] + m[2] * m[2] + m[3] * m[3])) * m[1];
    };
    auto heat = [](const double* m) { return m[1] * m[4]; };
    auto temperature = [](const double* m) { return m[4]; };
    if (kind_ < 0) {
      double J = 0.0;
      for (const lbmx_term& t : terms_) {
        switch (t.kind) {
          case LBMX_TERM_DENSITY: J += t.weight * sum(t.box, rho); break;
          case LBMX_TERM_VELOCITY_X: J += t.weight * sum(t.box, ux); break;
          case LBMX_TERM_TOTAL_PRESSURE_FLUX: J += t.weight * sum(t.box, power); break;
          case LBMX_TERM_HEAT_FLUX_X: J += t.weight * sum(t.box, heat); break;
          case LBMX_TERM_TEMPERATURE: J += t.weight * sum(t.box, temperature); break;
          default: break;
        }
      }
      return J;
    }
    switch (kind_) {
      case LBMX_OBJ_PRESSURE_DROP: return sum(a_, rho) / count(a_) - sum(b_, rho) / count(b_);
      case LBMX_OBJ_OUTFLOW: return -sum(a_, ux) / count(a_);
      case LBMX_OBJ_THROUGHFLOW: return -sum(a_, ux) / count(a_) - sum(b_, ux) / count(b_);
      case LBMX_OBJ_DISSIPATED_POWER: return sum(a_, power) - sum(b_, power);
      case LBMX_OBJ_HEAT_FLUX: return -sum(a_, heat);
      default: return 0.0;
    }
  }

 private:
  int kind_;
  lbmx_box a_, b_;
  std::vector<lbmx_term> terms_;
};
inline LBMacroFunctional PressureDrop2d(const Box2d& in, const Box2d& out) {  // pressureDrop2d.hh:14-52
  return LBMacroFunctional(LBMX_OBJ_PRESSURE_DROP, in.raw(), out.raw());
}
inline LBMacroFunctional PressureDrop3d(const Box3d& in, const Box3d& out) {
  return LBMacroFunctional(LBMX_OBJ_PRESSURE_DROP, in.raw(), out.raw());
}
inline LBMacroFunctional OutflowEast2d(const Box2d& out) {  // outflow2d.hh:15-54
  return LBMacroFunctional(LBMX_OBJ_OUTFLOW, out.raw());
}
inline LBMacroFunctional ThroughFlow2d(const Box2d& in, const Box2d& out) {  // throughFlow2d.hh:16-88
  return LBMacroFunctional(LBMX_OBJ_THROUGHFLOW, in.raw(), out.raw());
}
inline LBMacroFunctional DissipatedPower2d(const Box2d& in, const Box2d& out) {
  return LBMacroFunctional(LBMX_OBJ_DISSIPATED_POWER, in.raw(), out.raw());
}
inline LBMacroFunctional ThermalOutflowEast2d(const Box2d& out) {  // temperatureFlow2d.hh:18-54
  return LBMacroFunctional(LBMX_OBJ_HEAT_FLUX, out.raw());
}
// DomainTemperature2d (src/functionals/domainTemperature2d.hh): the mean temperature over the
// domain box, as one explicit term
inline LBMacroFunctional DomainTemperature2d(const Box2d& domain) {
  const lbmx_box b = domain.raw();
  const double area = (double)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1);
  return LBMacroFunctional(std::vector<lbmx_term>{lbmx_term{LBMX_TERM_TEMPERATURE, 0, 1.0 / area, b}});
}

// ---- binary design files (src/core/binaryIO.cc:3-34): PETSc's binary Vec format ----
inline PetscErrorCode writeVectorToBinary(const std::string& folder, const std::string& name,
                                          const std::vector<double>& v) {
  return lbmx_vec_write_binary((folder + name).c_str(), v.data(), v.size());
}
inline PetscErrorCode loadVectorFromBinary(const std::string& folder, const std::string& name,
                                           std::vector<double>& v) {
  return lbmx_vec_load_binary((folder + name).c_str(), v.data(), v.size());
}

// DMDAGlobalToNatural / DMDANaturalToGlobal for a DMDA split over lx.size() x ly.size() x
// lz.size() ranks (ownership ranges as DMDAGetOwnershipRanges gives them), dof values per node
inline PetscErrorCode DMDAGlobalToNatural(const int dims[3], const std::vector<int>& lx, const std::vector<int>& ly,
                                          const std::vector<int>& lz, int dof, const std::vector<double>& global,
                                          std::vector<double>& natural) {
  const int procs[3] = {(int)lx.size(), (int)ly.size(), lz.empty() ? 1 : (int)lz.size()};
  natural.resize(global.size());
  return lbmx_dmda_global_to_natural(dims, procs, lx.data(), ly.data(), lz.empty() ? nullptr : lz.data(), dof,
                                     global.data(), natural.data());
}
inline PetscErrorCode DMDANaturalToGlobal(const int dims[3], const std::vector<int>& lx, const std::vector<int>& ly,
                                          const std::vector<int>& lz, int dof, const std::vector<double>& natural,
                                          std::vector<double>& global) {
  const int procs[3] = {(int)lx.size(), (int)ly.size(), lz.empty() ? 1 : (int)lz.size()};
  global.resize(natural.size());
  return lbmx_dmda_natural_to_global(dims, procs, lx.data(), ly.data(), lz.empty() ? nullptr : lz.data(), dof,
                                     natural.data(), global.data());
}

class AdjointLBSolver;

// ---- solvers (src/latticeBoltzmann/LBSolver.hh:14-73, NewLBSolver2d.hh, isothermalLBSolver2d.hh) ----
class LBSolver {
 public:
  LBSolver() = default;
  LBSolver(const LBSolver&) = delete;
  LBSolver& operator=(const LBSolver&) = delete;
  virtual ~LBSolver() {
    if (h_) lbmx_destroy(h_);
  }
  PetscErrorCode addBoundaryCondition(const Box2d& b, const NewBoundaryDescriptor2d& bc) {
    const lbmx_box r = b.raw();
    return lbmx_add_boundary(h_, &r, &bc.desc);
  }
  PetscErrorCode addBoundaryCondition(const Box3d& b, const NewBoundaryDescriptor2d& bc) {
    const lbmx_box r = b.raw();
    return lbmx_add_boundary(h_, &r, &bc.desc);
  }
  // LBSolver::setFieldFromDesignVec (src/latticeBoltzmann/LBSolver.cc:333-348)
  PetscErrorCode setFieldFromDesignVec(const std::vector<double>& g) { return lbmx_set_design(h_, g.data(), g.size()); }
  // NewLBSolver2d/LBSolver3d::setUniformFieldValues (src/latticeBoltzmann/NewLBSolver2d.cc:43-65,
  // LBSolver3d.cc:43): the material field takes values[0] on the box (clipped to the grid)
  PetscErrorCode setUniformFieldValues(const Box2d& b, const PetscScalar* values) {
    return setUniformFieldValuesRaw(b.raw(), values);
  }
  PetscErrorCode setUniformFieldValues(const Box3d& b, const PetscScalar* values) {
    return setUniformFieldValuesRaw(b.raw(), values);
  }
  // LBSolver::createDesignVec (LBSolver.cc:318-331): a copy of the material field
  PetscErrorCode createDesignVec(std::vector<double>* v) const {
    v->assign((size_t)nodes(), 1.0);
    return lbmx_get_design(h_, v->data(), v->size());
  }
  // LBSolver::initializeAtEquilibrium (LBSolver.cc:234-249): rho = 1, u = 0 (T = 0)
  PetscErrorCode initializeAtEquilibrium() {
    const double u0[3] = {0, 0, 0};
    return lbmx_init_equilibrium(h_, 1.0, u0, 0.0);
  }
  // LBSolver::collideAndStream (LBSolver.cc:108-115): one step
  PetscErrorCode collideAndStream() {
    double j = 0.0;
    return lbmx_forward(h_, 1, &j);
  }
  // LBSolver::computeMacroFunctional: timeConst * J_t of the current state
  PetscErrorCode computeMacroFunctional(const LBMacroFunctional& f, PetscScalar timeConst, PetscScalar* val) const {
    std::vector<double> mac((size_t)nodes() * 5);
    PetscErrorCode e = lbmx_get_macros(h_, mac.data(), mac.size());
    if (e) return e;
    *val = timeConst * f.evaluate(mac, nx_, ny_);
    return 0;
  }
  // distributionsGlobal (AoS) and LBSolver::setDistributions + setCurrentTimestep (LBSolver.cc:65-97)
  PetscErrorCode getDistributions(std::vector<double>* f) const {
    f->resize((size_t)nodes() * numDOF_);
    return lbmx_get_state(h_, f->data(), f->size());
  }
  PetscErrorCode setDistributions(const std::vector<double>& f, PetscInt timestep) {
    return lbmx_set_state(h_, f.data(), f.size(), timestep);
  }
  // LBSolver::outputFields (src/latticeBoltzmann/LBSolver.cc:251-286): macros and design, .vts
  PetscErrorCode outputFields(const std::string& folder, const std::string& name) const {
    return lbmx_write_vts(h_, (folder + name).c_str());
  }
  // LBSolver::makeSourceAdjoint: the adjoint solver shares this engine's adjoint fields
  PetscErrorCode makeSourceAdjoint(AdjointLBSolver* adj);
  lbmx_solver* handle() const { return h_; }
  long long nodes() const { return (long long)nx_ * ny_ * nz_; }
  PetscInt nx() const { return nx_; }
  PetscInt ny() const { return ny_; }
  PetscInt nz() const { return nz_; }

 protected:
  PetscErrorCode setUniformFieldValuesRaw(const lbmx_box& b, const PetscScalar* values) {
    std::vector<double> g((size_t)nodes());
    PetscErrorCode e = lbmx_get_design(h_, g.data(), g.size());
    if (e) return e;
    for (int z = std::max(b.z0, 0); z <= std::min(b.z1, (int)nz_ - 1); ++z)
      for (int y = std::max(b.y0, 0); y <= std::min(b.y1, (int)ny_ - 1); ++y)
        for (int x = std::max(b.x0, 0); x <= std::min(b.x1, (int)nx_ - 1); ++x)
          g[((size_t)z * ny_ + y) * nx_ + x] = values[0];
    return lbmx_set_design(h_, g.data(), g.size());
  }
  PetscErrorCode create(lbmx_desc d) {
    nx_ = d.nx, ny_ = d.ny, nz_ = d.nz;
    numDOF_ = d.lattice == LBMX_D3Q19 ? 19 : (d.lattice == LBMX_D2Q9_D2Q5 ? 14 : 9);
    return lbmx_create(&d, &h_);
  }
  lbmx_solver* h_ = nullptr;
  PetscInt nx_ = 0, ny_ = 0, nz_ = 1, numDOF_ = 9;
};

template <class CollisionOperator>
class IsothermalLBSolver2d : public LBSolver {  // isothermalLBSolver2d.hh:20-223
 public:
  PetscErrorCode make(const NewLBSolverInfo2d& info, CollisionOperator op, int precision = LBMX_FP64) {
    lbmx_desc d{};
    d.nx = info.nx;
    d.ny = info.ny;
    d.nz = 1;
    d.precision = precision;
    op.fill(d);
    return create(d);
  }
  Box2d getBoundingBox() const { return Box2d(0, nx_ - 1, 0, ny_ - 1); }
};

template <class CollisionOperator>
class LBSolver3d : public LBSolver {  // src/latticeBoltzmann/LBSolver3d.cc
 public:
  PetscErrorCode make(const LBSolverInfo3d& info, CollisionOperator op, int precision = LBMX_FP64) {
    lbmx_desc d{};
    d.nx = info.nx;
    d.ny = info.ny;
    d.nz = info.nz;
    d.precision = precision;
    op.fill(d);
    return create(d);
  }
  Box3d getBoundingBox() const { return Box3d(0, nx_ - 1, 0, ny_ - 1, 0, nz_ - 1); }
};

// ---- adjoint (src/adjointLatticeBoltzmann/adjointLBSolver.hh:11-46) ----
class AdjointLBSolver {
 public:
  // AdjointLBSolver::getSensitivities: dJ/dgamma of the last forward + adjoint pass
  PetscErrorCode getSensitivities(std::vector<double>* sens) const {
    if (!solver_) return LBMX_ERR_STATE;
    sens->resize((size_t)solver_->nodes());
    return lbmx_get_sensitivity(solver_->handle(), sens->data(), sens->size());
  }
  // the adjoint populations nu_0 (AoS), adjointGlobal
  PetscErrorCode getAdjoints(std::vector<double>* adj, PetscInt numDOF) const {
    if (!solver_) return LBMX_ERR_STATE;
    adj->resize((size_t)solver_->nodes() * numDOF);
    return lbmx_get_adjoint(solver_->handle(), adj->data(), adj->size());
  }

  // The reverse loop step by step (applications/RobustPump2d/robustPump2d.cc:263-277), for a
  // caller that restores the forward states itself (a CheckpointingBase below).
  // resetAdjoints (adjointLBSolver.cc:138-145) and setCurrentTimestep (:147-151)
  PetscErrorCode resetAdjoints() {
    if (!solver_) return LBMX_ERR_STATE;
    source_ = false;
    return lbmx_adjoint_reset(solver_->handle(), (int)t_);
  }
  PetscErrorCode setCurrentTimestep(PetscInt t) {
    if (!solver_) return LBMX_ERR_STATE;
    t_ = t;
    return lbmx_adjoint_set_timestep(solver_->handle(), (int)t);
  }
  // adjointCollide(f, material) (adjointLBSolver.cc:28-48): f is the forward state of the
  // current step (getForwardSolution); the material is the solver's design field. The engine
  // fuses the reverse collision, the source and the transposed stream into one sweep, which
  // runs at adjointStream.
  PetscErrorCode adjointCollide(const std::vector<double>& f) {
    if (!solver_) return LBMX_ERR_STATE;
    source_ = false;
    return lbmx_set_state(solver_->handle(), f.data(), f.size(), (int)t_);
  }
  // computeCollideSource(objective, timeConst, f, material) (adjointLBSolver.cc:83-102)
  PetscErrorCode computeCollideSource(const LBMacroFunctional& obj, PetscScalar timeConst) {
    if (!solver_) return LBMX_ERR_STATE;
    PetscErrorCode e = obj.install(solver_->handle(), timeConst);
    if (!e) source_ = true;
    return e;
  }
  // adjointStream(f) (adjointLBSolver.cc:50-81), then the step counter moves to t - 1
  PetscErrorCode adjointStream() {
    if (!solver_) return LBMX_ERR_STATE;
    PetscErrorCode e = lbmx_adjoint_step(solver_->handle(), source_ ? 1 : 0);
    source_ = false;
    if (!e) --t_;
    return e;
  }
  PetscInt currentTimestep() const { return t_; }

 private:
  friend class LBSolver;
  LBSolver* solver_ = nullptr;
  PetscInt t_ = 0;
  bool source_ = false;
};

// ---- checkpointing plug-in interface (src/topOpt/checkpointingBase.hh:7-18) on host AoS
//      state vectors (getDistributions / setDistributions) ----
class CheckpointingBase {
 public:
  virtual ~CheckpointingBase() = default;
  virtual PetscErrorCode makeCheckpoint(PetscInt step, const std::vector<double>& f, void* userContext) = 0;
  virtual PetscErrorCode getForwardSolution(PetscInt step, std::vector<double>& f, void* userContext) = 0;
  virtual PetscErrorCode reset(void* userContext) = 0;

 protected:
  virtual PetscErrorCode advance(void* userContext) = 0;
  virtual PetscErrorCode setSolverState(PetscInt step, const std::vector<double>& f, void* userContext) = 0;
  virtual PetscErrorCode copyFromSolverState(std::vector<double>& f, void* userContext) = 0;
};

// Every state kept in host memory (the reference's ForwardSolutionStorage pattern): the
// simplest CheckpointingBase; the device engine's own Wang checkpointing is
// UnsteadyLBTopOpt::allocateMemoryForForwardSolution.
class LBNaiveCheckpointing : public CheckpointingBase {
 public:
  explicit LBNaiveCheckpointing(LBSolver& s) : solver_(s) {}
  PetscErrorCode makeCheckpoint(PetscInt step, const std::vector<double>& f, void*) override {
    if (step < 0) return LBMX_ERR_ARG;
    if ((size_t)step >= states_.size()) states_.resize((size_t)step + 1);
    states_[(size_t)step] = f;
    return 0;
  }
  PetscErrorCode getForwardSolution(PetscInt step, std::vector<double>& f, void*) override {
    if (step < 0 || (size_t)step >= states_.size() || states_[(size_t)step].empty()) return LBMX_ERR_STATE;
    f = states_[(size_t)step];
    return 0;
  }
  PetscErrorCode reset(void*) override {
    states_.clear();
    return 0;
  }

 protected:
  PetscErrorCode advance(void*) override { return solver_.collideAndStream(); }
  PetscErrorCode setSolverState(PetscInt step, const std::vector<double>& f, void*) override {
    return solver_.setDistributions(f, step);
  }
  PetscErrorCode copyFromSolverState(std::vector<double>& f, void*) override { return solver_.getDistributions(&f); }
  LBSolver& solver_;

 private:
  std::vector<std::vector<double>> states_;
};

inline PetscErrorCode LBSolver::makeSourceAdjoint(AdjointLBSolver* adj) {
  adj->solver_ = this;
  return 0;
}

// ---- design pipeline (src/topOpt/densityFilter2d.cc, projectionFilter2d.cc,
//      doubleProjectionFilter2d.cc): full-grid fields; the filter domain is a box ----
class Filter {
 public:
  Filter(const Filter&) = delete;
  Filter& operator=(const Filter&) = delete;
  virtual ~Filter() {
    if (h_) lbmx_filter_destroy(h_);
  }
  // filterDesign(design, filtered) on full-grid vectors
  PetscErrorCode filterDesign(const std::vector<double>& design, std::vector<double>& filtered) {
    filtered.resize(design.size());
    return lbmx_filter_design(h_, design.data(), design.size(), filtered.data(), filtered.size());
  }
  // filterSensitivities(sens, constraintSens, numConstraints) in place
  PetscErrorCode filterSensitivities(std::vector<double>& sens, std::vector<double>* constraintSens,
                                     PetscInt numConstraints) {
    std::vector<double> out(sens.size());
    LBM_CHKERRQ(lbmx_filter_sensitivities(h_, sens.data(), sens.size(), out.data(), out.size()));
    sens.swap(out);
    for (PetscInt c = 0; c < numConstraints; ++c) {
      std::vector<double> o(constraintSens[c].size());
      LBM_CHKERRQ(lbmx_filter_sensitivities(h_, constraintSens[c].data(), constraintSens[c].size(), o.data(),
                                            o.size()));
      constraintSens[c].swap(o);
    }
    return 0;
  }
  // the device filter handle (for lbmx_filter_design_to_solver and the solver-side sensitivities)
  lbmx_filter* handle() const { return h_; }

 protected:
  Filter() = default;
  PetscErrorCode build(const LBSolver& s, const lbmx_box& domain, int nstages, const int* radius, const int* project,
                       const double* beta, const double* eta) {
    lbmx_filter_desc d{};
    d.nx = s.nx();
    d.ny = s.ny();
    d.nz = s.nz();
    d.design_box = lbmx_box{0, s.nx() - 1, 0, s.ny() - 1, 0, s.nz() - 1};
    d.filter_box = domain;
    d.nstages = nstages;
    for (int k = 0; k < nstages; ++k) {
      d.radius[k] = radius[k];
      d.project[k] = project[k];
      d.beta[k] = beta[k];
      d.eta[k] = eta[k];
    }
    return lbmx_filter_create(&d, &h_);
  }
  lbmx_filter* h_ = nullptr;
};

class DensityFilter2d : public Filter {  // densityFilter2d.hh
 public:
  DensityFilter2d(PetscInt radius, const Box2d& domain, const LBSolver& s) {
    const int r[1] = {radius}, p[1] = {0};
    const double b[1] = {1.0}, e[1] = {0.5};
    status = build(s, domain.raw(), 1, r, p, b, e);
  }
  PetscErrorCode status = 0;
};

class ProjectionFilter2d : public Filter {  // projectionFilter2d.hh:113-131 (beta 1, eta 0.5)
 public:
  ProjectionFilter2d(PetscInt radius, const Box2d& domain, const LBSolver& s) {
    const int r[1] = {radius}, p[1] = {1};
    const double b[1] = {1.0}, e[1] = {0.5};
    status = build(s, domain.raw(), 1, r, p, b, e);
  }
  void setEta(PetscScalar eta) { eta_ = eta, lbmx_filter_set_projection(h_, 0, beta_, eta_); }
  void setBeta(PetscScalar beta) { beta_ = beta, lbmx_filter_set_projection(h_, 0, beta_, eta_); }
  void increaseBeta(PetscScalar inc) { beta_ *= inc, lbmx_filter_increase_beta(h_, inc); }
  PetscErrorCode status = 0;

 private:
  PetscScalar beta_ = 1.0, eta_ = 0.5;
};

class DoubleProjectionFilter2d : public Filter {  // doubleProjectionFilter2d.hh:64-86
 public:
  DoubleProjectionFilter2d(PetscInt r1, PetscInt r2, const Box2d& domain, const LBSolver& s) {
    const int r[2] = {r1, r2}, p[2] = {1, 1};
    const double b[2] = {1.0, 1.0}, e[2] = {0.5, 0.5};
    status = build(s, domain.raw(), 2, r, p, b, e);
  }
  void setFirstEta(PetscScalar v) { eta_[0] = v, lbmx_filter_set_projection(h_, 0, beta_[0], eta_[0]); }
  void setSecondEta(PetscScalar v) { eta_[1] = v, lbmx_filter_set_projection(h_, 1, beta_[1], eta_[1]); }
  void setFirstBeta(PetscScalar v) { beta_[0] = v, lbmx_filter_set_projection(h_, 0, beta_[0], eta_[0]); }
  void setSecondBeta(PetscScalar v) { beta_[1] = v, lbmx_filter_set_projection(h_, 1, beta_[1], eta_[1]); }
  void increaseBeta(PetscScalar inc) {
    beta_[0] *= inc, beta_[1] *= inc;
    lbmx_filter_increase_beta(h_, inc);
  }
  PetscErrorCode status = 0;

 private:
  PetscScalar beta_[2] = {1.0, 1.0}, eta_[2] = {0.5, 0.5};
};

// FullToDesignMapping2d (src/topOpt/fullToDesignMapping2d.cc:3-151) on host vectors
class FullToDesignMapping2d {
 public:
  FullToDesignMapping2d(const Box2d& designBox, const LBSolver& s) : box_(designBox), nx_(s.nx()) {}
  PetscErrorCode createDesignDomainVec(std::vector<double>* v) const {
    v->assign((size_t)box_.getNx() * box_.getNy(), 0.0);
    return 0;
  }
  PetscErrorCode mapDesignToFull(const std::vector<double>& design, std::vector<double>& full) const {
    for (PetscInt j = 0; j < box_.getNy(); ++j)
      for (PetscInt i = 0; i < box_.getNx(); ++i)
        full[(size_t)(j + box_.yRange.getBeginId()) * nx_ + i + box_.xRange.getBeginId()] =
            design[(size_t)j * box_.getNx() + i];
    return 0;
  }
  PetscErrorCode mapFullToDesign(const std::vector<double>& full, std::vector<double>& design) const {
    design.resize((size_t)box_.getNx() * box_.getNy());
    for (PetscInt j = 0; j < box_.getNy(); ++j)
      for (PetscInt i = 0; i < box_.getNx(); ++i)
        design[(size_t)j * box_.getNx() + i] =
            full[(size_t)(j + box_.yRange.getBeginId()) * nx_ + i + box_.xRange.getBeginId()];
    return 0;
  }

 private:
  Box2d box_;
  PetscInt nx_;
};

// ---- the forward + adjoint pass (src/topOpt/unsteadyLBTopOpt2d.hh:15-39) ----
class UnsteadyLBTopOpt {
 public:
  UnsteadyLBTopOpt(LBSolver& solver, AdjointLBSolver& adj, const LBMacroFunctional& obj, PetscInt numTimesteps,
                   PetscScalar timeConst, Filter* filter = nullptr)
      : solver_(solver), adj_(adj), obj_(obj), nt_(numTimesteps), tc_(timeConst), filter_(filter) {}
  // numCheckpoints = -1 stores every state (ForwardSolutionStorage); otherwise Wang
  // checkpointing with that many slots, placed online (LBDynamicCheckpointing,
  // src/topOpt/LBCheckpointing.cc:37-40) or, with staticPlacement, planned for the known
  // horizon (LBStaticCheckpointing, src/topOpt/staticCheckpointing.cc:3-55)
  PetscErrorCode allocateMemoryForForwardSolution(PetscInt numCheckpoints = -1, bool staticPlacement = false) {
    PetscErrorCode e = lbmx_checkpoint_config(solver_.handle(), numCheckpoints < 0 ? 0 : numCheckpoints);
    if (e) return e;
    return lbmx_checkpoint_mode(solver_.handle(), staticPlacement ? LBMX_CHECKPOINT_STATIC : LBMX_CHECKPOINT_DYNAMIC);
  }
  // Any placement of lbmx_checkpoint_mode_t, LBMX_CHECKPOINT_BINOMIAL included (binomial
  // checkpointing for the known horizon, which the reference does not have).
  PetscErrorCode allocateMemoryForForwardSolution(PetscInt numCheckpoints, lbmx_checkpoint_mode_t placement) {
    PetscErrorCode e = lbmx_checkpoint_config(solver_.handle(), numCheckpoints < 0 ? 0 : numCheckpoints);
    if (e) return e;
    return lbmx_checkpoint_mode(solver_.handle(), placement);
  }
  // J only (UnsteadyLBTopOpt2d::computeObjective)
  PetscErrorCode computeObjective() {
    LBM_CHKERRQ(obj_.install(solver_.handle(), tc_));
    LBM_CHKERRQ(solver_.initializeAtEquilibrium());
    return lbmx_forward(solver_.handle(), nt_, &J_);
  }
  // forward with checkpoints, reverse loop with recomputation, dJ/dgamma (and its filtered
  // form when a filter is attached: LBTopOpt2d's filter->filterSensitivities)
  PetscErrorCode forwardAndAdjoint() {
    LBM_CHKERRQ(obj_.install(solver_.handle(), tc_));
    const double u0[3] = {0, 0, 0};
    LBM_CHKERRQ(solver_.makeSourceAdjoint(&adj_));
    return lbmx_gradient(solver_.handle(), nt_, 1.0, u0, 0.0, &J_, nullptr, 0);
  }
  PetscScalar getObjective() const { return J_; }
  // dJ/dgamma mapped through the attached filter, on the device (full grid)
  PetscErrorCode getFilteredSensitivities(std::vector<double>* sens) const {
    if (!filter_) return adj_.getSensitivities(sens);
    sens->resize((size_t)solver_.nodes());
    return lbmx_filter_solver_sensitivities(filter_->handle(), solver_.handle(), sens->data(), sens->size());
  }

 private:
  LBSolver& solver_;
  AdjointLBSolver& adj_;
  LBMacroFunctional obj_;
  PetscInt nt_;
  PetscScalar tc_;
  Filter* filter_;
  PetscScalar J_ = 0.0;
};

// ---- method of moving asymptotes (src/topOpt/MMA.hh) ----
class MMA {
 public:
  MMA(PetscInt n, PetscInt m, const std::vector<double>& /*x*/) : n_(n), m_(m) { status = lbmx_mma_create(n, m, &h_); }
  MMA(const MMA&) = delete;
  MMA& operator=(const MMA&) = delete;
  ~MMA() {
    if (h_) lbmx_mma_destroy(h_);
  }
  void SetRobustConstants(const PetscScalar* a) { lbmx_mma_set_robust_constants(h_, a); }
  PetscErrorCode SetAsymptotes(PetscScalar init, PetscScalar dec, PetscScalar inc) {
    return lbmx_mma_set_asymptotes(h_, init, dec, inc);
  }
  PetscErrorCode ConstraintModification(bool on) {
    conMod_ = on ? 1 : 0;
    return lbmx_mma_set_options(h_, conMod_, asymType_);
  }
  PetscErrorCode SetRobustAsymptotesType(PetscInt v) {
    asymType_ = v;
    return lbmx_mma_set_options(h_, conMod_, asymType_);
  }
  // Update(design, sensitivities, constraintValues, constraintSensitivities, movlim)
  PetscErrorCode Update(std::vector<double>& x, const std::vector<double>& dfdx, const PetscScalar* gx,
                        const std::vector<double>* dgdx, PetscScalar movlim) {
    std::vector<const double*> rows((size_t)m_);
    for (PetscInt j = 0; j < m_; ++j) rows[j] = dgdx[j].data();
    return lbmx_mma_update(h_, x.data(), dfdx.data(), gx, rows.data(), movlim);
  }
  PetscErrorCode DesignChange(const std::vector<double>& x, PetscScalar& ch) {
    return lbmx_mma_design_change(h_, x.data(), &ch);
  }
  PetscErrorCode status = 0;

 private:
  PetscInt n_, m_;
  PetscInt conMod_ = 0, asymType_ = 0;
  lbmx_mma* h_ = nullptr;
};

}  // namespace lbm_topopt
