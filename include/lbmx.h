/*
 * lbmx.h -- C ABI of the MI355X lattice Boltzmann forward/adjoint engine.
 *
 * Every entry point returns an lbmx_status (0 = success), the counterpart of the
 * reference's PetscErrorCode convention (CHKERRQ propagation). Arrays crossing the
 * boundary are plain host pointers in the reference's single-rank PETSc DMDA order:
 *   design / sensitivity : natural order  (k*ny + j)*nx + i           (1 value per node)
 *   distributions        : array-of-structures [k][j][i][q]          (Q values per node)
 * so a PETSc adapter maps Vec <-> pointer with VecGetArrayRead/VecGetArray and nothing else.
 *
 * A handle is thread-compatible, not thread-safe: calls on one handle must not overlap.
 */
#ifndef LBMX_H
#define LBMX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBMX_VERSION_MAJOR 0
#define LBMX_VERSION_MINOR 1

/* ---- status codes (mirroring the PETSc error classes the reference raises) ---- */
typedef int lbmx_status;
#define LBMX_OK               0
#define LBMX_ERR_ARG          62   /* PETSC_ERR_ARG_OUTOFRANGE / bad argument            */
#define LBMX_ERR_SUP          56   /* PETSC_ERR_SUP: orientation/type not implemented    */
#define LBMX_ERR_MEM          55   /* PETSC_ERR_MEM: device allocation failed            */
#define LBMX_ERR_STATE        73   /* PETSC_ERR_ORDER: call sequence violated            */
#define LBMX_ERR_DEVICE       76   /* HIP runtime error (message in lbmx_last_error)     */

/* ---- lattices (src/latticeBoltzmann/lattices2d.hh:16-84, lattices3d.hh:34-53) ---- */
typedef enum {
  LBMX_D2Q9 = 0,        /* flow only, Q = 9                                   */
  LBMX_D3Q19 = 1,       /* flow only, Q = 19                                  */
  LBMX_D2Q9_D2Q5 = 2    /* thermal double distribution, 9 + 5 = 14 per node  */
} lbmx_lattice;

/* ---- flow collision operators ---- */
typedef enum {
  LBMX_BGK = 0,           /* IncompressibleBGK2d<D2Q9> (src/dynamics/incompressibleBGK2d.hh:94-153);
                             3D: incompressible BGK with IncompressibleFeq3d                 */
  LBMX_MRT = 1,           /* IncompressibleMRT2d<D2Q9> (src/dynamics/incompressibleMRT2d.hh:18-98),
                             IncompressibleMRT3d<D3Q19> (src/dynamics/MRT3d.hh:28-200)       */
  LBMX_CASCADED = 2,      /* IncompressibleCascaded2d<D2Q9> (src/dynamics/cascaded2d.hh:15-97) */
  LBMX_MRT_STANDARD = 3,  /* StandardMRT3d<D3Q19> (src/dynamics/MRT3d.hh:206-383), forward only */
  LBMX_BGK_STANDARD = 4   /* StandardBGK2d (src/dynamics/incompressibleBGK2d.hh:52-88),
                             StandardBGK3d (src/dynamics/BGK3d.hh:10-38), forward only       */
} lbmx_collision;

/* ---- how the design field gamma (1 = fluid, 0 = solid) enters the collision ---- */
typedef enum {
  LBMX_DESIGN_NONE = 0,
  LBMX_DESIGN_PBB = 1,      /* PartialBouncebackCollision<Base, BPInterpolation>
                               (src/dynamics/partialBouncebackCollision.hh:53-65):
                               imult = 0.5 * alpha_max * BP_q(gamma); alpha_max = 1 is the
                               reference operator exactly                                   */
  LBMX_DESIGN_BRINKMAN = 2, /* Brinkman drag F = -alpha_max * BP_q(gamma) * u with Guo forcing
                               (no reference counterpart for BGK/MRT3d; defined by this engine) */
  LBMX_DESIGN_DARCY = 3     /* DarcyMRT2d<D2Q9,BP> (src/dynamics/darcyMRT2d.hh:33-59), MRT only */
} lbmx_design;

typedef enum {
  LBMX_FP64 = 0,            /* plain form: FP64 storage, FP64 collision                      */
  LBMX_FP32 = 1,            /* FP32 storage, FP32 collision                                  */
  LBMX_FP16_SHIFTED = 2,    /* FP16 storage of (f_q - w_q)/w_q, FP32 collision, FP32 adjoint  */
  LBMX_BF16_SHIFTED = 3     /* BF16 storage of (f_q - w_q)/w_q, FP32 collision, FP32 adjoint  */
} lbmx_precision;

typedef struct {
  int nx, ny, nz;             /* nz = 1 for 2D lattices                                     */
  int lattice;                /* lbmx_lattice                                               */
  int collision;              /* lbmx_collision                                             */
  int design;                 /* lbmx_design                                                */
  int precision;              /* lbmx_precision                                             */
  int reserved0;
  double omega;               /* flow relaxation rate (1/tau)                               */
  double interp_q;            /* BPInterpolation penalisation q (src/core/interpolationFunctions.hh:7-28) */
  double alpha_max;           /* design coupling strength                                   */
  double body_force[3];       /* DarcyMRT2d constant body force (cfx, cfy)                  */
  /* thermal part (LBMX_D2Q9_D2Q5): ThermalObstacleCollision<ThermalMRT2d<D2Q5_TMRT>, Ramp> */
  double thermal_diffusivity; /* reference diffusivity D_ref (src/dynamics/thermalObstacleCollision.hh:47-53) */
  double ramp_p;              /* RampInterpolation penal (src/core/interpolationFunctions.hh:30-54) */
  double ramp_ck;             /* RampInterpolation Ck (k_solid/k_fluid = 1/Ck)               */
} lbmx_desc;

/* Inclusive index box, the reference's Box2d/Box3d (src/core/globalDefinitions.hh:25-71). */
typedef struct { int x0, x1, y0, y1, z0, z1; } lbmx_box;

/* ---- boundary conditions ---- */
typedef enum {
  LBMX_BC_BOUNCEBACK = 0,         /* OnGridBounceback2d/3d (src/boundaryConditions/bounceback2d.cc:7-193,
                                     bounceback3d.cc:8-511), all orientations                    */
  LBMX_BC_ZOUHE_VELOCITY = 1,     /* {Incompressible,Standard}ZouHeVelocity{West,North,East}2d
                                     (src/boundaryConditions/ZouHe2d.hh:23-195)                  */
  LBMX_BC_ZOUHE_PRESSURE = 2,     /* {Incompressible,Standard}ZouHePressure{East,West}2d
                                     (src/boundaryConditions/ZouHe2d.hh:240-390)                 */
  LBMX_BC_ZOUHE_NEUMANN = 3,      /* IncompressibleZouHeNeumann{East,West}2d<D2Q9>
                                     (src/boundaryConditions/ZouHe2d.hh:440-527; adjoint
                                     adjointZouHe2d.cc:114-172): u_x extrapolated from the two
                                     nodes inwards, u_y = 0; east/west, grid nx >= 4             */
  LBMX_BC_HECHT_VELOCITY = 4,     /* {Incompressible,Standard}HechtVelocityWest3d
                                     (src/boundaryConditions/Hecht3d.hh:26-110)                  */
  LBMX_BC_HECHT_PRESSURE = 5,     /* {Incompressible,Standard}HechtPressureEast3d
                                     (src/boundaryConditions/Hecht3d.hh:120-238)                 */
  LBMX_BC_THERMAL_BOUNCEBACK = 6, /* ThermalBounceback2d<D2Q9,D2Q5> (adiabatic),
                                     (src/boundaryConditions/thermalBounceback2d.hh:260-300)     */
  LBMX_BC_THERMAL_DIRICHLET = 7,  /* ThermalZouHeWest2d<D2Q9,D2Q5> (thermalZouHe2d.hh:62-93); the
                                     N/S/E orientations are the same construction, defined here  */
  LBMX_BC_THERMAL_NEUMANN = 8     /* ThermalZouHeNeumann{East,West}2d<D2Q9,D2Q5>
                                     (thermalZouHe2d.hh:229-300; adjoint adjointThermalZouHe2d.hh:
                                     161-227): T extrapolated from the two nodes inwards          */
} lbmx_bc_type;

typedef enum {
  LBMX_PROFILE_UNIFORM = 0,       /* value                                                      */
  LBMX_PROFILE_PARABOLIC = 1,     /* value*(R^2-(s-c0)^2)/R^2, s = tangential index (2D), as the
                                     InflowVelocity functors of applications/diffuser/diffuser.cc:21-31 */
  LBMX_PROFILE_PARABOLIC_DISK = 2,/* value*(1-r^2/R^2) inside r<=R, else 0; r^2=(s-c0)^2+(t-c1)^2
                                     (DevFlowYZ, applications/Diffuser3d/diffuser3d.cc:22-46)   */
  LBMX_PROFILE_PARABOLIC_DUCT = 3,/* value*(1-((s-c0)/R)^2)*(1-((t-c1)/R1)^2)                  */
  LBMX_PROFILE_TABLE = 4          /* value*a(i,j,k): per-node values handed to
                                     lbmx_add_boundary_table                                    */
} lbmx_profile_kind;

typedef enum {
  LBMX_TIME_CONSTANT = 0,
  LBMX_TIME_SINE = 1,             /* factor = t_offset + t_amp*sin(2*pi*t/t_period + t_phase),
                                     as applications/RobustPump2d/robustPump2d.cc:14-19           */
  LBMX_TIME_TABLE = 2             /* factor = b(t) from a table handed to lbmx_add_boundary_table */
} lbmx_time_kind;

typedef struct {
  int type;           /* lbmx_bc_type                                                      */
  int profile;        /* lbmx_profile_kind                                                 */
  int time_mode;      /* lbmx_time_kind                                                    */
  int reserved0;
  double value;       /* u_max (velocity), rho0 (pressure), T0 (thermal Dirichlet)         */
  double dir[3];      /* velocity direction multipliers (ux, uy, uz)                       */
  double center[2];   /* profile centre in the tangential index coordinates (s, t)         */
  double radius[2];   /* profile radii                                                     */
  double t_period, t_amp, t_offset, t_phase;
} lbmx_bc_desc;

/* ---- objective functionals: J = sum_t timeConst * sum_terms ---- */
typedef enum {
  LBMX_OBJ_NONE = 0,
  LBMX_OBJ_PRESSURE_DROP = 1,     /* PressureDrop2d (src/functionals/pressureDrop2d.hh:14-52):
                                     mean(rho, box_a) - mean(rho, box_b); also used in 3D       */
  LBMX_OBJ_OUTFLOW = 2,           /* OutflowEast2d (src/functionals/outflow2d.hh:15-54): -mean(u_x, box_a) */
  LBMX_OBJ_DISSIPATED_POWER = 3,  /* total-pressure flux through the ports, defined here:
                                     sum_box_a (p+|u|^2/2) u_x - sum_box_b (p+|u|^2/2) u_x      */
  LBMX_OBJ_HEAT_FLUX = 4,         /* ThermalOutflowEast2d (src/functionals/temperatureFlow2d.hh:18-54):
                                     -sum_box_a u_x * T                                          */
  LBMX_OBJ_THROUGHFLOW = 5        /* ThroughFlow2d (src/functionals/throughFlow2d.hh:16-88):
                                     -mean(u_x, box_a) - mean(u_x, box_b)                        */
} lbmx_objective_kind;

typedef struct {
  int kind;                 /* lbmx_objective_kind                                       */
  int reserved0;
  double time_const;        /* the reference's timeConst (e.g. 1/numSteps)               */
  lbmx_box box_a;
  lbmx_box box_b;
} lbmx_objective;

/* ---- linear box functionals: the extension point of LBMacroFunctional ----
   LBMacroFunctional (src/functionals/LBMacroFunctional.hh:6-14) lets an application supply any
   functional through evaluate / adjointCollideSource. This engine takes the functionals of the
   form J_t = timeConst * sum_k weight_k * sum_{x in box_k} m_k(x) over up to 4 terms, with the
   per-node quantities below; the adjoint source is their exact gradient. Every shipped
   functional is of this form, e.g. DomainTemperature2d (src/functionals/domainTemperature2d.hh:
   weight 1/area on LBMX_TERM_TEMPERATURE over the domain box). */
typedef enum {
  LBMX_TERM_DENSITY = 0,              /* rho                                                   */
  LBMX_TERM_VELOCITY_X = 1,           /* u_x (= j_x in the incompressible models)              */
  LBMX_TERM_TOTAL_PRESSURE_FLUX = 2,  /* (rho/3 + |u|^2/2) u_x                                 */
  LBMX_TERM_HEAT_FLUX_X = 3,          /* u_x T (thermal lattice)                               */
  LBMX_TERM_TEMPERATURE = 4           /* T = sum of the D2Q5 populations (thermal lattice)     */
} lbmx_term_kind;

typedef struct {
  int kind;                 /* lbmx_term_kind                                            */
  int reserved0;
  double weight;
  lbmx_box box;             /* nodes summed over (inclusive, inside the grid)            */
} lbmx_term;

typedef struct lbmx_solver lbmx_solver;

/* ---- lifetime ---- */
/* IsothermalLBSolver2d/3d::make (src/latticeBoltzmann/isothermalLBSolver2d.hh:20-46),
   ThermalLBSolver2d::make (src/latticeBoltzmann/thermalLBSolver2d.hh:15-110). The design
   field defaults to 1 (fluid), as VecSet(materialGlobal,1.) (src/latticeBoltzmann/NewLBSolver2d.cc:25). */
lbmx_status lbmx_create(const lbmx_desc* desc, lbmx_solver** out);
lbmx_status lbmx_destroy(lbmx_solver* s);
/* Bind a HIP stream (hipStream_t passed as void*); NULL = the legacy default stream. */
lbmx_status lbmx_set_stream(lbmx_solver* s, void* hip_stream);
/* Message of the last failing call on this thread. */
const char* lbmx_last_error(void);

/* ---- setup ---- */
/* NewLBSolver2d::addBoundaryCondition (src/latticeBoltzmann/NewLBSolver2d.cc:106-155),
   LBSolver3d::addBoundaryCondition (src/latticeBoltzmann/LBSolver3d.cc): the box is
   intersected with the 4 edges + 4 corners (2D) or 6 faces + 12 edges + 8 corners (3D)
   of the domain; each piece becomes one boundary loop, applied in insertion order. */
lbmx_status lbmx_add_boundary(lbmx_solver* s, const lbmx_box* box, const lbmx_bc_desc* bc);
/* The same for a boundary value given by any separable functor value * a(i,j,k) * b(t): the
   velocity / pressure / temperature functors the applications hand to their boundary
   descriptors (e.g. InflowVelocity, applications/diffuser/diffuser.cc:21-42; DevFlowYZ,
   applications/Diffuser3d/diffuser3d.cc:11-54; the modulated inlet of
   applications/RobustPump2d/robustPump2d.cc:13-25), tabulated once by the caller.
   profile (may be NULL: bc->profile applies) holds a(i,j,k) for every node of *box in natural
   order (i fastest), ncomp values per node: 1 = one factor (velocity: times bc->dir), or D (2 or
   3) = one factor per velocity component (u_a = value * dir[a] * a_a(i,j,k) * b(t)); it is copied
   to the device. time_factors (may be NULL: bc->time_mode applies) holds b(t) for the steps
   t = 0 .. nt-1 at which the condition is applied (the timestep argument of the reference's
   functors, the step the stream just reached); a forward or gradient call that would apply the
   condition at a step t >= nt fails with LBMX_ERR_ARG. bc->profile and bc->time_mode are
   ignored where a table is given. */
lbmx_status lbmx_add_boundary_table(lbmx_solver* s, const lbmx_box* box, const lbmx_bc_desc* bc, const double* profile,
                                    size_t nprofile, int ncomp, const double* time_factors, size_t nt);
/* LBSolver::setFieldFromDesignVec (src/latticeBoltzmann/LBSolver.cc:333-348). */
lbmx_status lbmx_set_design(lbmx_solver* s, const double* gamma, size_t n);
/* LBSolver::createDesignVec (src/latticeBoltzmann/LBSolver.cc:318-331): copy out the field. */
lbmx_status lbmx_get_design(lbmx_solver* s, double* gamma, size_t n);
/* The functional later used by lbmx_forward/lbmx_gradient. */
lbmx_status lbmx_set_objective(lbmx_solver* s, const lbmx_objective* obj);
/* The same from explicit terms (LBMacroFunctional, src/functionals/LBMacroFunctional.hh:6-14,
   for the linear box functionals described at lbmx_term): 0 < nterms <= 4; thermal kinds need
   the thermal lattice (LBMX_ERR_SUP otherwise). Replaces the functional set before. */
lbmx_status lbmx_set_objective_terms(lbmx_solver* s, const lbmx_term* terms, int nterms, double time_const);

/* ---- forward ---- */
/* LBSolver::initializeAtEquilibrium (src/latticeBoltzmann/LBSolver.cc:234-249) with uniform
   macros (rho0, u0) (uniformMacroInitialization), thermal part at temperature T0; t := 0. */
lbmx_status lbmx_init_equilibrium(lbmx_solver* s, double rho0, const double u0[3], double T0);
/* Per-node initial macros: macros[(k*ny+j)*nx+i][0..3] = (rho, ux, uy, uz) (getInitialMacroArray);
   thermal: T in [4]; stride = 5 doubles per node. */
lbmx_status lbmx_init_equilibrium_field(lbmx_solver* s, const double* macros, size_t n);
/* nsteps x LBSolver::collideAndStream (src/latticeBoltzmann/LBSolver.cc:108-115), with the
   objective of each new state accumulated as computeMacroFunctional does after the next
   collision; *J (may be NULL) receives the sum over these steps. */
lbmx_status lbmx_forward(lbmx_solver* s, int nsteps, double* J);
/* distributionsGlobal: the pre-collision state after the last stream+BC, AoS [node][q]. */
lbmx_status lbmx_get_state(lbmx_solver* s, double* f, size_t n);
/* LBSolver::setDistributions + setCurrentTimestep (src/latticeBoltzmann/LBSolver.cc:65-97). */
lbmx_status lbmx_set_state(lbmx_solver* s, const double* f, size_t n, int timestep);
/* Macros of the current state: (rho, ux, uy, uz[, T]) per node, 5 doubles per node. */
lbmx_status lbmx_get_macros(lbmx_solver* s, double* macros, size_t n);
lbmx_status lbmx_get_timestep(lbmx_solver* s, int* t);

/* ---- adjoint (AdjointLBSolver, src/adjointLatticeBoltzmann/adjointLBSolver.hh:11-46) ---- */
/* Wang dynamic checkpointing with nslots HBM slots (LBDynamicCheckpointing,
   src/topOpt/LBCheckpointing.cc:37-65); nslots <= 0 stores every step (NaiveCheckpointing). */
lbmx_status lbmx_checkpoint_config(lbmx_solver* s, int nslots);
/* Placement of the nslots checkpoints: LBMX_CHECKPOINT_DYNAMIC (default) places them online
   as the forward pass runs (LBDynamicCheckpointing, src/topOpt/LBCheckpointing.cc:37-65);
   LBMX_CHECKPOINT_STATIC keeps the checkpoints the same level-based placement gives when it
   plans the known horizon at once (LBStaticCheckpointing / StaticCheckpointing,
   src/topOpt/staticCheckpointing.cc:3-55, used by applications/RobustPump2d/robustPump2d.cc:166).
   LBMX_CHECKPOINT_BINOMIAL has no counterpart in the reference: binomial checkpointing for the
   known horizon (Griewank & Walther, ACM TOMS 26 (2000) 19-45, "revolve"), the fewest forward
   steps any placement of nslots checkpoints allows; for callers whose time to gradient is
   dominated by recomputation.
   The reverse pass and the results are the same; only the recompute count differs. */
typedef enum {
  LBMX_CHECKPOINT_DYNAMIC = 0,
  LBMX_CHECKPOINT_STATIC = 1,
  LBMX_CHECKPOINT_BINOMIAL = 2
} lbmx_checkpoint_mode_t;
lbmx_status lbmx_checkpoint_mode(lbmx_solver* s, int mode);
/* One design-gradient evaluation, the loop of applications/RobustPump2d/robustPump2d.cc:233-277:
   initializeAtEquilibrium(rho0,u0,T0), nsteps recorded forward steps, resetAdjoints, then
   nsteps reverse steps (adjointCollide, computeCollideSource, adjointStream) with checkpoint
   recomputation. J (may be NULL) = objective; sens (may be NULL, n = nodes) = dJ/dgamma. */
lbmx_status lbmx_gradient(lbmx_solver* s, int nsteps, double rho0, const double u0[3], double T0,
                          double* J, double* sens, size_t n);
/* AdjointLBSolver::getSensitivities after the last lbmx_gradient or lbmx_adjoint_step (the
   sensitivityGlobal Vec). */
lbmx_status lbmx_get_sensitivity(lbmx_solver* s, double* sens, size_t n);
/* The adjoint distributions after the last lbmx_gradient or lbmx_adjoint_step (adjointGlobal),
   AoS [node][q]. */
lbmx_status lbmx_get_adjoint(lbmx_solver* s, double* adj, size_t n);

/* The reverse loop step by step, for callers that drive the forward states themselves (their
   own checkpointing, a CheckpointingBase of src/topOpt/checkpointingBase.hh:7-18).
   AdjointLBSolver::resetAdjoints + setCurrentTimestep (src/adjointLatticeBoltzmann/adjointLBSolver.cc:138-145):
   zeroes the adjoint distributions and dJ/dgamma; the next reverse step is `timestep`. */
lbmx_status lbmx_adjoint_reset(lbmx_solver* s, int timestep);
/* AdjointLBSolver::setCurrentTimestep (src/adjointLatticeBoltzmann/adjointLBSolver.cc:147-151):
   the next reverse step, after lbmx_adjoint_reset. */
lbmx_status lbmx_adjoint_set_timestep(lbmx_solver* s, int timestep);
/* One reverse step t: adjointCollide (src/adjointLatticeBoltzmann/adjointLBSolver.cc:28-48),
   computeCollideSource of the functional set by lbmx_set_objective when with_source != 0
   (:83-102) and adjointStream (:50-81), in the order of applications/RobustPump2d/robustPump2d.cc:268-276;
   dJ/dgamma accumulates; then t := t - 1. The forward state must be the state of step t:
   after lbmx_forward reached t, or lbmx_set_state(..., t) (getForwardSolution); else
   LBMX_ERR_STATE. Step 0 (the initial state) is never reversed. */
lbmx_status lbmx_adjoint_step(lbmx_solver* s, int with_source);

/* Device-resident hand-off (HIP device pointers, FP64, natural node order). Both calls return
   with the device work done. setFieldFromDesignVec from a device array: */
lbmx_status lbmx_set_design_device(lbmx_solver* s, const double* gamma_dev, size_t n);
/* getSensitivities as a borrowed handle (src/adjointLatticeBoltzmann/adjointLBSolver.cc:153-157):
   the solver's own dJ/dgamma array, valid until the next gradient, reset or destroy. */
lbmx_status lbmx_get_sensitivity_device(lbmx_solver* s, const double** sens_dev, size_t* n);

/* ---- instrumentation ---- */
typedef struct {
  double fwd_kernel_ms;       /* device time of forward sweeps in the last call          */
  double adj_kernel_ms;       /* device time of adjoint sweeps in the last call          */
  double forward_phase_ms;    /* wall time of the recorded forward phase                 */
  double reverse_phase_ms;    /* wall time of the reverse phase incl. recomputation      */
  /* counters below are cumulative since lbmx_create (difference two reads for one call) */
  long long fwd_sweeps;       /* forward sweeps launched (incl. recomputation)           */
  long long adj_sweeps;       /* adjoint sweeps launched                                 */
  long long checkpoint_copies;/* state copies made for checkpoints (0: slots are swapped) */
  long long recompute_sweeps; /* forward sweeps re-run in reverse phases                 */
} lbmx_stats;
lbmx_status lbmx_get_stats(lbmx_solver* s, lbmx_stats* st);
/* Times nsweeps back-to-back sweeps with HIP events on the handle's stream. kind 0: forward
   sweeps from the current state (the state advances, objective accumulation on); kind 1:
   adjoint sweeps reading the current state as p_{t-1} with ping-ponged adjoint buffers.
   *ms = total device time of the sweeps. */
lbmx_status lbmx_time_sweeps(lbmx_solver* s, int kind, int nsweeps, double* ms);
/* Measured HBM ceiling: iters launches of a streaming copy of `bytes` (read + write);
   *gbs = (2 * bytes * iters) / device time. Also the calibration kernel for byte counters. */
lbmx_status lbmx_stream_copy(size_t bytes, int iters, double* gbs);
/* Diagnostic: GB/s (read + write) of a copy with the sweep's access pattern, planes = 9, 14 or
   19 FP64 planes of `nodes` values read and written at the same offset, one node per thread.
   No counterpart in the reference. */
lbmx_status lbmx_stream_copy_planes(int planes, size_t nodes, int iters, double* gbs);
/* Diagnostic: the same for 16-bit planes (the 16-bit storage forms), two nodes per thread, each
   plane's pair read and written as one 32-bit word. No counterpart in the reference. */
lbmx_status lbmx_stream_copy_planes16(int planes, size_t nodes, int iters, double* gbs);
/* Number of visible HIP devices (0 when the runtime finds none); no handle needed. */
lbmx_status lbmx_device_count(int* count);
/* Bytes of device memory held by the handle (state, adjoint, checkpoints). */
lbmx_status lbmx_device_bytes(lbmx_solver* s, size_t* bytes);

/* ---- host-only diagnostics (no device access) ---- */
/* Boundary pieces after box splitting, in application order: per piece 10 ints
   (type, nx, ny, nz, x0, x1, y0, y1, z0, z1). *count receives the number of pieces. */
lbmx_status lbmx_describe_boundaries(lbmx_solver* s, int* out, int max_pieces, int* count);
/* Replays the checkpoint schedule of lbmx_gradient for (nslots, nsteps) on labelled host
   buffers and verifies that every reverse step receives the right state. *advances = forward
   steps run in the reverse phase (recomputation), *max_slots = slots in use at most. */
lbmx_status lbmx_checkpoint_simulate(int nslots, int nsteps, long long* advances, int* max_slots);
/* The same for any placement (lbmx_checkpoint_mode). */
lbmx_status lbmx_checkpoint_simulate_mode(int nslots, int nsteps, int mode, long long* advances, int* max_slots);

/* ---- batched robust-ensemble form ----
   Several solver handles of one configuration (same lattice, operator, design coupling and
   precision; any grids, boundary pieces, functionals and designs) advanced in lockstep: every
   sweep of the design-gradient pass is one kernel launch over all members, and all members
   share one checkpoint schedule. Replaces the MPISplit communicators over which the reference
   runs robust realisations and design candidates side by side (src/core/MPISplit.cc:150-234,
   applications/RobustPump2d/robustPump2d.cc:47,146). Built for D2Q9 incompressible BGK, MRT and
   cascaded, D3Q19 incompressible BGK and MRT and the thermal D2Q9 + D2Q5 lattice (BGK or MRT
   flow part), in every precision form. Members stay owned by their handles and must outlive
   the batch; calls on members and on the batch must not overlap. */
typedef struct lbmx_batch lbmx_batch;
lbmx_status lbmx_batch_create(lbmx_solver* const* members, int count, lbmx_batch** out);
lbmx_status lbmx_batch_destroy(lbmx_batch* b);
/* checkpoint slots of the shared schedule (each slot holds every member's state) */
lbmx_status lbmx_batch_checkpoint_config(lbmx_batch* b, int nslots);
/* placement of the shared schedule's checkpoints (lbmx_checkpoint_mode) */
lbmx_status lbmx_batch_checkpoint_mode(lbmx_batch* b, int mode);
/* rows of the 64-wide workgroup tile of the batched sweeps: 1, 2 or 4 (default 4), with the
   forward of 2 and 4 rows taking two nodes per thread */
lbmx_status lbmx_batch_set_tile_rows(lbmx_batch* b, int rows);
/* the same with the forward's pairing chosen: paired != 0 takes rows y and y + rows/2 in one
   thread (2 and 4 rows), paired = 0 one node per thread. Results do not depend on the shape. */
lbmx_status lbmx_batch_set_tile(lbmx_batch* b, int rows, int paired);
/* the tile shape in use (set, or chosen by lbmx_batch_autotune); *paired = 0 for 1 row */
lbmx_status lbmx_batch_get_tile(lbmx_batch* b, int* rows, int* paired);
/* enable != 0: the first lbmx_batch_gradient captures its whole launch sequence into a HIP
   graph and later calls with the same steps, slots, tile height and buffers replay it (designs
   may change in between; they are read from the same device arrays). Forward/reverse phase
   times are then not separated in the stats. */
lbmx_status lbmx_batch_set_graph(lbmx_batch* b, int enable);
/* Autotuner: times nsweeps batched forward + adjoint sweeps for each tile shape (height 1, 2 or
   4; the forward of 2 and 4 rows paired or not), keeps the fastest; *rows = chosen height
   (lbmx_batch_get_tile gives the pairing), ms[3] (may be NULL) = ms per forward+adjoint sweep
   pair for heights 1, 2, 4, each with its better pairing. The members' current states, adjoints and dJ/dgamma arrays are discarded:
   until their next gradient, lbmx_get_sensitivity and the other sensitivity reads on a member
   return LBMX_ERR_STATE. */
lbmx_status lbmx_batch_autotune(lbmx_batch* b, int nsweeps, int* rows, double* ms);
/* lbmx_gradient for every member: J[count] (may be NULL) and sens[count] (array of pointers to
   per-member dJ/dgamma arrays in natural order; the array or any entry may be NULL). */
lbmx_status lbmx_batch_gradient(lbmx_batch* b, int nsteps, double rho0, const double u0[3], double T0,
                                double* J, double* const* sens);
lbmx_status lbmx_batch_get_stats(lbmx_batch* b, lbmx_stats* st);
lbmx_status lbmx_batch_device_bytes(lbmx_batch* b, size_t* bytes);

/* ---- design pipeline: the steps either side of gamma and dJ/dgamma ----
   Design-box mapping (FullToDesignMapping2d/3d, src/topOpt/fullToDesignMapping2d.cc:3-151,
   fullToDesignMapping3d.cc), cone density filter (DensityFilter2d/3d,
   src/topOpt/densityFilter2d.cc:3-146, densityFilter3d.cc:60-139), tanh projection
   (ProjectionFilter2d/3d, src/topOpt/projectionFilter2d.cc:27-87, projectionFilter3d.hh) and
   the two-stage chain (DoubleProjectionFilter2d, src/topOpt/doubleProjectionFilter2d.cc:28-48),
   on the device. Fields are FP64 in natural order; the full grid is non-periodic.

   Density filter of one stage with radius R: a row inside filter_box holds the weights
   R - |d| of the in-grid neighbours at offsets d with |d| <= R; every other row is the
   identity; the result is (W x) / (W 1). Its sensitivity map is W^T (s / (W 1)). A projecting
   stage then maps the filtered value x inside filter_box to
   (tanh(beta eta) + tanh(beta (x - eta))) / (tanh(beta eta) + tanh(beta (1 - eta))). */
#define LBMX_FILTER_MAX_RADIUS 16
typedef struct {
  int nx, ny, nz;          /* full lattice grid (nz = 1 in 2D) */
  lbmx_box design_box;     /* optimisation variables: natural order over this box */
  lbmx_box filter_box;     /* rows that are filtered and projected */
  int nstages;             /* 1 (DensityFilter / ProjectionFilter) or 2 (DoubleProjectionFilter) */
  int radius[2];           /* cone radius per stage, 1..LBMX_FILTER_MAX_RADIUS */
  int project[2];          /* nonzero: the stage projects after its density filter */
  double beta[2], eta[2];  /* projection sharpness and threshold per stage */
} lbmx_filter_desc;
typedef struct lbmx_filter lbmx_filter;
/* DensityFilter2d::DensityFilter2d + setUp, ProjectionFilter2d::ProjectionFilter2d,
   DoubleProjectionFilter2d::DoubleProjectionFilter2d, FullToDesignMapping2d::FullToDesignMapping2d.
   Allocates on the current device; the background field starts at 1 (fluid). */
lbmx_status lbmx_filter_create(const lbmx_filter_desc* desc, lbmx_filter** out);
lbmx_status lbmx_filter_destroy(lbmx_filter* f);
/* ProjectionFilter2d::setBeta/setEta (projectionFilter2d.hh:121-122),
   DoubleProjectionFilter2d::setFirst/SecondBeta/Eta (doubleProjectionFilter2d.hh:73-76) */
lbmx_status lbmx_filter_set_projection(lbmx_filter* f, int stage, double beta, double eta);
/* VariableProjectionFilter2d::setEtaValuesFromVector (src/topOpt/variableProjectionFilter2d.cc:123-128),
   VariableDoubleProjectionFilter2d::setSecondEtaValuesFromVector
   (src/topOpt/variableDoubleProjectionFilter2d.cc:28-36): the stage projects with a per-node
   threshold eta[i] (full grid, n = nx*ny*nz) in place of its constant eta; eta = NULL returns
   to the constant. */
lbmx_status lbmx_filter_set_eta_field(lbmx_filter* f, int stage, const double* eta, size_t n);
/* increaseBeta: every stage's beta *= factor (doubleProjectionFilter2d.hh:77-81) */
lbmx_status lbmx_filter_increase_beta(lbmx_filter* f, double factor);
/* Full-grid values outside design_box, n = nx*ny*nz: LBSolver::createDesignVec copies the
   solver's material field (src/latticeBoltzmann/LBSolver.cc:318-331) before
   mapDesignToFull inserts the design box. */
lbmx_status lbmx_filter_set_background(lbmx_filter* f, const double* full, size_t n);
/* mapDesignToFull + filterDesign: design (nd = design-box nodes) -> physical field (nf = full
   nodes; physical may be NULL). Caches each stage's density-filtered field for the next
   sensitivity call, as ProjectionFilter2d keeps densityFilteredField. */
lbmx_status lbmx_filter_design(lbmx_filter* f, const double* design, size_t nd, double* physical, size_t nf);
/* The same, then LBSolver::setFieldFromDesignVec (LBSolver.cc:333-348) on the device: the
   solver's design field is written in HBM without a host round trip. */
lbmx_status lbmx_filter_design_to_solver(lbmx_filter* f, const double* design, size_t nd, lbmx_solver* s);
/* filterSensitivities + mapFullToDesign for one full-grid field (objective or constraint):
   stages in reverse order, each the projection derivative at the cached filtered field, then
   the transposed density filter; out = the design-box entries (nd). */
lbmx_status lbmx_filter_sensitivities(lbmx_filter* f, const double* full_sens, size_t nf, double* design_sens,
                                      size_t nd);
/* The same on the solver's device-resident dJ/dgamma of its last lbmx_gradient. */
lbmx_status lbmx_filter_solver_sensitivities(lbmx_filter* f, lbmx_solver* s, double* design_sens, size_t nd);
/* The four calls above on device arrays: every double* argument is a HIP device pointer
   (physical may still be NULL). The design and its sensitivities then stay in HBM through
   the whole optimisation iteration. Inputs must be ready when the call is made; the calls
   return with the device work done. */
lbmx_status lbmx_filter_design_device(lbmx_filter* f, const double* design, size_t nd, double* physical, size_t nf);
lbmx_status lbmx_filter_design_to_solver_device(lbmx_filter* f, const double* design, size_t nd, lbmx_solver* s);
lbmx_status lbmx_filter_sensitivities_device(lbmx_filter* f, const double* full_sens, size_t nf, double* design_sens,
                                             size_t nd);
lbmx_status lbmx_filter_solver_sensitivities_device(lbmx_filter* f, lbmx_solver* s, double* design_sens, size_t nd);
/* Intermediate fields of the last lbmx_filter_design (outputIntermediateFields,
   projectionFilter2d.cc:89-103): which = 2*stage for the stage's density-filtered field,
   2*stage + 1 for its output; n = full nodes. */
lbmx_status lbmx_filter_get_field(lbmx_filter* f, int which, double* out, size_t n);
/* Milliseconds per call of the last lbmx_filter_design / sensitivity call on the device
   (HIP events around its kernels): ms[0] design, ms[1] sensitivities. */
lbmx_status lbmx_filter_timing(lbmx_filter* f, double ms[2]);

/* ---- file formats either side of the engine (host only) ---- */
/* writeVectorToBinary / loadVectorFromBinary (src/core/binaryIO.cc:3-34): PETSc's binary Vec
   format, big-endian int32 class id 1211214, int32 length, then the FP64 values in natural
   order; the write also leaves PETSc's `<path>.info` options file. The load fails unless the
   file holds exactly n values. */
lbmx_status lbmx_vec_write_binary(const char* path, const double* v, size_t n);
lbmx_status lbmx_vec_load_binary(const char* path, double* v, size_t n);
/* LBSolver::outputFields (src/latticeBoltzmann/LBSolver.cc:251-286): the current state's macros
   (rho, ux, uy[, uz][, T]) and the design field as a VTK XML StructuredGrid (.vts, ASCII). */
lbmx_status lbmx_write_vts(lbmx_solver* s, const char* path);
/* PETSc DMDA orderings (the reference's distributed Vecs, DMDACreate2d/3d in
   src/latticeBoltzmann/NewLBSolver2d.cc and LBSolver3d.cc): a global Vec of a DMDA split over
   procs[0] x procs[1] x procs[2] ranks holds the ranks' blocks one after another (rank
   i + procs[0] (j + procs[1] k)), each block in natural order over its own nodes, `dof` values
   per node. lx[procs[0]], ly[procs[1]], lz[procs[2]] are the nodes each rank owns along an axis
   (DMDAGetOwnershipRanges; lz may be NULL when procs[2] == 1). These convert such a Vec's
   array to the natural order the engine uses (DMDAGlobalToNatural) and back. */
lbmx_status lbmx_dmda_global_to_natural(const int dims[3], const int procs[3], const int* lx, const int* ly,
                                        const int* lz, int dof, const double* global, double* natural);
lbmx_status lbmx_dmda_natural_to_global(const int dims[3], const int procs[3], const int* lx, const int* ly,
                                        const int* lz, int dof, const double* natural, double* global);

/* ---- method of moving asymptotes (the design update of the optimisation loop) ----
   MMA of src/topOpt/MMA.cc (Aage's PETSc MMA with the reference's robust constants): each
   update builds Svanberg's convex separable subproblem around the design and solves it
   through its dual by an interior-point Newton iteration on the m multipliers. The n design
   variables live in HBM; design vectors cross the boundary as host arrays in natural order. */
#define LBMX_MMA_MAX_CONSTRAINTS 8
typedef struct lbmx_mma lbmx_mma;
/* MMA::MMA(n, m, x) (src/topOpt/MMA.cc:208-251): a_j = 0, c_j = 1000, asymptote factors
   0.5 / 0.7 / 1.2; m in 1..LBMX_MMA_MAX_CONSTRAINTS */
lbmx_status lbmx_mma_create(long long n, int m, lbmx_mma** out);
lbmx_status lbmx_mma_destroy(lbmx_mma* h);
/* MMA::SetRobustConstants (MMA.cc:290-296): a[m], the weights of the bound variable z in the
   constraints (1 for a robust realisation, 0 for the volume constraint) */
lbmx_status lbmx_mma_set_robust_constants(lbmx_mma* h, const double* a);
/* MMA::SetAsymptotes (MMA.cc:343-351) */
lbmx_status lbmx_mma_set_asymptotes(lbmx_mma* h, double init, double decrease, double increase);
/* MMA::ConstraintModification (MMA.hh:131) and MMA::SetRobustAsymptotesType (MMA.cc:353-367) */
lbmx_status lbmx_mma_set_options(lbmx_mma* h, int constraint_modification, int robust_asymptotes_type);
/* MMA::Update (MMA.cc:485-500): x[n] is the design in and the new design out; dfdx[n] the
   objective sensitivity, gx[m] the constraint values, dgdx[m] pointers to their sensitivities;
   move limits max(0, x - movlim) .. min(1, x + movlim). */
lbmx_status lbmx_mma_update(lbmx_mma* h, double* x, const double* dfdx, const double* gx, const double* const* dgdx,
                            double movlim);
/* MMA::DesignChange (MMA.cc:389-414): *ch = max |xnew - x before the last update|; xnew = NULL
   uses the design of the last update */
lbmx_status lbmx_mma_design_change(lbmx_mma* h, const double* xnew, double* ch);
/* The same two calls on device arrays (x, dfdx and every dgdx[j] are HIP device pointers; gx
   stays a host array of m values): the optimisation loop keeps the design and its
   sensitivities in HBM. They return with the device work done. */
lbmx_status lbmx_mma_update_device(lbmx_mma* h, double* x, const double* dfdx, const double* gx,
                                   const double* const* dgdx, double movlim);
lbmx_status lbmx_mma_design_change_device(lbmx_mma* h, const double* xnew, double* ch);
/* Diagnostics: asymptotes L[n], U[n], multipliers lam[m], dual Newton iterations and wall
   milliseconds of the last update (any pointer may be NULL). */
lbmx_status lbmx_mma_get_state(lbmx_mma* h, double* L, double* U, double* lam, int* newton_iters, double* ms);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the design change
   of lbmx_mma_design_change_device, max_i |xnew[i] - xold[i]| (MMA::DesignChange, MMA.cc:389-414),
   with the previous design passed in instead of an MMA handle. xnew, xold (n values) and ch (one
   value, the result) are HIP device pointers; the work is enqueued on hip_stream (NULL: the
   legacy default stream) and the call returns without waiting for it. It allocates nothing: the
   reduction's per-block partials live in a library-static scratch array of 2048 doubles (16 KB)
   on each device, kept until the process ends. Calls that overlap on one device race on that
   array. n must lie in 1..LBMX_NOALLOC_MAX_N (else LBMX_ERR_ARG). */
#define LBMX_NOALLOC_MAX_N (1LL << 40)
lbmx_status lbmx_design_change_noalloc(const double* xnew, const double* xold, long long n, double* ch,
                                       void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* LBMX_H */
