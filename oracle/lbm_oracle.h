/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement, in plain C, of the reference's lattice Boltzmann forward step,
 * reverse (adjoint) step, boundary conditions and functionals. Only tests/, the
 * smoke check and the bench's cpu_baseline leg may load it; the engine never does.
 * Layout is the reference's single-rank PETSc DMDA array-of-structures [k][j][i][q].
 *
 * Pinning status of each part is recorded in the header of lbm_oracle.c.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include "lbmx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_problem orc_problem;

orc_problem* orc_create(const lbmx_desc* d);
void orc_destroy(orc_problem* p);
int orc_add_boundary(orc_problem* p, const lbmx_box* box, const lbmx_bc_desc* bc);
/* the same with value tables (lbmx_add_boundary_table): prof (ncomp per node of box) and tfac
   (b(t), t = 0..nt-1) may be NULL */
int orc_add_boundary_table(orc_problem* p, const lbmx_box* box, const lbmx_bc_desc* bc, const double* prof,
                           long long nprof, int ncomp, const double* tfac, int nt);
int orc_num_boundaries(const orc_problem* p);
/* piece k: type, normal (3 ints), box (6 ints) */
int orc_boundary_info(const orc_problem* p, int k, int* type, int* nrm, int* box);
int orc_set_design(orc_problem* p, const double* gamma);
int orc_set_objective(orc_problem* p, const lbmx_objective* obj);
int orc_set_objective_terms(orc_problem* p, const lbmx_term* terms, int nterms, double time_const);
int orc_num_dof(const orc_problem* p);
void orc_set_threads(int n);
/* on != 0: the cascaded reverse collision is the reference's ReverseIncompressibleCascaded2d as
   coded (src/adjointDynamics/reverseCascaded2d.hh:27-133), which is not the transpose of the
   forward cascaded collision; for reporting the reference adjoint's deviation only */
void orc_set_reference_reverse(orc_problem* p, int on);

/* ---- node-level primitives (for unit tests) ---- */
void orc_lattice(const orc_problem* p, int* Q, int* ex, int* ey, int* ez, int* opp, double* w);
/* collision of one node: fin (Q) -> fout (Q), mac[0..4] = (rho, jx, jy, jz, T) */
void orc_collide_node(const orc_problem* p, const double* fin, double* fout, double* mac, double gamma);
/* transpose: lam_post (Q) -> lam_pre (Q) = (dC/df)^T lam_post, *dgamma += (dC/dgamma)^T lam_post */
void orc_reverse_collide_node(const orc_problem* p, const double* fin, double gamma,
                              const double* lam_post, double* lam_pre, double* dgamma);
/* apply boundary piece k at node (i,j,k) of the full post-stream state f (AoS), time t */
void orc_apply_bc(const orc_problem* p, int piece, int t, double* f);
void orc_apply_bc_adjoint(const orc_problem* p, int piece, int t, double* lam);
void orc_stream(const orc_problem* p, const double* fpost, double* fnew);
void orc_adjoint_stream(const orc_problem* p, const double* lam_in, double* lam_out);
/* objective of one state (sum over terms, times time_const) and its gradient (added into g) */
double orc_objective(const orc_problem* p, const double* f, int t);
void orc_objective_grad(const orc_problem* p, const double* f, int t, double* g);

/* ---- loops (the reference's solver step order) ---- */
void orc_init_equilibrium(const orc_problem* p, double* f, double rho0, const double* u0, double T0);
/* nsteps x (collide all, pull-stream, ++t, BCs in insertion order); J accumulates the objective
   of each new state. t is in/out. */
void orc_forward(const orc_problem* p, double* f, int nsteps, int* t, double* J);
/* Full gradient with every state stored: returns J, sens (nodes), adjoint (AoS) */
int orc_gradient(const orc_problem* p, int nsteps, double rho0, const double* u0, double T0,
                 double* J, double* sens, double* adj);
/* macros of a pre-collision state: (rho, jx, jy, jz, T) per node, 5 doubles */
void orc_macros(const orc_problem* p, const double* f, double* mac);

#ifdef __cplusplus
}
#endif
#endif
