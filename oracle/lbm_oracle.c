/*
 * ORACLE -- test infrastructure only (see lbm_oracle.h). The engine never links this.
 *
 * A plain-C restatement of the reference's lattice Boltzmann hot path. Each function
 * cites the reference file:line it follows. Array layout is the reference's DMDA AoS
 * order f[((k*ny + j)*nx + i)*Q + q].
 *
 * How each part is pinned (tests/test_oracle_*.py), independently of this code:
 *   lattice constants      : isotropy sums (sum w = 1, sum w c = 0, sum w cc = I/3), opp
 *   equilibrium            : the 9/19 equilibrium moments of the second-order polynomial
 *   BGK                    : affine in omega; omega=1 gives the equilibrium moments; the
 *                            compressible form against the textbook equilibrium at u = j/rho
 *   MRT 2D / 3D            : moment-space relaxation m' = m - S(m - m_eq) with the textbook
 *                            Lallemand-Luo (D2Q9) and d'Humieres 2002 (D3Q19) bases built in
 *                            the test from their polynomials; equal rates reduce to BGK
 *   thermal MRT (D2Q5)     : moment-space relaxation with the D2Q5 basis; T conserved
 *   PBB + BP               : closed-form 2x2 pair map; gamma=1 identity; gamma=0 pair mean
 *   Brinkman (Guo forcing) : alpha=0 reduces to the base operator; momentum change = F
 *   DarcyMRT2d forcing     : mass 0, momentum change F, moment rows (1 - s_k/2) x the Guo
 *                            source moments (q rows: the full source) -- the polynomials as coded
 *   pull stream            : brute-force permutation on tiny grids
 *   boundary conditions    : prescribed u / rho / T recovered at the node; bounce-back
 *                            against the rule "incoming <- opposite, tangential swapped", and
 *                            the (x0,y0,z0) D3Q19 corner against its as-coded copy-and-swap list;
 *                            every Zou-He / Hecht velocity, pressure and Neumann piece (each
 *                            orientation, incompressible and standard): only the unknowns
 *                            change, the prescribed moments hold, and the non-equilibrium
 *                            bounce-back holds along the normal and per tangential axis -- a
 *                            rank check shows these equations fix every unknown;
 *                            Zou-He Neumann (flow, D2Q5 temperature): u_x or T of the node equals
 *                            (4 m_1 - m_2)/3 of the two nodes inwards, u_y = 0, only the unknowns
 *                            change, whole-field dot test including the neighbour blocks
 *   reverse operators,
 *   adjoint BCs, stream    : dot-product tests against central differences of the forward
 *   functionals + gradient : closed forms on uniform states (pressure drop, outflow, through-
 *                            flow, total-pressure flux); central finite differences of J on
 *                            random cells, also with Neumann outlets
 *   cascaded collision     : conservation; shear moments relax with omega to rho u_a u_b for
 *                            any state; at u = 0 every moment reaches its equilibrium (rate 1);
 *                            for any rho and u = j, the central moments of orders 2-4 relax to
 *                            their equilibria (Geier, Greiner & Korvink 2006), which fixes every
 *                            term of k0..k5; its reverse: dot-product test (true transpose)
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define QMAX 24
#define ORC_PI 3.14159265358979323846
#define NTERM_MAX 4

/* ------------------------------------------------------------------------- */
/* lattices: src/latticeBoltzmann/lattices2d.hh:16-84, lattices3d.hh:34-53    */
/* ------------------------------------------------------------------------- */
static const int D2Q9_EX[9] = {0, 1, 0, -1, -1, -1, 0, 1, 1};
static const int D2Q9_EY[9] = {0, -1, -1, -1, 0, 1, 1, 1, 0};
static const int D2Q9_OPP[9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
static const double D2Q9_W[9] = {4. / 9., 1. / 36., 1. / 9., 1. / 36., 1. / 9.,
                                 1. / 36., 1. / 9., 1. / 36., 1. / 9.};
static const int D3Q19_EX[19] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
static const int D3Q19_EY[19] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
static const int D3Q19_EZ[19] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
static const double D3Q19_W[19] = {1. / 3., 1. / 18., 1. / 18., 1. / 18., 1. / 36., 1. / 36., 1. / 36.,
                                   1. / 36., 1. / 36., 1. / 36., 1. / 18., 1. / 18., 1. / 18.,
                                   1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
static const int D2Q5_EX[5] = {0, 0, -1, 0, 1};
static const int D2Q5_EY[5] = {0, -1, 0, 1, 0};
static const int D2Q5_OPP[5] = {0, 3, 4, 1, 2};
static const double D2Q5_W[5] = {1. / 3., 1. / 6., 1. / 6., 1. / 6., 1. / 6.};

/* equilibrium moment of one MRT row: m_eq = a*rho + b.v + (v^T C v)/rho_s  (flow), or
   m_eq = T*(a + b.u) (thermal). rate: 0 conserved, 1 the viscous rate omega, 2 constant s. */
typedef struct {
  double a, b[3], C[3][3];
  int rate;
  double s;
} orc_row;

typedef struct {
  int n;
  double M[QMAX][QMAX];
  double inv_norm2[QMAX];
  orc_row row[QMAX];
} orc_mrt;

typedef struct {
  int type;
  int nrm[3];
  lbmx_box box;
  lbmx_bc_desc d;
  /* value tables of lbmx_add_boundary_table (each piece owns a copy; NULL: none) */
  double* prof;      /* ncomp values per node of tbox, natural order */
  lbmx_box tbox;
  int ncomp;
  double* tfac;      /* b(t), t = 0..nt-1 */
  int nt;
} orc_bc;

typedef struct {
  lbmx_box box;
  double weight;
  int kind; /* 0 rho, 1 ux, 2 total pressure flux x, 3 ux*T, 4 T */
} orc_term;

struct orc_problem {
  lbmx_desc d;
  int D, Qf, Qt, Q;
  int ex[QMAX], ey[QMAX], ez[QMAX], opp[QMAX];
  double w[QMAX];
  int half;
  int incompressible;
  orc_mrt mrt, tmrt;
  int nbc, cap;
  orc_bc* bcs;
  int nterm;
  orc_term terms[NTERM_MAX];
  double time_const;
  double* gamma;
  size_t nodes;
  int reference_reverse; /* cascaded reverse as the reference codes it (test comparison only) */
};

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
void orc_set_reference_reverse(orc_problem* p, int on) { p->reference_reverse = on != 0; }

/* ------------------------------------------------------------------------- */
/* MRT bases                                                                  */
/* ------------------------------------------------------------------------- */
static void mrt_finish(orc_mrt* m, int Q) {
  m->n = Q;
  for (int k = 0; k < Q; ++k) {
    double s = 0;
    for (int q = 0; q < Q; ++q) s += m->M[k][q] * m->M[k][q];
    m->inv_norm2[k] = 1.0 / s;
  }
}

static void row_rate(orc_row* r, int rate, double s) { r->rate = rate; r->s = s; }

/* Lallemand & Luo (2000) D2Q9 basis with the rates and incompressible equilibria of
   IncompressibleMRT2d<D2Q9> (src/dynamics/incompressibleMRT2d.hh:44-97). */
static void build_mrt_d2q9(orc_mrt* m) {
  memset(m, 0, sizeof(*m));
  for (int q = 0; q < 9; ++q) {
    double cx = D2Q9_EX[q], cy = D2Q9_EY[q], c2 = cx * cx + cy * cy;
    m->M[0][q] = 1;
    m->M[1][q] = cx;
    m->M[2][q] = cy;
    m->M[3][q] = -4 + 3 * c2;
    m->M[4][q] = 4 - 10.5 * c2 + 4.5 * c2 * c2;
    m->M[5][q] = (-5 + 3 * c2) * cx;
    m->M[6][q] = (-5 + 3 * c2) * cy;
    m->M[7][q] = cx * cx - cy * cy;
    m->M[8][q] = cx * cy;
  }
  orc_row* r = m->row;
  r[0].a = 1; row_rate(&r[0], 0, 0);
  r[1].b[0] = 1; row_rate(&r[1], 0, 0);
  r[2].b[1] = 1; row_rate(&r[2], 0, 0);
  r[3].a = -2; r[3].C[0][0] = r[3].C[1][1] = 3; row_rate(&r[3], 2, 1.4);
  r[4].a = 1; r[4].C[0][0] = r[4].C[1][1] = -3; row_rate(&r[4], 2, 1.4);
  r[5].b[0] = -1; row_rate(&r[5], 2, 1.2);
  r[6].b[1] = -1; row_rate(&r[6], 2, 1.2);
  r[7].C[0][0] = 1; r[7].C[1][1] = -1; row_rate(&r[7], 1, 0);
  r[8].C[0][1] = r[8].C[1][0] = 0.5; row_rate(&r[8], 1, 0);
  mrt_finish(m, 9);
}

/* d'Humieres et al. (2002) D3Q19 basis; rates and equilibrium weights as
   MRT3d_Compile_constants (src/dynamics/MRT3d.hh:9-22, 92-109). */
static void build_mrt_d3q19(orc_mrt* m) {
  memset(m, 0, sizeof(*m));
  for (int q = 0; q < 19; ++q) {
    double cx = D3Q19_EX[q], cy = D3Q19_EY[q], cz = D3Q19_EZ[q];
    double c2 = cx * cx + cy * cy + cz * cz;
    m->M[0][q] = 1;
    m->M[1][q] = cx;
    m->M[2][q] = cy;
    m->M[3][q] = cz;
    m->M[4][q] = 19 * c2 - 30;
    m->M[5][q] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
    m->M[6][q] = (5 * c2 - 9) * cx;
    m->M[7][q] = (5 * c2 - 9) * cy;
    m->M[8][q] = (5 * c2 - 9) * cz;
    m->M[9][q] = 3 * cx * cx - c2;
    m->M[10][q] = (3 * c2 - 5) * (3 * cx * cx - c2);
    m->M[11][q] = cy * cy - cz * cz;
    m->M[12][q] = (3 * c2 - 5) * (cy * cy - cz * cz);
    m->M[13][q] = cx * cy;
    m->M[14][q] = cy * cz;
    m->M[15][q] = cx * cz;
    m->M[16][q] = (cy * cy - cz * cz) * cx;
    m->M[17][q] = (cz * cz - cx * cx) * cy;
    m->M[18][q] = (cx * cx - cy * cy) * cz;
  }
  const double s1 = 1.19, s2 = 1.4, s4 = 1.2, s10 = 1.4, s16 = 1.4;
  const double we = 3., wej = -5.5, wxx = -0.5;
  orc_row* r = m->row;
  r[0].a = 1; row_rate(&r[0], 0, 0);
  for (int a = 0; a < 3; ++a) { r[1 + a].b[a] = 1; row_rate(&r[1 + a], 0, 0); }
  r[4].a = -11; for (int a = 0; a < 3; ++a) r[4].C[a][a] = 19; row_rate(&r[4], 2, s1);
  r[5].a = we; for (int a = 0; a < 3; ++a) r[5].C[a][a] = wej; row_rate(&r[5], 2, s2);
  for (int a = 0; a < 3; ++a) { r[6 + a].b[a] = -2. / 3.; row_rate(&r[6 + a], 2, s4); }
  r[9].C[0][0] = 2; r[9].C[1][1] = -1; r[9].C[2][2] = -1; row_rate(&r[9], 1, 0);
  r[10].C[0][0] = 2 * wxx; r[10].C[1][1] = -wxx; r[10].C[2][2] = -wxx; row_rate(&r[10], 2, s10);
  r[11].C[1][1] = 1; r[11].C[2][2] = -1; row_rate(&r[11], 1, 0);
  r[12].C[1][1] = wxx; r[12].C[2][2] = -wxx; row_rate(&r[12], 2, s10);
  r[13].C[0][1] = r[13].C[1][0] = 0.5; row_rate(&r[13], 1, 0);
  r[14].C[1][2] = r[14].C[2][1] = 0.5; row_rate(&r[14], 1, 0);
  r[15].C[0][2] = r[15].C[2][0] = 0.5; row_rate(&r[15], 1, 0);
  for (int k = 16; k < 19; ++k) row_rate(&r[k], 2, s16);
  mrt_finish(m, 19);
}

/* ThermalMRT2d<D2Q5_TMRT> (src/dynamics/thermalMRT2d.hh:50-101): rows T, q_x, q_y, e, p_xx */
static void build_mrt_d2q5(orc_mrt* m) {
  memset(m, 0, sizeof(*m));
  for (int q = 0; q < 5; ++q) {
    double cx = D2Q5_EX[q], cy = D2Q5_EY[q], c2 = cx * cx + cy * cy;
    m->M[0][q] = 1;
    m->M[1][q] = cx;
    m->M[2][q] = cy;
    m->M[3][q] = 5 * c2 - 4;
    m->M[4][q] = cx * cx - cy * cy;
  }
  orc_row* r = m->row;
  r[0].a = 1; row_rate(&r[0], 0, 0);
  r[1].b[0] = 1; row_rate(&r[1], 1, 0);
  r[2].b[1] = 1; row_rate(&r[2], 1, 0);
  r[3].a = -2; row_rate(&r[3], 2, 1.5);
  row_rate(&r[4], 2, 1.5);
  mrt_finish(m, 5);
}

/* ------------------------------------------------------------------------- */
/* interpolation: src/core/interpolationFunctions.hh:7-54                      */
/* ------------------------------------------------------------------------- */
static double bp(double q, double s) { return 1. - s * (1. + q) / (q + s); }
static double bp_diff(double q, double s) {
  return -(1. + q) / (q + s) + s * (1. + q) / ((q + s) * (q + s));
}
static double ramp(double p, double ck, double s) {
  return (s * (ck * (1. + p) - 1.) + 1.) / (ck * (1. + p * s));
}
static double ramp_diff(double p, double ck, double s) {
  return (ck * (1. + p) - 1.) / (ck * (1. + p * s)) -
         ((s * (ck * (1. + p) - 1.) + 1.) * p) / (ck * (1. + p * s) * (1. + p * s));
}

/* ------------------------------------------------------------------------- */
/* problem setup                                                              */
/* ------------------------------------------------------------------------- */
orc_problem* orc_create(const lbmx_desc* d) {
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return NULL;
  /* operators restated so far */
  /* DarcyMRT2d: D2Q9 incompressible MRT only, forward only (its reverse is empty) */
  if (d->design == LBMX_DESIGN_DARCY && (d->lattice != LBMX_D2Q9 || d->collision != LBMX_MRT)) return NULL;
  /* cascaded: D2Q9 only, coupled to the design through partial bounce-back (or not at all) */
  if (d->collision == LBMX_CASCADED &&
      (d->lattice != LBMX_D2Q9 || d->design == LBMX_DESIGN_BRINKMAN)) return NULL;
  if (d->collision == LBMX_MRT_STANDARD && d->lattice != LBMX_D3Q19) return NULL;
  if (d->lattice != LBMX_D3Q19 && d->nz != 1) return NULL;
  orc_problem* p = (orc_problem*)calloc(1, sizeof(orc_problem));
  p->d = *d;
  p->incompressible = !(d->collision == LBMX_MRT_STANDARD || d->collision == LBMX_BGK_STANDARD);
  if (d->lattice == LBMX_D3Q19) {
    p->D = 3; p->Qf = 19; p->Qt = 0;
    for (int q = 0; q < 19; ++q) {
      p->ex[q] = D3Q19_EX[q]; p->ey[q] = D3Q19_EY[q]; p->ez[q] = D3Q19_EZ[q];
      p->w[q] = D3Q19_W[q]; p->opp[q] = q == 0 ? 0 : (q <= 9 ? q + 9 : q - 9);
    }
    build_mrt_d3q19(&p->mrt);
  } else {
    p->D = 2; p->Qf = 9;
    p->Qt = d->lattice == LBMX_D2Q9_D2Q5 ? 5 : 0;
    for (int q = 0; q < 9; ++q) {
      p->ex[q] = D2Q9_EX[q]; p->ey[q] = D2Q9_EY[q]; p->ez[q] = 0;
      p->w[q] = D2Q9_W[q]; p->opp[q] = D2Q9_OPP[q];
    }
    for (int q = 0; q < p->Qt; ++q) {
      p->ex[9 + q] = D2Q5_EX[q]; p->ey[9 + q] = D2Q5_EY[q]; p->ez[9 + q] = 0;
      p->w[9 + q] = D2Q5_W[q]; p->opp[9 + q] = 9 + D2Q5_OPP[q];
    }
    build_mrt_d2q9(&p->mrt);
    build_mrt_d2q5(&p->tmrt);
  }
  p->Q = p->Qf + p->Qt;
  p->half = (p->Qf - 1) / 2;
  p->nodes = (size_t)d->nx * d->ny * d->nz;
  p->gamma = (double*)malloc(p->nodes * sizeof(double));
  for (size_t n = 0; n < p->nodes; ++n) p->gamma[n] = 1.0;
  return p;
}

void orc_destroy(orc_problem* p) {
  if (!p) return;
  for (int k = 0; k < p->nbc; ++k) {
    free(p->bcs[k].prof);
    free(p->bcs[k].tfac);
  }
  free(p->gamma);
  free(p->bcs);
  free(p);
}

int orc_num_dof(const orc_problem* p) { return p->Q; }

void orc_lattice(const orc_problem* p, int* Q, int* ex, int* ey, int* ez, int* opp, double* w) {
  *Q = p->Q;
  for (int q = 0; q < p->Q; ++q) {
    ex[q] = p->ex[q]; ey[q] = p->ey[q]; ez[q] = p->ez[q]; opp[q] = p->opp[q]; w[q] = p->w[q];
  }
}

int orc_set_design(orc_problem* p, const double* g) {
  memcpy(p->gamma, g, p->nodes * sizeof(double));
  return 0;
}

static int box_size(const lbmx_box* b) {
  return (b->x1 - b->x0 + 1) * (b->y1 - b->y0 + 1) * (b->z1 - b->z0 + 1);
}

/* Functional terms of PressureDrop2d (src/functionals/pressureDrop2d.hh:14-52),
   OutflowEast2d (src/functionals/outflow2d.hh:15-54), ThermalOutflowEast2d
   (src/functionals/temperatureFlow2d.hh:18-54) and the total-pressure-flux objective. */
int orc_set_objective(orc_problem* p, const lbmx_objective* o) {
  p->nterm = 0;
  p->time_const = o->time_const;
  switch (o->kind) {
    case LBMX_OBJ_NONE: break;
    case LBMX_OBJ_PRESSURE_DROP:
      p->terms[0] = (orc_term){o->box_a, 1.0 / box_size(&o->box_a), 0};
      p->terms[1] = (orc_term){o->box_b, -1.0 / box_size(&o->box_b), 0};
      p->nterm = 2;
      break;
    case LBMX_OBJ_OUTFLOW:
      p->terms[0] = (orc_term){o->box_a, -1.0 / box_size(&o->box_a), 1};
      p->nterm = 1;
      break;
    case LBMX_OBJ_THROUGHFLOW: /* ThroughFlow2d (src/functionals/throughFlow2d.hh:16-88) */
      p->terms[0] = (orc_term){o->box_a, -1.0 / box_size(&o->box_a), 1};
      p->terms[1] = (orc_term){o->box_b, -1.0 / box_size(&o->box_b), 1};
      p->nterm = 2;
      break;
    case LBMX_OBJ_DISSIPATED_POWER:
      p->terms[0] = (orc_term){o->box_a, 1.0, 2};
      p->terms[1] = (orc_term){o->box_b, -1.0, 2};
      p->nterm = 2;
      break;
    case LBMX_OBJ_HEAT_FLUX:
      if (p->Qt == 0) return LBMX_ERR_SUP;
      p->terms[0] = (orc_term){o->box_a, -1.0, 3};
      p->nterm = 1;
      break;
    default: return LBMX_ERR_ARG;
  }
  return 0;
}

/* Explicit terms of a linear box functional (the LBMacroFunctional interface,
   src/functionals/LBMacroFunctional.hh:6-14), e.g. DomainTemperature2d
   (src/functionals/domainTemperature2d.hh: timeConst / area * sum_box T). */
int orc_set_objective_terms(orc_problem* p, const lbmx_term* t, int nterms, double time_const) {
  if (nterms <= 0 || nterms > NTERM_MAX) return LBMX_ERR_ARG;
  for (int m = 0; m < nterms; ++m) {
    if (t[m].kind < 0 || t[m].kind > 4) return LBMX_ERR_ARG;
    if ((t[m].kind == 3 || t[m].kind == 4) && p->Qt == 0) return LBMX_ERR_SUP;
  }
  p->time_const = time_const;
  for (int m = 0; m < nterms; ++m) p->terms[m] = (orc_term){t[m].box, t[m].weight, t[m].kind};
  p->nterm = nterms;
  return 0;
}

/* Box splitting of NewLBSolver2d::addBoundaryCondition (src/latticeBoltzmann/NewLBSolver2d.cc:106-155)
   with getBoxBoundaries (src/core/geometry2d.cc:31-53), and the 3D analogue
   (src/latticeBoltzmann/LBSolver3d.cc, src/core/geometry3d.cc:36-85). */
static int bc_supported(const orc_problem* p, int type, const int* n) {
  int nn = abs(n[0]) + abs(n[1]) + abs(n[2]);
  int inc = p->incompressible;
  switch (type) {
    case LBMX_BC_BOUNCEBACK: return 1;
    case LBMX_BC_ZOUHE_VELOCITY:
      if (p->D != 2 || nn != 1) return 0;
      if (inc) return n[0] == -1 || n[1] == 1 || n[0] == 1;
      return n[0] == -1;
    case LBMX_BC_ZOUHE_PRESSURE: return p->D == 2 && nn == 1 && n[1] == 0;
    case LBMX_BC_ZOUHE_NEUMANN: return p->D == 2 && nn == 1 && n[1] == 0 && inc;
    case LBMX_BC_HECHT_VELOCITY: return p->D == 3 && nn == 1 && n[0] == -1;
    case LBMX_BC_HECHT_PRESSURE: return p->D == 3 && nn == 1 && n[0] == 1;
    case LBMX_BC_THERMAL_BOUNCEBACK: return p->Qt > 0;
    case LBMX_BC_THERMAL_DIRICHLET: return p->Qt > 0 && nn == 1;
    case LBMX_BC_THERMAL_NEUMANN: return p->Qt > 0 && p->D == 2 && nn == 1 && n[1] == 0;
  }
  return 0;
}

static int intersect(const lbmx_box* a, const lbmx_box* b, lbmx_box* o) {
  o->x0 = a->x0 > b->x0 ? a->x0 : b->x0; o->x1 = a->x1 < b->x1 ? a->x1 : b->x1;
  o->y0 = a->y0 > b->y0 ? a->y0 : b->y0; o->y1 = a->y1 < b->y1 ? a->y1 : b->y1;
  o->z0 = a->z0 > b->z0 ? a->z0 : b->z0; o->z1 = a->z1 < b->z1 ? a->z1 : b->z1;
  return o->x1 >= o->x0 && o->y1 >= o->y0 && o->z1 >= o->z0;
}

int orc_add_boundary(orc_problem* p, const lbmx_box* box, const lbmx_bc_desc* bc) {
  return orc_add_boundary_table(p, box, bc, NULL, 0, 0, NULL, 0);
}

int orc_add_boundary_table(orc_problem* p, const lbmx_box* box, const lbmx_bc_desc* bc, const double* prof,
                           long long nprof, int ncomp, const double* tfac, int nt) {
  int X1 = p->d.nx - 1, Y1 = p->d.ny - 1, Z1 = p->d.nz - 1;
  long long bn = (long long)(box->x1 - box->x0 + 1) * (box->y1 - box->y0 + 1) * (box->z1 - box->z0 + 1);
  if (prof && (ncomp < 1 || nprof != bn * ncomp)) return LBMX_ERR_ARG;
  if (tfac && nt < 1) return LBMX_ERR_ARG;
  /* pieces in the reference's order; each as (box, normal) */
  lbmx_box pb[26];
  int pn[26][3];
  int np = 0;
#define PIECE(bx0, bx1, by0, by1, bz0, bz1, n0, n1, n2)                       \
  do {                                                                        \
    pb[np] = (lbmx_box){bx0, bx1, by0, by1, bz0, bz1};                        \
    pn[np][0] = n0; pn[np][1] = n1; pn[np][2] = n2; ++np;                      \
  } while (0)
  if (p->D == 2) {
    PIECE(1, X1 - 1, Y1, Y1, 0, 0, 0, 1, 0);   /* north */
    PIECE(1, X1 - 1, 0, 0, 0, 0, 0, -1, 0);    /* south */
    PIECE(0, 0, 1, Y1 - 1, 0, 0, -1, 0, 0);    /* west  */
    PIECE(X1, X1, 1, Y1 - 1, 0, 0, 1, 0, 0);   /* east  */
    PIECE(0, 0, Y1, Y1, 0, 0, -1, 1, 0);       /* NW    */
    PIECE(X1, X1, Y1, Y1, 0, 0, 1, 1, 0);      /* NE    */
    PIECE(0, 0, 0, 0, 0, 0, -1, -1, 0);        /* SW    */
    PIECE(X1, X1, 0, 0, 0, 0, 1, -1, 0);       /* SE    */
  } else {
    PIECE(1, X1 - 1, Y1, Y1, 1, Z1 - 1, 0, 1, 0);     /* north face */
    PIECE(1, X1 - 1, 0, 0, 1, Z1 - 1, 0, -1, 0);      /* south      */
    PIECE(0, 0, 1, Y1 - 1, 1, Z1 - 1, -1, 0, 0);      /* west       */
    PIECE(X1, X1, 1, Y1 - 1, 1, Z1 - 1, 1, 0, 0);     /* east       */
    PIECE(1, X1 - 1, 1, Y1 - 1, Z1, Z1, 0, 0, 1);     /* front      */
    PIECE(1, X1 - 1, 1, Y1 - 1, 0, 0, 0, 0, -1);      /* back       */
    PIECE(0, 0, Y1, Y1, 1, Z1 - 1, -1, 1, 0);         /* NW edge    */
    PIECE(X1, X1, Y1, Y1, 1, Z1 - 1, 1, 1, 0);        /* NE         */
    PIECE(1, X1 - 1, Y1, Y1, Z1, Z1, 0, 1, 1);        /* NF         */
    PIECE(1, X1 - 1, Y1, Y1, 0, 0, 0, 1, -1);         /* NB         */
    PIECE(0, 0, 0, 0, 1, Z1 - 1, -1, -1, 0);          /* SW         */
    PIECE(X1, X1, 0, 0, 1, Z1 - 1, 1, -1, 0);         /* SE         */
    PIECE(1, X1 - 1, 0, 0, Z1, Z1, 0, -1, 1);         /* SF         */
    PIECE(1, X1 - 1, 0, 0, 0, 0, 0, -1, -1);          /* SB         */
    PIECE(0, 0, 1, Y1 - 1, Z1, Z1, -1, 0, 1);         /* WF         */
    PIECE(0, 0, 1, Y1 - 1, 0, 0, -1, 0, -1);          /* WB         */
    PIECE(X1, X1, 1, Y1 - 1, Z1, Z1, 1, 0, 1);        /* EF         */
    PIECE(X1, X1, 1, Y1 - 1, 0, 0, 1, 0, -1);         /* EB         */
    PIECE(0, 0, Y1, Y1, 0, 0, -1, 1, -1);             /* N-face NW corner */
    PIECE(X1, X1, Y1, Y1, 0, 0, 1, 1, -1);            /* N-face NE        */
    PIECE(0, 0, Y1, Y1, Z1, Z1, -1, 1, 1);            /* N-face SW        */
    PIECE(X1, X1, Y1, Y1, Z1, Z1, 1, 1, 1);           /* N-face SE        */
    PIECE(0, 0, 0, 0, 0, 0, -1, -1, -1);              /* S-face NW        */
    PIECE(X1, X1, 0, 0, 0, 0, 1, -1, -1);             /* S-face NE        */
    PIECE(0, 0, 0, 0, Z1, Z1, -1, -1, 1);             /* S-face SW        */
    PIECE(X1, X1, 0, 0, Z1, Z1, 1, -1, 1);            /* S-face SE        */
  }
#undef PIECE
  lbmx_box clip = {0, X1, 0, Y1, 0, Z1}, b;
  if (!intersect(box, &clip, &b)) return 0;
  /* validate first so a failing call adds nothing */
  for (int k = 0; k < np; ++k) {
    lbmx_box o;
    if (intersect(&b, &pb[k], &o) && !bc_supported(p, bc->type, pn[k])) return LBMX_ERR_SUP;
  }
  for (int k = 0; k < np; ++k) {
    lbmx_box o;
    if (!intersect(&b, &pb[k], &o)) continue;
    if (p->nbc == p->cap) {
      p->cap = p->cap ? 2 * p->cap : 16;
      p->bcs = (orc_bc*)realloc(p->bcs, p->cap * sizeof(orc_bc));
    }
    orc_bc* e = &p->bcs[p->nbc++];
    e->type = bc->type;
    e->nrm[0] = pn[k][0]; e->nrm[1] = pn[k][1]; e->nrm[2] = pn[k][2];
    e->box = o;
    e->d = *bc;
    e->prof = NULL;
    e->tfac = NULL;
    e->tbox = *box;
    e->ncomp = ncomp;
    e->nt = nt;
    if (prof) {
      e->prof = (double*)malloc(nprof * sizeof(double));
      memcpy(e->prof, prof, nprof * sizeof(double));
    }
    if (tfac) {
      e->tfac = (double*)malloc(nt * sizeof(double));
      memcpy(e->tfac, tfac, nt * sizeof(double));
    }
  }
  return 0;
}

int orc_num_boundaries(const orc_problem* p) { return p->nbc; }

int orc_boundary_info(const orc_problem* p, int k, int* type, int* nrm, int* box) {
  if (k < 0 || k >= p->nbc) return LBMX_ERR_ARG;
  const orc_bc* e = &p->bcs[k];
  *type = e->type;
  nrm[0] = e->nrm[0]; nrm[1] = e->nrm[1]; nrm[2] = e->nrm[2];
  box[0] = e->box.x0; box[1] = e->box.x1; box[2] = e->box.y0;
  box[3] = e->box.y1; box[4] = e->box.z0; box[5] = e->box.z1;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* collision                                                                  */
/* ------------------------------------------------------------------------- */
static double row_meq(const orc_row* r, double rho, const double* v, double rho_s) {
  double m = r->a * rho + r->b[0] * v[0] + r->b[1] * v[1] + r->b[2] * v[2];
  double quad = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) quad += r->C[a][b] * v[a] * v[b];
  return m + quad / rho_s;
}

static double row_s(const orc_row* r, double omega) {
  return r->rate == 0 ? 0. : (r->rate == 1 ? omega : r->s);
}

/* incompressible (IncompressibleFeq2d/3d, src/dynamics/equilibriumDistributions2d.hh:24-40,
   equilibriumDistributions3d.hh:25-42) or standard (StandardFeq*, :6-22) equilibrium */
static double feq(const orc_problem* p, int q, double rho, const double* v) {
  double cu = p->ex[q] * v[0] + p->ey[q] * v[1] + p->ez[q] * v[2];
  double u2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  if (p->incompressible) return p->w[q] * (rho + 3. * cu + 4.5 * cu * cu - 1.5 * u2);
  return p->w[q] * rho * (1. + 3. * cu + 4.5 * cu * cu - 1.5 * u2);
}

/* Guo source with F = -alpha v: w_q[3(c-v).F + 9(c.v)(c.F)] */
static double guo_src(const orc_problem* p, int q, const double* v, double alpha) {
  double cu = p->ex[q] * v[0] + p->ey[q] * v[1] + p->ez[q] * v[2];
  double u2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  return -alpha * p->w[q] * (3. * cu + 9. * cu * cu - 3. * u2);
}

static void macros_flow(const orc_problem* p, const double* f, double* rho, double* j) {
  double r = 0, jx = 0, jy = 0, jz = 0;
  for (int q = 0; q < p->Qf; ++q) {
    r += f[q]; jx += p->ex[q] * f[q]; jy += p->ey[q] * f[q]; jz += p->ez[q] * f[q];
  }
  *rho = r; j[0] = jx; j[1] = jy; j[2] = jz;
}

static double design_alpha(const orc_problem* p, double gamma) {
  return p->d.alpha_max * bp(p->d.interp_q, gamma);
}

/* base collision of the flow part: BGK / MRT (incompressible or standard), with the
   equilibrium velocity v and an optional Guo source (Brinkman). out may alias nothing. */
/* IncompressibleCascaded2d<D2Q9> (src/dynamics/cascaded2d.hh:40-95): relaxation written as
   k = K(f, rho, u), f' = f + T k (momentToDistribution, :77-95), with the incompressible
   macros u = j (src/dynamics/isothermalMacros2d.hh:44-67). Shear moments k1, k2 relax with
   omega, the others with rate 1. Directions follow D2Q9_EX/EY above. */
static const double CASC_T[9][6] = {
    {-4, 0, 0, 0, 0, 4},  {2, 0, 1, 1, -1, 1},  {-1, -1, 0, -2, 0, -2},
    {2, 0, -1, 1, 1, 1},  {-1, 1, 0, 0, -2, -2}, {2, 0, 1, -1, 1, 1},
    {-1, -1, 0, 2, 0, -2}, {2, 0, -1, -1, -1, 1}, {-1, 1, 0, 0, 2, -2}};

/* k0..k5 as cascaded2d.hh:44-74 computes them */
static void casc_k(const double* f, double r, double x, double y, double om, double* k) {
  double u2 = x * x + y * y;
  k[0] = (r * u2 - f[8] - f[6] - f[2] - f[4] - 2. * (f[1] + f[3] + f[7] + f[5] - r / 3.)) / 12.;
  k[1] = om * (f[6] + f[2] - f[8] - f[4] + r * (x * x - y * y)) / 4.;
  k[2] = om * ((f[7] + f[3] - f[5] - f[1]) - r * x * y) / 4.;
  k[3] = (-((f[1] + f[3] - f[7] - f[5] - 2. * r * x * x * y + y * (r - f[6] - f[2] - f[0])) / 4. +
            0.5 * x * (f[7] - f[5] - f[1] + f[3]))) +
         0.5 * y * (-3. * k[0] - k[1]) + 2. * x * k[2];
  k[4] = (-((f[3] + f[5] - f[1] - f[7] - 2. * r * y * y * x + x * (r - f[4] - f[8] - f[0])) / 4. +
            0.5 * y * (f[7] + f[3] - f[1] - f[5]))) +
         0.5 * x * (-3. * k[0] + k[1]) + 2. * y * k[2];
  k[5] = (0.25 * (r / 9. - f[7] - f[5] - f[1] - f[3] +
                  2. * (x * (f[7] - f[5] + f[1] - f[3]) + y * (f[7] + f[5] - f[1] - f[3])) +
                  4. * x * y * (f[5] - f[7] + f[1] - f[3]) -
                  x * x * (f[6] + f[7] + f[5] + f[2] + f[1] + f[3]) +
                  y * y * (3. * r * x * x - f[8] - f[7] - f[5] - f[1] - f[3] - f[4]))) -
         2. * k[0] - 2. * x * k[4] - 2. * y * k[3] + 4. * x * y * k[2] - 1.5 * k[0] * u2 +
         0.5 * k[1] * (x * x - y * y);
}

/* Transpose of the cascaded Jacobian (reverse-mode derivative of casc_k, derived by hand;
   the reference's ReverseIncompressibleCascaded2d, src/adjointDynamics/reverseCascaded2d.hh,
   is not this transpose): lf = lg + (dk/df)^T T^T lg for the explicit f terms; the parts
   through rho and u are returned in *rb, vb[0..1]. */
static void casc_reverse(const double* f, double r, double x, double y, double om, int as_coded,
                         const double* lg, double* lf, double* rb, double* vb) {
  double k[6];
  casc_k(f, r, x, y, om, k);
  double kb[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 6; ++i)
    for (int q = 0; q < 9; ++q) kb[i] += CASC_T[q][i] * lg[q];
  double fb[9];
  for (int q = 0; q < 9; ++q) fb[q] = 0.;
  double r_b = 0, x_b = 0, y_b = 0;
  const double u2 = x * x + y * y;
  /* k5 */
  {
    const double b = kb[5];
    const double A5 = f[7] - f[5] + f[1] - f[3], B5 = f[7] + f[5] - f[1] - f[3];
    const double C5 = f[5] - f[7] + f[1] - f[3];
    const double E5 = f[6] + f[7] + f[5] + f[2] + f[1] + f[3];
    const double F5 = f[8] + f[7] + f[5] + f[1] + f[3] + f[4];
    const double Sb = 0.25 * b;
    r_b += Sb * (1. / 9. + 3. * x * x * y * y);
    x_b += Sb * (2. * A5 + 4. * y * C5 - 2. * x * E5 + 6. * r * x * y * y);
    y_b += Sb * (2. * B5 + 4. * x * C5 + 2. * y * (3. * r * x * x - F5));
    const double Ab = Sb * 2. * x, Bb = Sb * 2. * y, Cb = Sb * 4. * x * y, Eb = -Sb * x * x, Fb = -Sb * y * y;
    fb[7] += -Sb + Ab + Bb - Cb + Eb + Fb;
    fb[5] += -Sb - Ab + Bb + Cb + Eb + Fb;
    fb[1] += -Sb + Ab - Bb + Cb + Eb + Fb;
    fb[3] += -Sb - Ab - Bb - Cb + Eb + Fb;
    fb[6] += Eb; fb[2] += Eb;
    fb[8] += Fb; fb[4] += Fb;
    kb[0] += b * (-2. - 1.5 * u2);
    kb[4] += b * (-2. * x);
    kb[3] += b * (-2. * y);
    kb[2] += b * (4. * x * y);
    kb[1] += b * (0.5 * (x * x - y * y));
    x_b += b * (-2. * k[4] + 4. * y * k[2] - 3. * k[0] * x + k[1] * x);
    y_b += b * (-2. * k[3] + 4. * x * k[2] - 3. * k[0] * y - k[1] * y);
    if (as_coded) {
      /* ReverseIncompressibleCascaded2d (src/adjointDynamics/reverseCascaded2d.hh:27-133) is this
         transpose except in one place: the coefficients -2 k4 and -2 k3 of dk5/du_x and dk5/du_y
         (the adjoint statements at :65 and :69) use its own recomputation of k3 and k4 (:45-52),
         which lacks the forward's /4 on every term but the r u^2 one (cascaded2d.hh:53-62). That
         recomputation exceeds the forward's k3 and k4 by d3 and d4 below, so the reference's
         adjoint of u_x and u_y picks up -2 d4 b and -2 d3 b; its adjoint statements for k3 and
         k4 themselves carry the forward's /4 (:81, :94) and need no change. */
      const double d3 = -0.75 * (f[1] + f[3] - f[7] - f[5] + y * (r - f[0] - f[2] - f[6]));
      const double d4 = -0.75 * (f[3] + f[5] - f[1] - f[7] + x * (r - f[0] - f[4] - f[8]));
      x_b += b * (-2. * d4);
      y_b += b * (-2. * d3);
    }
  }
  /* k4 = -(P4/4 + y Q4/2) + x(-3 k0 + k1)/2 + 2 y k2 */
  {
    const double b = kb[4];
    const double Q4 = f[7] + f[3] - f[1] - f[5];
    const double Pb = -0.25 * b, Qb = -0.5 * y * b;
    fb[3] += Pb; fb[5] += Pb; fb[1] -= Pb; fb[7] -= Pb;
    r_b += Pb * (-2. * y * y * x + x);
    y_b += Pb * (-4. * r * x * y);
    x_b += Pb * (-2. * r * y * y + (r - f[4] - f[8] - f[0]));
    fb[4] -= Pb * x; fb[8] -= Pb * x; fb[0] -= Pb * x;
    fb[7] += Qb; fb[3] += Qb; fb[1] -= Qb; fb[5] -= Qb;
    y_b += b * (-0.5 * Q4 + 2. * k[2]);
    x_b += b * (0.5 * (-3. * k[0] + k[1]));
    kb[0] += b * (-1.5 * x);
    kb[1] += b * (0.5 * x);
    kb[2] += b * (2. * y);
  }
  /* k3 = -(P3/4 + x Q3/2) + y(-3 k0 - k1)/2 + 2 x k2 */
  {
    const double b = kb[3];
    const double Q3 = f[7] - f[5] - f[1] + f[3];
    const double Pb = -0.25 * b, Qb = -0.5 * x * b;
    fb[1] += Pb; fb[3] += Pb; fb[7] -= Pb; fb[5] -= Pb;
    r_b += Pb * (-2. * x * x * y + y);
    x_b += Pb * (-4. * r * x * y);
    y_b += Pb * (-2. * r * x * x + (r - f[6] - f[2] - f[0]));
    fb[6] -= Pb * y; fb[2] -= Pb * y; fb[0] -= Pb * y;
    fb[7] += Qb; fb[5] -= Qb; fb[1] -= Qb; fb[3] += Qb;
    x_b += b * (-0.5 * Q3 + 2. * k[2]);
    y_b += b * (0.5 * (-3. * k[0] - k[1]));
    kb[0] += b * (-1.5 * y);
    kb[1] += b * (-0.5 * y);
    kb[2] += b * (2. * x);
  }
  /* k2 */
  {
    const double b = 0.25 * om * kb[2];
    fb[7] += b; fb[3] += b; fb[5] -= b; fb[1] -= b;
    r_b -= b * x * y; x_b -= b * r * y; y_b -= b * r * x;
  }
  /* k1 */
  {
    const double b = 0.25 * om * kb[1];
    fb[6] += b; fb[2] += b; fb[8] -= b; fb[4] -= b;
    r_b += b * (x * x - y * y); x_b += b * 2. * r * x; y_b -= b * 2. * r * y;
  }
  /* k0 */
  {
    const double b = kb[0] / 12.;
    fb[8] -= b; fb[6] -= b; fb[2] -= b; fb[4] -= b;
    fb[1] -= 2. * b; fb[3] -= 2. * b; fb[7] -= 2. * b; fb[5] -= 2. * b;
    r_b += b * (u2 + 2. / 3.); x_b += b * 2. * r * x; y_b += b * 2. * r * y;
  }
  for (int q = 0; q < 9; ++q) lf[q] = lg[q] + fb[q];
  *rb += r_b;
  vb[0] += x_b;
  vb[1] += y_b;
}

/* DarcyMRT2d forcing (src/dynamics/darcyMRT2d.hh:38-58): added to the MRT output, with the
   unshifted incompressible macros u = j and F = -amax BP(gamma) rho u + body force. */
static void darcy_forcing(double om, double ux, double uy, double Fx, double Fy, double* f) {
  const double s1 = 1.4, s2 = 1.4;
  const double o3 = 1. / 3., o6 = 1. / 6., o12 = 1. / 12., o24 = 1. / 24.;
  f[0] += o3 * (s1 + s2 - 4.) * (Fx * ux + Fy * uy);
  f[1] += o24 * ((-4. * s1 + 2. * s2 + 4.) * ux + 2. + (3. * om - 6.) * uy) * Fx -
          o6 * Fy * ((-0.75 * om + 1.5) * ux + 0.5 + (s1 - 0.5 * s2 - 1.) * uy);
  f[2] += o12 * (-4. + (s1 - 2. * s2 - 3. * om + 8.) * uy) * Fy + o12 * Fx * ux * (s1 - 2. * s2 + 3. * om - 4.);
  f[3] += o24 * ((-4. * s1 + 2. * s2 + 4.) * ux - 2. + (-3. * om + 6.) * uy) * Fx -
          o6 * Fy * ((0.75 * om - 1.5) * ux + 0.5 + (s1 - 0.5 * s2 - 1.) * uy);
  f[4] += o12 * (-4. + (s1 - 2. * s2 - 3. * om + 8.) * ux) * Fx + o12 * Fy * uy * (s1 - 2. * s2 + 3. * om - 4.);
  f[5] += o24 * ((-4. * s1 + 2. * s2 + 4.) * ux - 2. + (3. * om - 6.) * uy) * Fx -
          o6 * Fy * ((-0.75 * om + 1.5) * ux - 0.5 + (s1 - 0.5 * s2 - 1.) * uy);
  f[6] += o12 * (4. + (s1 - 2. * s2 - 3. * om + 8.) * uy) * Fy + o12 * Fx * ux * (s1 - 2. * s2 + 3. * om - 4.);
  f[7] += o24 * ((-4. * s1 + 2. * s2 + 4.) * ux + 2. + (-3. * om + 6.) * uy) * Fx -
          o6 * Fy * ((0.75 * om - 1.5) * ux - 0.5 + (s1 - 0.5 * s2 - 1.) * uy);
  f[8] += o12 * (4. + (s1 - 2. * s2 - 3. * om + 8.) * ux) * Fx + o12 * Fy * uy * (s1 - 2. * s2 + 3. * om - 4.);
}

static void base_collide(const orc_problem* p, const double* f, double* out, double rho,
                         const double* v, double alpha, int forced) {
  const int Q = p->Qf;
  const double om = p->d.omega;
  double rho_s = p->incompressible ? 1. : rho;
  if (p->d.collision == LBMX_CASCADED) {
    double k[6];
    casc_k(f, rho, v[0], v[1], om, k);
    for (int q = 0; q < 9; ++q) {
      double acc = f[q];
      for (int i = 0; i < 6; ++i) acc += CASC_T[q][i] * k[i];
      out[q] = acc;
    }
    return;
  }
  if (p->d.collision == LBMX_BGK || p->d.collision == LBMX_BGK_STANDARD) {
    /* IncompressibleBGK2d (src/dynamics/incompressibleBGK2d.hh:38-47) */
    for (int q = 0; q < Q; ++q) {
      out[q] = f[q] * (1. - om) + om * feq(p, q, rho, v);
      if (forced) out[q] += (1. - 0.5 * om) * guo_src(p, q, v, alpha);
    }
    return;
  }
  /* MRT: f' = f - M^-1 S (M f - m_eq)   (src/dynamics/incompressibleMRT2d.hh:44-86,
     src/dynamics/MRT3d.hh:44-176, :222-359) */
  const orc_mrt* m = &p->mrt;
  double dm[QMAX], srcm[QMAX], src[QMAX];
  /* the equilibrium moments take the momentum: j itself (incompressible), rho v for
     StandardMRT3d, whose quadratic terms are j_a j_b / rho (src/dynamics/MRT3d.hh:270-285) */
  double vm[3];
  for (int a = 0; a < 3; ++a) vm[a] = p->incompressible ? v[a] : rho * v[a];
  if (forced)
    for (int q = 0; q < Q; ++q) src[q] = guo_src(p, q, v, alpha);
  for (int k = 0; k < Q; ++k) {
    double mk = 0;
    dm[k] = 0;
    srcm[k] = 0;
    if (m->row[k].rate != 0) {
      for (int q = 0; q < Q; ++q)
        if (m->M[k][q] != 0.) mk += m->M[k][q] * f[q];
      dm[k] = row_s(&m->row[k], om) * (mk - row_meq(&m->row[k], rho, vm, rho_s)) * m->inv_norm2[k];
    }
    if (forced) {
      double sk = 0;
      for (int q = 0; q < Q; ++q) sk += m->M[k][q] * src[q];
      srcm[k] = (1. - 0.5 * row_s(&m->row[k], om)) * sk * m->inv_norm2[k];
    }
  }
  for (int q = 0; q < Q; ++q) {
    double acc = f[q];
    for (int k = 0; k < Q; ++k)
      if (m->M[k][q] != 0. && (forced || m->row[k].rate != 0)) acc += m->M[k][q] * (srcm[k] - dm[k]);
    out[q] = acc;
  }
}

/* thermal part: ThermalObstacleCollision<ThermalMRT2d<D2Q5_TMRT>, Ramp>
   (src/dynamics/thermalObstacleCollision.hh:47-53, src/dynamics/thermalMRT2d.hh:76-101) */
static double thermal_omega(const orc_problem* p, double gamma) {
  double diff = p->d.thermal_diffusivity * ramp(p->d.ramp_p, p->d.ramp_ck, gamma);
  return 1. / (5. * diff + 0.5);
}

static void thermal_collide(const orc_problem* p, const double* g, double* out, const double* j,
                            double gamma, double* Tout) {
  const orc_mrt* m = &p->tmrt;
  double T = 0;
  for (int q = 0; q < 5; ++q) T += g[q];
  double omT = thermal_omega(p, gamma);
  double dm[5];
  for (int k = 0; k < 5; ++k) {
    double mk = 0;
    for (int q = 0; q < 5; ++q) mk += m->M[k][q] * g[q];
    const orc_row* r = &m->row[k];
    double meq = T * (r->a + r->b[0] * j[0] + r->b[1] * j[1]);
    dm[k] = row_s(r, omT) * (mk - meq) * m->inv_norm2[k];
  }
  for (int q = 0; q < 5; ++q) {
    double acc = g[q];
    for (int k = 0; k < 5; ++k) acc -= m->M[k][q] * dm[k];
    out[q] = acc;
  }
  *Tout = T;
}

void orc_collide_node(const orc_problem* p, const double* fin, double* fout, double* mac,
                      double gamma) {
  double rho, j[3];
  macros_flow(p, fin, &rho, j);
  mac[0] = rho; mac[1] = j[0]; mac[2] = j[1]; mac[3] = j[2]; mac[4] = 0;
  int design = p->d.design;
  if (design == LBMX_DESIGN_BRINKMAN) {
    double alpha = design_alpha(p, gamma);
    double v[3];
    double beta = p->incompressible ? 1. / (1. + 0.5 * alpha) : 1. / (rho * (1. + 0.5 * alpha));
    for (int a = 0; a < 3; ++a) v[a] = beta * j[a];
    base_collide(p, fin, fout, rho, v, alpha, 1);
  } else if (design == LBMX_DESIGN_DARCY) {
    base_collide(p, fin, fout, rho, j, 0., 0);
    const double im = -p->d.alpha_max * bp(p->d.interp_q, gamma) * rho;
    darcy_forcing(p->d.omega, j[0], j[1], im * j[0] + p->d.body_force[0], im * j[1] + p->d.body_force[1], fout);
  } else {
    double v[3];
    for (int a = 0; a < 3; ++a) v[a] = p->incompressible ? j[a] : j[a] / rho;
    base_collide(p, fin, fout, rho, v, 0., 0);
  }
  if (design == LBMX_DESIGN_PBB) {
    /* PartialBouncebackCollision (src/dynamics/partialBouncebackCollision.hh:53-65) */
    double imult = 0.5 * p->d.alpha_max * bp(p->d.interp_q, gamma);
    for (int d = 1; d <= p->half; ++d) {
      double ft = fout[d];
      fout[d] += imult * (fout[d + p->half] - fout[d]);
      fout[d + p->half] += imult * (ft - fout[d + p->half]);
    }
  }
  if (p->Qt) {
    double T;
    thermal_collide(p, fin + p->Qf, fout + p->Qf, j, gamma, &T);
    mac[4] = T;
  }
}

/* ------------------------------------------------------------------------- */
/* reverse collision: exact transpose of orc_collide_node's Jacobian           */
/* (role of ReversePartialBouncebackCollision, ReverseIncompressibleBGK2d/MRT2d/MRT3d,
    ReverseThermal*: src/adjointDynamics/reversePartialBounceback.hh:18-41,
    reverseIncompressibleBGK2d.hh:48-127, reverseIncompressibleMRT2d.hh:14-157,
    reverseMRT3d.hh:14-389, reverseThermalMRT2d.hh:13-94)                      */
/* ------------------------------------------------------------------------- */
static void dmeq(const orc_row* r, double rho, const double* v, double rho_s, int standard,
                 double* drho, double* dv) {
  *drho = r->a;
  double quad = 0;
  for (int a = 0; a < 3; ++a) {
    double s = 0;
    for (int b = 0; b < 3; ++b) { s += (r->C[a][b] + r->C[b][a]) * v[b]; quad += r->C[a][b] * v[a] * v[b]; }
    dv[a] = r->b[a] + s / rho_s;
  }
  if (standard) *drho -= quad / (rho_s * rho_s);
}

void orc_reverse_collide_node(const orc_problem* p, const double* fin, double gamma,
                              const double* lam_post, double* lam_pre, double* dgamma) {
  const int Q = p->Qf;
  const double om = p->d.omega;
  double rho, j[3];
  macros_flow(p, fin, &rho, j);
  double lg[QMAX];
  for (int q = 0; q < p->Q; ++q) lg[q] = lam_post[q];
  double gam_bar = 0;

  /* flow: design coupling */
  double alpha = 0, beta = 1., v[3];
  int forced = p->d.design == LBMX_DESIGN_BRINKMAN;
  if (forced) alpha = design_alpha(p, gamma);
  if (!p->incompressible) { /* no reverse for the standard operators (as the reference) */
    for (int q = 0; q < p->Q; ++q) lam_pre[q] = 0.;
    return;
  }
  beta = forced ? 1. / (1. + 0.5 * alpha) : 1.;
  for (int a = 0; a < 3; ++a) v[a] = beta * j[a];

  if (p->d.design == LBMX_DESIGN_PBB) {
    double g[QMAX];
    base_collide(p, fin, g, rho, v, 0., 0);
    double imult = 0.5 * p->d.alpha_max * bp(p->d.interp_q, gamma);
    double mbar = 0;
    for (int d = 1; d <= p->half; ++d) {
      double a = lg[d], b = lg[d + p->half];
      mbar += (a - b) * (g[d + p->half] - g[d]);
      lg[d] = a + imult * (b - a);
      lg[d + p->half] = b + imult * (a - b);
    }
    gam_bar += mbar * 0.5 * p->d.alpha_max * bp_diff(p->d.interp_q, gamma);
  }

  double rho_bar = 0, v_bar[3] = {0, 0, 0}, alpha_bar = 0;
  double lf[QMAX];
  if (p->d.collision == LBMX_CASCADED) {
    casc_reverse(fin, rho, v[0], v[1], om, p->reference_reverse, lg, lf, &rho_bar, v_bar);
  } else if (p->d.collision == LBMX_BGK) {
    for (int q = 0; q < Q; ++q) {
      double cq[3] = {p->ex[q], p->ey[q], p->ez[q]};
      double cu = cq[0] * v[0] + cq[1] * v[1] + cq[2] * v[2];
      double u2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      lf[q] = (1. - om) * lg[q];
      double t = om * p->w[q] * lg[q];
      rho_bar += t;
      for (int a = 0; a < 3; ++a) v_bar[a] += t * (3. * cq[a] + 9. * cu * cq[a] - 3. * v[a]);
      if (forced) {
        double ks = (1. - 0.5 * om) * lg[q] * p->w[q];
        alpha_bar += -ks * (3. * cu + 9. * cu * cu - 3. * u2);
        for (int a = 0; a < 3; ++a) v_bar[a] += -alpha * ks * (3. * cq[a] + 18. * cu * cq[a] - 6. * v[a]);
      }
    }
  } else { /* MRT */
    const orc_mrt* m = &p->mrt;
    double a[QMAX], sigma[QMAX];
    for (int k = 0; k < Q; ++k) {
      double s = 0;
      for (int q = 0; q < Q; ++q) s += m->M[k][q] * lg[q];
      a[k] = s * m->inv_norm2[k];
    }
    for (int q = 0; q < Q; ++q) {
      double acc = lg[q], sg = 0;
      for (int k = 0; k < Q; ++k) {
        acc -= m->M[k][q] * row_s(&m->row[k], om) * a[k];
        sg += m->M[k][q] * (1. - 0.5 * row_s(&m->row[k], om)) * a[k];
      }
      lf[q] = acc;
      sigma[q] = sg;
    }
    for (int k = 0; k < Q; ++k) {
      double b = row_s(&m->row[k], om) * a[k];
      if (b == 0.) continue;
      double dr, dv[3];
      dmeq(&m->row[k], rho, v, 1., 0, &dr, dv);
      rho_bar += b * dr;
      for (int c = 0; c < 3; ++c) v_bar[c] += b * dv[c];
    }
    if (forced) {
      for (int q = 0; q < Q; ++q) {
        double cq[3] = {p->ex[q], p->ey[q], p->ez[q]};
        double cu = cq[0] * v[0] + cq[1] * v[1] + cq[2] * v[2];
        double u2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        double ks = sigma[q] * p->w[q];
        alpha_bar += -ks * (3. * cu + 9. * cu * cu - 3. * u2);
        for (int c = 0; c < 3; ++c) v_bar[c] += -alpha * ks * (3. * cq[c] + 18. * cu * cq[c] - 6. * v[c]);
      }
    }
  }
  /* v = beta(alpha) j */
  double j_bar[3];
  for (int a = 0; a < 3; ++a) j_bar[a] = beta * v_bar[a];
  if (forced) {
    alpha_bar += -0.5 * beta * beta * (v_bar[0] * j[0] + v_bar[1] * j[1] + v_bar[2] * j[2]);
    gam_bar += alpha_bar * p->d.alpha_max * bp_diff(p->d.interp_q, gamma);
  }

  /* thermal part */
  if (p->Qt) {
    const orc_mrt* m = &p->tmrt;
    const double* g = fin + Q;
    const double* lt = lam_post + Q;
    double T = 0;
    for (int q = 0; q < 5; ++q) T += g[q];
    double omT = thermal_omega(p, gamma);
    double a[5], T_bar = 0, om_bar = 0;
    for (int k = 0; k < 5; ++k) {
      double s = 0;
      for (int q = 0; q < 5; ++q) s += m->M[k][q] * lt[q];
      a[k] = s * m->inv_norm2[k];
    }
    for (int q = 0; q < 5; ++q) {
      double acc = lt[q];
      for (int k = 0; k < 5; ++k) acc -= m->M[k][q] * row_s(&m->row[k], omT) * a[k];
      lam_pre[Q + q] = acc;
    }
    for (int k = 0; k < 5; ++k) {
      const orc_row* r = &m->row[k];
      double sk = row_s(r, omT);
      double b = sk * a[k];
      T_bar += b * (r->a + r->b[0] * j[0] + r->b[1] * j[1]);
      j_bar[0] += b * T * r->b[0];
      j_bar[1] += b * T * r->b[1];
      if (r->rate == 1) {
        double mk = 0;
        for (int q = 0; q < 5; ++q) mk += m->M[k][q] * g[q];
        double meq = T * (r->a + r->b[0] * j[0] + r->b[1] * j[1]);
        om_bar -= a[k] * (mk - meq);
      }
    }
    for (int q = 0; q < 5; ++q) lam_pre[Q + q] += T_bar;
    /* omega_T = 1/(5 D R(gamma) + 0.5) */
    double dR = ramp_diff(p->d.ramp_p, p->d.ramp_ck, gamma);
    gam_bar += om_bar * (-omT * omT * 5. * p->d.thermal_diffusivity * dR);
  }

  for (int q = 0; q < Q; ++q)
    lam_pre[q] = lf[q] + rho_bar + p->ex[q] * j_bar[0] + p->ey[q] * j_bar[1] + p->ez[q] * j_bar[2];
  *dgamma += gam_bar;
}

/* ------------------------------------------------------------------------- */
/* boundary conditions (node-local affine maps, applied in place)             */
/* ------------------------------------------------------------------------- */
static void tangential(const orc_problem* p, const int* n, int i, int j, int k, double* s, double* t) {
  if (p->D == 2) { *s = n[0] != 0 ? j : i; *t = 0; return; }
  if (n[0] != 0) { *s = j; *t = k; }
  else if (n[1] != 0) { *s = i; *t = k; }
  else { *s = i; *t = j; }
}

/* value * a(i,j,k) * b(t) of a piece; comp: the velocity component of a per-component table */
static double bc_value(const orc_problem* p, const orc_bc* e, int t, int i, int j, int k, int comp) {
  const lbmx_bc_desc* d = &e->d;
  double s, tt, shape = 1.;
  tangential(p, e->nrm, i, j, k, &s, &tt);
  if (e->prof) {
    long long n = ((long long)(k - e->tbox.z0) * (e->tbox.y1 - e->tbox.y0 + 1) + (j - e->tbox.y0)) *
                      (e->tbox.x1 - e->tbox.x0 + 1) + (i - e->tbox.x0);
    shape = e->prof[n * e->ncomp + (e->ncomp > 1 ? comp : 0)];
  } else switch (d->profile) {
    case LBMX_PROFILE_PARABOLIC: {
      double R = d->radius[0], x = s - d->center[0];
      shape = (R * R - x * x) / (R * R);
    } break;
    case LBMX_PROFILE_PARABOLIC_DISK: {
      double R = d->radius[0], x = s - d->center[0], y = tt - d->center[1];
      double r2 = x * x + y * y;
      shape = r2 <= R * R ? 1. - r2 / (R * R) : 0.;
    } break;
    case LBMX_PROFILE_PARABOLIC_DUCT: {
      double x = (s - d->center[0]) / d->radius[0], y = (tt - d->center[1]) / d->radius[1];
      shape = (1. - x * x) * (1. - y * y);
    } break;
    default: break;
  }
  double tf = 1.;
  if (e->tfac)
    tf = e->tfac[t < e->nt ? t : e->nt - 1];
  else if (d->time_mode == LBMX_TIME_SINE)
    tf = d->t_offset + d->t_amp * sin(2. * ORC_PI * t / d->t_period + d->t_phase);
  return d->value * shape * tf;
}

/* generic on-grid bounce-back rule; matches OnGridBounceback2d (src/boundaryConditions/
   bounceback2d.cc:7-193) and 25 of the 26 D3Q19 orientations of bounceback3d.cc:8-511;
   the corner at (x0,y0,z0) (bounceback3d.cc:434-451) leaves f9 and f13 as they came in
   (its swap of 9 and 18 follows the copy 18 <- 9) where the rule gives f9=f18 and f13=f4,
   reproduced here. */
static void bounceback_node(const orc_problem* p, const int* n, double* f) {
  int Q = p->Qf;
  double out[QMAX];
  for (int q = 0; q < Q; ++q) out[q] = f[q];
  int quirk = p->D == 3 && n[0] == -1 && n[1] == -1 && n[2] == -1;
  for (int q = 1; q < Q; ++q) {
    int cn = p->ex[q] * n[0] + p->ey[q] * n[1] + p->ez[q] * n[2];
    if (cn < 0) out[q] = f[p->opp[q]];
    else if (cn == 0) out[q] = f[p->opp[q]];
  }
  if (quirk) { out[13] = f[13]; out[18] = f[9]; out[9] = f[9]; }
  for (int q = 0; q < Q; ++q) f[q] = out[q];
}

static double sum_idx(const double* f, const int* idx, int n) {
  double s = 0;
  for (int a = 0; a < n; ++a) s += f[idx[a]];
  return s;
}

/* the node's populations: f; neighbours nb1, nb2 (one and two nodes inwards) for Neumann */
static void bc_node(const orc_problem* p, const orc_bc* e, int t, int i, int j, int k, double* f,
                    const double* nb1, const double* nb2) {
  const int inc = p->incompressible;
  const int* n = e->nrm;
  switch (e->type) {
    case LBMX_BC_BOUNCEBACK: bounceback_node(p, n, f); break;
    case LBMX_BC_ZOUHE_VELOCITY: {
      double ux = bc_value(p, e, t, i, j, k, 0) * e->d.dir[0], uy = bc_value(p, e, t, i, j, k, 1) * e->d.dir[1];
      if (n[0] == -1 && inc) { /* IncompressibleZouHeVelocityWest2d (ZouHe2d.hh:29-44) */
        f[1] = 0.5 * (f[6] - f[2]) + f[5] + ux / 6. - 0.5 * uy;
        f[7] = 0.5 * (f[2] - f[6]) + f[3] + ux / 6. + 0.5 * uy;
        f[8] = f[4] + 2. / 3. * ux;
      } else if (n[0] == -1) { /* StandardZouHeVelocityWest2d (ZouHe2d.hh:149-174) */
        double rho = (f[0] + f[2] + f[6] + 2. * (f[3] + f[4] + f[5])) / (1. - ux);
        f[1] = f[5] + 0.5 * (f[6] - f[2]) + rho * ux / 6. - 0.5 * rho * uy;
        f[7] = f[3] + 0.5 * (f[2] - f[6]) + rho * ux / 6. + 0.5 * rho * uy;
        f[8] = f[4] + 2. / 3. * rho * ux;
      } else if (n[1] == 1) { /* IncompressibleZouHeVelocityNorth2d (ZouHe2d.hh:69-84) */
        f[1] = 0.5 * (f[4] - f[8]) + f[5] - uy / 6. + 0.5 * ux;
        f[2] = f[6] - 2. / 3. * uy;
        f[3] = 0.5 * (f[8] - f[4]) + f[7] - uy / 6. - 0.5 * ux;
      } else { /* IncompressibleZouHeVelocityEast2d (ZouHe2d.hh:109-124) */
        f[3] = f[7] + 0.5 * (f[6] - f[2]) - ux / 6. - 0.5 * uy;
        f[4] = f[8] - 2. / 3. * ux;
        f[5] = f[1] + 0.5 * (f[2] - f[6]) - ux / 6. + 0.5 * uy;
      }
    } break;
    case LBMX_BC_ZOUHE_PRESSURE: {
      double r0 = bc_value(p, e, t, i, j, k, 0);
      if (n[0] == 1) { /* east: ZouHe2d.hh:252-270 (inc), standard variant */
        double s = f[0] + f[2] + f[6] + 2. * (f[1] + f[7] + f[8]);
        double u0 = inc ? -r0 + s : -1. + s / r0;
        double sc = inc ? 1. : r0;
        f[4] = f[8] - 2. / 3. * sc * u0;
        f[3] = f[7] - sc * u0 / 6. + 0.5 * (f[6] - f[2]);
        f[5] = f[1] - sc * u0 / 6. + 0.5 * (f[2] - f[6]);
      } else { /* west */
        double s = f[0] + f[2] + f[6] + 2. * (f[3] + f[4] + f[5]);
        double u0 = inc ? r0 - s : 1. - s / r0;
        double sc = inc ? 1. : r0;
        f[1] = f[5] + sc * u0 / 6. + 0.5 * (f[6] - f[2]);
        f[7] = f[3] + sc * u0 / 6. + 0.5 * (f[2] - f[6]);
        f[8] = f[4] + 2. / 3. * sc * u0;
      }
    } break;
    case LBMX_BC_ZOUHE_NEUMANN: { /* IncompressibleZouHeNeumann{East,West}2d (ZouHe2d.hh:445-520) */
      double u1 = nb1[1] + nb1[7] + nb1[8] - (nb1[3] + nb1[4] + nb1[5]);
      double u2 = nb2[1] + nb2[7] + nb2[8] - (nb2[3] + nb2[4] + nb2[5]);
      double ux = (4. * u1 - u2) / 3.;
      if (n[0] == 1) {
        f[3] = f[7] + 0.5 * (f[6] - f[2]) - ux / 6.;
        f[4] = f[8] - 2. / 3. * ux;
        f[5] = f[1] + 0.5 * (f[2] - f[6]) - ux / 6.;
      } else {
        f[1] = 0.5 * (f[6] - f[2]) + f[5] + ux / 6.;
        f[7] = 0.5 * (f[2] - f[6]) + f[3] + ux / 6.;
        f[8] = f[4] + 2. / 3. * ux;
      }
    } break;
    case LBMX_BC_HECHT_VELOCITY: { /* {Inc,Std}HechtVelocityWest3d (Hecht3d.hh:26-97) */
      double ux = bc_value(p, e, t, i, j, k, 0) * e->d.dir[0], uy = bc_value(p, e, t, i, j, k, 1) * e->d.dir[1],
             uz = bc_value(p, e, t, i, j, k, 2) * e->d.dir[2];
      static const int known[9] = {0, 1, 2, 4, 5, 10, 11, 13, 14};
      static const int outg[5] = {3, 6, 8, 16, 18};
      double rho = inc ? 1. : (sum_idx(f, known, 9) + 2. * sum_idx(f, outg, 5)) / (1. - ux);
      double Nyx = 0.5 * (f[5] + f[11] + f[13] - (f[2] + f[4] + f[14])) - rho * uy / 3.;
      double Nzx = 0.5 * (f[10] + f[13] + f[14] - (f[1] + f[4] + f[5])) - rho * uz / 3.;
      f[7] = f[16] + rho * (ux - uz) / 6. + Nzx;
      f[9] = f[18] + rho * (ux - uy) / 6. + Nyx;
      f[12] = f[3] + rho * ux / 3.;
      f[15] = f[6] + rho * (ux + uz) / 6. - Nzx;
      f[17] = f[8] + rho * (ux + uy) / 6. - Nyx;
    } break;
    case LBMX_BC_HECHT_PRESSURE: { /* {Inc,Std}HechtPressureEast3d (Hecht3d.hh:129-204) */
      double r0 = bc_value(p, e, t, i, j, k, 0);
      static const int known[9] = {0, 1, 2, 4, 5, 10, 11, 13, 14};
      static const int outg[5] = {7, 9, 12, 15, 17};
      double s = sum_idx(f, known, 9) + 2. * sum_idx(f, outg, 5);
      double ux = inc ? s - r0 : -1. + s / r0;
      double sc = inc ? 1. : r0;
      double Nyx = 0.5 * (f[5] + f[11] + f[13] - (f[2] + f[4] + f[14]));
      double Nzx = 0.5 * (f[10] + f[13] + f[14] - (f[1] + f[4] + f[5]));
      f[3] = f[12] - sc * ux / 3.;
      f[6] = f[15] - sc * ux / 6. + Nzx;
      f[8] = f[17] - sc * ux / 6. + Nyx;
      f[16] = f[7] - sc * ux / 6. - Nzx;
      f[18] = f[9] - sc * ux / 6. - Nyx;
    } break;
    case LBMX_BC_THERMAL_BOUNCEBACK: { /* ThermalBounceback2d<.,D2Q5> (thermalBounceback2d.hh:273-490) */
      double* g = f + p->Qf;
      for (int q = 1; q < 5; ++q) {
        int cn = D2Q5_EX[q] * n[0] + D2Q5_EY[q] * n[1];
        if (cn < 0) g[q] = g[D2Q5_OPP[q]];
      }
    } break;
    case LBMX_BC_THERMAL_NEUMANN: { /* ThermalZouHeNeumann{East,West}2d<.,D2Q5> (thermalZouHe2d.hh:229-300) */
      double* g = f + p->Qf;
      const double* g1 = nb1 + p->Qf;
      const double* g2 = nb2 + p->Qf;
      double T1 = g1[0] + g1[1] + g1[2] + g1[3] + g1[4];
      double T2 = g2[0] + g2[1] + g2[2] + g2[3] + g2[4];
      double T = (4. * T1 - T2) / 3.;
      if (n[0] == 1) g[2] = T - g[0] - g[1] - g[3] - g[4];
      else g[4] = T - g[0] - g[1] - g[2] - g[3];
    } break;
    case LBMX_BC_THERMAL_DIRICHLET: { /* ThermalZouHeWest2d<.,D2Q5> (thermalZouHe2d.hh:62-93) */
      double* g = f + p->Qf;
      double T = bc_value(p, e, t, i, j, k, 0);
      int unk = 0;
      for (int q = 1; q < 5; ++q)
        if (D2Q5_EX[q] * n[0] + D2Q5_EY[q] * n[1] < 0) unk = q;
      double s = 0;
      for (int q = 0; q < 5; ++q)
        if (q != unk) s += g[q];
      g[unk] = T - s;
    } break;
  }
}

static size_t nidx(const orc_problem* p, int i, int j, int k) {
  return ((size_t)k * p->d.ny + j) * p->d.nx + i;
}

void orc_apply_bc(const orc_problem* p, int piece, int t, double* f) {
  const orc_bc* e = &p->bcs[piece];
  const int Q = p->Q;
  for (int k = e->box.z0; k <= e->box.z1; ++k)
    for (int j = e->box.y0; j <= e->box.y1; ++j)
      for (int i = e->box.x0; i <= e->box.x1; ++i) {
        const double *nb1 = NULL, *nb2 = NULL;
        if (e->type == LBMX_BC_ZOUHE_NEUMANN || e->type == LBMX_BC_THERMAL_NEUMANN) {
          int dir = -e->nrm[0];
          nb1 = f + nidx(p, i + dir, j, k) * Q;
          nb2 = f + nidx(p, i + 2 * dir, j, k) * Q;
        }
        bc_node(p, e, t, i, j, k, f + nidx(p, i, j, k) * Q, nb1, nb2);
      }
}

/* Transpose of one boundary piece. Every boundary map is affine in the populations, so its
   Jacobian columns are probed with unit vectors (exact: the coefficients are dyadic or
   the same rounding as the forward) and lam <- A^T lam, with the neighbour blocks of the
   Neumann condition added into the neighbours' adjoints. */
void orc_apply_bc_adjoint(const orc_problem* p, int piece, int t, double* lam) {
  const orc_bc* e = &p->bcs[piece];
  const int Q = p->Q;
  const int nblk = (e->type == LBMX_BC_ZOUHE_NEUMANN || e->type == LBMX_BC_THERMAL_NEUMANN) ? 3 : 1;
  for (int k = e->box.z0; k <= e->box.z1; ++k)
    for (int j = e->box.y0; j <= e->box.y1; ++j)
      for (int i = e->box.x0; i <= e->box.x1; ++i) {
        double base[QMAX], zero[3][QMAX], probe[3][QMAX];
        double A[3][QMAX][QMAX]; /* A[blk][row][col] */
        memset(zero, 0, sizeof(zero));
        memcpy(base, zero[0], sizeof(base));
        bc_node(p, e, t, i, j, k, base, zero[1], zero[2]);
        for (int b = 0; b < nblk; ++b)
          for (int c = 0; c < Q; ++c) {
            double out[QMAX];
            memset(probe, 0, sizeof(probe));
            probe[b][c] = 1.;
            memcpy(out, probe[0], sizeof(out));
            bc_node(p, e, t, i, j, k, out, probe[1], probe[2]);
            for (int r = 0; r < Q; ++r) A[b][r][c] = out[r] - base[r];
          }
        double* ln = lam + nidx(p, i, j, k) * Q;
        double lin[QMAX];
        memcpy(lin, ln, Q * sizeof(double));
        for (int c = 0; c < Q; ++c) {
          double s = 0;
          for (int r = 0; r < Q; ++r) s += A[0][r][c] * lin[r];
          ln[c] = s;
        }
        if (nblk == 3) {
          int dir = -e->nrm[0];
          for (int b = 1; b < 3; ++b) {
            double* lb = lam + nidx(p, i + b * dir, j, k) * Q;
            for (int c = 0; c < Q; ++c) {
              double s = 0;
              for (int r = 0; r < Q; ++r) s += A[b][r][c] * lin[r];
              lb[c] += s;
            }
          }
        }
      }
}

/* ------------------------------------------------------------------------- */
/* streaming: IsoStreamLoop2d (src/dynamics/isothermalLBLoops2d.hh:65-86) and  */
/* AdjointIsoStreamingLoop2d (src/adjointDynamics/adjointIsothermalLBLoops2d.hh:139-163) */
/* ------------------------------------------------------------------------- */
void orc_stream(const orc_problem* p, const double* fpost, double* fnew) {
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz, Q = p->Q;
#pragma omp parallel for num_threads(g_threads) collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t n = nidx(p, i, j, k);
        for (int q = 0; q < Q; ++q) {
          if (p->ex[q] == 0 && p->ey[q] == 0 && p->ez[q] == 0) { fnew[n * Q + q] = fpost[n * Q + q]; continue; }
          int si = i - p->ex[q], sj = j - p->ey[q], sk = k - p->ez[q];
          if (si >= 0 && si < nx && sj >= 0 && sj < ny && sk >= 0 && sk < nz)
            fnew[n * Q + q] = fpost[nidx(p, si, sj, sk) * Q + q];
          else
            fnew[n * Q + q] = fpost[n * Q + p->opp[q]];
        }
      }
}

void orc_adjoint_stream(const orc_problem* p, const double* lin, double* lout) {
  const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz, Q = p->Q;
#pragma omp parallel for num_threads(g_threads) collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t n = nidx(p, i, j, k);
        for (int q = 0; q < Q; ++q) {
          if (p->ex[q] == 0 && p->ey[q] == 0 && p->ez[q] == 0) { lout[n * Q + q] = lin[n * Q + q]; continue; }
          int si = i + p->ex[q], sj = j + p->ey[q], sk = k + p->ez[q];
          if (si >= 0 && si < nx && sj >= 0 && sj < ny && sk >= 0 && sk < nz)
            lout[n * Q + q] = lin[nidx(p, si, sj, sk) * Q + q];
          else
            lout[n * Q + q] = lin[n * Q + p->opp[q]];
        }
      }
}

/* ------------------------------------------------------------------------- */
/* functionals                                                                */
/* ------------------------------------------------------------------------- */
static double term_value(const orc_problem* p, int kind, const double* f) {
  double rho, j[3];
  macros_flow(p, f, &rho, j);
  switch (kind) {
    case 0: return rho;
    case 1: return p->incompressible ? j[0] : j[0] / rho;
    case 2: return (rho / 3. + 0.5 * (j[0] * j[0] + j[1] * j[1] + j[2] * j[2])) * j[0];
    case 3: {
      double T = 0;
      for (int q = 0; q < p->Qt; ++q) T += f[p->Qf + q];
      return j[0] * T;
    }
    case 4: {
      double T = 0;
      for (int q = 0; q < p->Qt; ++q) T += f[p->Qf + q];
      return T;
    }
  }
  return 0;
}

static void term_grad(const orc_problem* p, int kind, const double* f, double wgt, double* g) {
  double rho, j[3];
  macros_flow(p, f, &rho, j);
  for (int q = 0; q < p->Qf; ++q) {
    double cx = p->ex[q], cy = p->ey[q], cz = p->ez[q];
    switch (kind) {
      case 0: g[q] += wgt; break;
      case 1: g[q] += p->incompressible ? wgt * cx : wgt * (cx - j[0] / rho) / rho; break;
      case 2: {
        double cj = cx * j[0] + cy * j[1] + cz * j[2];
        double pt = rho / 3. + 0.5 * (j[0] * j[0] + j[1] * j[1] + j[2] * j[2]);
        g[q] += wgt * (j[0] / 3. + cj * j[0] + pt * cx);
      } break;
      case 3: {
        double T = 0;
        for (int r = 0; r < p->Qt; ++r) T += f[p->Qf + r];
        g[q] += wgt * cx * T;
      } break;
    }
  }
  if (kind == 3)
    for (int q = 0; q < p->Qt; ++q) g[p->Qf + q] += wgt * j[0];
  if (kind == 4)
    for (int q = 0; q < p->Qt; ++q) g[p->Qf + q] += wgt;
}

double orc_objective(const orc_problem* p, const double* f, int t) {
  (void)t;
  double J = 0;
  for (int m = 0; m < p->nterm; ++m) {
    const orc_term* tm = &p->terms[m];
    for (int k = tm->box.z0; k <= tm->box.z1; ++k)
      for (int j = tm->box.y0; j <= tm->box.y1; ++j)
        for (int i = tm->box.x0; i <= tm->box.x1; ++i)
          J += p->time_const * tm->weight * term_value(p, tm->kind, f + nidx(p, i, j, k) * p->Q);
  }
  return J;
}

void orc_objective_grad(const orc_problem* p, const double* f, int t, double* g) {
  (void)t;
  for (int m = 0; m < p->nterm; ++m) {
    const orc_term* tm = &p->terms[m];
    for (int k = tm->box.z0; k <= tm->box.z1; ++k)
      for (int j = tm->box.y0; j <= tm->box.y1; ++j)
        for (int i = tm->box.x0; i <= tm->box.x1; ++i) {
          size_t n = nidx(p, i, j, k) * p->Q;
          term_grad(p, tm->kind, f + n, p->time_const * tm->weight, g + n);
        }
  }
}

/* ------------------------------------------------------------------------- */
/* loops                                                                      */
/* ------------------------------------------------------------------------- */
/* IsoEquilibriumInitializationLoop2d (src/dynamics/isothermalLBLoops2d.hh:264-285),
   ThermalMRTFeq2d (src/dynamics/thermalEquilibriumDistributions2d.hh:28-47) */
static void init_node(const orc_problem* p, double* f, double rho, const double* u, double T) {
  for (int q = 0; q < p->Qf; ++q) f[q] = feq(p, q, rho, u);
  if (p->Qt) {
    double meq[5] = {T, u[0] * T, u[1] * T, -2. * T, 0.};
    for (int q = 0; q < 5; ++q) {
      double s = 0;
      for (int k = 0; k < 5; ++k) s += p->tmrt.M[k][q] * p->tmrt.inv_norm2[k] * meq[k];
      f[p->Qf + q] = s;
    }
  }
}

void orc_init_equilibrium(const orc_problem* p, double* f, double rho0, const double* u0, double T0) {
  for (size_t n = 0; n < p->nodes; ++n) init_node(p, f + n * p->Q, rho0, u0, T0);
}

void orc_macros(const orc_problem* p, const double* f, double* mac) {
  for (size_t n = 0; n < p->nodes; ++n) {
    double rho, j[3];
    macros_flow(p, f + n * p->Q, &rho, j);
    mac[5 * n] = rho; mac[5 * n + 1] = j[0]; mac[5 * n + 2] = j[1]; mac[5 * n + 3] = j[2];
    double T = 0;
    for (int q = 0; q < p->Qt; ++q) T += f[n * p->Q + p->Qf + q];
    mac[5 * n + 4] = T;
  }
}

static void collide_all(const orc_problem* p, double* f) {
  const long N = (long)p->nodes;
  const int Q = p->Q;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (long n = 0; n < N; ++n) {
    double out[QMAX], mac[5];
    orc_collide_node(p, f + n * Q, out, mac, p->gamma[n]);
    memcpy(f + n * Q, out, Q * sizeof(double));
  }
}

/* LBSolver::collideAndStream (src/latticeBoltzmann/LBSolver.cc:108-115, 163-206, 288-293);
   the objective of the new state follows computeMacroFunctional after the next collision
   (applications/RobustPump2d/robustPump2d.cc:242-248). */
void orc_forward(const orc_problem* p, double* f, int nsteps, int* t, double* J) {
  size_t len = p->nodes * p->Q;
  double* tmp = (double*)malloc(len * sizeof(double));
  double *a = f, *b = tmp;
  for (int s = 0; s < nsteps; ++s) {
    collide_all(p, a);
    orc_stream(p, a, b);
    double* sw = a; a = b; b = sw;
    ++*t;
    for (int k = 0; k < p->nbc; ++k) orc_apply_bc(p, k, *t, a);
    if (J) *J += orc_objective(p, a, *t);
  }
  if (a != f) memcpy(f, a, len * sizeof(double));
  free(tmp);
}

/* The design-gradient loop of applications/RobustPump2d/robustPump2d.cc:233-277 with every
   state stored (NaiveCheckpointing, src/topOpt/naiveCheckpointing.cc). Reverse step per
   AdjointLBSolver (src/adjointLatticeBoltzmann/adjointLBSolver.cc:28-122):
   adjointCollide -> computeCollideSource -> adjoint BCs -> adjoint pull-stream -> --t.
   The adjoint BCs are applied in reverse insertion order (the transpose of the forward
   sequence; identical to the reference's order whenever pieces are disjoint). */
int orc_gradient(const orc_problem* p, int nsteps, double rho0, const double* u0, double T0,
                 double* J, double* sens, double* adj) {
  const int Q = p->Q;
  if (!p->incompressible || p->d.design == LBMX_DESIGN_DARCY) return LBMX_ERR_SUP;
  size_t len = p->nodes * Q;
  double* states = (double*)malloc((size_t)(nsteps + 1) * len * sizeof(double));
  if (!states) return LBMX_ERR_MEM;
  orc_init_equilibrium(p, states, rho0, u0, T0);
  int t = 0;
  double Jsum = 0;
  for (int s = 1; s <= nsteps; ++s) {
    memcpy(states + s * len, states + (s - 1) * len, len * sizeof(double));
    orc_forward(p, states + s * len, 1, &t, &Jsum);
  }
  if (J) *J = Jsum;
  double* lam = (double*)calloc(len, sizeof(double));
  double* tmp = (double*)malloc(len * sizeof(double));
  for (size_t n = 0; n < p->nodes; ++n) sens[n] = 0.;
  for (int step = nsteps; step > 0; --step) {
    const double* f = states + (size_t)step * len;
    const long N = (long)p->nodes;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long n = 0; n < N; ++n) {
      double lp[QMAX], dg = 0;
      orc_reverse_collide_node(p, f + n * Q, p->gamma[n], lam + n * Q, lp, &dg);
      memcpy(lam + n * Q, lp, Q * sizeof(double));
      sens[n] += dg;
    }
    orc_objective_grad(p, f, step, lam);
    for (int b = p->nbc - 1; b >= 0; --b) orc_apply_bc_adjoint(p, b, step, lam);
    orc_adjoint_stream(p, lam, tmp);
    memcpy(lam, tmp, len * sizeof(double));
  }
  if (adj) memcpy(adj, lam, len * sizeof(double));
  free(lam);
  free(tmp);
  free(states);
  return 0;
}
