/*
 * ORACLE -- test infrastructure only (tests/ and bench.py's cpu_baseline use it).
 *
 * Plain-C restatement, on one rank, of the method of moving asymptotes as the reference
 * runs it (src/topOpt/MMA.cc, Aage's PETSc MMA with the reference's robust constants):
 * Svanberg's convex separable subproblem around the current design, solved through its
 * dual by a primal-dual interior-point Newton iteration on the m multipliers.
 *
 * Pinned in tests/test_mma_oracle.py by what the method must deliver, independent of its
 * formulas: iterates converge to the KKT point of test problems solved independently (a
 * separable quadratic under a linear budget, by bisection on its multiplier); move limits
 * and bounds hold; the subproblem's dual residual at exit is below the stopping tolerance;
 * the first two updates place the asymptotes at x -/+ asyminit (xmax - xmin).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int n, m, k;
  double asyminit, asymdec, asyminc;
  int con_mod, asym_type;
  double *a, *c, *y, *lam, *mu, *s, *b, *grad, *hess;
  double z;
  double *L, *U, *alpha, *beta, *p0, *q0, *pij, *qij, *xo1, *xo2, *lower, *upper; /* pij, qij: m x n */
  int newton_iters;
  double last_residual;
} omma;

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* MMA::MMA(n, m, x) (src/topOpt/MMA.cc:208-251): a = 0, c = 1000, asymptote factors 0.5 / 0.7 / 1.2 */
omma* omma_create(int n, int m) {
  omma* o = (omma*)calloc(1, sizeof(omma));
  o->n = n;
  o->m = m;
  o->asyminit = 0.5;
  o->asymdec = 0.7;
  o->asyminc = 1.2;
  double** vm[] = {&o->a, &o->c, &o->y, &o->lam, &o->mu, &o->b, &o->grad};
  for (size_t i = 0; i < sizeof(vm) / sizeof(vm[0]); ++i) *vm[i] = (double*)calloc((size_t)m, sizeof(double));
  o->s = (double*)calloc(2 * (size_t)m, sizeof(double));
  o->hess = (double*)calloc((size_t)m * m, sizeof(double));
  for (int j = 0; j < m; ++j) o->c[j] = 1000.0;
  double** vn[] = {&o->L, &o->U, &o->alpha, &o->beta, &o->p0, &o->q0, &o->xo1, &o->xo2, &o->lower, &o->upper};
  for (size_t i = 0; i < sizeof(vn) / sizeof(vn[0]); ++i) *vn[i] = (double*)calloc((size_t)n, sizeof(double));
  o->pij = (double*)calloc((size_t)n * m, sizeof(double));
  o->qij = (double*)calloc((size_t)n * m, sizeof(double));
  return o;
}

void omma_destroy(omma* o) {
  double* all[] = {o->a, o->c, o->y, o->lam, o->mu, o->s, o->b, o->grad, o->hess, o->L, o->U, o->alpha, o->beta,
                   o->p0, o->q0, o->pij, o->qij, o->xo1, o->xo2, o->lower, o->upper};
  for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) free(all[i]);
  free(o);
}

/* SetRobustConstants (MMA.cc:290-296): the a_j of the subproblem */
void omma_set_a(omma* o, const double* a) { memcpy(o->a, a, (size_t)o->m * sizeof(double)); }
void omma_set_asymptotes(omma* o, double init, double dec, double inc) {
  o->asyminit = init;
  o->asymdec = dec;
  o->asyminc = inc;
}
void omma_set_options(omma* o, int constraint_modification, int robust_asymptotes_type) {
  o->con_mod = constraint_modification;
  o->asym_type = robust_asymptotes_type;
}

/* Subproblem around x (GenSub, MMA.cc:526-653) with move limits clipped to [0, 1]
   (SetOuterMovelimit, MMA.cc:369-387). dgdx: m rows of n. */
static void gen_sub(omma* o, const double* x, const double* dfdx, const double* gx, const double* dgdx, double movlim) {
  const int n = o->n, m = o->m;
  for (int i = 0; i < n; ++i) {
    o->upper[i] = dmin(1.0, x[i] + movlim);
    o->lower[i] = dmax(0.0, x[i] - movlim);
  }
  o->k += 1;
  if (o->k < 3) {
    for (int i = 0; i < n; ++i) {
      o->L[i] = (x[i] - o->asyminit * o->upper[i]) + o->asyminit * o->lower[i];
      o->U[i] = (x[i] + o->asyminit * o->upper[i]) - o->asyminit * o->lower[i];
    }
  } else {
    for (int i = 0; i < n; ++i) {
      /* asymptotes move out when the last two steps kept their direction, in when they oscillated */
      const double trend = (x[i] - o->xo1[i]) * (o->xo1[i] - o->xo2[i]);
      const double g = trend < 0.0 ? o->asymdec : (trend > 0.0 ? o->asyminc : 1.0);
      double Li = x[i] - g * (o->xo1[i] - o->L[i]);
      double Ui = x[i] + g * (o->U[i] - o->xo1[i]);
      double span = dmax(1.0e-5, o->upper[i] - o->lower[i]);
      if (o->asym_type == 0) {
        Li = dmin(dmax(Li, x[i] - 10.0 * span), x[i] - 0.01 * span);
        Ui = dmin(dmax(Ui, x[i] + 0.01 * span), x[i] + 10.0 * span);
      } else {
        Li = dmin(dmax(Li, x[i] - 100.0 * span), x[i] - 1.0e-4 * span);
        Ui = dmin(dmax(Ui, x[i] + 1.0e-4 * span), x[i] + 100.0 * span);
        const double lo = o->lower[i] - 1.0e-5, hi = o->upper[i] + 1.0e-5;
        if (x[i] < lo) {
          Li = x[i] - (hi - x[i]) / 0.9;
          Ui = x[i] + (hi - x[i]) / 0.9;
        }
        if (x[i] > hi) {
          Li = x[i] - (x[i] - lo) / 0.9;
          Ui = x[i] + (x[i] - lo) / 0.9;
        }
      }
      o->L[i] = Li;
      o->U[i] = Ui;
    }
  }
  const double feps = 1.0e-6;
  for (int i = 0; i < n; ++i) {
    const double ux = o->U[i] - x[i], xl = x[i] - o->L[i], ul = o->U[i] - o->L[i];
    o->alpha[i] = dmax(o->lower[i], 0.9 * o->L[i] + 0.1 * x[i]);
    o->beta[i] = dmin(o->upper[i], 0.9 * o->U[i] + 0.1 * x[i]);
    const double dp = dmax(0.0, dfdx[i]), dm = dmax(0.0, -dfdx[i]), ad = fabs(dfdx[i]);
    o->p0[i] = (ux * ux) * (dp + 0.001 * ad + 0.5 * feps / ul);
    o->q0[i] = (xl * xl) * (dm + 0.001 * ad + 0.5 * feps / ul);
    for (int j = 0; j < m; ++j) {
      const double g = dgdx[(size_t)j * n + i];
      const double gp = dmax(0.0, g), gm = dmax(0.0, -g);
      if (o->con_mod) {
        o->pij[(size_t)j * n + i] = (ux * ux) * (gp + 0.001 * fabs(g) + 0.5 * feps / ul);
        o->qij[(size_t)j * n + i] = (xl * xl) * (gm + 0.001 * fabs(g) + 0.5 * feps / ul);
      } else {
        o->pij[(size_t)j * n + i] = (ux * ux) * gp;
        o->qij[(size_t)j * n + i] = (xl * xl) * gm;
      }
    }
  }
  for (int j = 0; j < m; ++j) {
    double sum = 0.0;
    for (int i = 0; i < n; ++i)
      sum += o->pij[(size_t)j * n + i] / (o->U[i] - x[i]) + o->qij[(size_t)j * n + i] / (x[i] - o->L[i]);
    o->b[j] = sum - gx[j];
  }
}

/* primal point of the dual multipliers (XYZofLAMBDA, MMA.cc:697-748): y, z and x(lam) */
static void primal_of_lambda(omma* o, double* x) {
  const int n = o->n, m = o->m;
  double la = 0.0;
  for (int j = 0; j < m; ++j) {
    if (o->lam[j] < 0.0) o->lam[j] = 0.0;
    o->y[j] = dmax(0.0, o->lam[j] - o->c[j]);
    la += o->lam[j] * o->a[j];
  }
  o->z = dmax(0.0, 10.0 * (la - 1.0));
  for (int i = 0; i < n; ++i) {
    double pl = o->p0[i], ql = o->q0[i];
    for (int j = 0; j < m; ++j) {
      pl += o->pij[(size_t)j * n + i] * o->lam[j];
      ql += o->qij[(size_t)j * n + i] * o->lam[j];
    }
    const double sp = sqrt(pl), sq = sqrt(ql);
    double xi = (sp * o->L[i] + sq * o->U[i]) / (sp + sq);
    if (xi < o->alpha[i]) xi = o->alpha[i];
    if (xi > o->beta[i]) xi = o->beta[i];
    x[i] = xi;
  }
}

/* sum_i pij / (U - x) + qij / (x - L) for every j (the constraint approximations at x) */
static void con_sums(const omma* o, const double* x, double* out) {
  const int n = o->n;
  for (int j = 0; j < o->m; ++j) {
    double sum = 0.0;
    for (int i = 0; i < n; ++i)
      sum += o->pij[(size_t)j * n + i] / (o->U[i] - x[i]) + o->qij[(size_t)j * n + i] / (x[i] - o->L[i]);
    out[j] = sum;
  }
}

/* dual Hessian (DualHess, MMA.cc:786-886) */
static void dual_hessian(omma* o, const double* x) {
  const int n = o->n, m = o->m;
  for (int j = 0; j < m * m; ++j) o->hess[j] = 0.0;
  double* pq = (double*)calloc((size_t)m, sizeof(double));
  for (int i = 0; i < n; ++i) {
    double pl = o->p0[i], ql = o->q0[i];
    const double ux = o->U[i] - x[i], xl = x[i] - o->L[i];
    for (int j = 0; j < m; ++j) {
      pl += o->pij[(size_t)j * n + i] * o->lam[j];
      ql += o->qij[(size_t)j * n + i] * o->lam[j];
      pq[j] = o->pij[(size_t)j * n + i] / (ux * ux) - o->qij[(size_t)j * n + i] / (xl * xl);
    }
    double d2 = -1.0 / (2.0 * pl / pow(ux, 3.0) + 2.0 * ql / pow(xl, 3.0));
    const double sp = sqrt(pl), sq = sqrt(ql);
    const double xp = (sp * o->L[i] + sq * o->U[i]) / (sp + sq);
    if (xp < o->alpha[i] || xp > o->beta[i]) d2 = 0.0;  /* the bound is active: no curvature */
    for (int j = 0; j < m; ++j) {
      const double t = pq[j] * d2;
      for (int l = 0; l < m; ++l) o->hess[j * m + l] += t * pq[l];
    }
  }
  free(pq);
  double la = 0.0;
  for (int j = 0; j < m; ++j) {
    if (o->lam[j] < 0.0) o->lam[j] = 0.0;
    la += o->lam[j] * o->a[j];
    if (o->lam[j] > o->c[j]) o->hess[j * m + j] += -1.0;
    o->hess[j * m + j] += -o->mu[j] / o->lam[j];
  }
  if (la > 0.0)
    for (int j = 0; j < m; ++j)
      for (int l = 0; l < m; ++l) o->hess[j * m + l] += -10.0 * o->a[j] * o->a[l];
  double tr = 0.0;
  for (int j = 0; j < m; ++j) tr += o->hess[j * m + j];
  double corr = 1e-4 * tr / m;
  if (-corr < 1.0e-7) corr = -1.0e-7;
  for (int j = 0; j < m; ++j) o->hess[j * m + j] += corr;
}

/* in-place LU without pivoting and the two triangular solves (MMA.cc:952-986) */
static void lu_solve(double* K, double* rhs, int nn) {
  for (int p = 0; p < nn - 1; ++p)
    for (int i = p + 1; i < nn; ++i) {
      K[i * nn + p] = K[i * nn + p] / K[p * nn + p];
      for (int j = p + 1; j < nn; ++j) K[i * nn + j] = K[i * nn + j] - K[i * nn + p] * K[p * nn + j];
    }
  for (int i = 1; i < nn; ++i) {
    double acc = 0.0;
    for (int j = 0; j < i; ++j) acc = acc - K[i * nn + j] * rhs[j];
    rhs[i] = rhs[i] + acc;
  }
  rhs[nn - 1] = rhs[nn - 1] / K[(nn - 1) * nn + (nn - 1)];
  for (int i = nn - 2; i >= 0; --i) {
    double acc = rhs[i];
    for (int j = i + 1; j < nn; ++j) acc = acc - K[i * nn + j] * rhs[j];
    rhs[i] = acc / K[i * nn + i];
  }
}

/* dual interior-point solve of the subproblem (SolveDIP, DualLineSearch, DualResidual,
   MMA.cc:655-695, 888-950); x receives the new design */
static void solve_dual(omma* o, double* x) {
  const int n = o->n, m = o->m;
  for (int j = 0; j < m; ++j) {
    o->lam[j] = o->c[j] / 2.0;
    o->mu[j] = 1.0;
  }
  const double tol = 1.0e-9 * sqrt((double)(m + n));
  double epsi = 1.0, err = 1.0;
  double* res = (double*)calloc(2 * (size_t)m, sizeof(double));
  o->newton_iters = 0;
  while (epsi > tol) {
    int loop = 0;
    while (err > 0.9 * epsi && loop < 100) {
      ++loop;
      ++o->newton_iters;
      primal_of_lambda(o, x);
      con_sums(o, x, o->grad);
      for (int j = 0; j < m; ++j) o->grad[j] += -o->b[j] - o->a[j] * o->z - o->y[j];
      for (int j = 0; j < m; ++j) o->grad[j] = -1.0 * o->grad[j] - epsi / o->lam[j];
      dual_hessian(o, x);
      lu_solve(o->hess, o->grad, m);
      for (int j = 0; j < m; ++j) o->s[j] = o->grad[j];
      for (int j = 0; j < m; ++j) o->s[m + j] = -o->mu[j] + epsi / o->lam[j] - o->s[j] * o->mu[j] / o->lam[j];
      /* step to 1/1.005 of the way to the boundary of lam, mu > 0 */
      double theta = 1.005;
      for (int j = 0; j < m; ++j) {
        if (theta < -1.01 * o->s[j] / o->lam[j]) theta = -1.01 * o->s[j] / o->lam[j];
        if (theta < -1.01 * o->s[j + m] / o->mu[j]) theta = -1.01 * o->s[j + m] / o->mu[j];
      }
      theta = 1.0 / theta;
      for (int j = 0; j < m; ++j) {
        o->lam[j] = o->lam[j] + theta * o->s[j];
        o->mu[j] = o->mu[j] + theta * o->s[j + m];
      }
      primal_of_lambda(o, x);
      con_sums(o, x, res);
      double nr = 0.0;
      for (int j = 0; j < m; ++j) {
        res[j] += -o->b[j] - o->a[j] * o->z - o->y[j] + o->mu[j];
        res[j + m] = o->mu[j] * o->lam[j] - epsi;
      }
      for (int j = 0; j < 2 * m; ++j) nr = dmax(nr, fabs(res[j]));
      err = nr;
    }
    o->last_residual = err;
    epsi = epsi * 0.1;
  }
  free(res);
}

/* MMA::Update (MMA.cc:485-524): subproblem at x, history shift, dual solve; x is overwritten */
void omma_update(omma* o, double* x, const double* dfdx, const double* gx, const double* dgdx, double movlim) {
  gen_sub(o, x, dfdx, gx, dgdx, movlim);
  memcpy(o->xo2, o->xo1, (size_t)o->n * sizeof(double));
  memcpy(o->xo1, x, (size_t)o->n * sizeof(double));
  solve_dual(o, x);
}

/* MMA::DesignChange (MMA.cc:389-414): max |xnew - xo1| */
double omma_design_change(const omma* o, const double* xnew) {
  double ch = 0.0;
  for (int i = 0; i < o->n; ++i) ch = dmax(ch, fabs(xnew[i] - o->xo1[i]));
  return ch;
}

void omma_state(const omma* o, double* L, double* U, double* lam, int* k, int* newton, double* residual) {
  if (L) memcpy(L, o->L, (size_t)o->n * sizeof(double));
  if (U) memcpy(U, o->U, (size_t)o->n * sizeof(double));
  if (lam) memcpy(lam, o->lam, (size_t)o->m * sizeof(double));
  if (k) *k = o->k;
  if (newton) *newton = o->newton_iters;
  if (residual) *residual = o->last_residual;
}
