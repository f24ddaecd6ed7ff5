/*
 * ORACLE -- test infrastructure only (tests/, smoke() and bench.py's cpu_baseline use it).
 *
 * Plain-C restatement of the reference's design pipeline on one rank: the design-box
 * mapping, the cone density filter as an assembled sparse matrix, the tanh projection and
 * the sensitivity chain. Built with -ffp-contract=off: every multiply and add is rounded.
 *
 * Pinned in tests/test_design_oracle.py by closed forms (R = 1 identity, the R = 2 cone
 * weights and denominators in 2D and 3D, edge and corner rows), invariants (constants are
 * preserved inside the box, rows outside the box are the identity), the dot-product test
 * between the filter and its transpose, and central differences of the whole chain.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* compressed rows of the filter matrix W: row i holds (col, weight) in ascending column order */
typedef struct {
  long long n;
  long long* start; /* n + 1 */
  long long* col;
  double* val;
} odf_csr;

static int in_box(int x, int y, int z, const int* b) {
  return x >= b[0] && x <= b[1] && y >= b[2] && y <= b[3] && z >= b[4] && z <= b[5];
}

/* DensityFilter2d::setUp / DensityFilter3d::setUp (src/topOpt/densityFilter2d.cc:62-146,
   densityFilter3d.cc:60-139): identity rows (MatDiagonalSet), then for rows inside the
   filter box the weights R - dist of the in-grid columns with dist <= R, inserted over the
   diagonal. The grid is non-periodic, so the ghost region of one rank is the whole grid. */
static odf_csr odf_assemble(int nx, int ny, int nz, int R, const int* box) {
  odf_csr m;
  const long long n = (long long)nx * ny * nz;
  m.n = n;
  m.start = (long long*)malloc((size_t)(n + 1) * sizeof(long long));
  long long alloc = 0;
  /* count first to size the arrays */
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (!in_box(i, j, k, box)) { alloc += 1; continue; }
        for (int zz = -R; zz <= R; ++zz) {
          if (k + zz < 0 || k + zz >= nz) continue;
          for (int yy = -R; yy <= R; ++yy) {
            if (j + yy < 0 || j + yy >= ny) continue;
            for (int xx = -R; xx <= R; ++xx) {
              if (i + xx < 0 || i + xx >= nx) continue;
              if (sqrt((double)(xx * xx + yy * yy + zz * zz)) <= (double)R) alloc += 1;
            }
          }
        }
      }
  m.col = (long long*)malloc((size_t)alloc * sizeof(long long));
  m.val = (double*)malloc((size_t)alloc * sizeof(double));
  long long p = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long row = ((long long)k * ny + j) * nx + i;
        m.start[row] = p;
        if (!in_box(i, j, k, box)) {
          m.col[p] = row;
          m.val[p] = 1.0;
          ++p;
          continue;
        }
        for (int zz = -R; zz <= R; ++zz) {
          if (k + zz < 0 || k + zz >= nz) continue;
          for (int yy = -R; yy <= R; ++yy) {
            if (j + yy < 0 || j + yy >= ny) continue;
            for (int xx = -R; xx <= R; ++xx) {
              if (i + xx < 0 || i + xx >= nx) continue;
              double dist = sqrt((double)(xx * xx + yy * yy + zz * zz));
              if (dist <= (double)R) {
                dist = (double)R - dist;
                m.col[p] = ((long long)(k + zz) * ny + (j + yy)) * nx + (i + xx);
                m.val[p] = dist;
                ++p;
              }
            }
          }
        }
      }
  m.start[n] = p;
  return m;
}

static void odf_free(odf_csr* m) {
  free(m->start);
  free(m->col);
  free(m->val);
}

static double* filled(long long n, double v) {
  double* a = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  for (long long i = 0; i < n; ++i) a[i] = v;
  return a;
}

/* y = W x, row by row in column order (MatMult) */
static void csr_mult(const odf_csr* m, const double* x, double* y) {
  for (long long r = 0; r < m->n; ++r) {
    double sum = 0.0;
    for (long long p = m->start[r]; p < m->start[r + 1]; ++p) sum += m->val[p] * x[m->col[p]];
    y[r] = sum;
  }
}

/* y = W^T x by scattering each row in turn (MatMultTranspose) */
static void csr_mult_transpose(const odf_csr* m, const double* x, double* y) {
  for (long long r = 0; r < m->n; ++r) y[r] = 0.0;
  for (long long r = 0; r < m->n; ++r) {
    const double a = x[r];
    for (long long p = m->start[r]; p < m->start[r + 1]; ++p) y[m->col[p]] += a * m->val[p];
  }
}

/* Denominator W 1 (densityFilter2d.cc:139-140) */
void odf_denominator(int nx, int ny, int nz, int R, const int* box, double* den) {
  odf_csr m = odf_assemble(nx, ny, nz, R, box);
  double* ones = filled(m.n, 1.0);
  csr_mult(&m, ones, den);
  free(ones);
  odf_free(&m);
}

/* DensityFilter2d::filterDesign (densityFilter2d.cc:33-41): out = (W x) / (W 1) */
void odf_density_forward(int nx, int ny, int nz, int R, const int* box, const double* x, double* out) {
  odf_csr m = odf_assemble(nx, ny, nz, R, box);
  double* den = filled(m.n, 0.0);
  double* ones = filled(m.n, 1.0);
  csr_mult(&m, ones, den);
  csr_mult(&m, x, out);
  for (long long i = 0; i < m.n; ++i) out[i] = out[i] / den[i];
  free(den);
  free(ones);
  odf_free(&m);
}

/* DensityFilter2d::filterSensitivities (densityFilter2d.cc:43-60): s <- W^T (s / (W 1)) */
void odf_density_transpose(int nx, int ny, int nz, int R, const int* box, const double* s, double* out) {
  odf_csr m = odf_assemble(nx, ny, nz, R, box);
  double* den = filled(m.n, 0.0);
  double* tmp = filled(m.n, 1.0);
  csr_mult(&m, tmp, den);
  for (long long i = 0; i < m.n; ++i) tmp[i] = s[i] / den[i];
  csr_mult_transpose(&m, tmp, out);
  free(den);
  free(tmp);
  odf_free(&m);
}

/* ProjectionFilter2d::filterDesign (projectionFilter2d.cc:37-42), inside the box only;
   with a per-node threshold field etaf (may be NULL) as VariableProjectionFilter2d::filterDesign
   (variableProjectionFilter2d.cc:39-63) */
void odf_project(int nx, int ny, int nz, const int* box, double beta, double eta, const double* etaf,
                 const double* x, double* out) {
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long r = ((long long)k * ny + j) * nx + i;
        const double e = etaf ? etaf[r] : eta;
        out[r] = in_box(i, j, k, box)
                     ? (tanh(beta * e) + tanh(beta * (x[r] - e))) / (tanh(beta * e) + tanh(beta * (1. - e)))
                     : x[r];
      }
}

/* ProjectionFilter2d::filterSensitivities, first part (projectionFilter2d.cc:66-75):
   s *= projDiff inside the box, projDiff at the density-filtered field x; per-node thresholds
   as VariableProjectionFilter2d::filterSensitivities (variableProjectionFilter2d.cc:65-104) */
void odf_project_sens(int nx, int ny, int nz, const int* box, double beta, double eta, const double* etaf,
                      const double* x, double* s) {
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long r = ((long long)k * ny + j) * nx + i;
        if (!in_box(i, j, k, box)) continue;
        const double e = etaf ? etaf[r] : eta;
        const double pd = ((1.0 - pow(tanh(beta * (x[r] - e)), 2.0)) * beta) /
                          (tanh(beta * e) + tanh(beta * (1.0 - e)));
        s[r] *= pd;
      }
}

/* FullToDesignMapping2d::mapDesignToFull (fullToDesignMapping2d.cc:43-51 index maps,
   :118-128 scatter): design box entries inserted into a copy of the background */
void odf_map_to_full(int nx, int ny, int nz, const int* dbox, const double* design, const double* background,
                     double* full) {
  const int nxd = dbox[1] - dbox[0] + 1, nyd = dbox[3] - dbox[2] + 1;
  memcpy(full, background, (size_t)nx * ny * nz * sizeof(double));
  for (int k = dbox[4]; k <= dbox[5]; ++k)
    for (int j = dbox[2]; j <= dbox[3]; ++j)
      for (int i = dbox[0]; i <= dbox[1]; ++i)
        full[((long long)k * ny + j) * nx + i] =
            design[((long long)(k - dbox[4]) * nyd + (j - dbox[2])) * nxd + (i - dbox[0])];
}

/* FullToDesignMapping2d::mapFullToDesign (fullToDesignMapping2d.cc:107-116) */
void odf_extract(int nx, int ny, int nz, const int* dbox, const double* full, double* design) {
  const int nxd = dbox[1] - dbox[0] + 1, nyd = dbox[3] - dbox[2] + 1;
  (void)nz;
  for (int k = dbox[4]; k <= dbox[5]; ++k)
    for (int j = dbox[2]; j <= dbox[3]; ++j)
      for (int i = dbox[0]; i <= dbox[1]; ++i)
        design[((long long)(k - dbox[4]) * nyd + (j - dbox[2])) * nxd + (i - dbox[0])] =
            full[((long long)k * ny + j) * nx + i];
}
