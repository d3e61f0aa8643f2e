/*
 * CPU FP64 oracle of the coherent-potential solver.  TEST INFRASTRUCTURE
 * ONLY: the checker of the HIP CPA loop, never part of the product.
 *
 *   orc_tata    skkr_kpar_v6/nrsp-scf/fullcpa.f:2-24   initial coherent
 *               t^-1 (average t-matrix): mc = mb (c mb + (1-c) ma)^-1 ma
 *   orc_cpasol  fullcpa.f:73-153  one update of the coherent t^-1 after
 *               Ginatempo and Staunton, J. Phys. F 18, 1827 (1988), with the
 *               three error measures and the species-projected tau
 *   orc_cpacor  fullcpa.f:25-72   cpasol over the sites, skipping sites of
 *               concentration 1 or where mc equals ma or mb (compmat.f,
 *               matops.f:51-80)
 *
 * n x n matrices, column-major, leading dimension n.  Inversions use the
 * column-pivot Gauss-Jordan of gjinv.f (orc_gjinv), products the loop order of
 * tripmt (orc_tripmt, matops.f:294-342).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;
int orc_gjinv(cplx *a, int n, int lda);
void orc_tripmt(const cplx *u, cplx *b, const cplx *ust, int n1, int n2, int ld);

static int compmat(const cplx *x, const cplx *y, int n, double tol) {
  double xn = 0.0, dn = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const cplx xx = x[i + j * n], dx = x[i + j * n] - y[i + j * n];
      xn += creal(xx) * creal(xx) + cimag(xx) * cimag(xx);
      dn += creal(dx) * creal(dx) + cimag(dx) * cimag(dx);
    }
  return sqrt(dn) < tol * sqrt(xn);
}

int orc_tata(double conc, int n, cplx *mc, const cplx *ma, const cplx *mb) {
  const double ca = conc, cb = 1.0 - ca;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) mc[i + j * n] = ca * mb[i + j * n] + cb * ma[i + j * n];
  int st = orc_gjinv(mc, n, n);
  orc_tripmt(mb, mc, ma, n, n, n);
  return st;
}

/* err[0..2] = error1, error2, error3 of fullcpa.f */
int orc_cpasol(double conc, int n, cplx *mc, const cplx *ma, const cplx *mb, const cplx *tauc, cplx *taua,
               cplx *taub, double *err) {
  const double ca = conc, cb = 1.0 - ca;
  const size_t nn = (size_t)n * n;
  cplx *da = (cplx *)malloc(sizeof(cplx) * nn), *db = (cplx *)malloc(sizeof(cplx) * nn),
       *xa = (cplx *)malloc(sizeof(cplx) * nn), *xb = (cplx *)malloc(sizeof(cplx) * nn),
       *xc = (cplx *)malloc(sizeof(cplx) * nn), *h = (cplx *)malloc(sizeof(cplx) * nn);
  int st = 0;
  for (size_t e = 0; e < nn; ++e) {
    da[e] = mc[e] - ma[e];
    db[e] = mc[e] - mb[e];
    xa[e] = da[e];
    xb[e] = db[e];
  }
  st |= orc_gjinv(xa, n, n);
  st |= orc_gjinv(xb, n, n);
  for (size_t e = 0; e < nn; ++e) {
    xa[e] -= tauc[e];
    xb[e] -= tauc[e];
  }
  st |= orc_gjinv(xa, n, n);
  st |= orc_gjinv(xb, n, n);
  double e1 = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t e = i + (size_t)j * n;
      xc[e] = ca * xa[e] + cb * xb[e];
      if (cabs(xc[e]) > e1) e1 = cabs(xc[e]);
    }
  memcpy(h, xc, sizeof(cplx) * nn);
  st |= orc_gjinv(h, n, n);
  for (size_t e = 0; e < nn; ++e) h[e] += tauc[e];
  st |= orc_gjinv(h, n, n);
  for (size_t e = 0; e < nn; ++e) mc[e] -= h[e];
  memcpy(xc, tauc, sizeof(cplx) * nn);
  orc_tripmt(da, xc, db, n, n, n);
  double e2 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t e = i + (size_t)j * n;
      xc[e] = xc[e] - ca * da[e] - cb * db[e];
      if (cabs(xc[e]) > e2) e2 = cabs(xc[e]);
      d1 += creal(h[e] * conj(h[e]));
      d2 += creal(mc[e] * conj(mc[e]));
    }
  err[0] = e1;
  err[1] = e2;
  err[2] = sqrt(d1 / d2);
  memcpy(taua, xa, sizeof(cplx) * nn);
  memcpy(taub, xb, sizeof(cplx) * nn);
  orc_tripmt(tauc, taua, tauc, n, n, n);
  orc_tripmt(tauc, taub, tauc, n, n, n);
  for (size_t e = 0; e < nn; ++e) {
    taua[e] += tauc[e];
    taub[e] += tauc[e];
  }
  free(da);
  free(db);
  free(xa);
  free(xb);
  free(xc);
  free(h);
  return st;
}

/* all sites: arrays of nsite n x n blocks; returns the largest error3 of the
 * updated sites (the cpacor 'error') in *error */
int orc_cpacor(int nsite, int n, const double *conc, cplx *mc, const cplx *ma, const cplx *mb,
               const cplx *tauc, cplx *taua, cplx *taub, double *error) {
  const size_t nn = (size_t)n * n;
  int st = 0;
  *error = 0.0;
  for (int li = 0; li < nsite; ++li) {
    const size_t o = (size_t)li * nn;
    const int cpalay = 1.0 - conc[li] > 1.0e-6;
    const int ica = compmat(ma + o, mc + o, n, 1.0e-10), icb = compmat(mb + o, mc + o, n, 1.0e-10);
    if (!cpalay || ica + icb >= 1) {
      memcpy(taua + o, tauc + o, sizeof(cplx) * nn);
      memcpy(taub + o, tauc + o, sizeof(cplx) * nn);
      continue;
    }
    double err[3];
    st |= orc_cpasol(conc[li], n, mc + o, ma + o, mb + o, tauc + o, taua + o, taub + o, err);
    if (err[2] >= *error) *error = err[2];
  }
  return st;
}
