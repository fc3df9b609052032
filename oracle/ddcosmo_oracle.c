/*
 * ORACLE — test infrastructure only.
 *
 * Sequential FP64 restatement of the reference ddCOSMO path.  Every function
 * cites the reference routine it follows (file:line in filippolipparini/ddPCM).
 * The arithmetic keeps the reference's operation order so that, compiled
 * without FMA contraction (-ffp-contract=off, as the Makefile does), it agrees
 * with the compiled reference to a few ulps.
 *
 * Pinning (how each part is checked, tests/test_oracle_*.py):
 *   tables, Y_lm, dY_lm, fsw/dfsw ........ reference full-precision dump, the
 *       formaldehyde.log KATs, closed forms, finite differences, orthonormality
 *   neighbour list, fi/ui/zi, ccav ........ reference dump, brute force
 *   mkrhs / efld .......................... reference dump, long-double sums
 *   lx / lstarx ........................... reference dump of lx(v), lstarx(v);
 *       adjointness <u, L v> = <L* u, v>; dense solve of L; isolated ion
 *       (Born closed form)
 *   Jacobi/DIIS, cosmo .................... reference dump of sigma, S, E and
 *       the six shipped logs (energies, iteration counts)
 *   forces ................................ reference dump, shipped logs,
 *       sum_i F_i = 0, central differences of the energy
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ddcosmo_oracle.h"

#define IDX(l, m) ((l) * (l) + (l) + (m))

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: allocation failed\n"); abort(); }
  return p;
}

/* ---------------------------------------------------------------------- */
/* switching function chi and its derivative, ddcosmo.f90:566-607, 616-655 */
double orc_fsw(double t, double s, double eta) {
  double x = t - (s + 1.0) * eta / 2.0;
  double flow = 1.0 - eta;
  if (x >= 1.0) return 0.0;
  if (x <= flow) return 1.0;
  double a = 15.0 * eta - 12.0;
  double b = 10.0 * eta * eta - 15.0 * eta + 6.0;
  double e5 = eta * eta * eta * eta * eta;
  return ((x - 1.0) * (x - 1.0) * (1.0 - x) * (6.0 * x * x + a * x + b)) / e5;
}

double orc_dfsw(double t, double s, double eta) {
  double x = t - (s + 1.0) * eta / 2.0;
  double flow = 1.0 - eta;
  if (x >= 1.0) return 0.0;
  if (x <= flow) return 0.0;
  double e5 = eta * eta * eta * eta * eta;
  return 30.0 * (1.0 - x) * (x - 1.0) * (x - 1.0 + eta) * (x - 1.0 + eta) / e5;
}

/* ---------------------------------------------------------------------- */
/* associated Legendre recurrence, m outer / l inner, ddcosmo.f90:1056-1094 */
static void polleg(int lmax, double x, double y, double *plm) {
  double fact = 1.0, pmm = 1.0, somx2 = y;
  for (int m = 0; m <= lmax; ++m) {
    plm[IDX(m, m)] = pmm;
    if (m == lmax) return;
    double fm = (double)m;
    double pmm1 = x * (2.0 * fm + 1.0) * pmm;
    plm[IDX(m + 1, m)] = pmm1;
    double pmmo = pmm;
    for (int l = m + 2; l <= lmax; ++l) {
      double fl = (double)l;
      double pll = (x * (2.0 * fl - 1.0) * pmm1 - (fl + fm - 1.0) * pmm) / (fl - fm);
      plm[IDX(l, m)] = pll;
      pmm = pmm1;
      pmm1 = pll;
    }
    pmm = -pmmo * fact * somx2;
    fact = fact + 2.0;
  }
}

/* cos(m phi), sin(m phi) by Chebyshev recurrence, ddcosmo.f90:1103-1125 */
static void trgev(int lmax, double x, double y, double *cx, double *sx) {
  cx[0] = 1.0; sx[0] = 0.0;
  cx[1] = x;   sx[1] = y;
  for (int m = 2; m <= lmax; ++m) {
    cx[m] = 2.0 * x * cx[m - 1] - cx[m - 2];
    sx[m] = 2.0 * x * sx[m - 1] - sx[m - 2];
  }
}

/* real spherical harmonics at a unit vector, ddcosmo.f90:830-895 */
void orc_ylmbas(const orc_state *st, const double *x, double *basloc) {
  int lmax = st->lmax;
  double vplm[(32 + 1) * (32 + 1)], vcos[34], vsin[34];
  double cthe = x[2];
  double sthe = sqrt(1.0 - cthe * cthe);
  double cphi, sphi;
  if (sthe != 0.0) {
    cphi = x[0] / sthe;
    sphi = x[1] / sthe;
    trgev(lmax, cphi, sphi, vcos, vsin);
  } else {
    for (int m = 0; m <= lmax + 1; ++m) { vcos[m] = 1.0; vsin[m] = 0.0; }
  }
  polleg(lmax, cthe, sthe, vplm);
  for (int l = 0; l <= lmax; ++l) {
    int ind = IDX(l, 0);
    basloc[ind] = st->facs[ind] * vplm[ind];
    for (int m = 1; m <= l; ++m) {
      double plm = vplm[ind + m];
      basloc[ind + m] = st->facs[ind + m] * plm * vcos[m];
      basloc[ind - m] = st->facs[ind - m] * plm * vsin[m];
    }
  }
}

/* Y_lm and its tangential gradient (3, nylm), ddcosmo.f90:904-1043 */
void orc_dbasis(const orc_state *st, const double *x, double *basloc, double *dbsloc) {
  int lmax = st->lmax, nylm = st->nylm;
  double vplm[(32 + 1) * (32 + 1)], vcos[34], vsin[34];
  double cthe = x[2];
  double sthe = sqrt(1.0 - cthe * cthe);
  double cphi, sphi, et[3], ep[3];
  if (sthe != 0.0) { cphi = x[0] / sthe; sphi = x[1] / sthe; }
  else { cphi = 1.0; sphi = 0.0; }
  et[0] = cthe * cphi; et[1] = cthe * sphi; et[2] = -sthe;
  if (sthe != 0.0) { ep[0] = -sphi / sthe; ep[1] = cphi / sthe; ep[2] = 0.0; }
  else { ep[0] = 0.0; ep[1] = 1.0; ep[2] = 0.0; }
  if (sthe != 0.0) trgev(lmax, cphi, sphi, vcos, vsin);
  else for (int m = 0; m <= lmax + 1; ++m) { vcos[m] = 1.0; vsin[m] = 0.0; }
  /* integer pole factors of the reference (ddcosmo.f90:912,970-971) */
  int VC = 0;
  int VS = (int)cthe;
  polleg(lmax, cthe, sthe, vplm);
  for (int i = 0; i < nylm; ++i) basloc[i] = 0.0;
  for (int i = 0; i < 3 * nylm; ++i) dbsloc[i] = 0.0;
  for (int l = 0; l <= lmax; ++l) {
    int ind = IDX(l, 0);
    double fln = st->facs[ind];
    basloc[ind] = fln * vplm[ind];
    for (int d = 0; d < 3; ++d)
      dbsloc[3 * ind + d] = (l > 0) ? fln * vplm[ind + 1] * et[d] : 0.0;
    for (int m = 1; m <= l; ++m) {
      fln = st->facs[ind + m];
      double plm = fln * vplm[ind + m];
      double pp1 = 0.0;
      if (m < l) pp1 = -0.5 * vplm[ind + m + 1];
      double pm1 = 0.5 * ((double)(l + m) * (double)(l - m + 1) * vplm[ind + m - 1]);
      double pp = pp1 + pm1;
      basloc[ind + m] = plm * vcos[m];
      for (int d = 0; d < 3; ++d) {
        if (sthe != 0.0)
          dbsloc[3 * (ind + m) + d] = -fln * pp * vcos[m] * et[d] - (double)m * plm * vsin[m] * ep[d];
        else
          dbsloc[3 * (ind + m) + d] = -fln * pp * vcos[m] * et[d] - fln * pp * ep[d] * (double)VC;
      }
      basloc[ind - m] = plm * vsin[m];
      for (int d = 0; d < 3; ++d) {
        if (sthe != 0.0)
          dbsloc[3 * (ind - m) + d] = -fln * pp * vsin[m] * et[d] + (double)m * plm * vcos[m] * ep[d];
        else
          dbsloc[3 * (ind - m) + d] = -fln * pp * vsin[m] * et[d] - fln * pp * ep[d] * (double)VS;
      }
    }
  }
}

/* sum_l 4pi/(2l+1) t^l sum_m Y_lm sigma_lm, ddcosmo.f90:1140-1178 */
double orc_intmlp(const orc_state *st, double t, const double *sigma, const double *basloc) {
  double tt = 1.0, ss = 0.0;
  for (int l = 0; l <= st->lmax; ++l) {
    int ind = IDX(l, 0);
    double fac = tt / st->facl[ind];
    double dot = 0.0;
    for (int m = -l; m <= l; ++m) dot += basloc[ind + m] * sigma[ind + m];
    ss = ss + fac * dot;
    tt = tt * t;
  }
  return ss;
}

/* ---------------------------------------------------------------------- */
/* setup: ddinit, ddcosmo.f90:159-483 */
int orc_init(orc_state **out, int n, const double *x, const double *y,
             const double *z, const double *rvdw, int lmax, int ngrid,
             int iconv, int igrad, double eps, double eta) {
  if (n < 1 || lmax < 0 || lmax > 32) return -1;
  orc_state *st = (orc_state *)xcalloc(1, sizeof(orc_state));
  st->pi = 4.0 * atan(1.0);
  st->nsph = n;
  st->lmax = lmax;
  st->nylm = (lmax + 1) * (lmax + 1);
  st->ngrid = orc_snap_ngrid(ngrid);
  st->iconv = iconv;
  st->igrad = igrad;
  st->eps = eps;
  st->eta = eta;
  st->se = -1.0;
  int nylm = st->nylm, ng = st->ngrid;
  double pi = st->pi, sq2 = sqrt(2.0);

  /* factorials, facl, facs (ddcosmo.f90:225-243) */
  int nfact = 2 * lmax + 1 > 2 ? 2 * lmax + 1 : 2;
  double *fact = (double *)xcalloc(nfact, sizeof(double));
  fact[0] = 1.0; fact[1] = 1.0;
  for (int i = 3; i <= 2 * lmax + 1; ++i) fact[i - 1] = (double)(i - 1) * fact[i - 2];
  st->facl = (double *)xcalloc(nylm, sizeof(double));
  st->facs = (double *)xcalloc(nylm, sizeof(double));
  for (int l = 0; l <= lmax; ++l) {
    int ind = IDX(l, 0);
    double fl = (2.0 * (double)l + 1.0) / (4.0 * pi);
    double ffl = sqrt(fl);
    for (int m = -l; m <= l; ++m) st->facl[ind + m] = fl;
    st->facs[ind] = ffl;
    for (int m = 1; m <= l; ++m) {
      double fnorm = sq2 * ffl * sqrt(fact[l - m] / fact[l + m]);
      if (m % 2 == 1) fnorm = -fnorm;
      st->facs[ind + m] = fnorm;
      st->facs[ind - m] = fnorm;
    }
  }
  free(fact);

  st->csph = (double *)xcalloc(3 * (size_t)n, sizeof(double));
  st->rsph = (double *)xcalloc(n, sizeof(double));
  for (int i = 0; i < n; ++i) {
    st->csph[3 * i + 0] = x[i];
    st->csph[3 * i + 1] = y[i];
    st->csph[3 * i + 2] = z[i];
    st->rsph[i] = rvdw[i];
  }

  st->w = (double *)xcalloc(ng, sizeof(double));
  st->grid = (double *)xcalloc(3 * (size_t)ng, sizeof(double));
  if (orc_lebedev(ng, st->w, st->grid) != 0) { orc_free(st); return -2; }
  st->basis = (double *)xcalloc((size_t)nylm * ng, sizeof(double));
  for (int i = 0; i < ng; ++i) orc_ylmbas(st, st->grid + 3 * i, st->basis + (size_t)nylm * i);

  /* O(N^2) CSR neighbour list, increasing j (ddcosmo.f90:308-324) */
  st->inl = (int *)xcalloc(n + 1, sizeof(int));
  int cap = 16 * n + 16, cnt = 0;
  st->nl = (int *)xcalloc(cap, sizeof(int));
  for (int i = 0; i < n; ++i) {
    st->inl[i] = cnt;
    for (int j = 0; j < n; ++j) {
      if (i == j) continue;
      double dx = st->csph[3 * i] - st->csph[3 * j];
      double dy = st->csph[3 * i + 1] - st->csph[3 * j + 1];
      double dz = st->csph[3 * i + 2] - st->csph[3 * j + 2];
      double d2 = dx * dx + dy * dy + dz * dz;
      double r2 = (st->rsph[i] + st->rsph[j]) * (st->rsph[i] + st->rsph[j]);
      if (d2 <= r2) {
        if (cnt == cap) {
          cap *= 2;
          st->nl = (int *)realloc(st->nl, cap * sizeof(int));
          if (!st->nl) abort();
        }
        st->nl[cnt++] = j;
      }
    }
  }
  st->inl[n] = cnt;

  /* fi, ui, zi (ddcosmo.f90:359-411) */
  st->fi = (double *)xcalloc((size_t)ng * n, sizeof(double));
  st->ui = (double *)xcalloc((size_t)ng * n, sizeof(double));
  if (igrad) st->zi = (double *)xcalloc(3 * (size_t)ng * n, sizeof(double));
  double se = st->se;
  for (int i = 0; i < n; ++i) {
    for (int ig = 0; ig < ng; ++ig) {
      size_t p = (size_t)ng * i + ig;
      for (int ij = st->inl[i]; ij < st->inl[i + 1]; ++ij) {
        int j = st->nl[ij];
        double v[3];
        for (int d = 0; d < 3; ++d)
          v[d] = st->csph[3 * i + d] + st->rsph[i] * st->grid[3 * ig + d] - st->csph[3 * j + d];
        double vv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double t = vv / st->rsph[j];
        double xt = orc_fsw(t, se, eta);
        double swthr = 1.0 + (se + 1.0) * eta / 2.0;
        if (igrad && (t < swthr && t > swthr - eta)) {
          double fac = orc_dfsw(t, se, eta) / st->rsph[j];
          for (int d = 0; d < 3; ++d) st->zi[3 * p + d] += fac * v[d] / vv;
        }
        st->fi[p] += xt;
      }
      if (st->fi[p] <= 1.0) st->ui[p] = 1.0 - st->fi[p];
    }
  }

  /* exposed cavity points, sphere-major (ddcosmo.f90:422-468) */
  int ncav = 0;
  for (size_t p = 0; p < (size_t)ng * n; ++p)
    if (st->ui[p] > 0.0) ++ncav;
  st->ncav = ncav;
  st->ccav = (double *)xcalloc(3 * (size_t)ncav, sizeof(double));
  int ii = 0;
  for (int i = 0; i < n; ++i)
    for (int ig = 0; ig < ng; ++ig)
      if (st->ui[(size_t)ng * i + ig] > 0.0) {
        for (int d = 0; d < 3; ++d)
          st->ccav[3 * ii + d] = st->csph[3 * i + d] + st->rsph[i] * st->grid[3 * ig + d];
        ++ii;
      }
  *out = st;
  return 0;
}

void orc_free(orc_state *st) {
  if (!st) return;
  free(st->csph); free(st->rsph); free(st->w); free(st->grid); free(st->basis);
  free(st->facl); free(st->facs); free(st->inl); free(st->nl); free(st->fi);
  free(st->ui); free(st->zi); free(st->ccav);
  free(st);
}

/* ---------------------------------------------------------------------- */
/* solute potential and psi for point charges, mkrhs.f90:86-123 */
void orc_mkrhs(int n, const double *charge, const double *x, const double *y,
               const double *z, int ncav, const double *ccav, double *phi,
               int nylm, double *psi) {
  double pi = 4.0 * atan(1.0);
  double fac = sqrt(4.0 * pi);
  for (int ic = 0; ic < ncav; ++ic) {
    double v = 0.0;
    for (int j = 0; j < n; ++j) {
      double dx = ccav[3 * ic] - x[j];
      double dy = ccav[3 * ic + 1] - y[j];
      double dz = ccav[3 * ic + 2] - z[j];
      double d2 = dx * dx + dy * dy + dz * dz;
      double d = sqrt(d2);
      v = v + charge[j] / d;
    }
    phi[ic] = v;
  }
  memset(psi, 0, sizeof(double) * (size_t)nylm * n);
  for (int i = 0; i < n; ++i) psi[(size_t)nylm * i] = fac * charge[i];
}

/* electric field of point sources at targets, efld.f90:87-116 */
void orc_efld(int nsrc, const double *src, const double *csrc, int ntrg,
              const double *ctrg, double *ef) {
  for (int j = 0; j < ntrg; ++j) {
    double e0 = 0.0, e1 = 0.0, e2 = 0.0;
    for (int i = 0; i < nsrc; ++i) {
      double dx = ctrg[3 * j] - csrc[3 * i];
      double dy = ctrg[3 * j + 1] - csrc[3 * i + 1];
      double dz = ctrg[3 * j + 2] - csrc[3 * i + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      double rr = sqrt(r2);
      double r3 = r2 * rr;
      double f = src[i] / r3;
      e0 = e0 + f * dx;
      e1 = e1 + f * dy;
      e2 = e2 + f * dz;
    }
    ef[3 * j] = e0; ef[3 * j + 1] = e1; ef[3 * j + 2] = e2;
  }
}

/* g(n,i) = -ui(n,i) phi(ic), ddcosmo.f90:1187-1224 */
void orc_wghpot(const orc_state *st, const double *phi, double *g) {
  size_t np = (size_t)st->ngrid * st->nsph;
  int ic = 0;
  for (size_t p = 0; p < np; ++p) {
    g[p] = 0.0;
    if (st->ui[p] != 0.0) { g[p] = -st->ui[p] * phi[ic]; ++ic; }
  }
}

/* x_lm = sum_n Y_lm(s_n) w_n x(n), ddcosmo.f90:789-821 */
void orc_intrhs(const orc_state *st, const double *x, double *xlm) {
  int nylm = st->nylm;
  for (int k = 0; k < nylm; ++k) xlm[k] = 0.0;
  for (int ig = 0; ig < st->ngrid; ++ig) {
    const double *b = st->basis + (size_t)nylm * ig;
    for (int k = 0; k < nylm; ++k) xlm[k] = xlm[k] + b[k] * st->w[ig] * x[ig];
  }
}

/* off-diagonal action of L: calcv (ddcosmo.f90:1671-1746) + intrhs, negated
 * (matvec.f90:94-148) */
void orc_lx(const orc_state *st, const double *x, double *y) {
  int nylm = st->nylm, ng = st->ngrid;
  double *pot = (double *)xcalloc(ng, sizeof(double));
  double *basloc = (double *)xcalloc(nylm, sizeof(double));
  for (int i = 0; i < st->nsph; ++i) {
    for (int its = 0; its < ng; ++its) {
      size_t p = (size_t)ng * i + its;
      pot[its] = 0.0;
      if (!(st->ui[p] < 1.0)) continue;
      for (int ij = st->inl[i]; ij < st->inl[i + 1]; ++ij) {
        int j = st->nl[ij];
        double v[3];
        for (int d = 0; d < 3; ++d)
          v[d] = st->csph[3 * i + d] + st->rsph[i] * st->grid[3 * its + d] - st->csph[3 * j + d];
        double vv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double t = vv / st->rsph[j];
        if (t < 1.0) {
          double s[3] = {v[0] / vv, v[1] / vv, v[2] / vv};
          double xij = orc_fsw(t, st->se, st->eta);
          double oij = st->fi[p] > 1.0 ? xij / st->fi[p] : xij;
          orc_ylmbas(st, s, basloc);
          pot[its] = pot[its] + oij * orc_intmlp(st, t, x + (size_t)nylm * j, basloc);
        }
      }
    }
    double *yi = y + (size_t)nylm * i;
    orc_intrhs(st, pot, yi);
    for (int k = 0; k < nylm; ++k) yi[k] = -yi[k];
  }
  free(pot); free(basloc);
}

/* off-diagonal action of L*: xi = basis^T x, then adjrhs per sphere, negated
 * (matvec.f90:161-235, ddcosmo.f90:1301-1384) */
void orc_lstarx(const orc_state *st, const double *x, double *y) {
  int nylm = st->nylm, ng = st->ngrid, n = st->nsph;
  double *xi = (double *)xcalloc((size_t)ng * n, sizeof(double));
  double *basloc = (double *)xcalloc(nylm, sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int ig = 0; ig < ng; ++ig) {
      double s = 0.0;
      for (int k = 0; k < nylm; ++k) s += x[(size_t)nylm * i + k] * st->basis[(size_t)nylm * ig + k];
      xi[(size_t)ng * i + ig] = s;
    }
  double thr = 1.0 + (st->se + 1.0) / 2.0 * st->eta;
  for (int i = 0; i < n; ++i) {
    double *vlm = y + (size_t)nylm * i;
    for (int k = 0; k < nylm; ++k) vlm[k] = 0.0;
    for (int ij = st->inl[i]; ij < st->inl[i + 1]; ++ij) {
      int j = st->nl[ij];
      for (int ig = 0; ig < ng; ++ig) {
        double v[3];
        for (int d = 0; d < 3; ++d)
          v[d] = st->csph[3 * j + d] + st->rsph[j] * st->grid[3 * ig + d] - st->csph[3 * i + d];
        double vv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double tji = vv / st->rsph[i];
        if (tji < thr) {
          double s[3] = {v[0] / vv, v[1] / vv, v[2] / vv};
          double xji = orc_fsw(tji, st->se, st->eta);
          double fij = st->fi[(size_t)ng * j + ig];
          double oji = fij > 1.0 ? xji / fij : xji;
          orc_ylmbas(st, s, basloc);
          double t = 1.0;
          double fac = st->w[ig] * xi[(size_t)ng * j + ig] * oji;
          for (int l = 0; l <= st->lmax; ++l) {
            int ind = IDX(l, 0);
            double ffac = fac * t / st->facl[ind];
            for (int m = -l; m <= l; ++m) vlm[ind + m] = vlm[ind + m] + ffac * basloc[ind + m];
            t = t * tji;
          }
        }
      }
    }
    for (int k = 0; k < nylm; ++k) vlm[k] = -vlm[k];
  }
  free(xi); free(basloc);
}

/* inverse diagonal, matvec.f90:246-269 */
void orc_ldm1x(const orc_state *st, const double *x, double *y) {
  size_t nylm = st->nylm;
  for (int i = 0; i < st->nsph; ++i)
    for (size_t k = 0; k < nylm; ++k) y[nylm * i + k] = st->facl[k] * x[nylm * i + k];
}

/* rms over spheres of the per-sphere H-norm, matvec.f90:281-323,
 * ddcosmo.f90:1233-1271, jacobi_diis.f90:405-438 */
double orc_hnorm(const orc_state *st, const double *x) {
  double vrms = 0.0;
  for (int i = 0; i < st->nsph; ++i) {
    const double *u = x + (size_t)st->nylm * i;
    double un = 0.0;
    for (int l = 0; l <= st->lmax; ++l) {
      int ind = IDX(l, 0);
      double fac = 1.0 / (1.0 + (double)l);
      for (int m = -l; m <= l; ++m) un = un + fac * u[ind + m] * u[ind + m];
    }
    un = sqrt(un);
    vrms = vrms + un * un;
  }
  return sqrt(vrms / (double)st->nsph);
}

/* ---------------------------------------------------------------------- */
/* Gauss-Jordan with full pivoting on the bordered DIIS system, one right-hand
 * side (jacobi_diis.f90:299-400).  a is n x n column-major, 1-based macros. */
#define A_(r, c) a[((r) - 1) + (size_t)((c) - 1) * n]
static int gjinv(int n, double *a, double *b) {
  int *indxc = (int *)xcalloc(n + 1, sizeof(int));
  int *indxr = (int *)xcalloc(n + 1, sizeof(int));
  int *piv = (int *)xcalloc(n + 1, sizeof(int));
  int irow = 0, icol = 0, ok = 0;
  for (int i = 1; i <= n; ++i) {
    double big = 0.0;
    for (int j = 1; j <= n; ++j) {
      if (piv[j] != 1) {
        for (int k = 1; k <= n; ++k) {
          if (piv[k] == 0 && fabs(A_(j, k)) > big) {
            big = fabs(A_(j, k));
            irow = j;
            icol = k;
          }
        }
      }
    }
    if (icol < 1) goto done;
    piv[icol] += 1;
    if (piv[icol] > 1) goto done;
    if (irow != icol) {
      for (int c = 1; c <= n; ++c) { double s = A_(irow, c); A_(irow, c) = A_(icol, c); A_(icol, c) = s; }
      double s = b[irow - 1]; b[irow - 1] = b[icol - 1]; b[icol - 1] = s;
    }
    indxr[i] = irow;
    indxc[i] = icol;
    if (A_(icol, icol) == 0.0) goto done;
    double pinv = 1.0 / A_(icol, icol);
    A_(icol, icol) = 1.0;
    for (int c = 1; c <= n; ++c) A_(icol, c) = A_(icol, c) * pinv;
    b[icol - 1] = b[icol - 1] * pinv;
    for (int j = 1; j <= n; ++j) {
      if (j == icol) continue;
      double dum = A_(j, icol);
      A_(j, icol) = 0.0;
      for (int c = 1; c <= n; ++c) A_(j, c) = A_(j, c) - A_(icol, c) * dum;
      b[j - 1] = b[j - 1] - b[icol - 1] * dum;
    }
  }
  /* the column unscrambling of the reference only permutes the inverse,
   * which the DIIS step does not use */
  ok = 1;
done:
  free(indxc); free(indxr); free(piv);
  return ok;
}
#undef A_

/* Gram matrix update of the DIIS history, jacobi_diis.f90:258-297 (1-based b) */
#define B_(r, c) b[((r) - 1) + (size_t)((c) - 1) * lenb]
static double dot(size_t n, const double *u, const double *v) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += u[i] * v[i];
  return s;
}

static void makeb(size_t n, int nmat, int lenb, double *const *e, double *b) {
  if (nmat == 1) {
    B_(1, 1) = 0.0;
    B_(1, 2) = 1.0;
    B_(2, 1) = 1.0;
    B_(2, 2) = dot(n, e[0], e[0]);
  } else {
    B_(nmat + 1, 1) = 1.0;
    B_(1, nmat + 1) = 1.0;
    for (int i = 1; i <= nmat - 1; ++i) {
      double bij = dot(n, e[i - 1], e[nmat - 1]);
      B_(nmat + 1, i + 1) = bij;
      B_(i + 1, nmat + 1) = bij;
    }
    B_(nmat + 1, nmat + 1) = dot(n, e[nmat - 1], e[nmat - 1]);
  }
}

/* Pulay extrapolation with the drop-10 rule, jacobi_diis.f90:198-256.
 * x[], e[] are the history slots (pointer arrays so that the shift moves
 * pointers, which is the same as moving the vectors). */
static int diis(size_t n, int *nmat, int ndiis, double **x, double **e,
                double *b, double *xnew) {
  int lenb = ndiis + 1;
  if (*nmat >= ndiis) {
    for (int j = 2; j <= *nmat - 10; ++j)
      for (int k = 2; k <= *nmat - 10; ++k) B_(j, k) = B_(j + 10, k + 10);
    double *tx[64], *te[64];
    for (int j = 0; j < ndiis; ++j) { tx[j] = x[j]; te[j] = e[j]; }
    for (int j = 1; j <= *nmat - 10; ++j) { x[j - 1] = tx[j + 9]; e[j - 1] = te[j + 9]; }
    /* recycle the freed slots */
    for (int j = *nmat - 10 + 1; j <= ndiis; ++j) {
      x[j - 1] = tx[j - (*nmat - 10) - 1];
      e[j - 1] = te[j - (*nmat - 10) - 1];
    }
    *nmat -= 10;
  }
  int nmat1 = *nmat + 1;
  makeb(n, *nmat, lenb, e, b);
  double *bloc = (double *)xcalloc((size_t)nmat1 * nmat1, sizeof(double));
  double *cex = (double *)xcalloc(nmat1, sizeof(double));
  for (int c = 1; c <= nmat1; ++c)
    for (int r = 1; r <= nmat1; ++r) bloc[(r - 1) + (size_t)(c - 1) * nmat1] = B_(r, c);
  cex[0] = 1.0;
  int ok = gjinv(nmat1, bloc, cex);
  if (!ok) {
    *nmat = 1;
    free(bloc); free(cex);
    return 0;
  }
  for (size_t k = 0; k < n; ++k) xnew[k] = 0.0;
  for (int i = 1; i <= *nmat; ++i)
    for (size_t k = 0; k < n; ++k) xnew[k] = xnew[k] + cex[i] * x[i - 1][k];
  *nmat += 1;
  free(bloc); free(cex);
  return 1;
}
#undef B_

/* Jacobi iteration with DIIS, norm = 4 (user norm hnorm), jacobi_diis.f90:43-191.
 * Returns 1 when converged.  x holds the guess on entry. */
/* jacobi_diis with the reference's full plugin interface
 * (jacobi_diis.f90:43-191): callbacks with Fortran argument passing
 * (subroutine matvec(n,x,y), dm1vec(n,x,y), real*8 function u_norm(n,x)),
 * norm 1..4 with the reference's tests (:132-156: norm 1 tests the rms,
 * norm 2 the max, norm 3 both with 10 tol for the max, norm 4 u_norm), and
 * rmsvec (:405-438).  norms[2 (it-1)] = (rms or u_norm, max or -1) per
 * iteration, up to 256 iterations.  Returns n_iter as the reference leaves
 * it (n_iter + 1 when it did not converge); *ok = converged. */
int orc_jacobi_diis_ops(int n, int diis_max, int norm, double tol, const double *rhs,
                        double *x, int n_iter, orc_fmatvec_fn matvec,
                        orc_fmatvec_fn dm1vec, orc_fnorm_fn u_norm, int *ok_out,
                        double *norms) {
  size_t nn = (size_t)n;
  int dodiis = diis_max != 0, nmat = 1, ok = 0, it;
  int lenb = diis_max + 1;
  double tol_max = 10.0 * tol;
  double **xd = NULL, **ed = NULL, *b = NULL;
  if (dodiis) {
    xd = (double **)xcalloc(diis_max, sizeof(double *));
    ed = (double **)xcalloc(diis_max, sizeof(double *));
    for (int k = 0; k < diis_max; ++k) {
      xd[k] = (double *)xcalloc(nn, sizeof(double));
      ed[k] = (double *)xcalloc(nn, sizeof(double));
    }
    b = (double *)xcalloc((size_t)lenb * lenb, sizeof(double));
  }
  double *x_new = (double *)xcalloc(nn, sizeof(double));
  double *y = (double *)xcalloc(nn, sizeof(double));
  for (it = 1; it <= n_iter; ++it) {
    matvec(&n, x, y);
    for (size_t k = 0; k < nn; ++k) y[k] = rhs[k] - y[k];
    dm1vec(&n, y, x_new);
    if (dodiis) {
      for (size_t k = 0; k < nn; ++k) {
        xd[nmat - 1][k] = x_new[k];
        ed[nmat - 1][k] = x_new[k] - x[k];
      }
      diis(nn, &nmat, diis_max, xd, ed, b, x_new);
    }
    for (size_t k = 0; k < nn; ++k) x[k] = x_new[k] - x[k];
    double rms_norm = 0.0, max_norm = 0.0;
    if (norm <= 3) {
      for (size_t k = 0; k < nn; ++k) {
        max_norm = fmax(max_norm, fabs(x[k]));
        rms_norm = rms_norm + x[k] * x[k];
      }
      rms_norm = sqrt(rms_norm / (double)n);
      if (norm == 1) ok = rms_norm < tol;
      else if (norm == 2) ok = max_norm < tol;
      else ok = rms_norm < tol && max_norm < tol_max;
    } else {
      max_norm = -1.0;
      rms_norm = u_norm(&n, x);
      ok = rms_norm < tol;
    }
    if (norms && it <= 256) {
      norms[2 * (it - 1)] = rms_norm;
      norms[2 * (it - 1) + 1] = max_norm;
    }
    memcpy(x, x_new, nn * sizeof(double));
    if (ok) break;
  }
  if (dodiis) {
    for (int k = 0; k < diis_max; ++k) {
      free(xd[k]);
      free(ed[k]);
    }
    free(xd);
    free(ed);
    free(b);
  }
  free(x_new);
  free(y);
  *ok_out = ok;
  return ok ? it : n_iter + 1;
}

int orc_jacobi_diis(const orc_state *st, int n, int diis_max, double tol,
                    const double *rhs, double *x, int n_iter,
                    orc_matvec_fn matvec, orc_hist *hist) {
  size_t nn = (size_t)n;
  int dodiis = diis_max != 0, nmat = 1, ok = 0, it, nsing = 0;
  int lenb = diis_max + 1;
  double **xd = NULL, **ed = NULL, *b = NULL;
  if (dodiis) {
    xd = (double **)xcalloc(diis_max, sizeof(double *));
    ed = (double **)xcalloc(diis_max, sizeof(double *));
    for (int k = 0; k < diis_max; ++k) {
      xd[k] = (double *)xcalloc(nn, sizeof(double));
      ed[k] = (double *)xcalloc(nn, sizeof(double));
    }
    b = (double *)xcalloc((size_t)lenb * lenb, sizeof(double));
  }
  double *x_new = (double *)xcalloc(nn, sizeof(double));
  double *y = (double *)xcalloc(nn, sizeof(double));
  for (it = 1; it <= n_iter; ++it) {
    matvec(st, x, y);
    for (size_t k = 0; k < nn; ++k) y[k] = rhs[k] - y[k];
    orc_ldm1x(st, y, x_new);
    if (dodiis) {
      for (size_t k = 0; k < nn; ++k) {
        xd[nmat - 1][k] = x_new[k];
        ed[nmat - 1][k] = x_new[k] - x[k];
      }
      if (!diis(nn, &nmat, diis_max, xd, ed, b, x_new)) ++nsing;
    }
    for (size_t k = 0; k < nn; ++k) x[k] = x_new[k] - x[k];
    double rms_norm = orc_hnorm(st, x);
    ok = rms_norm < tol;
    if (hist && it <= 256) hist->norms[it - 1] = rms_norm;
    memcpy(x, x_new, nn * sizeof(double));
    if (ok) break;
  }
  if (hist) {
    hist->n_iter = ok ? it : n_iter + 1;
    hist->ok = ok;
    hist->n_singular = nsing;
  }
  if (dodiis) {
    for (int k = 0; k < diis_max; ++k) { free(xd[k]); free(ed[k]); }
    free(xd); free(ed); free(b);
  }
  free(x_new); free(y);
  return ok;
}

/* ---------------------------------------------------------------------- */
/* solver wrapper, cosmo.f90:1-226: RHS assembly (cart), guess D^-1 rhs,
 * Jacobi/DIIS with ndiis = 25, n_iter = 200, tol = 10^-iconv; energy. */
int orc_cosmo(const orc_state *st, int star, int cart, const double *phi,
              const double *glm, const double *psi, double *sigma,
              double *esolv, orc_hist *hist) {
  int nylm = st->nylm, n = st->nsph, ng = st->ngrid;
  size_t nn = (size_t)nylm * n;
  double tol = 1.0;
  for (int k = 0; k < st->iconv; ++k) tol *= 10.0;
  tol = 1.0 / tol;
  int ok;
  if (!star) {
    double *rhs = (double *)xcalloc(nn, sizeof(double));
    if (cart) {
      double *g = (double *)xcalloc((size_t)ng * n, sizeof(double));
      orc_wghpot(st, phi, g);
      for (int i = 0; i < n; ++i) orc_intrhs(st, g + (size_t)ng * i, rhs + (size_t)nylm * i);
      free(g);
    } else {
      memcpy(rhs, glm, nn * sizeof(double));
    }
    orc_ldm1x(st, rhs, sigma);
    ok = orc_jacobi_diis(st, (int)nn, 25, tol, rhs, sigma, 200, orc_lx, hist);
    if (esolv) *esolv = 0.5 * ((st->eps - 1.0) / st->eps) * dot(nn, sigma, psi);
    free(rhs);
  } else {
    orc_ldm1x(st, psi, sigma);
    ok = orc_jacobi_diis(st, (int)nn, 25, tol, psi, sigma, 200, orc_lstarx, hist);
  }
  return ok ? 0 : 1;
}

/* ---------------------------------------------------------------------- */
/* forces kernels, ddcosmo.f90:1436-1504 (fdoka) */
static void fdoka(const orc_state *st, int i, const double *sigma, const double *xi,
                  double *basloc, double *dbsloc, double *fx) {
  int ng = st->ngrid, nylm = st->nylm;
  double se = st->se, eta = st->eta;
  double tlow = 1.0 - 0.5 * (1.0 - se) * eta;
  double thigh = 1.0 + 0.5 * (1.0 + se) * eta;
  for (int ig = 0; ig < ng; ++ig) {
    size_t p = (size_t)ng * i + ig;
    double va[3] = {0, 0, 0};
    for (int ij = st->inl[i]; ij < st->inl[i + 1]; ++ij) {
      int j = st->nl[ij];
      double vij[3];
      for (int d = 0; d < 3; ++d)
        vij[d] = st->csph[3 * i + d] + st->rsph[i] * st->grid[3 * ig + d] - st->csph[3 * j + d];
      double vvij = sqrt(vij[0] * vij[0] + vij[1] * vij[1] + vij[2] * vij[2]);
      double tij = vvij / st->rsph[j];
      if (tij >= thigh) continue;
      double sij[3] = {vij[0] / vvij, vij[1] / vvij, vij[2] / vvij};
      orc_dbasis(st, sij, basloc, dbsloc);
      double alp[3] = {0, 0, 0};
      double t = 1.0;
      const double *sj = sigma + (size_t)nylm * j;
      for (int l = 1; l <= st->lmax; ++l) {
        int ind = IDX(l, 0);
        double fl = (double)l;
        double fac = t / st->facl[ind];
        for (int m = -l; m <= l; ++m) {
          double f2 = fac * sj[ind + m];
          double f1 = f2 * fl * basloc[ind + m];
          for (int d = 0; d < 3; ++d) alp[d] = alp[d] + f1 * sij[d] + f2 * dbsloc[3 * (ind + m) + d];
        }
        t = t * tij;
      }
      double beta = orc_intmlp(st, tij, sj, basloc);
      double xij = orc_fsw(tij, se, eta);
      double oij, f2;
      if (st->fi[p] > 1.0) { oij = xij / st->fi[p]; f2 = -oij / st->fi[p]; }
      else { oij = xij; f2 = 0.0; }
      double f1 = oij / st->rsph[j];
      for (int d = 0; d < 3; ++d) va[d] = va[d] + f1 * alp[d] + beta * f2 * st->zi[3 * p + d];
      if (tij > tlow) {
        double f3 = beta * orc_dfsw(tij, se, eta) / st->rsph[j];
        if (st->fi[p] > 1.0) f3 = f3 / st->fi[p];
        for (int d = 0; d < 3; ++d) va[d] = va[d] + f3 * sij[d];
      }
    }
    for (int d = 0; d < 3; ++d) fx[d] = fx[d] - st->w[ig] * xi[ig] * va[d];
  }
}

/* ddcosmo.f90:1511-1607 (fdokb), including the neighbour-of-neighbour sum */
static void fdokb(const orc_state *st, int i, const double *sigma, const double *xi,
                  double *basloc, double *dbsloc, double *fx) {
  int ng = st->ngrid, nylm = st->nylm;
  double se = st->se, eta = st->eta;
  double tlow = 1.0 - 0.5 * (1.0 - se) * eta;
  double thigh = 1.0 + 0.5 * (1.0 + se) * eta;
  const double *si = sigma + (size_t)nylm * i;
  for (int ig = 0; ig < ng; ++ig) {
    double vb[3] = {0, 0, 0}, vc[3] = {0, 0, 0};
    for (int ji = st->inl[i]; ji < st->inl[i + 1]; ++ji) {
      int j = st->nl[ji];
      double vji[3];
      for (int d = 0; d < 3; ++d)
        vji[d] = st->csph[3 * j + d] + st->rsph[j] * st->grid[3 * ig + d] - st->csph[3 * i + d];
      double vvji = sqrt(vji[0] * vji[0] + vji[1] * vji[1] + vji[2] * vji[2]);
      double tji = vvji / st->rsph[i];
      if (tji > thigh) continue;
      double sji[3] = {vji[0] / vvji, vji[1] / vvji, vji[2] / vvji};
      orc_dbasis(st, sji, basloc, dbsloc);
      double alp[3] = {0, 0, 0};
      double t = 1.0;
      for (int l = 1; l <= st->lmax; ++l) {
        int ind = IDX(l, 0);
        double fl = (double)l;
        double fac = t / st->facl[ind];
        for (int m = -l; m <= l; ++m) {
          double f2 = fac * si[ind + m];
          double f1 = f2 * fl * basloc[ind + m];
          for (int d = 0; d < 3; ++d) alp[d] = alp[d] + f1 * sji[d] + f2 * dbsloc[3 * (ind + m) + d];
        }
        t = t * tji;
      }
      size_t pj = (size_t)ng * j + ig;
      double xji = orc_fsw(tji, se, eta);
      double oji = st->fi[pj] > 1.0 ? xji / st->fi[pj] : xji;
      double f1 = oji / st->rsph[i];
      for (int d = 0; d < 3; ++d) vb[d] = vb[d] + f1 * alp[d] * xi[pj];
      if (tji > tlow) {
        double beta = orc_intmlp(st, tji, si, basloc);
        double di, fac;
        if (st->fi[pj] > 1.0) {
          di = 1.0 / st->fi[pj];
          fac = di * xji;
          int proc = 0;
          double bsum = 0.0;
          for (int jk = st->inl[j]; jk < st->inl[j + 1]; ++jk) {
            int k = st->nl[jk];
            double vjk[3];
            for (int d = 0; d < 3; ++d)
              vjk[d] = st->csph[3 * j + d] + st->rsph[j] * st->grid[3 * ig + d] - st->csph[3 * k + d];
            double vvjk = sqrt(vjk[0] * vjk[0] + vjk[1] * vjk[1] + vjk[2] * vjk[2]);
            double tjk = vvjk / st->rsph[k];
            if (k != i && tjk <= thigh) {
              proc = 1;
              double sjk[3] = {vjk[0] / vvjk, vjk[1] / vvjk, vjk[2] / vvjk};
              orc_ylmbas(st, sjk, basloc);
              double g1 = orc_intmlp(st, tjk, sigma + (size_t)nylm * k, basloc);
              double xjk = orc_fsw(tjk, se, eta);
              bsum = bsum + g1 * xjk;
            }
          }
          if (proc) {
            double g1 = di * di * orc_dfsw(tji, se, eta) / st->rsph[i];
            double g2 = g1 * xi[pj] * bsum;
            for (int d = 0; d < 3; ++d) vc[d] = vc[d] + g2 * sji[d];
          }
        } else {
          di = 1.0;
          fac = 0.0;
        }
        double f2 = (1.0 - fac) * di * orc_dfsw(tji, se, eta) / st->rsph[i];
        for (int d = 0; d < 3; ++d) vb[d] = vb[d] + f2 * xi[pj] * beta * sji[d];
      }
    }
    for (int d = 0; d < 3; ++d) fx[d] = fx[d] + st->w[ig] * (vb[d] - vc[d]);
  }
}

/* U^(x) Phi term, ddcosmo.f90:1614-1650 (fdoga) */
static void fdoga(const orc_state *st, int i, const double *xi, const double *phi, double *fx) {
  int ng = st->ngrid;
  double se = st->se, eta = st->eta;
  for (int ig = 0; ig < ng; ++ig) {
    size_t p = (size_t)ng * i + ig;
    double alp[3] = {0, 0, 0};
    if (st->ui[p] > 0.0 && st->ui[p] < 1.0)
      for (int d = 0; d < 3; ++d) alp[d] = alp[d] + phi[p] * xi[p] * st->zi[3 * p + d];
    for (int ji = st->inl[i]; ji < st->inl[i + 1]; ++ji) {
      int j = st->nl[ji];
      double vji[3];
      for (int d = 0; d < 3; ++d)
        vji[d] = st->csph[3 * j + d] + st->rsph[j] * st->grid[3 * ig + d] - st->csph[3 * i + d];
      double vvji = sqrt(vji[0] * vji[0] + vji[1] * vji[1] + vji[2] * vji[2]);
      double tji = vvji / st->rsph[i];
      double swthr = 1.0 + (se + 1.0) * eta / 2.0;
      size_t pj = (size_t)ng * j + ig;
      if (tji < swthr && tji > swthr - eta && st->ui[pj] > 0.0) {
        double sji[3] = {vji[0] / vvji, vji[1] / vvji, vji[2] / vvji};
        double fac = -orc_dfsw(tji, se, eta) / st->rsph[i];
        for (int d = 0; d < 3; ++d) alp[d] = alp[d] + fac * phi[pj] * xi[pj] * sji[d];
      }
    }
    for (int d = 0; d < 3; ++d) fx[d] = fx[d] - st->w[ig] * alp[d];
  }
}

/* ddCOSMO part of the forces, forces_dd.f90:85-164 */
void orc_forces_dd(const orc_state *st, const double *phi, const double *sigma,
                   const double *s, double *fx) {
  int ng = st->ngrid, n = st->nsph, nylm = st->nylm;
  double *xi = (double *)xcalloc((size_t)ng * n, sizeof(double));
  double *phiexp = (double *)xcalloc((size_t)ng * n, sizeof(double));
  double *basloc = (double *)xcalloc(nylm, sizeof(double));
  double *dbsloc = (double *)xcalloc(3 * (size_t)nylm, sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int ig = 0; ig < ng; ++ig)
      xi[(size_t)ng * i + ig] = dot(nylm, s + (size_t)nylm * i, st->basis + (size_t)nylm * ig);
  int ii = 0;
  for (size_t p = 0; p < (size_t)ng * n; ++p)
    if (st->ui[p] > 0.0) phiexp[p] = phi[ii++];
  for (int i = 0; i < 3 * n; ++i) fx[i] = 0.0;
  for (int i = 0; i < n; ++i) {
    fdoka(st, i, sigma, xi + (size_t)ng * i, basloc, dbsloc, fx + 3 * i);
    fdokb(st, i, sigma, xi, basloc, dbsloc, fx + 3 * i);
    fdoga(st, i, xi, phiexp, fx + 3 * i);
  }
  double fep = 0.5 * (st->eps - 1.0) / st->eps;
  for (int i = 0; i < 3 * n; ++i) fx[i] = fep * fx[i];
  free(xi); free(phiexp); free(basloc); free(dbsloc);
}

/* zeta intermediate in ncav order, ddcosmo.f90:1787-1806 */
void orc_ddmkzeta(const orc_state *st, const double *s, double *zeta) {
  int ng = st->ngrid, nylm = st->nylm, ii = 0;
  for (int i = 0; i < st->nsph; ++i)
    for (int its = 0; its < ng; ++its) {
      size_t p = (size_t)ng * i + its;
      if (st->ui[p] > 0.0) {
        zeta[ii] = st->w[its] * st->ui[p] * dot(nylm, st->basis + (size_t)nylm * its, s + (size_t)nylm * i);
        ++ii;
      }
    }
  for (int k = 0; k < ii; ++k) zeta[k] = 0.5 * ((st->eps - 1.0) / st->eps) * zeta[k];
}

/* the reference driver's calling sequence, main.f90:164-265 */
int orc_energy_forces(const orc_state *st, const double *charge, double *esolv,
                      double *fx, orc_hist *hd, orc_hist *ha) {
  int n = st->nsph, nylm = st->nylm, ncav = st->ncav, ng = st->ngrid;
  size_t nn = (size_t)nylm * n;
  double *x = (double *)xcalloc(n, sizeof(double));
  double *y = (double *)xcalloc(n, sizeof(double));
  double *z = (double *)xcalloc(n, sizeof(double));
  for (int i = 0; i < n; ++i) { x[i] = st->csph[3 * i]; y[i] = st->csph[3 * i + 1]; z[i] = st->csph[3 * i + 2]; }
  double *phi = (double *)xcalloc(ncav, sizeof(double));
  double *psi = (double *)xcalloc(nn, sizeof(double));
  double *sigma = (double *)xcalloc(nn, sizeof(double));
  orc_mkrhs(n, charge, x, y, z, ncav, st->ccav, phi, nylm, psi);
  int rc = orc_cosmo(st, 0, 1, phi, NULL, psi, sigma, esolv, hd);
  if (rc == 0 && st->igrad == 1 && fx) {
    double *s = (double *)xcalloc(nn, sizeof(double));
    rc = orc_cosmo(st, 1, 0, NULL, NULL, psi, s, NULL, ha);
    if (rc == 0) {
      orc_forces_dd(st, phi, sigma, s, fx);
      double *zeta = (double *)xcalloc(ncav, sizeof(double));
      orc_ddmkzeta(st, s, zeta);
      int ntmp = n > ncav ? n : ncav;
      double *ef = (double *)xcalloc(3 * (size_t)ntmp, sizeof(double));
      orc_efld(n, charge, st->csph, ncav, st->ccav, ef);
      int ii = 0;
      for (int i = 0; i < n; ++i)
        for (int ig = 0; ig < ng; ++ig)
          if (st->ui[(size_t)ng * i + ig] > 0.0) {
            for (int d = 0; d < 3; ++d) fx[3 * i + d] = fx[3 * i + d] - zeta[ii] * ef[3 * ii + d];
            ++ii;
          }
      orc_efld(ncav, zeta, st->ccav, n, st->csph, ef);
      for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) fx[3 * i + d] = fx[3 * i + d] - ef[3 * i + d] * charge[i];
      free(zeta); free(ef);
    }
    free(s);
  }
  free(x); free(y); free(z); free(phi); free(psi); free(sigma);
  return rc;
}

/* ---------------------------------------------------------------------- */
int orc_get_sizes(const orc_state *st, int *o) {
  o[0] = st->nsph; o[1] = st->ngrid; o[2] = st->lmax; o[3] = st->nylm;
  o[4] = st->ncav; o[5] = st->inl[st->nsph]; o[6] = st->igrad; o[7] = st->iconv;
  return 0;
}

int orc_copy_out(const orc_state *st, const char *name, double *dst) {
  size_t ng = st->ngrid, n = st->nsph, nylm = st->nylm;
  const double *src = NULL;
  size_t cnt = 0;
  if (!strcmp(name, "grid")) { src = st->grid; cnt = 3 * ng; }
  else if (!strcmp(name, "w")) { src = st->w; cnt = ng; }
  else if (!strcmp(name, "basis")) { src = st->basis; cnt = nylm * ng; }
  else if (!strcmp(name, "facs")) { src = st->facs; cnt = nylm; }
  else if (!strcmp(name, "facl")) { src = st->facl; cnt = nylm; }
  else if (!strcmp(name, "fi")) { src = st->fi; cnt = ng * n; }
  else if (!strcmp(name, "ui")) { src = st->ui; cnt = ng * n; }
  else if (!strcmp(name, "zi")) { src = st->zi; cnt = st->zi ? 3 * ng * n : 0; }
  else if (!strcmp(name, "ccav")) { src = st->ccav; cnt = 3 * (size_t)st->ncav; }
  else if (!strcmp(name, "csph")) { src = st->csph; cnt = 3 * n; }
  else if (!strcmp(name, "rsph")) { src = st->rsph; cnt = n; }
  else return -1;
  if (cnt) memcpy(dst, src, cnt * sizeof(double));
  return 0;
}

int orc_copy_out_int(const orc_state *st, const char *name, int *dst) {
  if (!strcmp(name, "inl")) memcpy(dst, st->inl, (st->nsph + 1) * sizeof(int));
  else if (!strcmp(name, "nl")) memcpy(dst, st->nl, st->inl[st->nsph] * sizeof(int));
  else return -1;
  return 0;
}
