/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, sequential FP64 restatement of the reference ddCOSMO path
 * (filippolipparini/ddPCM).  It is the checker for the HIP product: only the
 * tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it,
 * and never as the thing measured or shipped.
 *
 * Conventions follow the reference: FP64 everywhere, column-major arrays,
 * lm index fastest in (nylm, nsph) vectors with lm = l*l + l + m (0-based),
 * point index fastest in (ngrid, nsph) arrays, coordinates in bohr.
 * Neighbour lists and the cavity ordering are 0-based.
 *
 * Pinning status of each function is recorded at its definition.
 */
#ifndef DDCOSMO_ORACLE_H
#define DDCOSMO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state {
  int nsph, lmax, nylm, ngrid, ncav, igrad, iconv, iprint;
  double eps, eta, se, pi;
  double *csph;   /* (3, nsph) */
  double *rsph;   /* (nsph) */
  double *w;      /* (ngrid)  Lebedev weights x 4 pi */
  double *grid;   /* (3, ngrid) */
  double *basis;  /* (nylm, ngrid) */
  double *facl;   /* (nylm) (2l+1)/(4 pi) */
  double *facs;   /* (nylm) */
  int *inl;       /* (nsph+1) CSR offsets */
  int *nl;        /* (inl[nsph]) neighbours */
  double *fi, *ui;/* (ngrid, nsph) */
  double *zi;     /* (3, ngrid, nsph), only when igrad */
  double *ccav;   /* (3, ncav) */
} orc_state;

/* Jacobi/DIIS history written by orc_jacobi_diis when hist != NULL. */
typedef struct orc_hist {
  int n_iter;
  int ok;
  int n_singular;      /* number of DIIS resets from a singular B matrix */
  double norms[256];   /* increment norm per iteration (first 256) */
} orc_hist;

int    orc_snap_ngrid(int ngrid);
int    orc_lebedev(int ngrid, double *w, double *grid);

int    orc_init(orc_state **st, int n, const double *x, const double *y,
                const double *z, const double *rvdw, int lmax, int ngrid,
                int iconv, int igrad, double eps, double eta);
void   orc_free(orc_state *st);

double orc_fsw(double t, double s, double eta);
double orc_dfsw(double t, double s, double eta);
void   orc_ylmbas(const orc_state *st, const double *x, double *basloc);
void   orc_dbasis(const orc_state *st, const double *x, double *basloc,
                  double *dbsloc);
double orc_intmlp(const orc_state *st, double t, const double *sigma,
                  const double *basloc);

void   orc_mkrhs(int n, const double *charge, const double *x, const double *y,
                 const double *z, int ncav, const double *ccav, double *phi,
                 int nylm, double *psi);
void   orc_efld(int nsrc, const double *src, const double *csrc, int ntrg,
                const double *ctrg, double *ef);

void   orc_wghpot(const orc_state *st, const double *phi, double *g);
void   orc_intrhs(const orc_state *st, const double *x, double *xlm);
void   orc_lx(const orc_state *st, const double *x, double *y);
void   orc_lstarx(const orc_state *st, const double *x, double *y);
void   orc_ldm1x(const orc_state *st, const double *x, double *y);
double orc_hnorm(const orc_state *st, const double *x);

typedef void (*orc_matvec_fn)(const orc_state *, const double *, double *);
int    orc_jacobi_diis(const orc_state *st, int n, int diis_max, double tol,
                       const double *rhs, double *x, int n_iter,
                       orc_matvec_fn matvec, orc_hist *hist);

/* the reference's plugin interface (jacobi_diis.f90:43-55), Fortran-style
 * callbacks: arguments by reference */
typedef void (*orc_fmatvec_fn)(const int *n, const double *x, double *y);
typedef double (*orc_fnorm_fn)(const int *n, const double *x);
int    orc_jacobi_diis_ops(int n, int diis_max, int norm, double tol,
                           const double *rhs, double *x, int n_iter,
                           orc_fmatvec_fn matvec, orc_fmatvec_fn dm1vec,
                           orc_fnorm_fn u_norm, int *ok, double *norms);

int    orc_cosmo(const orc_state *st, int star, int cart, const double *phi,
                 const double *glm, const double *psi, double *sigma,
                 double *esolv, orc_hist *hist);
void   orc_forces_dd(const orc_state *st, const double *phi, const double *sigma,
                     const double *s, double *fx);
void   orc_ddmkzeta(const orc_state *st, const double *s, double *zeta);

/* The reference driver's whole calling sequence (main.f90:164-265) on
 * coordinates already in bohr.  fx may be NULL when igrad == 0. */
int    orc_energy_forces(const orc_state *st, const double *charge,
                         double *esolv, double *fx, orc_hist *hd,
                         orc_hist *ha);

/* accessors for the ctypes wrapper */
int    orc_get_sizes(const orc_state *st, int *out8);
int    orc_copy_out(const orc_state *st, const char *name, double *dst);
int    orc_copy_out_int(const orc_state *st, const char *name, int *dst);

#ifdef __cplusplus
}
#endif
#endif
