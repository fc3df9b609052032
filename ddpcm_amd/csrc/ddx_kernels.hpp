// Host-side launchers of the libddx kernels (one translation unit each).
#pragma once

#include "ddx_common.hpp"

namespace ddx {

// k_scan.hip — exclusive prefix sums (out has n+1 entries, out[n] = total)
int scan_exclusive(const int *in, int *out, size_t n, DevArray<int> &tmp,
                   hipStream_t s);

// k_setup.hip — ddinit on the device (ddcosmo.f90:308-468)
int run_setup(ddx_handle *h);

// k_gemm.hip — FP64 MFMA products with the harmonic tables
// y(lm, i) = alpha sum_n Yw(lm, n) P(n, i)
int project(ddx_handle *h, double alpha, const double *P, double *y, const DiisState *gate);
// xi(n, i) = [w_n] sum_lm Y(lm, n) S(lm, i)
int expand(ddx_handle *h, const double *S, double *xi, bool wscale, const DiisState *gate);

// k_sweep.hip — off-diagonal operators (matvec.f90:94-235)
int apply_lx(ddx_handle *h, const double *x, double *y, const DiisState *gate);
int apply_lstarx(ddx_handle *h, const double *x, double *y, const DiisState *gate);
// per-triple W_t sum_lm R_lm(t) X_{c,lm}, stored in point-slot order
int sweep_contract(ddx_handle *h, const double *X, double *rw, const DiisState *gate);

// k_solver.hip — Jacobi/DIIS (jacobi_diis.f90:43-438), RHS, norms
int solve_jacobi_diis(ddx_handle *h, int star, int diis_max, double tol,
                      const double *d_rhs, double *d_x, int max_iter,
                      int *n_iter, int *ok, std::vector<double> *hist);
int device_ldm1x(ddx_handle *h, const double *x, double *y);
int device_hnorm(ddx_handle *h, const double *x, double *norm_host);
int build_rhs(ddx_handle *h, const double *d_phi, double *d_rhs);
int energy_dot(ddx_handle *h, const double *a, const double *b, double scale,
               double *d_out);

// k_coulomb.hip — all-pairs potential and field (mkrhs.f90, efld.f90)
int coulomb_potential(ddx_handle *h, int nsrc, const double *q, const double4 *src,
                      int ntrg, const double *trg3, double *phi);
int coulomb_field(ddx_handle *h, int nsrc, const double *q, const double *src3,
                  int ntrg, const double *trg3, double *ef);
int make_psi(ddx_handle *h, const double *charge, double *psi);

// k_forces.hip — forces_dd, ddmkzeta and the field contractions
int forces_dd(ddx_handle *h, const double *d_phi, const double *d_sigma,
              const double *d_s, double *d_fx);
int make_zeta(ddx_handle *h, const double *d_s, double *d_zeta);
int field_forces(ddx_handle *h, const double *d_charge, const double *d_zeta,
                 double *d_fx);

}  // namespace ddx
