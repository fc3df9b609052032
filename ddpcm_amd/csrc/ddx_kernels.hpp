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
// y(lm, i) = alpha sum_n Yw(lm, n) P(n, i), as ysplit partial sums over
// disjoint point sets at y + j nylm nsph (j < ysplit)
int project(ddx_handle *h, hipStream_t s, double alpha, const double *P, double *y,
            const DiisState *gate, int ysplit);
// partial projections of L X inside the solver: up to kProjSplit, as many
// as keep about 1024 workgroups busy (h->ysplit, set with the block maps)
constexpr int kProjSplit = 4;
// xi(n, i) = [w_n] sum_lm Y(lm, n) S(lm, i)
int expand(ddx_handle *h, hipStream_t s, const double *S, double *xi, bool wscale,
           const DiisState *gate);
// the same two products in FP32 (v_mfma_f32_16x16x4f32, FP32 tables Yw32 /
// Y32) for the mixed precision: P and xi in float, S and y in double
int project32(ddx_handle *h, hipStream_t s, double alpha, const float *P, double *y,
              const DiisState *gate, int ysplit);
int expand32(ddx_handle *h, hipStream_t s, const double *S, float *xi, bool wscale,
             const DiisState *gate);
// the FP32 projection of L X summing the sweep's point terms rw itself
// (ptoff ranges, the gather's order) where the system is large enough for
// it; *done says whether it ran (else gather P and project32)
int project32_gathered(ddx_handle *h, hipStream_t s, double alpha, const float *rw, double *y,
                       const DiisState *gate, int ysplit, bool *done);
// float copies of the harmonic tables (made once per handle)
int make_tables32(ddx_handle *h);
// the zero-padded layouts of Y and Yw (handle set-up, after Y and Yw)
int make_ptables(ddx_handle *h);

// k_sweep.hip — off-diagonal operators (matvec.f90:94-235), on the scratch
// and stream of a lane.  L X reads X' = cf X from ln.xs: apply_lx makes it
// unless xs_ready says that ln.xs already holds cf x.
int apply_lx(ddx_handle *h, Lane &ln, const double *x, double *y, const DiisState *gate,
             bool xs_ready, int ysplit);
// (want_y = false leaves L* S as per-segment partial sums in ln.yp)
int apply_lstarx(ddx_handle *h, Lane &ln, const double *x, double *y, const DiisState *gate,
                 bool want_y);
// yd = L xd (ld.xs holds cf xd) and L* xa as per-segment partial sums in
// la.yp, in one pass over the triples, gated by the DIIS states of the lanes
int apply_pair(ddx_handle *h, Lane &ld, double *yd, Lane &la, const double *xa, int ysplit);
// local expansions A (n, (P+1)^2) of the far point charges q at every
// sphere (charges j with d_ij >= R_i / rho), P in {8, 12, 16}; cfP the cf
// factors facs 4pi/(2l+1) to order P
int fmm_p2l(ddx_handle *h, int P, double rho, const double *q, const double *cfP, double *A);
// xs = cf x
int cf_scale(ddx_handle *h, hipStream_t s, const double *x, double *xs);
// K(l, m), m >= 0: the scale of the FP32 sweep's basis entry (l, +-m)
double sweep_basis_scale(int l, int m);
// y = sum of the per-segment partial sums of L* S
int seg_sum(ddx_handle *h, hipStream_t s, const double *yp, double *y, const DiisState *gate);

// k_solver.hip — Jacobi/DIIS (jacobi_diis.f90:43-438), RHS, norms
// Solves on lanes: solve_begin sets one up (x holds the guess); solve_run
// iterates one solve, or a direct (star 0) and an adjoint (star 1) solve in
// lockstep on one stream, enqueueing groups of iterations and reading the
// DIIS states back between groups; solve_end reports each.
// solve_jacobi_diis runs the three for one solve.
// the state that gates a lane's kernels: the instance's own for a plain
// handle, the summary of a batch
inline const DiisState *solve_gate(const ddx_handle *h, const Lane &ln) {
  return ln.diis.p + (h->ninst > 1 ? h->ninst : 0);
}
int build_solver_maps(ddx_handle *h);
// releases the captured graphs of solver rounds (after a new setup)
void drop_round_graphs(ddx_handle *h);
int solve_begin(ddx_handle *h, Lane &ln, int star, int diis_max, double tol,
                const double *d_rhs, double *d_x, int max_iter);
int solve_run(ddx_handle *h, Lane *const *lanes, int count);
// solve_run in the handle's precision: FP64, or the mixed precision (FP32
// operator rounds on the increments, the operator image carried in FP64
// and refreshed by the FP64 operator)
int solve_lanes(ddx_handle *h, Lane *const *lanes, int count);
int solve_end(ddx_handle *h, Lane &ln, int *n_iter, int *ok, std::vector<double> *hist);
int solve_jacobi_diis(ddx_handle *h, Lane &ln, int star, int diis_max, double tol,
                      const double *d_rhs, double *d_x, int max_iter,
                      int *n_iter, int *ok, std::vector<double> *hist);
int device_ldm1x(ddx_handle *h, hipStream_t s, const double *x, double *y);
int device_hnorm(ddx_handle *h, const double *x, double *norm_host);
int build_rhs(ddx_handle *h, const double *d_phi, double *d_rhs);
// d_out[q] = scale <a, b> over the elements of instance q
int energy_dot(ddx_handle *h, const double *a, const double *b, double scale,
               double *d_out);

// k_coulomb.hip — all-pairs potential and field (mkrhs.f90, efld.f90) on
// tile tables (targets tb..te-1, sources sb..se-1 per workgroup)
void coulomb_tiles(const std::vector<int> &toff, const std::vector<int> &soff, bool split,
                   std::vector<int4> &tiles, std::vector<int> &grp);
// h->ctile (cavity points from the spheres) and h->ftile / h->fgrp (the
// nuclei from the cavity points), per instance
int build_coulomb_tiles(ddx_handle *h);
// phi; with ef != NULL also the field ef(3, .) of the same sources
int coulomb_potential(ddx_handle *h, const int4 *tiles, int ntile, const double *q,
                      const double4 *src, const double *trg3, double *phi, double *ef);
// ef(3, .) from per-chunk partials summed per target group
int coulomb_field(ddx_handle *h, const int4 *tiles, int ntile, const int *grp, int ngrp,
                  const double *q, const double *src3, const double *trg3, double *ef);
int coulomb_field_points(ddx_handle *h, int nsrc, const double *q, const double *src3, int ntrg,
                         const double *trg3, double *ef);
int make_psi(ddx_handle *h, const double *charge, double *psi);
// psi (nylm, n) of n charges: sqrt(4 pi) q_i in the l = 0 entry, zeros
int make_psi_n(ddx_handle *h, int n, int nylm, const double *charge, double *psi);
// device address of the scratch array of ddx_efld_noalloc on the current
// device (looked up once per device; out may be null)
int noalloc_setup(double **out);

// k_fmm.hip — far-field expansions of the Coulomb sums (ddx_set_coulomb)
bool use_fmm(const ddx_handle *h);
// phi at the cavity points and, with ef, the solute field there
int fmm_potential(ddx_handle *h, const double *q, double *phi, double *ef);
// ez(3, n): the field of the cavity charges zeta at the nuclei
int fmm_zeta_field(ddx_handle *h, const double *zeta, double *ez);

// k_forces.hip — forces_dd, ddmkzeta and the field contractions
// (xi_ready: h->xi already holds basis^T s)
int forces_dd(ddx_handle *h, const double *d_phi, const double *d_sigma,
              const double *d_s, double *d_fx, bool xi_ready = false);
// (xi_ready: h->xi already holds basis^T s, as forces_dd leaves it)
int make_zeta(ddx_handle *h, const double *d_s, double *d_zeta, bool xi_ready = false);
// solute_field: h->ef already holds the solute field at the cavity points
// (from coulomb_potential)
int field_forces(ddx_handle *h, const double *d_charge, const double *d_zeta,
                 double *d_fx, bool solute_field);
// its two parts: the fields (on h->stream), then fx -= the contractions
int field_forces_fields(ddx_handle *h, const double *d_charge, const double *d_zeta,
                        bool solute_field);
int field_forces_contract(ddx_handle *h, const double *d_charge, const double *d_zeta,
                          double *d_fx);
// size of h->ef that field_forces uses
size_t field_buffer_size(const ddx_handle *h);

// k_guess.hip — the learned initial guess
// (ws: room for split-K partials of products with few tiles and a long K)
int gemm_bf16(hipStream_t s, int M, int N, int K, const unsigned short *a, long ai, long ak,
              const unsigned short *b, long bk, long bj, float *c, long ldc, bool accumulate,
              float *ws = nullptr, size_t ws_floats = 0);
int convert_bf16(hipStream_t s, long n, const float *x, unsigned short *y);
int learned_guess(ddx_handle *h, int star, const double *v, double *x, bool delta);
int train_guess(ddx_handle *h, int star, const double *v, const double *x, int steps, double lr,
                double *loss);

}  // namespace ddx
