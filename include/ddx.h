/*
 * ddx — ddCOSMO on one AMD Instinct MI355X (gfx950).
 *
 * C-ABI of libddx.so: plain pointers, sizes and int status codes.  Every
 * entry point replaces one procedure of the reference ddCOSMO module
 * (filippolipparini/ddPCM) and keeps its argument meaning; the citation on
 * each line names the reference interface (file:line).
 *
 * Conventions (the reference's): FP64, column-major, harmonic index fastest
 * in (nylm, nsph) arrays with lm = l*l + l + m (0-based here), grid point
 * fastest in (ngrid, nsph) arrays, coordinates and radii in bohr.  Pointers
 * are host memory; all device state lives inside the handle.  One handle
 * belongs to one host thread at a time.  Every call returns DDX_OK (0) or a
 * non-zero status; ddx_last_error() gives the message.  Where the reference
 * prints and stops, ddx returns the status instead.
 */
#ifndef DDX_H
#define DDX_H

#ifdef __cplusplus
extern "C" {
#endif

#define DDX_OK 0
#define DDX_ERR_ARG 1        /* invalid argument (sizes, lmax, grid, name)      */
#define DDX_ERR_ALLOC 2      /* device or host allocation failed                */
#define DDX_ERR_NOCONV 3     /* Jacobi/DIIS did not converge in n_iter steps    */
#define DDX_ERR_CAPACITY 4   /* a sphere has more than 64 overlapping spheres   */
#define DDX_ERR_HIP 5        /* HIP runtime error                               */
#define DDX_ERR_STATE 6      /* call out of order (e.g. forces before solves)   */

/* precision of the solve (ddx_set_precision) */
#define DDX_PREC_FP64 0      /* FP64 throughout: the reference's arithmetic      */
#define DDX_PREC_MIXED 1     /* FP32 operator rounds, FP64-carried residual      */

typedef struct ddx_handle ddx_handle;

typedef struct ddx_stats {
  int niter_direct;         /* Jacobi/DIIS iterations of the last direct solve  */
  int niter_adjoint;        /* ... of the last adjoint solve                     */
  int ntriples;             /* (sphere, point, neighbour) triples with t < 1    */
  int nnl;                  /* neighbour pairs                                   */
  int ncav;                 /* exposed cavity points                             */
  int max_degree;           /* largest neighbour count                           */
  /* last run, HIP events; where the direct and adjoint solves run in
   * lockstep (energy + forces) ms_direct covers both and ms_adjoint is 0 */
  double ms_setup, ms_rhs, ms_direct, ms_adjoint, ms_forces;
} ddx_stats;

const char *ddx_last_error(void);

/* ddinit (ddcosmo.f90:159): tables, neighbour list, fi/ui/zi, cavity.
 * ngrid is snapped to the nearest Lebedev size (ddcosmo.f90:188-199). */
int ddx_init(ddx_handle **h, int n, const double *x, const double *y,
             const double *z, const double *rvdw, int iprint, int lmax,
             int ngrid, int iconv, int igrad, double eps, double eta);

/* memfree (ddcosmo.f90:491) */
int ddx_free(ddx_handle *h);

/* module scalars nsph, ncav, nylm, ngrid, lmax (ddcosmo.f90:106-115) */
int ddx_get_sizes(const ddx_handle *h, int *nsph, int *ncav, int *nylm,
                  int *ngrid, int *lmax);

/* module arrays (ddcosmo.f90:119-123): "ccav" (3,ncav), "csph" (3,nsph),
 * "rsph", "w", "grid" (3,ngrid), "basis" (nylm,ngrid), "facl", "facs",
 * "fi", "ui" (ngrid,nsph), "zi" (3,ngrid,nsph); and the intermediates the
 * reference prints at iprint >= 4: "rhs" (nylm,nsph) of the last direct
 * solve (cosmo.f90:131), "xi" (ngrid,nsph) of the last forces
 * (forces_dd.f90:116); and "sigma", "s" (nylm,nsph), the last direct and
 * adjoint solutions. */
int ddx_get_array(const ddx_handle *h, const char *name, double *dst);
/* inl (nsph+1, 1-based as in the reference) and nl (inl[nsph]-1 entries) */
int ddx_get_neighbours(const ddx_handle *h, int *inl, int *nl);

/* mkrhs (mkrhs.f90:1): potential of point charges at the cavity points and
 * the psi vector; phi(ncav), psi(nylm,nsph). */
int ddx_mkrhs(ddx_handle *h, const double *charge, double *phi, double *psi);

/* mkrhs (mkrhs.f90:1) with its whole argument list: the potential at the
 * ncav points ccav(3, ncav) of the n charges at (x, y, z), and psi(nylm, n)
 * (sqrt(4 pi) q_i in the l = 0 entries).  Any coordinates, not only the
 * handle's; the handle gives the stream. */
int ddx_mkrhs_at(ddx_handle *h, int n, const double *charge, const double *x,
                 const double *y, const double *z, int ncav, const double *ccav,
                 double *phi, int nylm, double *psi);

/* cosmo (cosmo.f90:1): star=0 solves L X = g (g from phi when cart=1, else
 * glm) and returns esolv; star=1 solves L* S = psi.  The guess is D^-1 rhs,
 * as in the reference, unless ddx_set_guess chose another. */
int ddx_cosmo(ddx_handle *h, int star, int cart, const double *phi,
              const double *glm, const double *psi, double *sigma,
              double *esolv);

/* forces_dd (forces_dd.f90:1): ddCOSMO part of the forces, fx(3,nsph). */
int ddx_forces_dd(ddx_handle *h, const double *phi, const double *sigma,
                  const double *s, double *fx);

/* ddmkzeta (ddcosmo.f90:1787): zeta(ncav). */
int ddx_mkzeta(ddx_handle *h, const double *s, double *zeta);

/* efld (efld.f90:1): field of nsrc point sources at ntrg targets; ef(3,ntrg). */
int ddx_efld(ddx_handle *h, int nsrc, const double *src, const double *csrc,
             int ntrg, const double *ctrg, double *ef);

/* efld without allocation (no counterpart in the reference; for callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops).  The same field as ddx_efld, with every pointer in device
 * memory, on `stream` (a hipStream_t; NULL = the null stream), and no handle.
 * It calls no hipMalloc, hipFree or allocating library routine: its partial
 * sums go to a 32 MiB scratch array of the library on each device, which
 * holds the last call's contents (every value a call reads, it first
 * writes).  Calls that overlap on one device race on that array.  The source
 * chunks are sized so that the partials always fit (DDX_ERR_CAPACITY
 * otherwise); results agree with ddx_efld to rounding. */
int ddx_efld_noalloc(int nsrc, const double *src, const double *csrc, int ntrg,
                     const double *ctrg, double *ef, void *stream);

/* operator layer (matvec.f90): lx (:94), lstarx (:161), ldm1x (:246),
 * hnorm (:281).  x, y are (nylm, nsph). */
int ddx_lx(ddx_handle *h, const double *x, double *y);
int ddx_lstarx(ddx_handle *h, const double *x, double *y);
int ddx_ldm1x(ddx_handle *h, const double *x, double *y);
int ddx_hnorm(ddx_handle *h, const double *x, double *norm);

/* jacobi_diis (jacobi_diis.f90:43-55), the reference's plugin interface:
 * jacobi_diis(n, lprint, diis_max, norm, tol, rhs, x, n_iter, ok, matvec,
 * dm1vec, u_norm).  The callbacks take the reference's Fortran argument
 * passing (everything by reference): subroutine matvec(n, x, y) and
 * dm1vec(n, x, y), real*8 function u_norm(n, x).  norm: 1 tests the rms of
 * the increment, 2 its max, 3 the rms and 10 tol for the max, 4 u_norm
 * (jacobi_diis.f90:132-156); lprint > 0 prints the reference's residual
 * lines.  x holds the guess on entry; n_iter: in = maximum iterations, out =
 * iterations done (n_iter + 1 with DDX_ERR_NOCONV, as the reference leaves
 * it); ok = converged; diis_max 0 (plain Jacobi) or 11..25.
 * With the handle's own operators (ddx_op_lx or ddx_op_lstarx, ddx_op_ldm1x,
 * norm 4 with ddx_op_hnorm) the whole solve runs on the device, in FP64.
 * Any other callbacks are called on the host with host vectors; the iterate
 * and the DIIS history stay in device memory, where the Gram rows, the
 * extrapolation and the norms are formed.  h may be NULL for those.
 * Where the reference prints and stops (norm 4 without u_norm), this
 * returns DDX_ERR_ARG. */
typedef void (*ddx_matvec_fn)(const int *n, const double *x, double *y);
typedef double (*ddx_norm_fn)(const int *n, const double *x);
int ddx_jacobi_diis_ops(ddx_handle *h, int n, int lprint, int diis_max, int norm,
                        double tol, const double *rhs, double *x, int *n_iter,
                        int *ok, ddx_matvec_fn matvec, ddx_matvec_fn dm1vec,
                        ddx_norm_fn u_norm);
/* The handle's operators lx (matvec.f90:94), lstarx (:161), ldm1x (:246)
 * and hnorm (:281) as such callbacks: they act on the handle of the
 * ddx_jacobi_diis_ops call running on the calling thread (and record an
 * error for that call outside one, or when n is not nylm * nsph). */
void ddx_op_lx(const int *n, const double *x, double *y);
void ddx_op_lstarx(const int *n, const double *x, double *y);
void ddx_op_ldm1x(const int *n, const double *x, double *y);
double ddx_op_hnorm(const int *n, const double *x);

/* The device solve of ddx_jacobi_diis_ops with the handle's operators,
 * selected by star (0: lx, 1: lstarx), dm1vec = ldm1x and norm 4 (hnorm),
 * lprint 0.  Same arguments otherwise; returns DDX_OK also when it did not
 * converge (ok = 0, n_iter = maximum + 1). */
int ddx_jacobi_diis(ddx_handle *h, int star, int diis_max, double tol,
                    const double *rhs, double *x, int *n_iter, int *ok);

/* residual history of the last solve: star 0 / 1 the last direct /
 * adjoint solve, 2 the last ddx_jacobi_diis_ops call; count = all norms,
 * of which up to cap are written */
int ddx_get_history(const ddx_handle *h, int star, int cap, double *norms,
                    int *count);

/* The reference driver's whole calling sequence for point charges
 * (main.f90:164-265): mkrhs, direct solve + energy, and with igrad the
 * adjoint solve, forces_dd, zeta and both field contractions.
 * esolv in hartree; fx(3,nsph) in hartree/bohr (may be NULL if igrad = 0). */
int ddx_energy_forces(ddx_handle *h, const double *charge, double *esolv,
                      double *fx);

/* Device-resident path (no counterpart in the reference: it splits
 * ddx_energy_forces so that a caller can keep inputs in HBM).
 * ddx_upload copies coordinates (bohr), radii and charges to the device;
 * ddx_run repeats setup (neighbours, fi/ui/zi, cavity), RHS, both solves
 * and the forces from them; ddx_result copies E and F back. */
int ddx_upload(ddx_handle *h, const double *x, const double *y,
               const double *z, const double *rvdw, const double *charge);
int ddx_run(ddx_handle *h);
int ddx_result(ddx_handle *h, double *esolv, double *fx);

/* Batches (no counterpart in the reference, which holds one system per
 * process in module state; each instance follows main.f90:164-265 as if
 * it ran alone, with its own neighbour list, potential, Jacobi/DIIS state,
 * convergence test and forces).  `count` molecules with nsph[q] spheres
 * each; coordinates, radii (and later charges) are concatenated in
 * instance order.  One handle, one pass of every kernel for all instances:
 * the sweeps, projections and Jacobi steps cover the whole batch, and each
 * instance stops updating once it has converged.
 *   ddx_batch_energy_forces  esolv[count]; fx(3, sum nsph).  If an instance
 *                            does not converge the others are still
 *                            complete and the call returns DDX_ERR_NOCONV.
 *   ddx_batch_result         after ddx_upload + ddx_run on a batch handle
 *   ddx_batch_iterations     direct / adjoint iterations per instance
 *                            (n_iter + 1 where it did not converge)
 *   ddx_batch_sizes          count, sphere and cavity offsets (count + 1)
 * The single-system calls ddx_cosmo, ddx_jacobi_diis, ddx_energy_forces and
 * ddx_result refuse batch handles (DDX_ERR_STATE); the operator calls apply
 * to all instances at once. */
int ddx_batch_init(ddx_handle **h, int count, const int *nsph, const double *x,
                   const double *y, const double *z, const double *rvdw, int lmax,
                   int ngrid, int iconv, int igrad, double eps, double eta);
int ddx_batch_energy_forces(ddx_handle *h, const double *charge, double *esolv,
                            double *fx);
int ddx_batch_result(ddx_handle *h, double *esolv, double *fx);
int ddx_batch_iterations(const ddx_handle *h, int *direct, int *adjoint);
int ddx_batch_sizes(const ddx_handle *h, int *count, int *sph_off, int *cav_off);

/* Precision of the solves (no counterpart in the reference, which is FP64
 * throughout).  DDX_PREC_MIXED runs the L X and L* S sweeps and the
 * harmonic products (projection, expansion) in FP32 with FP32 accumulation
 * inside an FP64 iteration: the operator's image L x of the iterate is
 * carried in FP64, the FP32 operator is applied to each increment x_new - x
 * and added to it, and every eighth round (the first included) applies the
 * FP64 operator to the iterate itself.  The residual g - (D + L) x, the
 * Jacobi/DIIS update, Gram row and bordered solve, the convergence test,
 * the energies and the forces stay FP64.  Results are as close to the exact
 * solution as the reference's own; iteration counts follow the FP64 path's
 * to within one or two.  ddx_lx / ddx_lstarx apply the FP32 operators while
 * it is set; ddx_jacobi_diis always runs FP64. */
int ddx_set_precision(ddx_handle *h, int prec);

/* Coulomb sums of the point-charge path (mkrhs.f90:102-114 and the two
 * field terms of main.f90:235-265, efld.f90:98-114; no counterpart in the
 * reference, which sums all pairs, as a choice).  DDX_COULOMB_DIRECT sums
 * all pairs.  DDX_COULOMB_FMM splits each sum into near pairs, summed
 * directly, and far pairs (sphere centres at least 5 radii apart) taken
 * through spherical-harmonic expansions of order 16 about the sphere
 * centres, which leave a relative error near 1e-12 per far term.
 * DDX_COULOMB_AUTO (the default) takes the expansions for single systems of
 * at least 8000 spheres with at least 100 exposed points per sphere, and
 * the direct sums otherwise.  Batch handles always sum directly
 * (DDX_ERR_STATE for DDX_COULOMB_FMM). */
#define DDX_COULOMB_AUTO 0
#define DDX_COULOMB_DIRECT 1
#define DDX_COULOMB_FMM 2
int ddx_set_coulomb(ddx_handle *h, int mode);

/* Initial guess of the solves (no counterpart in the reference beyond
 * DDX_GUESS_REFERENCE, which it always uses: cosmo.f90:136-138,174-176).
 *   DDX_GUESS_REFERENCE  D^-1 rhs
 *   DDX_GUESS_PREVIOUS   the handle's last solution of the same solve (the
 *                        previous frame of a trajectory); D^-1 rhs before
 *                        the first solve
 *   DDX_GUESS_LEARNED    D^-1 rhs + the correction of a per-sphere network
 *                        (inputs: D^-1 rhs of the sphere and 16 features of
 *                        its neighbourhood; two hidden layers of 128; BF16
 *                        weights, FP32 accumulation on MFMA), trained on the
 *                        handle's own converged solutions by
 *                        ddx_guess_train; D^-1 rhs until it is trained
 *   DDX_GUESS_LEARNED_PREVIOUS  the previous solution plus the change of the
 *                        network's map between the previous and the current
 *                        right-hand side (the previous solution alone until
 *                        the network is trained)
 * The Jacobi/DIIS convergence test is unchanged: the guess changes the
 * iteration count, not the tolerance.
 * ddx_guess_train: `steps` Adam steps (learning rate lr) of the direct and,
 * with forces, the adjoint network towards the last solve's solutions;
 * loss[2] = their mean squared errors (in units of the rms of D^-1 rhs).
 * ddx_guess_reset: fresh networks from `seed`. */
#define DDX_GUESS_REFERENCE 0
#define DDX_GUESS_PREVIOUS 1
#define DDX_GUESS_LEARNED 2
#define DDX_GUESS_LEARNED_PREVIOUS 3
int ddx_set_guess(ddx_handle *h, int mode);
int ddx_guess_train(ddx_handle *h, int steps, double lr, double *loss);
int ddx_guess_reset(ddx_handle *h, unsigned seed);

/* One forward and backward pass of the guess network with weights given in
 * host memory (no counterpart in the reference; for the tests of the
 * network's kernels against a CPU restatement).  F (n x in) BF16 bits; T (n x
 * out); wts = {W1 (in x 128), b1, W2 (128 x 128), b2, W3 (128 x out), b3}
 * FP32 row-major; returns Y (n x out), grads in the layouts of wts, loss. */
int ddx_mlp_grad(int n, int in, int out, const unsigned short *F, const float *T,
                 const float *const *wts, float *Y, float *const *grads, double *loss);

/* The BF16 product kernel of the learned guess on its own (no counterpart
 * in the reference; for its tests): C (M x N, FP32, row-major) = A (M x K)
 * B (K x N), A and B row-major BF16 given as raw 16-bit patterns, FP32
 * accumulation. */
int ddx_gemm_bf16(int M, int N, int K, const unsigned short *a, const unsigned short *b,
                  float *c);
int ddx_get_stats(const ddx_handle *h, ddx_stats *st);

/* Autotuner (no counterpart in the reference): times the whole pipeline of
 * the uploaded system (ddx_upload first) for 4, 8, 16, 32 and 64 sweep
 * workers per compute unit and then for 1, 2 and 4 point splits of the L X projection,
 * keeps the fastest of each on the handle, and leaves the results of a run
 * with them.  The worker count changes how the L* S sums are split into
 * partial sums, so results may differ from the untuned handle's in the last
 * bits. */
int ddx_autotune(ddx_handle *h, int *workers_per_cu, int *proj_split);

/* Graph replay of the solver (no counterpart in the reference): with it on,
 * after the first group of iterations of a solve every further group (two
 * rounds of sweep, Jacobi, DIIS and extrapolation kernels) is one launch of
 * a HIP graph captured on first use, which saves the host-side launch cost
 * of small systems.  Results are identical to the eager launches.  Kernel
 * timers, while on, force eager launches. */
int ddx_set_graphs(ddx_handle *h, int enable);

/* Kernel timing (instrumentation; no counterpart in the reference).  While
 * enabled for a kernel class, every launch of that class is bracketed by HIP
 * events on the handle's stream; ddx_kernel_timer_read returns the number of
 * launches and their summed duration since the last read and resets them. */
#define DDX_K_SWEEP_L 0      /* L X sweep over the triples               */
#define DDX_K_SWEEP_LSTAR 1  /* L* S sweep over the triples              */
#define DDX_K_PROJECT 2      /* MFMA projection onto harmonics           */
#define DDX_K_EXPAND 3       /* MFMA expansion on the grid (xi)          */
#define DDX_K_JACOBI 4       /* Jacobi update + DIIS Gram row            */
#define DDX_K_DIIS 5         /* DIIS bordered solve                      */
#define DDX_K_EXTRAP 6       /* DIIS extrapolation + increment norm      */
#define DDX_K_FIELD 7        /* all-pairs field                          */
#define DDX_K_POTENTIAL 8    /* all-pairs potential                      */
#define DDX_K_FORCES 9       /* per-triple forces terms                  */
#define DDX_K_SWEEP_PAIR 10  /* L X and L* S in one sweep (paired solves) */
#define DDX_K_COUNT 11
int ddx_kernel_timer(ddx_handle *h, int kernel, int enable);
int ddx_kernel_timer_read(ddx_handle *h, int kernel, int *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif
