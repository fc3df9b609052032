// Shared definitions of libddx: status handling, device buffers, the handle.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ddx.h"

namespace ddx {

constexpr int kMaxLmax = 16;       // template instantiations of the sweeps
constexpr int kMaxDegree = 64;     // neighbours per sphere (64-bit point masks)
constexpr int kDiisMax = 25;       // ndiis of the reference (ddcosmo.f90:87)
constexpr int kMaxIter = 200;      // n_iter of cosmo (cosmo.f90:81)
constexpr int kSweepWorkersPerCU = 8;  // one-wave sweep workers per compute unit (at least)
constexpr int kUnitsPerWorker = 12;    // units of 64 triples per sweep worker (default)
// one-wave sweep workers per workgroup: one, since a workgroup keeps its
// slot on the compute unit until its slowest wave ends and the workers'
// unit ranges end at different times (four to a workgroup: 6% slower)
constexpr int kSweepWaves = 1;

void set_error(const char *fmt, ...);

#define DDX_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::ddx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,            \
                       hipGetErrorString(e_));                               \
      return DDX_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#define DDX_TRY(call)                                                        \
  do {                                                                       \
    int s_ = (call);                                                         \
    if (s_ != DDX_OK) return s_;                                             \
  } while (0)

// Device buffer that only grows; contents are not preserved across growth.
template <typename T>
struct DevArray {
  T *p = nullptr;
  size_t cap = 0;
  DevArray() = default;
  DevArray(const DevArray &) = delete;
  DevArray &operator=(const DevArray &) = delete;
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  int reserve(size_t n) {
    if (n <= cap && p) return DDX_OK;
    release();
    size_t bytes = (n ? n : 1) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
      return DDX_ERR_ALLOC;
    }
    cap = n;
    return DDX_OK;
  }
  operator T *() const { return p; }
};

// Device-side Jacobi/DIIS bookkeeping, one per instance of a solve in
// flight.  A batch of several instances has one more state after theirs,
// the summary: done once every instance is done (the gate of the sweeps and
// of the host loop), it the largest iteration count, ok the conjunction.
struct DiisState {
  int nmat;                 // reference's nmat (1-based count, jacobi_diis.f90:89)
  int it;                   // iterations done
  int done;                 // converged or failed: later kernels return at once
  int ok;                   // converged
  int singular;             // DIIS resets from a singular B
  int pending_norm;         // the last k_extrap left increment-norm partials
  int slot[kDiisMax];       // logical history index -> physical slot
  int ncoef;                // coefficients valid in coef[] (0: keep Jacobi step)
  int remaining;            // batch summary only: instances not yet done
  double coef[kDiisMax];    // extrapolation weights, by logical index
  double b[(kDiisMax + 1) * (kDiisMax + 1)];  // bordered Gram matrix, rows of 25, 0-based
  double norm;              // increment norm of the last iteration
  double tol;               // convergence threshold of this instance
  double hist[kMaxIter + 1];
};

// The network of the learned guess (k_guess.hip): in -> 128 -> 128 -> out,
// widths padded to multiples of 16.  FP32 master weights with Adam moments,
// BF16 copies for the products.
constexpr int kGuessHidden = 128;
constexpr int kGuessGeo = 16;      // neighbourhood features per sphere
struct GuessNet {
  int in = 0, inp = 0, out = 0, outp = 0;
  DevArray<float> w[3], b[3], mw[3], vw[3], mb[3], vb[3], gw[3], gb[3];
  DevArray<unsigned short> wb[3];
  int t = 0;                 // Adam steps taken
  bool trained = false;
};
// Workspace of the network for n rows (spheres)
struct GuessWork {
  DevArray<unsigned short> F, H1, H2, dYb, dZb;
  DevArray<float> Z1, Z2, Y, T, dY, dH, dZ;
  DevArray<double> scale, part;
  DevArray<float> ksplit;   // split-K partials of the weight gradients
};

// Far-field Coulomb sums (k_fmm.hip): expansion order, far criterion
// d >= R / rho, and the systems that take them by default
constexpr int kFmmOrder = 16;        // the potential at the cavity points (enters E)
constexpr int kFmmFieldOrder = 10;   // the field of zeta at the nuclei (forces only)
constexpr double kFmmRho = 0.2;
// per-term truncation bound of the field's adaptive orders: a far pair
// (t = R / d) is expanded to the smallest order p with t^(p+1) below it (at
// most kFmmFieldOrder, 2e-8 = 0.2^11), a wave to the largest its pairs need
constexpr double kFmmFieldEps = 2e-10;
constexpr int kFmmMinSpheres = 8000;
constexpr double kFmmMinPoints = 100.0;   // exposed points per sphere
struct FmmState {
  int ngrid = 0;                        // grid the order-P tables were made for
  bool lists_ok = false;                // near lists of the current setup
  DevArray<double> Yp, facsP, cfP;      // Y (ngrid, (P+1)^2), facs and cf to order P
  DevArray<double> Yq, cfQ;             // Y and cf to the field's order
  DevArray<int> cnt, noff, nidx;        // near charges of each sphere (CSR)
  DevArray<int> toff, tidx;             // near spheres of each charge (CSR)
  DevArray<double> A, Xm, part;         // local expansions, cf-scaled multipoles, partials
  DevArray<int2> cmap;                  // per cavity point: (sphere, grid point)
  DevArray<int> perm;                   // the spheres in Morton order of their centres
};

// Event pairs around the launches of one kernel class (ddx_kernel_timer).
// A launch made for a Jacobi/DIIS solve carries the solve's serial number
// and round, so that rounds enqueued after convergence can be dropped.
struct KernelTimer {
  struct Rec {
    hipEvent_t b, e;
    int solve;                  // serial number of the solve, -1 outside solves
    int round;
  };
  bool on = false;
  std::vector<Rec> recs;        // launches since the last read
  std::vector<hipEvent_t> pool; // events free for reuse
};

// Workspace and progress of one Jacobi/DIIS solve.  Each lane owns a stream;
// the direct and adjoint solves of one energy + forces evaluation run on two
// lanes at once, so that the latency-bound DIIS steps of one overlap the
// sweeps of the other.
struct Lane {
  hipStream_t stream = nullptr;
  DevArray<double> y, xnew, xs, P, xiw, rw, yp, xd, ed, part;
  DevArray<DiisState> diis;       // ninst states (+ the summary of a batch)
  DevArray<int> misc;             // per instance: slot of the newest vector
  DevArray<int> res;              // per instance: iterations, ok (after the solve)
  std::vector<int> h_res;         // host copy of res (batches)
  DiisState *h_diis = nullptr;  // pinned host mirror of diis
  int *h_head = nullptr;        // pinned: two copies of (nmat, it, done, ok), one per group
  // the solve in flight (solve_begin .. solve_end)
  int serial = -1, star = 0, diis_max = 0, max_iter = 0, rounds = 0, enqueued = 0, group = 3;
  double tol = 0.0;
  const double *rhs = nullptr;
  double *x = nullptr;
  bool active = false;
  bool f32 = false;               // the operators of this apply run in FP32
  // mixed precision (solve_run_mixed): L x carried in FP64 in yacc, FP32
  // operator rounds on the increment inc
  bool incr = false;
  DevArray<double> yacc, inc;
};

}  // namespace ddx

struct ddx_handle {
  // parameters (ddcosmo.f90:106-115)
  int n = 0, lmax = 0, nylm = 0, ngrid = 0, iconv = 8, igrad = 0, iprint = 0;
  int prec = DDX_PREC_FP64;
  double eps = 78.39, eta = 0.1;
  size_t np = 0;            // n * ngrid points

  // host copies of tables and geometry
  std::vector<double> h_grid, h_w, h_basis, h_facl, h_facs;
  std::vector<double> h_x, h_y, h_z, h_r;

  // instances: a batch of independent systems, spheres soff[q] .. soff[q+1]-1
  // of instance q (one instance for a plain handle)
  int ninst = 1;
  std::vector<int> h_soff, h_cavoff;   // ninst + 1 sphere / cavity offsets

  // sizes produced by setup
  int ncav = 0, nnl = 0, ntri = 0, maxdeg = 0;
  bool have_setup = false, have_sigma = false, have_s = false, have_charge = false;
  bool conv_sigma = false, conv_s = false;   // the last solutions converged

  hipStream_t stream = nullptr;
  hipEvent_t ev[10] = {};
  // a second stream (default priority) for work that overlaps the first:
  // the field of zeta at the nuclei beside the force kernels; fork / join
  // events without timing
  hipStream_t stream2 = nullptr;
  hipEvent_t fev[2] = {};

  // tables
  ddx::DevArray<double4> grid;    // (x, y, z, w) per Lebedev point
  ddx::DevArray<double> Y, Yw;    // basis (nylm, ngrid) and basis * w
  ddx::DevArray<float> Y32, Yw32; // their float copies (mixed precision)
  bool have_tables32 = false;
  // zero-padded layouts for the products' coalesced operand loads: Yw with
  // lm fastest in columns of ldm = 16 * row tiles over ldnp points (whole
  // groups of 32), and Y lm-major, nylm4 rows (whole k-steps of 4) of ldnt
  // points (whole tiles of 16); double and float
  ddx::DevArray<double> YwP, YT;
  ddx::DevArray<float> YwP32, YT32;
  int ldm = 0, ldnp = 0, ldnt = 0, nylm4 = 0;
  ddx::DevArray<double> facl, cf; // (2l+1)/4pi and facs * 4pi/(2l+1)
  ddx::DevArray<double> facs;

  // geometry
  ddx::DevArray<double4> sph;     // (x, y, z, r)
  ddx::DevArray<double> charge;

  // neighbours, and the spatial hash that finds them
  ddx::DevArray<int> inl, nl, rev, prow, degcnt;
  ddx::DevArray<int> ckey, bcnt, boff, bsph, cand;
  ddx::DevArray<int> inst;        // instance of each sphere (empty: one instance)
  double rmax = 0.0;              // largest radius

  // per point
  ddx::DevArray<double> fi, ui, zi;
  ddx::DevArray<unsigned long long> tmask;
  ddx::DevArray<int> ptcnt, ptoff, cavflag, cavidx;
  ddx::DevArray<double> ccav;

  // triples grouped by the sphere that contains the point
  ddx::DevArray<int> paircnt, pairoff;
  // sweep work: units of 64 triples per container, segments per container
  ddx::DevArray<int> ucnt, uoff, segcnt, segoff;
  int nunit = 0, nworker = 0, nseg_cap = 0;
  int ncu = 0;                      // compute units of the device
  ddx::DevArray<double4> tgeo;    // (cos theta, sin theta, cos phi, sin phi)
  ddx::DevArray<double2> ttw;     // (t, W)
  ddx::DevArray<float4> tgeo32;   // the FP32 sweeps' records (geo32_of)
  ddx::DevArray<float> ttw32;     //   and W
  bool have_geo32 = false;        // the copies match tgeo / ttw
  ddx::DevArray<int> tp;          // owner point index i*ngrid + n
  ddx::DevArray<int> pttri;       // per point: its triples in neighbour order
  ddx::DevArray<int> tslot;       // per triple: its slot in pttri

  // block maps of the solver kernels: each workgroup of k_jacobi works on
  // a range of one instance's elements (jblk: instance, begin, end), each of
  // k_extrap on a range of its spheres (eblk); jboff / eboff: the blocks of
  // each instance.  Coulomb tiles: per workgroup a range of <= 256 targets
  // and of sources of one instance (ctile), and for the field at the
  // nuclei the source chunks of each target tile (ftile, fgrp: tiles of
  // each target group).
  int nbj = 0, nbe = 0, nctile = 0, nftile = 0, nfgrp = 0;
  ddx::DevArray<int4> jblk, eblk, ctile, ftile;
  ddx::DevArray<int> jboff, eboff, fgrp, soff, cavoff;
  ddx::DevArray<int4> xtile;      // tiles of a field between arbitrary point sets
  ddx::DevArray<int> xgrp;
  std::vector<int4> h_ctile, h_ftile, h_xtile;   // host copies of the tile tables
  std::vector<int> h_fgrp, h_xgrp;
  int maps_n = -1;                // sphere count the block maps were built for
  bool mute_timers = false;       // launches outside the timed classes' meaning
  // sweep workers: by default enough for about kUnitsPerWorker units each (at
  // least kSweepWorkersPerCU per CU); ddx_autotune fixes workers_per_cu
  int workers_per_cu = ddx::kSweepWorkersPerCU;
  bool workers_fixed = false;
  // groups of solver rounds captured as HIP graphs (ddx_set_graphs)
  bool use_graphs = false;
  struct RoundGraph {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec;
  };
  std::vector<RoundGraph> round_graphs;
  int ysplit = 1;                 // partial projections of L X in the solver
  int coulomb_mode = DDX_COULOMB_AUTO;   // ddx_set_coulomb
  ddx::FmmState fmm;

  // solver workspace: lane 0 runs on `stream`, lane 1 on a stream of its own
  ddx::Lane lane[2];
  ddx::DevArray<double> x, y, rhs, r, part;
  int solve_serial = 0;
  hipEvent_t gev[2] = {};           // completion of a group's state copy
  int tag_solve = -1, tag_round = 0;  // attached to timed launches

  // problem data / results
  ddx::DevArray<double> phi, psi, sigma, s, xi, fx, zeta, ef, energy, fpart;
  ddx::DevArray<double> fdd;        // owner and container sums of forces_dd
  ddx::DevArray<int> scan_tmp, misc;

  // initial guess of the solves (DDX_GUESS_*) and the learned guess's
  // networks for the direct (0) and adjoint (1) solve
  int guess_mode = 0;
  ddx::GuessNet gnet[2];
  ddx::GuessWork gwork;
  unsigned guess_seed = 1;
  ddx::DevArray<double> gprev[2];     // the learned map of the last right-hand side
  bool gprev_ok[2] = {false, false};

  std::vector<double> hist_direct, hist_adjoint, hist_ops;
  std::vector<int> iters_direct, iters_adjoint;   // per instance
  ddx_stats stats{};
  ddx::KernelTimer ktimer[DDX_K_COUNT];
};

namespace ddx {
// Start and stop events of one launch when its kernel class is being timed
// (null otherwise).  They go to hipExtLaunchKernelGGL, which records them as
// the kernel starts and ends, so the measured time is the kernel's own.
struct TimedLaunch {
  hipEvent_t b = nullptr, e = nullptr;
  TimedLaunch(ddx_handle *h, int kid, hipStream_t) {
    KernelTimer &t = h->ktimer[kid];
    if (!t.on || h->mute_timers) return;
    hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto &x : ev) {
      if (!t.pool.empty()) {
        x = t.pool.back();
        t.pool.pop_back();
      } else if (hipEventCreate(&x) != hipSuccess) {
        x = nullptr;
      }
    }
    if (!ev[0] || !ev[1]) {
      for (auto x : ev)
        if (x) t.pool.push_back(x);
      return;
    }
    t.recs.push_back({ev[0], ev[1], h->tag_solve, h->tag_round});
    b = ev[0];
    e = ev[1];
  }
};
}  // namespace ddx
