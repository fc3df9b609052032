// Jacobi iteration with Pulay DIIS, on the device (jacobi_diis.f90:43-438).
//
// One iteration is
//   y      = O x                              (matvec callback: lx or lstarx)
//   x_new  = D^-1 (rhs - y)                   (ldm1x, matvec.f90:246-269)
//   DIIS   store x_new, e = x_new - x; B row of dot products; bordered
//          system by Gauss-Jordan with full pivoting; x_new = sum c_i x_i
//          (diis/makeb/gjinv, jacobi_diis.f90:198-400, drop-10 rule included)
//   norm   hnorm(x_new - x) (matvec.f90:281-323); converged when < tol
// as four kernels: k_jacobi (update + Gram-row partial sums, fused),
// k_diis_small (one wave: B update, Gauss-Jordan, coefficients),
// k_extrap (extrapolation, increment norm partials, x <- x_new) and
// k_norm.  Their bookkeeping lives in a device DiisState; once it says
// done every later kernel returns at once, so the host can enqueue several
// iterations before it looks at the state.
//
// A handle may hold a batch of independent instances (molecules): each has
// its own DiisState and converges on its own, exactly as if solved alone.
// Every workgroup of k_jacobi and k_extrap works inside one instance (the
// block maps jblk, eblk), so that its partial sums belong to that instance;
// k_diis_small runs one workgroup per instance.  A plain handle is a batch
// of one.
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

#include <algorithm>
#include <cstring>
#include <iterator>

namespace ddx {
namespace {

constexpr int kJThreads = 256;
constexpr int kPartStride = kDiisMax + 1;   // 25 dots + self dot per block
// mixed precision: every kRefresh-th round applies the FP64 operator to the
// iterate itself (the FP64 image of the iterate restarts exactly there)
constexpr int kRefresh = 8;
// rounds per captured graph group (a multiple of kRefresh, so that every
// group of a mixed-precision solve starts at the same phase and one graph
// serves them all)
constexpr int kGraphRounds = 32;

// Pointers of one system's Jacobi/DIIS iteration; the kernels below take
// one or two systems (the paired direct and adjoint solves), selected by
// blockIdx.y (blockIdx.x for the single-workgroup kernels).
struct SysArgs {
  const double *rhs, *y;     // y: ny partial arrays of nylm nsph, summed here,
  int ny;                    // or with yp the per-segment sums of L* S:
  const double *yp;          //   y(lm, c) = sum of yp(lm, s), s in segoff[c] ..
  const int *segoff;
  double *x, *xs;            // xs = cf x, kept for the L X sweep (null for L*)
  double *xd, *ed, *xnew;    // DIIS history (ed: float in mixed precision) and the Jacobi step
  double *part, *part_e;     // Gram-row (26 rows of nbj) and increment-norm partial sums
  int *newslot;              // per instance
  DiisState *st;             // per instance, then the summary of a batch
  // mixed precision: the operator's image of the iterate carried in FP64
  double *yacc;              // L x (L* x): y adds to it (acc = 1) or replaces it
  int acc;                   //   (acc = 0, a round with the FP64 operator on x)
  double *inc;               // the last increment x_new - x (null: not kept)
  int xs_inc;                // xs = cf inc instead of cf x (the next round applies
                             //   the FP32 operator to the increment)
};
struct Systems {
  SysArgs s[2];
};

//
// a history value as double, converted where it is used: the empty asm keeps
// the compiler from moving the conversion (and the wait for its load) into
// the branch that loaded it, ahead of the element's other loads
__device__ __forceinline__ double widen_late(double v) { return v; }
__device__ __forceinline__ double widen_late(float v) {
  asm volatile("" : "+v"(v));
  return (double)v;
}

// E is the storage type of the residual history: double, or float in mixed
// precision, where the history only steers the extrapolation coefficients.
// Each residual is rounded to E once and that rounded value enters every dot
// product it is part of (its self dot included), so that B stays the Gram
// matrix of the stored vectors.
template <typename E>
__global__ __launch_bounds__(kJThreads) void k_jacobi(size_t nn, int nylm, int diis_max,
                                                      const int4 *__restrict__ jblk, int nbj,
                                                      const double *__restrict__ facl,
                                                      Systems sys) {
  const SysArgs &q0 = sys.s[blockIdx.y];
  const int4 blk = jblk[blockIdx.x];   // instance, elements blk.y .. blk.z - 1
  const DiisState *__restrict__ st = q0.st + blk.x;
  if (st->done) return;
  const double *__restrict__ rhs = q0.rhs;
  const double *__restrict__ y = q0.y;
  const int ny = q0.ny;
  const double *__restrict__ yp = q0.yp;
  const int *__restrict__ segoff = q0.segoff;
  const double *__restrict__ x = q0.x;
  double *__restrict__ xd = q0.xd;
  E *__restrict__ ed = reinterpret_cast<E *>(q0.ed);
  double *__restrict__ xnew = q0.xnew;
  double *__restrict__ part = q0.part;
  __shared__ double wpart[kJThreads / kWave][kPartStride];
  const int nmat = st->nmat;
  const int off = (diis_max > 0 && nmat >= diis_max) ? 10 : 0;
  const int ndot = diis_max > 0 ? nmat - off - 1 : 0;
  double *xo = diis_max > 0 ? xd + (size_t)st->slot[nmat - 1] * nn : xnew;
  E *eo = diis_max > 0 ? ed + (size_t)st->slot[nmat - 1] * nn : nullptr;
  const E *eh[kDiisMax];
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q)
    eh[q] = (q < ndot) ? ed + (size_t)st->slot[off + q] * nn : nullptr;
  if constexpr (sizeof(E) == 4) {
    // FP32 history: slots past ndot read the first slot's element again (a
    // cache hit, its value unused), so that every load is unconditional and
    // none waits at a branch for the one before it
    const E *fill = ndot > 0 ? eh[0] : (eo ? eo : reinterpret_cast<const E *>(x));
#pragma unroll
    for (int q = 0; q < kDiisMax; ++q)
      if (q >= ndot) eh[q] = fill;
  }
  double acc[kDiisMax];
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q) acc[q] = 0.0;
  double self = 0.0;
  for (size_t k = (size_t)blk.y + threadIdx.x; k < (size_t)blk.z; k += kJThreads) {
    // element indices fit in 32 bits (build_solver_maps checks nylm nsph)
    const unsigned c = (unsigned)k / (unsigned)nylm;
    const unsigned lm = (unsigned)k - c * (unsigned)nylm;
    // the history residuals are read first, with the element's other
    // operands, so that all of them are in flight together (kept as E until
    // the dot products, so that no conversion waits for its load early)
    E ev[kDiisMax];
#pragma unroll
    for (int q = 0; q < kDiisMax; ++q) {
      if constexpr (sizeof(E) == 4)
        ev[q] = eh[q][k];
      else
        ev[q] = (q < ndot) ? eh[q][k] : E(0);
    }
    double yk;
    if (yp) {
      yk = 0.0;
      for (int j = segoff[c]; j < segoff[c + 1]; ++j) yk += yp[(size_t)j * nylm + lm];
    } else {
      yk = y[k];
      for (int j = 1; j < ny; ++j) yk += y[(size_t)j * nn + k];
    }
    if (q0.yacc) {
      if (q0.acc) yk += q0.yacc[k];
      q0.yacc[k] = yk;
    }
    double xn = facl[lm] * (rhs[k] - yk);
    const E er = (E)(xn - x[k]);
    const double e = (double)er;
    xo[k] = xn;
    if (eo) {
      eo[k] = er;
#pragma unroll
      for (int q = 0; q < kDiisMax; ++q)
        if (q < ndot) acc[q] += widen_late(ev[q]) * e;
      self += e * e;
    }
  }
  if (diis_max == 0) return;
  // all 26 partial dots of the block: butterflies per wave, one barrier
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q) {
    if (q < ndot) {
      double v = wave_sum(acc[q]);
      if (lane_id() == 0) wpart[wave][q] = v;
    }
  }
  self = wave_sum(self);
  if (lane_id() == 0) wpart[wave][kDiisMax] = self;
  __syncthreads();
  const int q = threadIdx.x;
  if (q < ndot || q == kDiisMax)
    part[(size_t)q * nbj + blockIdx.x] =
        (wpart[0][q] + wpart[1][q]) + (wpart[2][q] + wpart[3][q]);
}

// The DIIS bookkeeping of one iteration on one wave, all in registers: lane r
// holds row r of the bordered Gram matrix B (0-based here; the reference's
// b(r+1, c+1)), n1 = nmat + 1 <= 25 rows.
//
//   drop-10 rule   when nmat reached diis_max the 10 oldest vectors go: rows
//                  and columns 11.. move to 1.. (jacobi_diis.f90:216-227)
//   makeb          new row/column of dot products (jacobi_diis.f90:258-297)
//   gjinv          Gauss-Jordan with full pivoting, right-hand side e_1
//                  (jacobi_diis.f90:299-400).  Pivot: the largest |a(j,k)|
//                  over rows j and columns k not yet pivoted (piv(j) != 1,
//                  piv(k) == 0), first in row-major order on ties, as the
//                  reference's nested loops pick it; the inverse's column
//                  unscrambling does not touch the solution and is skipped.
constexpr int kB = kDiisMax;      // register columns per row (n1 <= 25)

// value of v in lane `src`; src must be wave-uniform (it is moved to an SGPR)
__device__ __forceinline__ double lane_value(double v, int src) {
  const int s = __builtin_amdgcn_readfirstlane(src);
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), s);
  const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), s);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// one lane-pair stage of the pivot search: keep the larger value, the
// smaller row-major index on ties
template <int S>
__device__ __forceinline__ void pivot_stage(double &big, int &best) {
  const double pb = partner<S>(big);
  const int pi = partner<S>(best);
  const bool t = (pb > big) | ((pb == big) & (pi < best));
  big = t ? pb : big;
  best = t ? pi : best;
}

constexpr int kSmallThreads = 256;       // 4 workgroup partials per thread

// an instance of a batch has ended: fold it into the summary, which is done
// when the last instance is
__device__ __forceinline__ void instance_done(DiisState *sum, int it, int ok) {
  if (!sum) return;
  atomicMax(&sum->it, it);
  atomicAnd(&sum->ok, ok);
  if (atomicSub(&sum->remaining, 1) == 1) sum->done = 1;
}

// lanes per summed value in the Gram-row reduction of k_diis_small
constexpr int kTPQ = 8;

// partials r, r + kTPQ, ... of one value, added in that order; K loads per
// lane are in flight together (indices past nb re-read the last partial and
// are not added, so that no load is conditional)
template <int K>
__device__ __forceinline__ double sum_partials(const double *__restrict__ src, int nb, int r) {
  double s = 0.0;
  for (int b0 = r; b0 < nb; b0 += kTPQ * K) {
    double v[K];
#pragma unroll
    for (int i = 0; i < K; ++i) v[i] = src[min(b0 + kTPQ * i, nb - 1)];
#pragma unroll
    for (int i = 0; i < K; ++i)
      if (b0 + kTPQ * i < nb) s += v[i];
  }
  return s;
}

// Gauss-Jordan of k_diis_small: columns per lane (8 groups of lanes, n1 <= 25)
constexpr int kGJCols = 4;

// the columns of wave W's two groups, W + 8 j (lanes 0..31) and W + 4 + 8 j
// (lanes 32..63), of row lane & 31; lane L < kB holds row L of B in a
// (static register indices; the upper half takes its row from lane L - 32)
template <int W>
__device__ __forceinline__ void group_columns(const double (&a)[kB], double (&g)[kGJCols]) {
#pragma unroll
  for (int j = 0; j < kGJCols; ++j) {
    constexpr int lo0 = W, hi0 = W + 4;
    const int lo = lo0 + 8 * j, hi = hi0 + 8 * j;
    const double vl = (lo < kB) ? a[lo < kB ? lo : 0] : 0.0;
    const double vh = partner<0>((hi < kB) ? a[hi < kB ? hi : 0] : 0.0);
    g[j] = (lane_id() < 32) ? vl : vh;
  }
}

__global__ __launch_bounds__(kSmallThreads) void k_diis_small(
    int diis_max, int nbj, const int *__restrict__ jboff, const int *__restrict__ eboff,
    const int *__restrict__ soff, int ninst, int maxit, Systems sys) {
#pragma clang fp contract(off)
  const SysArgs &q0 = sys.s[blockIdx.y];
  const int inst = blockIdx.x;
  DiisState *__restrict__ st = q0.st + inst;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  // every state word and map entry of the iteration, loaded together
  const int done = st->done;
  const double tol = st->tol;
  int nmat = st->nmat;
  const int it0 = st->it;
  const bool pending = st->pending_norm != 0;
  const int jb = jboff[inst], eb = eboff[inst];
  const int nblk = jboff[inst + 1] - jb, nblk_e = eboff[inst + 1] - eb;
  const int nsph = soff[inst + 1] - soff[inst];
  if (done) return;
  DiisState *sum = ninst > 1 ? q0.st + ninst : nullptr;
  const double *__restrict__ part = q0.part + jb;
  const double *__restrict__ part_e = q0.part_e + eb;
  // lane r: row r of B and logical slot r, in flight with the partial sums
  double a[kB];
#pragma unroll
  for (int c = 0; c < kB; ++c) a[c] = (lane < kB) ? st->b[lane * kB + c] : 0.0;
  int myslot = (lane < diis_max) ? st->slot[lane] : 0;
  const int lastslot = st->slot[nmat - 1];
  __shared__ double red[kPartStride + 1];
  const int off = (nmat >= diis_max) ? 10 : 0;
  const int ndot = nmat - off - 1;
  // Gram-row dot products and the increment norm of the previous iteration
  // (value 26): eight consecutive lanes share a value, each sums every eighth
  // workgroup partial, so that all loads are in flight together, and the
  // eight partial sums meet in three lane-pair stages within the group
  {
    const int q = tid / kTPQ, r = tid % kTPQ;
    double s = 0.0;
    if (q <= kPartStride) {
      const bool need = (q < ndot || q == kDiisMax) || (q == kPartStride && pending);
      if (need) {
        const double *src = (q == kPartStride) ? part_e : part + (size_t)q * nbj;
        const int nb = (q == kPartStride) ? nblk_e : nblk;
        if (max(nblk, nblk_e) <= kTPQ * 8)
          s = sum_partials<8>(src, nb, r);
        else
          s = sum_partials<32>(src, nb, r);
      }
    }
    s += partner<3>(s);
    s += partner<4>(s);
    s += partner<5>(s);
    if (r == 0 && q <= kPartStride) red[q] = s;
  }
  __syncthreads();
  if (pending) {
    // rms of the per-sphere norms and the convergence test of the previous
    // iteration (rmsvec, jacobi_diis.f90:405-438; jacobi_diis.f90:153-156),
    // evaluated alike by every thread
    const double norm = sqrt(red[kPartStride] / (double)nsph);
    const bool stop = norm < tol || it0 + 1 >= maxit;
    if (tid == 0) {
      if (it0 <= kMaxIter) st->hist[it0] = norm;
      st->it = it0 + 1;
      st->norm = norm;
      st->ok = norm < tol;
      st->pending_norm = 0;
      if (stop) {
        st->done = 1;
        instance_done(sum, it0 + 1, norm < tol);
      }
    }
    if (stop) return;
  }
  // B update, by every wave alike (wave 0 persists it)
  {
  // lane i + 1 keeps dot i; every lane keeps the self dot
  const int qi = (lane >= 1 && lane <= kDiisMax) ? lane - 1 : 0;
  double mydot = red[qi];
  const double self = red[kDiisMax];
  if (!(lane >= 1 && lane - 1 < ndot)) mydot = 0.0;
  if (nmat >= diis_max) {
    const int m = nmat - 10;
    // rows 1..m-1 <- rows 11..m+9, then columns likewise
    double moved[kB];
#pragma unroll
    for (int c = 0; c < kB; ++c) moved[c] = __shfl(a[c], min(lane + 10, kWave - 1), kWave);
    if (lane >= 1 && lane <= m - 1) {
#pragma unroll
      for (int c = 0; c < kB; ++c) a[c] = moved[c];
#pragma unroll
      for (int c = 1; c + 10 < kB; ++c) a[c] = (c <= m - 1) ? a[c + 10] : a[c];
    }
    // slots: logical j (0-based) <- j + 10 for j < m, recycled ones after
    int src = (lane < m) ? lane + 10 : lane - m;
    myslot = __shfl(myslot, src < diis_max ? src : 0, kWave);
    nmat = m;
  }
  // makeb: row / column nmat (0-based) of the new vector
  if (nmat == 1) {
    if (lane == 0) { a[0] = 0.0; a[1] = 1.0; }
    if (lane == 1) { a[0] = 1.0; a[1] = self; }
  } else {
    double rowv[kB];   // rowv[c] = dot c-1, from the block's LDS copy
    rowv[0] = 0.0;
#pragma unroll
    for (int c = 1; c < kB; ++c) rowv[c] = red[c - 1];
#pragma unroll
    for (int c = 0; c < kB; ++c) {
      if (c == nmat) {
        if (lane == 0) a[c] = 1.0;
        else if (lane < nmat) a[c] = mydot;
        else if (lane == nmat) a[c] = self;
      }
    }
    if (lane == nmat) {
      a[0] = 1.0;
#pragma unroll
      for (int c = 1; c < kB; ++c)
        if (c < nmat) a[c] = rowv[c];
    }
  }
  }
  if (wave == 0) {
    // persist B and the slot map; k_extrap leaves this iteration's norm
    if (lane < kB) {
#pragma unroll
      for (int c = 0; c < kB; ++c) st->b[lane * kB + c] = a[c];
    }
    if (lane < diis_max) st->slot[lane] = myslot;
    if (lane == 0) {
      st->pending_norm = 1;
      q0.newslot[inst] = lastslot;
    }
  }

  // Gauss-Jordan with full pivoting on B, n1 = nmat + 1, by all four waves
  // and both halves of each: lane L works on row r = L & 31 (its current row
  // index is `pos`: the reference's row exchange becomes an exchange of two
  // pos values) and on the columns G, G + 8, G + 16, G + 24 of its group
  // G = wave + 4 (L >> 5); the right-hand side e_1 travels with wave 0.
  // Columns already pivoted hold the in-place inverse, which neither the
  // later pivots nor the solution read, so what they receive does not
  // matter; the arithmetic on every live element is the reference's.  A
  // candidate is coded row * 32 + column (n1 <= 25), which orders candidates
  // row-major as the reference's nested loops do; the pivot is the largest
  // value (an exact max), then the smallest code among its ties.
  // Per step and wave: its candidate, the lane holding that row, the signed
  // pivot value and the whole column it lies in, double-buffered so that
  // one barrier per step suffices.
  constexpr int kNW = kSmallThreads / kWave;
  constexpr int kNone = 0x7fffffff;
  struct alignas(16) Cand {
    double v, d;   // |pivot| (-1: none) and the signed pivot
    int k, l;      // candidate code and the lane holding its row
  };
  __shared__ Cand cand[2][kNW];
  __shared__ double ccol[2][kNW][32];
  const int n1 = nmat + 1;
  const int r = lane & 31, half = lane >> 5;
  double g[kGJCols];
  switch (__builtin_amdgcn_readfirstlane(wave)) {
    case 0: group_columns<0>(a, g); break;
    case 1: group_columns<1>(a, g); break;
    case 2: group_columns<2>(a, g); break;
    default: group_columns<3>(a, g); break;
  }
  const int grp = wave + 4 * half;
#pragma unroll
  for (int j = 0; j < kGJCols; ++j)
    if (!(r < n1 && grp + 8 * j < n1)) g[j] = 0.0;
  double bv = (lane == 0) ? 1.0 : 0.0;
  int pos = r;
  unsigned int live = (n1 >= 32) ? 0xffffffffu : ((1u << n1) - 1u);   // columns not pivoted
  bool ok = true;
  for (int it = 0; it < n1; ++it) {
    const int buf = it & 1;
    // pivot: largest |a(j, k)| > 0 over rows and columns not pivoted, first
    // in row-major order of the current row indices on ties
    const bool rowlive = r < n1 && ((live >> pos) & 1u);
    double big = -1.0;
    int best = kNone;
#pragma unroll
    for (int j = 0; j < kGJCols; ++j) {
      const int c = grp + 8 * j;
      const double v = (rowlive && c < n1 && ((live >> c) & 1u)) ? fabs(g[j]) : -1.0;
      const bool t = v > big;   // columns ascend: the first one wins ties
      big = t ? v : big;
      best = t ? pos * 32 + c : best;
    }
    if (!(big > 0.0)) best = kNone;
    {
      double m = big;
      m = fmax(m, partner<0>(m));
      m = fmax(m, partner<1>(m));
      m = fmax(m, partner<2>(m));
      m = fmax(m, partner<3>(m));
      m = fmax(m, partner<4>(m));
      m = fmax(m, partner<5>(m));
      int key = (big == m) ? best : kNone;
      key = min(key, partner<0>(key));
      key = min(key, partner<1>(key));
      key = min(key, partner<2>(key));
      key = min(key, partner<3>(key));
      key = min(key, partner<4>(key));
      key = min(key, partner<5>(key));
      big = m;
      best = __builtin_amdgcn_readfirstlane(key);
    }
    {
      // the column of this wave's candidate (the multipliers a(j, icol) if
      // it wins), held by the half of its group, and, from the lane holding
      // its row there, the signed pivot
      const int cc = best & 31;
      const int jj = (best == kNone) ? 0 : (cc >> 3);
      const int hc = (best == kNone) ? 0 : ((cc >> 2) & 1);
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < kGJCols; ++j) v = (j == jj) ? g[j] : v;
      if (half == hc) ccol[buf][wave][r] = v;
      int lr = __ffsll((unsigned long long)__ballot(lane < n1 && pos == (best >> 5))) - 1;
      lr = max(lr, 0);
      const double dv = lane_value(v, lr + 32 * hc);
      if (lane == 0) cand[buf][wave] = Cand{big, dv, best, lr};
    }
    __syncthreads();
    Cand cw[kNW];
#pragma unroll
    for (int w = 0; w < kNW; ++w) cw[w] = cand[buf][w];
    double m = cw[0].v;
#pragma unroll
    for (int w = 1; w < kNW; ++w) m = fmax(m, cw[w].v);
    best = kNone;
    int win = 0;
#pragma unroll
    for (int w = 0; w < kNW; ++w) {
      const bool t = (cw[w].v == m) & (cw[w].k < best);
      best = t ? cw[w].k : best;
      win = t ? w : win;
    }
    best = __builtin_amdgcn_readfirstlane(best);
    win = __builtin_amdgcn_readfirstlane(win);
    if (best == kNone) { ok = false; break; }
    double d = cw[0].d;
    int lr = cw[0].l;
#pragma unroll
    for (int w = 1; w < kNW; ++w) {
      d = (win == w) ? cw[w].d : d;
      lr = (win == w) ? cw[w].l : lr;
    }
    lr = __builtin_amdgcn_readfirstlane(lr);
    const int irow = best >> 5, icol = best & 31;
    if (!((live >> icol) & 1u) || d == 0.0) { ok = false; break; }
    // each row's multiplier a(j, icol) and the pivot row's values in this
    // lane's columns, read before the pivot row is scaled
    const double dum = ccol[buf][win][r];
    const int src = lr + 32 * half;
    double pr[kGJCols];
#pragma unroll
    for (int j = 0; j < kGJCols; ++j) pr[j] = __shfl(g[j], src, kWave);
    const double pinv = 1.0 / d;
    // the lane holding row irow becomes pivot row icol
    if (irow != icol) {
      int np = pos;
      if (r < n1 && pos == icol) np = irow;
      if (r == lr) np = icol;
      pos = np;
    }
    live &= ~(1u << icol);
    // scale the pivot row (a(icol, :) * pivinv) and eliminate column icol
    // from the other rows, in every column of the lane (no branch)
#pragma unroll
    for (int j = 0; j < kGJCols; ++j) {
      const double pc = pr[j] * pinv;
      const double el = g[j] - pc * dum;
      g[j] = (r == lr) ? pc : el;
    }
    if (wave == 0) {
      const double bp = lane_value(bv, lr) * pinv;
      bv = (r == lr) ? bp : bv - bp * dum;
    }
  }
  if (wave != 0) return;
  if (ok) {
    if (lane < n1 && pos >= 1) st->coef[pos - 1] = bv;
  }
  if (lane == 0) {
    if (!ok) {
      st->nmat = 1;
      st->ncoef = 0;
      st->singular += 1;
    } else {
      st->ncoef = nmat;
      st->nmat = nmat + 1;
    }
  }
}

// x_new = sum_i c_i x_i (or the plain Jacobi step), increment norm partials
// per sphere (hsnorm, ddcosmo.f90:1233-1271), x <- x_new, and xs = cf x_new
// where the system keeps it.  The workgroup streams its spheres' elements
// with all 256 threads (consecutive threads, consecutive elements) and
// leaves each element's norm term in LDS; then one thread per sphere adds
// its terms in lm order, as hsnorm does.  Spheres beyond the LDS stage are
// taken in chunks.
constexpr int kExtrapTerms = 2048;   // staged norm terms per workgroup (nylm <= 289)

__global__ __launch_bounds__(kJThreads) void k_extrap(int nylm, size_t nn, int diis_max,
                                                      const int4 *__restrict__ eblk,
                                                      const double *__restrict__ invl1,
                                                      const double *__restrict__ cf,
                                                      Systems sys) {
  const SysArgs &q0 = sys.s[blockIdx.y];
  const int4 blk = eblk[blockIdx.x];   // instance, spheres blk.y .. blk.z - 1
  const DiisState *__restrict__ st = q0.st + blk.x;
  if (st->done) return;
  const double *__restrict__ xd = q0.xd;
  const double *__restrict__ xnew = q0.xnew;
  const int *__restrict__ newslot = q0.newslot + blk.x;
  double *__restrict__ x = q0.x;
  double *__restrict__ xsc = q0.xs;
  double *__restrict__ incw = q0.inc;
  const bool xs_inc = q0.xs_inc != 0;
  double *__restrict__ part = q0.part_e;
  __shared__ double lds[kJThreads / kWave];
  __shared__ double terms[kExtrapTerms];
  const int ncoef = diis_max > 0 ? st->ncoef : 0;
  const double *xs[kDiisMax];
  double cs[kDiisMax];
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q) {
    xs[q] = (q < ncoef) ? xd + (size_t)st->slot[q] * nn : nullptr;
    cs[q] = (q < ncoef) ? st->coef[q] : 0.0;
  }
  const double *xj = diis_max > 0 ? xd + (size_t)(*newslot) * nn : xnew;
  const int chunk = kExtrapTerms / nylm;   // spheres per pass, at least 7
  const int step = kJThreads % nylm;       // lm advance per thread stride
  double usq = 0.0;
  for (int c0 = blk.y; c0 < blk.z; c0 += chunk) {
    const int c1 = min(blk.z, c0 + chunk);
    const unsigned ne = (unsigned)(c1 - c0) * (unsigned)nylm;
    const size_t e0 = (size_t)c0 * nylm;
    int lm = (int)(threadIdx.x % (unsigned)nylm);
    // two elements per step (j and j + 256), all their loads issued before
    // either one's stores
    for (unsigned j = threadIdx.x; j < ne; j += 2 * kJThreads) {
      const bool two = j + kJThreads < ne;
      int lm2 = lm + step;
      if (lm2 >= nylm) lm2 -= nylm;
      const size_t k = e0 + j, k2 = two ? k + kJThreads : k;
      double xn = 0.0, xn2 = 0.0;
      if (ncoef > 0) {
        double a[kDiisMax], b[kDiisMax];
#pragma unroll
        for (int q = 0; q < kDiisMax; ++q) {
          a[q] = (q < ncoef) ? xs[q][k] : 0.0;
          b[q] = (q < ncoef) ? xs[q][k2] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kDiisMax; ++q) {
          if (q < ncoef) {
            xn = xn + cs[q] * a[q];
            xn2 = xn2 + cs[q] * b[q];
          }
        }
      } else {
        xn = xj[k];
        xn2 = xj[k2];
      }
      const double x1 = x[k], x2 = x[k2];
      const double inc = xn - x1, inc2 = xn2 - x2;
      x[k] = xn;
      if (xsc) xsc[k] = cf[lm] * (xs_inc ? inc : xn);
      if (incw) incw[k] = inc;
      terms[j] = invl1[lm] * inc * inc;
      if (two) {
        x[k2] = xn2;
        if (xsc) xsc[k2] = cf[lm2] * (xs_inc ? inc2 : xn2);
        if (incw) incw[k2] = inc2;
        terms[j + kJThreads] = invl1[lm2] * inc2 * inc2;
      }
      lm = lm2 + step;
      if (lm >= nylm) lm -= nylm;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < c1 - c0; i += kJThreads) {
      const double *t = terms + (size_t)i * nylm;
      double sq = 0.0;
      for (int l = 0; l < nylm; ++l) sq += t[l];
      const double u = sqrt(sq);
      usq += u * u;
    }
    __syncthreads();
  }
  const double tot = block_sum_256(usq, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// rms of the per-sphere norms and the convergence test (rmsvec,
// jacobi_diis.f90:405-438; jacobi_diis.f90:153-156)
__global__ __launch_bounds__(kWave) void k_norm(const int *__restrict__ eboff,
                                                const int *__restrict__ soff, int ninst,
                                                int maxit, Systems sys) {
  const int inst = blockIdx.x;
  DiisState *__restrict__ st = sys.s[blockIdx.y].st + inst;
  if (st->done) return;
  const double tol = st->tol;
  const double *__restrict__ part = sys.s[blockIdx.y].part_e;
  double s = 0.0;
  for (int b = eboff[inst] + threadIdx.x; b < eboff[inst + 1]; b += kWave) s += part[b];
  s = wave_sum(s);
  if (threadIdx.x == 0) {
    double norm = sqrt(s / (double)(soff[inst + 1] - soff[inst]));
    int it = st->it;
    if (it <= kMaxIter) st->hist[it] = norm;
    st->it = it + 1;
    st->norm = norm;
    st->ok = norm < tol;
    st->done = (norm < tol) || (it + 1 >= maxit);
    if (st->done) instance_done(ninst > 1 ? sys.s[blockIdx.y].st + ninst : nullptr, it + 1,
                                norm < tol);
  }
}

// fresh states for the ninst instances and the summary of a batch
__global__ void k_diis_init(int ninst, double tol, DiisState *__restrict__ st) {
  const int q = blockIdx.x;
  DiisState *d = st + q;
  const int t = threadIdx.x;
  if (t < kDiisMax) d->slot[t] = t;
  if (t == 0) {
    d->nmat = 1;
    d->it = 0;
    d->done = 0;
    d->ok = q < ninst ? 0 : 1;
    d->singular = 0;
    d->pending_norm = 0;
    d->ncoef = 0;
    d->remaining = q < ninst ? 0 : ninst;
    d->norm = 0.0;
    d->tol = tol;
  }
}

// per instance: iterations as the reference reports them and the ok flag
__global__ void k_diis_results(int ninst, int maxit, const DiisState *__restrict__ st,
                               int *__restrict__ res) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ninst) return;
  res[2 * q] = st[q].ok ? st[q].it : maxit + 1;
  res[2 * q + 1] = st[q].ok;
}

__global__ void k_ldm1x(size_t nn, int nylm, const double *__restrict__ facl,
                        const double *__restrict__ x, double *__restrict__ y) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nn) y[k] = facl[k % nylm] * x[k];
}

__global__ __launch_bounds__(kJThreads) void k_hnorm_part(int nsph, int nylm,
                                                          const double *__restrict__ invl1,
                                                          const double *__restrict__ x,
                                                          double *__restrict__ part) {
  __shared__ double lds[kJThreads / kWave];
  const int wave = threadIdx.x / kWave, lane = lane_id(), wpb = kJThreads / kWave;
  double usq = 0.0;
  for (int i = blockIdx.x * wpb + wave; i < nsph; i += gridDim.x * wpb) {
    double s = 0.0;
    for (int lm = lane; lm < nylm; lm += kWave) {
      double v = x[(size_t)i * nylm + lm];
      s += invl1[lm] * v * v;
    }
    s = wave_sum(s);
    double u = sqrt(s);
    usq += u * u;
  }
  if (lane != 0) usq = 0.0;
  double tot = block_sum_256(usq, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kWave) void k_sum_parts(int nblk, double scale,
                                                     const double *__restrict__ part,
                                                     double *__restrict__ out) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kWave) s += part[b];
  s = wave_sum(s);
  if (threadIdx.x == 0) *out = scale * s;
}

// g(n,i) = -ui phi(ic) for exposed points (wghpot, ddcosmo.f90:1187-1224)
__global__ void k_wghpot(size_t np, const double *__restrict__ ui, const int *__restrict__ cavidx,
                         const double *__restrict__ phi, double *__restrict__ g) {
  size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double u = ui[p];
  g[p] = (u != 0.0) ? -u * phi[cavidx[p]] : 0.0;
}

// partial dot products over the elements of each k_jacobi block
__global__ __launch_bounds__(kJThreads) void k_dot_part(const int4 *__restrict__ jblk,
                                                        const double *__restrict__ a,
                                                        const double *__restrict__ b,
                                                        double *__restrict__ part) {
  __shared__ double lds[kJThreads / kWave];
  const int4 blk = jblk[blockIdx.x];
  double s = 0.0;
  for (size_t k = (size_t)blk.y + threadIdx.x; k < (size_t)blk.z; k += kJThreads)
    s += a[k] * b[k];
  double tot = block_sum_256(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// out[q] = scale * sum of the partials of instance q's blocks, one wave each
__global__ __launch_bounds__(kWave) void k_inst_sum(const int *__restrict__ boff, double scale,
                                                    const double *__restrict__ part,
                                                    double *__restrict__ out) {
  const int q = blockIdx.x;
  double s = 0.0;
  for (int b = boff[q] + threadIdx.x; b < boff[q + 1]; b += kWave) s += part[b];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[q] = scale * s;
}

int grid_blocks(size_t work, int per_block, int cap) {
  size_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > (size_t)cap) b = cap;
  return (int)b;
}

}  // namespace

int device_ldm1x(ddx_handle *h, hipStream_t s, const double *x, double *y) {
  size_t nn = (size_t)h->nylm * h->n;
  hipLaunchKernelGGL(k_ldm1x, dim3((nn + 255) / 256), dim3(256), 0, s, nn, h->nylm,
                     h->facl.p, x, y);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int device_hnorm(ddx_handle *h, const double *x, double *norm_host) {
  int nb = grid_blocks(h->n, kJThreads / kWave, 1024);
  DDX_TRY(h->part.reserve(1024 + 1));
  hipLaunchKernelGGL(k_hnorm_part, dim3(nb), dim3(kJThreads), 0, h->stream, h->n, h->nylm,
                     h->cf.p + h->nylm, x, h->part.p);
  DDX_HIP(hipGetLastError());
  DDX_TRY(h->energy.reserve(h->ninst + 4));
  double *out = h->energy.p + h->ninst + 1;
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kWave), 0, h->stream, nb, 1.0, h->part.p, out);
  DDX_HIP(hipGetLastError());
  double s = 0.0;
  DDX_HIP(hipMemcpyAsync(&s, out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  DDX_HIP(hipStreamSynchronize(h->stream));
  *norm_host = sqrt(s / (double)h->n);
  return DDX_OK;
}

int build_rhs(ddx_handle *h, const double *d_phi, double *d_rhs) {
  Lane &ln = h->lane[0];   // lane 0 runs on the handle's stream
  DDX_TRY(ln.P.reserve(h->np));
  hipLaunchKernelGGL(k_wghpot, dim3((h->np + 255) / 256), dim3(256), 0, h->stream, h->np,
                     h->ui.p, h->cavidx.p, d_phi, ln.P.p);
  DDX_HIP(hipGetLastError());
  // rhs(lm, i) = sum_n Y(lm, n) w_n g(n, i)   (intrhs per sphere, cosmo.f90:114-116)
  return project(h, h->stream, 1.0, ln.P.p, d_rhs, nullptr, 1);
}

int energy_dot(ddx_handle *h, const double *a, const double *b, double scale, double *d_out) {
  DDX_TRY(h->part.reserve((size_t)h->nbj + 1));
  hipLaunchKernelGGL(k_dot_part, dim3(h->nbj), dim3(kJThreads), 0, h->stream, h->jblk.p, a, b,
                     h->part.p);
  DDX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_inst_sum, dim3(h->ninst), dim3(kWave), 0, h->stream, h->jboff.p, scale,
                     h->part.p, d_out);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

// Block maps of the solver kernels, from the instances' sphere ranges: each
// instance's elements (spheres) are split evenly over
// ceil(elements / 512) (ceil(spheres / sp)) workgroups, at most 1024 each.
// sp, the spheres per k_extrap workgroup, is 4 up to
// about 8000 spheres in all and grows to 32 for larger systems and batches,
// where the per-workgroup reads of the DIIS state would dominate.
int build_solver_maps(ddx_handle *h) {
  std::vector<int4> jb, eb;
  std::vector<int> jo(h->ninst + 1), eo(h->ninst + 1);
  const long long ny = h->nylm;
  const int sp = std::max(4, std::min(32, h->n / 2048 / 4 * 4));
  for (int q = 0; q < h->ninst; ++q) {
    const int s0 = h->h_soff[q], s1 = h->h_soff[q + 1];
    const long long e0 = s0 * ny, e1 = s1 * ny;
    jo[q] = (int)jb.size();
    const int nj = grid_blocks((size_t)(e1 - e0), kJThreads * 2, 1024);
    for (int j = 0; j < nj; ++j)
      jb.push_back(make_int4(q, (int)(e0 + (e1 - e0) * j / nj), (int)(e0 + (e1 - e0) * (j + 1) / nj), 0));
    eo[q] = (int)eb.size();
    const int ne = grid_blocks((size_t)(s1 - s0), sp, 1024);
    for (int j = 0; j < ne; ++j)
      eb.push_back(make_int4(q, s0 + (int)((long long)(s1 - s0) * j / ne),
                             s0 + (int)((long long)(s1 - s0) * (j + 1) / ne), 0));
  }
  jo[h->ninst] = (int)jb.size();
  eo[h->ninst] = (int)eb.size();
  if ((long long)h->n * ny > 0x7fffffffll) {
    set_error("nylm * nsph = %lld exceeds the 32-bit element index", (long long)h->n * ny);
    return DDX_ERR_CAPACITY;
  }
  h->nbj = (int)jb.size();
  h->nbe = (int)eb.size();
  const int nblk = (h->n + 15) / 16;   // workgroups of one projection pass
  h->ysplit = nblk >= 1024 ? 1 : std::min(kProjSplit, (1024 + nblk - 1) / nblk);
  DDX_TRY(h->jblk.reserve(jb.size()));
  DDX_TRY(h->eblk.reserve(eb.size()));
  DDX_TRY(h->jboff.reserve(jo.size()));
  DDX_TRY(h->eboff.reserve(eo.size()));
  hipStream_t s = h->stream;
  DDX_HIP(hipMemcpyAsync(h->jblk.p, jb.data(), jb.size() * sizeof(int4), hipMemcpyHostToDevice, s));
  DDX_HIP(hipMemcpyAsync(h->eblk.p, eb.data(), eb.size() * sizeof(int4), hipMemcpyHostToDevice, s));
  DDX_HIP(hipMemcpyAsync(h->jboff.p, jo.data(), jo.size() * sizeof(int), hipMemcpyHostToDevice, s));
  DDX_HIP(hipMemcpyAsync(h->eboff.p, eo.data(), eo.size() * sizeof(int), hipMemcpyHostToDevice, s));
  DDX_HIP(hipStreamSynchronize(s));   // (built once per handle)
  return DDX_OK;
}

namespace {
SysArgs sys_args(const ddx_handle *h, Lane &ln) {
  SysArgs a;
  a.rhs = ln.rhs;
  a.y = ln.y.p;
  a.ny = ln.star ? 1 : h->ysplit;
  a.yp = ln.star ? ln.yp.p : nullptr;
  a.segoff = h->segoff.p;
  a.x = ln.x;
  a.xs = ln.star ? nullptr : ln.xs.p;
  a.xd = ln.xd.p;
  a.ed = ln.ed.p;
  a.xnew = ln.xnew.p;
  a.part = ln.part.p;
  a.part_e = ln.part.p + (size_t)h->nbj * kPartStride;
  a.newslot = ln.misc.p;
  a.st = ln.diis.p;
  a.yacc = ln.incr ? ln.yacc.p : nullptr;
  a.acc = 0;
  // the increment is read by the FP32 L* operator only: the L X sweep of
  // an increment round reads xs = cf inc, which k_extrap writes as well
  a.inc = (ln.incr && ln.star) ? ln.inc.p : nullptr;
  a.xs_inc = 0;
  return a;
}

// drops the timer records of one kernel class for rounds >= keep of a group
void drop_rounds(ddx_handle *h, int cls, int serial, int keep) {
  KernelTimer &t = h->ktimer[cls];
  size_t n = 0;
  for (size_t k = 0; k < t.recs.size(); ++k) {
    const KernelTimer::Rec &r = t.recs[k];
    if (r.solve == serial && r.round >= keep) {
      t.pool.push_back(r.b);
      t.pool.push_back(r.e);
    } else {
      t.recs[n++] = r;
    }
  }
  t.recs.resize(n);
}
}  // namespace

int solve_begin(ddx_handle *h, Lane &ln, int star, int diis_max, double tol, const double *d_rhs,
                double *d_x, int max_iter) {
  if (!(diis_max == 0 || (diis_max > 10 && diis_max <= kDiisMax))) {
    set_error("diis_max must be 0 or in 11..%d", kDiisMax);
    return DDX_ERR_ARG;
  }
  if (max_iter < 1 || max_iter > kMaxIter) {
    set_error("n_iter must be in 1..%d", kMaxIter);
    return DDX_ERR_ARG;
  }
  size_t nn = (size_t)h->nylm * h->n;
  DDX_TRY(ln.y.reserve(nn * kProjSplit));
  DDX_TRY(ln.xnew.reserve(nn));
  DDX_TRY(ln.xs.reserve(nn + 1));
  if (diis_max > 0) {
    DDX_TRY(ln.xd.reserve(nn * diis_max));
    DDX_TRY(ln.ed.reserve(nn * diis_max));
  }
  DDX_TRY(ln.part.reserve((size_t)h->nbj * kPartStride + h->nbe + 1));
  // the L* sweep leaves per-segment sums that k_jacobi reads: its buffer must
  // not move once the solve has started
  if (star) DDX_TRY(ln.yp.reserve((size_t)h->nseg_cap * h->nylm + 1));
  DDX_TRY(ln.diis.reserve(h->ninst + 1));
  DDX_TRY(ln.misc.reserve(h->ninst + 1));
  DDX_TRY(ln.res.reserve(2 * (size_t)h->ninst + 2));
  ln.star = star;
  ln.diis_max = diis_max;
  ln.max_iter = max_iter;
  ln.tol = tol;
  ln.rhs = d_rhs;
  ln.x = d_x;
  // with DIIS the norm of iteration k is evaluated by the DIIS kernel of
  // iteration k + 1, so up to max_iter + 1 rounds are enqueued
  ln.rounds = diis_max > 0 ? max_iter + 1 : max_iter;
  ln.enqueued = 0;
  ln.group = 3;
  ln.active = true;

  const int nst = h->ninst > 1 ? h->ninst + 1 : 1;
  hipLaunchKernelGGL(k_diis_init, dim3(nst), dim3(kWave), 0, ln.stream, h->ninst, tol,
                     ln.diis.p);
  DDX_HIP(hipGetLastError());
  // the L X sweep reads cf x: from here on k_extrap keeps it
  if (!star) DDX_TRY(cf_scale(h, ln.stream, d_x, ln.xs.p));
  return DDX_OK;
}

namespace {
// `nr` rounds of one solve, or of a direct and an adjoint solve in lockstep
int enqueue_rounds(ddx_handle *h, Lane *const *lanes, int count, int nr) {
  Lane &l0 = *lanes[0];
  const bool pair = count == 2;
  hipStream_t s = l0.stream;
  size_t nn = (size_t)h->nylm * h->n;
  const int nbj = h->nbj, nbe = h->nbe, ninst = h->ninst;
  const double *invl1 = h->cf.p + h->nylm;   // 1/(l+1) table stored after cf
  const int diis_max = l0.diis_max;
  Systems sys{};
  for (int k = 0; k < count; ++k) sys.s[k] = sys_args(h, *lanes[k]);
  for (int g = 0; g < nr; ++g, ++l0.enqueued) {
    h->tag_round = l0.enqueued;
    // mixed precision: rounds 0, kRefresh, 2 kRefresh, ... apply the FP64
    // operator to the iterate itself; the others the FP32 operator to the
    // last increment, added to the FP64 image
    const bool incr = l0.incr;
    const bool exact = !incr || l0.enqueued % kRefresh == 0;
    const bool next_exact = !incr || (l0.enqueued + 1) % kRefresh == 0;
    for (int k = 0; k < count; ++k) {
      lanes[k]->f32 = !exact;
      sys.s[k].acc = exact ? 0 : 1;
      sys.s[k].xs_inc = next_exact ? 0 : 1;
    }
    const double *xa = pair ? (exact ? lanes[1]->x : lanes[1]->inc.p) : nullptr;
    const double *x1 = exact ? l0.x : l0.inc.p;
    int rc;
    if (pair) {
      rc = apply_pair(h, l0, l0.y.p, *lanes[1], xa, h->ysplit);
    } else if (l0.star) {
      rc = apply_lstarx(h, l0, x1, l0.y.p, solve_gate(h, l0), false);
    } else {
      rc = apply_lx(h, l0, x1, l0.y.p, solve_gate(h, l0), true, h->ysplit);
    }
    for (int k = 0; k < count; ++k) lanes[k]->f32 = false;
    DDX_TRY(rc);
    {
      TimedLaunch tl(h, DDX_K_JACOBI, s);
      // mixed precision keeps the residual history in FP32
      if (l0.incr)
        hipExtLaunchKernelGGL(k_jacobi<float>, dim3(nbj, count), dim3(kJThreads), 0, s, tl.b, tl.e,
                              0, nn, h->nylm, diis_max, h->jblk.p, nbj, h->facl.p, sys);
      else
        hipExtLaunchKernelGGL(k_jacobi<double>, dim3(nbj, count), dim3(kJThreads), 0, s, tl.b,
                              tl.e, 0, nn, h->nylm, diis_max, h->jblk.p, nbj, h->facl.p, sys);
    }
    DDX_HIP(hipGetLastError());
    if (diis_max > 0) {
      TimedLaunch tl(h, DDX_K_DIIS, s);
      hipExtLaunchKernelGGL(k_diis_small, dim3(ninst, count), dim3(kSmallThreads), 0, s, tl.b,
                            tl.e, 0, diis_max, nbj, h->jboff.p, h->eboff.p, h->soff.p, ninst,
                            l0.max_iter, sys);
      DDX_HIP(hipGetLastError());
    }
    {
      TimedLaunch tl(h, DDX_K_EXTRAP, s);
      hipExtLaunchKernelGGL(k_extrap, dim3(nbe, count), dim3(kJThreads), 0, s, tl.b, tl.e, 0,
                            h->nylm, nn, diis_max, h->eblk.p, invl1, h->cf.p, sys);
    }
    DDX_HIP(hipGetLastError());
    if (diis_max == 0) {
      hipLaunchKernelGGL(k_norm, dim3(ninst, count), dim3(kWave), 0, s, h->eboff.p, h->soff.p,
                         ninst, l0.max_iter, sys);
      DDX_HIP(hipGetLastError());
    }
  }
  return DDX_OK;
}

// what a captured group of rounds depends on: every pointer and setting the
// launches take
std::vector<uintptr_t> round_key(const ddx_handle *h, Lane *const *lanes, int count, int nr) {
  std::vector<uintptr_t> k = {(uintptr_t)count, (uintptr_t)nr, (uintptr_t)h->ysplit,
                              (uintptr_t)h->nbj, (uintptr_t)h->nbe, (uintptr_t)h->nworker};
  for (int c = 0; c < count; ++c) {
    const Lane &l = *lanes[c];
    const uintptr_t v[] = {(uintptr_t)l.star, (uintptr_t)l.incr,
                           (uintptr_t)(l.incr ? lanes[0]->enqueued % kRefresh : 0),
                           (uintptr_t)l.yacc.p, (uintptr_t)l.inc.p, (uintptr_t)l.diis_max,
                           (uintptr_t)l.max_iter,
                           (uintptr_t)l.rhs, (uintptr_t)l.x, (uintptr_t)l.y.p, (uintptr_t)l.xnew.p,
                           (uintptr_t)l.xs.p, (uintptr_t)l.P.p, (uintptr_t)l.xiw.p,
                           (uintptr_t)l.rw.p, (uintptr_t)l.yp.p, (uintptr_t)l.xd.p,
                           (uintptr_t)l.ed.p, (uintptr_t)l.part.p, (uintptr_t)l.diis.p,
                           (uintptr_t)l.misc.p, (uintptr_t)l.stream};
    k.insert(k.end(), std::begin(v), std::end(v));
    uint64_t tb;
    std::memcpy(&tb, &l.tol, 8);
    k.push_back((uintptr_t)tb);
  }
  return k;
}

bool timers_on(const ddx_handle *h) {
  for (const auto &t : h->ktimer)
    if (t.on) return true;
  return false;
}

// the group of nr rounds as a HIP graph: captured on first use with these
// lanes, replayed afterwards (every round's state lives in device memory)
int launch_round_graph(ddx_handle *h, Lane *const *lanes, int count, int nr) {
  Lane &l0 = *lanes[0];
  const std::vector<uintptr_t> key = round_key(h, lanes, count, nr);
  hipGraphExec_t exec = nullptr;
  for (auto &g : h->round_graphs)
    if (g.key == key) exec = g.exec;
  if (!exec) {
    hipGraph_t graph = nullptr;
    const int first = l0.enqueued;
    DDX_HIP(hipStreamBeginCapture(l0.stream, hipStreamCaptureModeRelaxed));
    const int rc = enqueue_rounds(h, lanes, count, nr);
    const hipError_t ec = hipStreamEndCapture(l0.stream, &graph);
    l0.enqueued = first;
    if (rc != DDX_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    DDX_HIP(ec);
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DDX_HIP(ei);
    // keyed on the buffers as they stand after the capture
    h->round_graphs.push_back({round_key(h, lanes, count, nr), exec});
  }
  DDX_HIP(hipGraphLaunch(exec, l0.stream));
  l0.enqueued += nr;
  return DDX_OK;
}
}  // namespace

void drop_round_graphs(ddx_handle *h) {
  for (auto &g : h->round_graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  h->round_graphs.clear();
}

// Enqueues the next group of rounds and the copy of each lane's state header
// (nmat, it, done, ok) into slot `slot` of its pinned h_head, marked by
// h->gev[slot].  The first group of a solve (3 rounds) runs eagerly and
// makes every buffer the rounds use; eager groups after it have 2 rounds.
// With graphs on (and no kernel timer) each later group is one launch of a
// captured graph of kGraphRounds rounds, which covers the rest of a typical
// solve; the rounds after convergence return at once.
int solve_enqueue(ddx_handle *h, Lane *const *lanes, int count, int slot) {
  Lane &l0 = *lanes[0];
  const bool pair = count == 2;
  if (pair && (l0.star != 0 || lanes[1]->star != 1 || lanes[1]->stream != l0.stream ||
               lanes[1]->diis_max != l0.diis_max || lanes[1]->rounds != l0.rounds)) {
    set_error("paired solves need a direct and an adjoint solve with equal settings");
    return DDX_ERR_ARG;
  }
  hipStream_t s = l0.stream;
  h->tag_solve = l0.serial;
  const bool graph = h->use_graphs && l0.enqueued > 0 && !timers_on(h);
  const int nr = std::min(graph ? kGraphRounds : l0.group, l0.rounds - l0.enqueued);
  if (graph)
    DDX_TRY(launch_round_graph(h, lanes, count, nr));
  else
    DDX_TRY(enqueue_rounds(h, lanes, count, nr));
  h->tag_solve = -1;
  h->tag_round = 0;
  l0.group = 2;
  for (int k = 0; k < count; ++k)
    DDX_HIP(hipMemcpyAsync(lanes[k]->h_head + 4 * slot, solve_gate(h, *lanes[k]), 4 * sizeof(int),
                           hipMemcpyDeviceToHost, s));
  DDX_HIP(hipEventRecord(h->gev[slot], s));
  return DDX_OK;
}

int solve_end(ddx_handle *h, Lane &ln, int *n_iter, int *ok, std::vector<double> *hist) {
  const DiisState *hs = ln.h_diis;
  ln.active = false;
  // a batch: the summary's ok and largest count; per instance in h_res
  *n_iter = hs->ok ? hs->it : ln.max_iter + 1;
  *ok = hs->ok;
  if (hist) {
    if (h->ninst == 1) hist->assign(hs->hist, hs->hist + (hs->it < kMaxIter ? hs->it : kMaxIter));
    else hist->clear();
  }
  return DDX_OK;
}

int solve_run(ddx_handle *h, Lane *const *lanes, int count) {
  const int serial = h->solve_serial++;
  for (int k = 0; k < count; ++k) lanes[k]->serial = serial;
  // one eager group always runs ahead: while the host waits for the state
  // of group k, group k + 1 is already queued behind it, so the device does
  // not wait for the host between groups (a group queued after convergence
  // returns at once).  Of the graph groups only the first runs ahead (behind
  // the eager first group): a later one would mostly repeat finished rounds.
  int cur = 0, queued = 1, groups = 1;
  DDX_TRY(solve_enqueue(h, lanes, count, cur));
  while (true) {
    if (lanes[0]->enqueued < lanes[0]->rounds && queued == 1 &&
        (!h->use_graphs || groups == 1)) {
      DDX_TRY(solve_enqueue(h, lanes, count, cur ^ 1));
      ++queued;
      ++groups;
    }
    DDX_HIP(hipEventSynchronize(h->gev[cur]));
    --queued;
    bool all = true;
    for (int k = 0; k < count; ++k) all = all && lanes[k]->h_head[4 * cur + 2] != 0;
    if (all) break;
    if (queued == 0) {
      if (lanes[0]->enqueued >= lanes[0]->rounds) break;
      DDX_TRY(solve_enqueue(h, lanes, count, cur ^ 1));
      ++queued;
      ++groups;
    }
    cur ^= 1;
  }
  for (int k = 0; k < count; ++k) {
    Lane &ln = *lanes[k];
    DDX_HIP(hipMemcpyAsync(ln.h_diis, solve_gate(h, ln), sizeof(DiisState),
                           hipMemcpyDeviceToHost, ln.stream));
    if (h->ninst > 1) {
      hipLaunchKernelGGL(k_diis_results, dim3((h->ninst + 255) / 256), dim3(256), 0, ln.stream,
                         h->ninst, ln.max_iter, ln.diis.p, ln.res.p);
      DDX_HIP(hipGetLastError());
      ln.h_res.resize(2 * (size_t)h->ninst);
      DDX_HIP(hipMemcpyAsync(ln.h_res.data(), ln.res.p, 2 * sizeof(int) * h->ninst,
                             hipMemcpyDeviceToHost, ln.stream));
    }
  }
  DDX_HIP(hipStreamSynchronize(lanes[0]->stream));
  // launches of rounds after a solve's last iteration returned at once:
  // drop them from the timers
  int it[2] = {0, 0}, itmax = 0;
  for (int k = 0; k < count; ++k) {
    it[k] = lanes[k]->h_diis->it;
    if (it[k] > itmax) itmax = it[k];
  }
  const int shared[3] = {DDX_K_JACOBI, DDX_K_DIIS, DDX_K_EXTRAP};
  for (int c : shared) drop_rounds(h, c, serial, itmax);
  if (count == 2) {
    drop_rounds(h, DDX_K_SWEEP_PAIR, serial, itmax);
    drop_rounds(h, DDX_K_PROJECT, serial, it[0]);
    drop_rounds(h, DDX_K_EXPAND, serial, it[1]);
  } else {
    const bool star = lanes[0]->star;
    drop_rounds(h, star ? DDX_K_SWEEP_LSTAR : DDX_K_SWEEP_L, serial, it[0]);
    drop_rounds(h, star ? DDX_K_EXPAND : DDX_K_PROJECT, serial, it[0]);
  }
  return DDX_OK;
}

// Mixed precision: one Jacobi/DIIS iteration whose operator image L x
// (L* x) is carried in FP64 and refreshed exactly.  Round 0 and every
// kRefresh-th round apply the FP64 operator to the iterate; the rounds in
// between apply the FP32 sweeps and products to the last increment
// x_new - x and add the result to the FP64 image.  The Jacobi update, the
// residual rhs - L x, the iterate history and extrapolation, the norms and
// the convergence test are the FP64 path's (jacobi_diis.f90:102-183): the
// FP32 rounding enters only through increments, which shrink with the
// iteration, so the error carried in the image stays near 1e-7 of the
// increments since the last refresh and the iteration counts stay those of
// FP64 to within one or two.  The residual history, which only sets the
// extrapolation coefficients (its Gram matrix, accumulated in FP64, and the
// bordered solve), is stored in FP32: half the bytes of the Gram-row pass.
int solve_run_mixed(ddx_handle *h, Lane *const *lanes, int count) {
  DDX_TRY(make_tables32(h));
  const size_t nn = (size_t)h->nylm * h->n;
  for (int k = 0; k < count; ++k) {
    DDX_TRY(lanes[k]->yacc.reserve(nn + 1));
    DDX_TRY(lanes[k]->inc.reserve(nn + 1));
    lanes[k]->incr = true;
  }
  const int rc = solve_run(h, lanes, count);
  for (int k = 0; k < count; ++k) lanes[k]->incr = false;
  return rc;
}

int solve_lanes(ddx_handle *h, Lane *const *lanes, int count) {
  return h->prec == DDX_PREC_MIXED ? solve_run_mixed(h, lanes, count)
                                   : solve_run(h, lanes, count);
}

int solve_jacobi_diis(ddx_handle *h, Lane &ln, int star, int diis_max, double tol,
                      const double *d_rhs, double *d_x, int max_iter, int *n_iter, int *ok,
                      std::vector<double> *hist) {
  DDX_TRY(solve_begin(h, ln, star, diis_max, tol, d_rhs, d_x, max_iter));
  Lane *lp = &ln;
  DDX_TRY(solve_run(h, &lp, 1));
  return solve_end(h, ln, n_iter, ok, hist);
}

}  // namespace ddx
