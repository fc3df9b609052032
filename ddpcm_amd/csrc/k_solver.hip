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
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

constexpr int kJThreads = 256;
constexpr int kPartStride = kDiisMax + 1;   // 25 dots + self dot per block

__global__ __launch_bounds__(kJThreads) void k_jacobi(
    size_t nn, int nylm, int diis_max, const double *__restrict__ facl,
    const double *__restrict__ rhs, const double *__restrict__ y, const double *__restrict__ x,
    double *__restrict__ xd, double *__restrict__ ed, double *__restrict__ xnew,
    double *__restrict__ part, const DiisState *__restrict__ st) {
  if (st->done) return;
  __shared__ double lds[kJThreads / kWave];
  const int nmat = st->nmat;
  const int off = (diis_max > 0 && nmat >= diis_max) ? 10 : 0;
  const int ndot = diis_max > 0 ? nmat - off - 1 : 0;
  double *xo = diis_max > 0 ? xd + (size_t)st->slot[nmat - 1] * nn : xnew;
  double *eo = diis_max > 0 ? ed + (size_t)st->slot[nmat - 1] * nn : nullptr;
  const double *eh[kDiisMax];
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q)
    eh[q] = (q < ndot) ? ed + (size_t)st->slot[off + q] * nn : nullptr;
  double acc[kDiisMax];
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q) acc[q] = 0.0;
  double self = 0.0;
  for (size_t k = (size_t)blockIdx.x * kJThreads + threadIdx.x; k < nn;
       k += (size_t)gridDim.x * kJThreads) {
    double xn = facl[k % nylm] * (rhs[k] - y[k]);
    double e = xn - x[k];
    xo[k] = xn;
    if (eo) {
      eo[k] = e;
#pragma unroll
      for (int q = 0; q < kDiisMax; ++q)
        if (q < ndot) acc[q] += eh[q][k] * e;
      self += e * e;
    }
  }
  if (diis_max == 0) return;
  double *pb = part + (size_t)blockIdx.x * kPartStride;
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q) {
    if (q < ndot) {
      double s = block_sum_256(acc[q], lds);
      if (threadIdx.x == 0) pb[q] = s;
    }
  }
  double s = block_sum_256(self, lds);
  if (threadIdx.x == 0) pb[kDiisMax] = s;
}

// Gauss-Jordan with full pivoting on the (n1 x n1) bordered matrix in LDS,
// right-hand side e_1 (gjinv, jacobi_diis.f90:299-400).  Pivot choice: the
// largest |a(j,k)| over rows with piv != 1 and columns with piv == 0, ties to
// the first in (j, k) order, as the reference's nested loops pick it.
// Returns 1 on success; rhs holds the solution.
__device__ int gj_solve(int n1, double *a /* n1*n1, column-major */, double *bv,
                        int *piv, double *dj) {
  const int lane = threadIdx.x;
  for (int j = lane; j < n1; j += kWave) piv[j] = 0;
  __syncthreads();
  for (int it = 0; it < n1; ++it) {
    // parallel argmax over the linear index (j-major), first maximum wins
    double big = 0.0;
    int best = 0x7fffffff;
    for (int idx = lane; idx < n1 * n1; idx += kWave) {
      int j = idx / n1, k = idx % n1;
      if (piv[j] != 1 && piv[k] == 0) {
        double v = fabs(a[j + k * n1]);
        if (v > big) { big = v; best = idx; }
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      double ob = __shfl_xor(big, o, kWave);
      int oi = __shfl_xor(best, o, kWave);
      if (ob > big || (ob == big && oi < best)) { big = ob; best = oi; }
    }
    if (best == 0x7fffffff) return 0;   // nothing left that is non-zero
    const int irow = best / n1, icol = best % n1;
    __syncthreads();
    if (piv[icol] + 1 > 1) return 0;
    __syncthreads();
    if (lane == 0) piv[icol] += 1;
    if (irow != icol) {
      for (int c = lane; c < n1; c += kWave) {
        double t = a[irow + c * n1];
        a[irow + c * n1] = a[icol + c * n1];
        a[icol + c * n1] = t;
      }
      if (lane == 0) { double t = bv[irow]; bv[irow] = bv[icol]; bv[icol] = t; }
    }
    __syncthreads();
    double d = a[icol + icol * n1];
    if (d == 0.0) return 0;
    double pinv = 1.0 / d;
    __syncthreads();
    if (lane == 0) a[icol + icol * n1] = 1.0;
    __syncthreads();
    for (int c = lane; c < n1; c += kWave) a[icol + c * n1] *= pinv;
    if (lane == 0) bv[icol] *= pinv;
    __syncthreads();
    for (int j = lane; j < n1; j += kWave) dj[j] = a[j + icol * n1];
    __syncthreads();
    for (int idx = lane; idx < n1 * n1; idx += kWave) {
      int j = idx % n1, c = idx / n1;
      if (j == icol) continue;
      double aj = (c == icol) ? 0.0 : a[j + c * n1];
      a[j + c * n1] = aj - a[icol + c * n1] * dj[j];
    }
    __syncthreads();
    for (int j = lane; j < n1; j += kWave)
      if (j != icol) bv[j] = bv[j] - bv[icol] * dj[j];
    __syncthreads();
  }
  return 1;
}

__global__ __launch_bounds__(kWave) void k_diis_small(int nblk, int diis_max,
                                                      const double *__restrict__ part,
                                                      DiisState *__restrict__ st,
                                                      int *__restrict__ newslot_out) {
  if (st->done) return;
  constexpr int LB = kDiisMax + 1;
  __shared__ double dots[kDiisMax + 1];
  __shared__ double sb[LB * LB];          // B (1-based in the reference), 0-based here
  __shared__ double a[LB * LB];
  __shared__ double bv[LB];
  __shared__ double dj[LB];
  __shared__ int piv[LB];
  __shared__ int slots[kDiisMax];
  const int lane = threadIdx.x;
  const int lenb = diis_max + 1;
  int nmat = st->nmat;
  const int off = (nmat >= diis_max) ? 10 : 0;
  const int ndot = nmat - off - 1;
  // Gram-row dot products from the per-block partial sums (fixed order)
  if (lane <= kDiisMax && (lane < ndot || lane == kDiisMax)) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)b * kPartStride + lane];
    dots[lane] = s;
  }
  for (int idx = lane; idx < lenb * lenb; idx += kWave) sb[idx] = st->b[idx];
  for (int j = lane; j < diis_max; j += kWave) slots[j] = st->slot[j];
  __syncthreads();
  if (lane == 0) *newslot_out = slots[nmat - 1];
#define SB(r, c) sb[((r) - 1) + ((c) - 1) * lenb]
  if (nmat >= diis_max) {
    // drop the 10 oldest vectors (jacobi_diis.f90:216-227).  All sources of
    // the B shift lie at indices >= 12 and are read before any of them is
    // overwritten in the reference's loop order, so a gather-then-scatter
    // copy is the same operation.
    const int m = nmat - 10;
    double v[8];
    int cnt = 0;
    for (int idx = lane; idx < (m - 1) * (m - 1); idx += kWave) {
      int j = 2 + idx % (m - 1), k = 2 + idx / (m - 1);
      v[cnt++] = SB(j + 10, k + 10);
    }
    __syncthreads();
    cnt = 0;
    for (int idx = lane; idx < (m - 1) * (m - 1); idx += kWave) {
      int j = 2 + idx % (m - 1), k = 2 + idx / (m - 1);
      SB(j, k) = v[cnt++];
    }
    int sl = 0;
    if (lane < diis_max) {
      int j = lane + 1;
      sl = (j <= m) ? slots[j + 9] : slots[j - m - 1];
    }
    __syncthreads();
    if (lane < diis_max) slots[lane] = sl;
    nmat = m;
    __syncthreads();
  }
  // makeb (jacobi_diis.f90:258-297)
  if (nmat == 1) {
    if (lane == 0) {
      SB(1, 1) = 0.0;
      SB(1, 2) = 1.0;
      SB(2, 1) = 1.0;
      SB(2, 2) = dots[kDiisMax];
    }
  } else {
    if (lane == 0) {
      SB(nmat + 1, 1) = 1.0;
      SB(1, nmat + 1) = 1.0;
      SB(nmat + 1, nmat + 1) = dots[kDiisMax];
    }
    for (int i = 1 + lane; i <= nmat - 1; i += kWave) {
      SB(nmat + 1, i + 1) = dots[i - 1];
      SB(i + 1, nmat + 1) = dots[i - 1];
    }
  }
  __syncthreads();
  const int n1 = nmat + 1;
  for (int idx = lane; idx < n1 * n1; idx += kWave) {
    int r = idx % n1, c = idx / n1;
    a[idx] = SB(r + 1, c + 1);
  }
  for (int j = lane; j < n1; j += kWave) bv[j] = (j == 0) ? 1.0 : 0.0;
#undef SB
  __syncthreads();
  int ok = gj_solve(n1, a, bv, piv, dj);
  __syncthreads();
  // write back B and the slot map
  for (int idx = lane; idx < lenb * lenb; idx += kWave) st->b[idx] = sb[idx];
  for (int j = lane; j < diis_max; j += kWave) st->slot[j] = slots[j];
  if (ok) {
    for (int i = 1 + lane; i <= nmat; i += kWave) st->coef[i - 1] = bv[i];
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
// per sphere (hsnorm, ddcosmo.f90:1233-1271), x <- x_new.  One wave per sphere.
__global__ __launch_bounds__(kJThreads) void k_extrap(
    int nsph, int nylm, size_t nn, int diis_max, const double *__restrict__ xd,
    const double *__restrict__ xnew, const int *__restrict__ newslot,
    const double *__restrict__ invl1, double *__restrict__ x, double *__restrict__ part,
    const DiisState *__restrict__ st) {
  if (st->done) return;
  __shared__ double lds[kJThreads / kWave];
  const int ncoef = diis_max > 0 ? st->ncoef : 0;
  const double *xs[kDiisMax];
  double cs[kDiisMax];
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q) {
    xs[q] = (q < ncoef) ? xd + (size_t)st->slot[q] * nn : nullptr;
    cs[q] = (q < ncoef) ? st->coef[q] : 0.0;
  }
  const double *xj = diis_max > 0 ? xd + (size_t)(*newslot) * nn : xnew;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  const int wpb = kJThreads / kWave;
  double usq = 0.0;
  for (int i = blockIdx.x * wpb + wave; i < nsph; i += gridDim.x * wpb) {
    double s = 0.0;
    for (int lm = lane; lm < nylm; lm += kWave) {
      size_t k = (size_t)i * nylm + lm;
      double xn;
      if (ncoef > 0) {
        xn = 0.0;
#pragma unroll
        for (int q = 0; q < kDiisMax; ++q)
          if (q < ncoef) xn = xn + cs[q] * xs[q][k];
      } else {
        xn = xj[k];
      }
      double inc = xn - x[k];
      x[k] = xn;
      s += invl1[lm] * inc * inc;
    }
    s = wave_sum(s);
    double u = sqrt(s);
    usq += u * u;
  }
  // usq is wave-uniform: take lane 0 of each wave
  if (lane_id() != 0) usq = 0.0;
  double tot = block_sum_256(usq, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kWave) void k_norm(int nblk, int nsph, double tol, int maxit,
                                                const double *__restrict__ part,
                                                DiisState *__restrict__ st) {
  if (st->done) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kWave) s += part[b];
  s = wave_sum(s);
  if (threadIdx.x == 0) {
    double norm = sqrt(s / (double)nsph);
    int it = st->it;
    if (it <= kMaxIter) st->hist[it] = norm;
    st->it = it + 1;
    st->norm = norm;
    st->ok = norm < tol;
    st->done = (norm < tol) || (it + 1 >= maxit);
  }
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

__global__ __launch_bounds__(kJThreads) void k_dot_part(size_t nn, const double *__restrict__ a,
                                                        const double *__restrict__ b,
                                                        double *__restrict__ part) {
  __shared__ double lds[kJThreads / kWave];
  double s = 0.0;
  for (size_t k = (size_t)blockIdx.x * kJThreads + threadIdx.x; k < nn;
       k += (size_t)gridDim.x * kJThreads)
    s += a[k] * b[k];
  double tot = block_sum_256(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

int grid_blocks(size_t work, int per_block, int cap) {
  size_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > (size_t)cap) b = cap;
  return (int)b;
}

}  // namespace

int device_ldm1x(ddx_handle *h, const double *x, double *y) {
  size_t nn = (size_t)h->nylm * h->n;
  hipLaunchKernelGGL(k_ldm1x, dim3((nn + 255) / 256), dim3(256), 0, h->stream, nn, h->nylm,
                     h->facl.p, x, y);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int device_hnorm(ddx_handle *h, const double *x, double *norm_host) {
  int nb = grid_blocks(h->n, kJThreads / kWave, 1024);
  DDX_TRY(h->part.reserve((size_t)1024 * kPartStride + 2048));
  double *part = h->part.p + (size_t)1024 * kPartStride;
  hipLaunchKernelGGL(k_hnorm_part, dim3(nb), dim3(kJThreads), 0, h->stream, h->n, h->nylm,
                     h->cf.p + h->nylm, x, part);
  DDX_HIP(hipGetLastError());
  DDX_TRY(h->energy.reserve(4));
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kWave), 0, h->stream, nb, 1.0, part, h->energy.p + 1);
  DDX_HIP(hipGetLastError());
  double s = 0.0;
  DDX_HIP(hipMemcpyAsync(&s, h->energy.p + 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  DDX_HIP(hipStreamSynchronize(h->stream));
  *norm_host = sqrt(s / (double)h->n);
  return DDX_OK;
}

int build_rhs(ddx_handle *h, const double *d_phi, double *d_rhs) {
  DDX_TRY(h->P.reserve(h->np));
  hipLaunchKernelGGL(k_wghpot, dim3((h->np + 255) / 256), dim3(256), 0, h->stream, h->np,
                     h->ui.p, h->cavidx.p, d_phi, h->P.p);
  DDX_HIP(hipGetLastError());
  // rhs(lm, i) = sum_n Y(lm, n) w_n g(n, i)   (intrhs per sphere, cosmo.f90:114-116)
  return project(h, 1.0, h->P.p, nullptr, d_rhs, nullptr);
}

int energy_dot(ddx_handle *h, const double *a, const double *b, double scale, double *d_out) {
  size_t nn = (size_t)h->nylm * h->n;
  int nb = grid_blocks(nn, kJThreads * 4, 1024);
  DDX_TRY(h->part.reserve((size_t)1024 * kPartStride + 2048));
  double *part = h->part.p + (size_t)1024 * kPartStride;
  hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(kJThreads), 0, h->stream, nn, a, b, part);
  DDX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kWave), 0, h->stream, nb, scale, part, d_out);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int solve_jacobi_diis(ddx_handle *h, int star, int diis_max, double tol, const double *d_rhs,
                      double *d_x, int max_iter, int *n_iter, int *ok,
                      std::vector<double> *hist) {
  if (!(diis_max == 0 || (diis_max > 10 && diis_max <= kDiisMax))) {
    set_error("diis_max must be 0 or in 11..%d", kDiisMax);
    return DDX_ERR_ARG;
  }
  if (max_iter < 1 || max_iter > kMaxIter) {
    set_error("n_iter must be in 1..%d", kMaxIter);
    return DDX_ERR_ARG;
  }
  hipStream_t s = h->stream;
  size_t nn = (size_t)h->nylm * h->n;
  int nbj = grid_blocks(nn, kJThreads * 2, 1024);
  int nbe = grid_blocks(h->n, kJThreads / kWave, 1024);
  DDX_TRY(h->y.reserve(nn));
  DDX_TRY(h->xnew.reserve(nn));
  if (diis_max > 0) {
    DDX_TRY(h->xd.reserve(nn * diis_max));
    DDX_TRY(h->ed.reserve(nn * diis_max));
  }
  DDX_TRY(h->part.reserve((size_t)1024 * kPartStride + 2048));
  DDX_TRY(h->diis.reserve(2));
  DDX_TRY(h->misc.reserve(16));
  double *part_e = h->part.p + (size_t)1024 * kPartStride;
  int *newslot = h->misc.p;

  DiisState *hs = h->h_diis;
  hs->nmat = 1;
  hs->it = 0;
  hs->done = 0;
  hs->ok = 0;
  hs->singular = 0;
  hs->ncoef = 0;
  for (int q = 0; q < kDiisMax; ++q) hs->slot[q] = q;
  hs->norm = 0.0;
  DiisState *ds = h->diis.p;
  DDX_HIP(hipMemcpyAsync(ds, hs, sizeof(DiisState), hipMemcpyHostToDevice, s));

  const double *invl1 = h->cf.p + h->nylm;   // 1/(l+1) table stored after cf
  int enqueued = 0, group = 3;
  while (true) {
    for (int g = 0; g < group && enqueued < max_iter; ++g, ++enqueued) {
      if (star) DDX_TRY(apply_lstarx(h, d_x, h->y.p, ds));
      else DDX_TRY(apply_lx(h, d_x, h->y.p, ds));
      hipLaunchKernelGGL(k_jacobi, dim3(nbj), dim3(kJThreads), 0, s, nn, h->nylm, diis_max,
                         h->facl.p, d_rhs, h->y.p, d_x, h->xd.p, h->ed.p, h->xnew.p, h->part.p, ds);
      DDX_HIP(hipGetLastError());
      if (diis_max > 0) {
        hipLaunchKernelGGL(k_diis_small, dim3(1), dim3(kWave), 0, s, nbj, diis_max, h->part.p, ds,
                           newslot);
        DDX_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(k_extrap, dim3(nbe), dim3(kJThreads), 0, s, h->n, h->nylm, nn, diis_max,
                         h->xd.p, h->xnew.p, newslot, invl1, d_x, part_e, ds);
      DDX_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_norm, dim3(1), dim3(kWave), 0, s, nbe, h->n, tol, max_iter, part_e, ds);
      DDX_HIP(hipGetLastError());
    }
    DDX_HIP(hipMemcpyAsync(hs, ds, sizeof(DiisState), hipMemcpyDeviceToHost, s));
    DDX_HIP(hipStreamSynchronize(s));
    if (hs->done || enqueued >= max_iter) break;
    group = 2;
  }
  *n_iter = hs->ok ? hs->it : max_iter + 1;
  *ok = hs->ok;
  if (hist) {
    hist->assign(hs->hist, hs->hist + (hs->it < kMaxIter ? hs->it : kMaxIter));
  }
  return DDX_OK;
}

}  // namespace ddx
