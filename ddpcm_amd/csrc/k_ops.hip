// jacobi_diis behind the reference's plugin interface (jacobi_diis.f90:43-191):
// the caller brings the operator as callbacks with the reference's Fortran
// argument passing -- subroutine matvec(n, x, y), subroutine dm1vec(n, x, y),
// real*8 function u_norm(n, x) -- and picks the convergence norm (1..4).
//
// Two paths:
//   * the module's own operators (ddx_op_lx / ddx_op_lstarx, ddx_op_ldm1x,
//     ddx_op_hnorm with norm 4) run the device solver of the handle
//     (solve_jacobi_diis): operators, update, DIIS and norms all in HBM;
//   * any other callbacks run on the host, where the caller's operator lives.
//     The iterate and the DIIS history stay in HBM: each iteration uploads
//     the Jacobi step D^-1 (rhs - O x), and device kernels store it, form the
//     new Gram row, extrapolate and take the increment norms.  Only the
//     bordered (nmat+1)^2 system of makeb/gjinv is solved between the
//     callbacks, on the host, with the reference's pivoting.
// Iteration counts follow the reference: n_iter + 1 when the loop ran out.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

constexpr int kOT = 256;          // threads per workgroup
constexpr int kOMaxBlocks = 1024;
constexpr int kOVals = kDiisMax + 1;  // Gram-row dots + self dot

struct Hist25 {
  const double *p[kDiisMax];
};
struct Coef25 {
  const double *p[kDiisMax];
  double c[kDiisMax];
};

// the handle of the ddx_jacobi_diis_ops call running on this thread, for
// the ddx_op_* callbacks, and the first error one of them met
thread_local ddx_handle *t_ops_h = nullptr;
thread_local int t_ops_rc = DDX_OK;

__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) sh[w] = v;
  __syncthreads();
  double s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return s;
}

__device__ __forceinline__ double block_max(double v, double *sh) {
  for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) sh[w] = v;
  __syncthreads();
  double s = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
  __syncthreads();
  return s;
}

// x_diis(:, new) = xn, e_diis(:, new) = xn - x; partial dots of the new e
// with the ndot older ones (eh) and with itself: part[q * nblk + b]
__global__ __launch_bounds__(kOT) void k_ops_store(size_t n, const double *__restrict__ xn,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ xo,
                                                   double *__restrict__ eo, int ndot, Hist25 eh,
                                                   double *__restrict__ part) {
  __shared__ double sh[kOT / kWave];
  double acc[kDiisMax];
#pragma unroll
  for (int q = 0; q < kDiisMax; ++q) acc[q] = 0.0;
  double self = 0.0;
  for (size_t k = (size_t)blockIdx.x * kOT + threadIdx.x; k < n; k += (size_t)gridDim.x * kOT) {
    const double v = xn[k], e = v - x[k];
    xo[k] = v;
    eo[k] = e;
#pragma unroll
    for (int q = 0; q < kDiisMax; ++q)
      if (q < ndot) acc[q] += eh.p[q][k] * e;
    self += e * e;
  }
  for (int q = 0; q < kDiisMax; ++q)
    if (q < ndot) {
      const double s = block_sum(acc[q], sh);
      if (threadIdx.x == 0) part[(size_t)q * gridDim.x + blockIdx.x] = s;
    }
  const double s = block_sum(self, sh);
  if (threadIdx.x == 0) part[(size_t)kDiisMax * gridDim.x + blockIdx.x] = s;
}

// x_new = sum_i c_i x_diis(:, i) (ncoef > 0) or xn; inc = x_new - x;
// partial sum of inc^2 and max |inc| per block (part[b], part[nblk + b]);
// x <- x_new
__global__ __launch_bounds__(kOT) void k_ops_extrap(size_t n, int ncoef, Coef25 cf,
                                                    const double *__restrict__ xn,
                                                    double *__restrict__ x,
                                                    double *__restrict__ inc,
                                                    double *__restrict__ part) {
  __shared__ double sh[kOT / kWave];
  double sq = 0.0, mx = 0.0;
  for (size_t k = (size_t)blockIdx.x * kOT + threadIdx.x; k < n; k += (size_t)gridDim.x * kOT) {
    double v;
    if (ncoef > 0) {
      v = 0.0;
      for (int i = 0; i < ncoef; ++i) v = v + cf.c[i] * cf.p[i][k];
    } else {
      v = xn[k];
    }
    const double d = v - x[k];
    inc[k] = d;
    x[k] = v;
    sq += d * d;
    mx = fmax(mx, fabs(d));
  }
  const double s = block_sum(sq, sh);
  const double m = block_max(mx, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s;
    part[gridDim.x + blockIdx.x] = m;
  }
}

// out[q] = sum (or max, for q >= nsum) over the nblk partials of value q,
// in block order
__global__ void k_ops_reduce(int nblk, int nval, int nsum, const double *__restrict__ part,
                             double *__restrict__ out) {
  const int q = threadIdx.x;
  if (q >= nval) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b)
    s = q < nsum ? s + part[(size_t)q * nblk + b] : fmax(s, part[(size_t)q * nblk + b]);
  out[q] = s;
}

// Gauss-Jordan with full pivoting on the bordered system a (n x n,
// column-major) with right-hand side b (gjinv, jacobi_diis.f90:299-400):
// pivot = the largest |a(j,k)| over rows j and columns k not yet pivoted,
// first in row-major order on ties.  False on a singular matrix, with the
// reference's warning.
bool gjinv_host(int n, double *a, double *b) {
  auto A = [&](int r, int c) -> double & { return a[(r - 1) + (size_t)(c - 1) * n]; };
  int piv[kOVals + 1] = {0};
  int irow = 0, icol = 0;
  for (int i = 1; i <= n; ++i) {
    double big = 0.0;
    for (int j = 1; j <= n; ++j) {
      if (piv[j] == 1) continue;
      for (int k = 1; k <= n; ++k)
        if (piv[k] == 0 && std::fabs(A(j, k)) > big) {
          big = std::fabs(A(j, k));
          irow = j;
          icol = k;
        }
    }
    if (icol < 1) {
      std::printf(" warning: singular matrix in gjinv!\n");
      std::fflush(stdout);
      return false;
    }
    piv[icol] += 1;
    if (piv[icol] > 1) {
      std::printf(" warning: singular matrix in gjinv!\n");
      std::fflush(stdout);
      return false;
    }
    if (irow != icol) {
      for (int c = 1; c <= n; ++c) std::swap(A(irow, c), A(icol, c));
      std::swap(b[irow - 1], b[icol - 1]);
    }
    if (A(icol, icol) == 0.0) {
      std::printf(" warning: singular matrix in gjinv!\n");
      std::fflush(stdout);
      return false;
    }
    const double pinv = 1.0 / A(icol, icol);
    A(icol, icol) = 1.0;
    for (int c = 1; c <= n; ++c) A(icol, c) = A(icol, c) * pinv;
    b[icol - 1] = b[icol - 1] * pinv;
    for (int j = 1; j <= n; ++j) {
      if (j == icol) continue;
      const double dum = A(j, icol);
      A(j, icol) = 0.0;
      for (int c = 1; c <= n; ++c) A(j, c) = A(j, c) - A(icol, c) * dum;
      b[j - 1] = b[j - 1] - b[icol - 1] * dum;
    }
  }
  return true;
}

// Fortran Dw.d edit descriptor for |exponent| < 100: [-]0.dddD+ee, right-
// justified in w columns
void fortran_d(char *out, size_t cap, double v, int w, int d) {
  char mant[64];
  int e = 0;
  if (v == 0.0) {
    std::snprintf(mant, sizeof(mant), "0.%0*d", d, 0);
  } else {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.*e", d - 1, std::fabs(v));
    // buf = "D.DDDe+XX": the digits without the point, then the exponent
    char digits[32];
    int nd = 0;
    const char *p = buf;
    for (; *p && *p != 'e'; ++p)
      if (*p >= '0' && *p <= '9') digits[nd++] = *p;
    digits[nd] = 0;
    e = std::atoi(p + 1) + 1;
    std::snprintf(mant, sizeof(mant), "%s0.%s", v < 0.0 ? "-" : "", digits);
  }
  char full[96];
  std::snprintf(full, sizeof(full), "%sD%c%02d", mant, e < 0 ? '-' : '+', e < 0 ? -e : e);
  std::snprintf(out, cap, "%*s", w, full);
}

// one residual line of jacobi_diis.f90:161-174
void print_norm_line(int norm, int it, double rms_norm, double max_norm) {
  char a[32], b[32];
  if (norm == 1) {
    fortran_d(a, sizeof(a), max_norm, 14, 4);
    std::printf("  iter=%4d residual norm (max): %s\n", it, a);
  } else if (norm == 2) {
    fortran_d(a, sizeof(a), rms_norm, 14, 4);
    std::printf("  iter=%4d residual norm (rms): %s\n", it, a);
  } else if (norm == 3) {
    fortran_d(a, sizeof(a), rms_norm, 14, 4);
    fortran_d(b, sizeof(b), max_norm, 14, 4);
    std::printf("  iter=%4d residual norm (rms,max): %s%s\n", it, a, b);
  } else {
    fortran_d(a, sizeof(a), rms_norm, 14, 4);
    std::printf("  iter=%4d residual norm: %s\n", it, a);
  }
  std::fflush(stdout);
}

bool module_operators(ddx_matvec_fn matvec, ddx_matvec_fn dm1vec, int norm, ddx_norm_fn u_norm) {
  return (matvec == ddx_op_lx || matvec == ddx_op_lstarx) && dm1vec == ddx_op_ldm1x &&
         norm == 4 && u_norm == ddx_op_hnorm;
}

// the host-callback path
int ops_generic(ddx_handle *h, int n, int lprint, int diis_max, int norm, double tol,
                const double *rhs, double *x, int *n_iter, int *ok_out, ddx_matvec_fn matvec,
                ddx_matvec_fn dm1vec, ddx_norm_fn u_norm, std::vector<double> &hist) {
  hipStream_t s = h ? h->stream : nullptr;
  const size_t nn = (size_t)n;
  const int nblk = (int)std::min<size_t>(kOMaxBlocks, (nn + kOT - 1) / kOT);
  const bool dodiis = diis_max != 0;
  DevArray<double> dx, dxn, dinc, dxd, ded, part, red;
  DDX_TRY(dx.reserve(nn));
  DDX_TRY(dxn.reserve(nn));
  DDX_TRY(dinc.reserve(nn));
  DDX_TRY(part.reserve((size_t)kOVals * nblk));
  DDX_TRY(red.reserve(kOVals));
  if (dodiis) {
    DDX_TRY(dxd.reserve(nn * diis_max));
    DDX_TRY(ded.reserve(nn * diis_max));
  }
  std::vector<double> y(nn), xn(nn), inc(norm == 4 ? nn : 0), vals(kOVals);
  DDX_HIP(hipMemcpyAsync(dx.p, x, nn * sizeof(double), hipMemcpyHostToDevice, s));
  // DIIS bookkeeping (jacobi_diis.f90:198-297): logical index l (1-based)
  // of the history lives in physical slot phys[l - 1]; b is the bordered
  // Gram matrix, 1-based, lenb x lenb
  const int lenb = diis_max + 1;
  std::vector<double> b((size_t)lenb * lenb + 1, 0.0);
  auto B = [&](int r, int c) -> double & { return b[(size_t)(r - 1) + (size_t)(c - 1) * lenb]; };
  int phys[kDiisMax];
  for (int k = 0; k < kDiisMax; ++k) phys[k] = k;
  int nmat = 1;
  const double tol_max = 10.0 * tol;
  int it = 1, ok = 0;
  const int maxit = *n_iter;
  hist.clear();
  for (it = 1; it <= maxit; ++it) {
    t_ops_rc = DDX_OK;
    matvec(&n, x, y.data());
    for (size_t k = 0; k < nn; ++k) y[k] = rhs[k] - y[k];
    dm1vec(&n, y.data(), xn.data());
    DDX_TRY(t_ops_rc);
    DDX_HIP(hipMemcpyAsync(dxn.p, xn.data(), nn * sizeof(double), hipMemcpyHostToDevice, s));
    Coef25 cf{};
    int ncoef = 0;
    if (dodiis) {
      // store at logical nmat; with nmat >= diis_max the 10 oldest go
      // (jacobi_diis.f90:216-227), so the new row dots with the survivors
      const int off = nmat >= diis_max ? 10 : 0;
      const int ndot = nmat - off - 1;
      Hist25 eh{};
      for (int q = 0; q < ndot; ++q) eh.p[q] = ded.p + (size_t)phys[off + q] * nn;
      double *xo = dxd.p + (size_t)phys[nmat - 1] * nn, *eo = ded.p + (size_t)phys[nmat - 1] * nn;
      hipLaunchKernelGGL(k_ops_store, dim3(nblk), dim3(kOT), 0, s, nn, dxn.p, dx.p, xo, eo, ndot,
                         eh, part.p);
      DDX_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_ops_reduce, dim3(1), dim3(64), 0, s, nblk, kOVals, kOVals, part.p,
                         red.p);
      DDX_HIP(hipGetLastError());
      DDX_HIP(hipMemcpyAsync(vals.data(), red.p, kOVals * sizeof(double), hipMemcpyDeviceToHost,
                             s));
      DDX_HIP(hipStreamSynchronize(s));
      if (off) {
        for (int j = 2; j <= nmat - 10; ++j)
          for (int k = 2; k <= nmat - 10; ++k) B(j, k) = B(j + 10, k + 10);
        int rot[kDiisMax];
        for (int j = 0; j < nmat; ++j) rot[j] = phys[(j + 10) % nmat];
        for (int j = 0; j < nmat; ++j) phys[j] = rot[j];
        nmat -= 10;
      }
      // makeb (jacobi_diis.f90:258-297)
      if (nmat == 1) {
        B(1, 1) = 0.0;
        B(1, 2) = 1.0;
        B(2, 1) = 1.0;
        B(2, 2) = vals[kDiisMax];
      } else {
        B(nmat + 1, 1) = 1.0;
        B(1, nmat + 1) = 1.0;
        for (int i = 1; i <= nmat - 1; ++i) {
          B(nmat + 1, i + 1) = vals[i - 1];
          B(i + 1, nmat + 1) = vals[i - 1];
        }
        B(nmat + 1, nmat + 1) = vals[kDiisMax];
      }
      const int n1 = nmat + 1;
      double bloc[kOVals * kOVals], cex[kOVals];
      for (int c = 1; c <= n1; ++c)
        for (int r = 1; r <= n1; ++r) bloc[(r - 1) + (c - 1) * n1] = B(r, c);
      for (int r = 0; r < n1; ++r) cex[r] = 0.0;
      cex[0] = 1.0;
      if (!gjinv_host(n1, bloc, cex)) {
        nmat = 1;   // keep the Jacobi step (jacobi_diis.f90:236-239)
      } else {
        for (int i = 1; i <= nmat; ++i) {
          cf.p[i - 1] = dxd.p + (size_t)phys[i - 1] * nn;
          cf.c[i - 1] = cex[i];
        }
        ncoef = nmat;
        nmat += 1;
      }
    }
    hipLaunchKernelGGL(k_ops_extrap, dim3(nblk), dim3(kOT), 0, s, nn, ncoef, cf, dxn.p, dx.p,
                       dinc.p, part.p);
    DDX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ops_reduce, dim3(1), dim3(64), 0, s, nblk, 2, 1, part.p, red.p);
    DDX_HIP(hipGetLastError());
    DDX_HIP(hipMemcpyAsync(vals.data(), red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    DDX_HIP(hipMemcpyAsync(x, dx.p, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    if (norm == 4)
      DDX_HIP(hipMemcpyAsync(inc.data(), dinc.p, nn * sizeof(double), hipMemcpyDeviceToHost, s));
    DDX_HIP(hipStreamSynchronize(s));
    // the norm of the increment and the test (jacobi_diis.f90:128-156)
    double rms_norm, max_norm;
    if (norm <= 3) {
      rms_norm = std::sqrt(vals[0] / (double)n);
      max_norm = vals[1];
      ok = norm == 1 ? rms_norm < tol : norm == 2 ? max_norm < tol
                                                  : (rms_norm < tol && max_norm < tol_max);
    } else {
      max_norm = -1.0;
      t_ops_rc = DDX_OK;
      rms_norm = u_norm(&n, inc.data());
      DDX_TRY(t_ops_rc);
      ok = rms_norm < tol;
    }
    hist.push_back(rms_norm);
    if (lprint > 0) print_norm_line(norm, it, rms_norm, max_norm);
    if (ok) break;
  }
  *n_iter = ok ? it : maxit + 1;
  *ok_out = ok;
  return DDX_OK;
}

}  // namespace
}  // namespace ddx

using namespace ddx;

extern "C" {

void ddx_op_lx(const int *n, const double *x, double *y) {
  ddx_handle *h = t_ops_h;
  if (!h || !n || *n != h->nylm * h->n) {
    set_error("ddx_op_lx: no ddx_jacobi_diis_ops call of a handle of this size");
    t_ops_rc = DDX_ERR_STATE;
    return;
  }
  const int rc = ddx_lx(h, x, y);
  if (rc != DDX_OK) t_ops_rc = rc;
}

void ddx_op_lstarx(const int *n, const double *x, double *y) {
  ddx_handle *h = t_ops_h;
  if (!h || !n || *n != h->nylm * h->n) {
    set_error("ddx_op_lstarx: no ddx_jacobi_diis_ops call of a handle of this size");
    t_ops_rc = DDX_ERR_STATE;
    return;
  }
  const int rc = ddx_lstarx(h, x, y);
  if (rc != DDX_OK) t_ops_rc = rc;
}

void ddx_op_ldm1x(const int *n, const double *x, double *y) {
  ddx_handle *h = t_ops_h;
  if (!h || !n || *n != h->nylm * h->n) {
    set_error("ddx_op_ldm1x: no ddx_jacobi_diis_ops call of a handle of this size");
    t_ops_rc = DDX_ERR_STATE;
    return;
  }
  const int rc = ddx_ldm1x(h, x, y);
  if (rc != DDX_OK) t_ops_rc = rc;
}

double ddx_op_hnorm(const int *n, const double *x) {
  ddx_handle *h = t_ops_h;
  if (!h || !n || *n != h->nylm * h->n) {
    set_error("ddx_op_hnorm: no ddx_jacobi_diis_ops call of a handle of this size");
    t_ops_rc = DDX_ERR_STATE;
    return 0.0;
  }
  double v = 0.0;
  const int rc = ddx_hnorm(h, x, &v);
  if (rc != DDX_OK) t_ops_rc = rc;
  return v;
}

int ddx_jacobi_diis_ops(ddx_handle *h, int n, int lprint, int diis_max, int norm, double tol,
                        const double *rhs, double *x, int *n_iter, int *ok,
                        ddx_matvec_fn matvec, ddx_matvec_fn dm1vec, ddx_norm_fn u_norm) {
  if (!rhs || !x || !n_iter || !ok || !matvec || !dm1vec || n < 1 || *n_iter < 1 || norm < 1 ||
      norm > 4) {
    set_error("ddx_jacobi_diis_ops: invalid arguments");
    return DDX_ERR_ARG;
  }
  if (norm == 4 && !u_norm) {
    set_error("must provide a function norm(n,x) to evaluate the norm of the increment");
    return DDX_ERR_ARG;
  }
  if (!(diis_max == 0 || (diis_max > 10 && diis_max <= kDiisMax))) {
    set_error("diis_max must be 0 or in 11..%d", kDiisMax);
    return DDX_ERR_ARG;
  }
  *ok = 0;
  ddx_handle *const prev = t_ops_h;
  t_ops_h = h;
  int rc;
  if (h && h->ninst == 1 && n == h->nylm * h->n && *n_iter <= kMaxIter &&
      module_operators(matvec, dm1vec, norm, u_norm)) {
    // the module's own operators: the whole solve on the device, FP64
    rc = ddx_jacobi_diis(h, matvec == ddx_op_lstarx ? 1 : 0, diis_max, tol, rhs, x, n_iter, ok);
    if (rc == DDX_OK) {
      const std::vector<double> &v = matvec == ddx_op_lstarx ? h->hist_adjoint : h->hist_direct;
      h->hist_ops = v;
      if (lprint > 0)
        for (size_t k = 0; k < v.size(); ++k) print_norm_line(4, (int)k + 1, v[k], -1.0);
    }
  } else {
    std::vector<double> hist;
    rc = ops_generic(h, n, lprint, diis_max, norm, tol, rhs, x, n_iter, ok, matvec, dm1vec,
                     u_norm, hist);
    if (h) h->hist_ops = hist;
  }
  t_ops_h = prev;
  if (rc == DDX_OK && !*ok) {
    set_error("jacobi_diis did not converge in %d iterations", *n_iter - 1);
    rc = DDX_ERR_NOCONV;
  }
  return rc;
}

}  // extern "C"
