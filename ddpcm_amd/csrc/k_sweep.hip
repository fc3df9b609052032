// The off-diagonal ddCOSMO operators on the neighbour triples.
//
// For a triple t = (owner sphere o, grid point n, container sphere c) with
// t_n^oc < 1 define the vector
//     R_lm(t) = 4pi/(2l+1) * (t_n^oc)^l * Y_lm(s_n^oc)
// (the factors of intmlp, ddcosmo.f90:1140-1178, and of adjrhs,
// ddcosmo.f90:1359-1377).  Then
//     L X :  P(n,o)  = sum_{t at (o,n)} W_t R(t) . X_c        (calcv, ddcosmo.f90:1671-1746)
//            (LX)_o  = - sum_n w_n Y(:,n) P(n,o)              (intrhs + lx, matvec.f90:127-136)
//     L* S:  xi(n,o) = Y(:,n) . S_o                           (matvec.f90:196-204)
//            (L*S)_c = - sum_{t in c} w_n xi(n,o) W_t R(t)    (adjrhs, matvec.f90:213-223)
//
// Both sweeps walk the triples grouped by container c, so X_c / the output
// vector of c is uniform across the workgroup: X_c is read through the
// scalar cache, and the L* output is accumulated in registers per lane and
// reduced across lanes once per container through an LDS transpose.
// Y_lm follows ylmbas (ddcosmo.f90:830-895): Legendre recurrence of polleg
// (m outer, l inner, ddcosmo.f90:1056-1094), Chebyshev cos/sin(m phi) of
// trgev (ddcosmo.f90:1103-1125), normalisation facs; the per-(l,m) constant
// cf = facs / facl folds facs with 4pi/(2l+1).
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

__host__ __device__ constexpr int lmi(int l, int m) { return l * l + l + m; }

// sum_lm R_lm(t) X_lm for one triple (X uniform across the wave).
template <int LMAX>
__device__ __forceinline__ double contract_R(double cthe, double sthe, double cphi, double sphi,
                                             double t, const double *__restrict__ X,
                                             const double *__restrict__ cf) {
  double acc = 0.0;
  double pmm = 1.0, tm = 1.0;
  double cm = 1.0, sm = 0.0, cprev = 0.0, sprev = 0.0;
#pragma unroll
  for (int m = 0; m <= LMAX; ++m) {
    double a, b = 0.0;
    double p0 = pmm, p1 = 0.0, tl = tm;
    {
      double q = tl * p0;
      a = q * (cf[lmi(m, m)] * X[lmi(m, m)]);
      if (m > 0) b = q * (cf[lmi(m, m)] * X[lmi(m, -m)]);
    }
    if (m < LMAX) {
      p1 = cthe * (2.0 * m + 1.0) * p0;
      tl *= t;
      double q = tl * p1;
      a += q * (cf[lmi(m + 1, m)] * X[lmi(m + 1, m)]);
      if (m > 0) b += q * (cf[lmi(m + 1, m)] * X[lmi(m + 1, -m)]);
#pragma unroll
      for (int l = m + 2; l <= LMAX; ++l) {
        double p2 = (cthe * (2.0 * l - 1.0) * p1 - (double)(l + m - 1) * p0) * (1.0 / (double)(l - m));
        tl *= t;
        q = tl * p2;
        a += q * (cf[lmi(l, m)] * X[lmi(l, m)]);
        if (m > 0) b += q * (cf[lmi(l, m)] * X[lmi(l, -m)]);
        p0 = p1;
        p1 = p2;
      }
    }
    acc += (m == 0) ? a : (cm * a + sm * b);
    pmm = -pmm * (2.0 * m + 1.0) * sthe;
    tm *= t;
    double cn, sn;
    if (m == 0) { cn = cphi; sn = sphi; }
    else { cn = 2.0 * cphi * cm - cprev; sn = 2.0 * cphi * sm - sprev; }
    cprev = cm; sprev = sm; cm = cn; sm = sn;
  }
  return acc;
}

// acc_lm += coef * R_lm(t) for one triple (acc in registers).
template <int LMAX>
__device__ __forceinline__ void accum_R(double cthe, double sthe, double cphi, double sphi,
                                        double t, double coef, const double *__restrict__ cf,
                                        double *acc) {
  double pmm = 1.0, tm = 1.0;
  double cm = 1.0, sm = 0.0, cprev = 0.0, sprev = 0.0;
#pragma unroll
  for (int m = 0; m <= LMAX; ++m) {
    double qc = coef * cm, qs = coef * sm;
    double p0 = pmm, p1 = 0.0, tl = tm;
    {
      double z = tl * p0 * cf[lmi(m, m)];
      if (m == 0) acc[lmi(m, 0)] += z * coef;
      else { acc[lmi(m, m)] += z * qc; acc[lmi(m, -m)] += z * qs; }
    }
    if (m < LMAX) {
      p1 = cthe * (2.0 * m + 1.0) * p0;
      tl *= t;
      double z = tl * p1 * cf[lmi(m + 1, m)];
      if (m == 0) acc[lmi(m + 1, 0)] += z * coef;
      else { acc[lmi(m + 1, m)] += z * qc; acc[lmi(m + 1, -m)] += z * qs; }
#pragma unroll
      for (int l = m + 2; l <= LMAX; ++l) {
        double p2 = (cthe * (2.0 * l - 1.0) * p1 - (double)(l + m - 1) * p0) * (1.0 / (double)(l - m));
        tl *= t;
        z = tl * p2 * cf[lmi(l, m)];
        if (m == 0) acc[lmi(l, 0)] += z * coef;
        else { acc[lmi(l, m)] += z * qc; acc[lmi(l, -m)] += z * qs; }
        p0 = p1;
        p1 = p2;
      }
    }
    pmm = -pmm * (2.0 * m + 1.0) * sthe;
    tm *= t;
    double cn, sn;
    if (m == 0) { cn = cphi; sn = sphi; }
    else { cn = 2.0 * cphi * cm - cprev; sn = 2.0 * cphi * sm - sprev; }
    cprev = cm; sprev = sm; cm = cn; sm = sn;
  }
}

constexpr int kLThreads = 256;

// W_t R(t) . X_c for every triple t of container c = blockIdx.x, stored in
// the triple's slot of its point's list, so that each point's terms are
// contiguous for the gather.
template <int LMAX>
__global__ __launch_bounds__(kLThreads) void k_sweep_contract(
    const int *__restrict__ inl, const int *__restrict__ pairoff,
    const double4 *__restrict__ tgeo, const double2 *__restrict__ ttw,
    const int *__restrict__ tslot, const double *__restrict__ X,
    const double *__restrict__ cf, double *__restrict__ rw,
    const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  constexpr int NY = (LMAX + 1) * (LMAX + 1);
  const int c = blockIdx.x;
  const int beg = pairoff[inl[c]], end = pairoff[inl[c + 1]];
  const double *Xc = X + (size_t)c * NY;
  for (int k = beg + threadIdx.x; k < end; k += kLThreads) {
    double4 g = tgeo[k];
    double2 tw = ttw[k];
    rw[tslot[k]] = tw.y * contract_R<LMAX>(g.x, g.y, g.z, g.w, tw.x, Xc, cf);
  }
}

// P(n,o) = sum of the point's weighted terms, in neighbour order; zero where
// ui = 1 (calcv skips those points, ddcosmo.f90:1698).
__global__ __launch_bounds__(256) void k_gather_points(size_t np, const int *__restrict__ ptoff,
                                                       const double *__restrict__ rw,
                                                       const double *__restrict__ ui,
                                                       double *__restrict__ P,
                                                       const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double s = 0.0;
  const int b = ptoff[p], e = ptoff[p + 1];
  if (b < e && ui[p] < 1.0)
    for (int k = b; k < e; ++k) s += rw[k];
  P[p] = s;
}

constexpr int kRedChunk = 32;

// (L*S)_c for container c = blockIdx.x: one wave, lanes stride over the
// container's triples, registers accumulate, one LDS transpose reduces.
template <int LMAX>
__global__ __launch_bounds__(kWave) void k_sweep_adjoint(
    const int *__restrict__ inl, const int *__restrict__ pairoff,
    const double4 *__restrict__ tgeo, const double2 *__restrict__ ttw,
    const int *__restrict__ tp, const double *__restrict__ xiw,
    const double *__restrict__ cf, double *__restrict__ y,
    const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  constexpr int NY = (LMAX + 1) * (LMAX + 1);
  __shared__ double red[kWave][kRedChunk + 1];
  const int c = blockIdx.x, lane = threadIdx.x;
  const int beg = pairoff[inl[c]], end = pairoff[inl[c + 1]];
  double acc[NY];
#pragma unroll
  for (int k = 0; k < NY; ++k) acc[k] = 0.0;
  for (int k = beg + lane; k < end; k += kWave) {
    double4 g = tgeo[k];
    double2 tw = ttw[k];
    double coef = xiw[tp[k]] * tw.y;
    accum_R<LMAX>(g.x, g.y, g.z, g.w, tw.x, coef, cf, acc);
  }
  double *yc = y + (size_t)c * NY;
#pragma unroll
  for (int c0 = 0; c0 < NY; c0 += kRedChunk) {
#pragma unroll
    for (int k = 0; k < kRedChunk; ++k)
      if (c0 + k < NY) red[lane][k] = acc[c0 + k];
    __syncthreads();
    if (lane < kRedChunk && c0 + lane < NY) {
      double s = 0.0;
      for (int q = 0; q < kWave; ++q) s += red[q][lane];
      yc[c0 + lane] = -s;
    }
    __syncthreads();
  }
}

template <int LMAX>
int launch_contract(ddx_handle *h, hipStream_t s, const double *X, double *r,
                    const DiisState *gate) {
  TimedLaunch tl(h, DDX_K_SWEEP_L, s);
  hipLaunchKernelGGL(k_sweep_contract<LMAX>, dim3(h->n), dim3(kLThreads), 0, s,
                     h->inl.p, h->pairoff.p, h->tgeo.p, h->ttw.p, h->tslot.p, X, h->cf.p, r, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

template <int LMAX>
int launch_adjoint(ddx_handle *h, hipStream_t s, const double *xi, double *y,
                   const DiisState *gate) {
  TimedLaunch tl(h, DDX_K_SWEEP_LSTAR, s);
  hipLaunchKernelGGL(k_sweep_adjoint<LMAX>, dim3(h->n), dim3(kWave), 0, s, h->inl.p,
                     h->pairoff.p, h->tgeo.p, h->ttw.p, h->tp.p, xi, h->cf.p, y, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

#define DDX_LMAX_CASES(F, ...)                                                        \
  switch (h->lmax) {                                                                  \
    case 0: return F<0>(__VA_ARGS__);  case 1: return F<1>(__VA_ARGS__);              \
    case 2: return F<2>(__VA_ARGS__);  case 3: return F<3>(__VA_ARGS__);              \
    case 4: return F<4>(__VA_ARGS__);  case 5: return F<5>(__VA_ARGS__);              \
    case 6: return F<6>(__VA_ARGS__);  case 7: return F<7>(__VA_ARGS__);              \
    case 8: return F<8>(__VA_ARGS__);  case 9: return F<9>(__VA_ARGS__);              \
    case 10: return F<10>(__VA_ARGS__); case 11: return F<11>(__VA_ARGS__);           \
    case 12: return F<12>(__VA_ARGS__); case 13: return F<13>(__VA_ARGS__);           \
    case 14: return F<14>(__VA_ARGS__); case 15: return F<15>(__VA_ARGS__);           \
    case 16: return F<16>(__VA_ARGS__);                                               \
    default: set_error("lmax %d not supported (0..16)", h->lmax); return DDX_ERR_ARG; \
  }

}  // namespace

int sweep_contract(ddx_handle *h, hipStream_t s, const double *X, double *r,
                   const DiisState *gate) {
  if (h->n == 0) return DDX_OK;
  DDX_LMAX_CASES(launch_contract, h, s, X, r, gate);
}

int apply_lx(ddx_handle *h, Lane &ln, const double *x, double *y, const DiisState *gate) {
  hipStream_t s = ln.stream;
  DDX_TRY(ln.rw.reserve(h->ntri + 1));
  DDX_TRY(ln.P.reserve(h->np));
  DDX_TRY(sweep_contract(h, s, x, ln.rw.p, gate));
  hipLaunchKernelGGL(k_gather_points, dim3((h->np + 255) / 256), dim3(256), 0, s, h->np,
                     h->ptoff.p, ln.rw.p, h->ui.p, ln.P.p, gate);
  DDX_HIP(hipGetLastError());
  // y(lm, o) = - sum_n Yw(lm, n) P(n, o)
  return project(h, s, -1.0, ln.P.p, y, gate);
}

int apply_lstarx(ddx_handle *h, Lane &ln, const double *x, double *y, const DiisState *gate) {
  hipStream_t s = ln.stream;
  DDX_TRY(ln.xiw.reserve(h->np));
  // w_n xi(n, o) = w_n sum_lm Y(lm, n) x(lm, o)
  DDX_TRY(expand(h, s, x, ln.xiw.p, true, gate));
  if (h->n == 0) return DDX_OK;
  DDX_LMAX_CASES(launch_adjoint, h, s, ln.xiw.p, y, gate);
}

}  // namespace ddx
