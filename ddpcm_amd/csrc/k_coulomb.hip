// All-pairs Coulomb sums for point-charge solutes.
//   potential  phi(j) = sum_i q_i / |r_j - r_i|            (mkrhs.f90:102-114)
//   field      E(j)   = sum_i q_i (r_j - r_i) / |r_j - r_i|^3  (efld.f90:98-114)
// Each thread owns one target; sources stream through LDS in tiles of 256.
// The potential kernel can return the field of the same sources as well
// (the solute field at the cavity points that the forces contract with
// zeta), which costs one all-pairs pass less than a separate field kernel.
// 1/r comes from the hardware reciprocal square root refined by one
// third-order step (relative error a few ulp), which replaces an IEEE sqrt
// and divide per pair.
// When there are few targets (the field of the cavity charges at the nuclei)
// the sources are split over several workgroups and the partial fields are
// summed afterwards in a fixed order, so results are deterministic.
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

constexpr int kCT = 256;

// 1/sqrt(x) for x > 0: v_rsq_f64 (relative error below 2^-24 on gfx950)
// and one third-order correction: with e = 1 - x y^2, y (1 + e/2 + 3e^2/8)
// leaves a relative error of order e^3 plus rounding (about 1 ulp)
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y, e * fma(0.375, e, 0.5), y);
}

// phi(j) and, with FIELD, the field ef(:, j) of the same sources, from one
// evaluation of 1/r per pair
template <bool FIELD>
__global__ __launch_bounds__(kCT) void k_potential(int nsrc, const double *__restrict__ q,
                                                   const double4 *__restrict__ src, int ntrg,
                                                   const double *__restrict__ trg,
                                                   double *__restrict__ phi,
                                                   double *__restrict__ ef) {
  __shared__ double4 tile[kCT];
  int j = blockIdx.x * kCT + threadIdx.x;
  double tx = 0.0, ty = 0.0, tz = 0.0;
  if (j < ntrg) {
    tx = trg[3 * (size_t)j];
    ty = trg[3 * (size_t)j + 1];
    tz = trg[3 * (size_t)j + 2];
  }
  double v = 0.0, e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int i0 = 0; i0 < nsrc; i0 += kCT) {
    int i = i0 + threadIdx.x;
    if (i < nsrc) {
      double4 s = src[i];
      tile[threadIdx.x] = make_double4(s.x, s.y, s.z, q[i]);
    }
    __syncthreads();
    int cnt = min(kCT, nsrc - i0);
    for (int k = 0; k < cnt; ++k) {
      double4 s = tile[k];
      double dx = tx - s.x, dy = ty - s.y, dz = tz - s.z;
      double d2 = dx * dx + dy * dy + dz * dz;
      double y = rsqrt_nr(d2);
      double qy = s.w * y;
      v = v + qy;
      if (FIELD) {
        double f = qy * (y * y);
        e0 = e0 + f * dx;
        e1 = e1 + f * dy;
        e2 = e2 + f * dz;
      }
    }
    __syncthreads();
  }
  if (j < ntrg) {
    phi[j] = v;
    if (FIELD) {
      ef[3 * (size_t)j] = e0;
      ef[3 * (size_t)j + 1] = e1;
      ef[3 * (size_t)j + 2] = e2;
    }
  }
}

// partial fields of sources [split range) at targets; out[(split * ntrg + j) * 3]
__global__ __launch_bounds__(kCT) void k_field(int nsrc, const double *__restrict__ q,
                                               const double *__restrict__ src, int ntrg,
                                               const double *__restrict__ trg, int chunk,
                                               double *__restrict__ out) {
  __shared__ double4 tile[kCT];
  int j = blockIdx.x * kCT + threadIdx.x;
  int sbeg = blockIdx.y * chunk;
  int send = min(nsrc, sbeg + chunk);
  double tx = 0.0, ty = 0.0, tz = 0.0;
  if (j < ntrg) {
    tx = trg[3 * (size_t)j];
    ty = trg[3 * (size_t)j + 1];
    tz = trg[3 * (size_t)j + 2];
  }
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int i0 = sbeg; i0 < send; i0 += kCT) {
    int i = i0 + threadIdx.x;
    if (i < send)
      tile[threadIdx.x] = make_double4(src[3 * (size_t)i], src[3 * (size_t)i + 1],
                                       src[3 * (size_t)i + 2], q[i]);
    __syncthreads();
    int cnt = min(kCT, send - i0);
    for (int k = 0; k < cnt; ++k) {
      double4 s = tile[k];
      double dx = tx - s.x, dy = ty - s.y, dz = tz - s.z;
      double r2 = dx * dx + dy * dy + dz * dz;
      double y = rsqrt_nr(r2);
      double f = s.w * (y * y * y);
      e0 = e0 + f * dx;
      e1 = e1 + f * dy;
      e2 = e2 + f * dz;
    }
    __syncthreads();
  }
  if (j < ntrg) {
    size_t o = ((size_t)blockIdx.y * ntrg + j) * 3;
    out[o] = e0;
    out[o + 1] = e1;
    out[o + 2] = e2;
  }
}

__global__ void k_field_reduce(int nsplit, int ntrg, const double *__restrict__ part,
                               double *__restrict__ ef) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * ntrg) return;
  double s = 0.0;
  for (int b = 0; b < nsplit; ++b) s += part[(size_t)b * 3 * ntrg + k];
  ef[k] = s;
}

__global__ void k_psi(int n, int nylm, double fac, const double *__restrict__ q,
                      double *__restrict__ psi) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n * nylm) return;
  psi[k] = (k % nylm == 0) ? fac * q[k / nylm] : 0.0;
}

}  // namespace

int coulomb_potential(ddx_handle *h, int nsrc, const double *q, const double4 *src, int ntrg,
                      const double *trg3, double *phi, double *ef) {
  if (ntrg <= 0) return DDX_OK;
  TimedLaunch tl(h, DDX_K_POTENTIAL, h->stream);
  if (ef)
    hipExtLaunchKernelGGL(k_potential<true>, dim3((ntrg + kCT - 1) / kCT), dim3(kCT), 0,
                          h->stream, tl.b, tl.e, 0, nsrc, q, src, ntrg, trg3, phi, ef);
  else
    hipExtLaunchKernelGGL(k_potential<false>, dim3((ntrg + kCT - 1) / kCT), dim3(kCT), 0,
                          h->stream, tl.b, tl.e, 0, nsrc, q, src, ntrg, trg3, phi, nullptr);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int coulomb_field(ddx_handle *h, int nsrc, const double *q, const double *src3, int ntrg,
                  const double *trg3, double *ef) {
  if (ntrg <= 0) return DDX_OK;
  int tb = (ntrg + kCT - 1) / kCT;
  int nsplit = 1;
  if (tb < 1024 && nsrc > 4 * kCT) {
    nsplit = (1024 + tb - 1) / tb;
    int maxsplit = (nsrc + 4 * kCT - 1) / (4 * kCT);
    if (nsplit > maxsplit) nsplit = maxsplit;
    if (nsplit < 1) nsplit = 1;
  }
  int chunk = (nsrc + nsplit - 1) / nsplit;
  chunk = ((chunk + kCT - 1) / kCT) * kCT;
  nsplit = chunk > 0 ? (nsrc + chunk - 1) / chunk : 1;
  if (nsplit < 1) nsplit = 1;
  DDX_TRY(h->fpart.reserve((size_t)nsplit * 3 * ntrg));
  {
    TimedLaunch tl(h, DDX_K_FIELD, h->stream);
    hipExtLaunchKernelGGL(k_field, dim3(tb, nsplit), dim3(kCT), 0, h->stream, tl.b, tl.e, 0, nsrc,
                          q, src3, ntrg, trg3, chunk > 0 ? chunk : kCT, h->fpart.p);
  }
  DDX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_field_reduce, dim3((3 * ntrg + 255) / 256), dim3(256), 0, h->stream, nsplit,
                     ntrg, h->fpart.p, ef);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int make_psi(ddx_handle *h, const double *charge, double *psi) {
  size_t nn = (size_t)h->n * h->nylm;
  double fac = sqrt(4.0 * (4.0 * atan(1.0)));
  hipLaunchKernelGGL(k_psi, dim3((nn + 255) / 256), dim3(256), 0, h->stream, h->n, h->nylm, fac,
                     charge, psi);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

}  // namespace ddx
