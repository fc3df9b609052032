// FP64 matrix products on the gfx950 matrix cores (v_mfma_f64_16x16x4_f64),
// shaped for the two products of the ddCOSMO operators:
//
//   projection   y(lm, i) = alpha * sum_n Yw(lm, n) P(n, i)
//                (intrhs, ddcosmo.f90:789-821, for every sphere at once;
//                 Yw = basis * w).  P is either a dense (ngrid, nsph) array
//                 or, fused, the point values of the L sweep gathered from
//                 the triples: P(n, i) = sum_{t at (i, n)} W_t r_t where
//                 ui(n, i) < 1, else 0 (calcv, ddcosmo.f90:1695-1741).
//   expansion    xi(n, i) = [w_n] * sum_lm Y(lm, n) S(lm, i)
//                (matvec.f90:196-204, forces_dd.f90:108-113).
//
// Projection: one workgroup per 16 spheres holds all nylm rows (up to 8
// MFMA row tiles per pass); its four waves split the ngrid reduction and
// are summed in LDS in a fixed order.  Expansion: one wave per 16 x 16
// output tile.  Operands go straight from L2 to registers: Y (<= 1.2 MB at
// lmax 10 / 1202 points) stays cache-resident across workgroups.
//
// Operand layout of v_mfma_f64_16x16x4_f64 (one double per lane):
//   A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k,
//   D(i, j) in lane j + 16 (i % 4), element i / 4 of the 4-double result
//   (measured on gfx950).
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

constexpr int kMaxTiles = 8;   // row tiles of 16 per projection pass

template <bool FUSED>
__global__ __launch_bounds__(256) void k_project(
    int nylm, int ngrid, int nsph, double alpha, const double *__restrict__ Yw,
    const double *__restrict__ P, const int *__restrict__ ptoff,
    const int *__restrict__ pttri, const double2 *__restrict__ ttw,
    const double *__restrict__ r, const double *__restrict__ ui, double *__restrict__ y,
    const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  __shared__ v4d red[3][kMaxTiles][kWave];
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int i0 = blockIdx.x * 16;
  const int col = lane & 15, krow = lane >> 4;
  const int sph = i0 + col;
  const int ntile = (nylm + 15) / 16;
  for (int t0 = 0; t0 < ntile; t0 += kMaxTiles) {
    const int nt = min(kMaxTiles, ntile - t0);
    v4d acc[kMaxTiles];
#pragma unroll
    for (int t = 0; t < kMaxTiles; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 4 * wave; k0 < ngrid; k0 += 16) {
      const int n = k0 + krow;
      double b = 0.0;
      if (n < ngrid && sph < nsph) {
        const size_t p = (size_t)sph * ngrid + n;
        if (FUSED) {
          if (ui[p] < 1.0) {
            for (int k = ptoff[p]; k < ptoff[p + 1]; ++k) {
              const int tr = pttri[k];
              b += ttw[tr].y * r[tr];
            }
          }
        } else {
          b = P[p];
        }
      }
#pragma unroll
      for (int t = 0; t < kMaxTiles; ++t) {
        if (t < nt) {
          const int m = 16 * (t0 + t) + (lane & 15);
          const double a = (m < nylm && n < ngrid) ? Yw[m + (size_t)nylm * n] : 0.0;
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
        }
      }
    }
    if (wave > 0) {
#pragma unroll
      for (int t = 0; t < kMaxTiles; ++t)
        if (t < nt) red[wave - 1][t][lane] = acc[t];
    }
    __syncthreads();
    if (wave == 0 && sph < nsph) {
#pragma unroll
      for (int t = 0; t < kMaxTiles; ++t) {
        if (t < nt) {
          v4d s = acc[t] + red[0][t][lane];
          s = s + red[1][t][lane];
          s = s + red[2][t][lane];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int m = 16 * (t0 + t) + krow + 4 * q;
            if (m < nylm) y[m + (size_t)nylm * sph] = alpha * s[q];
          }
        }
      }
    }
    __syncthreads();
  }
}

// xi(n, i) = scale_n * sum_lm Y(lm, n) S(lm, i); one wave per 16 x 16 tile,
// blockIdx.x over point tiles of 64, blockIdx.y over sphere tiles of 16.
__global__ __launch_bounds__(256) void k_expand(int nylm, int ngrid, int nsph,
                                                const double *__restrict__ Y,
                                                const double *__restrict__ S,
                                                const double4 *__restrict__ grid, int wscale,
                                                double *__restrict__ xi,
                                                const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int n0 = blockIdx.x * 64 + 16 * wave;
  const int i0 = blockIdx.y * 16;
  if (n0 >= ngrid) return;
  const int row = lane & 15, kq = lane >> 4;
  const int na = n0 + row;          // A row: grid point
  const int sb = i0 + row;          // B column: sphere
  v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < nylm; k0 += 4) {
    const int k = k0 + kq;
    const double a = (na < ngrid && k < nylm) ? Y[k + (size_t)nylm * na] : 0.0;
    const double b = (sb < nsph && k < nylm) ? S[k + (size_t)nylm * sb] : 0.0;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  const int sph = i0 + row;
  if (sph >= nsph) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int n = n0 + kq + 4 * q;
    if (n < ngrid) {
      double v = acc[q];
      if (wscale) v *= grid[n].w;
      xi[n + (size_t)ngrid * sph] = v;
    }
  }
}

}  // namespace

int project(ddx_handle *h, double alpha, const double *P, const double *r, double *y,
            const DiisState *gate) {
  if (h->n <= 0) return DDX_OK;
  dim3 g((h->n + 15) / 16);
  if (P)
    hipLaunchKernelGGL(k_project<false>, g, dim3(256), 0, h->stream, h->nylm, h->ngrid, h->n,
                       alpha, h->Yw.p, P, nullptr, nullptr, nullptr, nullptr, nullptr, y, gate);
  else
    hipLaunchKernelGGL(k_project<true>, g, dim3(256), 0, h->stream, h->nylm, h->ngrid, h->n,
                       alpha, h->Yw.p, nullptr, h->ptoff.p, h->pttri.p, h->ttw.p, r, h->ui.p, y,
                       gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int expand(ddx_handle *h, const double *S, double *xi, bool wscale, const DiisState *gate) {
  if (h->n <= 0) return DDX_OK;
  dim3 g((h->ngrid + 63) / 64, (h->n + 15) / 16);
  hipLaunchKernelGGL(k_expand, g, dim3(256), 0, h->stream, h->nylm, h->ngrid, h->n, h->Y.p, S,
                     h->grid.p, wscale ? 1 : 0, xi, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

}  // namespace ddx
