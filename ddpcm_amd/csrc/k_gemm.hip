// FP64 matrix products on the gfx950 matrix cores (v_mfma_f64_16x16x4_f64),
// shaped for the two products of the ddCOSMO operators:
//
//   projection   y(lm, i) = alpha * sum_n Yw(lm, n) P(n, i)
//                (intrhs, ddcosmo.f90:789-821, for every sphere at once;
//                 Yw = basis * w) with P dense (ngrid, nsph).
//   expansion    xi(n, i) = [w_n] * sum_lm Y(lm, n) S(lm, i)
//                (matvec.f90:196-204, forces_dd.f90:108-113).
//
// Projection: one workgroup per 16 spheres holds all nylm rows (up to 8
// MFMA row tiles per pass); its four waves split the ngrid reduction and
// are summed in LDS in a fixed order; each wave loads the operands of four
// k-steps before issuing their MFMAs, to keep loads in flight.  Expansion: one wave per 16 x 16
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

constexpr int kUnroll = 4;   // k-steps whose operands are loaded before their MFMAs
constexpr int kExpandUnroll = 8;
constexpr size_t kExpandWide = 2048;   // blocks from which a wave takes 4 sphere tiles
constexpr int kUnroll32 = 2;   // k-steps in flight in the FP32 projection

__global__ __launch_bounds__(256) void k_project(int nylm, int ngrid, int nsph, double alpha,
                                                 const double *__restrict__ Yw,
                                                 const double *__restrict__ P,
                                                 double *__restrict__ y,
                                                 const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  __shared__ v4d red[3][kMaxTiles][kWave];
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  // blockIdx.y splits the points into contiguous ranges of whole k-steps
  // (so that each line of P is read by one workgroup); wave w of the split
  // takes its k-steps w, w + 4, ... and the split's sum goes to
  // y + blockIdx.y nylm nsph
  const int nsteps = (ngrid + 3) / 4;
  const int per = (nsteps + gridDim.y - 1) / gridDim.y;
  const int kbeg = 4 * per * blockIdx.y, kend = min(ngrid, 4 * per * (int)(blockIdx.y + 1));
  y += (size_t)blockIdx.y * nylm * nsph;
  const int i0 = blockIdx.x * 16;
  const int col = lane & 15, krow = lane >> 4;
  const int sph = i0 + col;
  const int ntile = (nylm + 15) / 16;
  const double *Pc = P + (size_t)min(sph, nsph - 1) * ngrid;
  for (int t0 = 0; t0 < ntile; t0 += kMaxTiles) {
    const int nt = min(kMaxTiles, ntile - t0);
    v4d acc[kMaxTiles];
#pragma unroll
    for (int t = 0; t < kMaxTiles; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    // k-steps of 4 points each
    for (int kb = kbeg + 4 * wave; kb < kend; kb += 16 * kUnroll) {
      double b[kUnroll];
      double a[kUnroll][kMaxTiles];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int n = kb + 16 * u + krow;
        const bool okn = n < kend;
        b[u] = (okn && sph < nsph) ? Pc[n] : 0.0;
#pragma unroll
        for (int t = 0; t < kMaxTiles; ++t) {
          const int m = 16 * (t0 + t) + col;
          a[u][t] = (t < nt && okn && m < nylm) ? Yw[m + (size_t)nylm * n] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
#pragma unroll
        for (int t = 0; t < kMaxTiles; ++t)
          if (t < nt) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][t], b[u], acc[t], 0, 0, 0);
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
          v4d sum = acc[t] + red[0][t][lane];
          sum = sum + red[1][t][lane];
          sum = sum + red[2][t][lane];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int m = 16 * (t0 + t) + krow + 4 * q;
            if (m < nylm) y[m + (size_t)nylm * sph] = alpha * sum[q];
          }
        }
      }
    }
    __syncthreads();
  }
}

// xi(n, i) = scale_n * sum_lm Y(lm, n) S(lm, i); one wave per 16 points x
// J tiles of 16 spheres (J independent accumulator chains that share the Y
// operand), blockIdx.x over point tiles of 64, blockIdx.y over sphere tiles
// of 16 J.  J = 4 pays where the grid has blocks to spare (large systems,
// batches), J = 1 keeps small systems' grids full.
template <int J>
__global__ __launch_bounds__(256) void k_expand(int nylm, int ngrid, int nsph,
                                                const double *__restrict__ Y,
                                                const double *__restrict__ S,
                                                const double4 *__restrict__ grid, int wscale,
                                                double *__restrict__ xi,
                                                const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int n0 = blockIdx.x * 64 + 16 * wave;
  const int i0 = blockIdx.y * 16 * J;
  if (n0 >= ngrid) return;
  const int row = lane & 15, kq = lane >> 4;
  const int na = n0 + row;          // A row: grid point
  v4d acc[J];
#pragma unroll
  for (int j = 0; j < J; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  // the operands of kExpandUnroll k-steps are loaded before their MFMAs
  for (int k0 = 0; k0 < nylm; k0 += 4 * kExpandUnroll) {
    double a[kExpandUnroll], b[kExpandUnroll][J];
#pragma unroll
    for (int u = 0; u < kExpandUnroll; ++u) {
      const int k = k0 + 4 * u + kq;
      a[u] = (na < ngrid && k < nylm) ? Y[k + (size_t)nylm * na] : 0.0;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int sb = i0 + 16 * j + row;   // B column: sphere
        b[u][j] = (sb < nsph && k < nylm) ? S[k + (size_t)nylm * sb] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kExpandUnroll; ++u)
#pragma unroll
      for (int j = 0; j < J; ++j)
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u][j], acc[j], 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int sph = i0 + 16 * j + row;
    if (sph < nsph) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int n = n0 + kq + 4 * q;
        if (n < ngrid) {
          double v = acc[j][q];
          if (wscale) v *= grid[n].w;
          xi[n + (size_t)ngrid * sph] = v;
        }
      }
    }
  }
}

// FP32 versions for the mixed precision, on v_mfma_f32_16x16x4f32:
//   A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k,
//   D(i, j) in lane j + 16 (i / 4), element i % 4.
typedef float v4f __attribute__((ext_vector_type(4)));

// projection with float tables and point values, FP32 accumulation; the
// same workgroup shape and point split as k_project, y in double
// (two k-steps in flight per wave: with four the compiler's schedule takes
// 361 registers, or spills 832 bytes per lane under a two-workgroup bound)
__global__ __launch_bounds__(256) void k_project32(int nylm, int ngrid, int nsph, double alpha,
                                                   const float *__restrict__ Yw,
                                                   const float *__restrict__ P,
                                                   double *__restrict__ y,
                                                   const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  __shared__ v4f red[3][kMaxTiles][kWave];
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int nsteps = (ngrid + 3) / 4;
  const int per = (nsteps + gridDim.y - 1) / gridDim.y;
  const int kbeg = 4 * per * blockIdx.y, kend = min(ngrid, 4 * per * (int)(blockIdx.y + 1));
  y += (size_t)blockIdx.y * nylm * nsph;
  const int i0 = blockIdx.x * 16;
  const int col = lane & 15, krow = lane >> 4;
  const int sph = i0 + col;
  const int ntile = (nylm + 15) / 16;
  const float *Pc = P + (size_t)min(sph, nsph - 1) * ngrid;
  for (int t0 = 0; t0 < ntile; t0 += kMaxTiles) {
    const int nt = min(kMaxTiles, ntile - t0);
    v4f acc[kMaxTiles];
#pragma unroll
    for (int t = 0; t < kMaxTiles; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int kb = kbeg + 4 * wave; kb < kend; kb += 16 * kUnroll32) {
      float b[kUnroll32];
      float a[kUnroll32][kMaxTiles];
#pragma unroll
      for (int u = 0; u < kUnroll32; ++u) {
        const int n = kb + 16 * u + krow;
        const bool okn = n < kend;
        b[u] = (okn && sph < nsph) ? Pc[n] : 0.f;
#pragma unroll
        for (int t = 0; t < kMaxTiles; ++t) {
          const int m = 16 * (t0 + t) + col;
          a[u][t] = (t < nt && okn && m < nylm) ? Yw[m + (size_t)nylm * n] : 0.f;
        }
      }
#pragma unroll
      for (int u = 0; u < kUnroll32; ++u)
#pragma unroll
        for (int t = 0; t < kMaxTiles; ++t)
          if (t < nt) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][t], b[u], acc[t], 0, 0, 0);
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
          v4f sum = acc[t] + red[0][t][lane];
          sum = sum + red[1][t][lane];
          sum = sum + red[2][t][lane];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int m = 16 * (t0 + t) + 4 * krow + q;
            if (m < nylm) y[m + (size_t)nylm * sph] = alpha * (double)sum[q];
          }
        }
      }
    }
    __syncthreads();
  }
}

// xi(n, i) = scale_n * sum_lm Y(lm, n) S(lm, i) in FP32 (S rounded on load),
// with the launch shapes of k_expand (J sphere tiles per wave)
template <int J>
__global__ __launch_bounds__(256) void k_expand32(int nylm, int ngrid, int nsph,
                                                  const float *__restrict__ Y,
                                                  const double *__restrict__ S,
                                                  const double4 *__restrict__ grid, int wscale,
                                                  float *__restrict__ xi,
                                                  const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int n0 = blockIdx.x * 64 + 16 * wave;
  const int i0 = blockIdx.y * 16 * J;
  if (n0 >= ngrid) return;
  const int row = lane & 15, kq = lane >> 4;
  const int na = n0 + row;
  v4f acc[J];
#pragma unroll
  for (int j = 0; j < J; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < nylm; k0 += 4 * kExpandUnroll) {
    float a[kExpandUnroll], b[kExpandUnroll][J];
#pragma unroll
    for (int u = 0; u < kExpandUnroll; ++u) {
      const int k = k0 + 4 * u + kq;
      a[u] = (na < ngrid && k < nylm) ? Y[k + (size_t)nylm * na] : 0.f;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int sb = i0 + 16 * j + row;
        b[u][j] = (sb < nsph && k < nylm) ? (float)S[k + (size_t)nylm * sb] : 0.f;
      }
    }
#pragma unroll
    for (int u = 0; u < kExpandUnroll; ++u)
#pragma unroll
      for (int j = 0; j < J; ++j)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[u][j], acc[j], 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int sph = i0 + 16 * j + row;
    if (sph < nsph) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int n = n0 + 4 * kq + q;
        if (n < ngrid) {
          float v = acc[j][q];
          if (wscale) v *= (float)grid[n].w;
          xi[n + (size_t)ngrid * sph] = v;
        }
      }
    }
  }
}

__global__ void k_to_f32(size_t n, const double *__restrict__ a, float *__restrict__ b) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = (float)a[k];
}

}  // namespace

int make_tables32(ddx_handle *h) {
  if (h->have_tables32) return DDX_OK;
  const size_t n = (size_t)h->nylm * h->ngrid;
  DDX_TRY(h->Y32.reserve(n));
  DDX_TRY(h->Yw32.reserve(n));
  hipLaunchKernelGGL(k_to_f32, dim3((n + 255) / 256), dim3(256), 0, h->stream, n, h->Y.p,
                     h->Y32.p);
  hipLaunchKernelGGL(k_to_f32, dim3((n + 255) / 256), dim3(256), 0, h->stream, n, h->Yw.p,
                     h->Yw32.p);
  DDX_HIP(hipGetLastError());
  h->have_tables32 = true;
  return DDX_OK;
}

int project32(ddx_handle *h, hipStream_t s, double alpha, const float *P, double *y,
              const DiisState *gate, int ysplit) {
  if (h->n <= 0) return DDX_OK;
  TimedLaunch tl(h, DDX_K_PROJECT, s);
  hipExtLaunchKernelGGL(k_project32, dim3((h->n + 15) / 16, ysplit), dim3(256), 0, s, tl.b, tl.e,
                        0, h->nylm, h->ngrid, h->n, alpha, h->Yw32.p, P, y, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int expand32(ddx_handle *h, hipStream_t s, const double *S, float *xi, bool wscale,
             const DiisState *gate) {
  if (h->n <= 0) return DDX_OK;
  const int pt = (h->ngrid + 63) / 64;
  const bool wide = (size_t)pt * ((h->n + 63) / 64) >= kExpandWide;
  dim3 g(pt, wide ? (h->n + 63) / 64 : (h->n + 15) / 16);
  TimedLaunch tl(h, DDX_K_EXPAND, s);
  hipExtLaunchKernelGGL(wide ? k_expand32<4> : k_expand32<1>, g, dim3(256), 0, s, tl.b, tl.e, 0, h->nylm, h->ngrid, h->n,
                        h->Y32.p, S, h->grid.p, wscale ? 1 : 0, xi, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int project(ddx_handle *h, hipStream_t s, double alpha, const double *P, double *y,
            const DiisState *gate, int ysplit) {
  if (h->n <= 0) return DDX_OK;
  TimedLaunch tl(h, DDX_K_PROJECT, s);
  hipExtLaunchKernelGGL(k_project, dim3((h->n + 15) / 16, ysplit), dim3(256), 0, s, tl.b, tl.e, 0,
                        h->nylm, h->ngrid, h->n, alpha, h->Yw.p, P, y, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int expand(ddx_handle *h, hipStream_t s, const double *S, double *xi, bool wscale,
           const DiisState *gate) {
  if (h->n <= 0) return DDX_OK;
  const int pt = (h->ngrid + 63) / 64;
  const bool wide = (size_t)pt * ((h->n + 63) / 64) >= kExpandWide;
  dim3 g(pt, wide ? (h->n + 63) / 64 : (h->n + 15) / 16);
  TimedLaunch tl(h, DDX_K_EXPAND, s);
  hipExtLaunchKernelGGL(wide ? k_expand<4> : k_expand<1>, g, dim3(256), 0, s, tl.b, tl.e, 0, h->nylm, h->ngrid, h->n,
                        h->Y.p, S, h->grid.p, wscale ? 1 : 0, xi, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

}  // namespace ddx
