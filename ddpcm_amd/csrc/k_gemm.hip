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
// k-steps before issuing their MFMAs, to keep loads in flight.  Expansion:
// one workgroup per 16 J spheres with their S rows staged in LDS
// (k_expand_t), its point tiles split over blockIdx.y when the sphere
// tiles alone give too few workgroups.  The Y tables (<= 1.2 MB at lmax 10 / 1202 points) stay
// cache-resident across workgroups.
//
// Operand layout of v_mfma_f64_16x16x4_f64 (one double per lane):
//   A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k,
//   D(i, j) in lane j + 16 (i % 4), element i / 4 of the 4-double result
//   (measured on gfx950).
#include <algorithm>

#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

constexpr int kMaxTiles = 8;   // row tiles of 16 per projection pass

constexpr int kUnroll = 4;   // k-steps whose operands are loaded before their MFMAs
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

// The A tile of a point group for the two projections below: columns
// 32 g .. 32 g + 31 of YwP (ldm = 16 NT lm each), one contiguous block, staged
// in LDS with 16-byte loads into rows of ldm + 16 (the pad puts the two
// half-waves' MFMA operand reads on distinct banks).  Loads past the block
// are clamped to its last vector, which is then written twice alike.
template <typename T, int NT>
struct ATile {
  typedef T v16 __attribute__((ext_vector_type(16 / sizeof(T))));
  static constexpr int kLdm = 16 * NT, kRow = kLdm + 16, kV = 16 / sizeof(T);
  static constexpr int kCount = 32 * kLdm / kV;          // vectors per tile
  static constexpr int kPer = (kCount + 255) / 256;      // per thread
  v16 r[kPer];
  __device__ __forceinline__ void load(const T *__restrict__ YwP, int g) {
    const T *src = YwP + (size_t)kLdm * 32 * g;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int f = min((int)threadIdx.x + 256 * j, kCount - 1) * kV;
      r[j] = *reinterpret_cast<const v16 *>(src + f);
    }
  }
  __device__ __forceinline__ void store(T *dst) const {
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int f = min((int)threadIdx.x + 256 * j, kCount - 1) * kV;
      *reinterpret_cast<v16 *>(dst + (f / kLdm) * kRow + f % kLdm) = r[j];
    }
  }
};

// The projection of large systems, both precisions: one workgroup per 64
// spheres, wave w owning spheres 16 w .. 16 w + 15 over all points and all
// NT row tiles.  Points go in groups of 32 (8 k-steps): the group's P tile
// (64 spheres x 32 points) is staged in LDS, each quarter-wave loading one
// sphere's 32 points as one contiguous run, and read back as the MFMA's B
// operand; A comes from YwP (lm fastest, zero-padded), 16 consecutive lm per
// quarter-wave.  The next group's tile is loaded while this one is used.
// In FP64 a wave keeps one accumulator per class of k-steps modulo 4 and
// adds them in class order at the end: the order of k_project's four waves,
// so that both kernels give the same bits.
template <typename T, int NT>
__global__ __launch_bounds__(256) void k_project_s(int nylm, int ngrid, int ldm, int nsph,
                                                   double alpha, const T *__restrict__ YwP,
                                                   const T *__restrict__ P,
                                                   double *__restrict__ y,
                                                   const DiisState *__restrict__ gate,
                                                   const int *__restrict__ inst, int ninst) {
  typedef T vt __attribute__((ext_vector_type(4)));
  typedef T v2t __attribute__((ext_vector_type(2)));
  constexpr int R = sizeof(T) == 8 ? 4 : 1;
  constexpr int kLd = 65;   // LDS row: 64 spheres and one pad
  // spheres whose molecules have all converged are not read again
  if (!spheres_active(gate, inst, ninst, blockIdx.x * 64, min(blockIdx.x * 64 + 64, nsph)))
    return;
  __shared__ T sP[2][32 * kLd];
  typedef ATile<T, NT> AT;
  __shared__ T sA[2][32 * AT::kRow];
  AT at;
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int i = lane & 15, kq = lane >> 4;
  const int s0 = blockIdx.x * 64;
  const int ngrp = (ngrid + 31) / 32;
  // staging: quarter-wave qd (0..15) loads spheres qd, qd + 16, qd + 32,
  // qd + 48; lane l of the quarter points 2 l and 2 l + 1 of the group
  const int qd = threadIdx.x >> 4, l = threadIdx.x & 15;
  const T *Pq[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) Pq[r] = P + (size_t)min(s0 + qd + 16 * r, nsph - 1) * ngrid;
  v2t pr[4];
  auto load_tile = [&](int g) {
    const int n = min(32 * g + 2 * l, ngrid - 2);
#pragma unroll
    for (int r = 0; r < 4; ++r) pr[r] = *reinterpret_cast<const v2t *>(Pq[r] + n);
  };
  auto store_tile = [&](T *dst) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      dst[(2 * l) * kLd + qd + 16 * r] = pr[r].x;
      dst[(2 * l + 1) * kLd + qd + 16 * r] = pr[r].y;
    }
  };
  vt acc[R][NT];
#pragma unroll
  for (int c = 0; c < R; ++c)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[c][t] = vt{T(0), T(0), T(0), T(0)};
  load_tile(0);
  at.load(YwP, 0);
  store_tile(sP[0]);
  at.store(sA[0]);
  __syncthreads();
  for (int g = 0; g < ngrp; ++g) {
    const int cur = g & 1;
    // the next group's tiles (the last group loads its own again: the loop
    // body stays one block, and that copy goes to the buffers no one reads)
    load_tile(min(g + 1, ngrp - 1));
    at.load(YwP, min(g + 1, ngrp - 1));
    // A of the group's k-steps, KH at a time: YwP(16 t + i, 32 g + 4 k + kq)
    constexpr int KH = sizeof(T) == 8 ? 4 : 8;
    const T *Ag = sA[cur] + i + AT::kRow * kq;
    const T *B = sP[cur] + kq * kLd + 16 * wave + i;
#pragma unroll
    for (int k0 = 0; k0 < 8; k0 += KH) {
      T a[KH][NT];
#pragma unroll
      for (int k = 0; k < KH; ++k)
#pragma unroll
        for (int t = 0; t < NT; ++t) a[k][t] = Ag[16 * t + AT::kRow * 4 * (k0 + k)];
#pragma unroll
      for (int k = 0; k < KH; ++k) {
        const T b = B[4 * (k0 + k) * kLd];
        const int c = (k0 + k) % R;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          if constexpr (sizeof(T) == 8)
            acc[c][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[k][t], b, acc[c][t], 0, 0, 0);
          else
            acc[c][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[k][t], b, acc[c][t], 0, 0, 0);
        }
      }
    }
    store_tile(sP[cur ^ 1]);
    at.store(sA[cur ^ 1]);
    __syncthreads();
  }
  const int sph = s0 + 16 * wave + i;
  if (sph >= nsph) return;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    vt sum = acc[0][t];
#pragma unroll
    for (int c = 1; c < R; ++c) sum = sum + acc[c][t];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // D(i, j) in lane j + 16 (i % 4), element i / 4 (f64), or lane
      // j + 16 (i / 4), element i % 4 (f32)
      const int m = 16 * t + (sizeof(T) == 8 ? kq + 4 * q : 4 * kq + q);
      if (m < nylm) y[m + (size_t)nylm * sph] = alpha * (double)sum[q];
    }
  }
}

// The FP32 projection of large systems fused with the point gather of L X:
// the P tile of a group (64 spheres x 32 points) is formed in LDS from the
// sweep's terms, P(n, o) = sum of rw[ptoff[p] .. ptoff[p + 1]) in order
// (k_gather_points), so that P never goes through memory.  Thread t sums
// point t % 32 of the spheres t / 32 + 8 r (r = 0..7); the first kTerms
// terms of each point are loaded unconditionally (indices clamped, values
// masked), further ones in a loop.  The term ranges of the group after next
// and the terms of the next group are loaded while this group's MFMAs run.
constexpr int kTerms = 4;
template <int NT>
__global__ __launch_bounds__(256) void k_project_g(int nylm, int ngrid, int ldm, int nsph,
                                                   int ntri, double alpha,
                                                   const float *__restrict__ YwP,
                                                   const int *__restrict__ ptoff,
                                                   const float *__restrict__ rw,
                                                   double *__restrict__ y,
                                                   const DiisState *__restrict__ gate,
                                                   const int *__restrict__ inst, int ninst) {
  constexpr int kLd = 65;
  if (!spheres_active(gate, inst, ninst, blockIdx.x * 64, min(blockIdx.x * 64 + 64, nsph)))
    return;
  __shared__ float sP[2][32 * kLd];
  typedef ATile<float, NT> AT;
  __shared__ float sA[2][32 * AT::kRow];
  AT at;
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int i = lane & 15, kq = lane >> 4;
  const int s0 = blockIdx.x * 64;
  const int ngrp = (ngrid + 31) / 32;
  const int pl = threadIdx.x & 31, q8 = threadIdx.x >> 5;
  const int last = max(ntri - 1, 0);
  size_t prow[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) prow[r] = (size_t)min(s0 + q8 + 8 * r, nsph - 1) * ngrid;
  int b[8], e[8];
  float tv[8][kTerms];
  auto load_ranges = [&](int g) {
    const int n = min(32 * g + pl, ngrid - 1);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      b[r] = ptoff[prow[r] + n];
      e[r] = ptoff[prow[r] + n + 1];
    }
  };
  auto load_terms = [&]() {
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int u = 0; u < kTerms; ++u) tv[r][u] = rw[min(b[r] + u, last)];
  };
  auto store_tile = [&](int g, float *dst) {
    const bool okn = 32 * g + pl < ngrid;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      float sum = 0.f;
#pragma unroll
      for (int u = 0; u < kTerms; ++u)
        if (b[r] + u < e[r]) sum += tv[r][u];
      for (int k = b[r] + kTerms; k < e[r]; ++k) sum += rw[k];
      dst[pl * kLd + q8 + 8 * r] = okn ? sum : 0.f;
    }
  };
  v4f acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  load_ranges(0);
  load_terms();
  at.load(YwP, 0);
  store_tile(0, sP[0]);
  at.store(sA[0]);
  load_ranges(min(1, ngrp - 1));
  __syncthreads();
  for (int g = 0; g < ngrp; ++g) {
    const int cur = g & 1;
    // terms of group g + 1 (ranges loaded in the previous step), then the
    // ranges of group g + 2 once the terms' indices are taken
    load_terms();
    int bn[8], en[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      bn[r] = b[r];
      en[r] = e[r];
    }
    load_ranges(min(g + 2, ngrp - 1));
    at.load(YwP, min(g + 1, ngrp - 1));
    const float *Ag = sA[cur] + i + AT::kRow * kq;
    const float *B = sP[cur] + kq * kLd + 16 * wave + i;
    float a[8][NT];
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int t = 0; t < NT; ++t) a[k][t] = Ag[16 * t + AT::kRow * 4 * k];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float bb = B[4 * k * kLd];
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[k][t], bb, acc[t], 0, 0, 0);
    }
    // the next group's tile from its terms (the last step forms its own
    // again, into the buffer no one reads)
    {
      int bs[8], es[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        bs[r] = b[r];
        es[r] = e[r];
        b[r] = bn[r];
        e[r] = en[r];
      }
      store_tile(min(g + 1, ngrp - 1), sP[cur ^ 1]);
      at.store(sA[cur ^ 1]);
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        b[r] = bs[r];
        e[r] = es[r];
      }
    }
    __syncthreads();
  }
  const int sph = s0 + 16 * wave + i;
  if (sph >= nsph) return;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int m = 16 * t + 4 * kq + q;
      if (m < nylm) y[m + (size_t)nylm * sph] = alpha * (double)acc[t][q];
    }
}

// The expansion, both precisions: xi(n, i) = [w_n] sum_lm Y(lm, n) S(lm, i)
// with T = double (v_mfma_f64_16x16x4f64) or float (v_mfma_f32_16x16x4f32,
// S rounded as it is staged).  One workgroup per 16 J spheres: their rows
// of S, contiguous in memory, are staged in LDS once; the four waves take
// the point tiles of 16 (w, w + 4, ...), each an MFMA chain over lm with A
// = the Y rows of the tile's points from YT (lm-major, zero-padded: each
// quarter-wave reads 16 consecutive points, and the A operands of 8
// k-steps are loaded before their MFMAs) and B from LDS.  Each finished
// tile goes back through LDS so that 16 lanes store one sphere's 16 points
// as one contiguous run.
template <typename T, int J>
__global__ __launch_bounds__(256) void k_expand_t(int nylm, int ngrid, int ldnt, int nsph,
                                                  const T *__restrict__ YT,
                                                  const double *__restrict__ S,
                                                  const double4 *__restrict__ grid, int wscale,
                                                  T *__restrict__ xi,
                                                  const DiisState *__restrict__ gate,
                                                  const int *__restrict__ inst, int ninst) {
  typedef T vt __attribute__((ext_vector_type(4)));
  extern __shared__ double smem_d[];
  T *sS = reinterpret_cast<T *>(smem_d);                 // 16 J spheres x nylm
  T *tr = sS + ((16 * J * nylm + 1) & ~1);               // 4 waves x J x 16 x 17
  const int i0 = blockIdx.x * 16 * J;
  const int nsp = min(16 * J, nsph - i0);
  // spheres whose molecules have all converged are not read again
  if (!spheres_active(gate, inst, ninst, i0, i0 + nsp)) return;
  for (int k = threadIdx.x; k < 16 * J * nylm; k += 256)
    sS[k] = k < nsp * nylm ? (T)S[(size_t)nylm * i0 + k] : T(0);
  __syncthreads();
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int row = lane & 15, kq = lane >> 4;
  T *trw = tr + wave * J * 16 * 17;
  const int ntile = (ngrid + 15) / 16;
  for (int pt = wave + 4 * blockIdx.y; pt < ntile; pt += 4 * gridDim.y) {
    const int n0 = pt * 16;
    const int na = n0 + row;
    vt acc[J];
#pragma unroll
    for (int j = 0; j < J; ++j) acc[j] = vt{T(0), T(0), T(0), T(0)};
    // YT(k, na): zero past nylm rows (nylm4 of them) and ngrid points
    const T *Ak = YT + (size_t)kq * ldnt + na;
    const int nk = (nylm + 3) / 4;
    for (int c0 = 0; c0 < nk; c0 += 8) {
      T a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = Ak[(size_t)4 * ldnt * min(c0 + u, nk - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (c0 + u < nk) {
          const int k = 4 * (c0 + u) + kq;
#pragma unroll
          for (int j = 0; j < J; ++j) {
            const T b = k < nylm ? sS[(16 * j + row) * nylm + k] : T(0);
            if constexpr (sizeof(T) == 8)
              acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b, acc[j], 0, 0, 0);
            else
              acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b, acc[j], 0, 0, 0);
          }
        }
      }
    }
    // D(i = point, j = sphere) sits in lane j + 16 (i % 4), element i / 4
    // (f64) or lane j + 16 (i / 4), element i % 4 (f32)
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = sizeof(T) == 8 ? kq + 4 * q : 4 * kq + q;
        trw[(j * 16 + row) * 17 + i] = acc[j][q];
      }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const int pl = lane & 15, sl = lane >> 4;
    const int n = n0 + pl;
    const T ws = (wscale && n < ngrid) ? (T)grid[n].w : T(1);
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int sp = 16 * j + sl + 4 * e;
        if (n < ngrid && sp < nsp) xi[n + (size_t)ngrid * (i0 + sp)] = ws * trw[sp * 17 + pl];
      }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
}

// sphere tiles per expansion workgroup: LDS for S and the transposes
template <typename T>
int expand_tiles(int nylm) {
  return nylm * 16 * 2 * (int)sizeof(T) <= 48 * 1024 ? 2 : 1;
}
template <typename T>
size_t expand_lds(int nylm, int J) {
  return (size_t)(((16 * J * nylm + 1) & ~1) + 4 * J * 16 * 17) * sizeof(T);
}
// point-tile split: about 2048 workgroups (8 per CU) where the sphere
// tiles give fewer, at least one point tile per wave
inline int expand_split(int nx, int ngrid) {
  const int ntile = (ngrid + 15) / 16;
  const int want = (2048 + nx - 1) / nx;
  return std::max(1, std::min(want, (ntile + 3) / 4));
}

// YwP(m, n) = Yw(m, n) in columns of ldm (lm fastest) and YT(k, n) = Y(k, n)
// in rows of ldnt (lm-major), zero in the padding; and their float copies
__global__ void k_pad_tables(int nylm, int ngrid, int ldm, int ldnp, int nylm4, int ldnt,
                             const double *__restrict__ Y, const double *__restrict__ Yw,
                             double *__restrict__ YwP, double *__restrict__ YT,
                             float *__restrict__ YwP32, float *__restrict__ YT32) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n1 = (size_t)ldm * ldnp, n2 = (size_t)nylm4 * ldnt;
  if (k < n1) {
    const int m = (int)(k % ldm), n = (int)(k / ldm);
    const double v = (m < nylm && n < ngrid) ? Yw[m + (size_t)nylm * n] : 0.0;
    YwP[k] = v;
    YwP32[k] = (float)v;
  } else if (k < n1 + n2) {
    const size_t q = k - n1;
    const int m = (int)(q / ldnt), n = (int)(q % ldnt);
    const double v = (m < nylm && n < ngrid) ? Y[m + (size_t)nylm * n] : 0.0;
    YT[q] = v;
    YT32[q] = (float)v;
  }
}

__global__ void k_to_f32(size_t n, const double *__restrict__ a, float *__restrict__ b) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) b[k] = (float)a[k];
}

}  // namespace

int make_ptables(ddx_handle *h) {
  h->ldm = (h->nylm + 15) / 16 * 16;
  h->ldnp = (h->ngrid + 31) / 32 * 32;
  h->nylm4 = (h->nylm + 3) / 4 * 4;
  h->ldnt = (h->ngrid + 15) / 16 * 16;
  const size_t n1 = (size_t)h->ldm * h->ldnp, n2 = (size_t)h->nylm4 * h->ldnt;
  DDX_TRY(h->YwP.reserve(n1));
  DDX_TRY(h->YwP32.reserve(n1));
  DDX_TRY(h->YT.reserve(n2));
  DDX_TRY(h->YT32.reserve(n2));
  hipLaunchKernelGGL(k_pad_tables, dim3((n1 + n2 + 255) / 256), dim3(256), 0, h->stream, h->nylm,
                     h->ngrid, h->ldm, h->ldnp, h->nylm4, h->ldnt, h->Y.p, h->Yw.p, h->YwP.p,
                     h->YT.p, h->YwP32.p, h->YT32.p);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

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

template <typename T, int NT>
void launch_ps(ddx_handle *h, hipStream_t s, double alpha, const T *YwP, const T *P, double *y,
               const DiisState *gate, hipEvent_t b, hipEvent_t e) {
  hipExtLaunchKernelGGL((k_project_s<T, NT>), dim3((h->n + 63) / 64), dim3(256), 0, s, b, e, 0,
                        h->nylm, h->ngrid, h->ldm, h->n, alpha, YwP, P, y, gate,
                        h->ninst > 1 ? h->inst.p : nullptr, h->ninst);
}

// the staged projection: systems of at least 1024 workgroups of 64 spheres,
// one point range, an even point count, up to NTMAX row tiles
template <typename T, int NTMAX>
bool launch_project_s(ddx_handle *h, hipStream_t s, double alpha, const T *YwP, const T *P,
                      double *y, const DiisState *gate, int ysplit, hipEvent_t b, hipEvent_t e) {
  const int ntile = (h->nylm + 15) / 16;
  if (ysplit != 1 || ntile > NTMAX || h->ngrid % 2 != 0 || h->ngrid < 2 ||
      (h->n + 63) / 64 < 1024)
    return false;
  switch (ntile) {
    case 1: launch_ps<T, 1>(h, s, alpha, YwP, P, y, gate, b, e); break;
    case 2: launch_ps<T, 2>(h, s, alpha, YwP, P, y, gate, b, e); break;
    case 3: launch_ps<T, 3>(h, s, alpha, YwP, P, y, gate, b, e); break;
    case 4: launch_ps<T, 4>(h, s, alpha, YwP, P, y, gate, b, e); break;
    case 5: if constexpr (NTMAX >= 5) launch_ps<T, 5>(h, s, alpha, YwP, P, y, gate, b, e); break;
    case 6: if constexpr (NTMAX >= 6) launch_ps<T, 6>(h, s, alpha, YwP, P, y, gate, b, e); break;
    case 7: if constexpr (NTMAX >= 7) launch_ps<T, 7>(h, s, alpha, YwP, P, y, gate, b, e); break;
    default: if constexpr (NTMAX >= 8) launch_ps<T, 8>(h, s, alpha, YwP, P, y, gate, b, e); break;
  }
  return true;
}

template <int NT>
void launch_pg(ddx_handle *h, hipStream_t s, double alpha, const float *rw, double *y,
               const DiisState *gate, hipEvent_t b, hipEvent_t e) {
  hipExtLaunchKernelGGL(k_project_g<NT>, dim3((h->n + 63) / 64), dim3(256), 0, s, b, e, 0,
                        h->nylm, h->ngrid, h->ldm, h->n, h->ntri, alpha, h->YwP32.p, h->ptoff.p,
                        rw, y, gate, h->ninst > 1 ? h->inst.p : nullptr, h->ninst);
}

int project32_gathered(ddx_handle *h, hipStream_t s, double alpha, const float *rw, double *y,
                       const DiisState *gate, int ysplit, bool *done) {
  *done = false;
  const int ntile = (h->nylm + 15) / 16;
  if (h->n <= 0 || ysplit != 1 || ntile > 8 || (h->n + 63) / 64 < 1024) return DDX_OK;
  TimedLaunch tl(h, DDX_K_PROJECT, s);
  switch (ntile) {
    case 1: launch_pg<1>(h, s, alpha, rw, y, gate, tl.b, tl.e); break;
    case 2: launch_pg<2>(h, s, alpha, rw, y, gate, tl.b, tl.e); break;
    case 3: launch_pg<3>(h, s, alpha, rw, y, gate, tl.b, tl.e); break;
    case 4: launch_pg<4>(h, s, alpha, rw, y, gate, tl.b, tl.e); break;
    case 5: launch_pg<5>(h, s, alpha, rw, y, gate, tl.b, tl.e); break;
    case 6: launch_pg<6>(h, s, alpha, rw, y, gate, tl.b, tl.e); break;
    case 7: launch_pg<7>(h, s, alpha, rw, y, gate, tl.b, tl.e); break;
    default: launch_pg<8>(h, s, alpha, rw, y, gate, tl.b, tl.e); break;
  }
  DDX_HIP(hipGetLastError());
  *done = true;
  return DDX_OK;
}

int project32(ddx_handle *h, hipStream_t s, double alpha, const float *P, double *y,
              const DiisState *gate, int ysplit) {
  if (h->n <= 0) return DDX_OK;
  TimedLaunch tl(h, DDX_K_PROJECT, s);
  if (launch_project_s<float, 8>(h, s, alpha, h->YwP32.p, P, y, gate, ysplit, tl.b, tl.e)) {
    DDX_HIP(hipGetLastError());
    return DDX_OK;
  }
  hipExtLaunchKernelGGL(k_project32, dim3((h->n + 15) / 16, ysplit), dim3(256), 0, s, tl.b, tl.e,
                        0, h->nylm, h->ngrid, h->n, alpha, h->Yw32.p, P, y, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int expand32(ddx_handle *h, hipStream_t s, const double *S, float *xi, bool wscale,
             const DiisState *gate) {
  if (h->n <= 0) return DDX_OK;
  const int J = expand_tiles<float>(h->nylm);
  const size_t lds = expand_lds<float>(h->nylm, J);
  TimedLaunch tl(h, DDX_K_EXPAND, s);
  const int nx = (h->n + 16 * J - 1) / (16 * J);
  hipExtLaunchKernelGGL(J == 2 ? k_expand_t<float, 2> : k_expand_t<float, 1>,
                        dim3(nx, expand_split(nx, h->ngrid)), dim3(256), lds, s, tl.b, tl.e, 0,
                        h->nylm, h->ngrid, h->ldnt, h->n, h->YT32.p, S, h->grid.p, wscale ? 1 : 0,
                        xi, gate, h->ninst > 1 ? h->inst.p : nullptr, h->ninst);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int project(ddx_handle *h, hipStream_t s, double alpha, const double *P, double *y,
            const DiisState *gate, int ysplit) {
  if (h->n <= 0) return DDX_OK;
  TimedLaunch tl(h, DDX_K_PROJECT, s);
  if (launch_project_s<double, 4>(h, s, alpha, h->YwP.p, P, y, gate, ysplit, tl.b, tl.e)) {
    DDX_HIP(hipGetLastError());
    return DDX_OK;
  }
  hipExtLaunchKernelGGL(k_project, dim3((h->n + 15) / 16, ysplit), dim3(256), 0, s, tl.b, tl.e, 0,
                        h->nylm, h->ngrid, h->n, alpha, h->Yw.p, P, y, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int expand(ddx_handle *h, hipStream_t s, const double *S, double *xi, bool wscale,
           const DiisState *gate) {
  if (h->n <= 0) return DDX_OK;
  const int J = expand_tiles<double>(h->nylm);
  const size_t lds = expand_lds<double>(h->nylm, J);
  TimedLaunch tl(h, DDX_K_EXPAND, s);
  const int nx = (h->n + 16 * J - 1) / (16 * J);
  hipExtLaunchKernelGGL(J == 2 ? k_expand_t<double, 2> : k_expand_t<double, 1>,
                        dim3(nx, expand_split(nx, h->ngrid)), dim3(256), lds, s, tl.b, tl.e, 0,
                        h->nylm, h->ngrid, h->ldnt, h->n, h->YT.p, S, h->grid.p, wscale ? 1 : 0,
                        xi, gate, h->ninst > 1 ? h->inst.p : nullptr, h->ninst);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

}  // namespace ddx
