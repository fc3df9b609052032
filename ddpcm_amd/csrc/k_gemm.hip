// FP64 matrix products on the gfx950 matrix cores (v_mfma_f64_16x16x4_f64).
//
// C(m, j) = alpha * sum_k A(m, k) B(k, j) with arbitrary element strides, so
// one kernel serves
//   the projection onto harmonics  y(lm, i) = -sum_n w_n Y(lm, n) P(n, i)
//     (intrhs, ddcosmo.f90:789-821, applied to every sphere at once), and
//   the expansion on the grid     xi(n, i) = sum_lm Y(lm, n) S(lm, i)
//     (matvec.f90:196-204, forces_dd.f90:108-113).
// Tile 64 x 64 per 256-thread workgroup, K staged through LDS 16 at a time;
// each wave owns a 32 x 32 quadrant = 2 x 2 MFMA 16x16 blocks.
//
// Operand layout of v_mfma_f64_16x16x4_f64 (one double per lane):
//   A(i, k) in lane i + 16 k, B(k, j) in lane j + 16 k,
//   D(i, j) in lane j + 16 (i % 4), element i / 4 of the 4-double result
//   (measured on gfx950).
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

constexpr int kTM = 64, kTN = 64, kTK = 16;

__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, double alpha,
                                                 const double *__restrict__ A, long sam, long sak,
                                                 const double *__restrict__ B, long sbk, long sbn,
                                                 double *__restrict__ C, long scm, long scn,
                                                 const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  __shared__ double As[kTK][kTM + 1];
  __shared__ double Bs[kTK][kTN + 1];
  const int m0 = blockIdx.x * kTM, n0 = blockIdx.y * kTN;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const bool a_kfast = (sak == 1);
  const bool b_kfast = (sbk == 1);
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  for (int k0 = 0; k0 < K; k0 += kTK) {
#pragma unroll
    for (int e = tid; e < kTK * kTM; e += 256) {
      int kk, mm;
      if (a_kfast) { kk = e % kTK; mm = e / kTK; }
      else { kk = e / kTM; mm = e % kTM; }
      int gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < K) ? A[gm * sam + gk * sak] : 0.0;
    }
#pragma unroll
    for (int e = tid; e < kTK * kTN; e += 256) {
      int kk, nn;
      if (b_kfast) { kk = e % kTK; nn = e / kTK; }
      else { kk = e / kTN; nn = e % kTN; }
      int gn = n0 + nn, gk = k0 + kk;
      Bs[kk][nn] = (gn < N && gk < K) ? B[gk * sbk + gn * sbn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 4) {
      const int r = lane & 15, kq = kk + (lane >> 4);
      double a0 = As[kq][wm + r], a1 = As[kq][wm + 16 + r];
      double b0 = Bs[kq][wn + r], b1 = Bs[kq][wn + 16 + r];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int gj = n0 + wn + b * 16 + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gi = m0 + wm + a * 16 + (lane >> 4) + 4 * q;
        if (gi < M && gj < N) C[gi * scm + gj * scn] = alpha * acc[a][b][q];
      }
    }
}

}  // namespace

int gemm_f64(int M, int N, int K, double alpha, const double *A, long sam, long sak,
             const double *B, long sbk, long sbn, double *C, long scm, long scn,
             const DiisState *gate, hipStream_t s) {
  if (M <= 0 || N <= 0) return DDX_OK;
  dim3 g((M + kTM - 1) / kTM, (N + kTN - 1) / kTN);
  hipLaunchKernelGGL(k_gemm_f64, g, dim3(256), 0, s, M, N, K, alpha, A, sam, sak, B, sbk, sbn,
                     C, scm, scn, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

}  // namespace ddx
