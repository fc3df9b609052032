// The learned initial guess (no counterpart in the reference, which always
// starts from D^-1 rhs, cosmo.f90:136-138,174-176).
//
// A small network maps, sphere by sphere, the scaled Jacobi guess
// u_i = D^-1 g_i / s (s: rms of D^-1 g over the system) and 16 features of
// the sphere's neighbourhood to a correction c_i, and the solver starts from
//     X_i = D^-1 g_i + s c_i.
// Two hidden layers of 128 with SiLU.  Weights are kept in FP32 (Adam) with
// BF16 copies; every product runs on v_mfma_f32_16x16x16bf16_1k with FP32
// accumulation.  It learns online from the solutions the solver converges
// to; the Jacobi/DIIS iteration still decides convergence, so a poor guess
// costs iterations, never accuracy.
//
// Layouts: activations row-major (sphere, feature), weights row-major
// (in, out); widths padded to multiples of 16 with zeros.
#include <cstdint>

#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

#include <cmath>
#include <random>

namespace ddx {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float sg = 1.0f / (1.0f + __expf(-z));
  return sg * (1.0f + z * (1.0f - sg));
}

// C(M x N) (=|+=) A B: A(i, k) = a[i ai + k ak], B(k, j) = b[k bk + j bj]
// (BF16), C(i, j) = c[i ldc + j] (FP32).  One wave per 16 x 16 tile of C:
// lane l holds A(l % 16, 4 (l / 16) .. + 3) and B(4 (l / 16) .. + 3, l % 16)
// of each 16-deep step, and D(4 (l / 16) + e, l % 16) in element e.
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K,
                                                   const unsigned short *__restrict__ a, long ai,
                                                   long ak, const unsigned short *__restrict__ b,
                                                   long bk, long bj, float *__restrict__ c,
                                                   long ldc, int accumulate, int kchunk) {
  const int wave = threadIdx.x / kWave, lane = lane_id();
  const int tiles_n = (N + 15) / 16;
  const long tile = (long)blockIdx.x * 4 + wave;
  const int ti = (int)(tile / tiles_n), tj = (int)(tile % tiles_n);
  if ((long)ti * 16 >= M) return;
  const int i = ti * 16 + (lane & 15), j = tj * 16 + (lane & 15), kq = (lane >> 4) * 4;
  // split K: this workgroup's range (kchunk a multiple of 16), partials of
  // split y at c + y M ldc
  const int kb = (int)blockIdx.y * kchunk, ke = min(K, kb + kchunk);
  if (gridDim.y > 1) c += (long)blockIdx.y * M * ldc;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 16) {
    bf16x4 av, bv;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = k0 + kq + e;
      av[e] = (i < M && k < ke) ? (short)a[(long)i * ai + (long)k * ak] : (short)0;
      bv[e] = (j < N && k < ke) ? (short)b[(long)k * bk + (long)j * bj] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int r = ti * 16 + (lane >> 4) * 4 + e;
    if (r < M && j < N) {
      float *o = c + (long)r * ldc + j;
      *o = (accumulate && gridDim.y == 1) ? *o + acc[e] : acc[e];
    }
  }
}

// The same product where A's rows are contiguous (ak = 1) and N <= 16 NT:
// one wave per 16 rows of C computes all NT column tiles, so that each row
// of A is read once (8-byte loads of four BF16 values) instead of once per
// column tile.  Every tile's chain of MFMAs is k_gemm_bf16's, so C is
// bitwise the same.
template <int NT>
__global__ __launch_bounds__(256) void k_gemm_bf16_rows(int M, int N, int K,
                                                        const unsigned short *__restrict__ a,
                                                        long ai,
                                                        const unsigned short *__restrict__ b,
                                                        long bk, long bj, float *__restrict__ c,
                                                        long ldc, int accumulate) {
  const int lane = lane_id();
  const long ti = (long)blockIdx.x * 4 + threadIdx.x / kWave;
  if (ti * 16 >= M) return;
  const int i = (int)ti * 16 + (lane & 15), col = lane & 15, kq = (lane >> 4) * 4;
  const unsigned short *ar = a + (long)min(i, M - 1) * ai;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int k = k0 + kq;
    bf16x4 av;
    if (i < M && k + 3 < K) {
      const uint2 w = *reinterpret_cast<const uint2 *>(ar + k);
      av[0] = (short)(w.x & 0xffffu);
      av[1] = (short)(w.x >> 16);
      av[2] = (short)(w.y & 0xffffu);
      av[3] = (short)(w.y >> 16);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) av[e] = (i < M && k + e < K) ? (short)ar[k + e] : (short)0;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int j = 16 * t + col;
      bf16x4 bv;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        bv[e] = (j < N && k + e < K) ? (short)b[(long)(k + e) * bk + (long)j * bj] : (short)0;
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int j = 16 * t + col;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r = (int)ti * 16 + kq + e;
      if (r < M && j < N) {
        float *o = c + (long)r * ldc + j;
        *o = accumulate ? *o + acc[t][e] : acc[t][e];
      }
    }
  }
}

// Z += bias (per column); H = act(Z) in BF16 (identity for the output layer)
__global__ void k_bias_act(long M, int N, float *__restrict__ z, const float *__restrict__ bias,
                           unsigned short *__restrict__ h, int act) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const float v = z[k] + bias[k % N];
  z[k] = v;
  if (h) h[k] = to_bf16(act ? silu(v) : v);
}

// dZ = dH act'(Z), FP32 and BF16
__global__ void k_act_grad(long M, int N, const float *__restrict__ dh, const float *__restrict__ z,
                           float *__restrict__ dz, unsigned short *__restrict__ dzb) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const float v = dh[k] * dsilu(z[k]);
  dz[k] = v;
  dzb[k] = to_bf16(v);
}

// mean squared error over the nout real outputs; dY = 2 (Y - T) / (M nout)
// (FP32 and BF16), partial losses per block
__global__ __launch_bounds__(256) void k_mse(long M, int nout, int ld, const float *__restrict__ y,
                                             const float *__restrict__ t, float *__restrict__ dy,
                                             unsigned short *__restrict__ dyb,
                                             double *__restrict__ part) {
  __shared__ double lds[4];
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (k < M * ld) {
    const int col = (int)(k % ld);
    float g = 0.f;
    if (col < nout) {
      const float d = y[k] - t[k];
      l = (double)d * d;
      g = 2.0f * d / (float)(M * nout);
    }
    dy[k] = g;
    dyb[k] = to_bf16(g);
  }
  const double s = block_sum_256(l, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// db(j) = sum_i dZ(i, j): one workgroup per column, thread t summing rows
// t, t + 256, ..., then a fixed-order tree over the threads
__global__ __launch_bounds__(256) void k_col_sum(long M, int N, const float *__restrict__ dz,
                                                 float *__restrict__ db) {
  __shared__ float red[256];
  const int j = blockIdx.x;
  float s = 0.f;
  for (long i = threadIdx.x; i < M; i += 256) s += dz[i * N + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] = red[0];
}

// C = sum of the nsplit partial products of a split-K product, in split
// order (part: nsplit blocks of M x ldc)
__global__ void k_splitk_sum(long M, int N, int nsplit, const float *__restrict__ part,
                             float *__restrict__ c, long ldc, int accumulate) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M * N) return;
  const long r = k / N, j = k % N;
  float s = 0.f;
  for (int q = 0; q < nsplit; ++q) s += part[(long)q * M * ldc + r * ldc + j];
  float *o = c + r * ldc + j;
  *o = accumulate ? *o + s : s;
}

// Adam on FP32 master weights, BF16 copy refreshed
__global__ void k_adam(long n, float *__restrict__ w, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float c1, float c2, unsigned short *__restrict__ wb) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.f - b1) * gk;
  const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  const float wk = w[k] - lr * (mk * c1) / (sqrtf(vk * c2) + 1e-8f);
  w[k] = wk;
  if (wb) wb[k] = to_bf16(wk);
}

__global__ void k_to_bf16(long n, const float *__restrict__ x, unsigned short *__restrict__ y) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k] = to_bf16(x[k]);
}

// scale = sqrt(mean over all elements of (facl v)^2): partial sums per
// workgroup (grid-stride), then one wave adds them in order
__global__ __launch_bounds__(256) void k_sumsq_part(size_t nn, int nylm,
                                                    const double *__restrict__ facl,
                                                    const double *__restrict__ v,
                                                    double *__restrict__ part) {
  __shared__ double lds[4];
  double s = 0.0;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < nn; k += (size_t)gridDim.x * 256) {
    const double u = facl[k % nylm] * v[k];
    s += u * u;
  }
  s = block_sum_256(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kWave) void k_rms(int nb, size_t nn, const double *__restrict__ part,
                                               double *__restrict__ scale) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += kWave) s += part[b];
  s = wave_sum(s);
  if (threadIdx.x == 0) *scale = nn ? sqrt(s / (double)nn) + 1e-300 : 1.0;
}

// Inputs of sphere i: u_i = D^-1 v_i / scale (nylm values), then kGuessGeo
// features of its neighbourhood, zero padding up to inp.  With
// w_ij = (r_i + r_j - d_ij) / (r_i + r_j), the overlap of i and j:
//   r_i / 4, degree / 16, exposed fraction and mean ui of i's points,
//   sum w_ij, min d_ij / (r_i + r_j), sum r_j / d_ij / 4,
//   sum w_ij e_ij (3), sum w_ij u_j,00, sum w_ij u_j,00 e_ij (3),
//   sum w_ij |u_j,1.|, u_i,00 degree / 16
// (e_ij the unit vector from i to j).  One wave per sphere: lanes over the
// points and over the neighbours (at most 64), wave sums.
__global__ __launch_bounds__(256) void k_features(
    int n, int nylm, int inp, int ngrid, const double *__restrict__ facl,
    const double *__restrict__ v, const double *__restrict__ scale, const double4 *__restrict__ sph,
    const int *__restrict__ inl, const int *__restrict__ nl, const double *__restrict__ ui,
    unsigned short *__restrict__ F) {
  const int i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const int lane = lane_id();
  if (i >= n) return;   // whole waves
  const double is = 1.0 / *scale;
  unsigned short *f = F + (size_t)i * inp;
  const double *vi = v + (size_t)i * nylm;
  for (int k = lane; k < nylm; k += kWave) f[k] = to_bf16((float)(facl[k] * vi[k] * is));
  double nexp = 0.0, usum = 0.0;
  for (int p = lane; p < ngrid; p += kWave) {
    const double u = ui[(size_t)i * ngrid + p];
    nexp += u > 0.0 ? 1.0 : 0.0;
    usum += u;
  }
  nexp = wave_sum(nexp);
  usum = wave_sum(usum);
  const double4 si = sph[i];
  const int b = inl[i], deg = inl[i + 1] - b;
  double g[kGuessGeo];
#pragma unroll
  for (int k = 0; k < kGuessGeo; ++k) g[k] = 0.0;
  double dmin = 1.0;
  if (lane < deg) {
    const int j = nl[b + lane];
    const double4 sj = sph[j];
    const double dx = sj.x - si.x, dy = sj.y - si.y, dz = sj.z - si.z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double rr = si.w + sj.w;
    const double w = (rr - d) / rr;
    const double id = d > 0.0 ? 1.0 / d : 0.0;
    const double uj0 = facl[0] * v[(size_t)j * nylm] * is;
    double uj1 = 0.0;
    for (int k = 1; k < 4 && k < nylm; ++k) {
      const double t = facl[k] * v[(size_t)j * nylm + k] * is;
      uj1 += t * t;
    }
    g[4] = w;
    dmin = d / rr;
    g[6] = sj.w * id * 0.25;
    g[7] = w * dx * id;
    g[8] = w * dy * id;
    g[9] = w * dz * id;
    g[10] = w * uj0;
    g[11] = w * uj0 * dx * id;
    g[12] = w * uj0 * dy * id;
    g[13] = w * uj0 * dz * id;
    g[14] = w * sqrt(uj1);
  }
#pragma unroll
  for (int k = 4; k < 15; ++k)
    if (k != 5) g[k] = wave_sum(g[k]);
#pragma unroll
  for (int st = 32; st >= 1; st >>= 1) dmin = fmin(dmin, __shfl_xor(dmin, st, kWave));
  g[0] = si.w * 0.25;
  g[1] = deg / 16.0;
  g[2] = nexp / ngrid;
  g[3] = usum / ngrid;
  g[5] = dmin;
  g[15] = facl[0] * vi[0] * is * deg / 16.0;
  if (lane < kGuessGeo) {
    double gv = 0.0;
#pragma unroll
    for (int k = 0; k < kGuessGeo; ++k) gv = (lane == k) ? g[k] : gv;
    f[nylm + lane] = to_bf16((float)gv);
  }
  for (int k = nylm + kGuessGeo + lane; k < inp; k += kWave) f[k] = 0;
}

// T(i, lm) = (x - D^-1 v) / scale, zero padding up to outp
__global__ void k_target(size_t n, int nylm, int outp, const double *__restrict__ facl,
                         const double *__restrict__ v, const double *__restrict__ x,
                         const double *__restrict__ scale, float *__restrict__ t) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * outp) return;
  const size_t i = k / outp;
  const int c = (int)(k % outp);
  float r = 0.f;
  if (c < nylm) {
    const size_t e = i * nylm + c;
    r = (float)((x[e] - facl[c] * v[e]) / *scale);
  }
  t[k] = r;
}

// p = D^-1 v + scale Y, the learned map of v; with pold (the map of the
// previous right-hand side) x = x + p - pold instead of x = p; pold = p
__global__ void k_apply_guess(size_t nn, int nylm, int outp, const double *__restrict__ facl,
                              const double *__restrict__ v, const float *__restrict__ y,
                              const double *__restrict__ scale, double *__restrict__ x,
                              double *__restrict__ pold, int delta) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const size_t i = k / nylm;
  const int c = (int)(k % nylm);
  const double p = facl[c] * v[k] + *scale * (double)y[i * outp + c];
  x[k] = delta ? x[k] + (p - pold[k]) : p;
  pold[k] = p;
}

}  // namespace

// ---------------------------------------------------------------- host side

namespace {
int grid_of(long n, int t = 256) { return (int)((n + t - 1) / t); }
}  // namespace

int gemm_bf16(hipStream_t s, int M, int N, int K, const unsigned short *a, long ai, long ak,
              const unsigned short *b, long bk, long bj, float *c, long ldc, bool accumulate,
              float *ws, size_t ws_floats) {
  if (M <= 0 || N <= 0) return DDX_OK;
  const long tiles = (long)((M + 15) / 16) * ((N + 15) / 16);
  // few output tiles over a long K (the weight gradients, K = the spheres):
  // split K so that about 2048 waves share it, partials summed in order
  int nsplit = 1, kchunk = ((K + 15) / 16) * 16;
  if (ws && tiles < 2048 && K >= 4096) {
    nsplit = (int)std::min<long>((2048 + tiles - 1) / tiles, (K + 1023) / 1024);
    nsplit = (int)std::min<long>(nsplit, (long)(ws_floats / ((size_t)M * ldc)));
    if (nsplit > 1) kchunk = ((K + nsplit - 1) / nsplit + 15) / 16 * 16;
    nsplit = std::max(1, (K + kchunk - 1) / kchunk);
  }
  // contiguous rows of A (the activations): one wave per 16 rows, all columns
  const int ntn = (N + 15) / 16;
  if (nsplit == 1 && ak == 1 && ntn <= 8 && (ai % 4) == 0 &&
      (reinterpret_cast<uintptr_t>(a) % 8) == 0) {
    const dim3 g((unsigned)(((M + 15) / 16 + 3) / 4));
    const int acc = accumulate ? 1 : 0;
    switch (ntn) {
#define DDX_GR(NT)                                                                          \
  case NT:                                                                                 \
    hipLaunchKernelGGL(k_gemm_bf16_rows<NT>, g, dim3(256), 0, s, M, N, K, a, ai, b, bk, bj, c, \
                       ldc, acc);                                                          \
    break;
      DDX_GR(1) DDX_GR(2) DDX_GR(3) DDX_GR(4) DDX_GR(5) DDX_GR(6) DDX_GR(7) DDX_GR(8)
#undef DDX_GR
    }
    DDX_HIP(hipGetLastError());
    return DDX_OK;
  }
  hipLaunchKernelGGL(k_gemm_bf16, dim3((unsigned)((tiles + 3) / 4), nsplit), dim3(256), 0, s, M,
                     N, K, a, ai, ak, b, bk, bj, nsplit > 1 ? ws : c, ldc, accumulate ? 1 : 0,
                     kchunk);
  DDX_HIP(hipGetLastError());
  if (nsplit > 1) {
    const long mn = (long)M * N;
    hipLaunchKernelGGL(k_splitk_sum, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, (long)M,
                       N, nsplit, ws, c, ldc, accumulate ? 1 : 0);
    DDX_HIP(hipGetLastError());
  }
  return DDX_OK;
}

int convert_bf16(hipStream_t s, long n, const float *x, unsigned short *y) {
  if (n <= 0) return DDX_OK;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_of(n)), dim3(256), 0, s, n, x, y);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

namespace {
constexpr int kH = kGuessHidden;

int pad16(int x) { return (x + 15) / 16 * 16; }

// fresh network: W ~ N(0, 1/fan_in) (the last layer scaled by 0.1, so that
// the first corrections are small), b = 0, Adam moments 0
int net_init(hipStream_t s, GuessNet &g, int in, int out, unsigned seed) {
  g.in = in;
  g.inp = pad16(in);
  g.out = out;
  g.outp = pad16(out);
  g.t = 0;
  g.trained = false;
  const int rows[3] = {g.inp, kH, kH}, cols[3] = {kH, kH, g.outp};
  const int fan[3] = {in, kH, kH}, real_rows[3] = {in, kH, kH}, real_cols[3] = {kH, kH, out};
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int l = 0; l < 3; ++l) {
    const size_t nw = (size_t)rows[l] * cols[l];
    std::vector<float> w(nw, 0.f);
    const double sd = (l == 2 ? 0.1 : 1.0) / std::sqrt((double)fan[l]);
    for (int r = 0; r < real_rows[l]; ++r)
      for (int c = 0; c < real_cols[l]; ++c) w[(size_t)r * cols[l] + c] = (float)(sd * nd(rng));
    DDX_TRY(g.w[l].reserve(nw));
    DDX_TRY(g.mw[l].reserve(nw));
    DDX_TRY(g.vw[l].reserve(nw));
    DDX_TRY(g.gw[l].reserve(nw));
    DDX_TRY(g.wb[l].reserve(nw));
    DDX_TRY(g.b[l].reserve(cols[l]));
    DDX_TRY(g.mb[l].reserve(cols[l]));
    DDX_TRY(g.vb[l].reserve(cols[l]));
    DDX_TRY(g.gb[l].reserve(cols[l]));
    DDX_HIP(hipMemcpyAsync(g.w[l].p, w.data(), nw * 4, hipMemcpyHostToDevice, s));
    DDX_HIP(hipMemsetAsync(g.mw[l].p, 0, nw * 4, s));
    DDX_HIP(hipMemsetAsync(g.vw[l].p, 0, nw * 4, s));
    DDX_HIP(hipMemsetAsync(g.b[l].p, 0, cols[l] * 4, s));
    DDX_HIP(hipMemsetAsync(g.mb[l].p, 0, cols[l] * 4, s));
    DDX_HIP(hipMemsetAsync(g.vb[l].p, 0, cols[l] * 4, s));
    DDX_TRY(convert_bf16(s, (long)nw, g.w[l].p, g.wb[l].p));
  }
  DDX_HIP(hipStreamSynchronize(s));
  return DDX_OK;
}

int work_reserve(GuessWork &w, const GuessNet &g, long n) {
  DDX_TRY(w.F.reserve((size_t)n * g.inp + 1));
  DDX_TRY(w.H1.reserve((size_t)n * kH + 1));
  DDX_TRY(w.H2.reserve((size_t)n * kH + 1));
  DDX_TRY(w.Z1.reserve((size_t)n * kH + 1));
  DDX_TRY(w.Z2.reserve((size_t)n * kH + 1));
  DDX_TRY(w.dH.reserve((size_t)n * kH + 1));
  DDX_TRY(w.dZ.reserve((size_t)n * kH + 1));
  DDX_TRY(w.dZb.reserve((size_t)n * kH + 1));
  DDX_TRY(w.Y.reserve((size_t)n * g.outp + 1));
  DDX_TRY(w.T.reserve((size_t)n * g.outp + 1));
  DDX_TRY(w.dY.reserve((size_t)n * g.outp + 1));
  DDX_TRY(w.dYb.reserve((size_t)n * g.outp + 1));
  DDX_TRY(w.scale.reserve(2));
  DDX_TRY(w.part.reserve((size_t)grid_of(n * g.outp) + 257));
  DDX_TRY(w.ksplit.reserve((size_t)64 * kH * kH));   // up to 64 split-K partials
  return DDX_OK;
}

// Y = W3 act(W2 act(W1 F + b1) + b2) + b3 on the n rows of w.F
int net_forward(hipStream_t s, const GuessNet &g, GuessWork &w, long n) {
  DDX_TRY(gemm_bf16(s, (int)n, kH, g.inp, w.F.p, g.inp, 1, g.wb[0].p, kH, 1, w.Z1.p, kH, false));
  hipLaunchKernelGGL(k_bias_act, dim3(grid_of(n * kH)), dim3(256), 0, s, n, kH, w.Z1.p, g.b[0].p,
                     w.H1.p, 1);
  DDX_TRY(gemm_bf16(s, (int)n, kH, kH, w.H1.p, kH, 1, g.wb[1].p, kH, 1, w.Z2.p, kH, false));
  hipLaunchKernelGGL(k_bias_act, dim3(grid_of(n * kH)), dim3(256), 0, s, n, kH, w.Z2.p, g.b[1].p,
                     w.H2.p, 1);
  DDX_TRY(gemm_bf16(s, (int)n, g.outp, kH, w.H2.p, kH, 1, g.wb[2].p, g.outp, 1, w.Y.p, g.outp,
                    false));
  hipLaunchKernelGGL(k_bias_act, dim3(grid_of(n * g.outp)), dim3(256), 0, s, n, g.outp, w.Y.p,
                     g.b[2].p, (unsigned short *)nullptr, 0);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

// gradients of the mean squared error between Y and w.T into g.gw, g.gb;
// the loss partials into w.part (nblocks of them)
int net_backward(hipStream_t s, GuessNet &g, GuessWork &w, long n, int *nblocks) {
  const int ob = grid_of(n * g.outp);
  *nblocks = ob;
  hipLaunchKernelGGL(k_mse, dim3(ob), dim3(256), 0, s, n, g.out, g.outp, w.Y.p, w.T.p, w.dY.p,
                     w.dYb.p, w.part.p);
  // layer 3: dW3 = H2^T dY, db3, dH2 = dY W3^T
  DDX_TRY(gemm_bf16(s, kH, g.outp, (int)n, w.H2.p, 1, kH, w.dYb.p, g.outp, 1, g.gw[2].p, g.outp,
                    false, w.ksplit.p, w.ksplit.cap));
  hipLaunchKernelGGL(k_col_sum, dim3(g.outp), dim3(256), 0, s, n, g.outp, w.dY.p, g.gb[2].p);
  DDX_TRY(gemm_bf16(s, (int)n, kH, g.outp, w.dYb.p, g.outp, 1, g.wb[2].p, 1, g.outp, w.dH.p, kH,
                    false));
  // layer 2
  hipLaunchKernelGGL(k_act_grad, dim3(grid_of(n * kH)), dim3(256), 0, s, n, kH, w.dH.p, w.Z2.p,
                     w.dZ.p, w.dZb.p);
  DDX_TRY(gemm_bf16(s, kH, kH, (int)n, w.H1.p, 1, kH, w.dZb.p, kH, 1, g.gw[1].p, kH, false,
                    w.ksplit.p, w.ksplit.cap));
  hipLaunchKernelGGL(k_col_sum, dim3(kH), dim3(256), 0, s, n, kH, w.dZ.p, g.gb[1].p);
  DDX_TRY(gemm_bf16(s, (int)n, kH, kH, w.dZb.p, kH, 1, g.wb[1].p, 1, kH, w.dH.p, kH, false));
  // layer 1
  hipLaunchKernelGGL(k_act_grad, dim3(grid_of(n * kH)), dim3(256), 0, s, n, kH, w.dH.p, w.Z1.p,
                     w.dZ.p, w.dZb.p);
  DDX_TRY(gemm_bf16(s, g.inp, kH, (int)n, w.F.p, 1, g.inp, w.dZb.p, kH, 1, g.gw[0].p, kH, false,
                    w.ksplit.p, w.ksplit.cap));
  hipLaunchKernelGGL(k_col_sum, dim3(kH), dim3(256), 0, s, n, kH, w.dZ.p, g.gb[0].p);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int net_adam(hipStream_t s, GuessNet &g, float lr) {
  const float b1 = 0.9f, b2 = 0.999f;
  g.t += 1;
  const float c1 = 1.f / (1.f - std::pow(b1, (float)g.t)), c2 = 1.f / (1.f - std::pow(b2, (float)g.t));
  const int rows[3] = {g.inp, kH, kH}, cols[3] = {kH, kH, g.outp};
  for (int l = 0; l < 3; ++l) {
    const long nw = (long)rows[l] * cols[l];
    hipLaunchKernelGGL(k_adam, dim3(grid_of(nw)), dim3(256), 0, s, nw, g.w[l].p, g.gw[l].p,
                       g.mw[l].p, g.vw[l].p, lr, b1, b2, c1, c2, g.wb[l].p);
    hipLaunchKernelGGL(k_adam, dim3(grid_of(cols[l])), dim3(256), 0, s, (long)cols[l], g.b[l].p,
                       g.gb[l].p, g.mb[l].p, g.vb[l].p, lr, b1, b2, c1, c2,
                       (unsigned short *)nullptr);
  }
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

double sum_parts(hipStream_t s, const GuessWork &w, int nb) {
  std::vector<double> h(nb);
  if (hipMemcpyAsync(h.data(), w.part.p, nb * sizeof(double), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return -1.0;
  double t = 0.0;
  for (double v : h) t += v;
  return t;
}

// the features of the handle's spheres from v (rhs or psi) into w.F
int features(ddx_handle *h, GuessNet &g, GuessWork &w, const double *v) {
  hipStream_t s = h->stream;
  const size_t nn = (size_t)h->nylm * h->n;
  const int nb = grid_of((long)nn) < 256 ? grid_of((long)nn) : 256;
  hipLaunchKernelGGL(k_sumsq_part, dim3(nb), dim3(256), 0, s, nn, h->nylm, h->facl.p, v, w.part.p);
  hipLaunchKernelGGL(k_rms, dim3(1), dim3(kWave), 0, s, nb, nn, w.part.p, w.scale.p);
  hipLaunchKernelGGL(k_features, dim3(grid_of(h->n, 256 / kWave)), dim3(256), 0, s, h->n, h->nylm,
                     g.inp,
                     h->ngrid, h->facl.p, v, w.scale.p, h->sph.p, h->inl.p, h->nl.p, h->ui.p,
                     w.F.p);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int ensure_net(ddx_handle *h, int star) {
  GuessNet &g = h->gnet[star];
  if (g.in == 0) DDX_TRY(net_init(h->stream, g, h->nylm + kGuessGeo, h->nylm, h->guess_seed + 7 * star));
  return work_reserve(h->gwork, g, h->n);
}
}  // namespace

// The learned guess of solve `star` (0 direct from rhs, 1 adjoint from psi):
// p = D^-1 v + scale * network(features of v).  x = p, or with `delta` (x
// holding the previous solution, whose right-hand side's map is kept)
// x = x + p - p_previous.  D^-1 v while the network has not been trained.
int learned_guess(ddx_handle *h, int star, const double *v, double *x, bool delta) {
  GuessNet &g = h->gnet[star];
  if (!g.trained) {
    h->gprev_ok[star] = false;
    return delta ? DDX_OK : device_ldm1x(h, h->stream, v, x);
  }
  DDX_TRY(ensure_net(h, star));
  GuessWork &w = h->gwork;
  const size_t nn = (size_t)h->nylm * h->n;
  DDX_TRY(h->gprev[star].reserve(nn + 1));
  DDX_TRY(features(h, g, w, v));
  DDX_TRY(net_forward(h->stream, g, w, h->n));
  const int d = delta && h->gprev_ok[star] ? 1 : 0;
  hipLaunchKernelGGL(k_apply_guess, dim3(grid_of((long)nn)), dim3(256), 0, h->stream, nn, h->nylm,
                     g.outp, h->facl.p, v, w.Y.p, w.scale.p, x, h->gprev[star].p, d);
  DDX_HIP(hipGetLastError());
  h->gprev_ok[star] = true;
  return DDX_OK;
}

// `steps` Adam steps of the network of solve `star` towards the solution x
// of right-hand side v; *loss = the mean squared error before the last step
int train_guess(ddx_handle *h, int star, const double *v, const double *x, int steps, double lr,
                double *loss) {
  DDX_TRY(ensure_net(h, star));
  GuessNet &g = h->gnet[star];
  GuessWork &w = h->gwork;
  hipStream_t s = h->stream;
  DDX_TRY(features(h, g, w, v));
  const size_t nt = (size_t)h->n * g.outp;
  hipLaunchKernelGGL(k_target, dim3(grid_of((long)nt)), dim3(256), 0, s, (size_t)h->n, h->nylm,
                     g.outp, h->facl.p, v, x, w.scale.p, w.T.p);
  DDX_HIP(hipGetLastError());
  double l = 0.0;
  for (int it = 0; it < steps; ++it) {
    int nb = 0;
    DDX_TRY(net_forward(s, g, w, h->n));
    DDX_TRY(net_backward(s, g, w, h->n, &nb));
    if (it == steps - 1) l = sum_parts(s, w, nb) / ((double)h->n * g.out);
    DDX_TRY(net_adam(s, g, (float)lr));
  }
  if (loss) *loss = l;
  g.trained = true;
  return DDX_OK;
}

}  // namespace ddx

using namespace ddx;

extern "C" {

// C (M x N, FP32, row-major) = A (M x K) B (K x N), BF16 operands in host
// memory given as raw 16-bit patterns; for the tests of the product kernel.
int ddx_gemm_bf16(int M, int N, int K, const unsigned short *a, const unsigned short *b,
                  float *c) {
  if (M < 0 || N < 0 || K < 0 || (M && K && !a) || (K && N && !b) || (M && N && !c))
    return DDX_ERR_ARG;
  DevArray<unsigned short> da, db;
  DevArray<float> dc;
  DDX_TRY(da.reserve((size_t)M * K + 1));
  DDX_TRY(db.reserve((size_t)K * N + 1));
  DDX_TRY(dc.reserve((size_t)M * N + 1));
  if ((size_t)M * K) DDX_HIP(hipMemcpy(da.p, a, (size_t)M * K * 2, hipMemcpyHostToDevice));
  if ((size_t)K * N) DDX_HIP(hipMemcpy(db.p, b, (size_t)K * N * 2, hipMemcpyHostToDevice));
  DDX_TRY(gemm_bf16(nullptr, M, N, K, da.p, K, 1, db.p, N, 1, dc.p, N, false));
  if ((size_t)M * N) DDX_HIP(hipMemcpy(c, dc.p, (size_t)M * N * 4, hipMemcpyDeviceToHost));
  DDX_HIP(hipDeviceSynchronize());
  return DDX_OK;
}

// One forward and backward pass of a network given in host memory, for the
// tests: F (n x in, BF16 bits), T (n x out), W1 (in x 128), b1, W2 (128 x
// 128), b2, W3 (128 x out), b3 (FP32, row-major, unpadded).  Returns Y (n x
// out), the gradients of the mean squared error in the same layouts as the
// weights, and the loss.
int ddx_mlp_grad(int n, int in, int out, const unsigned short *F, const float *T,
                 const float *const *wts, float *Y, float *const *grads, double *loss) {
  if (n < 1 || in < 1 || out < 1 || !F || !T || !wts || !Y || !grads) return DDX_ERR_ARG;
  GuessNet g;
  GuessWork w;
  hipStream_t s = nullptr;
  DDX_TRY(net_init(s, g, in, out, 1));
  DDX_TRY(work_reserve(w, g, n));
  const int rows[3] = {g.inp, kH, kH}, cols[3] = {kH, kH, g.outp};
  const int rr[3] = {in, kH, kH}, rc[3] = {kH, kH, out};
  for (int l = 0; l < 3; ++l) {
    std::vector<float> wp((size_t)rows[l] * cols[l], 0.f), bp(cols[l], 0.f);
    for (int r = 0; r < rr[l]; ++r)
      for (int c = 0; c < rc[l]; ++c) wp[(size_t)r * cols[l] + c] = wts[2 * l][(size_t)r * rc[l] + c];
    for (int c = 0; c < rc[l]; ++c) bp[c] = wts[2 * l + 1][c];
    DDX_HIP(hipMemcpy(g.w[l].p, wp.data(), wp.size() * 4, hipMemcpyHostToDevice));
    DDX_HIP(hipMemcpy(g.b[l].p, bp.data(), bp.size() * 4, hipMemcpyHostToDevice));
    DDX_TRY(convert_bf16(s, (long)wp.size(), g.w[l].p, g.wb[l].p));
  }
  std::vector<unsigned short> fp((size_t)n * g.inp, 0);
  std::vector<float> tp((size_t)n * g.outp, 0.f);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < in; ++k) fp[(size_t)i * g.inp + k] = F[(size_t)i * in + k];
    for (int k = 0; k < out; ++k) tp[(size_t)i * g.outp + k] = T[(size_t)i * out + k];
  }
  DDX_HIP(hipMemcpy(w.F.p, fp.data(), fp.size() * 2, hipMemcpyHostToDevice));
  DDX_HIP(hipMemcpy(w.T.p, tp.data(), tp.size() * 4, hipMemcpyHostToDevice));
  int nb = 0;
  DDX_TRY(net_forward(s, g, w, n));
  DDX_TRY(net_backward(s, g, w, n, &nb));
  if (loss) *loss = sum_parts(s, w, nb) / ((double)n * out);
  std::vector<float> yp((size_t)n * g.outp);
  DDX_HIP(hipMemcpy(yp.data(), w.Y.p, yp.size() * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < out; ++k) Y[(size_t)i * out + k] = yp[(size_t)i * g.outp + k];
  for (int l = 0; l < 3; ++l) {
    std::vector<float> gw((size_t)rows[l] * cols[l]), gb(cols[l]);
    DDX_HIP(hipMemcpy(gw.data(), g.gw[l].p, gw.size() * 4, hipMemcpyDeviceToHost));
    DDX_HIP(hipMemcpy(gb.data(), g.gb[l].p, gb.size() * 4, hipMemcpyDeviceToHost));
    for (int r = 0; r < rr[l]; ++r)
      for (int c = 0; c < rc[l]; ++c) grads[2 * l][(size_t)r * rc[l] + c] = gw[(size_t)r * cols[l] + c];
    for (int c = 0; c < rc[l]; ++c) grads[2 * l + 1][c] = gb[c];
  }
  DDX_HIP(hipDeviceSynchronize());
  return DDX_OK;
}

int ddx_set_guess(ddx_handle *h, int mode) {
  if (!h || mode < DDX_GUESS_REFERENCE || mode > DDX_GUESS_LEARNED_PREVIOUS) {
    set_error("ddx_set_guess: mode must be one of DDX_GUESS_*");
    return DDX_ERR_ARG;
  }
  h->guess_mode = mode;
  return DDX_OK;
}

int ddx_guess_reset(ddx_handle *h, unsigned seed) {
  if (!h) return DDX_ERR_ARG;
  h->guess_seed = seed;
  for (int k = 0; k < 2; ++k) h->gnet[k].in = 0;
  for (int k = 0; k < 2; ++k) h->gnet[k].trained = false;
  for (int k = 0; k < 2; ++k) h->gprev_ok[k] = false;
  return DDX_OK;
}

int ddx_guess_train(ddx_handle *h, int steps, double lr, double *loss) {
  if (!h || steps < 1 || !(lr > 0.0)) return DDX_ERR_ARG;
  if (!h->have_sigma || !h->conv_sigma) {
    set_error("ddx_guess_train needs a converged solve on the handle");
    return DDX_ERR_STATE;
  }
  double l[2] = {0.0, 0.0};
  DDX_TRY(train_guess(h, 0, h->rhs.p, h->sigma.p, steps, lr, &l[0]));
  if (h->have_s && h->conv_s) DDX_TRY(train_guess(h, 1, h->psi.p, h->s.p, steps, lr, &l[1]));
  if (loss) {
    loss[0] = l[0];
    loss[1] = l[1];
  }
  DDX_HIP(hipStreamSynchronize(h->stream));
  return DDX_OK;
}

}  // extern "C"
