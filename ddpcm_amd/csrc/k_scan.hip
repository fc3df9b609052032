// Exclusive prefix sum of int arrays, three-phase: per-block scan of
// 4096 items with block totals, recursive scan of the totals, offset add.
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const int *__restrict__ in,
                                                            int *__restrict__ out,
                                                            int *__restrict__ totals,
                                                            size_t n) {
  __shared__ int wsum[kScanThreads / kWave];
  size_t base = (size_t)blockIdx.x * kScanTile;
  int vals[kScanItems];
  int run = 0;
  // thread t owns items base + t*kScanItems ... (+kScanItems)
  size_t mine = base + (size_t)threadIdx.x * kScanItems;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    size_t g = mine + k;
    int v = (g < n) ? in[g] : 0;
    vals[k] = run;
    run += v;
  }
  // exclusive scan of per-thread totals across the block
  int lane = lane_id(), wave = threadIdx.x / kWave;
  int incl = run;
  for (int off = 1; off < kWave; off <<= 1) {
    int t = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += t;
  }
  if (lane == kWave - 1) wsum[wave] = incl;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < wave; ++w) wbase += wsum[w];
  int excl = wbase + incl - run;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    size_t g = mine + k;
    if (g < n) out[g] = excl + vals[k];
  }
  if (threadIdx.x == kScanThreads - 1 && totals) totals[blockIdx.x] = excl + run;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int *__restrict__ out,
                                                           const int *__restrict__ offs,
                                                           size_t n) {
  size_t base = (size_t)blockIdx.x * kScanTile;
  int o = offs[blockIdx.x];
  for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) {
    size_t g = base + k;
    if (g < n) out[g] += o;
  }
}

__global__ void k_scan_total(const int *__restrict__ in, int *__restrict__ out, size_t n) {
  // out[n] = out[n-1] + in[n-1]  (out[0] = 0 when n == 0)
  if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = n ? out[n - 1] + in[n - 1] : 0;
}

int scan_rec(const int *in, int *out, size_t n, int *tmp, size_t tmp_cap, hipStream_t s) {
  if (n == 0) return DDX_OK;
  size_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kScanThreads), 0, s, in, out, nullptr, n);
    DDX_HIP(hipGetLastError());
    return DDX_OK;
  }
  if (2 * nb + 1 > tmp_cap) {
    set_error("scan workspace too small");
    return DDX_ERR_ALLOC;
  }
  int *totals = tmp, *offs = tmp + nb;
  hipLaunchKernelGGL(k_scan_tile, dim3(nb), dim3(kScanThreads), 0, s, in, out, totals, n);
  DDX_HIP(hipGetLastError());
  DDX_TRY(scan_rec(totals, offs, nb, tmp + 2 * nb, tmp_cap - 2 * nb, s));
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanThreads), 0, s, out, offs, n);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

}  // namespace

int scan_exclusive(const int *in, int *out, size_t n, DevArray<int> &tmp, hipStream_t s) {
  size_t need = 0, m = n;
  while (m > (size_t)kScanTile) {
    m = (m + kScanTile - 1) / kScanTile;
    need += 2 * m + 1;
  }
  DDX_TRY(tmp.reserve(need + 16));
  DDX_TRY(scan_rec(in, out, n, tmp.p, tmp.cap, s));
  hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, s, in, out, n);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

}  // namespace ddx
