// Device-side helpers shared by the kernels of libddx.
#pragma once

#include <hip/hip_runtime.h>

#include "ddx_common.hpp"

namespace ddx {

constexpr int kWave = 64;

typedef double v4d __attribute__((ext_vector_type(4)));

// Switching function chi(t) of the interior variant (se = -1): 1 below 1-eta,
// 0 from 1 on, a quintic in between (fsw, ddcosmo.f90:566-607).
__device__ __forceinline__ double fsw(double t, double eta) {
#pragma clang fp contract(off)
  if (t >= 1.0) return 0.0;
  if (t <= 1.0 - eta) return 1.0;
  double a = 15.0 * eta - 12.0;
  double b = 10.0 * eta * eta - 15.0 * eta + 6.0;
  double e5 = eta * eta * eta * eta * eta;
  return ((t - 1.0) * (t - 1.0) * (1.0 - t) * (6.0 * t * t + a * t + b)) / e5;
}

// d chi / dt (dfsw, ddcosmo.f90:616-655)
__device__ __forceinline__ double dfsw(double t, double eta) {
#pragma clang fp contract(off)
  if (t >= 1.0 || t <= 1.0 - eta) return 0.0;
  double e5 = eta * eta * eta * eta * eta;
  return 30.0 * (1.0 - t) * (t - 1.0) * (t - 1.0 + eta) * (t - 1.0 + eta) / e5;
}

// Geometry is evaluated without FMA contraction, in the reference's operation
// order, so that every inside/outside decision (neighbour pairs, t < 1) is
// the reference's own and is identical in every kernel that recomputes it.
//
// v = c_o + r_o s_n - c_c : the point n of sphere o seen from sphere c
// (ddcosmo.f90:378, 1325, 1707).
__device__ __forceinline__ void point_vec(const double4 &so, const double4 &g,
                                          const double4 &sc, double v[3]) {
#pragma clang fp contract(off)
  v[0] = so.x + so.w * g.x - sc.x;
  v[1] = so.y + so.w * g.y - sc.y;
  v[2] = so.z + so.w * g.z - sc.z;
}

// |v| as sqrt(dot_product(v, v)) (ddcosmo.f90:379)
__device__ __forceinline__ double norm3(const double v[3]) {
#pragma clang fp contract(off)
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// |c_i - c_j| <= r_i + r_j, compared as squares (ddcosmo.f90:314-316)
__device__ __forceinline__ bool spheres_overlap(const double4 &a, const double4 &b) {
#pragma clang fp contract(off)
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  double d2 = dx * dx + dy * dy + dz * dz;
  double r2 = (a.w + b.w) * (a.w + b.w);
  return d2 <= r2;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ int wave_sum_int(int v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// number of set ballot bits below this lane
__device__ __forceinline__ int ballot_prefix(unsigned long long b) {
  unsigned long long below = (lane_id() == 0) ? 0ull : (~0ull >> (64 - lane_id()));
  return __popcll(b & below);
}

// Block-wide sum of one double per thread for blockDim.x == 256, fixed order.
__device__ __forceinline__ double block_sum_256(double v, double *lds4) {
  v = wave_sum(v);
  if (lane_id() == 0) lds4[threadIdx.x / kWave] = v;
  __syncthreads();
  double s = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  __syncthreads();
  return s;
}

}  // namespace ddx
