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

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// The FP32 sweep's record of a triple from its FP64 one (cos theta, sin theta,
// cos phi, sin phi) and (t, W): (t cos theta, t sin theta, cos phi, sin phi),
// with W apart; t^2 is the sum of the first two squared.
__device__ __forceinline__ float4 geo32_of(const double4 &g, const double2 &w) {
  return make_float4((float)(g.x * w.x), (float)(g.y * w.x), (float)g.z, (float)g.w);
}

// Orders the LDS accesses of one wave before and after this point: a wave's
// LDS operations complete in issue order, so only the compiler must not move
// them (for data one lane writes and others of the same wave read).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// Whether spheres s0 .. s1 - 1 (s1 > s0) still have work: gate is the
// solve's summary state (null: always); in a batch (inst: the instance of
// each sphere) the ninst instance states precede it, and the range is done
// when every instance it touches is.  Instances own contiguous sphere
// ranges, so a range touches inst[s0] .. inst[s1 - 1].
__device__ __forceinline__ bool spheres_active(const DiisState *gate, const int *inst, int ninst,
                                               int s0, int s1) {
  if (!gate) return true;
  if (gate->done) return false;
  if (!inst) return true;
  const int q1 = inst[s1 - 1];
  for (int q = inst[s0]; q <= q1; ++q)
    if (!gate[q - ninst].done) return true;
  return false;
}

// Value of this lane's partner at stage S of a six-stage cross-lane
// combine, without LDS: the lane 32 apart and the lane 16 apart (gfx950
// row swaps), then the mirror lane within 16 and within 8 and the quad
// swaps by 2 and by 1 (DPP).  Every stage pairs each lane with a distinct
// partner, so combining over the six stages covers all 64 lanes.
template <int S>
__device__ __forceinline__ unsigned partner_u32(unsigned v) {
  if constexpr (S == 0) {
    auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane_id() < 32) ? r[1] : r[0];
  } else if constexpr (S == 1) {
    auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane_id() & 16) ? r[0] : r[1];
  } else {
    constexpr int ctrl = S == 2 ? 0x140 : S == 3 ? 0x141 : S == 4 ? 0x4E : 0xB1;
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xf, 0xf, false);
  }
}

template <int S>
__device__ __forceinline__ double partner(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = partner_u32<S>((unsigned)b), hi = partner_u32<S>((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int S>
__device__ __forceinline__ int partner(int v) {
  return (int)partner_u32<S>((unsigned)v);
}

// Wave-wide sum; every lane receives the same bits (each stage adds two
// values that the pair of lanes holds alike, and addition commutes).
__device__ __forceinline__ double wave_sum(double v) {
  v += partner<0>(v);
  v += partner<1>(v);
  v += partner<2>(v);
  v += partner<3>(v);
  v += partner<4>(v);
  v += partner<5>(v);
  return v;
}

__device__ __forceinline__ int wave_sum_int(int v) {
  v += partner<0>(v);
  v += partner<1>(v);
  v += partner<2>(v);
  v += partner<3>(v);
  v += partner<4>(v);
  v += partner<5>(v);
  return v;
}

// largest value over the wave (every lane receives it)
__device__ __forceinline__ int wave_max_int(int v) {
  v = max(v, partner<0>(v));
  v = max(v, partner<1>(v));
  v = max(v, partner<2>(v));
  v = max(v, partner<3>(v));
  v = max(v, partner<4>(v));
  v = max(v, partner<5>(v));
  return v;
}

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
