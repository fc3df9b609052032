// Device setup of one ddCOSMO system (the per-geometry part of ddinit,
// ddcosmo.f90:308-468), plus the triple lists the sweeps run on.
//
//   neighbours  CSR, increasing j, |c_i - c_j| <= r_i + r_j (ddcosmo.f90:308-324),
//               found through a spatial hash in O(n) instead of the O(n^2) scan
//   per point   fi = sum_j chi(t_n^ij), ui = max(0, 1 - fi) when fi <= 1,
//               zi in the switch region (ddcosmo.f90:359-411), plus a 64-bit
//               mask of the neighbours j with t_n^ij < 1
//   cavity      exposed points ui > 0, sphere-major (ddcosmo.f90:422-468)
//   triples     every (owner sphere o, point n, container c) with t < 1,
//               grouped by container c (owners ascending, then points, the
//               loop order of adjrhs, ddcosmo.f90:1316-1381), with the
//               geometry the sweeps need: cos/sin theta, cos/sin phi of
//               s = v/|v| (ylmbas, ddcosmo.f90:846-866), t and the weight
//               W = chi(t)/max(fi, 1) (ddcosmo.f90:1722-1730, 1340-1348).
//               pttri lists the triples of each point in neighbour order
//               (the loop order of calcv, ddcosmo.f90:1701-1739); tslot is
//               its inverse, the slot of each triple in its point's list.
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

#include <algorithm>
#include <cmath>

namespace ddx {
namespace {

constexpr int kSetupThreads = 256;

// Neighbour search on a spatial hash.  Cubes of edge h >= 2 max r: two
// spheres that overlap lie in the same or in adjacent cubes.  Cube (ix, iy,
// iz) of instance k goes to bucket hash(k, ix, iy, iz) mod T; a bucket may
// hold several cubes, and the exact overlap test sorts them out.
__device__ __forceinline__ unsigned cell_bucket(long long ix, long long iy, long long iz,
                                                int inst, unsigned mask) {
  unsigned long long v = (unsigned long long)ix * 0x9E3779B97F4A7C15ull;
  v ^= (unsigned long long)iy * 0xC2B2AE3D27D4EB4Full;
  v ^= (unsigned long long)iz * 0x165667B19E3779F9ull;
  v ^= (unsigned long long)(unsigned)inst * 0xD6E8FEB86659FD93ull;
  v ^= v >> 31;
  return (unsigned)v & mask;
}

__device__ __forceinline__ long long cube_of(double x, double inv_h) {
  return (long long)floor(x * inv_h);
}

__device__ __forceinline__ unsigned sphere_bucket(const double4 &s, int inst, double inv_h,
                                                  unsigned mask, int d) {
  return cell_bucket(cube_of(s.x, inv_h) + d % 3 - 1, cube_of(s.y, inv_h) + (d / 3) % 3 - 1,
                     cube_of(s.z, inv_h) + d / 9 - 1, inst, mask);
}

// bucket of each sphere's own cube (d = 13: offsets 0, 0, 0) and the bucket sizes
__global__ void k_cell_count(const double4 *__restrict__ sph, const int *__restrict__ inst,
                             int n, double inv_h, unsigned mask, int *__restrict__ key,
                             int *__restrict__ bcnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned b = sphere_bucket(sph[i], inst ? inst[i] : 0, inv_h, mask, 13);
  key[i] = (int)b;
  atomicAdd(&bcnt[b], 1);
}

// spheres bucket by bucket (order inside a bucket is free; bcnt ends at 0)
__global__ void k_cell_fill(int n, const int *__restrict__ key, const int *__restrict__ boff,
                            int *__restrict__ bcnt, int *__restrict__ bsph) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = key[i];
  bsph[boff[b] + atomicSub(&bcnt[b], 1) - 1] = i;
}

// One wave per sphere i: the spheres j != i of its instance that overlap it,
// from the buckets of the 27 cubes around it, in increasing j, into
// cand(kMaxDegree, i); deg[i] = their number (also when above kMaxDegree,
// which the host turns into an error).
__global__ __launch_bounds__(kSetupThreads) void k_nbr_find(
    const double4 *__restrict__ sph, const int *__restrict__ inst, int n, double inv_h,
    unsigned mask, const int *__restrict__ boff, const int *__restrict__ bsph,
    int *__restrict__ deg, int *__restrict__ cand) {
  constexpr int W = kSetupThreads / kWave;
  __shared__ int list[W][kMaxDegree];
  const int wv = threadIdx.x / kWave, lane = lane_id();
  const int i = blockIdx.x * W + wv;
  int cnt = 0;
  if (i < n) {
    const double4 si = sph[i];
    const int ki = inst ? inst[i] : 0;
    // lane d < 27 takes cube d; a bucket that an earlier cube maps to is
    // left out
    const unsigned myb = lane < 27 ? sphere_bucket(si, ki, inv_h, mask, lane) : 0xffffffffu;
    bool dup = false;
    for (int e = 0; e < 27; ++e) {
      const unsigned be = (unsigned)__shfl((int)myb, e, kWave);
      dup = dup || (e < lane && be == myb);
    }
    unsigned long long todo = __ballot(lane < 27 && !dup);
    while (todo) {
      const int d = __ffsll(todo) - 1;
      todo &= todo - 1;
      const unsigned b = (unsigned)__shfl((int)myb, d, kWave);
      const int beg = boff[b], end = boff[b + 1];
      for (int k0 = beg; k0 < end; k0 += kWave) {
        const int k = k0 + lane;
        bool hit = false;
        int j = 0;
        if (k < end) {
          j = bsph[k];
          hit = j != i && (!inst || inst[j] == ki) && spheres_overlap(si, sph[j]);
        }
        const unsigned long long bal = __ballot(hit);
        const int pos = cnt + ballot_prefix(bal);
        if (hit && pos < kMaxDegree) list[wv][pos] = j;
        cnt += __popcll(bal);
      }
    }
  }
  __syncthreads();
  if (i < n) {
    if (lane == 0) deg[i] = cnt;
    if (cnt <= kMaxDegree && lane < cnt) {
      const int j = list[wv][lane];
      int rank = 0;
      for (int m = 0; m < cnt; ++m) rank += list[wv][m] < j ? 1 : 0;
      cand[(size_t)i * kMaxDegree + rank] = j;
    }
  }
}

// CSR from the sorted candidates
__global__ void k_nbr_compact(int n, const int *__restrict__ inl, const int *__restrict__ cand,
                              int *__restrict__ nl, int *__restrict__ prow) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t / kMaxDegree), k = (int)(t % kMaxDegree);
  if (i >= n) return;
  const int b = inl[i];
  if (k < inl[i + 1] - b) {
    nl[b + k] = cand[t];
    prow[b + k] = i;
  }
}

// rev[q] = position of sphere prow[q] in the neighbour list of nl[q]
__global__ void k_nbr_rev(int nnl, const int *__restrict__ inl, const int *__restrict__ nl,
                          const int *__restrict__ prow, int *__restrict__ rev) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnl) return;
  int c = prow[q], o = nl[q];
  int r = -1;
  for (int k = inl[o]; k < inl[o + 1]; ++k)
    if (nl[k] == c) { r = k - inl[o]; break; }
  rev[q] = r;
}

// fi, ui, zi of every point (ddcosmo.f90:441-498), the mask of the
// neighbours that contain it and the cavity flags.  One wave per sphere i:
// its neighbours' centres and radii are staged in LDS (lane k loads
// neighbour k; the degree is at most 64), then the lanes take the sphere's
// points n = lane, lane + 64, ..., each adding its terms over the
// neighbours in list order.  The exact terms (square root, divisions) are
// evaluated only for the neighbours that may contain the point, |v|^2 at
// most r^2 (1 + 1e-12): every t < 1 passes that test by far more than any
// rounding.
__global__ __launch_bounds__(kSetupThreads) void k_points(
    const double4 *__restrict__ sph, const double4 *__restrict__ grid, int ngrid,
    const int *__restrict__ inl, const int *__restrict__ nl, double eta, int igrad,
    double *__restrict__ fi, double *__restrict__ ui, double *__restrict__ zi,
    unsigned long long *__restrict__ tmask, int *__restrict__ ptcnt,
    int *__restrict__ cavflag, int n) {
  constexpr int kW = kSetupThreads / kWave;
  __shared__ double4 nb[kW][kWave];
  const int wave = threadIdx.x / kWave, lane = lane_id();
  const int i = blockIdx.x * kW + wave;
  if (i >= n) return;
  const int beg = inl[i], deg = inl[i + 1] - beg;
  if (lane < deg) nb[wave][lane] = sph[nl[beg + lane]];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  const double4 si = sph[i];
  for (int pn = lane; pn < ngrid; pn += kWave) {
    const size_t p = (size_t)i * ngrid + pn;
    const double4 g = grid[pn];
    double f = 0.0, z0 = 0.0, z1 = 0.0, z2 = 0.0;
    unsigned long long mask = 0, cand = 0;
    // the neighbours that may contain the point (the cheap test), then the
    // exact terms of those only, in neighbour order: the lanes of a wave
    // would otherwise all wait through the square root and divisions of
    // every neighbour that contains any one of them
    for (int k = 0; k < deg; ++k) {
      const double4 sj = nb[wave][k];
      double v[3];
      point_vec(si, g, sj, v);
      const double d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      if (d2 <= sj.w * sj.w * (1.0 + 1e-12)) cand |= 1ull << k;
    }
    for (; cand; cand &= cand - 1) {
      const int k = __ffsll((long long)cand) - 1;
      const double4 sj = nb[wave][k];
      double v[3];
      point_vec(si, g, sj, v);
      double vv = norm3(v);
      double t = vv / sj.w;
      if (t < 1.0) {
        f += fsw(t, eta);
        mask |= 1ull << k;
        if (igrad && t > 1.0 - eta) {
          double fac = dfsw(t, eta) / sj.w;
          z0 += fac * v[0] / vv;
          z1 += fac * v[1] / vv;
          z2 += fac * v[2] / vv;
        }
      }
    }
    double u = (f <= 1.0) ? 1.0 - f : 0.0;
    fi[p] = f;
    ui[p] = u;
    if (igrad) {
      zi[3 * p + 0] = z0;
      zi[3 * p + 1] = z1;
      zi[3 * p + 2] = z2;
    }
    tmask[p] = mask;
    ptcnt[p] = __popcll(mask);
    cavflag[p] = u > 0.0 ? 1 : 0;
  }
}

__global__ __launch_bounds__(kSetupThreads) void k_ccav(const double4 *__restrict__ sph,
                                                        const double4 *__restrict__ grid,
                                                        int ngrid, const int *__restrict__ cavflag,
                                                        const int *__restrict__ cavidx,
                                                        double *__restrict__ ccav, size_t np) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  if (!cavflag[p]) return;
  // 32-bit point indices (run_setup checks the count)
  const unsigned i = (unsigned)p / (unsigned)ngrid, n = (unsigned)p - i * (unsigned)ngrid;
  int ic = cavidx[p];
  double4 si = sph[i], g = grid[n];
  ccav[3 * (size_t)ic + 0] = si.x + si.w * g.x;
  ccav[3 * (size_t)ic + 1] = si.y + si.w * g.y;
  ccav[3 * (size_t)ic + 2] = si.z + si.w * g.z;
}

// OR of a 64-bit value over the wave (every lane receives it)
__device__ __forceinline__ unsigned long long wave_or64(unsigned long long v) {
  unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
  lo |= partner_u32<0>(lo); hi |= partner_u32<0>(hi);
  lo |= partner_u32<1>(lo); hi |= partner_u32<1>(hi);
  lo |= partner_u32<2>(lo); hi |= partner_u32<2>(hi);
  lo |= partner_u32<3>(lo); hi |= partner_u32<3>(hi);
  lo |= partner_u32<4>(lo); hi |= partner_u32<4>(hi);
  lo |= partner_u32<5>(lo); hi |= partner_u32<5>(hi);
  return ((unsigned long long)hi << 32) | lo;
}

// The pair (container nl[inl[o] + kr], owner o) of owner o's neighbour kr,
// as an index of the container-major pair list: the owner's rank in the
// container's list is rev of the mirrored pair.
__device__ __forceinline__ int pair_of(const int *__restrict__ inl, const int *__restrict__ nl,
                                       const int *__restrict__ rev, int o, int kr) {
  const int q = inl[o] + kr;
  return inl[nl[q]] + rev[q];
}

// One wave per owner sphere o: the number of its points inside each of its
// neighbours (lane kr counts neighbour kr), in one pass over its masks.
__global__ __launch_bounds__(kSetupThreads) void k_pair_count(
    int n, int ngrid, const int *__restrict__ inl, const int *__restrict__ nl,
    const int *__restrict__ rev, const unsigned long long *__restrict__ tmask,
    int *__restrict__ paircnt) {
  const int o = blockIdx.x * (kSetupThreads / kWave) + threadIdx.x / kWave;
  if (o >= n) return;
  const int lane = lane_id();
  const int deg = inl[o + 1] - inl[o];
  int cnt = 0;
  for (int n0 = 0; n0 < ngrid; n0 += kWave) {
    const int pn = n0 + lane;
    const unsigned long long m = pn < ngrid ? tmask[(size_t)o * ngrid + pn] : 0ull;
    unsigned long long u = wave_or64(m);
    while (u) {
      const int kr = __ffsll((long long)u) - 1;
      u &= u - 1;
      const int c = __popcll(__ballot((m >> kr) & 1ull));
      if (lane == kr) cnt += c;
    }
  }
  if (lane < deg) paircnt[pair_of(inl, nl, rev, o, lane)] = cnt;
}

// One wave per owner sphere o, in one pass over its masks: for each point n
// and each neighbour c containing it, the triple's place in c's run of
// owner o (points ascending), its point index (tp) and its slot in the
// point's list (pttri, tslot).  Lane kr keeps the next place of neighbour
// kr's run.
__global__ __launch_bounds__(kSetupThreads) void k_pair_fill(
    int n, int ngrid, const int *__restrict__ inl, const int *__restrict__ nl,
    const int *__restrict__ rev, const unsigned long long *__restrict__ tmask,
    const int *__restrict__ pairoff, const int *__restrict__ ptoff, int *__restrict__ tp,
    int *__restrict__ pttri, int *__restrict__ tslot) {
  const int o = blockIdx.x * (kSetupThreads / kWave) + threadIdx.x / kWave;
  if (o >= n) return;
  const int lane = lane_id();
  const int deg = inl[o + 1] - inl[o];
  int run = 0;
  if (lane < deg) run = pairoff[pair_of(inl, nl, rev, o, lane)];
  for (int n0 = 0; n0 < ngrid; n0 += kWave) {
    const int pn = n0 + lane;
    const size_t p = (size_t)o * ngrid + pn;
    const unsigned long long m = pn < ngrid ? tmask[p] : 0ull;
    unsigned long long u = wave_or64(m);
    while (u) {
      const int kr = __ffsll((long long)u) - 1;
      u &= u - 1;
      const bool hit = (m >> kr) & 1ull;
      const unsigned long long b = __ballot(hit);
      const int base = __shfl(run, kr, kWave);
      if (hit) {
        const int idx = base + ballot_prefix(b);
        const int slot = ptoff[p] + __popcll(m & ((1ull << kr) - 1ull));
        tp[idx] = (int)p;
        pttri[slot] = idx;
        tslot[idx] = slot;
      }
      if (lane == kr) run += __popcll(b);
    }
  }
}

// The record of triple (owner point p = o ngrid + n, container c), as calcv
// and adjrhs evaluate it (ddcosmo.f90:1707-1727, 1325-1345; the angles as
// ylmbas takes them, ddcosmo.f90:845-857, with cos phi = 1 on the axis).
__device__ __forceinline__ void triple_record(const double4 &so, const double4 &g,
                                              const double4 &sc, double f, double eta,
                                              double4 &geo, double2 &tw) {
  double v[3];
  point_vec(so, g, sc, v);
  const double vv = norm3(v);
  const double t = vv / sc.w;
  const double s0 = v[0] / vv, s1 = v[1] / vv, s2 = v[2] / vv;
  const double cthe = s2;
  const double sthe = sqrt(1.0 - cthe * cthe);
  double cphi = 1.0, sphi = 0.0;
  if (sthe != 0.0) {
    cphi = s0 / sthe;
    sphi = s1 / sthe;
  }
  const double chi = fsw(t, eta);
  const double w = (f > 1.0) ? chi / f : chi;
  geo = make_double4(cthe, sthe, cphi, sphi);
  tw = make_double2(t, w);
}

// One wave per container c: the records of its triples, all lanes busy, and
// with g32 / w32 the FP32 sweeps' records of them (geo32_of, and W).
__global__ __launch_bounds__(kSetupThreads) void k_triple_geo(
    int n, int ngrid, double eta, const double4 *__restrict__ sph,
    const double4 *__restrict__ grid, const int *__restrict__ inl,
    const int *__restrict__ pairoff, const int *__restrict__ tp, const double *__restrict__ fi,
    double4 *__restrict__ tgeo, double2 *__restrict__ ttw, float4 *__restrict__ g32,
    float *__restrict__ w32) {
  const int c = blockIdx.x * (kSetupThreads / kWave) + threadIdx.x / kWave;
  if (c >= n) return;
  const double4 sc = sph[c];
  const int end = pairoff[inl[c + 1]];
  for (int k = pairoff[inl[c]] + lane_id(); k < end; k += kWave) {
    const unsigned p = (unsigned)tp[k];
    const unsigned o = p / (unsigned)ngrid, pn = p - o * (unsigned)ngrid;
    double4 a;
    double2 b;
    triple_record(sph[o], grid[pn], sc, fi[p], eta, a, b);
    tgeo[k] = a;
    ttw[k] = b;
    if (g32) {
      g32[k] = geo32_of(a, b);
      w32[k] = (float)b.y;
    }
  }
}

// Work units of the sweeps: runs of kWave triples of one container.
// A launch of the sweep hands contiguous ranges of units to a fixed number
// of workers; a segment is the part of one worker's range inside one
// container, and each segment leaves one partial L* S vector.
__global__ void k_unit_count(int n, const int *__restrict__ inl, const int *__restrict__ pairoff,
                             int *__restrict__ ucnt) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int t = pairoff[inl[c + 1]] - pairoff[inl[c]];
  ucnt[c] = (t + kWave - 1) / kWave;
}

// segments per container: its first unit starts one, and so does every
// worker whose range begins inside the container (worker w starts at unit
// floor(w U / G))
__global__ void k_seg_count(int n, const int *__restrict__ uoff, long long U, long long G,
                            int *__restrict__ segcnt) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const long long a = uoff[c], b = uoff[c + 1];
  int cnt = 0;
  if (b > a) {
    const long long hi = (b * G + U - 1) / U, lo = ((a + 1) * G + U - 1) / U;
    cnt = 1 + (int)(hi > lo ? hi - lo : 0);
  }
  segcnt[c] = cnt;
}

// the largest neighbour count into *out (zeroed before): a block maximum,
// then one atomic maximum per workgroup
__global__ __launch_bounds__(256) void k_max_degree(int n, const int *__restrict__ deg,
                                                    int *__restrict__ out) {
  __shared__ int wm[256 / kWave];
  const int i = blockIdx.x * 256 + threadIdx.x;
  int m = wave_max_int(i < n ? deg[i] : 0);
  if (lane_id() == 0) wm[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 256 / kWave; ++w) m = max(m, wm[w]);
    atomicMax(out, max(m, wm[0]));
  }
}

// cavity offset of each instance: the cavity index of its first point
__global__ void k_cav_offsets(int ninst, int ngrid, const int *__restrict__ soff,
                              const int *__restrict__ cavidx, int *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q <= ninst) out[q] = cavidx[(size_t)soff[q] * ngrid];
}

}  // namespace

int run_setup(ddx_handle *h) {
  hipStream_t s = h->stream;
  int n = h->n, ng = h->ngrid;
  size_t np = (size_t)n * ng;
  h->np = np;
  drop_round_graphs(h);   // captured solver rounds hold the old buffers
  if (np > 0x7fffffffull) {
    // point indices are 32-bit in the per-point kernels
    set_error("nsph * ngrid = %zu points exceed the 32-bit point index", np);
    return DDX_ERR_CAPACITY;
  }

  // neighbour list on the spatial hash
  unsigned T = 64;
  while (T < 2u * (unsigned)n) T <<= 1;
  const double inv_h = 1.0 / (2.0 * h->rmax * (1.0 + 1e-12));
  const int *inst = h->inst.p;
  DDX_TRY(h->degcnt.reserve(n + 1));
  DDX_TRY(h->inl.reserve(n + 1));
  DDX_TRY(h->ckey.reserve(n));
  DDX_TRY(h->bcnt.reserve(T + 1));
  DDX_TRY(h->boff.reserve(T + 1));
  DDX_TRY(h->bsph.reserve(n));
  DDX_TRY(h->cand.reserve((size_t)n * kMaxDegree));
  DDX_HIP(hipMemsetAsync(h->bcnt.p, 0, (T + 1) * sizeof(int), s));
  hipLaunchKernelGGL(k_cell_count, dim3((n + 255) / 256), dim3(256), 0, s, h->sph.p, inst, n,
                     inv_h, T - 1, h->ckey.p, h->bcnt.p);
  DDX_HIP(hipGetLastError());
  DDX_TRY(scan_exclusive(h->bcnt.p, h->boff.p, T, h->scan_tmp, s));
  hipLaunchKernelGGL(k_cell_fill, dim3((n + 255) / 256), dim3(256), 0, s, n, h->ckey.p,
                     h->boff.p, h->bcnt.p, h->bsph.p);
  DDX_HIP(hipGetLastError());
  {
    const int W = kSetupThreads / kWave;
    hipLaunchKernelGGL(k_nbr_find, dim3((n + W - 1) / W), dim3(kSetupThreads), 0, s, h->sph.p,
                       inst, n, inv_h, T - 1, h->boff.p, h->bsph.p, h->degcnt.p, h->cand.p);
    DDX_HIP(hipGetLastError());
  }
  DDX_TRY(scan_exclusive(h->degcnt.p, h->inl.p, n, h->scan_tmp, s));
  // the largest degree into degcnt[n] (the scan reads the first n)
  DDX_HIP(hipMemsetAsync(h->degcnt.p + n, 0, sizeof(int), s));
  if (n > 0) {
    hipLaunchKernelGGL(k_max_degree, dim3((n + 255) / 256), dim3(256), 0, s, n, h->degcnt.p,
                       h->degcnt.p + n);
    DDX_HIP(hipGetLastError());
  }
  int nnl = 0, maxdeg = 0;
  DDX_HIP(hipMemcpyAsync(&maxdeg, h->degcnt.p + n, sizeof(int), hipMemcpyDeviceToHost, s));
  DDX_HIP(hipMemcpyAsync(&nnl, h->inl.p + n, sizeof(int), hipMemcpyDeviceToHost, s));
  DDX_HIP(hipStreamSynchronize(s));
  h->nnl = nnl;
  h->maxdeg = maxdeg;
  if (maxdeg > kMaxDegree) {
    set_error("a sphere overlaps %d others; at most %d are supported", maxdeg, kMaxDegree);
    return DDX_ERR_CAPACITY;
  }
  DDX_TRY(h->nl.reserve(nnl + 1));
  DDX_TRY(h->rev.reserve(nnl + 1));
  DDX_TRY(h->paircnt.reserve(nnl + 1));
  DDX_TRY(h->pairoff.reserve(nnl + 1));
  DDX_TRY(h->prow.reserve(nnl + 1));
  if (nnl > 0) {
    const size_t nt = (size_t)n * kMaxDegree;
    hipLaunchKernelGGL(k_nbr_compact, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, n,
                       h->inl.p, h->cand.p, h->nl.p, h->prow.p);
    DDX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_nbr_rev, dim3((nnl + 255) / 256), dim3(256), 0, s, nnl, h->inl.p,
                       h->nl.p, h->prow.p, h->rev.p);
    DDX_HIP(hipGetLastError());
  }

  // per-point quantities
  DDX_TRY(h->fi.reserve(np));
  DDX_TRY(h->ui.reserve(np));
  if (h->igrad) DDX_TRY(h->zi.reserve(3 * np));
  DDX_TRY(h->tmask.reserve(np));
  DDX_TRY(h->ptcnt.reserve(np + 1));
  DDX_TRY(h->ptoff.reserve(np + 1));
  DDX_TRY(h->cavflag.reserve(np + 1));
  DDX_TRY(h->cavidx.reserve(np + 1));
  const dim3 pg((unsigned)((np + kSetupThreads - 1) / kSetupThreads));
  hipLaunchKernelGGL(k_points, dim3((n + kSetupThreads / kWave - 1) / (kSetupThreads / kWave)),
                     dim3(kSetupThreads), 0, s, h->sph.p, h->grid.p, ng, h->inl.p, h->nl.p, h->eta,
                     h->igrad, h->fi.p, h->ui.p, h->igrad ? h->zi.p : nullptr, h->tmask.p,
                     h->ptcnt.p, h->cavflag.p, n);
  DDX_HIP(hipGetLastError());
  DDX_TRY(scan_exclusive(h->ptcnt.p, h->ptoff.p, np, h->scan_tmp, s));
  DDX_TRY(scan_exclusive(h->cavflag.p, h->cavidx.p, np, h->scan_tmp, s));
  int ntri = 0, ncav = 0;
  DDX_HIP(hipMemcpyAsync(&ntri, h->ptoff.p + np, sizeof(int), hipMemcpyDeviceToHost, s));
  DDX_HIP(hipMemcpyAsync(&ncav, h->cavidx.p + np, sizeof(int), hipMemcpyDeviceToHost, s));

  // triples per neighbour pair, grouped by container
  if (nnl > 0) {
    const int wpb = kSetupThreads / kWave;
    hipLaunchKernelGGL(k_pair_count, dim3((n + wpb - 1) / wpb), dim3(kSetupThreads), 0, s, n, ng,
                       h->inl.p, h->nl.p, h->rev.p, h->tmask.p, h->paircnt.p);
    DDX_HIP(hipGetLastError());
    DDX_TRY(scan_exclusive(h->paircnt.p, h->pairoff.p, nnl, h->scan_tmp, s));
  } else {
    DDX_HIP(hipMemsetAsync(h->pairoff.p, 0, sizeof(int), s));
  }
  DDX_TRY(h->ucnt.reserve(n + 1));
  DDX_TRY(h->uoff.reserve(n + 1));
  hipLaunchKernelGGL(k_unit_count, dim3((n + 255) / 256), dim3(256), 0, s, n, h->inl.p,
                     h->pairoff.p, h->ucnt.p);
  DDX_HIP(hipGetLastError());
  DDX_TRY(scan_exclusive(h->ucnt.p, h->uoff.p, n, h->scan_tmp, s));
  int nunit = 0;
  DDX_HIP(hipMemcpyAsync(&nunit, h->uoff.p + n, sizeof(int), hipMemcpyDeviceToHost, s));
  // the cavity offset of each instance, read back with the counts
  if (h->ninst > 1) {
    if (h->maps_n != n) {
      DDX_TRY(h->soff.reserve(h->ninst + 1));
      DDX_HIP(hipMemcpyAsync(h->soff.p, h->h_soff.data(), (h->ninst + 1) * sizeof(int),
                             hipMemcpyHostToDevice, s));
    }
    DDX_TRY(h->cavoff.reserve((size_t)h->ninst + 1));
    hipLaunchKernelGGL(k_cav_offsets, dim3((h->ninst + 256) / 256), dim3(256), 0, s, h->ninst,
                       ng, h->soff.p, h->cavidx.p, h->cavoff.p);
    DDX_HIP(hipGetLastError());
    h->h_cavoff.resize(h->ninst + 1);
    DDX_HIP(hipMemcpyAsync(h->h_cavoff.data(), h->cavoff.p, (h->ninst + 1) * sizeof(int),
                           hipMemcpyDeviceToHost, s));
  }
  DDX_HIP(hipStreamSynchronize(s));
  h->ntri = ntri;
  h->ncav = ncav;
  h->nunit = nunit;
  // workers: whole multiples of the R = 8 per CU one-wave workgroups that
  // the FP64 sweeps keep resident (two waves per SIMD; the lmax 6 FP32 sweep
  // holds four), as many as give about
  // kUnitsPerWorker units each (or as many per compute unit as ddx_autotune
  // chose).  Short ranges keep waves queued behind the resident ones, which
  // hides the memory latency of their triples; a partial last round of
  // workgroups would leave most CUs idle at the end.
  long long G = (long long)h->workers_per_cu * h->ncu;
  if (!h->workers_fixed) {
    const double rounds = (double)nunit / ((double)kUnitsPerWorker * (double)G);
    G *= std::max(1LL, std::llround(rounds));
  }
  if (G > nunit) G = nunit;
  h->nworker = (int)G;
  DDX_TRY(h->segcnt.reserve(n + 1));
  DDX_TRY(h->segoff.reserve(n + 1));
  // (a container with units implies U > 0 and G > 0)
  hipLaunchKernelGGL(k_seg_count, dim3((n + 255) / 256), dim3(256), 0, s, n, h->uoff.p,
                     (long long)nunit, G, h->segcnt.p);
  DDX_HIP(hipGetLastError());
  DDX_TRY(scan_exclusive(h->segcnt.p, h->segoff.p, n, h->scan_tmp, s));
  h->nseg_cap = (int)G + n;   // segments: at most one per container plus one per worker

  h->have_geo32 = false;   // the float copies are remade on their next use
  h->fmm.lists_ok = false;  // as are the near lists of the far-field sums
  DDX_TRY(h->tgeo.reserve(ntri + 1));
  DDX_TRY(h->ttw.reserve(ntri + 1));
  DDX_TRY(h->tp.reserve(ntri + 1));
  DDX_TRY(h->pttri.reserve(ntri + 1));
  DDX_TRY(h->tslot.reserve(ntri + 1));
  if (nnl > 0) {
    const int wpb = kSetupThreads / kWave;
    const dim3 grid_o((n + wpb - 1) / wpb);
    hipLaunchKernelGGL(k_pair_fill, grid_o, dim3(kSetupThreads), 0, s, n, ng, h->inl.p, h->nl.p,
                       h->rev.p, h->tmask.p, h->pairoff.p, h->ptoff.p, h->tp.p, h->pttri.p,
                       h->tslot.p);
    DDX_HIP(hipGetLastError());
    // the float copies come with the records when the handle runs the mixed
    // precision, else on the first FP32 sweep
    const bool f32 = h->prec == DDX_PREC_MIXED;
    if (f32) {
      DDX_TRY(h->tgeo32.reserve(ntri + 1));
      DDX_TRY(h->ttw32.reserve(ntri + 1));
    }
    hipLaunchKernelGGL(k_triple_geo, grid_o, dim3(kSetupThreads), 0, s, n, ng, h->eta, h->sph.p,
                       h->grid.p, h->inl.p, h->pairoff.p, h->tp.p, h->fi.p, h->tgeo.p,
                       h->ttw.p, f32 ? h->tgeo32.p : nullptr, f32 ? h->ttw32.p : nullptr);
    DDX_HIP(hipGetLastError());
    h->have_geo32 = f32;
  }
  DDX_TRY(h->ccav.reserve(3 * (size_t)ncav + 3));
  hipLaunchKernelGGL(k_ccav, pg, dim3(kSetupThreads), 0, s, h->sph.p, h->grid.p, ng,
                     h->cavflag.p, h->cavidx.p, h->ccav.p, np);
  DDX_HIP(hipGetLastError());
  // the solver's block maps (once per handle: they depend on the sphere
  // counts only; a batch's instance offsets went up with the counts above),
  // Coulomb tiles
  if (h->maps_n != n) {
    if (h->ninst == 1) {
      DDX_TRY(h->soff.reserve(h->ninst + 1));
      DDX_HIP(hipMemcpyAsync(h->soff.p, h->h_soff.data(), (h->ninst + 1) * sizeof(int),
                             hipMemcpyHostToDevice, s));
    }
    DDX_TRY(build_solver_maps(h));
    h->maps_n = n;
  }
  if (h->ninst == 1) h->h_cavoff.assign({0, ncav});
  DDX_TRY(build_coulomb_tiles(h));
  h->have_setup = true;
  h->stats.ntriples = ntri;
  h->stats.nnl = nnl;
  h->stats.ncav = ncav;
  h->stats.max_degree = maxdeg;
  return DDX_OK;
}

}  // namespace ddx
