// All-pairs Coulomb sums for point-charge solutes.
//   potential  phi(j) = sum_i q_i / |r_j - r_i|            (mkrhs.f90:102-114)
//   field      E(j)   = sum_i q_i (r_j - r_i) / |r_j - r_i|^3  (efld.f90:98-114)
// Each thread owns one or two targets; sources stream through LDS.  Work
// comes as a table of tiles (targets tb..te-1, at most 512 for the potential
// and 64 for the tiled field, and sources sb..se-1): for a batch of molecules
// every tile stays inside one molecule, so each sums only its own molecule's
// charges.
// The potential kernel can return the field of the same sources as well
// (the solute field at the cavity points that the forces contract with
// zeta), which costs one all-pairs pass less than a separate field kernel.
// 1/r comes from the hardware reciprocal square root refined by one
// third-order step (relative error a few ulp), which replaces an IEEE sqrt
// and divide per pair.
// When there are few targets (the field of the cavity charges at the nuclei)
// the sources of a target tile are split into chunks, one tile each, and
// the partial fields are summed afterwards in chunk order, so results are
// deterministic.
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

#include <algorithm>
#include <mutex>

namespace ddx {
namespace {

constexpr int kCT = 256;          // threads per workgroup, sources per LDS tile
constexpr int kTW = 2 * kCT;      // targets per tile: two per thread
constexpr int kFW = kWave;        // targets per tile of the tiled field: one per lane

// 1/sqrt(x) for x > 0: v_rsq_f64 (relative error below 2^-24 on gfx950)
// and one third-order correction: with e = 1 - x y^2, y (1 + e/2 + 3e^2/8)
// leaves a relative error of order e^3 plus rounding (about 1 ulp)
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y, e * fma(0.375, e, 0.5), y);
}

// phi(j) and, with FIELD, the field ef(:, j) of the same sources, from one
// evaluation of 1/r per pair; tile (tb, te, sb, se) per workgroup, targets
// tb + t and tb + t + kCT in thread t (each source read from LDS serves two)
template <bool FIELD>
__global__ __launch_bounds__(kCT) void k_potential(const int4 *__restrict__ tiles,
                                                   const double *__restrict__ q,
                                                   const double4 *__restrict__ src,
                                                   const double *__restrict__ trg,
                                                   double *__restrict__ phi,
                                                   double *__restrict__ ef) {
  __shared__ double4 tile[kCT];
  const int4 tl = tiles[blockIdx.x];
  const int j0 = tl.x + threadIdx.x, j1 = j0 + kCT;
  double ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
  if (j0 < tl.y) {
    ax = trg[3 * (size_t)j0];
    ay = trg[3 * (size_t)j0 + 1];
    az = trg[3 * (size_t)j0 + 2];
  }
  if (j1 < tl.y) {
    bx = trg[3 * (size_t)j1];
    by = trg[3 * (size_t)j1 + 1];
    bz = trg[3 * (size_t)j1 + 2];
  }
  double va = 0.0, ea0 = 0.0, ea1 = 0.0, ea2 = 0.0;
  double vb = 0.0, eb0 = 0.0, eb1 = 0.0, eb2 = 0.0;
  for (int i0 = tl.z; i0 < tl.w; i0 += kCT) {
    int i = i0 + threadIdx.x;
    if (i < tl.w) {
      double4 s = src[i];
      tile[threadIdx.x] = make_double4(s.x, s.y, s.z, q[i]);
    }
    __syncthreads();
    int cnt = min(kCT, tl.w - i0);
    for (int k = 0; k < cnt; ++k) {
      const double4 s = tile[k];
      {
        const double dx = ax - s.x, dy = ay - s.y, dz = az - s.z;
        const double y = rsqrt_nr(dx * dx + dy * dy + dz * dz);
        const double qy = s.w * y;
        va = va + qy;
        if (FIELD) {
          const double f = qy * (y * y);
          ea0 = ea0 + f * dx;
          ea1 = ea1 + f * dy;
          ea2 = ea2 + f * dz;
        }
      }
      {
        const double dx = bx - s.x, dy = by - s.y, dz = bz - s.z;
        const double y = rsqrt_nr(dx * dx + dy * dy + dz * dz);
        const double qy = s.w * y;
        vb = vb + qy;
        if (FIELD) {
          const double f = qy * (y * y);
          eb0 = eb0 + f * dx;
          eb1 = eb1 + f * dy;
          eb2 = eb2 + f * dz;
        }
      }
    }
    __syncthreads();
  }
  if (j0 < tl.y) {
    phi[j0] = va;
    if (FIELD) {
      ef[3 * (size_t)j0] = ea0;
      ef[3 * (size_t)j0 + 1] = ea1;
      ef[3 * (size_t)j0 + 2] = ea2;
    }
  }
  if (j1 < tl.y) {
    phi[j1] = vb;
    if (FIELD) {
      ef[3 * (size_t)j1] = eb0;
      ef[3 * (size_t)j1 + 1] = eb1;
      ef[3 * (size_t)j1 + 2] = eb2;
    }
  }
}

// fields of the sources sb..se-1 of tile tl at its targets tb..te-1 (two per
// thread, targets tb + t and tb + t + kCT); the sums of thread t come back in
// e[0..2] (first target) and e[3..5] (second), zero for targets past te
__device__ __forceinline__ void field_tile(const int4 tl, const double *__restrict__ q,
                                           const double *__restrict__ src,
                                           const double *__restrict__ trg, double4 *tile,
                                           double e[6]) {
  const int j0 = tl.x + threadIdx.x, j1 = j0 + kCT;
  double ax = 0.0, ay = 0.0, az = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
  if (j0 < tl.y) {
    ax = trg[3 * (size_t)j0];
    ay = trg[3 * (size_t)j0 + 1];
    az = trg[3 * (size_t)j0 + 2];
  }
  if (j1 < tl.y) {
    bx = trg[3 * (size_t)j1];
    by = trg[3 * (size_t)j1 + 1];
    bz = trg[3 * (size_t)j1 + 2];
  }
  double ea0 = 0.0, ea1 = 0.0, ea2 = 0.0, eb0 = 0.0, eb1 = 0.0, eb2 = 0.0;
  for (int i0 = tl.z; i0 < tl.w; i0 += kCT) {
    int i = i0 + threadIdx.x;
    if (i < tl.w)
      tile[threadIdx.x] = make_double4(src[3 * (size_t)i], src[3 * (size_t)i + 1],
                                       src[3 * (size_t)i + 2], q[i]);
    __syncthreads();
    int cnt = min(kCT, tl.w - i0);
    for (int k = 0; k < cnt; ++k) {
      const double4 s = tile[k];
      {
        const double dx = ax - s.x, dy = ay - s.y, dz = az - s.z;
        const double y = rsqrt_nr(dx * dx + dy * dy + dz * dz);
        const double f = s.w * (y * y * y);
        ea0 = ea0 + f * dx;
        ea1 = ea1 + f * dy;
        ea2 = ea2 + f * dz;
      }
      {
        const double dx = bx - s.x, dy = by - s.y, dz = bz - s.z;
        const double y = rsqrt_nr(dx * dx + dy * dy + dz * dz);
        const double f = s.w * (y * y * y);
        eb0 = eb0 + f * dx;
        eb1 = eb1 + f * dy;
        eb2 = eb2 + f * dz;
      }
    }
    __syncthreads();
  }
  if (j0 >= tl.y) ea0 = ea1 = ea2 = 0.0;
  if (j1 >= tl.y) eb0 = eb1 = eb2 = 0.0;
  e[0] = ea0;
  e[1] = ea1;
  e[2] = ea2;
  e[3] = eb0;
  e[4] = eb1;
  e[5] = eb2;
}

// ef of the targets of a target tile = sum of the partials of its chunk
// tiles b0 .. b1-1 (kTW targets each), in a fixed order: each
// target's chunks are split over kRq threads (chunk c to thread c mod kRq),
// whose sums meet in LDS in thread order.  blockIdx.y picks the slice of
// kCT / kRq targets.
constexpr int kRq = 8;
__device__ __forceinline__ void field_reduce(int b0, int b1, int tb, int te,
                                             const double *__restrict__ part,
                                             double *__restrict__ ef, double (*red)[3]) {
  const int r = threadIdx.x % kRq;
  const int t = blockIdx.y * (kCT / kRq) + threadIdx.x / kRq;
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int b = b0 + r; b < b1; b += kRq) {
    const size_t o = ((size_t)b * kTW + t) * 3;
    e0 += part[o];
    e1 += part[o + 1];
    e2 += part[o + 2];
  }
  red[threadIdx.x][0] = e0;
  red[threadIdx.x][1] = e1;
  red[threadIdx.x][2] = e2;
  __syncthreads();
  const int j = tb + t;
  if (r == 0 && j < te) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < kRq; ++k) {
      s0 += red[threadIdx.x + k][0];
      s1 += red[threadIdx.x + k][1];
      s2 += red[threadIdx.x + k][2];
    }
    ef[3 * (size_t)j] = s0;
    ef[3 * (size_t)j + 1] = s1;
    ef[3 * (size_t)j + 2] = s2;
  }
}

// The tiled field (few targets per molecule: the nuclei): one wave per tile
// of at most kFW targets, one per lane, the tile's sources through LDS kFW at
// a time; part[(b kFW + t) 3] holds tile b's partial field at target t.
__global__ __launch_bounds__(kFW) void k_field_w(const int4 *__restrict__ tiles,
                                                 const double *__restrict__ q,
                                                 const double *__restrict__ src,
                                                 const double *__restrict__ trg,
                                                 double *__restrict__ part) {
  __shared__ double4 tile[kFW];
  const int4 tl = tiles[blockIdx.x];
  const int j = tl.x + (int)threadIdx.x;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (j < tl.y) {
    ax = trg[3 * (size_t)j];
    ay = trg[3 * (size_t)j + 1];
    az = trg[3 * (size_t)j + 2];
  }
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int i0 = tl.z; i0 < tl.w; i0 += kFW) {
    const int i = i0 + (int)threadIdx.x;
    __syncthreads();   // the previous sources are read
    if (i < tl.w)
      tile[threadIdx.x] = make_double4(src[3 * (size_t)i], src[3 * (size_t)i + 1],
                                       src[3 * (size_t)i + 2], q[i]);
    __syncthreads();
    const int cnt = min(kFW, tl.w - i0);
    for (int k = 0; k < cnt; ++k) {
      const double4 sk = tile[k];
      const double dx = ax - sk.x, dy = ay - sk.y, dz = az - sk.z;
      const double y = rsqrt_nr(dx * dx + dy * dy + dz * dz);
      const double f = sk.w * (y * y * y);
      e0 = e0 + f * dx;
      e1 = e1 + f * dy;
      e2 = e2 + f * dz;
    }
  }
  if (j >= tl.y) e0 = e1 = e2 = 0.0;
  const size_t o = ((size_t)blockIdx.x * kFW + threadIdx.x) * 3;
  part[o] = e0;
  part[o + 1] = e1;
  part[o + 2] = e2;
}

// ef of the targets of group g = the partials of its tiles grp[g] ..
// grp[g+1]-1 (the source chunks of one target tile) in a fixed order: each
// target's chunks are split over kRw threads (chunk c to thread c mod kRw),
// whose sums meet in LDS in thread order.
constexpr int kRw = 4;
__global__ __launch_bounds__(kFW * kRw) void k_field_reduce_w(const int *__restrict__ grp,
                                                              const int4 *__restrict__ tiles,
                                                              const double *__restrict__ part,
                                                              double *__restrict__ ef) {
  __shared__ double red[kFW * kRw][3];
  const int b0 = grp[blockIdx.x], b1 = grp[blockIdx.x + 1];
  const int4 tl = tiles[b0];
  const int r = threadIdx.x % kRw, t = threadIdx.x / kRw;
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int b = b0 + r; b < b1; b += kRw) {
    const size_t o = ((size_t)b * kFW + t) * 3;
    e0 += part[o];
    e1 += part[o + 1];
    e2 += part[o + 2];
  }
  red[threadIdx.x][0] = e0;
  red[threadIdx.x][1] = e1;
  red[threadIdx.x][2] = e2;
  __syncthreads();
  const int j = tl.x + t;
  if (r == 0 && j < tl.y) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < kRw; ++k) {
      s0 += red[threadIdx.x + k][0];
      s1 += red[threadIdx.x + k][1];
      s2 += red[threadIdx.x + k][2];
    }
    ef[3 * (size_t)j] = s0;
    ef[3 * (size_t)j + 1] = s1;
    ef[3 * (size_t)j + 2] = s2;
  }
}

// The same field on a uniform tiling that needs no table: tile b covers
// target tile b / nchunk (kTW targets) and source chunk b % nchunk (`chunk`
// sources).  With one chunk the sums go straight to ef; with more, to
// part[(b kTW + t) 3], which k_field_reduce_uniform sums in chunk order.
__device__ __forceinline__ int4 uniform_tile(int b, int nchunk, int chunk, int ntrg, int nsrc) {
  const int tt = b / nchunk, c = b - tt * nchunk;
  const int tb = tt * kTW, sb = (int)min((long long)c * chunk, (long long)nsrc);
  return make_int4(tb, min(tb + kTW, ntrg), sb, (int)min((long long)sb + chunk, (long long)nsrc));
}

__global__ __launch_bounds__(kCT) void k_field_uniform(int ntrg, int nsrc, int nchunk, int chunk,
                                                       const double *__restrict__ q,
                                                       const double *__restrict__ src,
                                                       const double *__restrict__ trg,
                                                       double *__restrict__ part,
                                                       double *__restrict__ ef) {
  __shared__ double4 tile[kCT];
  double e[6];
  const int4 tl = uniform_tile(blockIdx.x, nchunk, chunk, ntrg, nsrc);
  field_tile(tl, q, src, trg, tile, e);
  if (nchunk == 1) {
    const int j0 = tl.x + threadIdx.x, j1 = j0 + kCT;
    if (j0 < tl.y) {
      ef[3 * (size_t)j0] = e[0];
      ef[3 * (size_t)j0 + 1] = e[1];
      ef[3 * (size_t)j0 + 2] = e[2];
    }
    if (j1 < tl.y) {
      ef[3 * (size_t)j1] = e[3];
      ef[3 * (size_t)j1 + 1] = e[4];
      ef[3 * (size_t)j1 + 2] = e[5];
    }
    return;
  }
  const size_t o = ((size_t)blockIdx.x * kTW + threadIdx.x) * 3, o1 = o + 3 * (size_t)kCT;
  part[o] = e[0];
  part[o + 1] = e[1];
  part[o + 2] = e[2];
  part[o1] = e[3];
  part[o1 + 1] = e[4];
  part[o1 + 2] = e[5];
}

__global__ __launch_bounds__(kCT) void k_field_reduce_uniform(int ntrg, int nchunk,
                                                              const double *__restrict__ part,
                                                              double *__restrict__ ef) {
  __shared__ double red[kCT][3];
  const int tb = blockIdx.x * kTW;
  field_reduce(blockIdx.x * nchunk, (blockIdx.x + 1) * nchunk, tb, min(tb + kTW, ntrg), part,
               ef, red);
}

__global__ void k_psi(int n, int nylm, double fac, const double *__restrict__ q,
                      double *__restrict__ psi) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n * nylm) return;
  psi[k] = (k % nylm == 0) ? fac * q[k / nylm] : 0.0;
}

}  // namespace

// Tiles of kTW targets, per instance q: targets toff[q] .. toff[q+1]-1,
// sources soff[q] .. soff[q+1]-1.  With `split`, the sources of each target
// tile are cut into chunks (multiples of kCT, at least 4 kCT) so that there
// are about 1024 tiles in all; grp then holds the first tile of each target
// tile's chunks.
void coulomb_tiles(const std::vector<int> &toff, const std::vector<int> &soff, bool split,
                   std::vector<int4> &tiles, std::vector<int> &grp) {
  // split: the tiled field, kFW targets per tile (one wave) and the sources
  // cut into chunks; otherwise kTW targets and all of the molecule's sources
  const int tw = split ? kFW : kTW;
  const size_t ninst = toff.size() - 1;
  long long work = 0;   // target tiles x sources
  for (size_t q = 0; q < ninst; ++q)
    work += (long long)((toff[q + 1] - toff[q] + tw - 1) / tw) * (soff[q + 1] - soff[q]);
  long long chunk = 0x7fffffff;
  if (split) {
    // about 32768 one-wave tiles: several rounds of the resident waves, so
    // that the uneven last chunks of each target tile do not leave CUs idle
    chunk = (work + 32767) / 32768;
    chunk = ((chunk + kFW - 1) / kFW) * kFW;
    if (chunk < 4 * kFW) chunk = 4 * kFW;
  }
  tiles.clear();
  grp.clear();
  for (size_t q = 0; q < ninst; ++q) {
    const int sb = soff[q], se = soff[q + 1];
    for (int tb = toff[q]; tb < toff[q + 1]; tb += tw) {
      const int te = std::min(tb + tw, toff[q + 1]);
      grp.push_back((int)tiles.size());
      if (se <= sb) {
        tiles.push_back(make_int4(tb, te, sb, sb));
        continue;
      }
      for (long long c = sb; c < se; c += chunk)
        tiles.push_back(make_int4(tb, te, (int)c, (int)std::min<long long>(c + chunk, se)));
    }
  }
  grp.push_back((int)tiles.size());
}

namespace {
int upload_tiles(ddx_handle *h, DevArray<int4> &dt, const std::vector<int4> &t) {
  DDX_TRY(dt.reserve(t.size() + 1));
  if (!t.empty())
    DDX_HIP(hipMemcpyAsync(dt.p, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice,
                           h->stream));
  return DDX_OK;
}
}  // namespace

int build_coulomb_tiles(ddx_handle *h) {
  // the host tables live in the handle until the next setup, so the copies
  // need no wait here
  std::vector<int> g;
  // potential (and solute field) at the cavity points from the spheres
  coulomb_tiles(h->h_cavoff, h->h_soff, false, h->h_ctile, g);
  h->nctile = (int)h->h_ctile.size();
  DDX_TRY(upload_tiles(h, h->ctile, h->h_ctile));
  // field of the cavity charges at the nuclei
  coulomb_tiles(h->h_soff, h->h_cavoff, true, h->h_ftile, h->h_fgrp);
  h->nftile = (int)h->h_ftile.size();
  h->nfgrp = (int)h->h_fgrp.size() - 1;
  DDX_TRY(upload_tiles(h, h->ftile, h->h_ftile));
  DDX_TRY(h->fgrp.reserve(h->h_fgrp.size()));
  DDX_HIP(hipMemcpyAsync(h->fgrp.p, h->h_fgrp.data(), h->h_fgrp.size() * sizeof(int),
                         hipMemcpyHostToDevice, h->stream));
  return DDX_OK;
}

int coulomb_potential(ddx_handle *h, const int4 *tiles, int ntile, const double *q,
                      const double4 *src, const double *trg3, double *phi, double *ef) {
  if (ntile <= 0) return DDX_OK;
  TimedLaunch tl(h, DDX_K_POTENTIAL, h->stream);
  if (ef)
    hipExtLaunchKernelGGL(k_potential<true>, dim3(ntile), dim3(kCT), 0, h->stream, tl.b, tl.e, 0,
                          tiles, q, src, trg3, phi, ef);
  else
    hipExtLaunchKernelGGL(k_potential<false>, dim3(ntile), dim3(kCT), 0, h->stream, tl.b, tl.e,
                          0, tiles, q, src, trg3, phi, nullptr);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int coulomb_field(ddx_handle *h, const int4 *tiles, int ntile, const int *grp, int ngrp,
                  const double *q, const double *src3, const double *trg3, double *ef) {
  if (ntile <= 0 || ngrp <= 0) return DDX_OK;
  DDX_TRY(h->fpart.reserve((size_t)ntile * kFW * 3));
  {
    TimedLaunch tl(h, DDX_K_FIELD, h->stream);
    hipExtLaunchKernelGGL(k_field_w, dim3(ntile), dim3(kFW), 0, h->stream, tl.b, tl.e, 0, tiles,
                          q, src3, trg3, h->fpart.p);
  }
  DDX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_field_reduce_w, dim3(ngrp), dim3(kFW * kRw), 0, h->stream, grp, tiles,
                     h->fpart.p, ef);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

// the field of nsrc arbitrary charges at ntrg arbitrary points (efld.f90:87-116)
int coulomb_field_points(ddx_handle *h, int nsrc, const double *q, const double *src3, int ntrg,
                         const double *trg3, double *ef) {
  if (ntrg <= 0) return DDX_OK;
  std::vector<int4> &t = h->h_xtile;
  std::vector<int> &g = h->h_xgrp;
  DDX_HIP(hipStreamSynchronize(h->stream));   // the previous tables' copies are done
  coulomb_tiles({0, ntrg}, {0, nsrc}, true, t, g);
  DDX_TRY(upload_tiles(h, h->xtile, t));
  DDX_TRY(h->xgrp.reserve(g.size()));
  DDX_HIP(hipMemcpyAsync(h->xgrp.p, g.data(), g.size() * sizeof(int), hipMemcpyHostToDevice,
                         h->stream));
  return coulomb_field(h, h->xtile.p, (int)t.size(), h->xgrp.p, (int)g.size() - 1, q, src3, trg3,
                       ef);
}

int make_psi_n(ddx_handle *h, int n, int nylm, const double *charge, double *psi) {
  size_t nn = (size_t)n * nylm;
  double fac = sqrt(4.0 * (4.0 * atan(1.0)));
  hipLaunchKernelGGL(k_psi, dim3((nn + 255) / 256), dim3(256), 0, h->stream, n, nylm, fac,
                     charge, psi);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int make_psi(ddx_handle *h, const double *charge, double *psi) {
  return make_psi_n(h, h->n, h->nylm, charge, psi);
}

// ---------------------------------------------------------------------------
// The field without allocation.  Its partial sums live in `scratch`, a
// file-scope device array of kScratchBytes on every GPU the library runs on
// (held until the process ends).  In host code the array's name is only a
// stand-in: its device address on the current GPU comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
static double *g_scratch[kMaxDevices];

int noalloc_setup(double **out) {
  int dev = 0;
  DDX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d beyond the %d devices the scratch table covers", dev, kMaxDevices);
    return DDX_ERR_ARG;
  }
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return DDX_ERR_HIP;
    }
    g_scratch[dev] = static_cast<double *>(p);
  }
  if (out) *out = g_scratch[dev];
  return DDX_OK;
}

}  // namespace ddx

extern "C" int ddx_efld_noalloc(int nsrc, const double *src, const double *csrc, int ntrg,
                                const double *ctrg, double *ef, void *stream) {
  using namespace ddx;
  if (nsrc < 0 || ntrg < 0 || (nsrc && (!src || !csrc)) || (ntrg && (!ctrg || !ef))) {
    set_error("ddx_efld_noalloc: invalid arguments");
    return DDX_ERR_ARG;
  }
  double *part = nullptr;
  DDX_TRY(noalloc_setup(&part));
  if (ntrg == 0) return DDX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // chunks as the table-driven path cuts them (about 8192 tiles, at least
  // 4 kCT sources each), then fewer where the partials would not fit
  const long long ntt = (ntrg + kTW - 1) / kTW;
  const long long max_tiles = (long long)(kScratchBytes / (sizeof(double) * 3 * kTW));
  long long chunk = (ntt * nsrc + 8191) / 8192;
  chunk = std::max<long long>(4 * kCT, ((chunk + kCT - 1) / kCT) * kCT);
  long long nchunk = std::max<long long>(1, (nsrc + chunk - 1) / chunk);
  if (ntt * nchunk > max_tiles) {
    nchunk = std::max<long long>(1, max_tiles / ntt);
    chunk = std::max<long long>(1, (nsrc + nchunk - 1) / nchunk);
    chunk = ((chunk + kCT - 1) / kCT) * kCT;
    nchunk = std::max<long long>(1, (nsrc + chunk - 1) / chunk);
  }
  const size_t need = nchunk > 1 ? (size_t)(ntt * nchunk) * kTW * 3 * sizeof(double) : 0;
  if (need > kScratchBytes || ntt * nchunk > 0x7fffffffll) {
    set_error("ddx_efld_noalloc: %lld partial tiles exceed the %zu-byte scratch",
              ntt * nchunk, kScratchBytes);
    return DDX_ERR_CAPACITY;
  }
  hipLaunchKernelGGL(k_field_uniform, dim3((unsigned)(ntt * nchunk)), dim3(kCT), 0, s, ntrg,
                     nsrc, (int)nchunk, (int)chunk, src, csrc, ctrg, part, ef);
  DDX_HIP(hipGetLastError());
  if (nchunk > 1) {
    hipLaunchKernelGGL(k_field_reduce_uniform, dim3((unsigned)ntt, kTW / (kCT / kRq)), dim3(kCT),
                       0, s, ntrg, (int)nchunk, part, ef);
    DDX_HIP(hipGetLastError());
  }
  return DDX_OK;
}
