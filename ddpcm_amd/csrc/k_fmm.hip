// Far-field Coulomb sums for large systems: the two all-pairs passes of the
// point-charge path (mkrhs.f90:102-114, efld.f90:98-114, main.f90:235-265),
// split into a near part summed directly and a far part through spherical-
// harmonic expansions about the sphere centres, to order P.
//
// The cavity points of sphere i lie on its surface, r_i + R_i s_n, and the
// charges sit at the sphere centres, so both passes expand about the
// centres.  A pair (sphere i, charge j) is far when d_ij >= R_i / rho.
//
//   potential and solute field at the cavity points of sphere i:
//     far   local expansion A_i (k_p2l in k_sweep.hip), evaluated with its
//           gradient at every exposed point by the Legendre rows of the
//           force kernels:  phi = sum_l S_l(s),  grad = (1/R) sum_l
//           (l S_l s + grad_s S_l),  S_l(s) = sum_m A_lm Y_lm(s)
//     near  direct sums over the charges of the sphere's near list
//   field of zeta at the nuclei:
//     far   multipole of each sphere's cavity charges, M_i,lm = sum_n zeta_n
//           Y_lm(s_n); the potential at u = x - r_i outside is (1/|u|)
//           sum_l t^l A_l(u^) with t = R_i/|u| and A_l the rows contracted
//           with cf M, whose gradient is (1/|u|^2) sum_l t^l (-(l+1) A_l u^
//           + grad_s A_l)
//     near  direct sums over the cavity points of the near spheres
// Truncated at order p, the expansion of a far pair errs by less than
// (charge / (d - R)) t^(p+1), t = R / d <= rho.  Each pass takes the order
// its error budget asks for.  The potential enters the energy (parity 1e-10
// relative) and takes P = 16: 1.3e-12 of each far term at t = 0.2.  The
// field of zeta at the nuclei enters only the forces (tolerance 1e-7
// Eh/bohr): its cap is P = 10 (2e-8 of a term at t = 0.2), and a pair needs
// only the smallest order with t^(p+1) below 2e-10; the nuclei are taken in
// Morton order of their centres, so that the 64 targets of a wave need
// similar orders, and the wave expands to the largest of them (most pairs
// need 5..7).  All sums run in a fixed order: results are deterministic.
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"
#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

namespace ddx {
namespace {

constexpr int kFT = 256;   // threads per workgroup

__host__ __device__ constexpr int lmi(int l, int m) { return l * l + l + m; }

// The Legendre rows of one direction contracted with the coefficients xs
// (lm order): for each l, f(l, A_l, Bt_l, Bp_l) with A_l = sum_m P_l^m
// u_c(m) and Bt_l, Bp_l the theta and phi components of the surface
// gradient of A_l (e_phi carrying 1/sin th), u_c = cos(m ph) xs_l,m +
// sin(m ph) xs_l,-m, u_s = cos(m ph) xs_l,-m - sin(m ph) xs_l,m.  The poles
// take the limits of dbasis (ddcosmo.f90:970-971).  Three rows of P_l^m in
// registers, l outer; rows past lcut (wave-uniform) are left out.
template <int P, typename F>
__device__ __forceinline__ void legendre_rows(double cthe, double sthe, double cphi, double sphi,
                                              const double *xs, F &&f, int lcut = P) {
  double cs[P + 1], sn[P + 1];
  cs[0] = 1.0;
  sn[0] = 0.0;
  cs[1] = cphi;
  sn[1] = sphi;
#pragma unroll
  for (int m = 2; m <= P; ++m) {
    cs[m] = 2.0 * cphi * cs[m - 1] - cs[m - 2];
    sn[m] = 2.0 * cphi * sn[m - 1] - sn[m - 2];
  }
  const bool pole = (sthe == 0.0);
  const double vs = (double)(int)cthe;
  // row l lives in rows[l % 2], computed in place over row l - 2
  double rows[2][P + 1];
#pragma unroll
  for (int l = 0; l <= P; ++l) {
    if (l > lcut) continue;   // (not a break: the loop stays unrolled)
    double *pc = rows[l & 1];
    const double *p1 = rows[(l + 1) & 1];
#pragma unroll
    for (int m = 0; m <= l; ++m) {
      if (l == 0) pc[0] = 1.0;
      else if (m == l) pc[m] = -(2.0 * l - 1.0) * sthe * p1[l - 1];
      else if (m == l - 1) pc[m] = cthe * (2.0 * m + 1.0) * p1[m];
      else pc[m] = (cthe * (2.0 * l - 1.0) * p1[m] - (double)(l + m - 1) * pc[m]) *
                   (1.0 / (double)(l - m));
    }
    double a = pc[0] * xs[lmi(l, 0)];
    double b = (l > 0) ? pc[1] * xs[lmi(l, 0)] : 0.0;
    double c = 0.0;
#pragma unroll
    for (int m = 1; m <= l; ++m) {
      const double xp = xs[lmi(l, m)], xm = xs[lmi(l, -m)];
      const double uc = cs[m] * xp + sn[m] * xm;
      const double us = cs[m] * xm - sn[m] * xp;
      const double dp = 0.5 * ((double)(l + m) * (double)(l - m + 1) * pc[m - 1]) +
                        ((m < l) ? -0.5 * pc[m < l ? m + 1 : m] : 0.0);
      a += pc[m] * uc;
      b -= dp * uc;
      c += (pole ? -dp * vs : (double)m * pc[m]) * us;
    }
    f(l, a, b, c);
    // keep the next row's LDS reads of xs after this row's use, and its
    // recurrence after this row's sums (scheduled ahead, the rows of all l
    // would be live at once and exceed the register file)
    asm volatile("" : "+v"(cthe), "+v"(sthe) : "v"(a), "v"(b), "v"(c) : "memory");
  }
}

// squared distance, evaluated the same way wherever a pair is classified
// near or far (no contraction, fixed order)
__device__ __forceinline__ double dist2(double ax, double ay, double az, double bx, double by,
                                        double bz) {
#pragma clang fp contract(off)
  const double dx = ax - bx, dy = ay - by, dz = az - bz;
  return (dx * dx + dy * dy) + dz * dz;
}

// direction angles of a vector of length r (r > 0)
__device__ __forceinline__ void angles(double x, double y, double z, double r, double &cthe,
                                       double &sthe, double &cphi, double &sphi) {
  const double rxy = sqrt(x * x + y * y);
  cthe = z / r;
  sthe = rxy / r;
  cphi = 1.0;
  sphi = 0.0;
  if (rxy > 0.0) {
    cphi = x / rxy;
    sphi = y / rxy;
  }
}

// unit vectors s, e_theta, e_phi of the force kernels' convention
__device__ __forceinline__ void frame(double cthe, double sthe, double cphi, double sphi,
                                      double *s, double *et, double *ep) {
  s[0] = sthe * cphi;
  s[1] = sthe * sphi;
  s[2] = cthe;
  et[0] = cthe * cphi;
  et[1] = cthe * sphi;
  et[2] = -sthe;
  ep[0] = 0.0;
  ep[1] = 1.0;
  ep[2] = 0.0;
  if (sthe != 0.0) {
    const double is = 1.0 / sthe;
    ep[0] = -sphi * is;
    ep[1] = cphi * is;
  }
}

// Near pairs (sphere i, charge j of its instance) with d_ij < R_i / rho:
// count per sphere (mode 0) or per charge (mode 1), one wave each.
__global__ __launch_bounds__(kFT) void k_near_count(int n, int mode, double rho,
                                                    const double4 *__restrict__ sph,
                                                    const int *__restrict__ inst,
                                                    const int *__restrict__ soff,
                                                    int *__restrict__ cnt) {
  const int w = blockIdx.x * (kFT / kWave) + threadIdx.x / kWave;
  if (w >= n) return;
  const double4 a = sph[w];
  int sb = 0, se = n;
  if (inst) {
    sb = soff[inst[w]];
    se = soff[inst[w] + 1];
  }
  int c = 0;
  for (int k = sb + lane_id(); k < se; k += kWave) {
    const double4 b = sph[k];
    const double rc = (mode == 0 ? a.w : b.w) / rho;
    c += (dist2(a.x, a.y, a.z, b.x, b.y, b.z) < rc * rc) ? 1 : 0;
  }
  c = (int)wave_sum((double)c);
  if (lane_id() == 0) cnt[w] = c;
}

// ... and the lists, in increasing index (ballot ranks within the wave)
__global__ __launch_bounds__(kFT) void k_near_fill(int n, int mode, double rho,
                                                   const double4 *__restrict__ sph,
                                                   const int *__restrict__ inst,
                                                   const int *__restrict__ soff,
                                                   const int *__restrict__ off,
                                                   int *__restrict__ idx) {
  const int w = blockIdx.x * (kFT / kWave) + threadIdx.x / kWave;
  if (w >= n) return;
  const double4 a = sph[w];
  int sb = 0, se = n;
  if (inst) {
    sb = soff[inst[w]];
    se = soff[inst[w] + 1];
  }
  int pos = off[w];
  const int lane = lane_id();
  for (int k0 = sb; k0 < se; k0 += kWave) {
    const int k = k0 + lane;
    bool hit = false;
    if (k < se) {
      const double4 b = sph[k];
      const double rc = (mode == 0 ? a.w : b.w) / rho;
      hit = dist2(a.x, a.y, a.z, b.x, b.y, b.z) < rc * rc;
    }
    const unsigned long long bal = __ballot(hit);
    if (hit) idx[pos + __popcll(bal & ((1ull << lane) - 1ull))] = k;
    pos += __popcll(bal);
  }
}

// (sphere, grid point) of every cavity point
__global__ void k_cav_map(size_t np, int ngrid, const int *__restrict__ cavidx,
                          int2 *__restrict__ cmap) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int c = cavidx[p];
  if (cavidx[p + 1] > c) cmap[c] = make_int2((int)(p / ngrid), (int)(p % ngrid));
}

// Far part of the potential (and with ef of the solute field) at the cavity
// points of sphere i = blockIdx.x: its local expansion (facs A in LDS) with
// its gradient, one wave taking the sphere's exposed points in turn.
template <int P>
__global__ __launch_bounds__(kWave) void k_l2p(int ngrid, const int *__restrict__ cavidx,
                                               const int2 *__restrict__ cmap,
                                               const double4 *__restrict__ sph,
                                               const double4 *__restrict__ grid,
                                               const double *__restrict__ A,
                                               const double *__restrict__ facsP,
                                               double *__restrict__ phi,
                                               double *__restrict__ ef) {
  constexpr int NY = (P + 1) * (P + 1);
  __shared__ double xs[NY];
  const int i = blockIdx.x;
  for (int k = threadIdx.x; k < NY; k += kWave) xs[k] = facsP[k] * A[(size_t)i * NY + k];
  __syncthreads();
  const int c0 = cavidx[(size_t)i * ngrid], c1 = cavidx[(size_t)(i + 1) * ngrid];
  const double ir = 1.0 / sph[i].w;
  for (int base = c0; base < c1; base += kWave) {
    // the barrier keeps the LDS reads of xs inside each point's Legendre
    // rows (hoisted out of the loop, all (P+1)^2 values would need registers)
    __syncthreads();
    const int ic = base + (int)threadIdx.x;
    if (ic >= c1) continue;
    const double4 g = grid[cmap[ic].y];
    double cthe, sthe, cphi, sphi;
    angles(g.x, g.y, g.z, 1.0, cthe, sthe, cphi, sphi);
    double v = 0.0, gr = 0.0, gt = 0.0, gp = 0.0;
    legendre_rows<P>(cthe, sthe, cphi, sphi, xs, [&](int l, double a, double b, double c) {
      v += a;
      gr += (double)l * a;
      gt += b;
      gp += c;
    });
    phi[ic] = v;
    if (ef) {
      double s[3], et[3], ep[3];
      frame(cthe, sthe, cphi, sphi, s, et, ep);
      // field = - grad phi
      ef[3 * (size_t)ic] = -ir * (gr * s[0] + gt * et[0] + gp * ep[0]);
      ef[3 * (size_t)ic + 1] = -ir * (gr * s[1] + gt * et[1] + gp * ep[1]);
      ef[3 * (size_t)ic + 2] = -ir * (gr * s[2] + gt * et[2] + gp * ep[2]);
    }
  }
}

// Near part: the charges of the sphere's near list added, in list order, to
// the far part k_l2p left at each cavity point.
__global__ __launch_bounds__(kFT) void k_near_pot(int ncav, const int2 *__restrict__ cmap,
                                                  const double4 *__restrict__ sph,
                                                  const int *__restrict__ noff,
                                                  const int *__restrict__ nidx,
                                                  const double *__restrict__ q,
                                                  const double *__restrict__ ccav,
                                                  double *__restrict__ phi,
                                                  double *__restrict__ ef) {
  const int ic = blockIdx.x * kFT + threadIdx.x;
  if (ic >= ncav) return;
  const int i = cmap[ic].x;
  const double cx = ccav[3 * (size_t)ic], cy = ccav[3 * (size_t)ic + 1],
               cz = ccav[3 * (size_t)ic + 2];
  double v = phi[ic], e0 = 0.0, e1 = 0.0, e2 = 0.0;
  if (ef) {
    e0 = ef[3 * (size_t)ic];
    e1 = ef[3 * (size_t)ic + 1];
    e2 = ef[3 * (size_t)ic + 2];
  }
  const int ne = noff[i + 1];
  for (int k = noff[i]; k < ne; ++k) {
    const int j = nidx[k];
    const double4 sj = sph[j];
    const double dx = cx - sj.x, dy = cy - sj.y, dz = cz - sj.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double ir = 1.0 / sqrt(r2);
    const double qr = q[j] * ir;
    v += qr;
    const double f = qr * ir * ir;
    e0 += f * dx;
    e1 += f * dy;
    e2 += f * dz;
  }
  phi[ic] = v;
  if (ef) {
    ef[3 * (size_t)ic] = e0;
    ef[3 * (size_t)ic + 1] = e1;
    ef[3 * (size_t)ic + 2] = e2;
  }
}

// X'_i = cf M_i, M_i,lm = sum over the exposed points n of sphere i of
// zeta(n) Y_lm(s_n): the multipole of the sphere's cavity charges
template <int P>
__global__ __launch_bounds__(kFT) void k_multipole(int ngrid, const int *__restrict__ cavidx,
                                                   const double *__restrict__ zeta,
                                                   const double *__restrict__ Yp,
                                                   const double *__restrict__ cfP,
                                                   double *__restrict__ Xm) {
  constexpr int NY = (P + 1) * (P + 1);
  extern __shared__ double zs[];
  const int i = blockIdx.x;
  for (int n = threadIdx.x; n < ngrid; n += kFT) {
    const size_t pt = (size_t)i * ngrid + n;
    zs[n] = cavidx[pt + 1] > cavidx[pt] ? zeta[cavidx[pt]] : 0.0;
  }
  __syncthreads();
  for (int lm = threadIdx.x; lm < NY; lm += kFT) {
    double m = 0.0;
    for (int n = 0; n < ngrid; ++n) m += zs[n] * Yp[(size_t)n * NY + lm];
    Xm[(size_t)i * NY + lm] = cfP[lm] * m;
  }
}

// The far field at the nuclei: targets perm[blockIdx.x * kFT + t] (the
// nuclei in Morton order, so that a wave's targets lie close together),
// source spheres of chunk blockIdx.y (spheres sb .. se - 1); partial fields
// per chunk.  A far pair (t = R_i / |u|) needs the order p with
// t^(p+1) <= exp(leps); each wave expands to the largest its pairs need.
template <int P>
__global__ __launch_bounds__(kFT) void k_m2p(int n, int nchunk, double rho,
                                             const double4 *__restrict__ sph,
                                             const int *__restrict__ perm, float leps,
                                             const double *__restrict__ Xm,
                                             double *__restrict__ part) {
  constexpr int NY = (P + 1) * (P + 1);
  __shared__ double xs[NY];
  __shared__ double4 si_sh;
  const int jj = blockIdx.x * kFT + threadIdx.x;
  const int j = jj < n ? perm[jj] : n;
  const int per = (n + nchunk - 1) / nchunk;
  const int sb = blockIdx.y * per, se = min(n, sb + per);
  double4 tj = make_double4(0.0, 0.0, 0.0, 0.0);
  if (j < n) tj = sph[j];
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int i = sb; i < se; ++i) {
    __syncthreads();
    for (int k = threadIdx.x; k < NY; k += kFT) xs[k] = Xm[(size_t)i * NY + k];
    if (threadIdx.x == 0) si_sh = sph[i];
    __syncthreads();
    const double4 si = si_sh;
    const double ux = tj.x - si.x, uy = tj.y - si.y, uz = tj.z - si.z;
    const double u2 = dist2(tj.x, tj.y, tj.z, si.x, si.y, si.z);
    const double rc = si.w / rho;
    const bool far = j < n && u2 >= rc * rc;
    double ru = 0.0, iu = 0.0, t = 0.0;
    int need = -1;
    if (far) {
      ru = sqrt(u2);
      iu = 1.0 / ru;
      t = si.w * iu;
      const int p = (int)ceilf(leps / __logf((float)t)) - 1;
      need = p < 0 ? 0 : (p > P ? P : p);
    }
    const int pw = wave_max_int(need);
    if (far) {
      double cthe, sthe, cphi, sphi;
      angles(ux, uy, uz, ru, cthe, sthe, cphi, sphi);
      double gr = 0.0, gt = 0.0, gp = 0.0, tl = 1.0;
      legendre_rows<P>(
          cthe, sthe, cphi, sphi, xs,
          [&](int l, double a, double b, double c) {
            gr += tl * (-(double)(l + 1) * a);
            gt += tl * b;
            gp += tl * c;
            tl *= t;
          },
          pw);
      double s[3], et[3], ep[3];
      frame(cthe, sthe, cphi, sphi, s, et, ep);
      const double w = iu * iu;   // field = - grad
      e0 -= w * (gr * s[0] + gt * et[0] + gp * ep[0]);
      e1 -= w * (gr * s[1] + gt * et[1] + gp * ep[1]);
      e2 -= w * (gr * s[2] + gt * et[2] + gp * ep[2]);
    }
  }
  if (j < n) {
    const size_t o = ((size_t)blockIdx.y * n + j) * 3;
    part[o] = e0;
    part[o + 1] = e1;
    part[o + 2] = e2;
  }
}

// The near field of zeta at nucleus j = blockIdx.x from the cavity points
// of its near spheres, plus the far partials of the chunks in order.
__global__ __launch_bounds__(kFT) void k_near_field(int n, int ngrid, int nchunk,
                                                    const double4 *__restrict__ sph,
                                                    const int *__restrict__ cavidx,
                                                    const double *__restrict__ ccav,
                                                    const double *__restrict__ zeta,
                                                    const int *__restrict__ toff,
                                                    const int *__restrict__ tidx,
                                                    const double *__restrict__ part,
                                                    double *__restrict__ ez) {
  __shared__ double red[3][kFT / kWave];
  const int j = blockIdx.x;
  const double4 tj = sph[j];
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (int b = toff[j]; b < toff[j + 1]; ++b) {
    const int i = tidx[b];
    const int c0 = cavidx[(size_t)i * ngrid], c1 = cavidx[(size_t)(i + 1) * ngrid];
    for (int c = c0 + (int)threadIdx.x; c < c1; c += kFT) {
      const double dx = tj.x - ccav[3 * (size_t)c], dy = tj.y - ccav[3 * (size_t)c + 1],
                   dz = tj.z - ccav[3 * (size_t)c + 2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double ir = 1.0 / sqrt(r2);
      const double f = zeta[c] * ir * ir * ir;
      e0 += f * dx;
      e1 += f * dy;
      e2 += f * dz;
    }
  }
  e0 = wave_sum(e0);
  e1 = wave_sum(e1);
  e2 = wave_sum(e2);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    red[0][w] = e0;
    red[1][w] = e1;
    red[2][w] = e2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    double v = (red[d][0] + red[d][1]) + (red[d][2] + red[d][3]);
    for (int c = 0; c < nchunk; ++c) v += part[((size_t)c * n + j) * 3 + d];
    ez[3 * (size_t)j + d] = v;
  }
}

}  // namespace

// whether this handle's Coulomb sums take the far-field expansions
bool use_fmm(const ddx_handle *h) {
  if (h->coulomb_mode == DDX_COULOMB_DIRECT || h->n == 0) return false;
  if (h->coulomb_mode == DDX_COULOMB_FMM) return true;
  // the expansions pay where the spheres hold many cavity points and most
  // pairs are far: single systems of thousands of spheres
  return h->ninst == 1 && h->n >= kFmmMinSpheres &&
         (double)h->ncav >= kFmmMinPoints * (double)h->n;
}

namespace {
// tables of order P and the near lists of the current setup
int fmm_prepare(ddx_handle *h) {
  FmmState &f = h->fmm;
  hipStream_t s = h->stream;
  const int P = kFmmOrder;
  const int NY = (P + 1) * (P + 1);
  if (f.ngrid != h->ngrid) {
    // Y on the grid, facs and cf = facs 4pi/(2l+1), to the orders of both passes
    for (int pass = 0; pass < 2; ++pass) {
      const int Q = pass == 0 ? P : kFmmFieldOrder;
      const int NQ = (Q + 1) * (Q + 1);
      std::vector<double> facl, facs, Y((size_t)h->ngrid * NQ), cf(NQ);
      harmonic_factors(Q, facl, facs);
      for (int n = 0; n < h->ngrid; ++n)
        real_harmonics(Q, facs.data(), &h->h_grid[3 * (size_t)n], &Y[(size_t)n * NQ]);
      for (int k = 0; k < NQ; ++k) cf[k] = facs[k] / facl[k];
      DevArray<double> &dY = pass == 0 ? f.Yp : f.Yq;
      DevArray<double> &dcf = pass == 0 ? f.cfP : f.cfQ;
      DDX_TRY(dY.reserve(Y.size()));
      DDX_TRY(dcf.reserve(NQ));
      DDX_HIP(hipMemcpyAsync(dY.p, Y.data(), Y.size() * 8, hipMemcpyHostToDevice, s));
      DDX_HIP(hipMemcpyAsync(dcf.p, cf.data(), NQ * 8, hipMemcpyHostToDevice, s));
      if (pass == 0) {
        DDX_TRY(f.facsP.reserve(NQ));
        DDX_HIP(hipMemcpyAsync(f.facsP.p, facs.data(), NQ * 8, hipMemcpyHostToDevice, s));
      }
      DDX_HIP(hipStreamSynchronize(s));
    }
    f.ngrid = h->ngrid;
  }
  if (f.lists_ok) return DDX_OK;
  const int n = h->n;
  const int *inst = h->ninst > 1 ? h->inst.p : nullptr;
  const int *soff = h->ninst > 1 ? h->soff.p : nullptr;
  const int nb = (n + kFT / kWave - 1) / (kFT / kWave);
  for (int mode = 0; mode < 2; ++mode) {
    DevArray<int> &off = mode == 0 ? f.noff : f.toff;
    DevArray<int> &idx = mode == 0 ? f.nidx : f.tidx;
    DDX_TRY(f.cnt.reserve((size_t)n + 1));
    DDX_TRY(off.reserve((size_t)n + 1));
    hipLaunchKernelGGL(k_near_count, dim3(nb), dim3(kFT), 0, s, n, mode, kFmmRho, h->sph.p, inst,
                       soff, f.cnt.p);
    DDX_HIP(hipGetLastError());
    DDX_TRY(scan_exclusive(f.cnt.p, off.p, (size_t)n, h->scan_tmp, s));
    int total = 0;
    DDX_HIP(hipMemcpyAsync(&total, off.p + n, sizeof(int), hipMemcpyDeviceToHost, s));
    DDX_HIP(hipStreamSynchronize(s));
    DDX_TRY(idx.reserve((size_t)total + 1));
    hipLaunchKernelGGL(k_near_fill, dim3(nb), dim3(kFT), 0, s, n, mode, kFmmRho, h->sph.p, inst,
                       soff, off.p, idx.p);
    DDX_HIP(hipGetLastError());
  }
  // the spheres in Morton order of their centres (30-bit codes over the
  // bounding box), for the adaptive orders of the field pass
  {
    std::vector<double4> c(n);
    DDX_HIP(hipMemcpyAsync(c.data(), h->sph.p, n * sizeof(double4), hipMemcpyDeviceToHost, s));
    DDX_HIP(hipStreamSynchronize(s));
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (const double4 &v : c) {
      const double x[3] = {v.x, v.y, v.z};
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], x[d]);
        hi[d] = std::max(hi[d], x[d]);
      }
    }
    const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    std::vector<std::pair<uint32_t, int>> key(n);
    for (int i = 0; i < n; ++i) {
      const double x[3] = {c[i].x, c[i].y, c[i].z};
      uint32_t code = 0;
      uint32_t b[3];
      for (int d = 0; d < 3; ++d)
        b[d] = (uint32_t)std::min(1023.0, std::max(0.0, (x[d] - lo[d]) / ext * 1024.0));
      for (int bit = 9; bit >= 0; --bit)
        for (int d = 0; d < 3; ++d) code = (code << 1) | ((b[d] >> bit) & 1u);
      key[i] = {code, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = key[i].second;
    DDX_TRY(f.perm.reserve((size_t)n));
    DDX_HIP(hipMemcpyAsync(f.perm.p, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    DDX_HIP(hipStreamSynchronize(s));
  }
  DDX_TRY(f.cmap.reserve((size_t)h->ncav + 1));
  hipLaunchKernelGGL(k_cav_map, dim3((unsigned)((h->np + 255) / 256)), dim3(256), 0, s, h->np,
                     h->ngrid, h->cavidx.p, f.cmap.p);
  DDX_HIP(hipGetLastError());
  f.lists_ok = true;
  return DDX_OK;
}
}  // namespace

// phi at the cavity points and, with ef, the solute field there
int fmm_potential(ddx_handle *h, const double *q, double *phi, double *ef) {
  DDX_TRY(fmm_prepare(h));
  FmmState &f = h->fmm;
  constexpr int P = kFmmOrder;
  constexpr int NY = (P + 1) * (P + 1);
  DDX_TRY(f.A.reserve((size_t)h->n * NY));
  TimedLaunch tl(h, DDX_K_POTENTIAL, h->stream);
  if (tl.b) DDX_HIP(hipEventRecord(tl.b, h->stream));
  DDX_TRY(fmm_p2l(h, P, kFmmRho, q, f.cfP.p, f.A.p));
  if (h->ncav > 0) {
    hipLaunchKernelGGL(k_l2p<P>, dim3(h->n), dim3(kWave), 0, h->stream, h->ngrid, h->cavidx.p,
                       f.cmap.p, h->sph.p, h->grid.p, f.A.p, f.facsP.p, phi, ef);
    DDX_HIP(hipGetLastError());
    const unsigned nb = (unsigned)((h->ncav + kFT - 1) / kFT);
    hipLaunchKernelGGL(k_near_pot, dim3(nb), dim3(kFT), 0, h->stream, h->ncav, f.cmap.p,
                       h->sph.p, f.noff.p, f.nidx.p, q, h->ccav.p, phi, ef);
    DDX_HIP(hipGetLastError());
  }
  if (tl.e) DDX_HIP(hipEventRecord(tl.e, h->stream));
  return DDX_OK;
}

// ez(3, n): the field of the cavity charges zeta at the nuclei
int fmm_zeta_field(ddx_handle *h, const double *zeta, double *ez) {
  DDX_TRY(fmm_prepare(h));
  FmmState &f = h->fmm;
  constexpr int P = kFmmFieldOrder;
  constexpr int NY = (P + 1) * (P + 1);
  const int n = h->n;
  // source chunks: about 2048 workgroups in all
  const int tb = (n + kFT - 1) / kFT;
  const int nchunk = std::max(1, std::min(n, 2048 / std::max(1, tb)));
  DDX_TRY(f.Xm.reserve((size_t)n * NY));
  DDX_TRY(f.part.reserve((size_t)nchunk * n * 3));
  hipStream_t s = h->stream;
  TimedLaunch tl(h, DDX_K_FIELD, s);
  if (tl.b) DDX_HIP(hipEventRecord(tl.b, s));
  hipLaunchKernelGGL(k_multipole<P>, dim3(n), dim3(kFT), h->ngrid * sizeof(double), s, h->ngrid,
                     h->cavidx.p, zeta, f.Yq.p, f.cfQ.p, f.Xm.p);
  DDX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_m2p<P>, dim3(tb, nchunk), dim3(kFT), 0, s, n, nchunk, kFmmRho, h->sph.p,
                     f.perm.p, (float)std::log(kFmmFieldEps), f.Xm.p, f.part.p);
  DDX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_near_field, dim3(n), dim3(kFT), 0, s, n, h->ngrid, nchunk, h->sph.p,
                     h->cavidx.p, h->ccav.p, zeta, f.toff.p, f.tidx.p, f.part.p, ez);
  DDX_HIP(hipGetLastError());
  if (tl.e) DDX_HIP(hipEventRecord(tl.e, s));
  return DDX_OK;
}

}  // namespace ddx
