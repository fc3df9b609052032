// Analytic ddCOSMO forces (forces_dd.f90:85-164) and the field terms of the
// point-charge driver (main.f90:212-265).
//
// fdoka, fdokb and fdoga (ddcosmo.f90:1436-1650) loop over the same
// (owner o, point n, container c) triples as the sweeps.  Per triple t two
// quantities depend only on the triple and the solution of its container:
//   beta_t = sum_l 4pi/(2l+1) t^l sum_m Y_lm(s) sigma_c,lm          (intmlp)
//   alp_t  = sum_{l>=1} 4pi/(2l+1) t^(l-1) sum_m sigma_c,lm (l Y_lm s + dY_lm(s))
// (ddcosmo.f90:1466-1481, 1544-1559; dY as in dbasis, ddcosmo.f90:904-1043).
// k_force_triple computes them once per triple; the owner pass (fdoka, the
// own-point term of fdoga) and the container pass (fdokb with its
// neighbour-of-neighbour sum b over the other triples of the same point,
// ddcosmo.f90:1575-1595, and the neighbour term of fdoga) then only combine
// scalars.  All sums are per-sphere block reductions in a fixed order.
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

__host__ __device__ constexpr int lmi(int l, int m) { return l * l + l + m; }

// Per triple, with X' = cf sigma_c (cf = facs 4pi/(2l+1)) and s the unit
// vector from the container to the point, e_theta = (cos th cos ph,
// cos th sin ph, -sin th), e_phi = (-sin ph, cos ph, 0) / sin th:
//   beta = sum_l t^l A_l,   alp = sum_{l>=1} t^(l-1) (l A_l s + Bt_l e_theta + Bp_l e_phi)
//   A_l  = sum_m P_l^m u_c(m),  u_c = cos(m ph) X'_l,m + sin(m ph) X'_l,-m
//   Bt_l = P_l^1 X'_l,0 - sum_{m>=1} dP_lm u_c(m),
//          dP_lm = (l+m)(l-m+1) P_l^(m-1) / 2 - P_l^(m+1) / 2
//   Bp_l = sum_{m>=1} m P_l^m u_s(m),  u_s = cos(m ph) X'_l,-m - sin(m ph) X'_l,m
// which regroups the l Y s + dY terms of dbasis (ddcosmo.f90:904-1043).  At
// the poles (sin th = 0) e_phi = (0, 1, 0) and the e_phi weight of m is
// -dP_lm int(cos th) (ddcosmo.f90:970-971).  The Legendre functions run
// l-outer, three rows in registers, with polleg's recurrence.
template <int LMAX, int BT>
__global__ __launch_bounds__(BT) void k_force_triple(
    const int *__restrict__ inl, const int *__restrict__ pairoff,
    const double4 *__restrict__ tgeo, const double2 *__restrict__ ttw,
    const double *__restrict__ sigma, const double *__restrict__ cf,
    double4 *__restrict__ fb) {
  constexpr int NY = (LMAX + 1) * (LMAX + 1);
  __shared__ double xs[NY];
  const int c = blockIdx.x;
  for (int k = threadIdx.x; k < NY; k += BT) xs[k] = cf[k] * sigma[(size_t)c * NY + k];
  __syncthreads();
  const int beg = pairoff[inl[c]], end = pairoff[inl[c + 1]];
  for (int k = beg + threadIdx.x; k < end; k += BT) {
    const double4 g = tgeo[k];
    const double t = ttw[k].x;
    const double cthe = g.x, sthe = g.y, cphi = g.z, sphi = g.w;
    double cs[LMAX + 1], sn[LMAX + 1];
    cs[0] = 1.0;
    sn[0] = 0.0;
    if constexpr (LMAX >= 1) {
      cs[1] = cphi;
      sn[1] = sphi;
#pragma unroll
      for (int m = 2; m <= LMAX; ++m) {
        cs[m] = 2.0 * cphi * cs[m - 1] - cs[m - 2];
        sn[m] = 2.0 * cphi * sn[m - 1] - sn[m - 2];
      }
    }
    const bool pole = (sthe == 0.0);
    const double vs = (double)(int)cthe;
    double p2[LMAX + 1], p1[LMAX + 1], pc[LMAX + 1];   // rows l-2, l-1, l
    double beta = 0.0, A = 0.0, Bt = 0.0, Bp = 0.0;
    double tl = 1.0, tlm1 = 1.0;
#pragma unroll
    for (int l = 0; l <= LMAX; ++l) {
#pragma unroll
      for (int m = 0; m <= l; ++m) {
        if (l == 0) pc[0] = 1.0;
        else if (m == l) pc[m] = -(2.0 * l - 1.0) * sthe * p1[l - 1];
        else if (m == l - 1) pc[m] = cthe * (2.0 * m + 1.0) * p1[m];
        else pc[m] = (cthe * (2.0 * l - 1.0) * p1[m] - (double)(l + m - 1) * p2[m]) *
                     (1.0 / (double)(l - m));
      }
      double al = pc[0] * xs[lmi(l, 0)];
      double bt = (l > 0) ? pc[1] * xs[lmi(l, 0)] : 0.0;
      double bp = 0.0;
#pragma unroll
      for (int m = 1; m <= l; ++m) {
        const double xp = xs[lmi(l, m)], xm = xs[lmi(l, -m)];
        const double uc = cs[m] * xp + sn[m] * xm;
        const double us = cs[m] * xm - sn[m] * xp;
        const double dp = 0.5 * ((double)(l + m) * (double)(l - m + 1) * pc[m - 1]) +
                          ((m < l) ? -0.5 * pc[m < l ? m + 1 : m] : 0.0);
        al += pc[m] * uc;
        bt -= dp * uc;
        bp += (pole ? -dp * vs : (double)m * pc[m]) * us;
      }
      beta += tl * al;
      if (l > 0) {
        A += tlm1 * ((double)l * al);
        Bt += tlm1 * bt;
        Bp += tlm1 * bp;
        tlm1 = tl;
      }
      tl *= t;
#pragma unroll
      for (int m = 0; m <= l; ++m) {
        p2[m] = p1[m];
        p1[m] = pc[m];
      }
    }
    const double s0 = sthe * cphi, s1 = sthe * sphi, s2 = cthe;
    const double et0 = cthe * cphi, et1 = cthe * sphi, et2 = -sthe;
    double ep0 = 0.0, ep1 = 1.0;
    if (!pole) {
      const double is = 1.0 / sthe;
      ep0 = -sphi * is;
      ep1 = cphi * is;
    }
    fb[k] = make_double4(A * s0 + Bt * et0 + Bp * ep0, A * s1 + Bt * et1 + Bp * ep1,
                         A * s2 + Bt * et2, beta);
  }
}

// Per-sphere reduction of the force kernels: one wave (BT = 64, large
// systems) or four (BT = 256, few spheres, where one wave per sphere would
// leave most SIMDs idle).
template <int BT>
__device__ __forceinline__ double force_sum(double v, double *lds) {
  if constexpr (BT == kWave) return wave_sum(v);
  else return block_sum_256(v, lds);
}

// fdoka and the own-point term of fdoga for sphere o = blockIdx.x, BT
// threads over its points.
template <int BT>
__global__ __launch_bounds__(BT) void k_force_owner(
    int ngrid, double eta, const int *__restrict__ inl, const int *__restrict__ nl,
    const double4 *__restrict__ sph, const double4 *__restrict__ grid,
    const int *__restrict__ ptoff, const int *__restrict__ pttri,
    const unsigned long long *__restrict__ tmask, const double2 *__restrict__ ttw,
    const double4 *__restrict__ fb, const double *__restrict__ fi,
    const double *__restrict__ ui, const double *__restrict__ zi,
    const double *__restrict__ xi, const double *__restrict__ phi,
    const int *__restrict__ cavidx, double *__restrict__ fo) {
  __shared__ double lds[4];
  const int o = blockIdx.x;
  const double4 so = sph[o];
  const double tlow = 1.0 - eta;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int n = threadIdx.x; n < ngrid; n += BT) {
    const size_t p = (size_t)o * ngrid + n;
    const double4 g = grid[n];
    const double fip = fi[p], xip = xi[p];
    const double z0 = zi[3 * p], z1 = zi[3 * p + 1], z2 = zi[3 * p + 2];
    double va0 = 0.0, va1 = 0.0, va2 = 0.0;
    unsigned long long m = tmask[p];
    for (int k = ptoff[p]; k < ptoff[p + 1]; ++k) {
      const int kk = __ffsll((long long)m) - 1;
      m &= m - 1;
      const int c = nl[inl[o] + kk];
      const int tr = pttri[k];
      const double4 sc = sph[c];
      double v[3];
      point_vec(so, g, sc, v);
      const double vv = norm3(v);
      const double s0 = v[0] / vv, s1 = v[1] / vv, s2 = v[2] / vv;
      const double2 tw = ttw[tr];
      const double4 ab = fb[tr];
      const double beta = ab.w;
      double fw2 = 0.0;
      if (fip > 1.0) fw2 = -tw.y / fip;
      const double fw1 = tw.y / sc.w;
      va0 += fw1 * ab.x + beta * fw2 * z0;
      va1 += fw1 * ab.y + beta * fw2 * z1;
      va2 += fw1 * ab.z + beta * fw2 * z2;
      if (tw.x > tlow) {
        double f3 = beta * dfsw(tw.x, eta) / sc.w;
        if (fip > 1.0) f3 = f3 / fip;
        va0 += f3 * s0;
        va1 += f3 * s1;
        va2 += f3 * s2;
      }
    }
    const double wx = g.w * xip;
    f0 -= wx * va0;
    f1 -= wx * va1;
    f2 -= wx * va2;
    const double u = ui[p];
    if (u > 0.0 && u < 1.0) {
      const double ph = phi[cavidx[p]];
      f0 -= g.w * (ph * xip * z0);
      f1 -= g.w * (ph * xip * z1);
      f2 -= g.w * (ph * xip * z2);
    }
  }
  f0 = force_sum<BT>(f0, lds);
  f1 = force_sum<BT>(f1, lds);
  f2 = force_sum<BT>(f2, lds);
  if (threadIdx.x == 0) {
    fo[3 * (size_t)o] = f0;
    fo[3 * (size_t)o + 1] = f1;
    fo[3 * (size_t)o + 2] = f2;
  }
}

// fdokb and the neighbour term of fdoga for container c = blockIdx.x, BT
// threads over its triples.
template <int BT>
__global__ __launch_bounds__(BT) void k_force_container(
    int ngrid, double eta, const int *__restrict__ inl, const int *__restrict__ pairoff,
    const double4 *__restrict__ sph, const double4 *__restrict__ grid,
    const int *__restrict__ ptoff, const int *__restrict__ pttri,
    const double2 *__restrict__ ttw, const int *__restrict__ tp,
    const double4 *__restrict__ fb, const double *__restrict__ fi,
    const double *__restrict__ ui, const double *__restrict__ xi,
    const double *__restrict__ phi, const int *__restrict__ cavidx,
    double *__restrict__ fc) {
  __shared__ double lds[4];
  const int c = blockIdx.x;
  const int beg = pairoff[inl[c]], end = pairoff[inl[c + 1]];
  const double4 sc = sph[c];
  const double tlow = 1.0 - eta;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int k = beg + threadIdx.x; k < end; k += BT) {
    const int p = tp[k];
    const int o = p / ngrid, n = p - o * ngrid;
    const double4 g = grid[n];
    double v[3];
    point_vec(sph[o], g, sc, v);
    const double vv = norm3(v);
    const double s0 = v[0] / vv, s1 = v[1] / vv, s2 = v[2] / vv;
    const double2 tw = ttw[k];
    const double t = tw.x;
    const double4 ab = fb[k];
    const double fip = fi[p], xip = xi[p];
    const double f1w = tw.y / sc.w;
    double vb0 = f1w * ab.x * xip, vb1 = f1w * ab.y * xip, vb2 = f1w * ab.z * xip;
    double vc0 = 0.0, vc1 = 0.0, vc2 = 0.0;
    if (t > tlow) {
      const double beta = ab.w;
      const double chi = fsw(t, eta);
      const double dchi = dfsw(t, eta);
      double di = 1.0, fac = 0.0;
      if (fip > 1.0) {
        di = 1.0 / fip;
        fac = di * chi;
        double bsum = 0.0;
        for (int q = ptoff[p]; q < ptoff[p + 1]; ++q) {
          const int tr = pttri[q];
          if (tr == k) continue;
          bsum = bsum + fb[tr].w * fsw(ttw[tr].x, eta);
        }
        const double g2 = di * di * dchi / sc.w * xip * bsum;
        vc0 = g2 * s0; vc1 = g2 * s1; vc2 = g2 * s2;
      }
      const double f2w = (1.0 - fac) * di * dchi / sc.w;
      vb0 += f2w * xip * beta * s0;
      vb1 += f2w * xip * beta * s1;
      vb2 += f2w * xip * beta * s2;
    }
    f0 += g.w * (vb0 - vc0);
    f1 += g.w * (vb1 - vc1);
    f2 += g.w * (vb2 - vc2);
    // fdoga, neighbour part: switch region and an exposed point of the owner
    if (t < 1.0 && t > 1.0 - eta && ui[p] > 0.0) {
      const double fa = dfsw(t, eta) / sc.w * phi[cavidx[p]] * xip;
      f0 += g.w * (fa * s0);
      f1 += g.w * (fa * s1);
      f2 += g.w * (fa * s2);
    }
  }
  f0 = force_sum<BT>(f0, lds);
  f1 = force_sum<BT>(f1, lds);
  f2 = force_sum<BT>(f2, lds);
  if (threadIdx.x == 0) {
    fc[3 * (size_t)c] = f0;
    fc[3 * (size_t)c + 1] = f1;
    fc[3 * (size_t)c + 2] = f2;
  }
}

__global__ void k_force_combine(int n, double fep, const double *__restrict__ fo,
                                const double *__restrict__ fc, double *__restrict__ fx) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 3 * n) fx[k] = fep * (fo[k] + fc[k]);
}

// zeta(ic) = 1/2 f(eps) w_n ui xi(n,i) for exposed points (ddcosmo.f90:1787-1806)
__global__ void k_zeta(size_t np, int ngrid, double scale, const double4 *__restrict__ grid,
                       const double *__restrict__ ui, const int *__restrict__ cavidx,
                       const double *__restrict__ xi, double *__restrict__ zeta) {
  size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double u = ui[p];
  // 32-bit point indices (run_setup checks the count)
  if (u > 0.0) zeta[cavidx[p]] = scale * (grid[(unsigned)p % (unsigned)ngrid].w * u * xi[p]);
}

// fx_i -= sum over the cavity points of sphere i of zeta E  (main.f90:245-253)
// and fx_i -= q_i E_zeta(r_i)  (main.f90:263-265)
__global__ __launch_bounds__(256) void k_field_contract(int ngrid, const int *__restrict__ cavidx,
                                                        const double *__restrict__ zeta,
                                                        const double *__restrict__ ef,
                                                        const double *__restrict__ q,
                                                        const double *__restrict__ ezeta,
                                                        double *__restrict__ fx) {
  __shared__ double lds[4];
  const int i = blockIdx.x;
  const int b = cavidx[(size_t)i * ngrid], e = cavidx[(size_t)(i + 1) * ngrid];
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int ic = b + threadIdx.x; ic < e; ic += 256) {
    double z = zeta[ic];
    f0 += z * ef[3 * (size_t)ic];
    f1 += z * ef[3 * (size_t)ic + 1];
    f2 += z * ef[3 * (size_t)ic + 2];
  }
  f0 = block_sum_256(f0, lds);
  f1 = block_sum_256(f1, lds);
  f2 = block_sum_256(f2, lds);
  if (threadIdx.x == 0) {
    double qi = q[i];
    fx[3 * (size_t)i] = fx[3 * (size_t)i] - f0 - ezeta[3 * (size_t)i] * qi;
    fx[3 * (size_t)i + 1] = fx[3 * (size_t)i + 1] - f1 - ezeta[3 * (size_t)i + 1] * qi;
    fx[3 * (size_t)i + 2] = fx[3 * (size_t)i + 2] - f2 - ezeta[3 * (size_t)i + 2] * qi;
  }
}

__global__ void k_sph3(int n, const double4 *__restrict__ sph, double *__restrict__ c3) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 s = sph[i];
  c3[3 * i] = s.x;
  c3[3 * i + 1] = s.y;
  c3[3 * i + 2] = s.z;
}

// waves per sphere of the force kernels: one where the spheres alone fill
// the device several times over, else four
int force_waves(const ddx_handle *h) { return h->n >= 16384 ? 1 : 4; }

template <int LMAX>
int launch_triple(ddx_handle *h, const double *sigma, double4 *fb) {
  TimedLaunch tl(h, DDX_K_FORCES, h->stream);
  if (force_waves(h) == 1)
    hipExtLaunchKernelGGL((k_force_triple<LMAX, kWave>), dim3(h->n), dim3(kWave), 0, h->stream,
                          tl.b, tl.e, 0, h->inl.p, h->pairoff.p, h->tgeo.p, h->ttw.p, sigma,
                          h->cf.p, fb);
  else
    hipExtLaunchKernelGGL((k_force_triple<LMAX, 256>), dim3(h->n), dim3(256), 0, h->stream, tl.b,
                          tl.e, 0, h->inl.p, h->pairoff.p, h->tgeo.p, h->ttw.p, sigma, h->cf.p,
                          fb);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int force_triple(ddx_handle *h, const double *sigma, double4 *fb) {
  switch (h->lmax) {
    case 0: return launch_triple<0>(h, sigma, fb);
    case 1: return launch_triple<1>(h, sigma, fb);
    case 2: return launch_triple<2>(h, sigma, fb);
    case 3: return launch_triple<3>(h, sigma, fb);
    case 4: return launch_triple<4>(h, sigma, fb);
    case 5: return launch_triple<5>(h, sigma, fb);
    case 6: return launch_triple<6>(h, sigma, fb);
    case 7: return launch_triple<7>(h, sigma, fb);
    case 8: return launch_triple<8>(h, sigma, fb);
    case 9: return launch_triple<9>(h, sigma, fb);
    case 10: return launch_triple<10>(h, sigma, fb);
    case 11: return launch_triple<11>(h, sigma, fb);
    case 12: return launch_triple<12>(h, sigma, fb);
    case 13: return launch_triple<13>(h, sigma, fb);
    case 14: return launch_triple<14>(h, sigma, fb);
    case 15: return launch_triple<15>(h, sigma, fb);
    case 16: return launch_triple<16>(h, sigma, fb);
    default: set_error("lmax %d not supported (0..16)", h->lmax); return DDX_ERR_ARG;
  }
}

}  // namespace

// forces_dd: xi = basis^T s, per-triple beta/alp, owner and container passes,
// scale by 1/2 (eps-1)/eps (forces_dd.f90:160-161).
int forces_dd(ddx_handle *h, const double *d_phi, const double *d_sigma, const double *d_s,
              double *d_fx, bool xi_ready) {
  if (!h->igrad) {
    set_error("forces need igrad = 1 at ddx_init (zi is built only then)");
    return DDX_ERR_STATE;
  }
  hipStream_t s = h->stream;
  int n = h->n;
  DDX_TRY(h->xi.reserve(h->np));
  if (!xi_ready) DDX_TRY(expand(h, h->stream, d_s, h->xi.p, false, nullptr));
  DDX_TRY(h->r.reserve(4 * (size_t)h->ntri + 4));
  double4 *fb = reinterpret_cast<double4 *>(h->r.p);
  DDX_TRY(h->fdd.reserve(6 * (size_t)n + 6));
  double *fo = h->fdd.p, *fc = h->fdd.p + 3 * (size_t)n;
  if (n > 0) {
    DDX_TRY(force_triple(h, d_sigma, fb));
    const bool one = force_waves(h) == 1;
    hipLaunchKernelGGL(one ? k_force_owner<kWave> : k_force_owner<256>, dim3(n),
                       dim3(one ? kWave : 256), 0, s, h->ngrid, h->eta, h->inl.p, h->nl.p,
                       h->sph.p, h->grid.p, h->ptoff.p, h->pttri.p, h->tmask.p, h->ttw.p, fb,
                       h->fi.p, h->ui.p, h->zi.p, h->xi.p, d_phi, h->cavidx.p, fo);
    DDX_HIP(hipGetLastError());
    hipLaunchKernelGGL(one ? k_force_container<kWave> : k_force_container<256>, dim3(n),
                       dim3(one ? kWave : 256), 0, s, h->ngrid, h->eta, h->inl.p, h->pairoff.p,
                       h->sph.p, h->grid.p, h->ptoff.p, h->pttri.p, h->ttw.p, h->tp.p, fb,
                       h->fi.p, h->ui.p, h->xi.p, d_phi, h->cavidx.p, fc);
    DDX_HIP(hipGetLastError());
    double fep = 0.5 * (h->eps - 1.0) / h->eps;
    hipLaunchKernelGGL(k_force_combine, dim3((3 * n + 255) / 256), dim3(256), 0, s, n, fep, fo,
                       fc, d_fx);
    DDX_HIP(hipGetLastError());
  }
  return DDX_OK;
}

int make_zeta(ddx_handle *h, const double *d_s, double *d_zeta, bool xi_ready) {
  hipStream_t s = h->stream;
  DDX_TRY(h->xi.reserve(h->np));
  // xi = basis^T s, unless forces_dd left it for the same s
  if (!xi_ready) DDX_TRY(expand(h, h->stream, d_s, h->xi.p, false, nullptr));
  double scale = 0.5 * ((h->eps - 1.0) / h->eps);
  hipLaunchKernelGGL(k_zeta, dim3((h->np + 255) / 256), dim3(256), 0, s, h->np, h->ngrid, scale,
                     h->grid.p, h->ui.p, h->cavidx.p, h->xi.p, d_zeta);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

size_t field_buffer_size(const ddx_handle *h) {
  return 3 * ((size_t)h->ncav + h->n) + 6 + 3 * (size_t)h->n + (size_t)h->ncav + 1;
}

int field_forces_fields(ddx_handle *h, const double *d_charge, const double *d_zeta,
                        bool solute_field) {
  hipStream_t s = h->stream;
  int n = h->n;
  size_t nef = 3 * ((size_t)h->ncav + n) + 6;
  DDX_TRY(h->ef.reserve(field_buffer_size(h)));
  double *c3 = h->ef.p + nef;
  double *ef = h->ef.p, *ez = h->ef.p + 3 * (size_t)h->ncav + 3;
  hipLaunchKernelGGL(k_sph3, dim3((n + 255) / 256), dim3(256), 0, s, n, h->sph.p, c3);
  DDX_HIP(hipGetLastError());
  // solute field at the cavity points (unless the potential kernel made
  // it), and the field of zeta at the nuclei
  const bool fmm = use_fmm(h);
  if (!solute_field) {
    double *phi_scratch = c3 + 3 * (size_t)n;
    if (fmm)
      DDX_TRY(fmm_potential(h, d_charge, phi_scratch, ef));
    else
      DDX_TRY(coulomb_potential(h, h->ctile.p, h->nctile, d_charge, h->sph.p, h->ccav.p,
                                phi_scratch, ef));
  }
  if (fmm)
    DDX_TRY(fmm_zeta_field(h, d_zeta, ez));
  else
    DDX_TRY(coulomb_field(h, h->ftile.p, h->nftile, h->fgrp.p, h->nfgrp, d_zeta, h->ccav.p, c3,
                          ez));
  return DDX_OK;
}

int field_forces_contract(ddx_handle *h, const double *d_charge, const double *d_zeta,
                          double *d_fx) {
  const int n = h->n;
  double *ef = h->ef.p, *ez = h->ef.p + 3 * (size_t)h->ncav + 3;
  hipLaunchKernelGGL(k_field_contract, dim3(n), dim3(256), 0, h->stream, h->ngrid, h->cavidx.p,
                     d_zeta, ef, d_charge, ez, d_fx);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int field_forces(ddx_handle *h, const double *d_charge, const double *d_zeta, double *d_fx,
                 bool solute_field) {
  DDX_TRY(field_forces_fields(h, d_charge, d_zeta, solute_field));
  return field_forces_contract(h, d_charge, d_zeta, d_fx);
}

}  // namespace ddx
