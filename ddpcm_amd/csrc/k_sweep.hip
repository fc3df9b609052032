// The off-diagonal ddCOSMO operators on the neighbour triples.
//
// For a triple t = (owner sphere o, grid point n, container sphere c) with
// t_n^oc < 1 define the vector
//     R_lm(t) = 4pi/(2l+1) * (t_n^oc)^l * Y_lm(s_n^oc)
// (the factors of intmlp, ddcosmo.f90:1140-1178, and of adjrhs,
// ddcosmo.f90:1359-1377).  Then
//     L X :  P(n,o)  = sum_{t at (o,n)} W_t R(t) . X_c        (calcv, ddcosmo.f90:1671-1746)
//            (LX)_o  = - sum_n w_n Y(:,n) P(n,o)              (intrhs + lx, matvec.f90:127-136)
//     L* S:  xi(n,o) = Y(:,n) . S_o                           (matvec.f90:196-204)
//            (L*S)_c = - sum_{t in c} w_n xi(n,o) W_t R(t)    (adjrhs, matvec.f90:213-223)
//
// One sweep kernel serves both operators, and both at once: the triples are
// grouped by container c, one wave per container.  The per-(l, m) constant
// cf = facs * 4pi/(2l+1) is applied outside the triple loop: L X contracts
// with X' = cf X (uniform across the wave, read through the scalar cache),
// L* S accumulates sum_t coef_t p~ trig in registers per lane and multiplies
// by cf after one LDS transpose reduction per container.  When the direct
// and the adjoint solve run in lockstep a single pass over the triples
// builds each basis once for both.
// The basis follows ylmbas (ddcosmo.f90:830-895) with t^l folded into the
// Legendre recurrence of polleg (m outer, l inner, ddcosmo.f90:1056-1094):
//     p~(m, m)   = t^m P_m^m,   p~(m+1, m) = (2m+1) (t cos) p~(m, m)
//     p~(l, m)   = [(2l-1) (t cos) p~(l-1, m) - (l+m-1) t^2 p~(l-2, m)] / (l-m)
// and cos/sin(m phi) by the Chebyshev recurrence of trgev
// (ddcosmo.f90:1103-1125).
#include "ddx_device.hpp"
#include "ddx_kernels.hpp"

namespace ddx {
namespace {

__host__ __device__ constexpr int lmi(int l, int m) { return l * l + l + m; }

// The adjoint accumulators of a lane live in registers.  Where all nylm of
// them do not fit they are split into blocks of consecutive m, each built by
// its own pass over the chunk's triples: block b holds the entries (l, +-m)
// with mblock_begin(b) <= m < mblock_begin(b + 1).  The first pass of the
// paired sweep also evaluates L X and gets a smaller budget.  The budgets
// are the largest that the register allocator fits in 256 VGPRs without
// spilling (its resource report, at two waves per SIMD).
// FP32 accumulators (F) take one register each instead of two.
__host__ __device__ constexpr int acc_cap(int lmax, bool F) {
  return F ? 128 : ((lmax == 13 || lmax == 14) ? 64 : 72);
}
__host__ __device__ constexpr int acc_pair_less(bool F) { return F ? 32 : 16; }  // taken by L X
__host__ __device__ constexpr int acc_one(bool F) { return F ? 144 : 81; }  // one block up to here

// entries (l, +-m), l = m..lmax
__host__ __device__ constexpr int mcount(int lmax, int m) {
  return (m == 0 ? 1 : 2) * (lmax + 1 - m);
}
// accumulators of m0 <= m < m1
__host__ __device__ constexpr int mspan(int lmax, int m0, int m1) {
  int c = 0;
  for (int m = m0; m < m1; ++m) c += mcount(lmax, m);
  return c;
}
// first m of block b; lmax + 1 once the blocks are used up
__host__ __device__ constexpr int mblock_begin(int lmax, bool dir, int b, bool F) {
  if ((lmax + 1) * (lmax + 1) <= acc_one(F)) return b == 0 ? 0 : lmax + 1;
  int m = 0;
  for (int k = 0; k < b && m <= lmax; ++k) {
    const int cap = acc_cap(lmax, F) - ((k == 0 && dir) ? acc_pair_less(F) : 0);
    int cnt = 0;
    while (m <= lmax && cnt + mcount(lmax, m) <= cap) cnt += mcount(lmax, m++);
  }
  return m;
}
// position of entry (l, m) (neg: (l, -m)) among the accumulators of the
// block m0 <= m < m1: lm order when the block holds all entries, else m by
// m, and inside m: l, +m before -m
__host__ __device__ constexpr int aidx(int lmax, int m0, int m1, int l, int m, bool neg) {
  if (m0 == 0 && m1 == lmax + 1) return l * l + l + (neg ? -m : m);
  return mspan(lmax, m0, m) + (m == 0 ? l : 2 * (l - m) + (neg ? 1 : 0));
}

// X' of the container in the wave's LDS row
template <typename R>
using LdsPtr = __attribute__((address_space(3))) const R *;

// One triple: returns sum_lm p~_lm trig_m Xc'_lm when DIR and adds
// coef p~_lm trig_m to the accumulators of AM0 <= |m| < AM1 (trig_m =
// cos(m phi) for m >= 0, sin(|m| phi) for m < 0).  Without DIR the
// recurrences in m only step up to AM0 before the l recurrence starts.
template <typename R, int LMAX, bool DIR, int AM0, int AM1>
__device__ __forceinline__ R sweep_triple(R cthe, R sthe, R cphi, R sphi, R t, LdsPtr<R> Xc, R coef,
                                          R *acc) {
  constexpr bool ADJ = AM0 < AM1;
  constexpr int MB = DIR ? 0 : AM0;          // first m with the l recurrence
  constexpr int ME = DIR ? LMAX + 1 : AM1;   // one past the last m needed
  const R ct = cthe * t, st = sthe * t, t2 = t * t;
  R s = 0.0;
  R pmm = 1.0;   // t^m P_m^m
  R cm = 1.0, sm = 0.0, cprev = 0.0, sprev = 0.0;
#pragma unroll
  for (int m = 0; m < ME; ++m) {
    if (m >= MB) {
      const bool acc_m = ADJ && m >= AM0 && m < AM1;
      R qc = 0.0, qs = 0.0;
      if (acc_m) {
        qc = (m == 0) ? coef : coef * cm;
        qs = (m == 0) ? R(0) : coef * sm;
      }
      R a = 0.0, b = 0.0;
      // X' of this m is read from LDS only after the previous m's sum: the
      // compiler would otherwise keep all nylm values in registers
      LdsPtr<R> Xm = Xc;
      if (DIR) asm volatile("" : "+v"(Xm) : "v"(s));
      auto visit = [&](int l, R p) {
        if (DIR) {
          a += p * Xm[lmi(l, m)];
          if (m > 0) b += p * Xm[lmi(l, -m)];
        }
        if (acc_m) {
          acc[aidx(LMAX, AM0, AM1, l, m, false)] += p * qc;
          if (m > 0) acc[aidx(LMAX, AM0, AM1, l, m, true)] += p * qs;
        }
      };
      R p0 = pmm;
      visit(m, p0);
      if (m < LMAX) {
        R p1 = ((R)(2 * m + 1) * ct) * p0;
        visit(m + 1, p1);
#pragma unroll
        for (int l = m + 2; l <= LMAX; ++l) {
          constexpr R one = 1;
          const R ca = (R)(2 * l - 1) * (one / (R)(l - m));
          const R cb = (R)(l + m - 1) * (one / (R)(l - m));
          const R p2 = ca * (ct * p1) - cb * (t2 * p0);
          visit(l, p2);
          p0 = p1;
          p1 = p2;
        }
      }
      if (DIR) s += (m == 0) ? a : (cm * a + sm * b);
    }
    if (m + 1 < ME) {
      pmm = (-(R)(2 * m + 1) * st) * pmm;
      R cn, sn;
      if (m == 0) { cn = cphi; sn = sphi; }
      else { cn = R(2) * cphi * cm - cprev; sn = R(2) * cphi * sm - sprev; }
      cprev = cm; sprev = sm; cm = cn; sm = sn;
    }
  }
  return s;
}

// Ratio K(l, m) / K(l - 1, m) of the scales of the FP32 basis (l > m), and
// the coefficient of t^2 q(l - 2, m) in its recurrence (l >= m + 2).
__host__ __device__ constexpr double basis_ratio(int l, int m) {
  return l == m + 1 ? 2.0 * m + 1.0 : (2.0 * l - 1.0) / (double)(l - m);
}
__host__ __device__ constexpr double basis_beta(int l, int m) {
  return ((double)(l + m - 1) / (double)(l - m)) / (basis_ratio(l, m) * basis_ratio(l - 1, m));
}

// sweep_triple in FP32 on the scaled basis q_lm = p~_lm / K_lm (ct = t cos
// theta, st = t sin theta, t^2 = ct^2 + st^2), where
//     K(m, m) = -(2m - 1) K(m - 1, m - 1),  K(m + 1, m) = (2m + 1) K(m, m),
//     K(l, m) = (2l - 1) / (l - m) K(l - 1, m),
// so that every recurrence step has one product less:
//     q(m, m) = st q(m - 1, m - 1),  q(m + 1, m) = ct q(m, m),
//     q(l, m) = ct q(l - 1, m) - beta(l, m) t^2 q(l - 2, m).
// Xc holds K X' and the caller multiplies the adjoint sums by K.
template <int LMAX, bool DIR, int AM0, int AM1>
__device__ __forceinline__ float sweep_triple32(float ct, float st, float cphi, float sphi,
                                                LdsPtr<float> Xc, float coef, float *acc) {
  constexpr bool ADJ = AM0 < AM1;
  constexpr int MB = DIR ? 0 : AM0;          // first m with the l recurrence
  constexpr int ME = DIR ? LMAX + 1 : AM1;   // one past the last m needed
  const float t2 = fmaf(ct, ct, st * st), c2 = 2.0f * cphi;
  float s = 0.0f;
  float qmm = 1.0f;
  float cm = 1.0f, sm = 0.0f, cprev = 0.0f, sprev = 0.0f;
#pragma unroll
  for (int m = 0; m < ME; ++m) {
    if (m >= MB) {
      const bool acc_m = ADJ && m >= AM0 && m < AM1;
      float qc = 0.0f, qs = 0.0f;
      if (acc_m) {
        qc = (m == 0) ? coef : coef * cm;
        qs = (m == 0) ? 0.0f : coef * sm;
      }
      float a = 0.0f, b = 0.0f;
      LdsPtr<float> Xm = Xc;
      if (DIR) asm volatile("" : "+v"(Xm) : "v"(s));
      auto visit = [&](int l, float q) {
        if (DIR) {
          if (l == m) {
            a = q * Xm[lmi(l, m)];
            if (m > 0) b = q * Xm[lmi(l, -m)];
          } else {
            a = fmaf(q, Xm[lmi(l, m)], a);
            if (m > 0) b = fmaf(q, Xm[lmi(l, -m)], b);
          }
        }
        if (acc_m) {
          float &ap = acc[aidx(LMAX, AM0, AM1, l, m, false)];
          ap = fmaf(q, qc, ap);
          if (m > 0) {
            float &an = acc[aidx(LMAX, AM0, AM1, l, m, true)];
            an = fmaf(q, qs, an);
          }
        }
      };
      float q0 = qmm;
      visit(m, q0);
      if (m < LMAX) {
        float q1 = ct * q0;
        visit(m + 1, q1);
#pragma unroll
        for (int l = m + 2; l <= LMAX; ++l) {
          const float q2 = fmaf(-(float)basis_beta(l, m), t2 * q0, ct * q1);
          visit(l, q2);
          q0 = q1;
          q1 = q2;
        }
      }
      if (DIR) {
        if (m == 0) {
          s = a;
        } else {
          s = fmaf(cm, a, s);
          s = fmaf(sm, b, s);
        }
      }
    }
    if (m + 1 < ME) {
      qmm = st * qmm;
      float cn, sn;
      if (m == 0) {
        cn = cphi;
        sn = sphi;
      } else {
        cn = fmaf(c2, cm, -cprev);
        sn = fmaf(c2, sm, -sprev);
      }
      cprev = cm; sprev = sm; cm = cn; sm = sn;
    }
  }
  return s;
}

// List length per lane after s halvings of n values.
__host__ __device__ constexpr int reduced_len(int n, int s) {
  return s == 0 ? n : reduced_len((n + 1) / 2, s - 1);
}

// 64-bit value of the partner lane of stage S of wave_reduce_split (S >= 2):
// mirror within 16 lanes, mirror within 8, then quad swaps by 2 and by 1
template <int S>
__device__ __forceinline__ double dpp_partner(double v) {
  constexpr int ctrl = S == 2 ? 0x140 : S == 3 ? 0x141 : S == 4 ? 0x4E : 0xB1;
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), ctrl, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Stage S of a reduce-scatter across the wave: each lane splits its list of
// N values into halves x = v[0, H) and y = v[H, N), pairs with a lane that
// differs in lane bit 5 - S (and holds the same columns), keeps one half
// (bit clear: x, set: y) and adds the partner's copy of that half.  Stage 0
// pairs lanes 32 apart and stage 1 lanes 16 apart by the gfx950 row swaps,
// stages 2..5 use DPP.  After six stages each lane holds the wave sums of
// columns base .. base + len - 1.
// FP32 values move as one 32-bit word.
template <int S>
__device__ __forceinline__ float dpp_partner(float v) {
  constexpr int ctrl = S == 2 ? 0x140 : S == 3 ? 0x141 : S == 4 ? 0x4E : 0xB1;
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xf, 0xf, false));
}

template <int S, int N, typename R>
__device__ __forceinline__ void wave_reduce_split(R *v, int &base, int &len) {
  constexpr int H = (N + 1) / 2;
  const bool side = (lane_id() >> (5 - S)) & 1;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const R x = v[k];
    const R y = (k + H < N) ? v[k + H] : R(0);
    if constexpr (S < 2 && sizeof(R) == 4) {
      const unsigned xb = __float_as_uint(x), yb = __float_as_uint(y);
      unsigned nxb, nyb;
      if constexpr (S == 0) {
        auto r = __builtin_amdgcn_permlane32_swap(xb, yb, false, false);
        nxb = r[0]; nyb = r[1];
      } else {
        auto r = __builtin_amdgcn_permlane16_swap(xb, yb, false, false);
        nxb = r[0]; nyb = r[1];
      }
      v[k] = __uint_as_float(nxb) + __uint_as_float(nyb);
    } else if constexpr (S < 2) {
      const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
      const unsigned xl = (unsigned)(bx & 0xffffffffll), xh = (unsigned)(bx >> 32);
      const unsigned yl = (unsigned)(by & 0xffffffffll), yh = (unsigned)(by >> 32);
      unsigned nxl, nyl, nxh, nyh;
      if constexpr (S == 0) {
        auto rl = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
        auto rh = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
        nxl = rl[0]; nyl = rl[1]; nxh = rh[0]; nyh = rh[1];
      } else {
        auto rl = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
        auto rh = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
        nxl = rl[0]; nyl = rl[1]; nxh = rh[0]; nyh = rh[1];
      }
      // the swap leaves, in every lane, its own half and the partner's copy
      // of it, one in each register
      const double nx = __longlong_as_double(((long long)nxh << 32) | nxl);
      const double ny = __longlong_as_double(((long long)nyh << 32) | nyl);
      v[k] = nx + ny;
    } else {
      const double keep = side ? y : x;
      const double send = side ? x : y;
      v[k] = keep + dpp_partner<S>(send);
    }
  }
  if (side) {
    base += H;
    len = len > H ? len - H : 0;
  } else {
    len = len < H ? len : H;
  }
  if constexpr (S < 5) wave_reduce_split<S + 1, H, R>(v, base, len);
}

// The per-triple records in the sweep's arithmetic: (cos theta, sin theta,
// cos phi, sin phi) and (t, W) as double4 / double2 for the FP64 sweep;
// (t cos theta, t sin theta, cos phi, sin phi) as float4 and W as float for
// the FP32 sweep (20 bytes: t enters only through those products and t^2).
template <typename R>
struct Geo {
  typedef double4 G4;
  typedef double2 G2;
};
template <>
struct Geo<float> {
  typedef float4 G4;   // (t cos theta, t sin theta, cos phi, sin phi)
  typedef float G2;    // W
};

// One pass over triples beg..end-1 (end > beg) of a container: rw[tslot] =
// W (R . X_c) when DIR (X' in xsh), and when AM0 < AM1 the (l, +-m),
// AM0 <= m < AM1, entries of the chunk's share of (L*S)_c, -cf * sum_t w_n
// xi(n, o) W p~ trig, into the segment row yp.  The FP32 pass works on the
// scaled basis of sweep_triple32: xsh holds K X' and the sums are scaled by
// cf K (the table after cf, 1/(l+1), 1/facl and K in cf).
template <typename R, int LMAX, bool DIR, int AM0, int AM1>
__device__ __forceinline__ void sweep_pass(
    int beg, int end, const typename Geo<R>::G4 *__restrict__ tgeo,
    const typename Geo<R>::G2 *__restrict__ ttw,
    const int *__restrict__ tslot, const int *__restrict__ tp, LdsPtr<R> xsh,
    R *__restrict__ rw, int spare, const R *__restrict__ xiw, const double *__restrict__ cf,
    double *__restrict__ yp) {
  constexpr bool F = sizeof(R) == 4;
  constexpr bool ADJ = AM0 < AM1;
  constexpr int NA = ADJ ? mspan(LMAX, AM0, AM1) : 1;
  R acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = R(0);
  // software pipeline, two triples deep: geometry, slot and owner point of
  // triple k + 128 and xi of triple k + 64 load while triple k is evaluated
  typedef typename Geo<R>::G4 G4;
  typedef typename Geo<R>::G2 G2;
  int k = beg + lane_id();
  if constexpr (F) {
    // Indices past the chunk read its last triple instead, so that no load
    // is conditional, and the loop is unrolled by three so that the three
    // stages rotate by name (a register copy of a load in flight waits for
    // it).  The rounds have no exit between their steps, which would let the
    // compiler sink the loads to their uses: a lane past the chunk's end
    // evaluates its clamped triple with weight 0 and stores into the spare
    // slot rw[spare].
    struct Stage {
      G4 g;
      G2 w;   // W
      int s, p;
      R x;
    };
    const int last = end - 1;
    auto fetch = [&](Stage &st, int j) {
      j = min(j, last);
      st.g = tgeo[j];
      st.w = ttw[j];
      if constexpr (DIR) st.s = tslot[j];
      if constexpr (ADJ) st.p = tp[j];
    };
    Stage A, B, C;
    fetch(A, k);
    if constexpr (ADJ) A.x = xiw[A.p];
    fetch(B, k + kWave);
    // triple k from cur; triple k + 128 into fut, xi of triple k + 64 into nxt
    auto step = [&](Stage &cur, Stage &nxt, Stage &fut) {
      fetch(fut, k + 2 * kWave);
      if constexpr (ADJ) nxt.x = xiw[nxt.p];
      const bool live = k < end;
      const R coef = ADJ ? (live ? cur.x * cur.w : R(0)) : R(0);
      const R v = sweep_triple32<LMAX, DIR, AM0, AM1>(cur.g.x, cur.g.y, cur.g.z, cur.g.w, xsh,
                                                      coef, acc);
      if constexpr (DIR) rw[live ? cur.s : spare] = cur.w * v;
      k += kWave;
    };
    const int nit = (end - beg + kWave - 1) / kWave;   // the same in all lanes
    int it = 0;
    for (; it + 3 <= nit; it += 3) {
      step(A, B, C);
      step(B, C, A);
      step(C, A, B);
    }
    if (it < nit) {
      step(A, B, C);
      if (it + 1 < nit) step(B, C, A);
    }
  } else {
    // FP64: loads under their bounds tests, stages copied down after each
    // triple (the passes of lmax >= 8 need the registers a third stage takes)
    G4 z4;
    z4.x = z4.y = z4.z = z4.w = R(0);
    G2 z2;
    z2.x = z2.y = R(0);
    G4 g0 = z4, g1 = z4;
    G2 w0 = z2, w1 = z2;
    int s0 = 0, s1 = 0, p1 = 0;
    R x0 = R(0);
    if (k < end) {
      g0 = tgeo[k];
      w0 = ttw[k];
      if (DIR) s0 = tslot[k];
      if (ADJ) x0 = xiw[tp[k]];
    }
    if (k + kWave < end) {
      g1 = tgeo[k + kWave];
      w1 = ttw[k + kWave];
      if (DIR) s1 = tslot[k + kWave];
      if (ADJ) p1 = tp[k + kWave];
    }
    for (; k < end; k += kWave) {
      const int k2 = k + 2 * kWave;
      G4 g2 = z4;
      G2 w2 = z2;
      int s2 = 0, p2 = 0;
      if (k2 < end) {
        g2 = tgeo[k2];
        w2 = ttw[k2];
        if (DIR) s2 = tslot[k2];
        if (ADJ) p2 = tp[k2];
      }
      R x1 = R(0);
      if (ADJ && k + kWave < end) x1 = xiw[p1];
      const R coef = ADJ ? x0 * (R)w0.y : R(0);
      const R v = sweep_triple<R, LMAX, DIR, AM0, AM1>((R)g0.x, (R)g0.y, (R)g0.z, (R)g0.w,
                                                       (R)w0.x, xsh, coef, acc);
      if (DIR) rw[s0] = (R)w0.y * v;
      g0 = g1; g1 = g2;
      w0 = w1; w1 = w2;
      s0 = s1; s1 = s2;
      x0 = x1; p1 = p2;
    }
  }
  if constexpr (ADJ) {
    // sum of acc over the 64 lanes, entry by entry, spread over the lanes
    int base = 0, len = NA;
    wave_reduce_split<0, NA, R>(acc, base, len);
    constexpr int NF = reduced_len(NA, 6);
    constexpr int NY = (LMAX + 1) * (LMAX + 1);
    const double *cs = F ? cf + 4 * NY : cf;   // cf K or cf
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      if (j < len) {
        // lm of block entry q
        const int q = base + j;
        int lm = q;
#pragma unroll
        for (int m = AM0; m < ((AM0 == 0 && AM1 == LMAX + 1) ? AM0 : AM1); ++m) {
          const int o = mspan(LMAX, AM0, m);
          if (q >= o && q < o + mcount(LMAX, m)) {
            const int r = q - o;
            const int l = m + (m == 0 ? r : (r >> 1));
            lm = lmi(l, (m > 0 && (r & 1)) ? -m : m);
          }
        }
        yp[lm] = -(cs[lm] * (double)acc[j]);
      }
    }
  }
}

// The adjoint passes of blocks B.. (the first also evaluates L X when DIR)
template <typename R, int LMAX, bool DIR, int B>
__device__ __forceinline__ void sweep_blocks(
    int beg, int end, const typename Geo<R>::G4 *__restrict__ tgeo,
    const typename Geo<R>::G2 *__restrict__ ttw,
    const int *__restrict__ tslot, const int *__restrict__ tp, LdsPtr<R> xsh,
    R *__restrict__ rw, int spare, const R *__restrict__ xiw, const double *__restrict__ cf,
    double *__restrict__ yp) {
  constexpr bool F = sizeof(R) == 4;
  constexpr int M0 = mblock_begin(LMAX, DIR, B, F), M1 = mblock_begin(LMAX, DIR, B + 1, F);
  sweep_pass<R, LMAX, DIR && B == 0, M0, M1>(beg, end, tgeo, ttw, tslot, tp, xsh, rw, spare, xiw,
                                             cf, yp);
  if constexpr (M1 <= LMAX) sweep_blocks<R, LMAX, DIR, B + 1>(beg, end, tgeo, ttw, tslot, tp,
                                                             xsh, rw, spare, xiw, cf, yp);
}

// Triples beg..end-1 of container c: rw[tslot] = W (R . X_c) when DIR
// (rw[spare]: scratch), and when ADJ the chunk's share of (L*S)_c into the
// segment row yp.
template <typename R, int LMAX, bool DIR, bool ADJ>
__device__ __forceinline__ void sweep_chunk(
    int c, int beg, int end, const typename Geo<R>::G4 *__restrict__ tgeo,
    const typename Geo<R>::G2 *__restrict__ ttw,
    const int *__restrict__ tslot, const int *__restrict__ tp, const double *__restrict__ Xs,
    R *__restrict__ rw, int spare, const R *__restrict__ xiw, const double *__restrict__ cf,
    double *__restrict__ yp) {
  constexpr int NY = (LMAX + 1) * (LMAX + 1);
  // X'_c in the wave's LDS row (rounded to R; K X' in FP32, for the scaled
  // basis of sweep_triple32): every lane reads the same word (broadcast)
  __shared__ R xsh_all[kSweepWaves][DIR ? NY : 1];
  R *xsh = xsh_all[__builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave))];
  if (DIR) {
    wave_lds_sync();   // reads of the previous segment's X' are done
    for (int k = lane_id(); k < NY; k += kWave) {
      const double x = Xs[(size_t)c * NY + k];
      xsh[k] = sizeof(R) == 4 ? (R)(x * cf[3 * NY + k]) : (R)x;
    }
    wave_lds_sync();
  }
  const LdsPtr<R> xl = (LdsPtr<R>)xsh;
  if constexpr (ADJ)
    sweep_blocks<R, LMAX, DIR, 0>(beg, end, tgeo, ttw, tslot, tp, xl, rw, spare, xiw, cf, yp);
  else
    sweep_pass<R, LMAX, true, 0, 0>(beg, end, tgeo, ttw, tslot, tp, xl, rw, spare, xiw, cf, yp);
}

// One wave per worker w (kSweepWaves to a workgroup, w < G): units
// floor(w U / G) .. floor((w+1) U / G) - 1, container by container.  DIR:
// L X terms (rw, slot order); ADJ: one partial L* S per segment (yp).  A side whose solve
// has ended (gate done) is skipped; in a batch (inst: the instance of each
// sphere, gates: the summaries after ninst instance states) so is every
// container of an instance that has converged.
// (two waves per SIMD: the register budget that keeps a block of
// accumulators and the two-deep load pipeline without spilling)
template <typename R, int LMAX, bool DIR, bool ADJ>
__global__ __launch_bounds__(kWave * kSweepWaves, 2) void k_sweep(
    long long U, long long G, const int *__restrict__ uoff, const int *__restrict__ segoff,
    const int *__restrict__ inl, const int *__restrict__ pairoff, int n,
    const typename Geo<R>::G4 *__restrict__ tgeo,
    const typename Geo<R>::G2 *__restrict__ ttw,
    const int *__restrict__ tslot, const int *__restrict__ tp, const double *__restrict__ Xs,
    R *__restrict__ rw, int spare, const R *__restrict__ xiw, const double *__restrict__ cf,
    double *__restrict__ yp, const DiisState *__restrict__ gdir,
    const DiisState *__restrict__ gadj, const int *__restrict__ inst, int ninst) {
  const bool dir = DIR && !(gdir && gdir->done);
  const bool adj = ADJ && !(gadj && gadj->done);
  if (!dir && !adj) return;
  constexpr int NY = (LMAX + 1) * (LMAX + 1);
  // (the wave index through readfirstlane: the compiler does not see that
  // threadIdx.x / 64 is uniform across a wave, and every quantity below
  // would otherwise live in vector registers and load per lane)
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
  const long long w = (long long)blockIdx.x * kSweepWaves + wave;
  if (w >= G) return;
  const int u0 = (int)(w * U / G), u1 = (int)((w + 1) * U / G);
  // container of unit u0: the last c with uoff[c] <= u0
  int lo = 0, hi = n;   // uoff[lo] <= u0 < uoff[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (uoff[mid] <= u0) lo = mid;
    else hi = mid;
  }
  int c = lo;
  while (uoff[c + 1] <= u0) ++c;   // skip containers without triples
  // global index of this worker's first segment
  int seg = segoff[c];
  if (u0 > uoff[c]) {
    const long long wlo = (((long long)uoff[c] + 1) * G + U - 1) / U;
    seg += 1 + (int)(w - wlo);
  }
  for (int u = u0; u < u1;) {
    while (uoff[c + 1] <= u) {
      ++c;
      seg = segoff[c];
    }
    const int ue = min(u1, uoff[c + 1]);
    const int cb = pairoff[inl[c]], ce = pairoff[inl[c + 1]];
    const int beg = cb + (u - uoff[c]) * kWave;
    const int end = min(ce, cb + (ue - uoff[c]) * kWave);
    double *ys = yp + (size_t)seg * NY;
    bool dc = dir, ac = adj;
    if (inst) {
      const int q = inst[c] - ninst;   // the instance's state, before the summary
      dc = dc && !(gdir && gdir[q].done);
      ac = ac && !(gadj && gadj[q].done);
    }
    if constexpr (DIR && ADJ) {
      if (dc && ac)
        sweep_chunk<R, LMAX, true, true>(c, beg, end, tgeo, ttw, tslot, tp, Xs, rw, spare, xiw, cf, ys);
      else if (dc)
        sweep_chunk<R, LMAX, true, false>(c, beg, end, tgeo, ttw, tslot, tp, Xs, rw, spare, xiw, cf, ys);
      else if (ac)
        sweep_chunk<R, LMAX, false, true>(c, beg, end, tgeo, ttw, tslot, tp, Xs, rw, spare, xiw, cf, ys);
    } else if constexpr (DIR) {
      if (dc)
        sweep_chunk<R, LMAX, true, false>(c, beg, end, tgeo, ttw, tslot, tp, Xs, rw, spare, xiw, cf, ys);
    } else {
      if (ac)
        sweep_chunk<R, LMAX, false, true>(c, beg, end, tgeo, ttw, tslot, tp, Xs, rw, spare, xiw, cf, ys);
    }
    u = ue;
    ++seg;
  }
}

// P(n,o) = sum of the point's weighted terms, in neighbour order.  Points
// with ui = 1, which calcv skips (ddcosmo.f90:1698), are exactly the points
// without triples (ui < 1 iff some t < 1), so their P is zero here as well.
template <typename R>
__global__ __launch_bounds__(256) void k_gather_points(size_t np, int ngrid,
                                                       const int *__restrict__ ptoff,
                                                       const R *__restrict__ rw,
                                                       R *__restrict__ P,
                                                       const DiisState *__restrict__ gate,
                                                       const int *__restrict__ inst, int ninst) {
  // points of spheres whose molecules have all converged are not read again
  {
    const size_t p0 = (size_t)blockIdx.x * blockDim.x;
    const size_t p1 = min(np, p0 + blockDim.x);
    if (!spheres_active(gate, inst, ninst, (int)(p0 / ngrid), (int)((p1 - 1) / ngrid) + 1))
      return;
  }
  size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int b = ptoff[p], e = ptoff[p + 1];
  // the first four terms are loaded together (most points have at most
  // four), then added in order like the rest
  R v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = (b + k < e) ? rw[b + k] : R(0);
  R s = R(0);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (b + k < e) s += v[k];
  for (int k = b + 4; k < e; ++k) s += rw[k];
  P[p] = s;
}

// (L*S)_c = sum of its segments' shares, in segment order
__global__ void k_seg_sum(int n, int nylm, const int *__restrict__ segoff,
                          const double *__restrict__ yp, double *__restrict__ y,
                          const DiisState *__restrict__ gate) {
  if (gate && gate->done) return;
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)n * nylm) return;
  const int c = (int)(k / nylm), lm = (int)(k % nylm);
  double r = 0.0;
  for (int j = segoff[c]; j < segoff[c + 1]; ++j) r += yp[(size_t)j * nylm + lm];
  y[k] = r;
}

// xs = cf * x, elementwise over (nylm, nsph)
__global__ void k_cf_scale(size_t nn, int nylm, const double *__restrict__ cf,
                           const double *__restrict__ x, double *__restrict__ xs) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nn) xs[k] = cf[k % nylm] * x[k];
}

// ---------------------------------------------------------------------------
// Far-field expansions of the Coulomb sums (used by k_fmm.hip).  The local
// expansion at target sphere i of the point charges j far from it:
//   A_i,lm = cf_lm sum_{j: d_ij^2 >= rc_i^2} (q_j / d) t^l p~_l^m(cos) trig_m,
//   t = R_i / d,  rc_i = R_i / rho
// (addition theorem: 1/|R s - v| = sum_l 4pi/(2l+1) R^l / d^(l+1)
// Y_lm(s) Y_lm(v / d) for R < d), so that the far charges' potential at the
// point R_i s of sphere i is sum_lm A_lm Y_lm(s), with a truncation error
// below (q / (d - R)) rho^(P+1) per charge.  One wave per target sphere,
// lanes over the charges of its instance, one pass per block of
// accumulators (the sweep's m-blocks), each lane's sums reduce-scattered.
template <int P, int AM0, int AM1>
__device__ __forceinline__ void p2l_pass(int i, double4 si, double rc2, int sb, int se,
                                         const double4 *__restrict__ sph,
                                         const double *__restrict__ q,
                                         const double *__restrict__ cf, double *__restrict__ A) {
  constexpr int NA = mspan(P, AM0, AM1);
  constexpr int NY = (P + 1) * (P + 1);
  double acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0.0;
  for (int j = sb + (int)threadIdx.x; j < se; j += kWave) {
    const double4 sj = sph[j];
    const double vx = sj.x - si.x, vy = sj.y - si.y, vz = sj.z - si.z;
    double d2;
    {
#pragma clang fp contract(off)
      const double ax = si.x - sj.x, ay = si.y - sj.y, az = si.z - sj.z;
      d2 = (ax * ax + ay * ay) + az * az;   // as the near lists classify the pair
    }
    double coef = 0.0, cthe = 1.0, sthe = 0.0, cphi = 1.0, sphi = 0.0, t = 0.0;
    if (d2 >= rc2 && j != i) {
      const double id = 1.0 / sqrt(d2);
      const double rxy = sqrt(vx * vx + vy * vy);
      coef = q[j] * id;
      t = si.w * id;
      cthe = vz * id;
      sthe = rxy * id;
      if (rxy > 0.0) {
        cphi = vx / rxy;
        sphi = vy / rxy;
      }
    }
    sweep_triple<double, P, false, AM0, AM1>(cthe, sthe, cphi, sphi, t, nullptr, coef, acc);
  }
  int base = 0, len = NA;
  wave_reduce_split<0, NA, double>(acc, base, len);
  constexpr int NF = reduced_len(NA, 6);
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    if (j < len) {
      const int qi = base + j;
      int lm = qi;
#pragma unroll
      for (int m = AM0; m < ((AM0 == 0 && AM1 == P + 1) ? AM0 : AM1); ++m) {
        const int o = mspan(P, AM0, m);
        if (qi >= o && qi < o + mcount(P, m)) {
          const int r = qi - o;
          const int l = m + (m == 0 ? r : (r >> 1));
          lm = lmi(l, (m > 0 && (r & 1)) ? -m : m);
        }
      }
      A[(size_t)i * NY + lm] = cf[lm] * acc[j];
    }
  }
}

template <int P, int B>
__device__ __forceinline__ void p2l_blocks(int i, double4 si, double rc2, int sb, int se,
                                           const double4 *__restrict__ sph,
                                           const double *__restrict__ q,
                                           const double *__restrict__ cf,
                                           double *__restrict__ A) {
  constexpr int M0 = mblock_begin(P, false, B, false), M1 = mblock_begin(P, false, B + 1, false);
  p2l_pass<P, M0, M1>(i, si, rc2, sb, se, sph, q, cf, A);
  if constexpr (M1 <= P) p2l_blocks<P, B + 1>(i, si, rc2, sb, se, sph, q, cf, A);
}

// sources of target sphere i: the spheres of its instance (soff, inst) or
// all n spheres
template <int P>
__global__ __launch_bounds__(kWave, 2) void k_p2l(int n, const double4 *__restrict__ sph,
                                                  const double *__restrict__ q,
                                                  const int *__restrict__ inst,
                                                  const int *__restrict__ soff, double rho,
                                                  const double *__restrict__ cf,
                                                  double *__restrict__ A) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const double4 si = sph[i];
  const double rc = si.w / rho;
  int sb = 0, se = n;
  if (inst) {
    sb = soff[inst[i]];
    se = soff[inst[i] + 1];
  }
  p2l_blocks<P, 0>(i, si, rc * rc, sb, se, sph, q, cf, A);
}

// the FP32 sweeps' per-triple records (geo32_of, and W), made once per
// setup (on the first FP32 sweep after it)
__global__ void k_geo_to_f32(long long n, const double4 *__restrict__ g,
                             const double2 *__restrict__ w, float4 *__restrict__ g32,
                             float *__restrict__ w32) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double4 a = g[k];
  const double2 b = w[k];
  g32[k] = geo32_of(a, b);
  w32[k] = (float)b.y;
}

int geometry32(ddx_handle *h, hipStream_t s) {
  if (h->have_geo32) return DDX_OK;
  DDX_TRY(h->tgeo32.reserve((size_t)h->ntri + 1));
  DDX_TRY(h->ttw32.reserve((size_t)h->ntri + 1));
  if (h->ntri > 0) {
    hipLaunchKernelGGL(k_geo_to_f32, dim3((unsigned)(((long long)h->ntri + 255) / 256)),
                       dim3(256), 0, s, (long long)h->ntri, h->tgeo.p, h->ttw.p, h->tgeo32.p,
                       h->ttw32.p);
    DDX_HIP(hipGetLastError());
  }
  h->have_geo32 = true;
  return DDX_OK;
}

template <typename R, int LMAX, bool DIR, bool ADJ>
int launch_sweep(ddx_handle *h, hipStream_t s, const double *Xs, R *rw, const R *xiw, double *yp,
                 const DiisState *gdir, const DiisState *gadj) {
  if (h->nworker == 0) return DDX_OK;
  const typename Geo<R>::G4 *tgeo;
  const typename Geo<R>::G2 *ttw;
  if constexpr (sizeof(R) == 4) {
    DDX_TRY(geometry32(h, s));
    tgeo = h->tgeo32.p;
    ttw = h->ttw32.p;
  } else {
    tgeo = h->tgeo.p;
    ttw = h->ttw.p;
  }
  TimedLaunch tl(h, (DIR && ADJ) ? DDX_K_SWEEP_PAIR : (DIR ? DDX_K_SWEEP_L : DDX_K_SWEEP_LSTAR), s);
  hipExtLaunchKernelGGL((k_sweep<R, LMAX, DIR, ADJ>),
                        dim3((h->nworker + kSweepWaves - 1) / kSweepWaves),
                        dim3(kWave * kSweepWaves), 0, s, tl.b,
                        tl.e, 0, (long long)h->nunit, (long long)h->nworker, h->uoff.p,
                        h->segoff.p, h->inl.p, h->pairoff.p, h->n, tgeo, ttw,
                        h->tslot.p, h->tp.p, Xs, rw, (int)h->ntri, xiw, h->cf.p, yp, gdir, gadj,
                        h->ninst > 1 ? h->inst.p : nullptr, h->ninst);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

template <typename R, int LMAX>
int sweep_lmax(ddx_handle *h, hipStream_t s, int mode, const double *Xs, R *rw, const R *xiw,
               double *y, const DiisState *gdir, const DiisState *gadj) {
  if (mode == 3) return launch_sweep<R, LMAX, true, true>(h, s, Xs, rw, xiw, y, gdir, gadj);
  if (mode == 1) return launch_sweep<R, LMAX, true, false>(h, s, Xs, rw, xiw, y, gdir, gadj);
  return launch_sweep<R, LMAX, false, true>(h, s, Xs, rw, xiw, y, gdir, gadj);
}

// mode: 1 = L X terms, 2 = L* S shares per segment, 3 = both.  R: the
// arithmetic of the sweep (double, or float in the increment rounds of the
// mixed precision); X' and the L* S shares stay double.
template <typename R>
int sweep(ddx_handle *h, hipStream_t s, int mode, const double *Xs, R *rw, const R *xiw,
          double *y, const DiisState *gdir, const DiisState *gadj) {
  if (h->n == 0) return DDX_OK;
  switch (h->lmax) {
    case 0: return sweep_lmax<R, 0>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 1: return sweep_lmax<R, 1>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 2: return sweep_lmax<R, 2>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 3: return sweep_lmax<R, 3>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 4: return sweep_lmax<R, 4>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 5: return sweep_lmax<R, 5>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 6: return sweep_lmax<R, 6>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 7: return sweep_lmax<R, 7>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 8: return sweep_lmax<R, 8>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 9: return sweep_lmax<R, 9>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 10: return sweep_lmax<R, 10>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 11: return sweep_lmax<R, 11>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 12: return sweep_lmax<R, 12>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 13: return sweep_lmax<R, 13>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 14: return sweep_lmax<R, 14>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 15: return sweep_lmax<R, 15>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    case 16: return sweep_lmax<R, 16>(h, s, mode, Xs, rw, xiw, y, gdir, gadj);
    default: set_error("lmax %d not supported (0..16)", h->lmax); return DDX_ERR_ARG;
  }
}

// P = gathered L X terms of each point; y = - Yw^T-projection of P
template <typename R>
int finish_lx(ddx_handle *h, Lane &ln, double *y, const DiisState *gate, int ysplit) {
  hipStream_t s = ln.stream;
  R *rw = reinterpret_cast<R *>(ln.rw.p);
  R *P = reinterpret_cast<R *>(ln.P.p);
  if constexpr (sizeof(R) == 4) {
    bool done = false;
    DDX_TRY(project32_gathered(h, s, -1.0, rw, y, gate, ysplit, &done));
    if (done) return DDX_OK;
  }
  hipLaunchKernelGGL(k_gather_points<R>, dim3((h->np + 255) / 256), dim3(256), 0, s, h->np,
                     h->ngrid, h->ptoff.p, rw, P, gate, h->ninst > 1 ? h->inst.p : nullptr,
                     h->ninst);
  DDX_HIP(hipGetLastError());
  // y(lm, o) = - sum_n Yw(lm, n) P(n, o)
  if constexpr (sizeof(R) == 4) return project32(h, s, -1.0, P, y, gate, ysplit);
  else return project(h, s, -1.0, P, y, gate, ysplit);
}

// the lane's point values of the L* S sweep: w_n xi(n, o) in R
template <typename R>
int expand_xi(ddx_handle *h, Lane &ln, const double *x, const DiisState *gate) {
  R *xiw = reinterpret_cast<R *>(ln.xiw.p);
  if constexpr (sizeof(R) == 4) return expand32(h, ln.stream, x, xiw, true, gate);
  else return expand(h, ln.stream, x, xiw, true, gate);
}

template <typename R>
int lx_r(ddx_handle *h, Lane &ln, const double *x, double *y, const DiisState *gate,
         bool xs_ready, int ysplit) {
  if (!xs_ready) DDX_TRY(cf_scale(h, ln.stream, x, ln.xs.p));
  DDX_TRY(sweep<R>(h, ln.stream, 1, ln.xs.p, reinterpret_cast<R *>(ln.rw.p), nullptr, nullptr,
                   gate, nullptr));
  return finish_lx<R>(h, ln, y, gate, ysplit);
}

template <typename R>
int lstarx_r(ddx_handle *h, Lane &ln, const double *x, double *y, const DiisState *gate,
             bool want_y) {
  // w_n xi(n, o) = w_n sum_lm Y(lm, n) x(lm, o)
  DDX_TRY(expand_xi<R>(h, ln, x, gate));
  DDX_TRY(sweep<R>(h, ln.stream, 2, nullptr, nullptr, reinterpret_cast<const R *>(ln.xiw.p),
                   ln.yp.p, nullptr, gate));
  return want_y ? seg_sum(h, ln.stream, ln.yp.p, y, gate) : DDX_OK;
}

template <typename R>
int pair_r(ddx_handle *h, Lane &ld, double *yd, const DiisState *gd, Lane &la, const double *xa,
           const DiisState *ga, int ysplit) {
  DDX_TRY(expand_xi<R>(h, la, xa, ga));
  DDX_TRY(sweep<R>(h, ld.stream, 3, ld.xs.p, reinterpret_cast<R *>(ld.rw.p),
                   reinterpret_cast<const R *>(la.xiw.p), la.yp.p, gd, ga));
  return finish_lx<R>(h, ld, yd, gd, ysplit);
}

int reserve_lx(ddx_handle *h, Lane &ln) {
  DDX_TRY(ln.rw.reserve(h->ntri + 1));
  DDX_TRY(ln.P.reserve(h->np));
  return ln.xs.reserve((size_t)h->nylm * h->n + 1);
}

int reserve_lstarx(ddx_handle *h, Lane &ln) {
  DDX_TRY(ln.xiw.reserve(h->np));
  return ln.yp.reserve((size_t)h->nseg_cap * h->nylm + 1);
}

}  // namespace

double sweep_basis_scale(int l, int m) {
  double K = 1.0;
  for (int j = 1; j <= m; ++j) K *= -(2.0 * j - 1.0);
  for (int j = m + 1; j <= l; ++j) K *= basis_ratio(j, m);
  return K;
}

int cf_scale(ddx_handle *h, hipStream_t s, const double *x, double *xs) {
  size_t nn = (size_t)h->nylm * h->n;
  if (nn == 0) return DDX_OK;
  hipLaunchKernelGGL(k_cf_scale, dim3((nn + 255) / 256), dim3(256), 0, s, nn, h->nylm, h->cf.p, x,
                     xs);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

int seg_sum(ddx_handle *h, hipStream_t s, const double *yp, double *y, const DiisState *gate) {
  size_t nn = (size_t)h->nylm * h->n;
  if (nn == 0) return DDX_OK;
  hipLaunchKernelGGL(k_seg_sum, dim3((nn + 255) / 256), dim3(256), 0, s, h->n, h->nylm,
                     h->segoff.p, yp, y, gate);
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

// The applies below run in the lane's precision: ln.f32 selects the FP32
// sweeps, gather and MFMA products of the mixed precision.
int apply_lx(ddx_handle *h, Lane &ln, const double *x, double *y, const DiisState *gate,
             bool xs_ready, int ysplit) {
  DDX_TRY(reserve_lx(h, ln));
  return ln.f32 ? lx_r<float>(h, ln, x, y, gate, xs_ready, ysplit)
                : lx_r<double>(h, ln, x, y, gate, xs_ready, ysplit);
}

int apply_lstarx(ddx_handle *h, Lane &ln, const double *x, double *y, const DiisState *gate,
                 bool want_y) {
  DDX_TRY(reserve_lstarx(h, ln));
  return ln.f32 ? lstarx_r<float>(h, ln, x, y, gate, want_y)
                : lstarx_r<double>(h, ln, x, y, gate, want_y);
}

int apply_pair(ddx_handle *h, Lane &ld, double *yd, Lane &la, const double *xa, int ysplit) {
  DDX_TRY(reserve_lx(h, ld));
  DDX_TRY(reserve_lstarx(h, la));
  const DiisState *gd = solve_gate(h, ld), *ga = solve_gate(h, la);
  return ld.f32 ? pair_r<float>(h, ld, yd, gd, la, xa, ga, ysplit)
                : pair_r<double>(h, ld, yd, gd, la, xa, ga, ysplit);
}

// A (n, (P+1)^2) of the far charges of every sphere (k_p2l); cfP: the cf
// factors to order P
int fmm_p2l(ddx_handle *h, int P, double rho, const double *q, const double *cfP, double *A) {
  if (h->n == 0) return DDX_OK;
  const int *inst = h->ninst > 1 ? h->inst.p : nullptr;
  const int *soff = h->ninst > 1 ? h->soff.p : nullptr;
  hipStream_t s = h->stream;
  switch (P) {
    case 8:
      hipLaunchKernelGGL(k_p2l<8>, dim3(h->n), dim3(kWave), 0, s, h->n, h->sph.p, q, inst, soff,
                         rho, cfP, A);
      break;
    case 12:
      hipLaunchKernelGGL(k_p2l<12>, dim3(h->n), dim3(kWave), 0, s, h->n, h->sph.p, q, inst,
                         soff, rho, cfP, A);
      break;
    case 16:
      hipLaunchKernelGGL(k_p2l<16>, dim3(h->n), dim3(kWave), 0, s, h->n, h->sph.p, q, inst,
                         soff, rho, cfP, A);
      break;
    default:
      set_error("far-field order %d not built (8, 12, 16)", P);
      return DDX_ERR_ARG;
  }
  DDX_HIP(hipGetLastError());
  return DDX_OK;
}

}  // namespace ddx
