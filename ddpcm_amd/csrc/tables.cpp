// One-time tables of a ddCOSMO system: Lebedev grid, normalisation factors and
// the real spherical harmonics at the grid points (ddinit, ddcosmo.f90:173-267).
#include <cmath>
#include <vector>

#include "tables.hpp"

namespace ddx {
namespace {
#include "lebedev_table.inc"

// Octahedral orbit of one Lebedev parameter set, in the point order of the
// reference's gen_oh (llgnew.f:93-680): base permutations in a fixed order,
// then the signs of the non-zero coordinates with the first flipping fastest.
int expand_orbit(int code, double a, double b, double v, double *xyzw) {
  double base[6][3];
  int nb = 0;
  switch (code) {
    case 1: {
      double t[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      for (int i = 0; i < 3; ++i)
        for (int d = 0; d < 3; ++d) base[i][d] = t[i][d];
      nb = 3;
      break;
    }
    case 2: {
      double q = std::sqrt(0.5);
      double t[3][3] = {{0, q, q}, {q, 0, q}, {q, q, 0}};
      for (int i = 0; i < 3; ++i)
        for (int d = 0; d < 3; ++d) base[i][d] = t[i][d];
      nb = 3;
      break;
    }
    case 3: {
      double q = std::sqrt(1.0 / 3.0);
      base[0][0] = base[0][1] = base[0][2] = q;
      nb = 1;
      break;
    }
    case 4: {
      double c = std::sqrt(1.0 - 2.0 * a * a);
      double t[3][3] = {{a, a, c}, {a, c, a}, {c, a, a}};
      for (int i = 0; i < 3; ++i)
        for (int d = 0; d < 3; ++d) base[i][d] = t[i][d];
      nb = 3;
      break;
    }
    case 5: {
      double c = std::sqrt(1.0 - a * a);
      double t[6][3] = {{a, c, 0}, {c, a, 0}, {a, 0, c}, {c, 0, a}, {0, a, c}, {0, c, a}};
      for (int i = 0; i < 6; ++i)
        for (int d = 0; d < 3; ++d) base[i][d] = t[i][d];
      nb = 6;
      break;
    }
    case 6: {
      double c = std::sqrt(1.0 - a * a - b * b);
      double t[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
      for (int i = 0; i < 6; ++i)
        for (int d = 0; d < 3; ++d) base[i][d] = t[i][d];
      nb = 6;
      break;
    }
    default:
      return -1;
  }
  int np = 0;
  for (int ib = 0; ib < nb; ++ib) {
    int nz[3], k = 0;
    for (int d = 0; d < 3; ++d)
      if (base[ib][d] != 0.0) nz[k++] = d;
    for (int mask = 0; mask < (1 << k); ++mask) {
      double p[3] = {base[ib][0], base[ib][1], base[ib][2]};
      for (int q = 0; q < k; ++q)
        if (mask & (1 << q)) p[nz[q]] = -p[nz[q]];
      xyzw[4 * np + 0] = p[0];
      xyzw[4 * np + 1] = p[1];
      xyzw[4 * np + 2] = p[2];
      xyzw[4 * np + 3] = v;
      ++np;
    }
  }
  return np;
}

}  // namespace

int snap_ngrid(int ngrid) {
  int best = -1, dist = 1 << 30;
  for (int g = 0; g < kLebedevNumGrids; ++g) {
    int d = kLebedevGrids[g][0] - ngrid;
    if (d < 0) d = -d;
    if (d < dist) { dist = d; best = kLebedevGrids[g][0]; }
  }
  return best;
}

int lebedev_grid(int ngrid, std::vector<double> &grid, std::vector<double> &w) {
  for (int g = 0; g < kLebedevNumGrids; ++g) {
    if (kLebedevGrids[g][0] != ngrid) continue;
    grid.assign(3 * (size_t)ngrid, 0.0);
    w.assign(ngrid, 0.0);
    const double pi = 4.0 * std::atan(1.0);
    double buf[4 * 48];
    int n = 0;
    for (int o = 0; o < kLebedevGrids[g][2]; ++o) {
      const auto &orb = kLebedevOrbits[kLebedevGrids[g][1] + o];
      int np = expand_orbit(orb.code, orb.a, orb.b, orb.v, buf);
      if (np < 0 || n + np > ngrid) return -1;
      for (int q = 0; q < np; ++q, ++n) {
        grid[3 * n + 0] = buf[4 * q + 0];
        grid[3 * n + 1] = buf[4 * q + 1];
        grid[3 * n + 2] = buf[4 * q + 2];
        w[n] = buf[4 * q + 3] * 4.0 * pi;   // weights sum to 4 pi (llgnew.f:81-87)
      }
    }
    return n == ngrid ? 0 : -1;
  }
  return -1;
}

void harmonic_factors(int lmax, std::vector<double> &facl, std::vector<double> &facs) {
  const int nylm = (lmax + 1) * (lmax + 1);
  const double pi = 4.0 * std::atan(1.0), sq2 = std::sqrt(2.0);
  std::vector<double> fact(2 * lmax + 2, 1.0);
  for (int i = 2; i <= 2 * lmax; ++i) fact[i] = (double)i * fact[i - 1];   // fact[i] = i!
  facl.assign(nylm, 0.0);
  facs.assign(nylm, 0.0);
  for (int l = 0; l <= lmax; ++l) {
    int ind = l * l + l;
    double fl = (2.0 * l + 1.0) / (4.0 * pi);
    double ffl = std::sqrt(fl);
    for (int m = -l; m <= l; ++m) facl[ind + m] = fl;
    facs[ind] = ffl;
    for (int m = 1; m <= l; ++m) {
      double fnorm = sq2 * ffl * std::sqrt(fact[l - m] / fact[l + m]);
      if (m % 2 == 1) fnorm = -fnorm;   // cancels the (-1)^m of the Legendre recurrence
      facs[ind + m] = fnorm;
      facs[ind - m] = fnorm;
    }
  }
}

// Y_lm at a unit vector: Legendre recurrence m outer / l inner, Chebyshev
// cos/sin(m phi), normalisation facs (ylmbas, ddcosmo.f90:830-895).
void real_harmonics(int lmax, const double *facs, const double *s, double *y) {
  const double cthe = s[2];
  const double sthe = std::sqrt(1.0 - cthe * cthe);
  std::vector<double> cs(lmax + 2, 1.0), sn(lmax + 2, 0.0);
  if (sthe != 0.0) {
    double cphi = s[0] / sthe, sphi = s[1] / sthe;
    cs[0] = 1.0;
    sn[0] = 0.0;
    if (lmax >= 1) { cs[1] = cphi; sn[1] = sphi; }
    for (int m = 2; m <= lmax; ++m) {
      cs[m] = 2.0 * cphi * cs[m - 1] - cs[m - 2];
      sn[m] = 2.0 * cphi * sn[m - 1] - sn[m - 2];
    }
  }
  std::vector<double> plm((lmax + 1) * (lmax + 1), 0.0);
  double pmm = 1.0, fact = 1.0;
  for (int m = 0; m <= lmax; ++m) {
    plm[m * m + 2 * m] = pmm;
    if (m == lmax) break;
    double p0 = pmm, p1 = cthe * (2.0 * m + 1.0) * pmm;
    plm[(m + 1) * (m + 1) + (m + 1) + m] = p1;
    for (int l = m + 2; l <= lmax; ++l) {
      double p2 = (cthe * (2.0 * l - 1.0) * p1 - (l + m - 1.0) * p0) / (double)(l - m);
      plm[l * l + l + m] = p2;
      p0 = p1;
      p1 = p2;
    }
    pmm = -pmm * fact * sthe;
    fact += 2.0;
  }
  for (int l = 0; l <= lmax; ++l) {
    int ind = l * l + l;
    y[ind] = facs[ind] * plm[ind];
    for (int m = 1; m <= l; ++m) {
      double p = plm[ind + m];
      y[ind + m] = facs[ind + m] * p * cs[m];
      y[ind - m] = facs[ind - m] * p * sn[m];
    }
  }
}

}  // namespace ddx
