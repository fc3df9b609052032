/*
 * ORACLE — test infrastructure only.
 *
 * Lebedev-Laikov grids for the oracle: expands the published orbit
 * parameters (shared data table, lebedev_table.inc) into points and weights,
 * in the point order of the reference's gen_oh (llgnew.f:93-680) and with the
 * weights scaled by 4 pi as llgrid does (llgnew.f:81-87).
 *
 * Orbit ordering: for each orbit code the base coordinate permutations below
 * are visited in order; for each base point the signs of its non-zero
 * coordinates are enumerated with the first non-zero coordinate flipping
 * fastest.
 *
 * Pinned by: the full-precision grid/weight dump of the compiled reference
 * (tests/golden), the printed grid of tests/formaldehyde.log:1275-1500,
 * sum(w) = 4 pi and exact integration of products of spherical harmonics up
 * to the grid's degree.
 */
#include <math.h>
#include "ddcosmo_oracle.h"
#include "../ddpcm_amd/csrc/lebedev_table.inc"

static int expand_orbit(int code, double a, double b, double v,
                        double *x, double *y, double *z, double *wt) {
  double base[6][3];
  int nb = 0;
  double c;
  switch (code) {
    case 1:
      base[0][0] = 1; base[0][1] = 0; base[0][2] = 0;
      base[1][0] = 0; base[1][1] = 1; base[1][2] = 0;
      base[2][0] = 0; base[2][1] = 0; base[2][2] = 1;
      nb = 3; break;
    case 2:
      a = sqrt(0.5);
      base[0][0] = 0; base[0][1] = a; base[0][2] = a;
      base[1][0] = a; base[1][1] = 0; base[1][2] = a;
      base[2][0] = a; base[2][1] = a; base[2][2] = 0;
      nb = 3; break;
    case 3:
      a = sqrt(1.0 / 3.0);
      base[0][0] = a; base[0][1] = a; base[0][2] = a;
      nb = 1; break;
    case 4:
      b = sqrt(1.0 - 2.0 * a * a);
      base[0][0] = a; base[0][1] = a; base[0][2] = b;
      base[1][0] = a; base[1][1] = b; base[1][2] = a;
      base[2][0] = b; base[2][1] = a; base[2][2] = a;
      nb = 3; break;
    case 5:
      b = sqrt(1.0 - a * a);
      base[0][0] = a; base[0][1] = b; base[0][2] = 0;
      base[1][0] = b; base[1][1] = a; base[1][2] = 0;
      base[2][0] = a; base[2][1] = 0; base[2][2] = b;
      base[3][0] = b; base[3][1] = 0; base[3][2] = a;
      base[4][0] = 0; base[4][1] = a; base[4][2] = b;
      base[5][0] = 0; base[5][1] = b; base[5][2] = a;
      nb = 6; break;
    case 6:
      c = sqrt(1.0 - a * a - b * b);
      base[0][0] = a; base[0][1] = b; base[0][2] = c;
      base[1][0] = a; base[1][1] = c; base[1][2] = b;
      base[2][0] = b; base[2][1] = a; base[2][2] = c;
      base[3][0] = b; base[3][1] = c; base[3][2] = a;
      base[4][0] = c; base[4][1] = a; base[4][2] = b;
      base[5][0] = c; base[5][1] = b; base[5][2] = a;
      nb = 6; break;
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
      x[np] = p[0]; y[np] = p[1]; z[np] = p[2]; wt[np] = v;
      ++np;
    }
  }
  return np;
}

int orc_lebedev(int ngrid, double *w, double *grid) {
  for (int g = 0; g < kLebedevNumGrids; ++g) {
    if (kLebedevGrids[g][0] != ngrid) continue;
    double x[48], y[48], z[48], wt[48];
    int off = kLebedevGrids[g][1], no = kLebedevGrids[g][2], n = 0;
    double pi = 4.0 * atan(1.0);
    for (int o = 0; o < no; ++o) {
      int np = expand_orbit(kLebedevOrbits[off + o].code, kLebedevOrbits[off + o].a,
                            kLebedevOrbits[off + o].b, kLebedevOrbits[off + o].v,
                            x, y, z, wt);
      if (np < 0) return -1;
      for (int p = 0; p < np; ++p, ++n) {
        grid[3 * n + 0] = x[p];
        grid[3 * n + 1] = y[p];
        grid[3 * n + 2] = z[p];
        w[n] = wt[p] * 4.0 * pi;
      }
    }
    return n == ngrid ? 0 : -1;
  }
  return -1;
}

/* nearest of the 32 grid sizes, first one on ties (ddcosmo.f90:188-199) */
int orc_snap_ngrid(int ngrid) {
  int best = -1, dist = 100000;
  for (int g = 0; g < kLebedevNumGrids; ++g) {
    int d = kLebedevGrids[g][0] - ngrid;
    if (d < 0) d = -d;
    if (d < dist) { dist = d; best = kLebedevGrids[g][0]; }
  }
  return best;
}
