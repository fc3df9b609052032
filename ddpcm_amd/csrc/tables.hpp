// Host-side one-time tables (Lebedev grid, harmonic normalisation, Y_lm).
#pragma once

#include <vector>

namespace ddx {

// nearest of the 32 Lebedev sizes, first one on ties (ddcosmo.f90:188-199)
int snap_ngrid(int ngrid);
// points (3, ngrid) and weights (sum 4 pi); 0 on success
int lebedev_grid(int ngrid, std::vector<double> &grid, std::vector<double> &w);
// facl = (2l+1)/4pi and facs (ddcosmo.f90:225-243)
void harmonic_factors(int lmax, std::vector<double> &facl, std::vector<double> &facs);
// real spherical harmonics y[nylm] at unit vector s (ylmbas)
void real_harmonics(int lmax, const double *facs, const double *s, double *y);

}  // namespace ddx
